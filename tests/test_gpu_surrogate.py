"""The learned NFW deflection surrogate on the GPU: its training gradients against the CPU
oracle (oracle/mlp.py), HIP-graph training against eager training, its deflection error against
FP64 quadrature, the quadrature fallback outside its training box, and what it gains over
quadrature in the ray shooting (speed, at the error it has).

Accuracy contract of the network (4 -> 128 x4 -> 5, BF16 weights): median relative deflection
error below 3e-3 and 99th percentile below 3e-2 inside the training box.  At the noise level of
the 200x200 workload that error is far above what |delta ln E| < 0.1 allows, so evidence
evaluations keep quadrature unless a lens asks for the surrogate."""
import numpy as np
import pytest

from oracle import mlp
from qlens_amd import qlx
from qlens_amd import workloads as W

pytestmark = pytest.mark.gpu

OFFS = [0, 512, 640, 17024, 17152, 33536, 33664, 50048, 50176, 50816, 50821]
NAMES = ["W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4", "W5", "b5"]


def test_surrogate_gradient_matches_oracle(qlx_ctx):
    rng = np.random.default_rng(3)
    P = (rng.standard_normal(mlp.NPARAM) * 0.1).astype(np.float32)
    qlx_ctx.set_surrogate_params(P)
    assert np.array_equal(qlx_ctx.surrogate_params(), P)
    X = rng.uniform(-1, 1, (200, 4)).astype(np.float32)  # 200: a partial last tile of 64 samples
    T = rng.standard_normal((200, 5)).astype(np.float32)
    lg, gg = qlx_ctx.surrogate_grad(X, T)
    lo, go = mlp.loss_grad(P.astype(np.float64), X.astype(np.float64), T.astype(np.float64), emulate=True)
    assert abs(lg - lo) <= 2e-6 * abs(lo), (lg, lo)
    for k in range(10):
        a, b = OFFS[k], OFFS[k + 1]
        err = np.linalg.norm(gg[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        assert err < 2e-3, (NAMES[k], err)


def test_surrogate_graph_training_matches_eager(qlx_ctx):
    """One optimizer step captured in a HIP graph and replayed gives the eager parameters exactly,
    for several replay counts with eager work and synchronization between the runs."""
    for steps in (1, 6, 17):
        la = qlx_ctx.surrogate_train(n_samples=4096, steps=steps, batch=1024, seed=11, graph=False)
        pa = qlx_ctx.surrogate_params()
        lb = qlx_ctx.surrogate_train(n_samples=4096, steps=steps, batch=1024, seed=11, graph=True)
        pb = qlx_ctx.surrogate_params()
        assert la == lb and np.array_equal(pa, pb), steps


@pytest.fixture(scope="module")
def trained(qlx_ctx):
    loss = qlx_ctx.surrogate_train(n_samples=2000000, steps=12000, batch=8192, lr=2e-3, seed=7)
    return loss


def _quad_reference(inv, n, seed):
    rng = np.random.default_rng(seed)
    lnr = rng.uniform(np.log(1e-3), np.log(20), n)
    phi = rng.uniform(0, 2 * np.pi, n)
    q = rng.uniform(0.3, 1, n)
    lrs = rng.uniform(0, np.log(20), n)
    ref = []
    for i in range(n):
        r = np.exp(lnr[i])
        xy = np.array([[r * np.cos(phi[i]), r * np.sin(phi[i])]])
        ref.append(inv.nfw_def_hess((1.0, np.exp(lrs[i]), q[i], 0.0, 0.0, 0.0), xy, tol=1e-10)[0])
    return np.stack([lnr, phi, q, lrs], 1), np.array(ref)


def test_surrogate_deflection_error(qlx_ctx, trained):
    x4, ref = _quad_reference(qlx_ctx, 300, 5)
    ys = qlx_ctx.surrogate_eval(x4).astype(np.float64)
    ae = np.hypot(ys[:, 0] - ref[:, 0], ys[:, 1] - ref[:, 1]) / np.hypot(ref[:, 0], ref[:, 1])
    he = np.abs(ys[:, 2:] - ref[:, 2:]).max(1) / np.abs(ref[:, 2:]).max(1)
    print("surrogate deflection rel err: median %.3g p99 %.3g; Hessian median %.3g" % (
        np.median(ae), np.percentile(ae, 99), np.median(he)))
    assert np.median(ae) < 3e-3 and np.percentile(ae, 99) < 3e-2
    assert np.median(he) < 1e-2


def test_surrogate_per_pixel_deflection_error_c3(qlx_ctx, trained):
    """The network judged by quadrature on every ray of the 400x400 / 4x4 image (2.56 M rays, the
    NFW lens of the numerical-profile workload): the source-plane points of an evaluation with the
    learned NFW deflection against the quadrature evaluation's (the SIE and shear parts are the
    same closed forms in both, so the difference is the NFW deflection error), relative to the
    NFW deflection itself (FP64 quadrature at the same points); per pixel the worst of its 16 rays.
    The network's stated contract, median < 3e-3 and 99th percentile < 3e-2, over the image."""
    case = W.c3()
    nx = len(case["xvals"]) - 1
    qlx_ctx.load_image(case["xvals"], case["yvals"], np.zeros((nx, nx)), sigma=case["sigma"])
    qlx_ctx.set_psf(case["psf"])
    qlx_ctx.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"])
    data = qlx_ctx.mock_image(case["lenses"], case["sources"], case["sigma"], case["seed"])
    qlx_ctx.load_workload(case, data)
    rays = {}
    for mode in (qlx.QLX_DEF_QUADRATURE, qlx.QLX_DEF_SURROGATE):
        qlx_ctx.invert_image_surface_brightness_map(case["lenses"], def_mode=mode)
        rays[mode] = qlx_ctx.get_rays(0)
    ns = case["nsplit"]
    xv, yv = case["xvals"], case["yvals"]
    u = (1 + 2 * np.arange(ns)) / (2 * ns)
    sx = (1 - u)[None, :] * xv[:-1, None] + u[None, :] * xv[1:, None]  # [i, ii]
    sy = (1 - u)[None, :] * yv[:-1, None] + u[None, :] * yv[1:, None]  # [j, jj]
    pts = np.zeros((nx, nx, ns, ns, 2))
    pts[..., 0] = sx[:, None, :, None]
    pts[..., 1] = sy[None, :, None, :]
    pts = pts.reshape(-1, 2)
    nfw = [lz for lz in case["lenses"] if lz[0] == W.NFW]
    a_nfw = pts - qlx_ctx.raytrace(nfw, pts)
    d = np.hypot(*(rays[qlx.QLX_DEF_SURROGATE] - rays[qlx.QLX_DEF_QUADRATURE]).reshape(-1, 2).T)
    rel = (d / np.hypot(*a_nfw.T)).reshape(nx * nx, ns * ns).max(1)
    print("C3 per-pixel NFW deflection error (worst of 16 rays): median %.3g, p99 %.3g, max %.3g; "
          "absolute median %.3g arcsec" % (np.median(rel), np.percentile(rel, 99), rel.max(), np.median(d)))
    assert np.median(rel) < 3e-3 and np.percentile(rel, 99) < 3e-2


def test_surrogate_falls_back_to_quadrature_outside_box(qlx_ctx, golden_dir, trained):
    g = np.load(golden_dir + "/inversion_C2s.npz")
    case = W.c2_small()
    qlx_ctx.load_workload(case, g["data"])
    (t0, p0), rest = case["lenses"][0], case["lenses"][1:]
    model = [(t0, (p0[0], 30.0) + tuple(p0[2:]))] + list(rest)  # rs = 30: outside the box
    rq = qlx_ctx.invert_image_surface_brightness_map(model)[0]
    rs = qlx_ctx.invert_image_surface_brightness_map(model, def_mode=qlx.QLX_DEF_SURROGATE)[0]
    assert rq["minus2lnE"] == rs["minus2lnE"] and rq["status"] == 0


def test_mixed_precision_at_scale_c5(qlx_ctx, trained):
    """400x400 image at 0.02"/pix, 100x100 source grid (F is 10000 x 10000): FP64 with quadrature
    deflections (A), mixed precision with quadrature (B) and mixed precision with the learned
    deflection (C) on a batch of 16 proposals (the configuration's batch size); B's FP64
    refinement must reach the configured 1e-10 relative residual on every item.  B against A is the precision error, asserted within the
    stated |delta lnE| < 0.1 (A is the FP64 path pinned to the reference at 1e-10 on C1/C2s; the
    CPU oracle would need hours at this size); C against A is stated, not asserted: the network's
    ~1e-3 deflection error dominates it."""
    case = W.c5()
    nx = len(case["xvals"]) - 1
    qlx_ctx.load_image(case["xvals"], case["yvals"], np.zeros((nx, nx)), sigma=case["sigma"])
    qlx_ctx.set_psf(case["psf"])
    qlx_ctx.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"])
    data = qlx_ctx.mock_image(case["lenses"], case["sources"], case["sigma"], case["seed"])
    props = W.proposals(case, 16, seed=55)
    out = {}
    for tag, prec, mode in (("A", qlx.QLX_PREC_FP64, qlx.QLX_DEF_QUADRATURE),
                            ("B", qlx.QLX_PREC_MIXED, qlx.QLX_DEF_QUADRATURE),
                            ("C", qlx.QLX_PREC_MIXED, qlx.QLX_DEF_SURROGATE)):
        qlx_ctx.load_workload(case, data, precision=prec)
        res = qlx_ctx.invert_image_surface_brightness_map(props, def_mode=mode)
        src = np.stack([qlx_ctx.get_source(b)["src"] for b in range(len(props))])
        out[tag] = (res, src, qlx_ctx.kernel_time("total")[0])
        for r in res:
            assert r["status"] == 0 and np.isfinite(r["minus2lnE"]), (tag, r)
            if tag == "B":
                assert r["refine_resid"] <= 1e-10, r
    A, B, Cs = out["A"], out["B"], out["C"]
    assert A[0][0]["n_src_active"] > 8000, A[0][0]["n_src_active"]
    dB = max(abs(b["minus2lnE"] - a["minus2lnE"]) / 2 for a, b in zip(A[0], B[0]))
    dC = max(abs(c["minus2lnE"] - a["minus2lnE"]) / 2 for a, c in zip(A[0], Cs[0]))
    scale = np.abs(A[1]).max()
    eB = np.abs(B[1] - A[1]).max() / scale
    eC = np.abs(Cs[1] - A[1]).max() / scale
    print("C5 n_src %d n_img %d: FP64 %.1f ms, mixed %.1f ms, mixed+surrogate %.1f ms per 16 evaluations"
          % (A[0][0]["n_src_active"], A[0][0]["n_img_active"], A[2], B[2], Cs[2]))
    print("C5 mixed: |dlnE| %.3g, source max err %.3g (relative to max |s|); mixed+surrogate: |dlnE| %.3g, "
          "source max err %.3g" % (dB, eB, dC, eC))
    assert dB < 0.1
    assert eB < 1e-3


def test_required_deflection_accuracy(qlx_ctx):
    """What |delta lnE| < 0.1 asks of any deflection approximation on the 200x200 / 60x60
    workload: the NFW deflection is linear in ks, so ks x (1 + eps) is a systematic relative
    deflection error of exactly eps.  Measured: |delta lnE| ~ 1e5 eps (eps = 1e-3 -> ~96, 1e-5 ->
    ~1.0, 1e-6 -> ~0.09), so the contract needs a systematic error at the 1e-6 level, three orders
    below the network's ~1e-3 (test_surrogate_deflection_error) -- the reason the evidence keeps
    quadrature."""
    case = W.c2()
    nx = len(case["xvals"]) - 1
    qlx_ctx.load_image(case["xvals"], case["yvals"], np.zeros((nx, nx)), sigma=case["sigma"])
    qlx_ctx.set_psf(case["psf"])
    qlx_ctx.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"])
    data = qlx_ctx.mock_image(case["lenses"], case["sources"], case["sigma"], case["seed"])
    qlx_ctx.load_workload(case, data)
    props = W.proposals(case, 8, seed=5)
    base = qlx_ctx.invert_image_surface_brightness_map(props)
    worst = {}
    epss = (1e-3, 1e-4, 1e-5, 1e-6)
    for eps in epss:
        models = [[(t, ((p[0] * (1 + eps),) + tuple(p[1:])) if t == W.NFW else p) for t, p in m] for m in props]
        res = qlx_ctx.invert_image_surface_brightness_map(models)
        worst[eps] = max(abs(a["minus2lnE"] - b["minus2lnE"]) / 2 for a, b in zip(res, base))
        print("systematic deflection error %g: max |dlnE| %.4g" % (eps, worst[eps]))
    # roughly linear in eps over three decades, and the 0.1 contract sits near eps = 1e-6
    for e, e10 in zip(epss[:-1], epss[1:]):
        assert 3 < worst[e] / worst[e10] < 30, (e, worst)
    assert worst[1e-5] > 0.1 and worst[1e-3] > 10


def test_surrogate_gain_over_quadrature(qlx_ctx, trained):
    """2.56 M rays of the numerical-profile lens (400x400 pixels, 4x4 supersampling): the time of
    the surrogate deflection stage plus the remaining ray shooting, against quadrature ray
    shooting, on the same evaluations.  A documented comparison: on MI355X the tabulated-node
    quadrature runs ~3.6 G rays/s and the network stage (512 SiLUs per ray, two quarter-rate
    transcendentals each) ~3 G rays/s, so the gain is about 0.85x; asserted is the network's own
    throughput floor (a regression guard), and the ratio is printed."""
    case = W.c3()
    nx = len(case["xvals"]) - 1
    qlx_ctx.load_image(case["xvals"], case["yvals"], np.zeros((nx, nx)), sigma=case["sigma"])
    qlx_ctx.set_psf(case["psf"])
    qlx_ctx.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"])
    data = qlx_ctx.mock_image(case["lenses"], case["sources"], case["sigma"], case["seed"])
    qlx_ctx.load_workload(case, data)
    models = [case["lenses"]] * 4
    t = {}
    for mode in (qlx.QLX_DEF_QUADRATURE, qlx.QLX_DEF_SURROGATE):
        qlx_ctx.invert_image_surface_brightness_map(models, def_mode=mode)
        qlx_ctx.invert_image_surface_brightness_map(models, def_mode=mode)
        ms = qlx_ctx.kernel_time("raytrace")[0]
        if mode == qlx.QLX_DEF_SURROGATE:
            ms += qlx_ctx.kernel_time("surrogate")[0]
        t[mode] = ms
    gain = t[qlx.QLX_DEF_QUADRATURE] / t[qlx.QLX_DEF_SURROGATE]
    print("ray shooting 4 x 2.56 M rays: quadrature %.3f ms, surrogate %.3f ms, gain %.2fx" % (
        t[qlx.QLX_DEF_QUADRATURE], t[qlx.QLX_DEF_SURROGATE], gain))
    rays = 4 * nx * nx * 16
    assert rays / (t[qlx.QLX_DEF_SURROGATE] * 1e-3) > 1.5e9  # network + remaining ray work, rays/s
