"""GPU parity: the HIP path (through the C ABI) against the compiled reference's golden vectors
and the oracle.  Bar: integer/index outputs identical; FP64 results within 1e-10 relative.

Source amplitudes and the model image are compared as whole vectors: ||x - x_ref|| / ||x_ref||
within 1e-10.  A per-cell relative bound would be meaningless for cells whose amplitude is ~0
(the solution is regularized, so edge cells sit at noise level and their relative error is
unbounded); the vector norm is what the evidence terms (chi^2, s^T R s) are built from."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from qlens_amd import qlx
from qlens_amd import workloads as W

pytestmark = pytest.mark.gpu

RTOL_EVIDENCE = 1e-10


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _mfma_got(out, row_of):
    got = np.zeros((16, 16))
    for lane in range(64):
        for v in range(4):
            got[row_of(lane, v), lane % 16] = out[lane * 4 + v]
    return got


def test_mfma_f64_accumulator_layout(qlx_ctx):
    out = qlx_ctx.selftest_mfma(0)
    A = np.array([[i + 1 + 0.01 * k for k in range(4)] for i in range(16)])
    B = np.array([[1 + k + 0.1 * j for j in range(16)] for k in range(4)])
    got = _mfma_got(out, lambda l, v: l // 16 + 4 * v)
    assert np.allclose(got, A @ B, rtol=1e-14), "row = lane/16 + 4*reg, col = lane%16 layout does not hold"


def test_mfma_f32_accumulator_layout(qlx_ctx):
    out = qlx_ctx.selftest_mfma(1)
    A = np.array([[np.float32(i + 1 + np.float32(0.01) * k) for k in range(4)] for i in range(16)], dtype=np.float64)
    B = np.array([[np.float32(1 + k + np.float32(0.1) * j) for j in range(16)] for k in range(4)], dtype=np.float64)
    got = _mfma_got(out, lambda l, v: 4 * (l // 16) + v)
    assert np.allclose(got, A @ B, rtol=1e-6), "row = 4*(lane/16) + reg, col = lane%16 layout does not hold"


def test_mfma_bf16_operand_and_accumulator_layout(qlx_ctx):
    out = qlx_ctx.selftest_mfma(2)
    A = np.array([[(i + 1) * (k % 4 + 1) / 4 for k in range(32)] for i in range(16)])
    B = np.array([[(j + 1) * ((k // 4) % 8 + 1) / 8 for j in range(16)] for k in range(32)])
    got = _mfma_got(out, lambda l, v: 4 * (l // 16) + v)
    # every operand value is exact in bf16 and every partial sum exact in fp32
    assert np.array_equal(got, A @ B), "bf16 16x16x32 operand / accumulator layout does not hold"


def test_mfma_f16_operand_and_accumulator_layout(qlx_ctx):
    # the mixed-path SYRK (FP16 pool) relies on this layout
    out = qlx_ctx.selftest_mfma(3)
    A = np.array([[(i + 1) * (k % 4 + 1) / 4 for k in range(32)] for i in range(16)])
    B = np.array([[(j + 1) * ((k // 4) % 8 + 1) / 8 for j in range(16)] for k in range(32)])
    got = _mfma_got(out, lambda l, v: 4 * (l // 16) + v)
    assert np.array_equal(got, A @ B), "f16 16x16x32 operand / accumulator layout does not hold"


def test_raytrace_matches_oracle(qlx_ctx, golden_dir):
    g = np.load(os.path.join(golden_dir, "deflection.npz"))
    sets = json.load(open(os.path.join(golden_dir, "deflection_lenses.json")))
    pts = g["points"]
    keep = np.abs(pts).sum(1) > 0  # exact SIE centre is 0/0 in the reference
    qlx_ctx.set_settings(2, (-1, 1, -1, 1), (10, 10), 1e3, 1e-6)
    for name, lenses in sets.items():
        lenses = [(int(t), tuple(p)) for t, p in lenses]
        beta = qlx_ctx.raytrace(lenses, pts[keep])
        ref = pts[keep] - g[name + "_total"][keep]
        assert np.allclose(beta, ref, rtol=1e-12, atol=1e-13), name


def test_nfw_hessian_matches_reference(qlx_ctx, golden_dir):
    """Lens-frame deflection + Hessian from the GPU quadrature kernel, rotated to the sky frame,
    against the reference's LensProfile::deflection / hessian (psi_xy included).  Deflections agree
    to 1e-10.  The Hessian's K integrals evaluate kappa'(x^2) = -(ks/r^2)(x^2(2 - 3F) + 1)/(x^2 - 1)^2
    (models.cpp:986-994), which loses up to 10 digits to cancellation at nodes with x^2 near 1 (down
    to the series window |x^2 - 1| < 1e-5); the device atan/atanh differ from the host libm's in the
    last place, so the bound there is the quadrature tolerance (1e-6) rather than 1e-10.  The
    evidence never uses kappa' (only the surrogate's training targets do)."""
    g = np.load(os.path.join(golden_dir, "deflection.npz"))
    sets = json.load(open(os.path.join(golden_dir, "deflection_lenses.json")))
    deg = 0.017453292519943295
    for name, lenses in sets.items():
        pts = g["points"]
        for l, (t, p) in enumerate(lenses):
            if t != W.NFW or p[2] == 1.0:
                continue
            ks, rs, q, th, xc, yc = p
            sinth, costh = np.cos(deg * th), -np.sin(deg * th)
            R = np.array([[costh, sinth], [-sinth, costh]])
            x, y = pts[:, 0] - xc, pts[:, 1] - yc
            xy = np.stack([x * costh + y * sinth, -x * sinth + y * costh], 1)
            keep = np.abs(xy).max(1) > 1e-14
            out = qlx_ctx.nfw_def_hess(p, xy[keep], tol=1e-6)
            Hl = np.zeros((len(out), 2, 2))
            Hl[:, 0, 0], Hl[:, 0, 1], Hl[:, 1, 0], Hl[:, 1, 1] = out[:, 2], out[:, 3], out[:, 3], out[:, 4]
            H = np.einsum("ji,njk,kl->nil", R, Hl, R)
            ref_h = g[name + "_hess"][l][keep]
            eh = np.abs(H - ref_h).max(axis=(1, 2)) / np.abs(ref_h).max(axis=(1, 2))
            print("%s: Hessian max rel err %.3g (median %.3g, %d of %d points above 1e-10)" % (
                name, eh.max(), np.median(eh), int((eh > 1e-10).sum()), len(eh)))
            assert eh.max() < 1e-6, name
            assert np.allclose(out[:, :2] @ R, g[name + "_per_lens"][l][keep], rtol=1e-10, atol=1e-12), name


def test_nfw_def_hess_matches_oracle(qlx_ctx):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-6, 6, size=(300, 2))
    lens = (0.2, 5.0, 0.75, 0.0, 0.0, 0.0)
    got = qlx_ctx.nfw_def_hess(lens, pts, tol=1e-6)
    ref = O.nfw_def_hess(lens, pts, tol=1e-6)
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-13)


def cns_case(g):
    """The non-square masked workload with the reference's generated anisotropic PSF."""
    case = W.c_ns()
    case["psf"] = g["psf_used"]
    return case


CASES = {"C1": W.c1, "C2s": W.c2_small, "C2": W.c2, "C4s": lambda: W.c1_geometry_200(20), "CSPL": W.c_sple,
         "CSPS": W.c_sple_steep}


def _case(name, g):
    return cns_case(g) if name == "CNS" else CASES[name]()


def vrel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", ["C1", "C2s", "CNS", "C2", "C4s", "CSPL", "CSPS"])
def test_mock_image_matches_reference(qlx_ctx, golden_dir, name):
    """The GPU mock generator (ray-traced Sersic SB per subpixel, PSF, NR3 noise at seed 10) against
    the reference's own mock of the same workload (the fixture's data)."""
    g = np.load(os.path.join(golden_dir, "inversion_%s.npz" % name))
    case = _case(name, g)
    print("case", name, flush=True)
    qlx_ctx.load_workload(case, g["data"])
    img = qlx_ctx.mock_image(case["lenses"], case["sources"], case["sigma"], case["seed"])
    assert np.allclose(img, g["data"], rtol=1e-12, atol=1e-14)


def test_gaussian_psf_matches_reference(qlx_ctx, golden_dir):
    """qlx_set_psf_gaussian against QLens::generate_PSF_matrix on the same pixel steps, bit for bit
    (isotropic, anisotropic, non-square pixels)."""
    g = np.load(os.path.join(golden_dir, "psf_gaussian.npz"))
    for k, (sx, sy, th, dx, dy) in enumerate(g["cases"]):
        xv, yv = W.corners(0.0, 40 * dx, 40), W.corners(0.0, 30 * dy, 30)
        qlx_ctx.load_image(xv, yv, np.zeros((40, 30)), sigma=1.0)
        qlx_ctx.set_psf_gaussian(sx, sy, th)
        assert np.array_equal(qlx_ctx.get_psf(), g["psf_%d" % k]), k


@pytest.mark.parametrize("name", ["C1", "C2s", "CNS", "C2", "CSPL", "CSPS"])
def test_inversion_matches_reference(qlx_ctx, golden_dir, name):
    g = np.load(os.path.join(golden_dir, "inversion_%s.npz" % name))
    case = _case(name, g)
    print("case", name, flush=True)
    qlx_ctx.load_workload(case, g["data"])
    if name == "CNS":  # the PSF built by the library itself from the widths and threshold
        qlx_ctx.set_psf_gaussian(case["psf_sigma"][0], case["psf_sigma"][1], case["psf_threshold"])
        assert np.array_equal(qlx_ctx.get_psf(), g["psf_used"])
    r = qlx_ctx.invert_image_surface_brightness_map(case["lenses"])[0]
    assert r["status"] == 0
    assert r["n_img_active"] == int(g["n_img"]) and r["n_src_active"] == int(g["n_src"])
    for k in ("minus2lnE", "chisq0", "logdetF", "logdetR"):
        assert rel(r[k], float(g[k])) < RTOL_EVIDENCE, (k, r[k], float(g[k]))
    sol = qlx_ctx.get_source(0)
    es, em = vrel(sol["src"], g["src"]), vrel(sol["model"], g["model"])
    print("%s: -2lnE rel %.2e, |ds|/|s| %.2e, |dmodel|/|model| %.2e" % (
        name, rel(r["minus2lnE"], float(g["minus2lnE"])), es, em))
    assert es < RTOL_EVIDENCE and em < RTOL_EVIDENCE
    assert np.array_equal(sol["active_index"], g["active"])
    # subpixel source-plane points of the in-mask pixels against the reference's
    rays = qlx_ctx.get_rays(0)
    nx, ny = rays.shape[:2]
    got = rays.reshape(nx * ny, -1, 2)[g["ray_pixels"]]
    d = np.abs(got - g["rays"])
    print("%s rays: %d compared, max |d beta| %.3g, %d above 1e-13" % (name, d.size // 2, d.max(), int((d > 1e-13).sum())))
    assert np.allclose(got, g["rays"], rtol=1e-12, atol=1e-13)


def test_batched_grid_sizes_match_reference(qlx_ctx, golden_dir):
    """The 200 x 200 batched geometry on config 1's truth: one call with one item per source-grid
    side (20..60, uneven grids in one submission) against the reference's evaluation of each, and
    each item's own source grid read back."""
    g = np.load(os.path.join(golden_dir, "inversion_C4s.npz"))
    case = W.c1_geometry_200(30)
    qlx_ctx.load_workload(case, g["data"])
    sides = [int(v) for v in g["sides"]]
    res = qlx_ctx.invert_image_surface_brightness_map([case["lenses"]] * len(sides), src_n=[(k, k) for k in sides])
    for b, (k, r) in enumerate(zip(sides, res)):
        assert r["status"] == 0, (k, r)
        assert r["n_src_active"] == int(g["n_src_%d" % k]) and r["n_img_active"] == int(g["n_img_%d" % k]), k
        for f in ("minus2lnE", "chisq0", "logdetF", "logdetR"):
            assert rel(r[f], float(g["%s_%d" % (f, k)])) < RTOL_EVIDENCE, (k, f, r[f])
        sol = qlx_ctx.get_source(b)
        assert sol["src"].shape == (k, k)
        es = vrel(sol["src"], g["src_%d" % k])
        em = vrel(sol["model"].reshape(-1)[g["model_pixels"]], g["model_%d" % k])
        print("side %d: -2lnE rel %.2e |ds|/|s| %.2e |dmodel|/|model| %.2e" % (
            k, rel(r["minus2lnE"], float(g["minus2lnE_%d" % k])), es, em))
        assert es < RTOL_EVIDENCE and em < RTOL_EVIDENCE, k


def test_inversion_is_deterministic(qlx_ctx, golden_dir):
    g = np.load(os.path.join(golden_dir, "inversion_C1.npz"))
    case = W.c1()
    qlx_ctx.load_workload(case, g["data"])
    a = qlx_ctx.invert_image_surface_brightness_map(case["lenses"])[0]
    b = qlx_ctx.invert_image_surface_brightness_map(case["lenses"])[0]
    assert a["minus2lnE"] == b["minus2lnE"] and a["logdetF"] == b["logdetF"]


def test_batch_of_perturbed_models_matches_oracle(qlx_ctx, golden_dir):
    g = np.load(os.path.join(golden_dir, "inversion_C1.npz"))
    case = W.c1()
    qlx_ctx.load_workload(case, g["data"])
    rng = np.random.default_rng(11)
    models = []
    for _ in range(3):
        (t0, p0), (t1, p1) = case["lenses"]
        p0 = tuple(v * (1 + 0.01 * rng.standard_normal()) if i in (0, 1, 2) else v for i, v in enumerate(p0))
        p1 = tuple(v * (1 + 0.01 * rng.standard_normal()) if i in (0, 1) else v for i, v in enumerate(p1))
        models.append([(t0, p0), (t1, p1)])
    res = qlx_ctx.invert_image_surface_brightness_map(models)
    for m, r in zip(models, res):
        ro, _ = O.invert(case, g["data"], lenses=m, want_arrays=False)
        assert r["n_src_active"] == ro["n_src"] and r["n_img_active"] == ro["n_img"]
        assert rel(r["minus2lnE"], ro["minus2lnE"]) < RTOL_EVIDENCE


def test_failed_items_carry_sentinels(qlx_ctx, golden_dir):
    """Items with invalid lens parameters (NaN, q out of range) or an oversize grid fail alone:
    their status is an error and minus2lnE the 2e30 sentinel, the other items are unaffected, and
    the call returns QLX_ITEMS_FAILED rather than raising."""
    g = np.load(os.path.join(golden_dir, "inversion_C1.npz"))
    case = W.c1()
    qlx_ctx.load_workload(case, g["data"])
    (t0, p0), s = case["lenses"]
    bad_nan = [(t0, (float("nan"),) + p0[1:]), s]
    bad_q = [(t0, (p0[0], 1.5) + p0[2:]), s]
    res = qlx_ctx.invert_image_surface_brightness_map([bad_nan, case["lenses"], bad_q], src_n=[(30, 30), (30, 30), (1, 5)])
    assert res[0]["status"] == -1 and res[0]["minus2lnE"] == 2e30
    assert res[2]["status"] == -6 and res[2]["minus2lnE"] == 2e30
    assert res[1]["status"] == 0 and rel(res[1]["minus2lnE"], float(g["minus2lnE"])) < RTOL_EVIDENCE
    res = qlx_ctx.invert_image_surface_brightness_map([bad_q])
    assert res[0]["status"] == -1 and res[0]["minus2lnE"] == 2e30
    with pytest.raises(qlx.QlxError):
        qlx_ctx.get_source(0)  # a failed item keeps no solution


@pytest.mark.parametrize("precision", [0, 1])
def test_singular_F_is_reported(qlx_ctx, golden_dir, precision):
    """lambda = 0, a mask of six pixels and a 60 x 60 source grid: F = Lb^T C Lb has rank <= 6 with
    37 active cells, so 31 pivots of its Cholesky are rounding noise of either sign (the chance
    that all are positive is ~2^-31).  The reference dies there (pixelgrid.cpp:17384); here the
    item reports QLX_ERR_NOT_PD with the 2e30 sentinel.  The oracle flags the same case."""
    from oracle import oracle as O
    g = np.load(os.path.join(golden_dir, "inversion_C1.npz"))
    case = W.c1()
    mask = np.zeros((100, 100), dtype=np.uint8)
    mask[62:65, 48:50] = 1
    case["mask"] = mask
    case["regparam"] = 0.0
    case["src_n"] = (60, 60)
    ro, _ = O.invert(case, g["data"], want_arrays=False)
    assert ro["status"] != 0 and ro["n_src"] > 30
    qlx_ctx.load_workload(case, g["data"], precision=precision)
    r = qlx_ctx.invert_image_surface_brightness_map(case["lenses"])[0]
    assert r["status"] == -5 and r["minus2lnE"] == 2e30, r
    assert r["n_src_active"] == ro["n_src"] and r["n_img_active"] == ro["n_img"]


def test_multi_wave_equals_single_wave(golden_dir):
    """A small HBM budget splits a call into several waves (admission); every item's result and
    its kept solution equal the single-wave call's bit for bit."""
    from qlens_amd import qlx as Q
    g = np.load(os.path.join(golden_dir, "inversion_C2s.npz"))
    case = W.c2_small()
    models = _perturbed(case, 7, 21)
    grids = [(24, 24), (20, 28), (24, 24), (16, 16), (30, 30), (24, 24), (22, 18)]
    out = {}
    for budget in (0, 48 << 20):
        inv = Q.PixelInversion(0, hbm_budget=budget)
        try:
            inv.load_workload(case, g["data"])
            res = inv.invert_image_surface_brightness_map(models, src_n=grids)
            waves = Q.get_stats(inv)["n_waves"]
            sols = [inv.get_source(b)["src"] for b in range(len(models))]
        finally:
            inv.close()
        out[budget] = (res, waves, sols)
    (r1, w1, s1), (r2, w2, s2) = out[0], out[48 << 20]
    print("waves: default budget %d, small budget %d" % (w1, w2))
    assert w1 == 1 and w2 >= 2
    for a, b in zip(r1, r2):
        assert a == b
    for a, b in zip(s1, s2):
        assert np.array_equal(a, b)


def test_sentinels(qlx_ctx, golden_dir):
    g = np.load(os.path.join(golden_dir, "inversion_C1.npz"))
    case = W.c1()
    case["srcgrid"] = (50.0, 51.0, 50.0, 51.0)
    qlx_ctx.load_workload(case, g["data"])
    r = qlx_ctx.invert_image_surface_brightness_map(case["lenses"])[0]
    assert r["minus2lnE"] == 2e30 and r["n_src_active"] == 0
    case = W.c1()
    qlx_ctx.load_workload(case, g["data"])
    r = qlx_ctx.invert_image_surface_brightness_map(case["lenses"], src_n=[(1, 30)])[0]
    assert r["minus2lnE"] == 2e30


def _perturbed(case, n, seed):
    rng = np.random.default_rng(seed)
    models = []
    for _ in range(n):
        m = []
        for t, p in case["lenses"]:
            p = list(p)
            # SIE (b, q), NFW (ks, rs, q), shear gamma, SPLE (b, alpha, q), shear components
            for i in {1: (0, 1), 2: (0, 1, 2), 3: (0,), 4: (0, 1, 2), 5: (0, 1)}[t]:
                p[i] *= 1 + 0.01 * rng.standard_normal()
            if t in (1, 2, 4):
                k = {1: 1, 2: 2, 4: 2}[t]
                p[k] = min(p[k], 0.999)
            m.append((t, tuple(p)))
        models.append(m)
    return models


def test_batch_equals_single_bitwise(qlx_ctx, golden_dir):
    """Items of one batch share launches but not arithmetic: each equals its own solo evaluation."""
    g = np.load(os.path.join(golden_dir, "inversion_C2s.npz"))
    case = W.c2_small()
    qlx_ctx.load_workload(case, g["data"])
    models = _perturbed(case, 6, 5)
    batch = qlx_ctx.invert_image_surface_brightness_map(models)
    for m, rb in zip(models, batch):
        rs = qlx_ctx.invert_image_surface_brightness_map(m)[0]
        assert rb["minus2lnE"] == rs["minus2lnE"] and rb["logdetR"] == rs["logdetR"]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_shared_regularization_factorizations(qlx_ctx, golden_dir, prec):
    """Items whose regularization matrix R is the same (same grid and active cells) share one
    factorization of it; items with other active cells get their own.  A batch mixing both kinds
    (repeated models, models with another Einstein radius and so other mapped cells) must give
    every item its solo result bit for bit."""
    from qlens_amd import qlx
    g = np.load(os.path.join(golden_dir, "inversion_C1.npz"))
    case = dict(W.c1())
    case["srcgrid"] = (-1.6, 1.6, -1.6, 1.6)  # wider than the mapped region: the active cells vary
    qlx_ctx.load_workload(case, g["data"], precision=qlx.QLX_PREC_FP64 if prec == "fp64" else qlx.QLX_PREC_MIXED)
    base = case["lenses"]

    def scaled(f):
        (t, p), rest = base[0], base[1:]
        return [(t, (p[0] * f,) + tuple(p[1:]))] + list(rest)

    models = [base, scaled(1.15), base, scaled(0.85), scaled(1.15), base, scaled(1.3)]
    batch = qlx_ctx.invert_image_surface_brightness_map(models)
    nsrc = [r["n_src_active"] for r in batch]
    assert len(set(nsrc)) >= 3, nsrc  # several different active sets in the batch
    for m, rb in zip(models, batch):
        rs = qlx_ctx.invert_image_surface_brightness_map(m)[0]
        assert rb["logdetR"] == rs["logdetR"] and rb["minus2lnE"] == rs["minus2lnE"], (rb["n_src_active"], rs["n_src_active"])
        assert rb["status"] == rs["status"] == 0


def test_uneven_source_grids_in_one_batch(qlx_ctx, golden_dir):
    """One submission with per-item source grids (and a too-small grid) against the oracle."""
    g = np.load(os.path.join(golden_dir, "inversion_C1.npz"))
    case = W.c1()
    qlx_ctx.load_workload(case, g["data"])
    sizes = [(20, 20), (1, 7), (30, 24), (12, 40), (24, 24)]
    res = qlx_ctx.invert_image_surface_brightness_map([case["lenses"]] * len(sizes), src_n=sizes)
    for (nx, ny), r in zip(sizes, res):
        if nx < 2 or ny < 2:
            assert r["minus2lnE"] == 2e30
            continue
        c = dict(case)
        c["src_n"] = (nx, ny)
        ro, _ = O.invert(c, g["data"], want_arrays=False)
        assert r["n_src_active"] == ro["n_src"] and r["n_img_active"] == ro["n_img"]
        assert rel(r["minus2lnE"], ro["minus2lnE"]) < RTOL_EVIDENCE, (nx, ny)


# Mixed precision (BF16-split F, FP32 Cholesky, FP64 refinement and log-dets): the accuracy
# contract is |delta ln E| < 0.1 against the FP64 evidence; the refined solve must converge.
@pytest.mark.parametrize("name", ["C1", "C2s", "CNS", "C2", "CSPL"])
def test_mixed_precision_evidence(qlx_ctx, golden_dir, name):
    from qlens_amd import qlx
    g = np.load(os.path.join(golden_dir, "inversion_%s.npz" % name))
    case = _case(name, g)
    qlx_ctx.load_workload(case, g["data"], precision=qlx.QLX_PREC_MIXED)
    models = [case["lenses"]] + _perturbed(case, 3, 9)
    res = qlx_ctx.invert_image_surface_brightness_map(models)
    qlx_ctx.load_workload(case, g["data"])
    ref = qlx_ctx.invert_image_surface_brightness_map(models)
    assert abs(res[0]["minus2lnE"] - float(g["minus2lnE"])) / 2 < 0.1
    for r, f in zip(res, ref):
        assert r["status"] == 0 and r["n_src_active"] == f["n_src_active"]
        assert abs(r["minus2lnE"] - f["minus2lnE"]) / 2 < 0.1, (r["minus2lnE"], f["minus2lnE"])
        assert r["logdetR"] == f["logdetR"]  # R stays FP64 in both modes
        assert r["refine_resid"] <= 1e-10, r["refine_resid"]  # the configured refine_tol


AUTOGRID = [("C1_both", "C1", True, True), ("CNS_box", "CNS", True, False), ("CNS_npix", "CNS", False, True),
            ("CNS_both", "CNS", True, True)]


@pytest.mark.parametrize("tag,wl,ab,an", AUTOGRID)
def test_auto_source_grid_matches_reference(qlx_ctx, golden_dir, tag, wl, ab, an):
    """auto_srcgrid / auto_src_npixels: each item's box and size found on the device from its own
    rays, against the reference's (box within 1e-12 relative: the rays agree to ~1e-13; size
    exact), and the evaluation on that grid against the reference's."""
    g = np.load(os.path.join(golden_dir, "autogrid.npz"))
    gi = np.load(os.path.join(golden_dir, "inversion_%s.npz" % wl))
    case = dict(_case(wl, gi), auto_srcgrid=ab, auto_src_npixels=an)
    print("case", tag, flush=True)
    qlx_ctx.load_workload(case, gi["data"])
    r = qlx_ctx.invert_image_surface_brightness_map(case["lenses"])[0]
    box = np.array([r["src_xmin"], r["src_xmax"], r["src_ymin"], r["src_ymax"]])
    print("%s: grid %dx%d box %s, -2lnE rel %.2e" % (tag, r["src_nx"], r["src_ny"], box,
                                                      rel(r["minus2lnE"], float(g["minus2lnE_" + tag]))))
    assert [r["src_nx"], r["src_ny"]] == list(g["size_" + tag])
    assert np.allclose(box, g["box_" + tag], rtol=1e-12, atol=0)
    assert r["status"] == 0
    assert r["n_img_active"] == int(g["n_img_" + tag]) and r["n_src_active"] == int(g["n_src_" + tag])
    for k in ("minus2lnE", "chisq0", "logdetF", "logdetR"):
        assert rel(r[k], float(g["%s_%s" % (k, tag)])) < RTOL_EVIDENCE, (k, r[k])
    sol = qlx_ctx.get_source(0)
    assert sol["src"].shape == tuple(g["size_" + tag])
    assert vrel(sol["src"], g["src_" + tag]) < RTOL_EVIDENCE
    assert np.array_equal(sol["active_index"], g["active_" + tag])


def test_auto_source_grid_per_item_in_a_batch(qlx_ctx, golden_dir):
    """A batch of perturbed models with the automatic grid: every item has its own box and size,
    each equal, with the item's whole result, to the item evaluated alone."""
    gi = np.load(os.path.join(golden_dir, "inversion_CNS.npz"))
    case = dict(_case("CNS", gi), auto_srcgrid=True, auto_src_npixels=True)
    qlx_ctx.load_workload(case, gi["data"])
    models = [case["lenses"]] + _perturbed(case, 5, 21)
    res = qlx_ctx.invert_image_surface_brightness_map(models)
    sizes = set()
    for k, m in enumerate(models):
        one = qlx_ctx.invert_image_surface_brightness_map(m)[0]
        assert res[k]["status"] == 0
        for f in ("minus2lnE", "src_nx", "src_ny", "src_xmin", "src_xmax", "src_ymin", "src_ymax", "n_src_active"):
            assert res[k][f] == one[f], (k, f)
        sizes.add((res[k]["src_nx"], res[k]["src_ny"], res[k]["src_xmin"]))
    assert len(sizes) > 1  # the perturbed models do get their own grids


def test_cholesky_panel_autotune_is_bitwise_neutral(qlx_ctx, golden_dir):
    """The mixed path's FP32 Cholesky panel width (512, 1024, 2048 columns, or chosen by the
    autotuner on the first wave of a size class) changes the schedule only: every result is the
    same bit for bit, and the autotuner reports one of the three widths."""
    g = np.load(os.path.join(golden_dir, "inversion_C2.npz"))
    case = W.c2()
    qlx_ctx.load_workload(case, g["data"], precision=qlx.QLX_PREC_MIXED)
    models = [case["lenses"]] + _perturbed(case, 3, 17)
    out = {}
    for panel in (512, 1024, 2048, 0):
        qlx_ctx.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"],
                             precision=qlx.QLX_PREC_MIXED, chol_panel=panel)
        res = qlx_ctx.invert_image_surface_brightness_map(models)
        st = qlx.get_stats(qlx_ctx)
        out[panel] = [(r["minus2lnE"], r["logdetF"], r["refine_resid"]) for r in res]
        print("chol_panel %d: used %d columns, -2lnE[0] %.12g" % (panel, st["chol_panel"], res[0]["minus2lnE"]))
        assert st["chol_panel"] == (panel or st["chol_panel"]) and st["chol_panel"] in (512, 1024, 2048)
        assert all(r["status"] == 0 for r in res)
    for panel in (1024, 2048, 0):
        assert out[panel] == out[512], panel


def test_factorization_graphs_replay_bitwise(qlx_ctx, golden_dir):
    """The factorizations run as captured HIP graphs, replayed on later calls with eager work and
    synchronizations in between: every replay gives the results of launching the kernels one by
    one, bit for bit, in FP64 (one proposal per call, and a batch) and in mixed precision."""
    g = np.load(os.path.join(golden_dir, "inversion_C2s.npz"))
    case = W.c2_small()
    models = [case["lenses"]] + _perturbed(case, 3, 29)
    for prec in (qlx.QLX_PREC_FP64, qlx.QLX_PREC_MIXED):
        ref = {}
        for graphs in (False, True):
            qlx_ctx.load_workload(case, g["data"], precision=prec)
            qlx_ctx.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"],
                                 precision=prec, graphs=graphs)
            before = qlx.get_stats(qlx_ctx)
            runs = []
            for rep in range(4):  # the first call captures, the next ones replay
                one = qlx_ctx.invert_image_surface_brightness_map(models[rep % len(models)])[0]
                batch = qlx_ctx.invert_image_surface_brightness_map(models)
                qlx_ctx.get_source(0)  # eager work and a synchronization between replays
                runs.append((one, batch))
            after = qlx.get_stats(qlx_ctx)
            if graphs:
                assert after["graph_replays"] - before["graph_replays"] >= 8, (before, after)
            ref[graphs] = runs
        for (o0, b0), (o1, b1) in zip(ref[False], ref[True]):
            assert o0 == o1 and b0 == b1, prec
    print("graphs: %d captured, %d replayed" % (after["graph_captures"], after["graph_replays"]))
