"""Regularization-parameter optimisation (X1), the likelihood wrapper (A1) and the sampler batching
adapter on the GPU, against the compiled reference and the plain evaluation.

X1: 'optimize_regparam on' in the reference (Brent over log10 lambda in [-3, 5], tol 0.01, at most 20
evaluations; pixelgrid.cpp:15865-15979, 16526-16699, 16849-16932) on C1 and on the non-square masked
CNS workload.  The optimum lambda depends on every objective value through Brent's parabolic steps,
so it is compared at 1e-8 relative; -2 ln E at the optimum (a minimum: first-order insensitive to
lambda) and the log-determinants at 1e-10 / 1e-9."""
import os

import numpy as np
import pytest

from qlens_amd import qlx
from qlens_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _cns(golden_dir):
    g = np.load(os.path.join(golden_dir, "inversion_CNS.npz"))
    case = W.c_ns()
    case["psf"] = g["psf_used"]
    return case


@pytest.mark.parametrize("name", ["C1", "CNS"])
def test_regparam_optimisation_matches_reference(qlx_ctx, golden_dir, name):
    g = np.load(os.path.join(golden_dir, "regopt.npz"))
    case = W.c1() if name == "C1" else _cns(golden_dir)
    print("case", name, flush=True)
    qlx_ctx.load_workload(case, g["data_" + name], optimize_regparam=True)
    (t0, p0) = case["lenses"][0]
    other = [(t0, (p0[0] * 1.01,) + tuple(p0[1:]))] + list(case["lenses"][1:])
    res = qlx_ctx.invert_image_surface_brightness_map([case["lenses"], other, case["lenses"]])
    r = res[0]
    print("%s: regparam %.12g (reference %.12g), -2lnE rel %.2e, %d objective evaluations" % (
        name, r["regparam"], float(g["regparam_" + name]), rel(r["minus2lnE"], float(g["minus2lnE_" + name])),
        r["regopt_evals"]))
    assert r["status"] == 0 and res[1]["status"] == 0
    assert rel(r["regparam"], float(g["regparam_" + name])) < 1e-8
    assert rel(r["minus2lnE"], float(g["minus2lnE_" + name])) < 1e-10
    assert rel(r["chisq0"], float(g["chisq0_" + name])) < 1e-9
    assert rel(r["logdetF"], float(g["logdetF_" + name])) < 1e-9
    assert rel(r["logdetR"], float(g["logdetR_" + name])) < 1e-10
    src = qlx_ctx.get_source(0)["src"]
    ref = g["src_" + name]
    assert np.linalg.norm(src - ref) / np.linalg.norm(ref) < 1e-8
    # an item's optimisation does not depend on its batch neighbours
    assert res[2] == r
    one = qlx_ctx.invert_image_surface_brightness_map(other)[0]
    assert one == res[1]
    assert 3 <= r["regopt_evals"] <= 21


def test_regparam_optimisation_needs_fp64(qlx_ctx):
    case = W.c1()
    qlx_ctx.load_image(case["xvals"], case["yvals"], np.zeros((100, 100)), sigma=case["sigma"])
    with pytest.raises(qlx.QlxError):
        qlx_ctx.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"],
                             precision=qlx.QLX_PREC_MIXED, optimize_regparam=True)


def _fit_setup(inv, golden_dir):
    g = np.load(os.path.join(golden_dir, "inversion_C1.npz"))
    case = W.c1()
    inv.load_workload(case, g["data"])
    # vary SIE b and q (limits), shear gamma as log10 with its Jacobian, a Gaussian prior on b
    params = [dict(lens=0, param=0, lo=0.5, hi=2.0, prior=("gauss", 1.2, 0.1, 2.5)),
              dict(lens=0, param=1, lo=0.3, hi=1.0),
              dict(lens=1, param=0, lo=1e-4, hi=0.3, transform="log10", jacobian=True)]
    inv.set_fit_model(case["lenses"], params)
    return case


def test_loglike_wrapper_matches_the_evidence(qlx_ctx, golden_dir):
    """-ln L = (-2 ln E)/2 + prior + Jacobian terms of the updated model (lens.cpp:10657-10748);
    1e30 outside the penalty limits or for an invalid model."""
    case = _fit_setup(qlx_ctx, golden_dir)
    (t0, p0), (t1, p1) = case["lenses"]
    x = np.array([[1.2, 0.8, np.log10(0.05)],      # the truth
                  [1.25, 0.78, np.log10(0.045)],
                  [2.5, 0.8, np.log10(0.05)],       # b above its limit
                  [1.2, 0.8, np.log10(0.5)],        # gamma above its limit
                  [1.19, 0.81, np.log10(0.051)]])
    ll, res = qlx_ctx.loglike(x)
    assert ll[2] == 1e30 and ll[3] == 1e30 and res[2]["status"] == -1
    for k in (0, 1, 4):
        b, q, lg = x[k]
        model = [(t0, (b, q) + tuple(p0[2:])), (t1, (10 ** lg,) + tuple(p1[1:]))]
        r = qlx_ctx.invert_image_surface_brightness_map(model)[0]
        expect = r["minus2lnE"] / 2 + (np.log(2.5) + ((b - 1.2) / 0.1) ** 2 / 2) - np.log(10 ** lg)
        assert rel(ll[k], expect) < 1e-12, (k, ll[k], expect)
        assert res[k]["minus2lnE"] == r["minus2lnE"]


def test_batched_sampler_equals_one_proposal_loop(qlx_ctx, golden_dir):
    """The nested sampler's rejection phase with 32 proposals per GPU call accepts exactly the points
    (and -ln L) of the one-proposal-per-call loop on the same draws."""
    from qlens_amd.sampler import BatchedMonoSampler
    _fit_setup(qlx_ctx, golden_dir)
    lo, hi = [1.1, 0.7, np.log10(0.03)], [1.3, 0.9, np.log10(0.07)]
    runs = {}
    for batch in (1, 32):
        smp = BatchedMonoSampler(qlx_ctx, lo, hi, batch=batch, rng=np.random.default_rng(99))
        live_u, live_ll = smp.initial_live_points(10)
        hist = []
        for _ in range(6):
            hist.append(smp.replace_worst(live_u, live_ll))
        runs[batch] = (live_u.copy(), live_ll.copy(), hist, smp.calls)
    (u1, l1, h1, c1), (u32, l32, h32, c32) = runs[1], runs[32]
    print("one-proposal loop: %d calls; batched: %d calls" % (c1, c32))
    assert np.array_equal(u1, u32) and np.array_equal(l1, l32) and h1 == h32
    assert c32 < c1


@pytest.mark.parametrize("tag", ["true", "start"])
def test_shipped_demo_end_to_end(qlx_ctx, golden_dir, tag):
    """The reference's shipped pixel-fit demo (workloads.demo) read from its FITS files by the
    library (image at data_pixel_size 0.049", mask, PSF cut at 1e-2), split_imgpixels off, the
    automatic source grid, optimize_regparam over [-5, 5] and the SPLE (alpha 1.17) + shear
    components lens, against the reference (fed the same files through its text-image route).
    At the true lens chi^2 is close to the number of fitted pixels (the reference's own mock noise
    level): the FITS axes are read the right way round."""
    g = np.load(os.path.join(golden_dir, "demo.npz"))
    case = W.demo()
    fits = os.path.join(golden_dir, "fits")
    print("demo", tag, flush=True)
    qlx_ctx.load_image_fits(os.path.join(fits, case["files"][0]), pixel_size=case["pixel_size"], sigma=case["sigma"])
    qlx_ctx.load_mask_fits(os.path.join(fits, case["files"][1]))
    qlx_ctx.load_psf_fits(os.path.join(fits, case["files"][2]), case["psf_threshold"])
    qlx_ctx.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"],
                         optimize_regparam=True, regopt_bounds=case["regopt_bounds"], auto_srcgrid=True,
                         auto_src_npixels=True)
    lenses = W.DEMO_TRUE if tag == "true" else W.DEMO_START
    r = qlx_ctx.invert_image_surface_brightness_map(lenses)[0]
    print("demo %s: grid %dx%d, regparam %.10g (reference %.10g), -2lnE %.10g (reference %.10g), chisq %.6g over %d pixels"
          % (tag, r["src_nx"], r["src_ny"], r["regparam"], float(g["regparam_" + tag]), r["minus2lnE"],
             float(g["minus2lnE_" + tag]), r["chisq0"], r["n_img_active"]))
    assert r["status"] == 0
    assert [r["src_nx"], r["src_ny"]] == list(g["size_" + tag])
    box = [r["src_xmin"], r["src_xmax"], r["src_ymin"], r["src_ymax"]]
    assert np.allclose(box, g["box_" + tag], rtol=1e-12, atol=0)
    assert r["n_img_active"] == int(g["n_img_" + tag]) and r["n_src_active"] == int(g["n_src_" + tag])
    assert rel(r["regparam"], float(g["regparam_" + tag])) < 1e-8
    assert rel(r["minus2lnE"], float(g["minus2lnE_" + tag])) < 1e-10
    assert rel(r["logdetF"], float(g["logdetF_" + tag])) < 1e-9
    assert rel(r["logdetR"], float(g["logdetR_" + tag])) < 1e-10
    sol = qlx_ctx.get_source(0)
    assert np.linalg.norm(sol["src"] - g["src_" + tag]) / np.linalg.norm(g["src_" + tag]) < 1e-8
    assert np.array_equal(sol["active_index"], g["active_" + tag])
    if tag == "true":
        assert 0.8 < r["chisq0"] / r["n_img_active"] < 1.5
