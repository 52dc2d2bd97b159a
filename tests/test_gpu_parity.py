"""GPU parity: the HIP path (through the C ABI) against the oracle and the reference's golden
vectors.  Bar: integer/index outputs identical; FP64 results within 1e-10 relative."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
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


def test_nfw_def_hess_matches_oracle(qlx_ctx):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-6, 6, size=(300, 2))
    lens = (0.2, 5.0, 0.75, 0.0, 0.0, 0.0)
    got = qlx_ctx.nfw_def_hess(lens, pts, tol=1e-6)
    ref = O.nfw_def_hess(lens, pts, tol=1e-6)
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-13)


def test_mock_image_matches_oracle(qlx_ctx, golden_dir):
    case = W.c1()
    g = np.load(os.path.join(golden_dir, "inversion_C1.npz"))
    qlx_ctx.load_workload(case, g["data"])
    img = qlx_ctx.mock_image(case["lenses"], case["sources"], case["sigma"], case["seed"])
    assert np.allclose(img, g["data"], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name,builder", [("C1", W.c1), ("C2s", W.c2_small)])
def test_inversion_matches_reference(qlx_ctx, golden_dir, name, builder):
    g = np.load(os.path.join(golden_dir, "inversion_%s.npz" % name))
    case = builder()
    qlx_ctx.load_workload(case, g["data"])
    r = qlx_ctx.invert_image_surface_brightness_map(case["lenses"])[0]
    assert r["status"] == 0
    assert r["n_img_active"] == int(g["n_img"]) and r["n_src_active"] == int(g["n_src"])
    for k, gk in [("minus2lnE", "minus2lnE"), ("chisq0", "chisq0"), ("logdetF", "logdetF"), ("logdetR", "logdetR")]:
        assert rel(r[k], float(g[gk])) < RTOL_EVIDENCE, (k, r[k], float(g[gk]))
    src, model = qlx_ctx.get_solution()
    assert np.allclose(src, g["src"], rtol=1e-8, atol=1e-9 * np.abs(g["src"]).max())
    assert np.allclose(model, g["model"], rtol=1e-8, atol=1e-9 * np.abs(g["model"]).max())


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
            for i in {1: (0, 1), 2: (0, 1, 2), 3: (0,)}[t]:
                p[i] *= 1 + 0.01 * rng.standard_normal()
            if t in (1, 2):
                k = 1 if t == 1 else 2
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
@pytest.mark.parametrize("name,builder", [("C1", W.c1), ("C2s", W.c2_small)])
def test_mixed_precision_evidence(qlx_ctx, golden_dir, name, builder):
    from qlens_amd import qlx
    g = np.load(os.path.join(golden_dir, "inversion_%s.npz" % name))
    case = builder()
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
        assert r["refine_resid"] < 1e-6, r["refine_resid"]
