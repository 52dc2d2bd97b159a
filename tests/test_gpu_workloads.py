"""Batched evidence over uneven source grids: 256 proposals around the 200x200 / 60x60 workload's
truth, each with its own square source grid of 20..60 cells per side, in one call.

Properties checked at full size (the oracle would need hours for 256 evaluations; per-item parity
with the reference is pinned on the C1 and C2s fixtures in test_gpu_parity.py):
  * every item succeeds,
  * an item's batched result equals its stand-alone evaluation bit for bit,
  * the mixed-precision mode stays within its stated |delta lnE| < 0.1 of FP64 on every item.
"""
import numpy as np
import pytest

from qlens_amd import qlx, workloads as W

pytestmark = pytest.mark.gpu


def _load(inv, case, precision):
    nx = len(case["xvals"]) - 1
    inv.load_image(case["xvals"], case["yvals"], np.zeros((nx, nx)), sigma=case["sigma"])
    inv.set_psf(case["psf"])
    inv.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"])
    data = inv.mock_image(case["lenses"], case["sources"], case["sigma"], case["seed"])
    inv.load_workload(case, data, precision=precision)
    return data


def test_batched_uneven_source_grids_256(qlx_ctx):
    case, models, grids = W.c4()
    data = _load(qlx_ctx, case, qlx.QLX_PREC_FP64)
    res = qlx_ctx.invert_image_surface_brightness_map(models, src_n=grids)
    assert len(res) == 256
    for k, r in enumerate(res):
        assert r["status"] == 0, (k, grids[k], r["status"])
        assert np.isfinite(r["minus2lnE"]) and r["n_src_active"] > 0, (k, r)
    print("grid sides", sorted(set(g[0] for g in grids)), "-2lnE range %.6g .. %.6g" % (
        min(r["minus2lnE"] for r in res), max(r["minus2lnE"] for r in res)))
    # one item of each grid size, evaluated alone
    picked = {}
    for k, g in enumerate(grids):
        picked.setdefault(g[0], k)
    for side, k in sorted(picked.items()):
        one = qlx_ctx.invert_image_surface_brightness_map(models[k], src_n=[grids[k]])[0]
        assert one["minus2lnE"] == res[k]["minus2lnE"], (side, k)
        assert one["logdetF"] == res[k]["logdetF"] and one["chisq0"] == res[k]["chisq0"], (side, k)
    # the same batch in mixed precision
    qlx_ctx.load_workload(case, data, precision=qlx.QLX_PREC_MIXED)
    resm = qlx_ctx.invert_image_surface_brightness_map(models, src_n=grids)
    d = [abs(a["minus2lnE"] - b["minus2lnE"]) / 2 for a, b in zip(resm, res)]
    print("mixed vs FP64 over 256 uneven items: max |dlnE| %.3g, median %.3g" % (max(d), float(np.median(d))))
    assert max(d) < 0.1
