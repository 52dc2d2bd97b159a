"""The C ABI library builds, loads and exports every entry point include/qlx.h declares (CPU)."""
import ctypes
import os
import re

import pytest

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LIB = os.path.join(REPO, "qlens_amd", "libqlx.so")


def declared_symbols():
    src = open(os.path.join(REPO, "include", "qlx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qlx_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "qlx_eval" in syms and "qlx_ctx_create" in syms and len(syms) >= 20


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqlx.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqlx.so not built")
def test_structs_match_header_layout():
    # sizes of the ctypes mirrors equal the C structs (checked by field arithmetic here)
    from qlens_amd import qlx
    assert ctypes.sizeof(qlx.QlxLens) == 8 + 6 * 8
    assert ctypes.sizeof(qlx.QlxLensModel) == 8 + qlx.QLX_MAX_LENSES * ctypes.sizeof(qlx.QlxLens)
    assert ctypes.sizeof(qlx.QlxResult) == 11 * 8 + 8 * 4
    # and the library's own sizeof of every struct (a host-only call, no device needed)
    lib = ctypes.CDLL(LIB)
    sizes = (ctypes.c_size_t * 8)()
    assert lib.qlx_struct_sizes(sizes, 8) == 8
    mirrors = [qlx.QlxLens, qlx.QlxLensModel, qlx.QlxOpts, qlx.QlxSettings, qlx.QlxResult, qlx.QlxFitParam,
               qlx.QlxSurrogateCfg, qlx.QlxStats]
    assert list(sizes) == [ctypes.sizeof(m) for m in mirrors]


def test_no_cpu_fallback_in_product():
    # the product package never imports, loads or runs anything under oracle/
    for root, _, files in os.walk(os.path.join(REPO, "qlens_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(root, f)).read()
                assert "from oracle" not in txt and "import oracle" not in txt, f
                assert "liboracle" not in txt and "ref_harness" not in txt, f


def test_batch_packing_matches_field_by_field():
    # make_models / the numpy result view give the same bytes and values as the per-field ctypes
    # forms, for uniform and ragged lens lists (SHEAR has four parameters, the others six)
    import numpy as np
    from qlens_amd import qlx, workloads as W
    props = W.proposals(W.c2(), 16)
    models = props + [[props[0][0]], props[1][:2], []]
    ref = (qlx.QlxLensModel * len(models))()
    for k, lm in enumerate(models):
        ref[k] = qlx.make_model(lm, qlx.QLX_DEF_SURROGATE)
    assert bytes(qlx.make_models(models, qlx.QLX_DEF_SURROGATE)) == bytes(ref)
    # models of one layout (the per-lens-position fill) and one that differs from the first only
    # in a lens type or a parameter count (the general fill)
    for dm in (qlx.QLX_DEF_QUADRATURE, qlx.QLX_DEF_SURROGATE):
        for ms in (props, props[:5] + [[props[5][1], props[5][0], props[5][2]]],
                   props[:5] + [[props[5][0], props[5][1], (props[5][2][0], props[5][2][1][:2])]]):
            ref = (qlx.QlxLensModel * len(ms))()
            for k, lm in enumerate(ms):
                ref[k] = qlx.make_model(lm, dm)
            assert bytes(qlx.make_models(ms, dm)) == bytes(ref)
    res = (qlx.QlxResult * 3)()
    res[1].minus2lnE, res[1].status, res[2].refine_steps = -3.25, -7, 4
    rows = np.frombuffer(res, dtype=qlx._RESULT_DT).tolist()
    got = [dict(zip(qlx._RESULT_DT.names, r)) for r in rows]
    assert got == [{f: getattr(res[k], f) for f, _ in qlx.QlxResult._fields_} for k in range(3)]
