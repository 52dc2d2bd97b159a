"""The mixed path's NFW convergence table (qlens_amd/csrc/nfw_gtab.h) against the closed form.

kappa = 2 ks g(x^2) with g(t) = (1 - F(t)) / (t - 1), F(t) = atan(sqrt(t-1))/sqrt(t-1) (t > 1) or
atanh(sqrt(1-t))/sqrt(1-t) (t < 1) (models.cpp:978-999).  This test evaluates the header's
polynomials the way the device does (piece from the exponent and the top five mantissa bits, u from
the rest, Horner) and compares with g computed here in 80-bit long double (closed form, and the
series sum_k (-1)^k (t-1)^k / (2k+3) near t = 1), plus g(1) = 1/3 and FP64 values at large t."""
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(__file__), "..", "qlens_amd", "csrc", "nfw_gtab.h")
LD = np.longdouble


def _load():
    txt = open(HDR).read()
    d = {k: int(v) for k, v in re.findall(r"#define QLX_GTAB_(\w+) (-?\d+)", txt)}
    body = txt[txt.index("{") + 1:txt.rindex("}")]
    coef = np.array([float(x) for x in body.replace("\n", " ").split(",") if x.strip()])
    assert len(coef) == (d["EMAX"] - d["EMIN"]) * d["NP"] * d["NC"]
    return d, coef


def _g_ref(t):
    t = LD(t)
    d = t - 1
    if abs(d) < LD(0.01):
        return sum(((-d) ** k) / LD(2 * k + 3) for k in range(40))
    if t > 1:
        s = np.sqrt(d)
        return (1 - np.arctan(s) / s) / d
    s = np.sqrt(-d)
    return (1 - np.arctanh(s) / s) / d


def _g_dev(d, coef, t):
    b = int(np.float64(t).view(np.int64))
    e = ((b >> 52) & 0x7FF) - 1023
    p = (b >> 47) & (d["NP"] - 1)
    u = float(np.int64(((b & ((1 << 47) - 1)) << 5) | (1023 << 52)).view(np.float64)) - 1.0
    c = coef[((e - d["EMIN"]) * d["NP"] + p) * d["NC"]:][:d["NC"]]
    g = LD(c[-1])
    for k in range(d["NC"] - 2, -1, -1):
        g = LD(float(g * LD(u) + LD(c[k])))  # one rounding per step (the device's fma)
    return float(g)


def test_table_matches_closed_form():
    d, coef = _load()
    rng = np.random.default_rng(11)
    ts = np.exp(rng.uniform(np.log(2.0 ** d["EMIN"]), np.log(2.0 ** d["EMAX"] * 0.999), 4000))
    ts = np.concatenate([ts, 1 + rng.uniform(-0.02, 0.02, 500), [0.5, 2.0, 1e-6, 3.999]])
    err = max(abs((LD(_g_dev(d, coef, t)) - _g_ref(t)) / _g_ref(t)) for t in ts)
    assert err < 1e-12, err


def test_table_special_values():
    d, coef = _load()
    assert abs(_g_dev(d, coef, 1.0) * 3.0 - 1.0) < 1e-12  # a piece boundary: interpolation error only
    # large t: no cancellation in the closed form, so plain FP64 is a second, independent reference
    for t in (37.5, 2.0 ** 20 + 3.0, 2.0 ** 28):
        s = np.sqrt(t - 1)
        g64 = (1 - np.arctan(s) / s) / (t - 1)
        assert abs(_g_dev(d, coef, t) / g64 - 1) < 1e-12, t
