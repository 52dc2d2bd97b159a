"""CPU restatement of the deflection-surrogate network (test infrastructure: the checker of the
GPU training kernels; never part of the product path).

MLP 4 -> 128 x4 (SiLU) -> 5, parameters flattened as W1[128][4] b1[128] W2[128][128] b2 W3 b3 W4
b4 W5[5][128] b5[5] (50821 values).  The loss is the mean squared error over n samples x 5
outputs.  With emulate=True the arithmetic is rounded where the HIP kernels round: every weight to
BF16, every hidden activation h = silu(z) to BF16, the back-propagated dz of hidden layers 2-4 to
BF16 (they feed MFMA instructions); everything else, and every sum, is carried in float64 here
(FP32 on the GPU).  With emulate=False it is the exact float64 network, pinned against finite
differences in tests/test_oracle.py.

The network replaces the reference's tabulated deflection option (models.cpp:4054-5148); its
targets are the integrals of LensProfile::deflection_and_hessian_numerical (profile.cpp:2532-2560).
"""
import numpy as np

NH = 128
NPARAM = 50821
_SHAPES = [("W1", (NH, 4)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)), ("b3", (NH,)),
           ("W4", (NH, NH)), ("b4", (NH,)), ("W5", (5, NH)), ("b5", (5,))]


def bf16(a):
    """Round to bfloat16 (round to nearest, ties to even) and return float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    assert P.size == NPARAM
    out, o = {}, 0
    for name, shp in _SHAPES:
        n = int(np.prod(shp))
        out[name] = P[o:o + n].reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in _SHAPES])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * _sig(z)


def dsilu(z):
    s = _sig(z)
    return s * (1.0 + z * (1.0 - s))


def forward(P, X, emulate=True):
    """Network outputs y[n, 5] and the cached (z, h) of the hidden layers."""
    p = unpack(P)
    rw = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    rh = bf16 if emulate else (lambda a: a)
    X = np.asarray(X, dtype=np.float64)
    zs, hs = [], []
    z = X @ rw(p["W1"]).T + p["b1"]
    h = rh(silu(z))
    zs.append(z)
    hs.append(h)
    for k in ("2", "3", "4"):
        z = h @ rw(p["W" + k]).T + p["b" + k]
        h = rh(silu(z))
        zs.append(z)
        hs.append(h)
    y = h @ rw(p["W5"]).T + p["b5"]
    return y, zs, hs


def loss_grad(P, X, T, emulate=True):
    """Mean-squared-error loss and its gradient (flattened like P)."""
    p = unpack(P)
    rw = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    rd = bf16 if emulate else (lambda a: a)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    n = X.shape[0]
    y, zs, hs = forward(P, X, emulate)
    r = y - T
    loss = float(np.sum(r * r) / (5.0 * n))
    dy = 2.0 * r / (5.0 * n)
    g = {"W5": dy.T @ hs[3], "b5": dy.sum(0)}
    dz = rd((dy @ rw(p["W5"])) * dsilu(zs[3]))
    for layer, k in ((3, "4"), (2, "3"), (1, "2")):
        g["W" + k] = dz.T @ hs[layer - 1]
        g["b" + k] = dz.sum(0)
        dh = dz @ rw(p["W" + k])
        dz = dh * dsilu(zs[layer - 1])
        if layer > 1:
            dz = rd(dz)
    g["W1"] = dz.T @ X
    g["b1"] = dz.sum(0)
    return loss, pack(g)
