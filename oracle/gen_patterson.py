"""Generate the nested Gauss-Patterson rules (1, 3, 7, ..., 511 points) from first principles.

The rules are the ones QLens tabulates in gauss.cpp:521-2099 and uses in
LensIntegral::PattersonIntegrate (profile.cpp:3830-3881): level 0 is the 1-point
midpoint rule, level 1 its Kronrod extension (= 3-point Gauss-Legendre), and each
further level is Patterson's optimal extension (T. N. L. Patterson, "The optimum
addition of points to quadrature formulae", Math. Comp. 22 (1968) 847-856): the n+1
new nodes are the zeros of the degree-(n+1) polynomial orthogonal to all polynomials of
degree <= n under the (sign-changing) weight pi_n(x) = prod (x - x_i) of the existing n
nodes.  Weights are the interpolatory weights of the full node set.

Everything is computed in mpmath at 120 significant digits and rounded once to binary64.
Output: qlens_amd/csrc/patterson_table.h (hex-float constants, shared by the HIP library
and by the CPU restatement in oracle/).  Table layout matches the reference's indexing:
nodes x[0..510] ascending on [-1,1]; level L (order n = 2^(L+1)-1) uses node indices
i = (k+1)*(512/(n+1)) - 1 for k = 0..n-1 and weight w[L][k].

Run:  python oracle/gen_patterson.py            (about a minute)
"""
import os
import sys
import mpmath as mp

mp.mp.dps = 120
NLEV = 9


def legendre_all(x, nmax):
    """P_0..P_nmax at x (three-term recurrence)."""
    p = [mp.mpf(1), x]
    for n in range(1, nmax):
        p.append(((2 * n + 1) * x * p[n] - n * p[n - 1]) / (n + 1))
    return p[: nmax + 1]


def gauss_legendre(m):
    """m-point Gauss-Legendre nodes/weights by Newton on P_m."""
    xs, ws = [], []
    for i in range(1, m + 1):
        x = mp.cos(mp.pi * (i - mp.mpf(1) / 4) / (m + mp.mpf(1) / 2))
        for _ in range(100):
            p0, p1 = mp.mpf(1), x
            for n in range(1, m):
                p0, p1 = p1, ((2 * n + 1) * x * p1 - n * p0) / (n + 1)
            dp = m * (x * p1 - p0) / (x * x - 1)
            dx = p1 / dp
            x -= dx
            if abs(dx) < mp.mpf(10) ** (-mp.mp.dps + 5):
                break
        p0, p1 = mp.mpf(1), x
        for n in range(1, m):
            p0, p1 = p1, ((2 * n + 1) * x * p1 - n * p0) / (n + 1)
        dp = m * (x * p1 - p0) / (x * x - 1)
        xs.append(x)
        ws.append(2 / ((1 - x * x) * dp * dp))
    return xs, ws


def clenshaw_even(c, x):
    """sum_m c[m] P_{2m}(x), evaluated by forward recurrence (degree <= 512, fine in mp)."""
    nmax = 2 * (len(c) - 1)
    p = legendre_all(x, nmax)
    return mp.fsum(c[m] * p[2 * m] for m in range(len(c)))


def extend(nodes):
    """Patterson extension of a symmetric odd-size rule: returns the n+1 new nodes."""
    n = len(nodes)
    deg = n + 1  # degree of the new node polynomial q (even, since n is odd)
    # q = sum_{m even <= deg} c_m P_m with c_deg = 1.  Conditions: int pi_n q P_k = 0 for k <= n;
    # pi_n is odd and q even, so only odd k give nonzero integrals.
    G = (3 * n + 3) // 2 + 2
    gx, gw = gauss_legendre(G)
    pis = []
    for x in gx:
        v = mp.mpf(1)
        for t in nodes:
            v *= x - t
        pis.append(v)
    evens = list(range(0, deg + 1, 2))
    odds = list(range(1, n + 1, 2))
    P = [legendre_all(x, deg) for x in gx]
    A = mp.matrix(len(odds), len(evens) - 1)
    rhs = mp.matrix(len(odds), 1)
    for r, k in enumerate(odds):
        row = [gw[g] * pis[g] * P[g][k] for g in range(G)]
        for cidx, m in enumerate(evens[:-1]):
            A[r, cidx] = mp.fsum(row[g] * P[g][m] for g in range(G))
        rhs[r] = -mp.fsum(row[g] * P[g][deg] for g in range(G))
    sol = mp.lu_solve(A, rhs)
    c = [sol[i] for i in range(len(evens) - 1)] + [mp.mpf(1)]
    # new positive roots interlace the existing nonnegative nodes and 1
    pos_old = sorted([t for t in nodes if t >= 0])
    brackets = pos_old + [mp.mpf(1)]
    roots = []
    for a, b in zip(brackets[:-1], brackets[1:]):
        fa, fb = clenshaw_even(c, a), clenshaw_even(c, b)
        if fa * fb > 0:
            raise RuntimeError("no sign change in bracket (%s, %s)" % (a, b))
        r = mp.findroot(lambda t: clenshaw_even(c, t), (a, b), solver="anderson")
        roots.append(r)
    newn = sorted([-r for r in roots] + roots)
    if len(newn) != n + 1:
        raise RuntimeError("wrong number of new nodes")
    return newn


def interp_weights(nodes):
    """Interpolatory weights: w_i = (1/w'(x_i)) * sum_g W_g w(y_g)/(y_g - x_i), exact Gauss-Legendre."""
    N = len(nodes)
    G = N // 2 + 4
    G += G % 2  # even order: no Gauss-Legendre node at 0, where every Patterson rule has one
    gx, gw = gauss_legendre(G)
    omega = []
    for y in gx:
        v = mp.mpf(1)
        for t in nodes:
            v *= y - t
        omega.append(v)
    ws = []
    for i, xi in enumerate(nodes):
        d = mp.mpf(1)
        for j, xj in enumerate(nodes):
            if j != i:
                d *= xi - xj
        s = mp.fsum(gw[g] * omega[g] / (gx[g] - xi) for g in range(G))
        ws.append(s / d)
    return ws


def build():
    levels_nodes = [[mp.mpf(0)]]
    for L in range(1, NLEV):
        prev = levels_nodes[-1]
        new = extend(prev)
        levels_nodes.append(sorted(prev + new))
        print("level %d: %d nodes" % (L, len(levels_nodes[-1])), file=sys.stderr, flush=True)
    full = levels_nodes[-1]
    assert len(full) == 511
    weights = [interp_weights(nodes) for nodes in levels_nodes]
    return full, levels_nodes, weights


def hexf(v):
    return float(v).hex()


def main():
    full, levels_nodes, weights = build()
    # consistency of the nested indexing with the reference's stride pattern
    for L, nodes in enumerate(levels_nodes):
        n = len(nodes)
        step = 512 // (n + 1)
        for k in range(n):
            assert full[(k + 1) * step - 1] == nodes[k]
    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "..", "qlens_amd", "csrc", "patterson_table.h")
    with open(out, "w") as f:
        f.write("// Generated by oracle/gen_patterson.py: nested Gauss-Patterson rules on [-1,1],\n")
        f.write("// computed from Patterson (1968) extensions at 120 digits and rounded to binary64.\n")
        f.write("// Nodes ascending; level L (order 2^(L+1)-1) uses node (k+1)*(512/(order+1))-1 with\n")
        f.write("// weight QLX_PAT_W[QLX_PAT_WOFF[L] + k].  Same indexing as gauss.cpp:521-2099.\n")
        f.write("#pragma once\n\n")
        f.write("#define QLX_PAT_N 511\n#define QLX_PAT_NLEV 9\n\n")
        f.write("static const double QLX_PAT_X[511] = {\n")
        for i in range(0, 511, 4):
            f.write("  " + ", ".join(hexf(v) for v in full[i:i + 4]) + ",\n")
        f.write("};\n\n")
        offs, o = [], 0
        for w in weights:
            offs.append(o)
            o += len(w)
        f.write("static const int QLX_PAT_WOFF[10] = {%s};\n\n" % ", ".join(str(v) for v in offs + [o]))
        f.write("static const double QLX_PAT_W[%d] = {\n" % o)
        allw = [v for w in weights for v in w]
        for i in range(0, len(allw), 4):
            f.write("  " + ", ".join(hexf(v) for v in allw[i:i + 4]) + ",\n")
        f.write("};\n")
    print("wrote", out, file=sys.stderr)


if __name__ == "__main__":
    main()
