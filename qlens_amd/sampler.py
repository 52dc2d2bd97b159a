"""Batched proposal evaluation for QLens's samplers (the sampler batching adapter).

The reference's nested sampler UCMC::MonoSample (mcmchdr.cpp:2901-3180) calls LOGLIKE once per
proposal:
  * the N initial live points one by one (mcmchdr.cpp:2963-2993; a point whose -ln L is NaN or
    infinite is drawn again),
  * then, to replace the worst live point (largest -ln L, likeMin in the reference's naming), fresh
    uniform draws in the prior box until one has -ln L below it (3057-3091).  Under MPI each group
    evaluates one draw and the first acceptable draw in group order wins.

Here those loops read their proposals from a stream of draws that is evaluated `batch` at a time
by qlx_loglike (one batched GPU evaluation, the A1 wrapper fitmodel_loglike_extended_source).  The
loops consume the stream in draw order, as the one-proposal loop does; evaluated draws not used by
one replacement are kept for the next.  So the accepted points, their -ln L and the number of draws
consumed are exactly those of the one-proposal-per-call loop on the same draws (no evaluation is
wasted either), while the GPU sees `batch` proposals per launch.  Draws are mapped to parameters by
UCMC::Convert (mcmchdr.cpp:2837-2843): x = lo + u (hi - lo).

The T-Walk kernel (mcmchdr.cpp, TWalk) advances one pair of walkers per likelihood call; its
batched form here evaluates the current proposals of many independent chains in one call
(`evaluate_chains`).
"""
import numpy as np


class BatchedMonoSampler:
    """Rejection phase of MonoSample with batched likelihood calls.

    inv     a qlx.PixelInversion with image, PSF, settings and fit model (set_fit_model) loaded
    lo, hi  lower / upper limits of the sampler coordinates (the reference's lowerLimits/upperLimits)
    batch   proposals per GPU call
    rng     numpy Generator for the unit-cube draws
    """

    def __init__(self, inv, lo, hi, batch=256, rng=None):
        self.inv = inv
        self.lo = np.asarray(lo, dtype=np.float64)
        self.hi = np.asarray(hi, dtype=np.float64)
        self.ma = len(self.lo)
        self.batch = int(batch)
        self.rng = rng if rng is not None else np.random.default_rng(0)
        self.calls = 0        # batched likelihood calls
        self.evaluations = 0  # proposals evaluated

    def convert(self, u):
        """UCMC::Convert: unit-cube point(s) -> sampler coordinates."""
        return self.lo + np.asarray(u) * (self.hi - self.lo)

    def loglike(self, u):
        """-ln L of unit-cube points u[n, ma] in one batched call."""
        ll, _ = self.inv.loglike(self.convert(u))
        self.calls += 1
        self.evaluations += len(u)
        return ll

    def _draw(self):
        """The next (u, -ln L) of the draw stream, evaluating `batch` new draws when it runs dry."""
        if not getattr(self, "_buf", None):
            u = self.rng.random((self.batch, self.ma))
            ll = self.loglike(u)
            self._buf = list(zip(u, ll))[::-1]
        return self._buf.pop()

    def initial_live_points(self, n):
        """n live points with finite -ln L (mcmchdr.cpp:2963-2993): a draw whose -ln L is not finite
        is replaced by the next."""
        pts, lls = [], []
        while len(pts) < n:
            u, ll = self._draw()
            if np.isfinite(ll):
                pts.append(u)
                lls.append(ll)
        return np.array(pts), np.array(lls)

    def replace_worst(self, live_u, live_ll):
        """Replace the worst live point (largest -ln L) by the first draw whose -ln L is below it
        (mcmchdr.cpp:3057-3091).  Returns (index replaced, new -ln L, draws consumed)."""
        imin = int(np.argmax(live_ll))
        like_min = live_ll[imin]
        used = 0
        while True:
            u, ll = self._draw()
            used += 1
            if ll < like_min:
                live_u[imin] = u
                live_ll[imin] = ll
                return imin, ll, used


def evaluate_chains(inv, proposals):
    """The current proposals of many independent chains (e.g. T-Walk chains, one proposal each) in
    one batched likelihood call; returns their -ln L."""
    ll, _ = inv.loglike(np.asarray(proposals, dtype=np.float64))
    return ll
