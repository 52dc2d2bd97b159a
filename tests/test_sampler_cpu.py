"""The sampler batching adapter's draw-stream logic on the CPU, with a stand-in likelihood."""
import numpy as np

from qlens_amd.sampler import BatchedMonoSampler


class _Stub:
    """-ln L of a Gaussian bump, with infinite values in a corner (they are drawn again)."""

    def __init__(self):
        self.calls = 0

    def loglike(self, x):
        x = np.asarray(x)
        self.calls += 1
        ll = 0.5 * ((x - 0.3) ** 2).sum(1) / 0.01
        ll[(x[:, 0] > 0.95)] = np.inf
        return ll, None


def test_batch_size_does_not_change_the_chain():
    out = {}
    for batch in (1, 3, 50):
        stub = _Stub()
        smp = BatchedMonoSampler(stub, [0, 0], [1, 1], batch=batch, rng=np.random.default_rng(5))
        u, ll = smp.initial_live_points(20)
        hist = [smp.replace_worst(u, ll) for _ in range(40)]
        out[batch] = (u, ll, hist, stub.calls)
    for b in (3, 50):
        assert np.array_equal(out[1][0], out[b][0]) and np.array_equal(out[1][1], out[b][1])
        assert out[1][2] == out[b][2]
    assert out[50][3] < out[3][3] < out[1][3]
    # the live set only improves: every accepted -ln L was below the worst it replaced
    assert np.all(np.isfinite(out[1][1]))
