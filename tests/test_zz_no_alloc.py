"""itp_propagate_noalloc: the propagation entry point that allocates nothing
(its intermediate fields live in the library's fixed device scratch array).
Checked through the C ABI against itp_propagate on the same inputs, bit for bit
(the same kernels on the same data), on two calls in a row, the second with a
smaller set; and a set beyond the array's size is rejected before any work."""
import numpy as np
import pytest

import itp2d_amd as I
from itp2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def _states(n, sx, sy, dx, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n, sy, sx)) + 1j * rng.standard_normal((n, sy, sx))
    return f / (np.sqrt((np.abs(f) ** 2).sum(axis=(1, 2), keepdims=True)) * dx)


@pytest.mark.parametrize("precision", [I.FP64, I.FP32])
def test_propagate_noalloc_matches_propagate(ctx, precision):
    sx = sy = 64
    L = 12.0
    dx = L / sx
    g = I.Grid(ctx, sx, sy, dx, I.DIRICHLET)
    pot = I.Potential(g, W.harmonic(sx, sy, dx))
    sp = I.MultiProductSplit(pot, 2, 0.05, 0.7)
    for call, N in enumerate([12, 5]):  # the second call with a smaller input
        say(f"noalloc propagate call {call}: N={N} precision={precision}")
        f = _states(N, sx, sy, dx, 10 + N)
        a, b = I.StateSet(g, N, precision), I.StateSet(g, N, precision)
        a.upload(f)
        b.upload(f)
        sp.propagate(a)
        sp.propagate_noalloc(b)
        ra, rb = a.download(), b.download()
        assert np.array_equal(ra, rb)
        assert np.abs(ra - f).max() > 1e-3  # it did propagate
        a.close()
        b.close()


def test_propagate_noalloc_rejects_large_sets(ctx):
    sx = sy = 256
    dx = 20.0 / sx
    g = I.Grid(ctx, sx, sy, dx, I.PERIODIC)
    pot = I.Potential(g, W.harmonic(sx, sy, dx))
    sp = I.MultiProductSplit(pot, 1, 0.05, 0.0)
    N = I.noalloc_max_bytes() // (sx * sy * 16) + 1
    say(f"noalloc propagate rejects N={N} on {sx}x{sy}")
    s = I.StateSet(g, N)
    f = _states(2, sx, sy, dx, 3)
    s.upload(np.concatenate([f] * ((N + 1) // 2))[:N])
    before = s.download()
    with pytest.raises(I.ITPError) as e:
        sp.propagate_noalloc(s)
    assert e.value.code == 6  # ITP_E_UNSUPPORTED
    assert np.array_equal(s.download(), before)
    s.close()
