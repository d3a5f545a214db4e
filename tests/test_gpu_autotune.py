"""Launch-shape autotuning of the overlap's split-K (the counterpart of the
reference's measured FFTW plans and wisdom file, itp2d.cpp:85-93, 124-128):
a tuned overlap equals the untuned one to rounding (only the order of the
chunk sums changes), a shape is measured once per process, and the wisdom
text round-trips, rejects malformed lines and, once imported, replaces the
measurement."""
import numpy as np
import pytest

import itp2d_amd as I
from itp2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def _set(ctx, sx, N, prec, seed):
    g = I.Grid(ctx, sx, sx, 10.0 / sx, I.PERIODIC)
    s = I.StateSet(g, N, prec)
    s.upload(I.init_states(seed, N, sx, sx, 10.0 / sx))
    return g, s


def _wisdom(ctx):
    return {tuple(int(x) for x in line.split()[1:4]): int(line.split()[4])
            for line in ctx.wisdom_export().splitlines()}


@pytest.mark.parametrize("prec,tol", [(I.FP64, 1e-13), (I.FP32, 1e-5)])
def test_tuned_overlap_matches_untuned(ctx, prec, tol):
    say(f"autotune overlap prec {prec}")
    sx, N = 192, 136
    g, s = _set(ctx, sx, N, prec, W.SEED_BASE + 3)
    ref = s.overlap()
    t0 = ctx.tunings()
    ctx.set_autotune(1)
    try:
        a = s.overlap()   # tunes the shape (unless an earlier test of this process did)
        b = s.overlap()   # the tuned count, no new measurement
        g2, s2 = _set(ctx, sx, N, prec, W.SEED_BASE + 3)
        c = s2.overlap()  # a new set of the same shape: the wisdom's count
    finally:
        ctx.set_autotune(0)
    assert ctx.tunings() - t0 <= 1
    key = (8 if prec == I.FP64 else 4, sx * sx, 192)
    assert key in _wisdom(ctx)
    scale = np.abs(ref).max()
    for x in (a, b, c):
        assert np.abs(x - ref).max() / scale < tol
        assert np.allclose(x, x.conj().T, rtol=0, atol=tol * scale)
    assert np.array_equal(a, b) and np.array_equal(b, c)


def test_wisdom_round_trip_and_import(ctx):
    say("autotune wisdom text")
    # a shape no workload has (the wisdom is the process's: other tests add theirs)
    with pytest.raises(I.ITPError):
        ctx.wisdom_import("gram 8 12345 64\n")  # a field short
    with pytest.raises(I.ITPError):
        ctx.wisdom_import("plan 8 12345 64 2\n")
    with pytest.raises(I.ITPError):
        ctx.wisdom_import("gram 16 12345 64 2\n")  # 4 or 8 bytes per real
    with pytest.raises(I.ITPError):
        ctx.wisdom_import("gram 8 12345 100 2\n")  # padded state counts are multiples of 64
    with pytest.raises(I.ITPError):
        ctx.wisdom_import("gram 8 12345 64 2\ngram 8 12345 64 0\n")  # nothing of a bad text is added
    assert (8, 12345, 64) not in _wisdom(ctx)
    # an imported shape is used as is: no measurement, the count it names
    sx, N = 64, 40
    ctx.wisdom_import(f"gram 8 {sx * sx} 64 1\n")
    assert _wisdom(ctx)[(8, sx * sx, 64)] == 1
    g, s = _set(ctx, sx, N, I.FP64, W.SEED_BASE + 5)
    ref = s.overlap()
    t0 = ctx.tunings()
    ctx.set_autotune(1)
    try:
        a = s.overlap()
    finally:
        ctx.set_autotune(0)
    assert ctx.tunings() == t0
    assert np.abs(a - ref).max() / np.abs(ref).max() < 1e-13
    # export -> import gives the same wisdom
    text = ctx.wisdom_export()
    before = _wisdom(ctx)
    ctx.wisdom_import(text)
    assert _wisdom(ctx) == before


def test_batch_with_autotune_matches(ctx):
    """The same batch with and without tuning: the same steps and converged
    counts, energies to 1e-10 relative."""
    say("autotune batch")
    cfgs = []
    for name in ("c1_ho64", "c2_small"):
        cfg, V = W.config(name)
        cfgs.append((cfg, V))
    out = []
    for tune in (0, 1):
        b = I.Batch(0, 2)
        b.set_autotune(tune)
        for cfg, V in cfgs:
            b.add(cfg, V)
        assert b.run() == 0
        out.append([(b.result(i), b.sorted(i)) for i in range(len(cfgs))])
        b.close()
    for (cfg, _), (r0, s0), (r1, s1) in zip(cfgs, *out):
        nw = cfg["n_wanted"]
        assert r0["steps"] == r1["steps"] and r0["converged"] == r1["converged"]
        assert np.abs(s0[0][:nw] - s1[0][:nw]).max() / np.abs(s0[0][:nw]).max() < 1e-10
