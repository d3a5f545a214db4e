"""The batched configuration against the oracle: eight instances of
workloads.c4_instances (all three grid sizes, 53-138 states, B from 0.008 to
1.86) with committed oracle goldens (tests/golden/itp_c4_<i>.npz, made by
tests/golden/make_itp_golden.py). The FP64 path reproduces each golden run:
same step and converged counts, eigenvalues to 1e-10 relative. The mixed path,
run on all eight together through the batch API, converges every state the
golden converges, with energies within 1e-8 relative."""
import os

import numpy as np
import pytest

import itp2d_amd as I
from itp2d_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
INSTANCES = [0, 27, 41, 25, 46, 1, 5, 11]


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def _gold(i):
    return np.load(os.path.join(GOLD, f"itp_c4_{i}.npz"))


@pytest.mark.parametrize("i", INSTANCES)
def test_c4_instance_fp64_matches_oracle(ctx, i):
    cfg, V = W.c4_instances(64)[i]
    say(f"c4 instance {i}: {cfg['sx']}x{cfg['sy']} N={cfg['N']} B={cfg['B']:.3f} fp64")
    ref = _gold(i)
    s = I.ITPSystem(ctx, cfg, V)
    s.run()
    assert not s.error_flag
    assert s.total_steps == int(ref["steps"])
    assert s.num_converged == int(ref["converged"])
    E = s.sorted()[0]
    assert np.array_equal(s.converged(), ref["conv_flags"].astype(bool))
    c = ref["conv_flags"].astype(bool)
    assert np.abs(E[c] - ref["E"][c]).max() / np.abs(ref["E"][c]).max() < 1e-10
    s.close()


def test_c4_batch_mixed_matches_oracle():
    insts = W.c4_instances(64)
    say(f"c4 batch of {len(INSTANCES)} instances, mixed path")
    b = I.Batch(0, workers=4)
    for i in INSTANCES:
        cfg, V = insts[i]
        b.add(dict(cfg, precision=I.FP32), V)
    assert b.run() == 0
    for k, i in enumerate(INSTANCES):
        ref = _gold(i)
        r = b.result(k)
        assert r["status"] == 0 and not r["error_flag"]
        E, _, conv = b.sorted(k)
        c = ref["conv_flags"].astype(bool)
        assert conv[c].all(), i  # every state the oracle converges
        err = np.abs(E[c] - ref["E"][c]) / np.abs(ref["E"][c])
        assert err.max() < 1e-8, (i, err.max())
        # the FP32 phase did the work: a few FP64 steps at most beyond the hand-off
        st = b.mixed_stats(k)
        assert st[0] > 0 and r["steps"] - st[0] <= 6, (i, st, r["steps"])
    b.close()
