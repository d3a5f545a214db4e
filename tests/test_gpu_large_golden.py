"""The reduced forms of the large configurations, and the 256x256 Dirichlet
configuration at its full size, against oracle goldens
(tests/golden/itp_<name>.npz, made by tests/golden/make_itp_golden.py):
c3_ho128 is the 512x512 oscillator's physics at 128x128 with N = 200,
c5_stadium256 the magnetic hard-wall stadium (B = 2, 6th-order split,
eps 0.02, recovery) at 256x256 with N = 100, and c2_dirichlet256 the full
256x256, N = 500, B = 1 workload (the oracle's 55 steps took about an hour on
8 cores). The FP64 path reproduces each golden run (steps, converged count,
recoveries' outcome, eigenvalues to 1e-10); the mixed path converges every
state the golden converges, within 1e-8."""
import os

import numpy as np
import pytest

import itp2d_amd as I
from itp2d_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
NAMES = ["c3_ho128", "c5_stadium256", "c2_dirichlet256"]


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def _gold(name):
    f = os.path.join(GOLD, f"itp_{name}.npz")
    if not os.path.exists(f):
        pytest.skip(f"{f} not generated")
    return np.load(f)


@pytest.mark.parametrize("name", NAMES)
def test_reduced_fp64_matches_oracle(ctx, name):
    ref = _gold(name)
    cfg, V = W.config(name)
    say(f"{name} fp64")
    s = I.ITPSystem(ctx, cfg, V)
    s.run()
    assert s.error_flag == bool(ref["error"])
    assert s.total_steps == int(ref["steps"])
    assert s.num_converged == int(ref["converged"])
    c = ref["conv_flags"].astype(bool)
    E = s.sorted()[0]
    assert np.abs(E[c] - ref["E"][c]).max() / np.abs(ref["E"][c]).max() < 1e-10
    s.close()


@pytest.mark.parametrize("name", NAMES)
def test_reduced_mixed_matches_oracle(ctx, name):
    ref = _gold(name)
    cfg, V = W.config(name)
    say(f"{name} mixed")
    s = I.ITPSystem(ctx, dict(cfg, precision=I.FP32), V)
    s.run()
    assert not s.error_flag
    c = ref["conv_flags"].astype(bool)
    conv = s.converged()
    assert conv[c].all()
    E = s.sorted()[0]
    assert (np.abs(E[c] - ref["E"][c]) / np.abs(ref["E"][c])).max() < 1e-8
    s.close()
