"""HIP graph replay of the fixed launch sequences (propagation, overlap,
Cholesky factor): the same solves with graphs on and off give bit-identical
results. Every step of a solve replays the propagation graph (and, on the
mixed path, the overlap and Cholesky graphs) with the driver's own eager work
(energies, eigensolver, rotation) and host synchronizations between the
replays, and every step's energies are compared."""
import numpy as np
import pytest

import itp2d_amd as I
from itp2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def _history(ctx, cfg, V, graphs):
    ctx.set_graphs(graphs)
    try:
        s = I.ITPSystem(ctx, cfg, V)
        s.run()
        E, sd = s.history()
        out = (s.total_steps, s.num_converged, E, sd, s.states())
        s.close()
        return out
    finally:
        ctx.set_graphs(0)


@pytest.mark.parametrize("name,prec", [("c1_ho64", I.FP64), ("c2_small", I.FP64), ("c2_small", I.FP32)])
def test_graph_replay_matches_direct_launches(ctx, name, prec):
    say(f"graphs {name} prec {prec}")
    cfg, V = W.config(name)
    cfg = dict(cfg, precision=prec)
    a = _history(ctx, cfg, V, 1)
    b = _history(ctx, cfg, V, 0)
    assert a[0] == b[0] and a[1] == b[1]
    assert np.array_equal(a[2], b[2], equal_nan=True)
    assert np.array_equal(a[3], b[3], equal_nan=True)
    assert np.array_equal(a[4], b[4])


def test_graph_replay_follows_time_step_changes(ctx):
    """A time-step change rebuilds the split's tables: the propagation graph
    of the old tables must not be replayed (eps list and divisor case)."""
    say("graphs eps changes")
    cfg, V = W.config("c1_ho64")
    cfg = dict(cfg, eps_values=[0.3, 0.2, 0.1], timestep_test=("relchange", 1e-7), final_test=("relchange", 1e-9))
    a = _history(ctx, cfg, V, 1)
    b = _history(ctx, cfg, V, 0)
    assert a[0] == b[0] and a[1] == b[1]
    assert np.array_equal(a[2], b[2], equal_nan=True)
