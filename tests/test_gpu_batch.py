"""Batches of independent instances (include/itp2d_batch.h) on the GPU.

Every instance of a batch runs the ITPSystem loop on its worker's own
context and stream while other instances run beside it; the kernels are
deterministic, so each instance must reproduce its single-instance run
exactly (same steps, same energies and deviations to the bit, same
convergence flags). The single-instance loop itself is pinned against the
oracle in test_gpu_parity.py.
"""
import numpy as np
import pytest

import itp2d_amd as I
from itp2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def small_instances():
    out = []
    c1, V1 = W.config("c1_ho64")
    out.append((c1, V1))
    c2, V2 = W.config("c2_small")
    out.append((c2, V2))
    out.append((dict(c2, N=24, n_wanted=18, B=0.4, halforder=1, seed=c2["seed"] + 5), V2))
    out.append((dict(c1, boundary=1, B=0.9, N=16, n_wanted=12, seed=c1["seed"] + 9), V1))
    out.append((dict(c2, precision=I.FP32, seed=c2["seed"] + 3), V2))  # mixed path
    return out


def test_batch_matches_single(ctx):
    insts = small_instances()
    say(f"batch of {len(insts)} instances, 3 workers")
    b = I.Batch(0, workers=3)
    for cfg, V in insts:
        b.add(cfg, V)
    assert len(b) == len(insts)
    assert b.run() == 0
    assert b.wall_time > 0
    for i, (cfg, V) in enumerate(insts):
        say(f"single instance {i}")
        s = I.ITPSystem(ctx, cfg, V)
        assert s.run() == 0
        E, sd, _ = s.sorted()
        r = b.result(i)
        assert r["status"] == 0 and not r["error_flag"]
        assert r["steps"] == s.total_steps
        assert r["converged"] == s.num_converged
        bE, bsd, bf = b.sorted(i)
        assert np.array_equal(bE, E), i
        assert np.array_equal(bsd, sd), i
        assert np.array_equal(bf, s.converged()), i
        assert r["converged"] >= cfg["n_wanted"]


def test_batch_c4_subset():
    """The first instances of the batched configuration (128^2 to 256^2 grids,
    up to 400 states, B in [0, 2)) all converge their wanted states."""
    insts = W.c4_instances(4)
    say("c4 subset: " + ", ".join(f"{c['sx']}^2 N={c['N']} B={c['B']:.2f}" for c, _ in insts))
    b = I.Batch(0, workers=4)
    for cfg, V in insts:
        b.add(cfg, V)
    assert b.run() == 0
    for i, (cfg, _) in enumerate(insts):
        r = b.result(i)
        assert r["status"] == 0 and not r["error_flag"], (i, r)
        assert r["converged"] >= cfg["n_wanted"], (i, r)
        E, sd, f = b.sorted(i)
        assert np.all(np.diff(E) >= 0)


def test_destroyed_objects_memory_is_reused_and_cleared():
    """A destroyed set's device memory serves the next set of its context
    (no hipFree, which would wait for every stream of the device): the new
    set gets the old field buffer and reads as zeros."""
    say("device memory reuse")
    c = I.Context(0)
    try:
        g = I.Grid(c, 64, 64, 0.2)
        a = I.StateSet(g, 40)
        a.upload(I.init_states(W.SEED_BASE, 40, 64, 64, 0.2))
        pa = I.lib().itp_stateset_device_ptr(a.h)
        a.close()
        b = I.StateSet(g, 40)
        assert I.lib().itp_stateset_device_ptr(b.h) == pa
        assert not np.any(b.download())
        b.close()
    finally:
        c.close()


def test_context_released_by_its_last_object():
    """Objects may be destroyed after their context's handle (the reference's
    objects have no such order): the context, its stream and its cached
    memory go with the last object."""
    import ctypes as C
    say("context outlived by its objects")
    L = I.lib()
    ctx, g, s = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert L.itp_ctx_create(0, C.byref(ctx)) == 0
    assert L.itp_grid_create(ctx, 32, 32, C.c_double(0.3), I.PERIODIC, C.byref(g)) == 0
    assert L.itp_stateset_create(g, 8, I.FP64, C.byref(s)) == 0
    L.itp_ctx_destroy(ctx)
    L.itp_grid_destroy(g)
    L.itp_stateset_destroy(s)
