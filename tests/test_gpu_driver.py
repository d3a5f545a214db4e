"""The driver's schedule and bail-out semantics (itpsystem.cpp:202-233, 258-289,
291-372, 407-470; convergence.cpp:75-200) on the GPU path against the oracle's
driver on the same seeded inputs: time-step lists (used on convergence, or
exhausted one per step), the divisor, the minimum-time-step and maximum-step
bail-outs, the absstdev test, and recovery from a non-positive overlap
eigenvalue (a zero initial state, --recover). Same step count, error
flag, final time step and converged count; energies to 1e-10 relative."""
import numpy as np
import pytest

import itp2d_amd as I
from itp2d_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def _base():
    sx = sy = 32
    L = 10.0
    V = W.harmonic(sx, sy, L / sx)
    cfg = dict(sx=sx, sy=sy, lenx=L, boundary=I.PERIODIC, N=10, n_wanted=6, halforder=2, B=0.5,
               eps_values=[0.1], eps_divisor=5.0, max_steps=300, min_time_step=1e-9, seed=W.SEED_BASE,
               timestep_test=("relchange", 1e-9), final_test=("relchange", 1e-9))
    return cfg, V


CASES = {
    "divisor": {"eps_values": [0.2], "timestep_test": ("relchange", 1e-7), "final_test": ("relchange", 1e-10)},
    "min_eps_bailout": {"eps_values": [0.1], "min_time_step": 0.05, "final_test": ("relchange", 1e-14)},
    "max_steps_bailout": {"max_steps": 9},
    "eps_list": {"eps_values": [0.3, 0.2, 0.1], "timestep_test": ("relchange", 1e-7),
                 "final_test": ("relchange", 1e-12)},
    "eps_list_exhausted": {"eps_values": [0.3, 0.2, 0.1, 0.05], "exhaust_eps": 1},
    "absstdev": {"timestep_test": ("absstdev", 1e-3, 1e-7), "final_test": ("absstdev", 1e-3, 1e-9)},
    "dirichlet_magnetic": {"boundary": I.DIRICHLET, "B": 1.0},
}


@pytest.mark.parametrize("case", list(CASES) + ["recover", "recover_eigenvalue_check"])
def test_driver_semantics_match_oracle(ctx, case):
    """recover_eigenvalue_check: the recovery case with the eigenvalue check
    (Cholesky test, then tridiagonalization and bisection before the
    eigendecomposition) on for every matrix size: the rejected first overlap
    skips the eigenvectors and the run is the same."""
    say(f"driver semantics {case}")
    cfg, V = _base()
    init = None
    if case == "recover_eigenvalue_check":
        ctx.set_eig_check_min(1)
        checks0 = ctx.eig_check_stats()
    if case.startswith("recover"):
        cfg.update(recover=1, eps_values=[0.1], max_steps=1500)
        init = I.init_states(cfg["seed"] + 7, cfg["N"], cfg["sx"], cfg["sy"], cfg["lenx"] / cfg["sx"])
        init[3] = 0  # an exactly zero overlap eigenvalue: the first orthonormalization fails
    else:
        cfg.update(CASES[case])
    g = I.ITPSystem(ctx, cfg, V, init_states=init)
    g.run()
    o = O.System(cfg, V, init_states=init)
    o.run()
    assert g.total_steps == o.total_steps
    assert g.error_flag == o.error_flag
    assert g.eps == pytest.approx(o.eps, rel=1e-15)
    assert g.num_converged == o.num_converged
    Eg, Eo = g.sorted()[0], o.sorted()[0]
    nw = cfg["n_wanted"]
    assert np.abs(Eg[:nw] - Eo[:nw]).max() / np.abs(Eo[:nw]).max() < 1e-10
    if case.startswith("recover"):
        assert g.recoveries == 1
        assert g.eps == pytest.approx(0.1 / 5 ** (len(g.time_step_history()) - 1))
    if case == "recover_eigenvalue_check":
        checks, rejected = (a - b for a, b in zip(ctx.eig_check_stats(), checks0))
        ctx.set_eig_check_min(1024)
        # only the first overlap (with a zero state) fails the Cholesky test and gets the check
        assert rejected == 1 and checks == 1
    if case.endswith("bailout"):
        assert g.error_flag
    if case == "eps_list_exhausted":
        assert [e for _, e in g.time_step_history()][:4] == [0.3, 0.2, 0.1, 0.05]


def test_save_flag_checked_and_cleared(ctx):
    """The save flag (the reference's SIGUSR1 flag, itpsystem.cpp:235-241) is
    checked after propagation, orthonormalization, the energies and at the end
    of every step: a flag raised before the run calls the save callback once,
    right after the first propagation, and is cleared; raising it again from
    the callback's own step calls it at the next check point. The run itself
    is unchanged (same steps and energies as without the flag)."""
    say("driver save flag")
    cfg, V = _base()
    ref = I.ITPSystem(ctx, cfg, V)
    ref.run()
    flag = __import__("ctypes").c_int(1)
    calls = []

    def save(sys_):
        st = sys_.states()
        assert st.shape == (cfg["N"], cfg["sy"], cfg["sx"])
        calls.append(sys_.total_steps)

    g = I.ITPSystem(ctx, cfg, V)
    g.set_save_flag(flag, save)
    g.step()
    assert calls == [1] and flag.value == 0
    flag.value = 1
    g.step()
    assert calls == [1, 2] and flag.value == 0
    g.run()
    assert calls == [1, 2]
    assert g.total_steps == ref.total_steps and g.num_converged == ref.num_converged
    assert np.array_equal(g.sorted()[0], ref.sorted()[0])
