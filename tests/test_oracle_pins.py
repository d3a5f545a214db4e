"""More pins of the FP64 CPU oracle: the reference's own known answers for
State (test_state.cpp:139-183), EigenSolver (test_eigensolver.cpp:25-74),
Laplacian (test_laplacian.cpp:38-58) and the Hamiltonian stability test
(test_hamiltonian.cpp:40-106), restated; and the energy deviation sigma
against what the mathematics fixes: for a normalised state and a Hermitian H,
<psi|H H psi> - <psi|H psi>^2 = ||(H - E) psi||^2 (evaluated with a dense H
and numpy), and sigma vanishes for an eigenvector of that dense H.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from itp2d_amd import workloads as W

EPS = np.finfo(float).eps
NaN = float("nan")


# --- test_state.cpp: 2x2 fields, dx = 1; the reference's A(x, y) is [y][x] here
def _A():
    return np.full((2, 2), 1 + 2j)


def _B():
    B = np.zeros((2, 2), complex)
    B[0, 0] = 1 + 0j  # B(0,0)
    B[1, 0] = 1 + 2j  # B(0,1)
    B[0, 1] = 1 + 0j  # B(1,0)
    B[1, 1] = 1 + 2j  # B(1,1)
    return B


def test_reference_state_dot_and_norm():
    z, nA = O.dot_norm(_A(), _B(), 1.0)
    assert z == 12 - 4j
    assert nA == np.sqrt(20)
    zb, nB = O.dot_norm(_B(), _A(), 1.0)
    assert zb == np.conj(z)
    assert nB == np.sqrt(12)


def test_reference_pointwise_multiply():
    out = O.pointwise(_A(), "multiply", [1.0, 2.0, 3.0, 4.0])
    # A(0,0), A(0,1), A(1,0), A(1,1)
    assert (out[0, 0], out[1, 0], out[0, 1], out[1, 1]) == (1 + 2j, 3 + 6j, 2 + 4j, 4 + 8j)


def test_reference_pointwise_multiply_imaginary_shiftx():
    """new(x) = old(x-1) * i v(x-1) for x >= 1 and new(0) = 0 (state.hpp:194-201): the NaN
    entries sit where v(x-1) is never read (the top x column), so any read of them shows."""
    out = O.pointwise(_A(), "multiply_imaginary_shiftx", [1.0, NaN, 2.0, NaN])
    assert (out[0, 0], out[1, 0], out[0, 1], out[1, 1]) == (0, 0, -2 + 1j, -4 + 2j)
    assert np.isfinite(out).all()


def test_reference_pointwise_multiply_y():
    out = O.pointwise(_A(), "multiply_y", [1.0, 2.0, NaN])
    assert (out[0, 0], out[1, 0], out[0, 1], out[1, 1]) == (1 + 2j, 2 + 4j, 1 + 2j, 2 + 4j)


def test_reference_pointwise_multiply_and_add():
    out = O.pointwise(_A(), "multiply_and_add", [1.0, 2.0, 3.0, 4.0], _B())
    assert (out[0, 0], out[1, 0], out[0, 1], out[1, 1]) == (2 + 2j, 4 + 8j, 3 + 4j, 5 + 10j)


def test_shiftx_on_a_wide_grid_is_a_shift():
    """On a 5x3 field the shift map moves column x-1 to x times i v(x-1) and zeroes column 0."""
    rng = np.random.default_rng(1)
    f = rng.standard_normal((3, 5)) + 1j * rng.standard_normal((3, 5))
    v = rng.standard_normal((3, 5))
    v[:, -1] = NaN
    out = O.pointwise(f, "multiply_imaginary_shiftx", v.ravel())
    expect = np.zeros_like(f)
    expect[:, 1:] = f[:, :-1] * 1j * v[:, :-1]
    assert np.array_equal(out, expect)


# --- test_eigensolver.cpp:25-74
def test_reference_eigensolver_2x2_real():
    """[[1, 2], [2, 4]]: eigenvalues 0 and 5 with eigenvectors (-2, 1)/sqrt5 and
    (1, 2)/sqrt5. The reference compares zheev's output exactly; the oracle's
    Jacobi solver is held to 4 ulp, and its eigenvectors equal those up to a
    unit phase factor (the phase is LAPACK's convention, not the algorithm's)."""
    A = np.array([[1, 2], [2, 4]], complex)
    w, V = O.eigh(A)
    assert abs(w[0]) < 4 * EPS * 5 and abs(w[1] - 5) < 4 * EPS * 5
    for n, e in enumerate([np.array([-2, 1]) / np.sqrt(5), np.array([1, 2]) / np.sqrt(5)]):
        ph = np.vdot(e, V[:, n])
        assert abs(abs(ph) - 1) < 4 * EPS
        assert np.abs(V[:, n] - ph * e).max() < 4 * EPS


def _reference_8x8():
    N = 8
    A = np.zeros((N, N), complex)
    for i in range(N):
        for j in range(N):
            A[i, j] = 1 if i == j else (1j * abs(i - j) if j > i else -1j * abs(i - j))
    return A


def test_reference_eigensolver_spectral_decomposition_8x8():
    A = _reference_8x8()
    w, V = O.eigh(A)
    assert np.linalg.norm(A - (V * w) @ V.conj().T) < 300 * EPS
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - np.linalg.eigvalsh(A)).max() < 100 * EPS * np.abs(w).max()


# --- test_laplacian.cpp:38-58: the Laplacian of exp(-r^2) is 4 (r^2 - 1) exp(-r^2)
def test_reference_laplacian_of_gaussian():
    n, L = 42, 12.0
    dx = L / n
    px, py = W.positions(n, n, dx)
    X, Y = np.meshgrid(px, py)
    G = np.exp(-(X * X + Y * Y)) + 0j
    R = 4 * (X * X + Y * Y - 1) * np.exp(-(X * X + Y * Y))
    # Kinetic at B = 0 is -1/2 Laplacian (kinetic.cpp:25-60)
    lap = -2 * O.apply_kinetic(G[None], dx, O.PERIODIC, 0.0)[0]
    assert np.sqrt(np.mean(np.abs(lap - R) ** 2)) < 100 * EPS
    assert np.abs(lap - R).max() < 1000 * EPS


# --- test_hamiltonian.cpp:40-106: the ITP ground state is stable under H / E0
@pytest.mark.parametrize("case", ["box", "harmonic", "harmonic_magnetic"])
def test_reference_hamiltonian_ground_state_stable(case):
    n = 40
    if case == "box":
        L, bt, B, V, h = np.pi, O.DIRICHLET, 0.0, None, 5
    else:
        L, bt, h = 10.0, O.PERIODIC, 5
        B = 1.0 if case == "harmonic_magnetic" else 0.0
        V = W.harmonic(n, n, L / n)
    cfg = dict(sx=n, sy=n, lenx=L, boundary=bt, N=1, n_wanted=1, halforder=h, B=B, eps_values=[1.0],
               eps_divisor=5.0, max_steps=50, min_time_step=1e-9, seed=W.SEED_BASE,
               timestep_test=("relstdev", 1e-3, 1e-4), final_test=("relstdev", 1e-3, 0.0))
    s = O.System(cfg, V)
    s.run()
    assert not s.error_flag
    E, _, idx = s.sorted()
    g = s.states()[idx[0]]
    t = g.copy()
    for _ in range(3):
        t = O.apply_hamiltonian(t[None], L / n, bt, B, V)[0] / E[0]
        assert np.sqrt(np.mean(np.abs(t - g) ** 2)) < 5e-3


# --- sigma against the dense Hamiltonian
def _dense_h(sx, sy, dx, bt, B, V):
    n = sx * sy
    H = np.zeros((n, n), complex)
    for k in range(n):
        e = np.zeros(n, complex)
        e[k] = 1
        H[:, k] = O.apply_hamiltonian(e.reshape(1, sy, sx), dx, bt, B, V)[0].ravel()
    return H


BRANCHES = [(O.PERIODIC, 0.0), (O.DIRICHLET, 0.0), (O.PERIODIC, 0.8), (O.DIRICHLET, 0.8)]


@pytest.mark.parametrize("bt,B", BRANCHES)
def test_sigma_against_dense_hamiltonian(bt, B):
    """E = Re <psi|H psi>, sigma = Re sqrt(<psi|H H psi> - <psi|H psi>^2) with the complex
    principal square root (operators.cpp:40-50), from the oracle's two H applications,
    against numpy on the dense H: on the Hermitian branches sigma^2 also equals
    ||(H - E) psi||^2 dx^2, which fixes the formula without restating it."""
    sx = sy = 12
    L = 8.0
    dx = L / sx
    V = W.harmonic(sx, sy, dx) + 0.3 * np.cos(W.positions(sx, sy, dx)[0])[None, :]
    H = _dense_h(sx, sy, dx, bt, B, V)
    rng = np.random.default_rng(21)
    # smooth random states (their energies stay well inside the grid's spectrum)
    w, U = np.linalg.eig(H)
    order = np.argsort(w.real)
    lo = U[:, order[:20]]
    psi = (lo @ (rng.standard_normal((20, 4)) + 1j * rng.standard_normal((20, 4)))).T
    psi /= np.sqrt((np.abs(psi) ** 2).sum(axis=1, keepdims=True)) * dx
    E, sd, mean, msq = O.energies(psi.reshape(4, sy, sx), dx, bt, B, V)
    m1 = np.einsum("ni,ij,nj->n", psi.conj(), H, psi) * dx * dx
    m2 = np.einsum("ni,ij,jk,nk->n", psi.conj(), H, H, psi) * dx * dx
    assert np.abs(mean - m1).max() < 1e-12 * np.abs(m1).max()
    assert np.abs(msq - m2).max() < 1e-12 * np.abs(m2).max()
    assert np.abs(E - m1.real).max() < 1e-12 * np.abs(m1).max()
    assert np.abs(sd - np.sqrt(m2 - m1 * m1).real).max() < 1e-9 * np.abs(sd).max()
    if not (bt == O.DIRICHLET and B != 0):  # Hermitian branches
        assert np.abs(H - H.conj().T).max() < 1e-12 * np.abs(H).max()
        r = (psi @ H.T) - E[:, None] * psi
        res = np.sqrt((np.abs(r) ** 2).sum(axis=1)) * dx
        assert np.abs(sd - res).max() < 1e-9 * res.max()


@pytest.mark.parametrize("bt,B", [b for b in BRANCHES if not (b[0] == O.DIRICHLET and b[1] != 0)])
def test_sigma_vanishes_for_eigenvectors(bt, B):
    sx = sy = 12
    L = 8.0
    dx = L / sx
    V = W.harmonic(sx, sy, dx)
    H = _dense_h(sx, sy, dx, bt, B, V)
    Hh = 0.5 * (H + H.conj().T)
    w, U = np.linalg.eigh(Hh)
    psi = U[:, :6].T / dx
    E, sd, _, _ = O.energies(psi.reshape(6, sy, sx), dx, bt, B, V)
    assert np.abs(E - w[:6]).max() < 1e-12 * np.abs(w[:6]).max()
    # sqrt of a difference of rounded numbers of size E^2: at most ~sqrt(eps) E
    assert sd.max() < 1e-6 * np.abs(w[:6]).max()
    # and a state that is not an eigenvector has sigma of order its spread
    mix = (psi[0] + psi[1]) / np.sqrt(2)
    E2, sd2, _, _ = O.energies(mix.reshape(1, sy, sx), dx, bt, B, V)
    assert sd2[0] == pytest.approx(0.5 * (w[1] - w[0]), rel=1e-9)
