"""The gfx950 sweep kernels' own pass code (itp2d_amd/csrc/line_ops.hpp) and
sweep sequencing (engine.hpp), run on the CPU one emulated workgroup at a
time, against the oracle. Checks indexing, Makhoul reorderings, the magnetic
Dirichlet split and the epilogues without a GPU."""
import numpy as np
import pytest

from oracle import pyoracle as O
from itp2d_amd import workloads as W
from tests.emu import pyemu as E


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(params=["reg", "lds"])
def engine(request):
    """Run a test on the register-resident sweep engine and on the LDS-resident one."""
    E.set_reg(request.param == "reg")
    yield request.param
    E.set_reg(True)


def test_radix_plans_even():
    for n in [8, 16, 40, 64, 128, 192, 256, 512, 1024]:
        p = E.plan(n)
        assert p is not None and len(p) % 2 == 0 and int(np.prod(p)) == n and max(p) <= 8


@pytest.mark.parametrize("sx,sy", [(16, 16), (64, 64), (40, 40), (192, 64), (64, 128), (32, 256), (256, 32),
                                   (512, 64), (1024, 32)])
def test_transforms(sx, sy, engine):
    rng = np.random.default_rng(sx * sy)
    f = rng.standard_normal((2, sy, sx)) + 1j * rng.standard_normal((2, sy, sx))
    for k, name in enumerate(O.TRANSFORMS):
        a = E.transform(f, 0.1, 0, k)
        b = np.stack([O.transform(x, name) for x in f])
        assert rel(a, b) < 1e-14, name


def _setup(sx, sy, L=12.0):
    dx = L / sx
    V = W.harmonic(sx, sy, dx) + 0.3 * np.exp(-((W.positions(sx, sy, dx)[0][None, :] - 1) ** 2))
    rng = np.random.default_rng(sx + 7 * sy)
    f = rng.standard_normal((3, sy, sx)) + 1j * rng.standard_normal((3, sy, sx))
    f /= np.sqrt((np.abs(f) ** 2).sum(axis=(1, 2), keepdims=True)) * dx
    return dx, V, f


BRANCHES = [(0, 0.0), (1, 0.0), (0, 1.3), (1, 1.3)]


@pytest.mark.parametrize("bt,B", BRANCHES)
@pytest.mark.parametrize("sx,sy", [(64, 64), (40, 40), (64, 128), (256, 32), (192, 32), (32, 192)])
def test_operators(bt, B, sx, sy, engine):
    dx, V, f = _setup(sx, sy)
    assert rel(E.expkinetic(f, dx, bt, B, 0.07, -1.0, 0.9), O.apply_expkinetic(f, dx, bt, B, 0.07, -1.0, 0.9)) < 1e-14
    assert rel(E.exppotential(f, dx, bt, V, 0.07, -0.5, 0.8), O.apply_exppotential(f, dx, V, 0.07, -0.5, 0.8)) < 1e-14
    assert rel(E.exppotential(f, dx, bt, None, 0.07, -0.5, 0.8), O.apply_exppotential(f, dx, None, 0.07, -0.5, 0.8)) < 1e-14
    for ex in (1, 2, 3):
        a = E.split(f, dx, bt, B, V, 0.05, 0.7, ex)
        b = O.apply_secondorder(f, dx, bt, B, V, 0.05, 0.7, ex)
        assert rel(a, b) < 1e-14
    assert rel(E.split(f, dx, bt, B, None, 0.05, 1.0, 1), O.apply_secondorder(f, dx, bt, B, None, 0.05, 1.0, 1)) < 1e-14
    for h in (1, 2, 3):
        assert rel(E.split(f, dx, bt, B, V, 0.05, halforder=h), O.propagate(f, dx, bt, B, V, h, 0.05)) < 1e-14


@pytest.mark.parametrize("bt,B", BRANCHES)
@pytest.mark.parametrize("sx,sy", [(64, 64), (256, 32)])
def test_hamiltonian_and_energies(bt, B, sx, sy, engine):
    dx, V, f = _setup(sx, sy)
    hf, Ee, se = E.hamiltonian(f, dx, bt, B, V)
    assert rel(hf, O.apply_hamiltonian(f, dx, bt, B, V)) < 1e-14
    Eo, so, _, _ = O.energies(f, dx, bt, B, V)
    assert np.abs(Ee - Eo).max() < 1e-12 * np.abs(Eo).max()
    assert np.abs(se - so).max() < 1e-10 * np.abs(so).max()
    _, En, _ = E.hamiltonian(f, dx, bt, B, None)
    En_o, _, _, _ = O.energies(f, dx, bt, B, None)
    assert np.abs(En - En_o).max() < 1e-12 * np.abs(En_o).max()


@pytest.mark.parametrize("bt,B", BRANCHES)
def test_fp32_propagation_tolerance(bt, B, engine):
    """complex-FP32 fields: one MPS step within FP32 rounding of the FP64 result."""
    dx, V, f = _setup(64, 64)
    a = E.split(f, dx, bt, B, V, 0.05, halforder=2, fp32=True)
    b = O.propagate(f, dx, bt, B, V, 2, 0.05)
    assert rel(a, b) < 2e-6


@pytest.mark.parametrize("n", [2, 7, 33, 64, 100])
def test_block_jacobi_emulated(n):
    """The block Jacobi phases (block_jacobi.hpp) run on the CPU: eigenpairs of random Hermitian matrices."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.conj().T
    w, V, sweeps = E.block_jacobi(A)
    assert sweeps > 0
    scale = np.abs(np.linalg.eigvalsh(A)).max()
    assert np.abs(w - np.linalg.eigvalsh(A)).max() < 1e-13 * n * scale
    assert np.abs(V @ np.diag(w) @ V.conj().T - A).max() < 1e-13 * n * scale
    assert np.abs(V.conj().T @ V - np.eye(n)).max() < 1e-14 * n


@pytest.mark.parametrize("k", [2, 7, 12, 48, 63, 64])
def test_cluster_solve_emulated(k):
    """Cluster start (cluster_eig.hpp): the exact eigenvectors of a principal
    sub-matrix on scattered indices, embedded in the identity, diagonalise that
    block (eigenvalues of the block as numpy's to rounding), are unitary, and
    leave every other row and column alone."""
    rng = np.random.default_rng(k)
    n = 64
    d = np.exp(-0.05 * np.sort(rng.uniform(1, 40, n)))
    Eps = 1e-9 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = np.diag(d) + Eps + Eps.conj().T
    idx = np.sort(rng.choice(n, size=k, replace=False))
    # a strongly coupled, nearly degenerate cluster
    B = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
    A[np.ix_(idx, idx)] += 1e-3 * (B + B.conj().T) + np.diag(np.full(k, 0.5 - d[idx]))
    X = E.cluster_solve(A, idx)
    U = X[np.ix_(idx, idx)]
    assert np.abs(U.conj().T @ U - np.eye(k)).max() < 1e-14 * k
    Acc = A[np.ix_(idx, idx)]
    D = U.conj().T @ Acc @ U
    off = D - np.diag(np.diag(D))
    assert np.abs(off).max() < 1e-15 * k * np.abs(Acc).max()
    assert np.abs(np.sort(np.diag(D).real) - np.linalg.eigvalsh(Acc)).max() < 1e-15 * k * np.abs(Acc).max()
    rest = np.setdiff1d(np.arange(n), idx)
    assert np.array_equal(X[np.ix_(rest, np.arange(n))], np.eye(n)[rest])
    assert np.array_equal(X[np.ix_(idx, rest)], np.zeros((k, n - k)))


def test_cluster_solve_updates_columns_emulated():
    """With a general X the cluster rotation multiplies X's cluster columns:
    X[:, idx] U, every other column untouched."""
    rng = np.random.default_rng(9)
    n, k = 80, 9
    d = np.sort(rng.uniform(0.1, 1.0, n))
    Eps = 1e-3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = np.diag(d) + Eps + Eps.conj().T
    idx = np.sort(rng.choice(n, size=k, replace=False))
    X0 = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    U = E.cluster_solve(A, idx)[np.ix_(idx, idx)]
    X = E.cluster_solve(A, idx, X0)
    assert np.abs(X[:, idx] - X0[:, idx] @ U).max() < 1e-14
    rest = np.setdiff1d(np.arange(n), idx)
    assert np.array_equal(X[:, rest], X0[:, rest])


def test_cluster_pairs_emulated():
    """Strong-pair listing: exactly the pairs i < j < N with |a_ij| > tau |a_ii - a_jj|
    (brute force over the matrix), padded rows and columns ignored."""
    rng = np.random.default_rng(5)
    np_, N, tau = 96, 90, 0.1
    d = np.sort(rng.uniform(0, 1, np_))
    Eps = 2e-3 * (rng.standard_normal((np_, np_)) + 1j * rng.standard_normal((np_, np_)))
    A = np.diag(d) + Eps + Eps.conj().T
    cnt, pairs, mx = E.cluster_pairs(A, N, tau)
    want = set()
    rho_max = 0.0
    for i in range(N):
        for j in range(i + 1, N):
            rho = abs(A[i, j]) / abs(A[i, i].real - A[j, j].real)
            rho_max = max(rho_max, rho)
            if rho > tau:
                want.add((i, j))
    assert cnt == len(want) and set(map(tuple, pairs.tolist())) == want
    assert abs(mx - rho_max) <= 1e-15 * rho_max
    assert cnt > 10


def test_block_jacobi_near_diagonal_converges_fast():
    rng = np.random.default_rng(2)
    n = 150
    d = np.exp(-0.1 * np.sort(rng.uniform(1, 30, n)))
    Eps = 1e-7 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = np.diag(d) + Eps + Eps.conj().T
    w, V, sweeps = E.block_jacobi(A)
    assert sweeps <= 3
    assert np.abs(w - np.linalg.eigvalsh(A)).max() < 5e-14
