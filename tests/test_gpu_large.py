"""The large configurations at full size, through the C ABI, by properties
that do not depend on the size: the 512x512 oscillator with N = 2000 states
(both precisions) and the 1024x1024 magnetic stadium with N = 4000 (FP32
fields, as the mixed path holds them). Three ITP iterations from seeded noise
must leave an orthonormal set (1e-11 in FP64; 2e-5 in FP32, the precision of
the fields) with finite energies; and the eigensolver takes the N x N problems
of these sizes, checked against LAPACK (numpy.linalg.eigvalsh). The reduced
forms of both configurations run against oracle goldens in
test_gpu_large_golden.py."""
import numpy as np
import pytest

import itp2d_amd as I
from itp2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def _setup(ctx, name, N=None):
    cfg, V = W.config(name)
    dx = cfg["lenx"] / cfg["sx"]
    g = I.Grid(ctx, cfg["sx"], cfg["sy"], dx, cfg["boundary"])
    pot = I.Potential(g, V)
    return cfg, g, pot, dx


def _orthonormality(S):
    return np.abs(S.overlap() - np.eye(S.N)).max()


@pytest.mark.parametrize("prec", [I.FP64, I.FP32])
def test_c3_full_size_iterations(ctx, prec):
    cfg, g, pot, dx = _setup(ctx, "c3_ho512")
    N = cfg["N"]
    say(f"c3 512x512 N={N} prec={prec}: 3 iterations")
    sp = I.MultiProductSplit(pot, cfg["halforder"], cfg["eps_values"][0], cfg["B"])
    S = I.StateSet(g, N, prec)
    S.init_noise(cfg["seed"], dx)
    for _ in range(3):
        sp.propagate(S)
        if prec == I.FP64:
            S.orthonormalize()
        else:
            assert S.orthonormalize_cholesky() is False
    # the orthonormalization's error grows with the condition number of the overlap it
    # inverts and with the overlap's own rounding (sums over M grid points, ~u sqrt(M)):
    # u sqrt(M) kappa for the canonical (eigendecomposition) route; FP32 fields hold 2e-5
    if prec == I.FP64:
        lam = S.overlap_eigenvalues()
        kappa = lam.max() / lam.min()
        bound = max(1e-12, 5e-15 * np.sqrt(cfg["sx"] * cfg["sy"]) * kappa)
    else:
        kappa = S.cholesky_condition()
        bound = max(2e-5, 1e-11 * kappa)
    err = _orthonormality(S)
    say(f"  condition {kappa:.3e}, orthonormality {err:.3e} (bound {bound:.1e})")
    assert err < bound
    H = I.Hamiltonian(pot, cfg["B"])
    dst = I.StateSet(g, N)
    E = H.rayleigh_quotients(S, dst)
    # variational: no Rayleigh quotient below the discrete ground level (~1, the oscillator's)
    assert np.isfinite(E).all() and E.min() > 0.99
    for o in (dst, S, H, sp):
        o.close()


def test_c5_full_size_orthonormalization(ctx):
    cfg, g, pot, dx = _setup(ctx, "c5_stadium1024")
    N = cfg["N"]
    say(f"c5 1024x1024 N={N} fp32: propagate and CholeskyQR")
    sp = I.MultiProductSplit(pot, cfg["halforder"], cfg["eps_values"][0], cfg["B"])
    assert sp.halforder == 3
    S = I.StateSet(g, N, I.FP32)
    S.init_noise(cfg["seed"], dx)
    assert S.orthonormalize_cholesky() is False
    assert _orthonormality(S) < 2e-5
    sp.propagate(S)
    # one propagation of noise at eps = 0.02 spans exp(-2 eps E) far below FP32 (and FP64)
    # resolution: CholeskyQR cannot factor it, and the eigendecomposition route reports the
    # non-positive eigenvalue the reference's recovery acts on
    with pytest.raises(I.NonPositiveEigenvalue):
        S.orthonormalize_cholesky()
    sp.close()
    S.close()


def _overlap_like(n, seed):
    """Hermitian matrices like the overlap matrices of an ITP run: a decaying
    spectrum with coupled states (and a few near-degenerate pairs)."""
    rng = np.random.default_rng(seed)
    d = np.exp(-0.02 * np.sort(rng.uniform(1, 200, n)))
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = np.diag(d) + 1e-4 * (X + X.conj().T) * np.sqrt(np.outer(d, d))
    return 0.5 * (A + A.conj().T)


@pytest.mark.parametrize("n", [2000, 4000])
def test_eigh_large(ctx, n):
    say(f"eigh n={n}")
    A = _overlap_like(n, n)
    w, V = I.eigh(ctx, A)
    wl = np.linalg.eigvalsh(A)
    assert np.abs(w - wl).max() < 1e-13 * np.abs(wl).max()
    r = A @ V[:, :64] - V[:, :64] * w[:64]
    assert np.abs(r).max() < 1e-12 * np.abs(A).max()


def _hermitian(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "overlap":
        return _overlap_like(n, seed)
    if kind == "diag":  # already diagonal: every reflector is the identity
        return np.diag(rng.uniform(-1, 1, n)).astype(np.complex128)
    if kind == "phase":  # tridiagonal with complex couplings: zero x, complex alpha
        A = np.diag(rng.uniform(-1, 1, n)).astype(np.complex128)
        for k in range(n - 1):
            A[k + 1, k] = np.exp(1j * rng.uniform(0, 2 * np.pi)) * rng.uniform(0.1, 1)
            A[k, k + 1] = np.conj(A[k + 1, k])
        return A
    if kind == "block":  # two uncoupled blocks and a repeated eigenvalue
        X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        A = X + X.conj().T
        h = n // 2
        A[:h, h:] = 0
        A[h:, :h] = 0
        A[h:, h:] = A[:n - h, :n - h] if n % 2 == 0 else A[h:, h:]
        return A
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return 0.5 * (X + X.conj().T)


@pytest.mark.parametrize("n,kind", [(1, "rand"), (2, "rand"), (3, "phase"), (17, "rand"), (64, "diag"), (130, "block"),
                                    (300, "phase"), (500, "overlap"), (2000, "overlap"), (4000, "rand")])
def test_eigvalsh_tridiagonal(ctx, n, kind):
    """Eigenvalues by Householder tridiagonalization and bisection (the
    eigenvalue check of large overlaps) against LAPACK's eigvalsh."""
    say(f"eigvalsh n={n} {kind}")
    A = _hermitian(kind, n, n + 7)
    w = I.eigvalsh(ctx, A)
    wl = np.linalg.eigvalsh(A)
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - wl).max() < 1e-13 * np.abs(wl).max()
