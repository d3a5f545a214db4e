"""Cholesky orthonormalization (itp_orthonormalize_cholesky, the CholeskyQR
step of the mixed path) and the FP64 Rayleigh quotients of a set
(itp_rayleigh_quotients), through the C ABI, against numpy: the factor R of
S = R^H R from numpy.linalg.cholesky, Psi R^-1 to 1e-10 on FP64 sets,
orthonormal to FP32 accuracy on FP32 sets with the same span, the fallback to
the eigendecomposition route on nearly dependent and on zero states, and the
Rayleigh quotients against the oracle's energies."""
import numpy as np
import pytest

import itp2d_amd as I
from itp2d_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def rand_states(n, sx, sy, dx, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n, sy, sx)) + 1j * rng.standard_normal((n, sy, sx))
    return f / (np.sqrt((np.abs(f) ** 2).sum(axis=(1, 2), keepdims=True)) * dx)


def numpy_cholqr(f, dx):
    A = f.reshape(f.shape[0], -1)
    S = (A.conj() @ A.T) * dx * dx          # S_ij = <psi_i|psi_j>
    R = np.linalg.cholesky(S).conj().T      # S = L L^H = R^H R with R = L^H upper
    Q = np.linalg.solve(R.T, A)             # rows: (Psi R^-1)_j = sum_i Psi_i (R^-1)_ij
    return Q.reshape(f.shape)


@pytest.mark.parametrize("N,sx", [(5, 32), (40, 64), (130, 32), (300, 64)])
def test_cholesky_fp64_matches_numpy(ctx, N, sx):
    say(f"cholesky fp64 N={N} sx={sx}")
    dx = 10.0 / sx
    g = I.Grid(ctx, sx, sx, dx, I.PERIODIC)
    f = rand_states(N, sx, sx, dx, N)
    # a scaled, correlated set, like propagated states
    f = f * np.exp(-5.0 * np.arange(N) / N)[:, None, None] + 0.3 * f[0][None]
    S = I.StateSet(g, N)
    S.upload(f)
    assert S.orthonormalize_cholesky() is False
    out = S.download()
    ref = numpy_cholqr(f, dx)
    assert np.abs(out - ref).max() < 1e-10 * np.abs(ref).max()
    # CholeskyQR loses orthogonality in proportion to the condition number of
    # S: no worse than numpy's own CholeskyQR of the same set
    def loss(q):
        A = q.reshape(N, -1)
        return np.abs((A.conj() @ A.T) * dx * dx - np.eye(N)).max()
    assert loss(out) < 10 * loss(ref) + 1e-13


@pytest.mark.parametrize("N,sx", [(40, 64), (500, 64), (70, 20)])
def test_cholesky_fp32_orthonormal_same_span(ctx, N, sx):
    # 20 x 20: the rotation's last tile of 64 grid points is partial (M = 400)
    say(f"cholesky fp32 N={N}")
    dx = 12.0 / sx
    g = I.Grid(ctx, sx, sx, dx, I.PERIODIC)
    f = rand_states(N, sx, sx, dx, 7 + N)
    f = f * np.exp(-0.01 * np.arange(N))[:, None, None]
    S = I.StateSet(g, N, I.FP32)
    S.upload(f)
    assert S.orthonormalize_cholesky() is False
    out = S.download()
    A = out.reshape(N, -1)
    assert np.abs((A.conj() @ A.T) * dx * dx - np.eye(N)).max() < 2e-5
    B = f.reshape(N, -1)
    proj = (A.conj() @ B.T) * dx * dx
    assert np.abs(B - proj.T @ A).max() < 2e-5 * np.abs(B).max()
    # the basis is the Cholesky one: state j lies in the span of inputs 0..j
    ref = numpy_cholqr(f, dx)
    assert np.abs(out - ref).max() < 1e-4 * np.abs(ref).max()


def test_cholesky_falls_back_on_nearly_dependent_states(ctx):
    say("cholesky fallback, nearly dependent")
    sx, N = 32, 12
    dx = 8.0 / sx
    g = I.Grid(ctx, sx, sx, dx, I.DIRICHLET)
    f = rand_states(N, sx, sx, dx, 99)
    extra = rand_states(1, sx, sx, dx, 100)[0]
    f[5] = f[2] + 1e-3 * extra  # pivot ratio ~1e-6, below a threshold of 1e-4
    S = I.StateSet(g, N)
    S.upload(f)
    try:
        ctx.set_cholesky_tol(1e-4)
        assert S.orthonormalize_cholesky() is True
    finally:
        ctx.set_cholesky_tol(0.0)
    A = S.download().reshape(N, -1)
    assert np.abs((A.conj() @ A.T) * dx * dx - np.eye(N)).max() < 1e-8
    # the fallback is the eigendecomposition route: the same result as orthonormalize()
    S2 = I.StateSet(g, N)
    S2.upload(f)
    S2.orthonormalize()
    assert np.array_equal(S2.download().reshape(N, -1), A)


def test_cholesky_zero_state_raises_like_orthonormalize(ctx):
    say("cholesky fallback, zero state")
    sx = 16
    dx = 0.5
    g = I.Grid(ctx, sx, sx, dx, I.PERIODIC)
    f = rand_states(4, sx, sx, dx, 4)
    f[2] = 0
    S = I.StateSet(g, 4)
    S.upload(f)
    with pytest.raises(I.NonPositiveEigenvalue) as e:
        S.orthonormalize_cholesky()
    assert e.value.n == 0 and e.value.value == 0.0
    assert np.array_equal(S.download(), f)


@pytest.mark.parametrize("bt,B", [(I.PERIODIC, 0.0), (I.DIRICHLET, 0.0), (I.PERIODIC, 0.7), (I.DIRICHLET, 0.7)])
@pytest.mark.parametrize("prec", [I.FP64, I.FP32])
def test_rayleigh_quotients(ctx, bt, B, prec):
    say(f"rayleigh quotients bt={bt} B={B} prec={prec}")
    sx, N = 64, 6
    L = 12.0
    dx = L / sx
    V = W.harmonic(sx, sx, dx)
    g = I.Grid(ctx, sx, sx, dx, bt)
    pot = I.Potential(g, V)
    H = I.Hamiltonian(pot, B)
    f = O.propagate(rand_states(N, sx, sx, dx, 3), dx, bt, B, V, 2, 0.3)  # smooth states
    scale = np.linspace(0.5, 2.0, N)[:, None, None]
    src = I.StateSet(g, N, prec)
    src.upload(f * scale)
    dst = I.StateSet(g, N)
    E = H.rayleigh_quotients(src, dst)
    if prec == I.FP32:
        f = (f * scale).astype(np.complex64).astype(np.complex128) / scale
    Eo = O.energies(f / (np.sqrt((np.abs(f) ** 2).sum(axis=(1, 2), keepdims=True)) * dx), dx, bt, B, V)[0]
    # FP32 fields: the quotient is evaluated in FP64 from the same rounded fields the oracle
    # sees, so only the FP64 evaluation order differs (1e-12); against the FP32 fields'
    # rounding the error is first order in (H - E) psi, small for these smooth states (1e-8)
    assert np.abs(E - Eo).max() < 1e-12 * np.abs(Eo).max()
    assert np.array_equal(dst.download(), src.download())  # dst holds src's fields
