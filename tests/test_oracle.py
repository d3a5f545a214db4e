"""Pins the FP64 CPU oracle before it is trusted as the checker.

Sources of truth: scipy/numpy golden vectors (tests/golden/transforms.npz),
the FFTW definitions evaluated directly, the reference's own known-answer
tests restated (test_transformer.cpp, test_eigensolver.cpp, test_stateset.cpp,
test_itp.cpp), and closed forms (MPS coefficient sums, free evolution, the
dense discrete Hamiltonian on a tiny grid).
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from itp2d_amd import workloads as W

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "transforms.npz"))
EPS = np.finfo(float).eps


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("n", [8, 16, 40, 64, 192, 256])
def test_1d_transforms_vs_scipy(n):
    x = GOLD[f"x{n}"]
    assert rel(O.fft1d(x, -1), GOLD[f"fft{n}"]) < 1e-14
    assert rel(O.fft1d(x, +1), GOLD[f"ifft{n}"]) < 1e-14
    for kind in ["RODFT10", "RODFT01", "REDFT10", "REDFT01"]:
        assert rel(O.r2r1d(x, kind), GOLD[f"{kind}{n}"]) < 1e-14, kind


@pytest.mark.parametrize("n", [8, 16, 40])
def test_1d_r2r_vs_definition(n):
    x = GOLD[f"x{n}"]
    for kind in ["RODFT10", "RODFT01", "REDFT10", "REDFT01"]:
        assert rel(O.r2r1d(x, kind), O.r2r1d(x, kind, direct=True)) < 1e-13, kind


@pytest.mark.parametrize("key", ["16x16", "24x40"])
def test_2d_transforms_vs_scipy(key):
    f = GOLD[f"field{key}"]
    for kind in ["FFT", "iFFT", "FFTx", "FFTy", "DST", "iDST", "DSTx", "iDSTy", "DCT", "iDCTx"]:
        assert rel(O.transform(f, kind), GOLD[f"{kind}{key}"]) < 1e-14, kind


# test_transformer.cpp:25-149 (16x16, dx = 0.5)
SX = SY = 16
DX = 0.5


def _field(fn):
    g = O.grid_tables(SX, SY, DX)
    X, Y = np.meshgrid(g["posx"], g["posy"])
    return fn(X, Y)


def test_reference_kvalues():
    g = O.grid_tables(SX, SY, DX)
    L = SX * DX
    for x in range(SX):
        assert g["fft_kx"][x] == (2 * np.pi / L) * ((x if x < SX // 2 else x - SX))
        assert g["dsct_kx"][x] == pytest.approx((np.pi / L) * (x + 1), rel=1e-15)


@pytest.mark.parametrize("kind", ["FFT", "FFTx", "FFTy", "DST", "DSTx", "DSTy", "DCT", "DCTx", "DCTy"])
def test_reference_inverse_roundtrip(kind):
    A = _field(lambda X, Y: np.exp(-(X * X + Y * Y)) * (1 - 1j))
    inv = O.TRANSFORMS[O.TRANSFORMS.index(kind) + 1]
    T = O.transform(O.transform(A, kind), inv)
    g = O.grid_tables(SX, SY, DX)
    norm = {"FFT": 0, "FFTx": 1, "FFTy": 2, "DST": 3, "DSTx": 4, "DSTy": 5, "DCT": 3, "DCTx": 4, "DCTy": 5}[kind]
    T = T * g["norms"][norm]
    assert np.sqrt(np.mean(np.abs(A - T) ** 2)) < EPS


def test_reference_dirac_delta_idst_and_sine_blob_dst():
    blob = _field(lambda X, Y: 4 * np.cos(X * np.pi / (SX * DX)) * np.cos(Y * np.pi / (SY * DX)) * (1 - 1j))
    T = np.zeros((SY, SX), complex)
    T[0, 0] = 1 - 1j
    assert np.sqrt(np.mean(np.abs(O.transform(T, "iDST") - blob) ** 2)) < 10 * EPS
    g = O.grid_tables(SX, SY, DX)
    D = O.transform(blob, "DST") * g["norms"][3]
    assert np.sqrt(np.mean(np.abs(D - T) ** 2)) < 10 * EPS


def test_reference_dirac_delta_ifft_and_unit_wave_fft():
    def wave(X, Y):
        px = 2 * np.pi * X / (SX * DX) - np.pi / SX
        py = 2 * np.pi * Y / (SY * DX) - np.pi / SY
        return np.exp(1j * px) * np.exp(1j * py)
    W_ = _field(wave)
    T = np.zeros((SY, SX), complex)
    T[1, 1] = 1
    assert np.sqrt(np.mean(np.abs(O.transform(T, "iFFT") - W_) ** 2)) < 10 * EPS
    g = O.grid_tables(SX, SY, DX)
    assert np.sqrt(np.mean(np.abs(O.transform(W_, "FFT") * g["norms"][0] - T) ** 2)) < 10 * EPS


# a complex Hermitian 2x2 with eigenvalues 0 and 2 (the reference's own real 2x2 and
# 8x8 known answers, test_eigensolver.cpp:25-74, are in test_oracle_pins.py)
def test_eigensolver_2x2_complex():
    A = np.array([[1, 1j], [-1j, 1]])
    w, V = O.eigh(A)
    assert w == pytest.approx([0, 2], abs=1e-15)
    for k in range(2):
        assert np.abs(A @ V[:, k] - w[k] * V[:, k]).max() < 1e-15


@pytest.mark.parametrize("n", [2, 8, 33])
def test_eigensolver_vs_lapack(n):
    A = GOLD[f"herm{n}"]
    w, V = O.eigh(A)
    assert np.abs(w - GOLD[f"eig{n}"]).max() < 300 * EPS * np.abs(w).max()
    # spectral reconstruction and orthonormal eigenvectors
    assert np.abs(V @ np.diag(w) @ V.conj().T - A).max() < 300 * EPS * np.abs(A).max()
    assert np.abs(V.conj().T @ V - np.eye(n)).max() < 100 * EPS


def test_eigensolver_degenerate():
    Q, _ = np.linalg.qr(np.random.default_rng(3).standard_normal((12, 12)) + 1j * np.random.default_rng(4).standard_normal((12, 12)))
    d = np.array([1, 1, 1, 2, 2, 3, 3, 3, 3, 5, 8, 8], float)
    A = Q @ np.diag(d) @ Q.conj().T
    w, V = O.eigh(A)
    assert np.abs(w - d).max() < 1e-13
    assert np.abs(V @ np.diag(w) @ V.conj().T - A).max() < 1e-13


# test_stateset.cpp:26-42: 8 random states on 16x16, orthonormal to 16 eps
def test_reference_stateset_orthonormalize():
    rng = np.random.default_rng(8)
    s = rng.standard_normal((8, 16, 16)) + 1j * rng.standard_normal((8, 16, 16))
    rc, out, _ = O.orthonormalize(s, 0.5)
    assert rc == 0
    assert O.how_orthonormal(out, 0.5) < 16 * EPS


def test_orthonormalize_zero_state_fails():
    """A zero state gives an exactly zero overlap eigenvalue: NonPositiveEigenvalue(0, 0)
    (stateset.cpp:164-168), and the states are left untouched."""
    rng = np.random.default_rng(9)
    s = rng.standard_normal((4, 16, 16)) + 1j * rng.standard_normal((4, 16, 16))
    s[2] = 0
    rc, out, (n, lam) = O.orthonormalize(s, 0.5)
    assert rc == 2
    assert n == 0 and lam == 0.0
    assert np.array_equal(out, s)


def test_mps_coefficients():
    # multiproductsplit.cpp:49-58: h=1 -> 1; h=2 -> -1/3, 4/3; h=3 -> 1/24, -16/15, 81/40
    assert O.mps_coefficients(1) == pytest.approx([1.0])
    assert O.mps_coefficients(2) == pytest.approx([-1 / 3, 4 / 3])
    assert O.mps_coefficients(3) == pytest.approx([1 / 24, -16 / 15, 81 / 40])
    for h in range(1, 7):
        assert O.mps_coefficients(h).sum() == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("bt", [O.PERIODIC, O.DIRICHLET])
def test_free_evolution_is_exact(bt):
    """V = 0, B = 0: propagation multiplies each transform mode by exp(-eps k^2 / 2)."""
    sx, sy, dx, eps = 16, 24, 0.3, 0.07
    g = O.grid_tables(sx, sy, dx)
    kx = g["fft_kx"] if bt == O.PERIODIC else g["dsct_kx"]
    ky = g["fft_ky"] if bt == O.PERIODIC else g["dsct_ky"]
    KX, KY = np.meshgrid(kx, ky)
    rng = np.random.default_rng(5)
    c = rng.standard_normal((sy, sx)) + 1j * rng.standard_normal((sy, sx))
    inv = "iFFT" if bt == O.PERIODIC else "iDST"
    psi = O.transform(c, inv)
    out = O.propagate(psi[None], dx, bt, 0.0, None, 3, eps)[0]
    expect = O.transform(c * np.exp(-0.5 * eps * (KX ** 2 + KY ** 2)), inv)
    assert rel(out, expect) < 1e-13


def _dense_h(sx, sy, dx, bt, B, V):
    n = sx * sy
    H = np.zeros((n, n), complex)
    for k in range(n):
        e = np.zeros(n, complex)
        e[k] = 1
        H[:, k] = O.apply_hamiltonian(e.reshape(1, sy, sx), dx, bt, B, V)[0].ravel()
    return H


@pytest.mark.parametrize("bt,B", [(O.PERIODIC, 0.0), (O.DIRICHLET, 0.0), (O.PERIODIC, 0.8), (O.DIRICHLET, 0.8)])
def test_itp_fixed_point_is_discrete_spectrum(bt, B):
    """At convergence the ITP energies are the lowest eigenvalues of the
    discrete Hamiltonian the restated operators define (to the splitting error)."""
    sx = sy = 12
    L = 8.0
    dx = L / sx
    V = W.harmonic(sx, sy, dx)
    H = _dense_h(sx, sy, dx, bt, B, V)
    ev = np.sort(np.linalg.eigvals(H).real)
    cfg = dict(sx=sx, sy=sy, lenx=L, boundary=bt, N=8, n_wanted=5, halforder=4, B=B, eps_values=[0.05],
               eps_divisor=5.0, max_steps=3000, min_time_step=1e-9, seed=11,
               timestep_test=("relchange", 1e-13), final_test=("relchange", 1e-13))
    s = O.System(cfg, V)
    s.run()
    E, _, _ = s.sorted()
    assert not s.error_flag
    assert np.abs(E[:5] - ev[:5]).max() < 1e-6


HO = W.ho_levels(8)


def _run(sx, L, bt, B, N, nw, eps, tt, ft, V):
    cfg = dict(sx=sx, sy=sx, lenx=L, boundary=bt, N=N, n_wanted=nw, halforder=5, B=B, eps_values=[eps],
               eps_divisor=5.0, max_steps=500, min_time_step=1e-9, seed=0x20120131, timestep_test=tt, final_test=ft)
    s = O.System(cfg, V)
    s.run()
    assert not s.error_flag
    return s.sorted()[0][:nw]


def _fock_darwin(N, B):
    U = np.sqrt(1 + 0.25 * B * B)
    return np.sort([(2 * n + abs(m) + 1) * U - 0.5 * m * B for n in range(N) for m in range(-N, N + 1)])[:N]


# test_itp.cpp:56-165: harmonic oscillator, periodic and Dirichlet, 40x40, L=12
@pytest.mark.parametrize("bt", [O.PERIODIC, O.DIRICHLET])
def test_reference_harmonic_oscillator(bt):
    V = W.harmonic(40, 40, 12.0 / 40)
    E = _run(40, 12.0, bt, 0.0, 14, 8, 1.0, ("relstdev", 1e-4, 1e-5), ("relstdev", 1e-4), V)
    assert np.abs(E - HO).max() < 1e-4


# test_itp.cpp:266-412: Fock-Darwin, periodic B=10 L=7 and Dirichlet B=1 L=10
@pytest.mark.parametrize("bt,B,L", [(O.PERIODIC, 10.0, 7.0), (O.DIRICHLET, 1.0, 10.0)])
def test_reference_fock_darwin(bt, B, L):
    V = W.harmonic(40, 40, L / 40)
    E = _run(40, L, bt, B, 14, 8, 1.0, ("relstdev", 1e-3, 1e-4), ("relstdev", 1e-3), V)
    assert np.abs(E - _fock_darwin(8, B)).max() < 1e-3


# test_itp.cpp:414-443: magnetic particle in a pi-box, 19 known-good energies
def test_reference_magnetic_box():
    known = np.array([1.077814, 2.134673, 3.089442, 3.551456, 5.075289, 5.499240, 5.729555, 7.042548])
    E = _run(64, np.pi, O.DIRICHLET, 1.0, 16, 8, 0.3, ("abschange", 1e-3), ("onestep",), None)
    assert np.abs(E - known).max() < 2e-3


def test_init_states_normalised_and_seeded():
    a = O.init_states(7, 3, 8, 6, 0.25)
    b = O.init_states(7, 3, 8, 6, 0.25)
    assert np.array_equal(a, b)
    for s in a:
        assert np.sqrt(np.sum(np.abs(s) ** 2)) * 0.25 == pytest.approx(1.0, abs=1e-14)
