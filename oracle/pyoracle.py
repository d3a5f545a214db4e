"""ctypes wrapper for the FP64 CPU oracle (oracle/itp2d_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg, as the checker. The product package never
imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

TRANSFORMS = ["FFT", "iFFT", "FFTx", "iFFTx", "FFTy", "iFFTy", "DST", "iDST", "DSTx",
              "iDSTx", "DSTy", "iDSTy", "DCT", "iDCT", "DCTx", "iDCTx", "DCTy", "iDCTy"]
R2R = {"RODFT10": 0, "RODFT01": 1, "REDFT10": 2, "REDFT01": 3}
PERIODIC, DIRICHLET = 0, 1
TESTS = {"none": 0, "onestep": 1, "abschange": 2, "relchange": 3, "absstdev": 4, "relstdev": 5}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _setup(_lib)
    return _lib


class OrcTest(C.Structure):
    _fields_ = [("kind", C.c_int), ("p0", C.c_double), ("p1", C.c_double)]


class OrcParams(C.Structure):
    _fields_ = [("sx", C.c_size_t), ("sy", C.c_size_t), ("lenx", C.c_double),
                ("boundary", C.c_int), ("N", C.c_size_t), ("n_wanted", C.c_size_t),
                ("ignore_lowest", C.c_size_t), ("halforder", C.c_int), ("B", C.c_double),
                ("eps_values", C.POINTER(C.c_double)), ("n_eps", C.c_size_t),
                ("eps_divisor", C.c_double), ("exhaust_eps", C.c_int), ("max_steps", C.c_int),
                ("min_time_step", C.c_double), ("recover", C.c_int), ("seed", C.c_uint64),
                ("timestep_test", OrcTest), ("final_test", OrcTest)]


_dp = C.POINTER(C.c_double)
_sz = C.c_size_t


def _setup(L):
    L.orc_fft1d.argtypes = [_dp, _sz, _sz, C.c_int]
    L.orc_r2r1d.argtypes = [_dp, _sz, _sz, C.c_int]
    L.orc_r2r1d_direct.argtypes = [_dp, _sz, _sz, C.c_int]
    L.orc_transform.argtypes = [_sz, _sz, _dp, C.c_int]
    L.orc_grid_tables.argtypes = [_sz, _sz, C.c_double] + [_dp] * 7
    L.orc_apply_exppotential.argtypes = [_sz, _sz, C.c_double, _dp, C.c_double, C.c_double,
                                         C.c_double, _dp, _sz]
    L.orc_apply_expkinetic.argtypes = [_sz, _sz, C.c_double, C.c_int, C.c_double, C.c_double,
                                       C.c_double, C.c_double, _dp, _sz]
    L.orc_apply_kinetic.argtypes = [_sz, _sz, C.c_double, C.c_int, C.c_double, _dp, _sz]
    L.orc_apply_hamiltonian.argtypes = [_sz, _sz, C.c_double, C.c_int, C.c_double, _dp, _dp, _sz]
    L.orc_apply_secondorder.argtypes = [_sz, _sz, C.c_double, C.c_int, C.c_double, _dp,
                                        C.c_double, C.c_double, C.c_int, _dp, _sz]
    L.orc_propagate.argtypes = [_sz, _sz, C.c_double, C.c_int, C.c_double, _dp, C.c_int,
                                C.c_double, _dp, _sz]
    L.orc_mps_coefficients.argtypes = [C.c_int, _dp]
    L.orc_eigh.argtypes = [_sz, _dp, _dp]
    L.orc_orthonormalize.argtypes = [_sz, _sz, C.c_double, _dp, _sz, C.POINTER(_sz), _dp]
    L.orc_overlap.argtypes = [_sz, _sz, C.c_double, _dp, _sz, _dp]
    L.orc_energies.argtypes = [_sz, _sz, C.c_double, C.c_int, C.c_double, _dp, _dp, _sz,
                               _dp, _dp, _dp, _dp]
    L.orc_how_orthonormal.argtypes = [_sz, _sz, C.c_double, _dp, _sz]
    L.orc_how_orthonormal.restype = C.c_double
    L.orc_init_states.argtypes = [C.POINTER(C.c_uint64), _sz, _sz, C.c_double, _sz, _dp]
    L.orc_system_create.argtypes = [C.POINTER(OrcParams), _dp, _dp]
    L.orc_system_create.restype = C.c_void_p
    for name in ["orc_system_destroy", "orc_system_step", "orc_system_run", "orc_system_finished",
                 "orc_system_error_flag", "orc_system_total_steps", "orc_system_history_len"]:
        getattr(L, name).argtypes = [C.c_void_p]
    L.orc_system_eps.argtypes = [C.c_void_p]
    L.orc_system_eps.restype = C.c_double
    L.orc_system_total_time.argtypes = [C.c_void_p]
    L.orc_system_total_time.restype = C.c_double
    L.orc_system_num_converged.argtypes = [C.c_void_p]
    L.orc_system_num_converged.restype = C.c_size_t
    L.orc_system_num_timestep_converged.argtypes = [C.c_void_p]
    L.orc_system_num_timestep_converged.restype = C.c_size_t
    L.orc_system_sorted.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(_sz)]
    L.orc_system_states.argtypes = [C.c_void_p, _dp]
    L.orc_system_converged.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.orc_pointwise.argtypes = [C.c_int, _sz, _sz, _dp, _dp, _dp]
    L.orc_dot_norm.argtypes = [_sz, C.c_double, _dp, _dp, _dp]
    L.orc_dot_norm.restype = C.c_double
    L.orc_set_num_threads.argtypes = [C.c_int]


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


def fft1d(x, sign=-1):
    x = _c128(x).copy()
    _check(lib().orc_fft1d(_ptr(x.view(np.float64)), x.size, 1, sign), "fft1d")
    return x


def r2r1d(x, kind, direct=False):
    x = _c128(x).copy()
    f = lib().orc_r2r1d_direct if direct else lib().orc_r2r1d
    _check(f(_ptr(x.view(np.float64)), x.size, 1, R2R[kind]), "r2r1d")
    return x


def transform(field, kind):
    f = _c128(field).copy()
    sy, sx = f.shape
    _check(lib().orc_transform(sx, sy, _ptr(f.view(np.float64)), TRANSFORMS.index(kind)), "transform")
    return f


def grid_tables(sx, sy, dx):
    out = [np.zeros(n) for n in (sx, sy, sx, sy, sx, sy, 6)]
    lib().orc_grid_tables(sx, sy, dx, *[_ptr(a) for a in out])
    keys = ["posx", "posy", "fft_kx", "fft_ky", "dsct_kx", "dsct_ky", "norms"]
    return dict(zip(keys, out))


def _vptr(V):
    return None if V is None else _ptr(np.ascontiguousarray(V, dtype=np.float64))


def apply_exppotential(states, dx, V, e, coeff=-0.5, prefactor=1.0):
    s = _c128(states).copy()
    n, sy, sx = s.shape
    Vc = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    _check(lib().orc_apply_exppotential(sx, sy, dx, _ptr(Vc), e, coeff, prefactor,
                                        _ptr(s.view(np.float64)), n), "exppotential")
    return s


def apply_expkinetic(states, dx, boundary, B, e, coeff=-1.0, prefactor=1.0):
    s = _c128(states).copy()
    n, sy, sx = s.shape
    _check(lib().orc_apply_expkinetic(sx, sy, dx, boundary, B, e, coeff, prefactor,
                                      _ptr(s.view(np.float64)), n), "expkinetic")
    return s


def apply_kinetic(states, dx, boundary, B):
    s = _c128(states).copy()
    n, sy, sx = s.shape
    _check(lib().orc_apply_kinetic(sx, sy, dx, boundary, B, _ptr(s.view(np.float64)), n), "kinetic")
    return s


def apply_hamiltonian(states, dx, boundary, B, V):
    s = _c128(states).copy()
    n, sy, sx = s.shape
    Vc = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    _check(lib().orc_apply_hamiltonian(sx, sy, dx, boundary, B, _ptr(Vc),
                                       _ptr(s.view(np.float64)), n), "hamiltonian")
    return s


def apply_secondorder(states, dx, boundary, B, V, e, prefactor=1.0, exponent=1):
    s = _c128(states).copy()
    n, sy, sx = s.shape
    Vc = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    _check(lib().orc_apply_secondorder(sx, sy, dx, boundary, B, _ptr(Vc), e, prefactor, exponent,
                                       _ptr(s.view(np.float64)), n), "secondorder")
    return s


def propagate(states, dx, boundary, B, V, halforder, eps):
    s = _c128(states).copy()
    n, sy, sx = s.shape
    Vc = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    _check(lib().orc_propagate(sx, sy, dx, boundary, B, _ptr(Vc), halforder, eps,
                               _ptr(s.view(np.float64)), n), "propagate")
    return s


def mps_coefficients(h):
    out = np.zeros(max(h, 1))
    lib().orc_mps_coefficients(h, _ptr(out))
    return out


def eigh(A):
    """Hermitian eigendecomposition like zheev('V','U'): ascending w, vectors in columns."""
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    a = np.asfortranarray(A).ravel(order="K").copy()  # column-major buffer
    a = np.ascontiguousarray(np.asfortranarray(A).T).ravel()  # [col][row] = column-major
    w = np.zeros(n)
    _check(lib().orc_eigh(n, _ptr(a.view(np.float64)), _ptr(w)), "eigh")
    V = a.reshape(n, n).T  # a[r*n + c] = element c of eigenvector r
    return w, np.ascontiguousarray(V)


def orthonormalize(states, dx):
    s = _c128(states).copy()
    n, sy, sx = s.shape
    bn = _sz(0)
    bl = C.c_double(0)
    rc = lib().orc_orthonormalize(sx, sy, dx, _ptr(s.view(np.float64)), n, C.byref(bn), C.byref(bl))
    return rc, s, (bn.value, bl.value)


def overlap(states, dx):
    s = _c128(states)
    n, sy, sx = s.shape
    S = np.zeros(n * n, dtype=np.complex128)
    _check(lib().orc_overlap(sx, sy, dx, _ptr(s.view(np.float64)), n, _ptr(S.view(np.float64))), "overlap")
    return S.reshape(n, n).T  # column-major -> S[i, j] = <psi_i|psi_j>


def energies(states, dx, boundary, B, V):
    s = _c128(states)
    n, sy, sx = s.shape
    E = np.zeros(n)
    sd = np.zeros(n)
    mc = np.zeros(2 * n)
    ms = np.zeros(2 * n)
    Vc = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    _check(lib().orc_energies(sx, sy, dx, boundary, B, _ptr(Vc), _ptr(s.view(np.float64)), n,
                              _ptr(E), _ptr(sd), _ptr(mc), _ptr(ms)), "energies")
    return E, sd, mc.view(np.complex128), ms.view(np.complex128)


POINTWISE = {"multiply": 0, "multiply_imaginary_shiftx": 1, "multiply_y": 2, "multiply_and_add": 3}


def pointwise(state, kind, values, addstate=None):
    """A State pointwise operation (state.hpp:180-211) on one [sy][sx] field; returns the new field."""
    s = _c128(state).copy()
    v = np.ascontiguousarray(values, dtype=np.float64)
    a = None if addstate is None else _c128(addstate)
    sy, sx = s.shape
    _check(lib().orc_pointwise(POINTWISE[kind], sx, sy, _ptr(s.view(np.float64)), _ptr(v),
                               None if a is None else _ptr(a.view(np.float64))), "pointwise")
    return s


def dot_norm(a, b, dx):
    """(State::dot(a, b), State::norm(a)) (state.hpp:213-240)."""
    x, y = _c128(a), _c128(b)
    out = np.zeros(2)
    nrm = lib().orc_dot_norm(x.size, dx, _ptr(x.view(np.float64)), _ptr(y.view(np.float64)), _ptr(out))
    return complex(out[0], out[1]), nrm


def how_orthonormal(states, dx):
    s = _c128(states)
    n, sy, sx = s.shape
    return lib().orc_how_orthonormal(sx, sy, dx, _ptr(s.view(np.float64)), n)


def init_states(seed, N, sx, sy, dx):
    st = C.c_uint64(seed)
    out = np.zeros((N, sy, sx), dtype=np.complex128)
    lib().orc_init_states(C.byref(st), sx, sy, dx, N, _ptr(out.view(np.float64)))
    return out


def make_params(cfg):
    """cfg: dict with the fields of OrcParams (tests given as ('relchange', 1e-8, 0))."""
    p = OrcParams()
    eps = np.ascontiguousarray(cfg.get("eps_values", []), dtype=np.float64)
    for k in ["sx", "sy", "lenx", "boundary", "N", "n_wanted", "halforder", "B", "eps_divisor",
              "max_steps", "min_time_step", "seed"]:
        setattr(p, k, cfg[k])
    p.ignore_lowest = cfg.get("ignore_lowest", 0)
    p.exhaust_eps = int(cfg.get("exhaust_eps", 0))
    p.recover = int(cfg.get("recover", 0))
    p.eps_values = _ptr(eps) if eps.size else None
    p.n_eps = eps.size
    for name in ["timestep_test", "final_test"]:
        t = cfg[name]
        setattr(p, name, OrcTest(TESTS[t[0]], t[1] if len(t) > 1 else 0.0, t[2] if len(t) > 2 else 0.0))
    return p, eps


class System:
    """The oracle's ITP driver (itpsystem.cpp restated)."""

    def __init__(self, cfg, V=None, init_states=None):
        self.cfg = cfg
        self._params, self._eps = make_params(cfg)
        self._V = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
        self._init = None if init_states is None else _c128(init_states)
        self.h = lib().orc_system_create(C.byref(self._params), _ptr(self._V),
                                         None if self._init is None else _ptr(self._init.view(np.float64)))
        if not self.h:
            raise RuntimeError("orc_system_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_system_destroy(self.h)
            self.h = None

    def step(self):
        lib().orc_system_step(self.h)

    def run(self):
        return lib().orc_system_run(self.h)

    @property
    def finished(self):
        return bool(lib().orc_system_finished(self.h))

    @property
    def error_flag(self):
        return bool(lib().orc_system_error_flag(self.h))

    @property
    def total_steps(self):
        return lib().orc_system_total_steps(self.h)

    @property
    def eps(self):
        return lib().orc_system_eps(self.h)

    @property
    def num_converged(self):
        return lib().orc_system_num_converged(self.h)

    @property
    def total_time(self):
        return lib().orc_system_total_time(self.h)

    def sorted(self):
        N = self.cfg["N"]
        E = np.zeros(N)
        sd = np.zeros(N)
        idx = (_sz * N)()
        lib().orc_system_sorted(self.h, _ptr(E), _ptr(sd), idx)
        return E, sd, np.array(list(idx))

    def converged(self):
        """Final-convergence flag of each sorted position."""
        f = np.zeros(int(self.cfg["N"]), dtype=np.int32)
        lib().orc_system_converged(self.h, f.ctypes.data_as(C.POINTER(C.c_int)))
        return f.astype(bool)

    def states(self):
        N, sy, sx = self.cfg["N"], self.cfg["sy"], self.cfg["sx"]
        out = np.zeros((N, sy, sx), dtype=np.complex128)
        lib().orc_system_states(self.h, _ptr(out.view(np.float64)))
        return out
