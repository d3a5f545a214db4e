"""ctypes wrapper for the CPU emulation of the sweep kernels (tests only)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libemu.so")
_lib = None
_dp = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
        _lib = C.CDLL(LIB)
        _lib.emu_transform.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _dp, C.c_int, C.c_int, C.c_int]
        _lib.emu_expkinetic.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_double, _dp, C.c_int]
        _lib.emu_split.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _dp, C.c_double,
                                   C.c_double, C.c_int, C.c_int, _dp, C.c_int, C.c_int]
        _lib.emu_hamiltonian.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _dp, _dp,
                                         C.c_int, _dp, _dp, C.c_int]
        _lib.emu_plan.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        _lib.emu_block_jacobi.argtypes = [C.c_int, _dp, _dp, C.c_int]
        _lib.emu_set_reg.argtypes = [C.c_int]
        _lib.emu_cluster_solve.argtypes = [C.c_int, _dp, C.POINTER(C.c_int), C.c_int, _dp]
        _lib.emu_cluster_pairs.argtypes = [C.c_int, C.c_int, _dp, C.c_double, C.c_int, C.POINTER(C.c_int), _dp]
        _lib.emu_exppotential.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _dp, C.c_double, C.c_double,
                                          C.c_double, _dp, C.c_int]
    return _lib


def set_reg(on):
    """Select the register-resident sweep engine (True, default) or the LDS-resident one."""
    lib().emu_set_reg(1 if on else 0)


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def plan(n):
    r = (C.c_int * 8)()
    nst = C.c_int(0)
    if lib().emu_plan(n, r, C.byref(nst)):
        return None
    return list(r)[: nst.value]


def transform(fields, dx, bt, kind, fp32=False):
    f = np.ascontiguousarray(fields, dtype=np.complex128).copy()
    n, sy, sx = f.shape
    assert lib().emu_transform(sx, sy, dx, bt, _p(f.view(np.float64)), n, kind, int(fp32)) == 0
    return f


def expkinetic(fields, dx, bt, B, e, coeff=-1.0, pref=1.0):
    f = np.ascontiguousarray(fields, dtype=np.complex128).copy()
    n, sy, sx = f.shape
    assert lib().emu_expkinetic(sx, sy, dx, bt, B, e, coeff, pref, _p(f.view(np.float64)), n) == 0
    return f


def exppotential(fields, dx, bt, V, e, coeff=-0.5, pref=1.0):
    f = np.ascontiguousarray(fields, dtype=np.complex128).copy()
    n, sy, sx = f.shape
    Vc = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    assert lib().emu_exppotential(sx, sy, dx, bt, _p(Vc), e, coeff, pref, _p(f.view(np.float64)), n) == 0
    return f


def split(fields, dx, bt, B, V, eps, prefactor=1.0, exponent=1, halforder=0, fp32=False):
    f = np.ascontiguousarray(fields, dtype=np.complex128).copy()
    n, sy, sx = f.shape
    Vc = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    assert lib().emu_split(sx, sy, dx, bt, B, _p(Vc), eps, prefactor, exponent, halforder,
                           _p(f.view(np.float64)), n, int(fp32)) == 0
    return f


def hamiltonian(fields, dx, bt, B, V, fp32=False):
    f = np.ascontiguousarray(fields, dtype=np.complex128).copy()
    n, sy, sx = f.shape
    Vc = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    E = np.zeros(n)
    sd = np.zeros(n)
    assert lib().emu_hamiltonian(sx, sy, dx, bt, B, _p(Vc), _p(f.view(np.float64)), n, _p(E), _p(sd),
                                 int(fp32)) == 0
    return f, E, sd


def block_jacobi(A, max_inner=8):
    """Emulated block Jacobi: (ascending eigenvalues, eigenvectors in columns, sweeps)."""
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    a = np.ascontiguousarray(A.T).ravel().copy()
    w = np.zeros(n)
    sweeps = lib().emu_block_jacobi(n, _p(a.view(np.float64)), _p(w), max_inner)
    return w, np.ascontiguousarray(a.reshape(n, n).T), sweeps


def cluster_solve(A, idx, X=None):
    """Emulated cluster solve: X (identity by default) with the eigenvectors of A[idx][:, idx] at idx."""
    A = np.ascontiguousarray(A, dtype=np.complex128)
    n = A.shape[0]
    X = np.eye(n, dtype=np.complex128) if X is None else np.ascontiguousarray(X, dtype=np.complex128).copy()
    ix = np.ascontiguousarray(idx, dtype=np.int32)
    rc = lib().emu_cluster_solve(n, _p(A.view(np.float64)), ix.ctypes.data_as(C.POINTER(C.c_int)), len(ix),
                                 _p(X.view(np.float64)))
    assert rc == 0
    return X


def cluster_pairs(A, N, tau, cap=8192):
    """Emulated strong-pair listing: (count, pairs (as listed, up to cap), largest ratio)."""
    A = np.ascontiguousarray(A, dtype=np.complex128)
    np_ = A.shape[0]
    pairs = np.zeros(2 * cap, dtype=np.int32)
    mx = np.zeros(1)
    cnt = lib().emu_cluster_pairs(np_, N, _p(A.view(np.float64)), tau, cap, pairs.ctypes.data_as(C.POINTER(C.c_int)),
                                  _p(mx))
    return cnt, pairs[: 2 * min(cnt, cap)].reshape(-1, 2), float(mx[0])
