"""ctypes binding of libitp2d_amd.so (C ABI: include/itp2d.h, include/itp2d_system.h).

Host arrays are numpy complex128 of shape [N, sy, sx] (the reference's
StateArray order, statearray.cpp:21-75); potentials are float64 [sy, sx].
"""
import ctypes as C
import os
import sys
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(HERE, "lib", "libitp2d_amd.so")

TRANSFORMS = ["FFT", "iFFT", "FFTx", "iFFTx", "FFTy", "iFFTy", "DST", "iDST", "DSTx",
              "iDSTx", "DSTy", "iDSTy", "DCT", "iDCT", "DCTx", "iDCTx", "DCTy", "iDCTy"]
PERIODIC, DIRICHLET = 0, 1
FP64, FP32 = 0, 1
TESTS = {"none": 0, "onestep": 1, "abschange": 2, "relchange": 3, "absstdev": 4, "relstdev": 5}
PHASES = ["propagate", "overlap", "eigensolve", "lincomb", "energies"]

ITP_OK, ITP_E_INVALID, ITP_E_NONPOSITIVE_EIG, ITP_E_NONNORMAL_EIG, ITP_E_EIGENSOLVER = 0, 1, 2, 3, 4


class ITPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class NonPositiveEigenvalue(ITPError):
    """exceptions.hpp:55-71: the overlap matrix has an eigenvalue <= 0."""

    def __init__(self, code, msg, n, value):
        super().__init__(code, msg)
        self.n = n
        self.value = value


class NonNormalEigenvalue(NonPositiveEigenvalue):
    """exceptions.hpp:73-86"""


class EigensolverError(ITPError):
    """exceptions.hpp:237-244"""


_lib = None
_dp = C.POINTER(C.c_double)
_sz = C.c_size_t
_vp = C.c_void_p
# void (*itps_save_fn)(void* user)
_SaveFn = C.CFUNCTYPE(None, C.c_void_p)


class _Test(C.Structure):
    _fields_ = [("kind", C.c_int), ("p0", C.c_double), ("p1", C.c_double)]


class _Params(C.Structure):
    _fields_ = [("sx", _sz), ("sy", _sz), ("lenx", C.c_double), ("boundary", C.c_int), ("N", _sz),
                ("n_wanted", _sz), ("ignore_lowest", _sz), ("halforder", C.c_int), ("B", C.c_double),
                ("eps_values", _dp), ("n_eps", _sz), ("eps_divisor", C.c_double),
                ("exhaust_eps", C.c_int), ("max_steps", C.c_int), ("min_time_step", C.c_double),
                ("recover", C.c_int), ("seed", C.c_uint64), ("timestep_test", _Test),
                ("final_test", _Test), ("precision", C.c_int), ("mixed_tol", C.c_double),
                ("accuracy", C.c_double), ("mixed_check_every", C.c_int), ("mixed_skip_ortho", C.c_int)]


def lib():
    """Loads the gfx950 library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise ImportError(f"{library_path} is missing: run python itp2d_amd/build.py "
                              "(there is no CPU fallback)")
        L = C.CDLL(library_path)
        sig = {
            "itp_ctx_create": [C.c_int, C.POINTER(_vp)],
            "itp_ctx_destroy": [_vp],
            "itp_last_error": [_vp],
            "itp_ctx_synchronize": [_vp],
            "itp_ctx_num_cus": [_vp],
            "itp_ctx_enable_timing": [_vp, C.c_int],
            "itp_ctx_phase_times": [_vp, _dp, C.POINTER(C.c_longlong)],
            "itp_ctx_reset_phase_times": [_vp],
            "itp_ctx_eig_stats": [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
            "itp_ctx_set_eig_inner": [_vp, C.c_int],
            "itp_ctx_set_eig_grid": [_vp, C.c_int],
            "itp_ctx_set_eig_spin_limit": [_vp, C.c_ulonglong],
            "itp_ctx_set_cholesky_tol": [_vp, C.c_double],
            "itp_ctx_set_eig_mode": [_vp, C.c_int],
            "itp_ctx_set_eig_handoff": [_vp, C.c_double],
            "itp_ctx_oa_stats": [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
            "itp_ctx_set_eig_cluster": [_vp, C.c_double],
            "itp_ctx_set_fp32_rotation": [_vp, C.c_int],
            "itp_ctx_cluster_stats": [_vp, C.POINTER(C.c_longlong)],
            "itp_grid_create": [_vp, _sz, _sz, C.c_double, C.c_int, C.POINTER(_vp)],
            "itp_grid_destroy": [_vp],
            "itp_grid_kvalues": [_vp, _dp, _dp, _dp, _dp],
            "itp_grid_positions": [_vp, _dp, _dp],
            "itp_grid_normalization_factor": [_vp, C.c_int],
            "itp_transform_host": [_vp, _dp, _sz, C.c_int],
            "itp_transform_device": [_vp, _vp, _sz, C.c_int],
            "itp_potential_create": [_vp, _dp, C.POINTER(_vp)],
            "itp_potential_destroy": [_vp],
            "itp_stateset_create": [_vp, _sz, C.c_int, C.POINTER(_vp)],
            "itp_stateset_destroy": [_vp],
            "itp_stateset_upload": [_vp, _dp],
            "itp_stateset_download": [_vp, _dp],
            "itp_stateset_device_ptr": [_vp],
            "itp_stateset_size": [_vp],
            "itp_split_create": [_vp, C.c_int, C.c_double, C.c_double, C.POINTER(_vp)],
            "itp_split_destroy": [_vp],
            "itp_split_set_time_step": [_vp, C.c_double],
            "itp_split_halforder": [_vp],
            "itp_propagate": [_vp, _vp],
            "itp_propagate_noalloc": [_vp, _vp],
            "itp_propagate_noalloc_max_bytes": [],
            "itp_orthonormalize": [_vp, C.POINTER(_sz), _dp],
            "itp_overlap": [_vp, _dp],
            "itp_hamiltonian_create": [_vp, C.c_double, C.POINTER(_vp)],
            "itp_hamiltonian_destroy": [_vp],
            "itp_energies": [_vp, _vp, _dp, _dp],
            "itp_rayleigh_ritz": [_vp, _vp, _dp, _dp],
            "itp_stateset_convert": [_vp, _vp],
            "itp_stateset_init_noise": [_vp, C.POINTER(C.c_uint64), C.c_double],
            "itp_stateset_overlap_eigenvalues": [_vp, _dp],
            "itp_energies_complex": [_vp, _vp, _dp, _dp],
            "itp_apply_expkinetic_host": [_vp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _sz],
            "itp_apply_exppotential_host": [_vp, C.c_double, C.c_double, C.c_double, _dp, _sz],
            "itp_apply_secondorder_host": [_vp, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _sz],
            "itp_apply_hamiltonian_host": [_vp, _dp, _sz],
            "itp_eigh_host": [_vp, _sz, _dp, _dp],
            "itp_eigvalsh_host": [_vp, _sz, _dp, _dp],
            "itp_ctx_set_eig_check_min": [_vp, C.c_int],
            "itp_ctx_set_graphs": [_vp, C.c_int],
            "itp_ctx_set_autotune": [_vp, C.c_int],
            "itp_ctx_wisdom_export": [_vp, C.c_char_p, _sz, C.POINTER(_sz)],
            "itp_ctx_wisdom_import": [_vp, C.c_char_p],
            "itp_ctx_tunings": [_vp],
            "itp_ctx_eig_check_stats": [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
            "itp_init_states": [C.POINTER(C.c_uint64), _sz, _sz, C.c_double, _sz, _dp],
            "itps_create": [_vp, C.POINTER(_Params), _dp, _dp, C.POINTER(_vp)],
            "itps_destroy": [_vp],
            "itps_step": [_vp],
            "itps_run": [_vp],
            "itps_finished": [_vp],
            "itps_error_flag": [_vp],
            "itps_total_steps": [_vp],
            "itps_eps": [_vp],
            "itps_num_converged": [_vp],
            "itps_num_timestep_converged": [_vp],
            "itps_sorted": [_vp, _dp, _dp, C.POINTER(_sz)],
            "itps_states": [_vp, _dp],
            "itps_total_time": [_vp],
            "itps_recoveries": [_vp],
            "itps_converged": [_vp, C.POINTER(C.c_int)],
            "itps_history_steps": [_vp],
            "itps_history_row": [_vp, _sz, _dp, _dp],
            "itps_time_step_history": [_vp, C.POINTER(C.c_int), _dp, _sz],
            "itps_rayleigh_ritz": [_vp],
            "itps_mixed_steps": [_vp],
            "itps_mixed_stats": [_vp, C.POINTER(C.c_int)],
            "itp_orthonormalize_cholesky": [_vp, C.POINTER(C.c_int), C.POINTER(_sz), _dp],
            "itp_rayleigh_quotients": [_vp, _vp, _vp, _dp],
            "itp_stateset_cholesky_condition": [_vp, _dp],
            "itp_stateset_cholesky_diagonal": [_vp, _dp],
            "itps_set_abort_flag": [_vp, C.POINTER(C.c_int)],
            "itps_set_save_flag": [_vp, C.POINTER(C.c_int), _SaveFn, _vp],
            "itps_install_signal_flags": [C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "itps_set_verbosity": [_vp, C.c_int],
            "itp_potential_create_type": [_vp, C.c_char_p, _dp, _sz, C.POINTER(_vp)],
            "itp_potential_type_params": [C.c_char_p, _dp, _sz, C.POINTER(C.c_int), _dp],
            "itp_potential_add_impurities": [_vp, C.c_int, _dp, _sz, _dp],
            "itp_potential_values": [_vp, _dp],
            "itp_batch_create": [C.c_int, C.c_int, C.POINTER(_vp)],
            "itp_batch_destroy": [_vp],
            "itp_batch_set_autotune": [_vp, C.c_int],
            "itp_batch_add": [_vp, C.POINTER(_Params), _dp, C.POINTER(_sz)],
            "itp_batch_size": [_vp],
            "itp_batch_run": [_vp],
            "itp_batch_wall_time": [_vp],
            "itp_batch_result": [_vp, _sz, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(_sz),
                                 C.POINTER(C.c_int), _dp],
            "itp_batch_sorted": [_vp, _sz, _dp, _dp, C.POINTER(C.c_int)],
            "itp_batch_mixed_stats": [_vp, _sz, C.POINTER(C.c_int)],
        }
        for name, args in sig.items():
            getattr(L, name).argtypes = args
        L.itp_last_error.restype = C.c_char_p
        L.itp_grid_normalization_factor.restype = C.c_double
        L.itp_stateset_device_ptr.restype = _vp
        L.itp_stateset_size.restype = _sz
        L.itp_propagate_noalloc_max_bytes.restype = _sz
        L.itps_eps.restype = C.c_double
        L.itps_total_time.restype = C.c_double
        L.itps_num_converged.restype = _sz
        L.itps_num_timestep_converged.restype = _sz
        L.itps_history_steps.restype = _sz
        L.itps_time_step_history.restype = _sz
        L.itp_batch_size.restype = _sz
        L.itp_batch_wall_time.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


class _Owned:
    """A handle of the library with an explicit close().

    close() frees the handle at once (the owning Context closes its children
    first, newest first). __del__ calls close() too, except while the
    interpreter is finalizing: objects are then collected in no set order
    and the HIP runtime may already be shutting down, so the handles are left
    to the process exit instead of being freed through HIP calls.
    """

    _destroy = None  # name of the library's destroy function

    def _register(self, ctx):
        kids = ctx._children
        if len(kids) >= 256:
            kids[:] = [r for r in kids if r() is not None]
        kids.append(weakref.ref(self))

    def close(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            getattr(_lib, self._destroy)(h)
        self.h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        self.close()


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class Context:
    """One device and one stream. close() (or leaving a `with` block) frees
    every object made on it, newest first, then the context itself."""

    def __init__(self, device=0):
        self._children = []
        self.h = _vp()
        rc = lib().itp_ctx_create(device, C.byref(self.h))
        if rc:
            raise ITPError(rc, f"itp_ctx_create({device}) failed (no GPU?)")

    def error(self):
        return lib().itp_last_error(self.h).decode()

    def check(self, rc, what="", bad=None):
        if rc == ITP_OK:
            return
        msg = f"{what}: {self.error()}"
        if rc == ITP_E_NONPOSITIVE_EIG:
            raise NonPositiveEigenvalue(rc, msg, *(bad or (0, 0.0)))
        if rc == ITP_E_NONNORMAL_EIG:
            raise NonNormalEigenvalue(rc, msg, *(bad or (0, 0.0)))
        if rc == ITP_E_EIGENSOLVER:
            raise EigensolverError(rc, msg)
        raise ITPError(rc, msg)

    def synchronize(self):
        self.check(lib().itp_ctx_synchronize(self.h), "synchronize")

    @property
    def num_cus(self):
        return lib().itp_ctx_num_cus(self.h)

    def enable_timing(self, on=True):
        lib().itp_ctx_enable_timing(self.h, int(on))

    def phase_times(self):
        ms = (C.c_double * 5)()
        cnt = (C.c_longlong * 5)()
        lib().itp_ctx_phase_times(self.h, ms, cnt)
        return {PHASES[i]: (ms[i], cnt[i]) for i in range(5)}

    def reset_phase_times(self):
        lib().itp_ctx_reset_phase_times(self.h)

    def eig_stats(self):
        calls, sweeps = C.c_longlong(0), C.c_longlong(0)
        lib().itp_ctx_eig_stats(self.h, C.byref(calls), C.byref(sweeps))
        return calls.value, sweeps.value

    def set_eig_inner(self, inner):
        self.check(lib().itp_ctx_set_eig_inner(self.h, int(inner)), "set_eig_inner")

    def set_eig_grid(self, workgroups):
        self.check(lib().itp_ctx_set_eig_grid(self.h, int(workgroups)), "set_eig_grid")

    def set_eig_spin_limit(self, polls):
        """Polls of a block-Jacobi grid barrier before the wait gives up (default 2**24); a wait that
        gives up ends the call with EigensolverError."""
        self.check(lib().itp_ctx_set_eig_spin_limit(self.h, int(polls)), "set_eig_spin_limit")

    def set_cholesky_tol(self, tol):
        """Pivot threshold of the Cholesky orthonormalization relative to the diagonal (0: default)."""
        self.check(lib().itp_ctx_set_cholesky_tol(self.h, float(tol)), "set_cholesky_tol")

    def set_eig_mode(self, mode):
        """0: Jacobi handing off to the refinement iteration (default); 1: Jacobi only."""
        self.check(lib().itp_ctx_set_eig_mode(self.h, int(mode)), "set_eig_mode")

    def set_eig_handoff(self, ratio):
        """Coupling-to-gap ratio below which the Jacobi sweeps hand off to the refinement."""
        self.check(lib().itp_ctx_set_eig_handoff(self.h, float(ratio)), "set_eig_handoff")

    def oa_stats(self):
        """(calls finished by the refinement iteration, calls that fell back to Jacobi)."""
        a, b = C.c_longlong(0), C.c_longlong(0)
        lib().itp_ctx_oa_stats(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_fp32_rotation(self, kind):
        """Rotation of FP32 sets: 1 split-BF16x3 (default), 0 FP32 MFMA."""
        self.check(lib().itp_ctx_set_fp32_rotation(self.h, int(kind)), "set_fp32_rotation")

    def set_graphs(self, on):
        """1: propagation, overlap and Cholesky factor launch sequences replayed as HIP graphs; 0 (default):
        launched one by one."""
        self.check(lib().itp_ctx_set_graphs(self.h, int(on)), "set_graphs")

    def set_autotune(self, on):
        """1: the first overlap of each shape (precision, grid points, padded states) times the Gram
        product's split-K chunk counts and later sets of that shape use the fastest; 0 (default): the
        formula's count."""
        self.check(lib().itp_ctx_set_autotune(self.h, int(on)), "set_autotune")

    def tunings(self):
        """Shapes tuned by measurement in this context."""
        return lib().itp_ctx_tunings(self.h)

    def wisdom_export(self):
        """The process's tuned shapes as text, one "gram <bytes per real> <grid points> <padded
        states> <chunks>" line each."""
        n = _sz(0)
        self.check(lib().itp_ctx_wisdom_export(self.h, None, 0, C.byref(n)), "wisdom_export")
        buf = C.create_string_buffer(n.value + 1)
        self.check(lib().itp_ctx_wisdom_export(self.h, buf, n.value + 1, C.byref(n)), "wisdom_export")
        return buf.value.decode()

    def wisdom_import(self, text):
        """Adds the shapes of a wisdom text; ITPError (nothing added) on a malformed line."""
        self.check(lib().itp_ctx_wisdom_import(self.h, text.encode()), "wisdom_import")

    def set_eig_check_min(self, n):
        """Overlap matrices of at least n states get the eigenvalue check (tridiagonalization and
        bisection) before the eigendecomposition (default 1024)."""
        self.check(lib().itp_ctx_set_eig_check_min(self.h, int(n)), "set_eig_check_min")

    def eig_check_stats(self):
        """(eigenvalue checks run, matrices they rejected)"""
        a, b = C.c_longlong(0), C.c_longlong(0)
        lib().itp_ctx_eig_check_stats(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_eig_cluster(self, tau):
        """Cluster start threshold on |a_ij| / |a_ii - a_jj| (0: off)."""
        self.check(lib().itp_ctx_set_eig_cluster(self.h, float(tau)), "set_eig_cluster")

    def cluster_stats(self):
        """(calls finished by the cluster start, clusters solved, largest cluster, fallbacks to Jacobi)"""
        out = (C.c_longlong * 4)()
        self.check(lib().itp_ctx_cluster_stats(self.h, out), "cluster_stats")
        return tuple(out)

    def close(self):
        for ref in reversed(getattr(self, "_children", [])):
            obj = ref()
            if obj is not None:
                obj.close()
        self._children = []
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.itp_ctx_destroy(h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if sys.is_finalizing():
            return
        self.close()


class Grid(_Owned):
    """DataLayout + Transformer (datalayout.hpp:30-54, transformer.hpp:30-63)."""

    _destroy = "itp_grid_destroy"

    def __init__(self, ctx, sx, sy, dx, boundary=PERIODIC):
        self.ctx, self.sx, self.sy, self.dx, self.boundary = ctx, sx, sy, dx, boundary
        self.h = _vp()
        ctx.check(lib().itp_grid_create(ctx.h, sx, sy, dx, boundary, C.byref(self.h)), "itp_grid_create")
        self._register(ctx)

    def kvalues(self):
        a = [np.zeros(n) for n in (self.sx, self.sy, self.sx, self.sy)]
        lib().itp_grid_kvalues(self.h, *[_p(x) for x in a])
        return dict(zip(["fft_kx", "fft_ky", "dsct_kx", "dsct_ky"], a))

    def positions(self):
        px, py = np.zeros(self.sx), np.zeros(self.sy)
        lib().itp_grid_positions(self.h, _p(px), _p(py))
        return px, py

    def normalization_factor(self, kind):
        return lib().itp_grid_normalization_factor(self.h, TRANSFORMS.index(kind) if isinstance(kind, str) else kind)

    def transform(self, fields, kind):
        f = _c128(fields).copy()
        squeeze = f.ndim == 2
        if squeeze:
            f = f[None]
        k = TRANSFORMS.index(kind) if isinstance(kind, str) else kind
        self.ctx.check(lib().itp_transform_host(self.h, _p(f.view(np.float64)), f.shape[0], k), "transform")
        return f[0] if squeeze else f

    def apply_expkinetic(self, fields, B, e, coeff=-1.0, prefactor=1.0):
        f = _c128(fields).copy()
        self.ctx.check(lib().itp_apply_expkinetic_host(self.h, B, e, coeff, prefactor, _p(f.view(np.float64)),
                                                        f.shape[0]), "expkinetic")
        return f



POTENTIAL_TYPES = ["zero", "harmonic", "elliptic", "prettyhardsquare", "softpentagon", "henonheiles", "gaussian",
                   "quartic", "squareoscillator", "power", "ring", "cosh", "softstadium", "powerstadium"]
IMP_GAUSSIAN, IMP_COULOMB, IMP_HEMISPHERE, IMP_DELTA = 0, 1, 2, 3


def potential_type_params(name, params=()):
    """Kind index and the 5 resolved parameters of a named potential type (host only);
    raises ITPError for parameters the reference rejects."""
    a = np.ascontiguousarray(params, dtype=np.float64)
    kind = C.c_int(-1)
    out = np.zeros(5)
    rc = lib().itp_potential_type_params(name.encode(), _p(a) if a.size else None, a.size, C.byref(kind), _p(out))
    if rc:
        raise ITPError(rc, f"potential type {name}{tuple(params)}")
    return kind.value, out


class Potential(_Owned):
    """Potential (potential.hpp:31-58); V=None is the null potential."""

    _destroy = "itp_potential_destroy"

    def __init__(self, grid, V=None, _handle=None):
        self.grid = grid
        self.h = _vp()
        if _handle is not None:
            self.h = _handle
            self.V = None
            self._register(grid.ctx)
            return
        self.V = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
        grid.ctx.check(lib().itp_potential_create(grid.h, _p(self.V), C.byref(self.h)), "itp_potential_create")
        self._register(grid.ctx)

    @classmethod
    def of_type(cls, grid, name, params=()):
        """A named potential type of the reference, evaluated on the device."""
        a = np.ascontiguousarray(params, dtype=np.float64)
        h = _vp()
        grid.ctx.check(lib().itp_potential_create_type(grid.h, name.encode(), _p(a) if a.size else None, a.size,
                                                       C.byref(h)), f"potential {name}")
        return cls(grid, _handle=h)

    def add_impurities(self, kind, data, kind_params=None):
        """Adds realised impurities (rows x, y[, per-impurity parameters]) of a kind."""
        d = np.ascontiguousarray(data, dtype=np.float64)
        kp = None if kind_params is None else np.ascontiguousarray(kind_params, dtype=np.float64)
        n = d.shape[0] if d.size else 0
        self.grid.ctx.check(lib().itp_potential_add_impurities(self.h, int(kind), _p(d) if n else None, n, _p(kp)),
                            "add_impurities")

    def values(self):
        out = np.zeros((self.grid.sy, self.grid.sx))
        self.grid.ctx.check(lib().itp_potential_values(self.h, _p(out)), "potential values")
        return out

    def apply_exppotential(self, fields, e, coeff=-0.5, prefactor=1.0):
        f = _c128(fields).copy()
        self.grid.ctx.check(lib().itp_apply_exppotential_host(self.h, e, coeff, prefactor,
                                                               _p(f.view(np.float64)), f.shape[0]), "exppotential")
        return f

    def apply_secondorder(self, fields, B, e, prefactor=1.0, exponent=1):
        f = _c128(fields).copy()
        self.grid.ctx.check(lib().itp_apply_secondorder_host(self.h, B, e, prefactor, exponent,
                                                              _p(f.view(np.float64)), f.shape[0]), "secondorder")
        return f



class StateSet(_Owned):
    """StateSet (stateset.hpp): N fields on the device, FP64 or FP32."""

    _destroy = "itp_stateset_destroy"

    def __init__(self, grid, N, precision=FP64):
        self.grid, self.N, self.precision = grid, N, precision
        self.h = _vp()
        grid.ctx.check(lib().itp_stateset_create(grid.h, N, precision, C.byref(self.h)), "itp_stateset_create")
        self._register(grid.ctx)

    def upload(self, states):
        s = _c128(states)
        assert s.shape == (self.N, self.grid.sy, self.grid.sx)
        self.grid.ctx.check(lib().itp_stateset_upload(self.h, _p(s.view(np.float64))), "upload")

    def download(self):
        out = np.zeros((self.N, self.grid.sy, self.grid.sx), dtype=np.complex128)
        self.grid.ctx.check(lib().itp_stateset_download(self.h, _p(out.view(np.float64))), "download")
        return out

    def orthonormalize(self):
        n = _sz(0)
        v = C.c_double(0)
        rc = lib().itp_orthonormalize(self.h, C.byref(n), C.byref(v))
        self.grid.ctx.check(rc, "orthonormalize", bad=(n.value, v.value))

    def orthonormalize_cholesky(self):
        """Cholesky orthonormalization Psi <- Psi R^-1 (S = R^H R); returns True when it fell back
        to the eigendecomposition route."""
        used = C.c_int(0)
        n = _sz(0)
        v = C.c_double(0)
        rc = lib().itp_orthonormalize_cholesky(self.h, C.byref(used), C.byref(n), C.byref(v))
        self.grid.ctx.check(rc, "orthonormalize_cholesky", bad=(n.value, v.value))
        return bool(used.value)

    def assign(self, other):
        """Copy the fields of another set of the same grid and size (converting precision)."""
        self.grid.ctx.check(lib().itp_stateset_convert(self.h, other.h), "convert")

    def init_noise(self, seed, dx):
        """Fill with the seeded noise of init_states, generated on the device; returns the advanced seed."""
        st = C.c_uint64(seed)
        self.grid.ctx.check(lib().itp_stateset_init_noise(self.h, C.byref(st), C.c_double(dx)), "init_noise")
        return st.value

    def cholesky_condition(self):
        """max R_ii^2 / min R_ii^2 of the last Cholesky orthonormalization's factor (0: none)."""
        k = C.c_double(0)
        self.grid.ctx.check(lib().itp_stateset_cholesky_condition(self.h, C.byref(k)), "cholesky_condition")
        return k.value

    def overlap_eigenvalues(self):
        lam = np.zeros(self.N)
        self.grid.ctx.check(lib().itp_stateset_overlap_eigenvalues(self.h, _p(lam)), "overlap_eigenvalues")
        return lam

    def overlap(self):
        out = np.zeros(self.N * self.N, dtype=np.complex128)
        self.grid.ctx.check(lib().itp_overlap(self.h, _p(out.view(np.float64))), "overlap")
        return out.reshape(self.N, self.N).T  # column-major -> S[i, j]



class MultiProductSplit(_Owned):
    """The evolution operator T (multiproductsplit.cpp:23-58)."""

    _destroy = "itp_split_destroy"

    def __init__(self, potential, halforder, eps, B=0.0):
        self.pot = potential
        self.h = _vp()
        potential.grid.ctx.check(lib().itp_split_create(potential.h, halforder, eps, B, C.byref(self.h)),
                                 "itp_split_create")
        self._register(potential.grid.ctx)

    @property
    def halforder(self):
        return lib().itp_split_halforder(self.h)

    def set_time_step(self, eps):
        self.pot.grid.ctx.check(lib().itp_split_set_time_step(self.h, eps), "set_time_step")

    def propagate(self, states):
        self.pot.grid.ctx.check(lib().itp_propagate(self.h, states.h), "propagate")

    def propagate_noalloc(self, states):
        """propagate() with its intermediate fields in the library's fixed device scratch array
        (no allocation; sets larger than noalloc_max_bytes() are rejected)."""
        self.pot.grid.ctx.check(lib().itp_propagate_noalloc(self.h, states.h), "propagate_noalloc")



class Hamiltonian(_Owned):
    """H = T + V (itpsystem.cpp:117-119)."""

    _destroy = "itp_hamiltonian_destroy"

    def __init__(self, potential, B=0.0):
        self.pot = potential
        self.h = _vp()
        potential.grid.ctx.check(lib().itp_hamiltonian_create(potential.h, B, C.byref(self.h)),
                                 "itp_hamiltonian_create")
        self._register(potential.grid.ctx)

    def energies(self, states):
        E = np.zeros(states.N)
        sd = np.zeros(states.N)
        self.pot.grid.ctx.check(lib().itp_energies(self.h, states.h, _p(E), _p(sd)), "energies")
        return E, sd

    def rayleigh_ritz(self, states):
        """Rayleigh-Ritz on the set: (ascending Ritz values, residual norms); slot n becomes Ritz vector n."""
        E = np.zeros(states.N)
        sd = np.zeros(states.N)
        self.pot.grid.ctx.check(lib().itp_rayleigh_ritz(self.h, states.h, _p(E), _p(sd)), "rayleigh_ritz")
        return E, sd

    def rayleigh_quotients(self, states, dst):
        """FP64 Rayleigh quotients of the states (either precision), evaluated on a copy in dst
        (an FP64 set of the same grid and size)."""
        E = np.zeros(states.N)
        self.pot.grid.ctx.check(lib().itp_rayleigh_quotients(self.h, states.h, dst.h, _p(E)), "rayleigh_quotients")
        return E

    def energies_complex(self, states):
        m = np.zeros(2 * states.N)
        q = np.zeros(2 * states.N)
        self.pot.grid.ctx.check(lib().itp_energies_complex(self.h, states.h, _p(m), _p(q)), "energies")
        return m.view(np.complex128), q.view(np.complex128)

    def apply(self, fields):
        f = _c128(fields).copy()
        self.pot.grid.ctx.check(lib().itp_apply_hamiltonian_host(self.h, _p(f.view(np.float64)), f.shape[0]),
                                "hamiltonian")
        return f



_signal_flags = None


def install_signal_flags():
    """The reference program's SIGINT and SIGUSR1 handlers (itp2d.cpp:40-66),
    installed in the library: returns (abort_flag, save_flag) as ctypes.c_int
    for ITPSystem.set_abort_flag and set_save_flag. Installed once per process."""
    global _signal_flags
    if _signal_flags is None:
        flags = (C.c_int(0), C.c_int(0))
        rc = lib().itps_install_signal_flags(C.byref(flags[0]), C.byref(flags[1]))
        if rc:
            raise ITPError(rc, "itps_install_signal_flags")
        _signal_flags = flags
    return _signal_flags


def noalloc_max_bytes():
    """Size of the fixed device scratch array of propagate_noalloc."""
    return lib().itp_propagate_noalloc_max_bytes()


def eigh(ctx, A):
    """EigenSolver::solve on the GPU: ascending w, eigenvectors in columns."""
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    a = np.ascontiguousarray(A.T).ravel().copy()  # column-major
    w = np.zeros(n)
    ctx.check(lib().itp_eigh_host(ctx.h, n, _p(a.view(np.float64)), _p(w)), "eigh")
    return w, np.ascontiguousarray(a.reshape(n, n).T)


def eigvalsh(ctx, A):
    """Ascending eigenvalues of the Hermitian A by tridiagonalization and bisection on the GPU."""
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    a = np.ascontiguousarray(A.T).ravel().copy()  # column-major
    w = np.zeros(n)
    ctx.check(lib().itp_eigvalsh_host(ctx.h, n, _p(a.view(np.float64)), _p(w)), "eigvalsh")
    return w


def init_states(seed, N, sx, sy, dx):
    """Seeded uniform-noise initial states, each normalised (host-side setup)."""
    st = C.c_uint64(seed)
    out = np.zeros((N, sy, sx), dtype=np.complex128)
    lib().itp_init_states(C.byref(st), sx, sy, dx, N, _p(out.view(np.float64)))
    return out


def _params(cfg):
    p = _Params()
    eps = np.ascontiguousarray(cfg.get("eps_values", []), dtype=np.float64)
    for k in ["sx", "sy", "lenx", "boundary", "N", "n_wanted", "halforder", "B", "eps_divisor",
              "max_steps", "min_time_step", "seed"]:
        setattr(p, k, cfg[k])
    p.ignore_lowest = cfg.get("ignore_lowest", 0)
    p.exhaust_eps = int(cfg.get("exhaust_eps", 0))
    p.recover = int(cfg.get("recover", 0))
    p.precision = int(cfg.get("precision", FP64))
    p.mixed_tol = float(cfg.get("mixed_tol", 0.0))
    p.accuracy = float(cfg.get("accuracy", 0.0))
    p.mixed_check_every = int(cfg.get("mixed_check_every", 0))
    p.mixed_skip_ortho = int(cfg.get("mixed_skip_ortho", 0))
    p.eps_values = _p(eps) if eps.size else None
    p.n_eps = eps.size
    for name in ["timestep_test", "final_test"]:
        t = cfg[name]
        setattr(p, name, _Test(TESTS[t[0]], t[1] if len(t) > 1 else 0.0, t[2] if len(t) > 2 else 0.0))
    return p, eps


class ITPSystem(_Owned):
    """The ITP driver (itpsystem.cpp) on the GPU engine."""

    _destroy = "itps_destroy"

    def __init__(self, ctx, cfg, V=None, init_states=None):
        self.ctx, self.cfg = ctx, cfg
        self._params, self._eps = _params(cfg)
        self._V = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
        self._init = None if init_states is None else _c128(init_states)
        self.h = _vp()
        rc = lib().itps_create(ctx.h, C.byref(self._params), _p(self._V),
                               None if self._init is None else _p(self._init.view(np.float64)), C.byref(self.h))
        ctx.check(rc, "itps_create")
        self._register(ctx)

    def step(self):
        return lib().itps_step(self.h)

    def run(self):
        return lib().itps_run(self.h)

    @property
    def finished(self):
        return bool(lib().itps_finished(self.h))

    @property
    def error_flag(self):
        return bool(lib().itps_error_flag(self.h))

    @property
    def total_steps(self):
        return lib().itps_total_steps(self.h)

    @property
    def eps(self):
        return lib().itps_eps(self.h)

    @property
    def num_converged(self):
        return lib().itps_num_converged(self.h)

    @property
    def total_time(self):
        return lib().itps_total_time(self.h)

    @property
    def recoveries(self):
        return lib().itps_recoveries(self.h)

    @property
    def mixed_steps(self):
        """Steps of the mixed path's FP32 phase."""
        return lib().itps_mixed_steps(self.h)

    def mixed_stats(self):
        """(FP32-phase steps, FP64 checks, CholeskyQR steps that fell back to the eigendecomposition,
        steps whose orthonormalization was skipped)."""
        out = (C.c_int * 4)()
        self.ctx.check(lib().itps_mixed_stats(self.h, out), "itps_mixed_stats")
        return tuple(out)

    def set_abort_flag(self, flag):
        """Poll a ctypes.c_int at the start of every step; non-zero ends the run with the error flag
        (the reference's SIGINT abort flag). None removes it."""
        self._abort = flag
        self.ctx.check(lib().itps_set_abort_flag(self.h, None if flag is None else C.byref(flag)), "set_abort_flag")

    def set_save_flag(self, flag, callback):
        """Check a ctypes.c_int after propagation, orthonormalization and the energies of every step and at
        its end; when non-zero, callback(self) runs and the flag is cleared (the reference's SIGUSR1 save
        flag, itpsystem.cpp:235-241). None removes it."""
        if flag is None:
            self._save = None
            self.ctx.check(lib().itps_set_save_flag(self.h, None, _SaveFn(), None), "set_save_flag")
            return
        fn = _SaveFn(lambda user: callback(self))
        self._save = (flag, fn)  # kept alive while the driver may call it
        self.ctx.check(lib().itps_set_save_flag(self.h, C.byref(flag), fn, None), "set_save_flag")

    def set_verbosity(self, level):
        """Progress printout levels 0-3 as the reference's -v."""
        self.ctx.check(lib().itps_set_verbosity(self.h, int(level)), "set_verbosity")

    def rayleigh_ritz(self):
        """Rayleigh-Ritz pass on the current states (bias-free Ritz values and residual norms)."""
        self.ctx.check(lib().itps_rayleigh_ritz(self.h), "itps_rayleigh_ritz")

    def converged(self):
        """Final-convergence flag of each sorted position."""
        f = np.zeros(int(self.cfg["N"]), dtype=np.int32)
        lib().itps_converged(self.h, f.ctypes.data_as(C.POINTER(C.c_int)))
        return f.astype(bool)

    def history(self):
        """Step history: (E, sigma) arrays [steps][N] of the sorted energies and
        deviations of every step (NaN rows for steps that computed none)."""
        steps, N = lib().itps_history_steps(self.h), int(self.cfg["N"])
        E, sd = np.zeros((steps, N)), np.zeros((steps, N))
        for k in range(steps):
            self.ctx.check(lib().itps_history_row(self.h, k, _p(E[k]), _p(sd[k])), "itps_history_row")
        return E, sd

    def time_step_history(self):
        """(first step using it, eps) pairs of every time step the run used."""
        n = lib().itps_time_step_history(self.h, None, None, 0)
        steps, eps = (C.c_int * max(n, 1))(), np.zeros(max(n, 1))
        lib().itps_time_step_history(self.h, steps, _p(eps), n)
        return [(int(steps[k]), float(eps[k])) for k in range(n)]

    def sorted(self):
        N = self.cfg["N"]
        E, sd = np.zeros(N), np.zeros(N)
        idx = (_sz * N)()
        lib().itps_sorted(self.h, _p(E), _p(sd), idx)
        return E, sd, np.array(list(idx))

    def states(self):
        N, sy, sx = self.cfg["N"], self.cfg["sy"], self.cfg["sx"]
        out = np.zeros((N, sy, sx), dtype=np.complex128)
        self.ctx.check(lib().itps_states(self.h, _p(out.view(np.float64))), "itps_states")
        return out



class Batch:
    """Independent ITP problems run together (include/itp2d_batch.h): worker
    host threads, each with its own context and stream, take the instances
    largest first; every instance runs the loop of ITPSystem."""

    def __init__(self, device=0, workers=4):
        self.h = _vp()
        rc = lib().itp_batch_create(int(device), int(workers), C.byref(self.h))
        if rc:
            raise ITPError(rc, "itp_batch_create")
        self._cfgs = []

    def set_autotune(self, on):
        """Autotuning (Context.set_autotune) in every worker context of later runs."""
        rc = lib().itp_batch_set_autotune(self.h, int(on))
        if rc:
            raise ITPError(rc, "itp_batch_set_autotune")

    def add(self, cfg, V=None):
        p, eps = _params(cfg)
        Vc = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
        idx = _sz(0)
        rc = lib().itp_batch_add(self.h, C.byref(p), _p(Vc), C.byref(idx))
        if rc:
            raise ITPError(rc, "itp_batch_add")
        self._cfgs.append(cfg)
        return idx.value

    def __len__(self):
        return lib().itp_batch_size(self.h)

    def run(self):
        """Runs every instance; returns the first failing status (0 if none)."""
        return lib().itp_batch_run(self.h)

    @property
    def wall_time(self):
        return lib().itp_batch_wall_time(self.h)

    def result(self, i):
        st, steps, err = C.c_int(0), C.c_int(0), C.c_int(0)
        conv = _sz(0)
        t = np.zeros(1)
        rc = lib().itp_batch_result(self.h, int(i), C.byref(st), C.byref(steps), C.byref(conv), C.byref(err), _p(t))
        if rc:
            raise ITPError(rc, "itp_batch_result")
        return dict(status=st.value, steps=steps.value, converged=conv.value, error_flag=bool(err.value),
                    time=float(t[0]))

    def mixed_stats(self, i):
        """ITPSystem.mixed_stats of instance i."""
        out = (C.c_int * 4)()
        rc = lib().itp_batch_mixed_stats(self.h, int(i), out)
        if rc:
            raise ITPError(rc, "itp_batch_mixed_stats")
        return tuple(out)

    def sorted(self, i):
        N = int(self._cfgs[i]["N"])
        E, sd = np.zeros(N), np.zeros(N)
        f = np.zeros(N, dtype=np.int32)
        rc = lib().itp_batch_sorted(self.h, int(i), _p(E), _p(sd), f.ctypes.data_as(C.POINTER(C.c_int)))
        if rc:
            raise ITPError(rc, "itp_batch_sorted")
        return E, sd, f.astype(bool)

    def close(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.itp_batch_destroy(h)
        self.h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        self.close()
