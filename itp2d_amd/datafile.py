"""HDF5 result files in the reference's layout (Datafile, datafile.cpp:32-467).

`Datafile` wraps the C ABI of include/itp2d_datafile.h (libitp2d_datafile.so, a
host-only library linked to the HDF5 C library). `save_run` writes a finished
ITPSystem run as ITPSystem writes it with its default FinalStates save mode
(itpsystem.cpp:70-114, 221-226, 399-403, 512-536): the run attributes, the
potential, the time-step, energy and deviation histories, the sorted final
states with their state history, and the final energies and deviations.
"""
import ctypes as C
import os
import time

import numpy as np

_dp = C.POINTER(C.c_double)
_sz = C.c_size_t
_vp = C.c_void_p
_lib = None

PROGRAM_VERSION = "itp2d_amd 1.0"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libitp2d_datafile.so")


class DatafileError(RuntimeError):
    pass


def lib():
    """The datafile library (raises if it was not built: it needs the HDF5 C library)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DatafileError(f"{LIB_PATH} is not built (it needs the HDF5 C library; see itp2d_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        sigs = {
            "itp_datafile_create": [C.c_char_p, _sz, _sz, C.c_double, C.c_int, C.POINTER(_vp)],
            "itp_datafile_close": [_vp],
            "itp_datafile_flush": [_vp],
            "itp_datafile_attr_int": [_vp, C.c_char_p, C.c_int],
            "itp_datafile_attr_ulong": [_vp, C.c_char_p, C.c_ulong],
            "itp_datafile_attr_double": [_vp, C.c_char_p, C.c_double],
            "itp_datafile_attr_string": [_vp, C.c_char_p, C.c_char_p],
            "itp_datafile_write_states": [_vp, C.c_int, _dp, _sz, C.POINTER(_sz)],
            "itp_datafile_write_time_step": [_vp, C.c_int, C.c_double],
            "itp_datafile_write_history_row": [_vp, C.c_int, _sz, _dp, _sz],
            "itp_datafile_write_final": [_vp, C.c_int, _dp, _sz],
            "itp_datafile_write_potential": [_vp, _dp],
            "itp_datafile_write_noise": [_vp, _dp, _sz],
            "itp_datafile_read_doubles": [C.c_char_p, C.c_char_p, _dp, _sz, C.POINTER(_sz)],
            "itp_datafile_last_error": [],
        }
        for name, args in sigs.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_int
        L.itp_datafile_close.restype = None
        L.itp_datafile_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise DatafileError(f"{what}: {lib().itp_datafile_last_error().decode()}")


def _d(a):
    return a.ctypes.data_as(_dp)


class Datafile:
    """An HDF5 result file (Datafile, datafile.hpp:40-123)."""

    def __init__(self, path, sx, sy, dx, clobber=False):
        self.h = _vp()
        self.sx, self.sy = int(sx), int(sy)
        _check(lib().itp_datafile_create(os.fsencode(path), self.sx, self.sy, float(dx), int(bool(clobber)),
                                         C.byref(self.h)), f"create {path}")

    def attr(self, name, value):
        """Root attribute: bool and int as int, float as double, str as a string
        (Datafile::add_attribute overloads)."""
        n = name.encode()
        if isinstance(value, (bool, np.bool_, int, np.integer)):
            _check(lib().itp_datafile_attr_int(self.h, n, int(value)), name)
        elif isinstance(value, (float, np.floating)):
            _check(lib().itp_datafile_attr_double(self.h, n, float(value)), name)
        else:
            _check(lib().itp_datafile_attr_string(self.h, n, str(value).encode()), name)

    def attr_ulong(self, name, value):
        _check(lib().itp_datafile_attr_ulong(self.h, name.encode(), int(value)), name)

    def write_states(self, step, states, order=None):
        """One slot of /states (complex [N][sy][sx]), in `order` if given, and its /state_history pair."""
        a = np.ascontiguousarray(states, dtype=np.complex128)
        if a.ndim != 3 or a.shape[1:] != (self.sy, self.sx):
            raise ValueError(f"states must be [N][{self.sy}][{self.sx}]")
        o = None
        if order is not None:
            o = np.ascontiguousarray(order, dtype=np.uint64)
            if o.shape != (a.shape[0],):
                raise ValueError("order must list every state")
        _check(lib().itp_datafile_write_states(self.h, int(step), _d(a.view(np.float64)), a.shape[0],
                                               None if o is None else o.ctypes.data_as(C.POINTER(_sz))),
               "write_states")

    def write_time_step(self, step, eps):
        _check(lib().itp_datafile_write_time_step(self.h, int(step), float(eps)), "write_time_step")

    def write_history_row(self, which, step, values):
        """Row `step` of /energy_history (which 0) or /deviation_history (which 1)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(lib().itp_datafile_write_history_row(self.h, int(which), int(step), _d(v), v.size), "history row")

    def write_final(self, which, values):
        """/final_energies (which 0) or /final_energy_standard_deviations (which 1)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(lib().itp_datafile_write_final(self.h, int(which), _d(v), v.size), "final values")

    def write_potential(self, V):
        v = np.ascontiguousarray(V, dtype=np.float64)
        if v.shape != (self.sy, self.sx):
            raise ValueError(f"potential must be [{self.sy}][{self.sx}]")
        _check(lib().itp_datafile_write_potential(self.h, _d(v)), "write_potential")

    def write_noise(self, data):
        v = np.ascontiguousarray(data, dtype=np.float64).ravel()
        _check(lib().itp_datafile_write_noise(self.h, _d(v), v.size), "write_noise")

    def flush(self):
        _check(lib().itp_datafile_flush(self.h), "flush")

    def close(self):
        if self.h:
            lib().itp_datafile_close(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_doubles(path, dataset):
    """A double dataset of a result file as a flat array (for checks)."""
    n = _sz(0)
    _check(lib().itp_datafile_read_doubles(os.fsencode(path), dataset.encode(), None, 0, C.byref(n)), dataset)
    out = np.zeros(n.value)
    if n.value:
        _check(lib().itp_datafile_read_doubles(os.fsencode(path), dataset.encode(), _d(out), n.value, C.byref(n)),
               dataset)
    return out


def convergence_test_description(t):
    """ConvergenceTest description strings (convergence.hpp:57-68, convergence.cpp:105-206)."""
    kind = t[0]
    p = list(t[1:]) + [0.0, 0.0]
    if kind == "none":
        return "none"
    if kind == "onestep":
        return "one-step convergence"
    if kind == "relchange":
        return f"relative energy change < {p[0]:g}"
    if kind == "abschange":
        return f"absolute energy change < {p[0]:g}"
    if kind == "absstdev":
        return f"absolute energy deviation < {p[0]:g} or absolute energy deviation change < {p[1]:g}"
    if kind == "relstdev":
        return f"relative energy deviation < {p[0]:g} or relative energy deviation change < {p[1]:g}"
    raise ValueError(kind)


def save_run(system, path, V=None, save="final_states", clobber=False, attributes=None, noise_data=None):
    """Writes a finished ITPSystem run to `path` in the reference's layout.

    save: "final_states" (the reference's default FinalStates: histories, final
    energies and the sorted final states) or "energies" (OnlyEnergies: no states).
    attributes: extra root attributes, e.g. the reference's "potential",
    "noise", "impurity_type", "impurity_distribution", "impurity_constraint"
    descriptions (written as "none" when not given).
    """
    if save not in ("final_states", "energies"):
        raise ValueError("save must be 'final_states' or 'energies'")
    cfg = system.cfg
    with Datafile(path, int(cfg["sx"]), int(cfg["sy"]), float(cfg["lenx"]) / int(cfg["sx"]), clobber) as f:
        write_run_attributes(f, system, V, attributes, noise_data)
        write_run_results(f, system, final_states=save == "final_states")
    return path


def _save_on_flag(system, f, save_flag):
    """SIGUSR1-style saves (itpsystem.cpp:235-241, 374-388): when save_flag (a
    ctypes.c_int) is raised during the run, the current states are written at
    the current step, sorted by the last computed energies (unsorted before
    the first energies)."""
    if save_flag is None:
        return

    def save(sys_):
        order = sys_.sorted()[2] if sys_.history()[0].shape[0] > 0 else None
        f.write_states(sys_.total_steps, sys_.states(), order=order)

    system.set_save_flag(save_flag, save)


def run_and_save(system, path, V=None, save="final_states", clobber=False, attributes=None, noise_data=None,
                 save_flag=None):
    """Runs `system` to the end with the result file open, as the reference's
    ITPSystem does: the run description first, states written whenever
    save_flag (a ctypes.c_int, e.g. raised by SIGUSR1) is raised, then the
    results as save_run writes them. Returns the driver's status."""
    if save not in ("final_states", "energies"):
        raise ValueError("save must be 'final_states' or 'energies'")
    cfg = system.cfg
    with Datafile(path, int(cfg["sx"]), int(cfg["sy"]), float(cfg["lenx"]) / int(cfg["sx"]), clobber) as f:
        write_run_attributes(f, system, V, attributes, noise_data)
        _save_on_flag(system, f, save_flag)
        try:
            rc = system.run()
        finally:
            if save_flag is not None:
                system.set_save_flag(None, None)
        write_run_results(f, system, final_states=save == "final_states")
    return rc


def run_saving_every_step(system, path, V=None, clobber=False, attributes=None, noise_data=None, save_flag=None):
    """Runs `system` to the end, saving as the reference's Everything mode does:
    the initial states (unsorted, step 0, itpsystem.cpp:132-133), the sorted
    states after every step (itpsystem.cpp:451-453), the histories and the
    final energies; and, when save_flag is raised, the states at that point
    as well. Returns the driver's status."""
    cfg = system.cfg
    rc = 0
    with Datafile(path, int(cfg["sx"]), int(cfg["sy"]), float(cfg["lenx"]) / int(cfg["sx"]), clobber) as f:
        write_run_attributes(f, system, V, attributes, noise_data)
        _save_on_flag(system, f, save_flag)
        f.write_states(0, system.states())
        while not system.finished:
            r = system.step()
            rc = rc or r
            if system.total_steps > 0 and not (system.finished and system.error_flag):
                f.write_states(system.total_steps, system.states(), order=system.sorted()[2])
        if save_flag is not None:
            system.set_save_flag(None, None)
        write_run_results(f, system, final_states=False)
    return rc


def write_run_attributes(f, system, V=None, attributes=None, noise_data=None):
    """The run description written when the reference's ITPSystem starts
    (itpsystem.cpp:70-114): run attributes, the potential, the noise data."""
    cfg = system.cfg
    sx, sy = int(cfg["sx"]), int(cfg["sy"])
    extra = dict(attributes or {})
    f.attr("program_version", extra.pop("program_version", PROGRAM_VERSION))
    f.attr_ulong("random_seed", int(cfg["seed"]))
    f.attr("start_time", extra.pop("start_time", time.strftime("%Y-%m-%d %H:%M:%S")))
    f.attr("num_threads", int(extra.pop("num_threads", 1)))
    f.attr("num_states", int(cfg["N"]))
    f.attr("num_wanted_to_converge", int(cfg["n_wanted"]))
    f.attr("ignore_lowest", int(cfg.get("ignore_lowest", 0)))
    f.attr("grid_length", float(cfg["lenx"]))
    f.attr("grid_boundary_type", "dirichlet" if int(cfg["boundary"]) else "periodic")
    f.attr("operator_splitting_order", 2 * int(cfg["halforder"]))
    for name in ("potential", "noise", "impurity_type", "impurity_distribution", "impurity_constraint"):
        f.attr(name, extra.pop(name, "none"))
    f.attr("timestep_convergence_test", convergence_test_description(cfg["timestep_test"]))
    f.attr("final_convergence_test", convergence_test_description(cfg["final_test"]))
    f.attr("magnetic_field_strength", float(cfg["B"]))
    for name, value in extra.items():
        f.attr(name, value)
    if V is not None and np.any(np.asarray(V) != 0):  # a null potential is not written
        f.write_potential(np.asarray(V, dtype=np.float64).reshape(sy, sx))
    if noise_data is not None:
        f.write_noise(noise_data)


def write_run_results(f, system, final_states=True):
    """What the reference writes during and at the end of the run
    (itpsystem.cpp:113-114, 221-226, 399-403, 512-536)."""
    ts = system.time_step_history()
    if ts:
        f.attr("initial_time_step", ts[0][1])
    for step, eps in ts:
        f.write_time_step(step, eps)
    # per-step energies and deviations
    E_hist, sd_hist = system.history()
    for k in range(E_hist.shape[0]):
        f.write_history_row(0, k, E_hist[k])
        f.write_history_row(1, k, sd_hist[k])
    E, sd, idx = system.sorted()
    if final_states:
        f.write_states(system.total_steps, system.states(), order=idx)
    f.write_final(0, E)
    f.write_final(1, sd)
    f.attr("num_converged", int(system.num_converged))
    f.attr("error_flag", int(system.error_flag))
    f.attr("total_steps_done", int(system.total_steps))
    times = system.ctx.phase_times()
    for name, phase in (("propagation_time", "propagate"), ("orthonormalization_time", None),
                        ("dotproduct_time", "overlap"), ("eigensolve_time", "eigensolve"),
                        ("lincomb_time", "lincomb"), ("convtest_time", "energies")):
        if phase is None:
            ms = sum(times[p][0] for p in ("overlap", "eigensolve", "lincomb"))
        else:
            ms = times[phase][0]
        f.attr(name, ms / 1e3)
    f.attr("io_time", 0.0)
    f.attr("total_time", float(system.total_time))
