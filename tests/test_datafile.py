"""HDF5 result file in the reference's layout (datafile.cpp:32-467).

CPU: the file the writer makes is read back with h5dump (the HDF5 tools'
own reader, independent of this library): dataset names, element types
(the {r, i} compound, the [sy][sx] array types, the {step, index} and
{step, time_step} compounds), extents, values and root attributes; the
reference's error behaviour (no overwrite without clobber, no attribute
written twice). GPU: a C1 run saved by save_run holds the run's sorted final
energies, its step history and its potential.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from itp2d_amd import datafile as D

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

def _lib_loads():
    try:
        D.lib()
        return True
    except (OSError, D.DatafileError):
        return False


needs_lib = pytest.mark.skipif(not _lib_loads(), reason="datafile library not built or no HDF5 runtime to load it")
needs_dump = pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")


def dump(path, *args):
    return subprocess.run([H5DUMP, *args, path], check=True, capture_output=True, text=True).stdout


def numbers(text):
    body = text[text.index("DATA {"):]
    return np.array([float(x) for x in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", re.sub(r"\(\d+(,\d+)*\):", "", body))])


def small_file(path, clobber=False):
    rng = np.random.default_rng(7)
    sx, sy, N = 6, 4, 3
    states = rng.standard_normal((N, sy, sx)) + 1j * rng.standard_normal((N, sy, sx))
    V = rng.uniform(0, 2, (sy, sx))
    E = np.array([0.5, 1.5, 2.5])
    with D.Datafile(path, sx, sy, 0.25, clobber) as f:
        f.attr("num_states", N)
        f.attr("grid_length", 1.5)
        f.attr("grid_boundary_type", "periodic")
        f.attr("error_flag", False)
        f.attr_ulong("random_seed", 2 ** 40 + 3)
        f.write_potential(V)
        f.write_time_step(1, 0.1)
        f.write_time_step(7, 0.02)
        for k in range(3):
            f.write_history_row(0, k, E + 0.1 * (3 - k))
            f.write_history_row(1, k, np.full(N, 0.01 * (k + 1)))
        f.write_states(3, states)
        f.write_states(5, states, order=[2, 0, 1])
        f.write_final(0, E)
        f.write_final(1, [1e-9, 2e-9, 3e-9])
    return states, V, E


@needs_lib
@needs_dump
def test_layout_read_by_h5dump(tmp_path):
    p = str(tmp_path / "run.h5")
    states, V, E = small_file(p)
    head = dump(p, "-H")
    for name in ["states", "state_history", "final_energies", "final_energy_standard_deviations", "energy_history",
                 "deviation_history", "time_step_history", "potential_values"]:
        assert f'DATASET "{name}"' in head, name
    flat = " ".join(head.split())
    assert ('DATASET "states" { DATATYPE H5T_ARRAY { [4][6] H5T_COMPOUND { H5T_IEEE_F64LE "r"; '
            'H5T_IEEE_F64LE "i"; } } DATASPACE SIMPLE { ( 2, 3 ) / ( H5S_UNLIMITED, H5S_UNLIMITED ) }') in flat
    assert ('DATASET "state_history" { DATATYPE H5T_COMPOUND { H5T_STD_I32LE "step"; H5T_STD_I32LE "index"; } '
            'DATASPACE SIMPLE { ( 2 ) / ( H5S_UNLIMITED ) }') in flat
    assert ('DATASET "time_step_history" { DATATYPE H5T_COMPOUND { H5T_STD_I32LE "step"; '
            'H5T_IEEE_F64LE "time_step"; } DATASPACE SIMPLE { ( 2 ) / ( H5S_UNLIMITED ) }') in flat
    assert ('DATASET "potential_values" { DATATYPE H5T_ARRAY { [4][6] H5T_IEEE_F64LE } DATASPACE SCALAR') in flat
    assert 'DATASET "energy_history" { DATATYPE H5T_IEEE_F64LE DATASPACE SIMPLE { ( 3, 3 ) / ( H5S_UNLIMITED, H5S_UNLIMITED ) }' in flat
    assert 'DATASET "final_energies" { DATATYPE H5T_IEEE_F64LE DATASPACE SIMPLE { ( 3 ) / ( H5S_UNLIMITED ) }' in flat
    # values through h5dump, full precision
    fmt = ["-A", "0", "-m", "%.17g"]
    assert np.array_equal(numbers(dump(p, "-d", "/final_energies", *fmt)), E)
    got = numbers(dump(p, "-d", "/energy_history", *fmt)).reshape(3, 3)
    assert np.array_equal(got, np.array([E + 0.1 * (3 - k) for k in range(3)]))
    assert np.array_equal(numbers(dump(p, "-d", "/potential_values", *fmt)), V.ravel())
    assert np.array_equal(numbers(dump(p, "-d", "/time_step_history", *fmt)), [1, 0.1, 7, 0.02])
    assert np.array_equal(numbers(dump(p, "-d", "/state_history", *fmt)), [3, 0, 5, 1])
    s = numbers(dump(p, "-d", "/states", *fmt)).reshape(2, 3, 4, 6, 2)
    c = s[..., 0] + 1j * s[..., 1]
    assert np.array_equal(c[0], states) and np.array_equal(c[1], states[[2, 0, 1]])
    # root attributes
    a = " ".join(dump(p, "-A", "-a", "/random_seed").split())
    assert "H5T_STD_U64LE" in a and str(2 ** 40 + 3) in a
    a = " ".join(dump(p, "-a", "/grid_boundary_type").split())
    assert 'STRSIZE 8' in a and '"periodic"' in a
    for name, want in [("grid_sizex", 6), ("grid_sizey", 4), ("num_states", 3), ("error_flag", 0)]:
        a = " ".join(dump(p, "-a", f"/{name}").split())
        assert "H5T_STD_I32LE" in a and f"(0): {want} " in a + " ", name
    a = " ".join(dump(p, "-a", "/grid_delta").split())
    assert "H5T_IEEE_F64LE" in a and "0.25" in a
    assert "description" in dump(p, "-a", "/states/description")


@needs_lib
def test_reader_and_errors(tmp_path):
    p = str(tmp_path / "run.h5")
    states, V, E = small_file(p)
    assert np.array_equal(D.read_doubles(p, "/final_energies"), E)
    assert np.array_equal(D.read_doubles(p, "/potential_values"), V.ravel())
    assert np.array_equal(D.read_doubles(p, "/final_energy_standard_deviations"), [1e-9, 2e-9, 3e-9])
    with pytest.raises(D.DatafileError):
        D.read_doubles(p, "/states")  # complex data
    with pytest.raises(D.DatafileError):
        D.read_doubles(p, "/no_such")
    # an existing file is not overwritten without clobber (Datafile ctor, H5F_ACC_EXCL)
    with pytest.raises(D.DatafileError):
        D.Datafile(p, 6, 4, 0.25)
    assert np.array_equal(D.read_doubles(p, "/final_energies"), E)
    small_file(p, clobber=True)
    with D.Datafile(str(tmp_path / "b.h5"), 6, 4, 0.25) as f:
        f.attr("x", 1)
        with pytest.raises(D.DatafileError):
            f.attr("x", 2)  # attributes are created once
        with pytest.raises(D.DatafileError):
            f.write_states(1, states, order=[0, 1, 5])
        with pytest.raises(ValueError):
            f.write_potential(np.zeros((6, 4)))
    with pytest.raises(D.DatafileError):
        D.Datafile(str(tmp_path / "no_dir" / "c.h5"), 6, 4, 0.25)


def test_convergence_test_descriptions():
    """The reference's ConvergenceTest descriptions (convergence.hpp:57-68, convergence.cpp:105-206)."""
    d = D.convergence_test_description
    assert d(("none",)) == "none"
    assert d(("onestep",)) == "one-step convergence"
    assert d(("relchange", 1e-9)) == "relative energy change < 1e-09"
    assert d(("abschange", 0.001)) == "absolute energy change < 0.001"
    assert d(("absstdev", 1e-8, 1e-10)) == "absolute energy deviation < 1e-08 or absolute energy deviation change < 1e-10"
    assert d(("relstdev", 0.5, 2.0)) == "relative energy deviation < 0.5 or relative energy deviation change < 2"


@pytest.mark.gpu
@needs_lib
def test_save_run(ctx, tmp_path):
    import itp2d_amd as I
    from itp2d_amd import workloads as W
    print("[gpu] save_run c1_ho64", flush=True)
    cfg, V = W.config("c1_ho64")
    s = I.ITPSystem(ctx, cfg, V)
    assert s.run() == 0
    p = str(tmp_path / "c1.h5")
    D.save_run(s, p, V=V, attributes={"potential": "harmonic"})
    E, sd, _ = s.sorted()
    assert np.array_equal(D.read_doubles(p, "/final_energies"), E)
    assert np.array_equal(D.read_doubles(p, "/final_energy_standard_deviations"), sd)
    h = D.read_doubles(p, "/energy_history").reshape(-1, cfg["N"])
    assert h.shape[0] == s.total_steps and np.array_equal(h[-1], E)
    assert np.array_equal(D.read_doubles(p, "/potential_values"), np.asarray(V, dtype=float).ravel())
    ts = s.time_step_history()
    assert ts[0] == (1, cfg["eps_values"][0]) and all(a[0] < b[0] for a, b in zip(ts, ts[1:]))
