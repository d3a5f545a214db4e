"""The command line (commandlineparser.cpp:222-364, parser.cpp:26-59,
convergence.cpp:28-206): defaults, derived values and checks on the CPU; on
the GPU, a run through the command line matches the same run through the
API bit for bit and writes the result file in every save mode."""
import math

import numpy as np
import pytest

from itp2d_amd import cli


def test_defaults():
    cfg, opts = cli.parse(["--rngseed", "5"])
    assert (cfg["sx"], cfg["sy"], cfg["lenx"], cfg["N"], cfg["n_wanted"], cfg["halforder"]) == (64, 64, 12.0, 25, 16, 5)
    assert cfg["boundary"] == 0 and cfg["B"] == 0.0 and cfg["eps_values"] == [] and cfg["eps_divisor"] == 5.0
    assert cfg["max_steps"] == 50 and cfg["min_time_step"] == 1e-9 and cfg["seed"] == 5 and cfg["recover"] == 0
    assert cfg["timestep_test"] == ("relstdev", 1e-3, 1e-4) and cfg["final_test"] == ("relstdev", 1e-3, 0.0)
    assert opts["save"] == "final_states" and opts["datafile"] == "data/itp2d.h5" and opts["verbosity"] == 1
    assert opts["potential"] == ("harmonic", [])


def test_derived_values():
    cfg, _ = cli.parse(["-n", "40"])
    assert cfg["N"] == 50 and cfg["n_wanted"] == 40          # N = 1.25 n when only n is given
    cfg, _ = cli.parse(["-N", "30"])
    assert cfg["N"] == 30 and cfg["n_wanted"] == 30          # n = N when only N is given
    cfg, _ = cli.parse(["-N", "30", "-n", "20", "--ignore-lowest", "4"])
    assert (cfg["N"], cfg["n_wanted"], cfg["ignore_lowest"]) == (30, 20, 4)
    cfg, _ = cli.parse(["-s", "128", "--pi", "-l", "2", "--dirichlet", "-B", "1.5", "-d", "4"])
    assert (cfg["sx"], cfg["sy"], cfg["boundary"], cfg["B"], cfg["halforder"]) == (128, 128, 1, 1.5, 2)
    assert cfg["lenx"] == 2 * math.pi
    cfg, opts = cli.parse(["-x", "96", "-y", "48", "-e", "0.1", "-e", "0.02", "--exhaust-timestep-list", "-vv", "-q",
                           "--save-only-energies", "-f", "-o", "x.h5", "--recover", "--maxsteps", "7", "--mineps", "0"])
    assert (cfg["sx"], cfg["sy"], cfg["eps_values"], cfg["exhaust_eps"], cfg["recover"]) == (96, 48, [0.1, 0.02], 1, 1)
    assert (opts["verbosity"], opts["save"], opts["clobber"], opts["datafile"]) == (2, "energies", True, "x.h5")
    assert (cfg["max_steps"], cfg["min_time_step"]) == (7, 0.0)
    assert cli.parse(["--save-everything"])[1]["save"] == "everything"
    assert cli.parse(["--save-nothing"])[1]["save"] == "nothing"
    assert cli.parse(["--highmem-orthonormalization", "-t", "8", "--wisdomfile", "w"])[0]["N"] == 25


@pytest.mark.parametrize("argv", [
    ["--save-nothing", "--save-everything"], ["-s", "64", "-x", "32"], ["-d", "3"], ["-d", "0"], ["-N", "10", "-n", "11"],
    ["-N", "10", "-n", "8", "--ignore-lowest", "3"], ["-e", "0"], ["-e", "-0.1"], ["-D", "0"], ["-l", "-1"],
    ["-t", "0"], ["--maxsteps", "0"], ["--mineps", "-1"], ["-T", "relchange"], ["-T", "relstdev(1,2,3)"],
    ["-F", "nosuchtest"], ["-p", "nosuchpotential"], ["-p", "harmonic(1,2)"], ["-p", "harmonic 1"],
    ["--noise", "impurities", "--impurity-type", "gaussian", "--impurity-distribution", "uniform(1)"],
    ["--impurity-type", "gaussian"], ["--noise", "unknown"], ["--copy-states", "a.h5"], ["--no-such-option"],
    ["-s", "x"], ["--ignore-lowest", "-1"], ["--rngseed", "-5"]])
def test_rejected(argv):
    with pytest.raises(cli.CLIError):
        cli.parse(argv)
    assert cli.main(argv) == 2


@pytest.mark.parametrize("s,want", [
    ("harmonic", ("harmonic", [])), ("  harmonic  ", ("harmonic", [])), ("harmonic()", ("harmonic", [])),
    ("gaussian(1, 2.5,-3,4e-2)", ("gaussian", [1.0, 2.5, -3.0, 0.04])), ("ring (3,1,2)", ("ring", [3.0, 1.0, 2.0])),
    ("relstdev(1e-3,1e-4)", ("relstdev", [1e-3, 1e-4]))])
def test_parameter_strings(s, want):
    assert cli.parse_parameter_string(s) == want


@pytest.mark.parametrize("s", ["", "   ", "gaussian(1,2", "gaussian 1", "gaussian(1,x)", "gaussian(1,,2)"])
def test_parameter_strings_rejected(s):
    with pytest.raises(cli.CLIError):
        cli.parse_parameter_string(s)


@pytest.mark.parametrize("s,want", [
    ("none", ("none",)), ("no", ("none",)), ("null()", ("none",)), ("onestep", ("onestep",)),
    ("one-step", ("onestep",)), ("abschange(1e-6)", ("abschange", 1e-6)), ("absEdelta(2)", ("abschange", 2.0)),
    ("relchange(1e-9)", ("relchange", 1e-9)), ("relEdelta(3)", ("relchange", 3.0)),
    ("absstdev(1e-8)", ("absstdev", 1e-8, 0.0)), ("absresidual(1,2)", ("absstdev", 1.0, 2.0)),
    ("relstdev(1e-3,1e-4)", ("relstdev", 1e-3, 1e-4)), ("relresidual(0.5)", ("relstdev", 0.5, 0.0))])
def test_convergence_strings(s, want):
    assert cli.parse_convergence_test(s) == want


@pytest.mark.gpu
def test_cli_run_matches_api(ctx, tmp_path):
    """A run through the command line equals the same run through the API bit
    for bit (both with launch-shape tuning, so both use the tuned shape), its
    result file holds that run's energies in each save mode, and the wisdom
    file it writes holds the tuned shape."""
    import itp2d_amd as I
    from itp2d_amd import datafile as D
    print("[gpu] command-line run", flush=True)
    argv = ["-s", "32", "-l", "10", "-n", "6", "-N", "8", "-d", "2", "-e", "0.1", "-p", "harmonic(1.5)",
            "-T", "relchange(1e-10)", "-F", "relchange(1e-10)", "--maxsteps", "2000", "--rngseed", "11", "-q",
            "--wisdomfile", str(tmp_path / "wisdom")]
    cfg, opts = cli.parse(argv)
    g = I.Grid(ctx, 32, 32, 10 / 32, 0)
    V = I.Potential.of_type(g, "harmonic", [1.5]).values()
    ctx.set_autotune(1)
    try:
        s = I.ITPSystem(ctx, dict(cfg, precision=I.FP64), V)
        assert s.run() == 0 and not s.error_flag
    finally:
        ctx.set_autotune(0)
    E, sd, _ = s.sorted()
    for mode in ([], ["--save-only-energies"], ["--save-everything"]):
        p = str(tmp_path / f"run{len(mode)}{'e' if '--save-everything' in mode else ''}.h5")
        assert cli.main(argv + mode + ["-o", p]) == 0
        assert np.array_equal(D.read_doubles(p, "/final_energies"), E)
        assert np.array_equal(D.read_doubles(p, "/final_energy_standard_deviations"), sd)
        assert D.read_doubles(p, "/energy_history").size == s.total_steps * 8
    # an existing datafile is kept without --force, overwritten with it
    assert cli.main(argv + ["-o", p]) == 4
    assert cli.main(argv + ["-o", p, "-f"]) == 0
    assert cli.main(argv + ["--save-nothing", "-o", str(tmp_path / "none.h5")]) == 0
    assert not (tmp_path / "none.h5").exists()
    assert any(line.split()[:4] == ["gram", "8", "1024", "64"] for line in (tmp_path / "wisdom").read_text().splitlines())


@pytest.mark.gpu
def test_cli_run_matches_oracle_golden(tmp_path):
    """The c1 workload (64x64 periodic oscillator on [-8,8]^2, N = 20, 16
    wanted, 2nd-order split, eps 0.1, relchange(1e-12) tests) through the
    command line: the result file written by run_and_save holds the oracle
    golden's step count (energy history rows) and sorted final energies to
    1e-10 (tests/golden/itp_c1_ho64.npz)."""
    import os
    from itp2d_amd import datafile as D
    from itp2d_amd import workloads as W
    print("[gpu] command-line run against the c1 golden", flush=True)
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "itp_c1_ho64.npz"))
    cfg, _ = W.config("c1_ho64")
    p = str(tmp_path / "c1.h5")
    argv = ["-s", "64", "-l", "16", "-n", "16", "-N", "20", "-d", "2", "-e", "0.1", "-D", "5", "-p", "harmonic",
            "-T", "relchange(1e-12)", "-F", "relchange(1e-12)", "--maxsteps", "2000", "--mineps", "1e-9",
            "--rngseed", str(cfg["seed"]), "-q", "--wisdomfile", str(tmp_path / "wisdom"), "-o", p]
    assert cli.main(argv) == 0
    N = cfg["N"]
    assert D.read_doubles(p, "/energy_history").size == int(ref["steps"]) * N
    E = D.read_doubles(p, "/final_energies")
    nw = cfg["n_wanted"]
    assert np.abs(E[:nw] - ref["E"][:nw]).max() / np.abs(ref["E"][:nw]).max() < 1e-10
    assert D.read_doubles(p, "/final_energy_standard_deviations").size == N


@pytest.mark.gpu
def test_cli_wisdom_file_of_another_format_is_replaced(tmp_path, capsys):
    """A wisdom file this engine did not write (FFTW's own format, as the
    reference leaves it, itp2d.cpp:86-92) is not used: the run goes on with a
    note and the file is written back with the tuned shapes."""
    print("[gpu] command-line run with an FFTW wisdom file", flush=True)
    w = tmp_path / "fftw_wisdom"
    w.write_text("(fftw-3.3.10 fftw_wisdom #x4be12fff #x7b2df9b2 #xa9bdb5b5 #xc8a0e8ba\n)\n")
    argv = ["-s", "32", "-l", "10", "-n", "4", "-N", "6", "-e", "0.1", "-p", "harmonic", "-T", "relchange(1e-8)",
            "-F", "relchange(1e-8)", "--maxsteps", "500", "--rngseed", "3", "--save-nothing", "--wisdomfile", str(w)]
    assert cli.main(argv) == 0
    assert "not used" in capsys.readouterr().err
    lines = w.read_text().splitlines()
    assert lines and all(line.split()[0] == "gram" and len(line.split()) == 5 for line in lines)


def test_signal_flags_installed_like_the_reference():
    """SIGUSR1 raises the save flag and SIGINT the abort flag, each with the
    reference's note on stderr (itp2d.cpp:40-66); run in a child process so
    the handlers stay out of the test runner."""
    import subprocess
    import sys
    code = ("import os, signal, time; from itp2d_amd import api as I; a, s = I.install_signal_flags(); "
            "os.kill(os.getpid(), signal.SIGUSR1); time.sleep(0.2); print('save', s.value, 'abort', a.value); "
            "os.kill(os.getpid(), signal.SIGINT); time.sleep(0.2); print('save', s.value, 'abort', a.value)")
    root = __import__("os").path.dirname(__import__("os").path.dirname(__import__("os").path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split("\n")[:2] == ["save 1 abort 0", "save 1 abort 1"]
    assert "Caught SIGUSR1. Saving states at next convenient stop and continuing." in r.stderr
    assert "Caught SIGINT. Saving data and quitting at next convenient spot." in r.stderr
