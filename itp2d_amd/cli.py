"""Command line of the reference's itp2d program, running on the gfx950 engine.

`python -m itp2d_amd.cli [options]` accepts the reference's options
(CommandLineParser, commandlineparser.cpp:222-364) with their defaults
(parameters.cpp:25-55), checks and derived values, builds the named
potential on the device, runs the ITP driver and writes the HDF5 result file
in the reference's layout. Exit status as itp2d.cpp:73-134: 0 done, 1 the
run ended with its error flag, 2 invalid arguments, 3 set-up failed, 4 the
run failed.

Options with no effect on this engine, accepted for compatibility:
--highmem-orthonormalization (the rotation is always one out-of-place
product) and --threads (one GPU stream).
--wisdomfile names the file of measured launch shapes, the counterpart of the
reference's FFTW wisdom (itp2d.cpp:85-93, 124-128; plans are measured there
by default, FFTW_PATIENT, itp2d_common.hpp:56): runs tune the overlap's
launch shape on first use of a shape, read the file first if it exists (a
file in another format is ignored with a note, as fftw_import_wisdom_from_file
fails on it) and write the tuned shapes back after the run.
Not supported (an argument error): --noise impurities (the reference draws
impurity realisations with its own generator; pass realisations through
itp_potential_add_impurities instead) and --copy-states.
--save-everything saves the initial states and the sorted states after
every step, as the reference does.
Added: --mixed runs the mixed-precision path (FP32 fields, FP64 finish).
"""
import argparse
import math
import os
import re
import sys

WHITESPACE = " \f\n\r\t\v"

DEFAULTS = dict(rngseed=0x20120131, datafile="data/itp2d.h5", verbosity=1, threads=2, sizex=64, sizey=64, lenx=12.0,
                N=25, potential="harmonic", timestep_test="relstdev(1e-3,1e-4)", final_test="relstdev(1e-3,0)",
                noise="none", impurity_type="N/A", impurity_distribution="N/A", impurity_constraint="none", B=0.0,
                order=10, eps=0.5, eps_divisor=5.0, states=16, ignore_lowest=0, max_steps=50, min_time_step=1e-9)

# convergence-test names and aliases -> (kind, accepted parameter counts) (convergence.cpp:28-53)
CONVERGENCE = {"none": ("none", {0}), "no": ("none", {0}), "null": ("none", {0}),
               "onestep": ("onestep", {0}), "one-step": ("onestep", {0}),
               "abschange": ("abschange", {1}), "absEdelta": ("abschange", {1}),
               "relchange": ("relchange", {1}), "relEdelta": ("relchange", {1}),
               "absstdev": ("absstdev", {1, 2}), "absresidual": ("absstdev", {1, 2}),
               "relstdev": ("relstdev", {1, 2}), "relresidual": ("relstdev", {1, 2})}


class CLIError(ValueError):
    """An invalid command line (exit status 2)."""


def parse_parameter_string(s):
    """`name` or `name(v1, v2, ...)` -> (name, [values]) (parser.cpp:26-59)."""
    t = s.strip(WHITESPACE)
    if not t:
        raise CLIError(f"cannot parse '{s}'")
    m = re.match(r"[^" + re.escape(WHITESPACE) + r"(]*", t)
    name, rest = m.group(0), t[m.end():].strip(WHITESPACE)
    if not rest:
        return name, []
    if not (rest.startswith("(") and rest.endswith(")")):
        raise CLIError(f"cannot parse '{s}'")
    inner = rest[1:-1]
    if not inner:
        return name, []
    values = []
    for part in inner.split(","):
        try:
            values.append(float(part.strip(WHITESPACE)))
        except ValueError:
            raise CLIError(f"cannot parse '{s}'") from None
    return name, values


def parse_convergence_test(s):
    """A convergence-test description -> the driver's test tuple (convergence.cpp:28-206)."""
    name, p = parse_parameter_string(s)
    if name not in CONVERGENCE:
        raise CLIError(f"unknown convergence test '{s}'")
    kind, counts = CONVERGENCE[name]
    if len(p) not in counts:
        raise CLIError(f"convergence test '{s}' takes {' or '.join(str(c) for c in sorted(counts))} parameters")
    if kind in ("absstdev", "relstdev") and len(p) == 1:
        p = p + [0.0]
    return (kind, *p)


def argument_parser():
    a = argparse.ArgumentParser(prog="itp2d", description="Imaginary-time propagation of 2D single-particle states "
                                "on an MI355X GPU (the reference itp2d's options).")
    add = a.add_argument
    add("--highmem-orthonormalization", action="store_true", help="accepted; no effect here")
    add("--wisdomfile", default="fftw_wisdom", help="file of measured launch shapes (read if present, "
        "written after the run)")
    add("--noise", default=DEFAULTS["noise"])
    add("--impurity-type", default=None)
    add("--impurity-distribution", default=None)
    add("--impurity-constraint", default=DEFAULTS["impurity_constraint"])
    add("--recover", action="store_true")
    add("--rngseed", type=int, default=None)
    add("--mineps", type=float, default=DEFAULTS["min_time_step"])
    add("--maxsteps", type=int, default=DEFAULTS["max_steps"])
    add("--exhaust-timestep-list", action="store_true")
    add("-D", "--timestep-divisor", type=float, default=DEFAULTS["eps_divisor"])
    add("-e", "--timestep", type=float, action="append", default=None)
    add("-d", "--order", type=int, default=DEFAULTS["order"])
    add("-N", "--totalstates", type=int, default=None)
    add("--dirichlet", action="store_true")
    add("--pi", action="store_true")
    add("-l", "--lenx", type=float, default=DEFAULTS["lenx"])
    add("-y", "--sizey", type=int, default=None)
    add("-x", "--sizex", type=int, default=None)
    add("-s", "--size", type=int, default=None)
    add("-B", "--magneticfield", type=float, default=DEFAULTS["B"])
    add("--ignore-lowest", type=int, default=DEFAULTS["ignore_lowest"])
    add("-n", "--states", type=int, default=None)
    add("-t", "--threads", type=int, default=DEFAULTS["threads"], help="accepted; no effect here")
    add("-q", "--quiet", action="count", default=0)
    add("-v", "--verbose", action="count", default=0)
    add("--save-nothing", action="store_true")
    add("--save-only-energies", action="store_true")
    add("--save-everything", action="store_true")
    add("-f", "--force", action="store_true")
    add("--copy-states", default=None)
    add("-o", "--datafile", default=DEFAULTS["datafile"])
    add("-T", "--timestep-convergence-test", default=DEFAULTS["timestep_test"])
    add("-F", "--final-convergence-test", default=DEFAULTS["final_test"])
    add("-p", "--potential", default=DEFAULTS["potential"])
    add("--mixed", action="store_true", help="mixed-precision path (FP32 fields, FP64 finish); not in the reference")
    return a


def _positive(name, v):
    if v is not None and not v > 0:
        raise CLIError(f"{name} has to be positive")


def parse(argv):
    """Command line -> (driver configuration, run options) with the reference's
    checks and derived values (commandlineparser.cpp:264-364)."""
    a = argument_parser()
    try:
        ns = a.parse_args(argv)
    except SystemExit as e:
        raise CLIError("invalid arguments") if e.code else e
    if sum([ns.save_nothing, ns.save_only_energies, ns.save_everything]) > 1:
        raise CLIError("--save-nothing, --save-only-energies and --save-everything cannot be set together")
    _positive("--threads", ns.threads)
    if ns.size is not None and (ns.sizex is not None or ns.sizey is not None):
        raise CLIError("--size and (--sizey or --sizex) cannot be set together")
    for name, v in (("--size", ns.size), ("--sizex", ns.sizex), ("--sizey", ns.sizey), ("--lenx", ns.lenx),
                    ("--timestep-divisor", ns.timestep_divisor), ("--totalstates", ns.totalstates),
                    ("--states", ns.states), ("--order", ns.order), ("--maxsteps", ns.maxsteps)):
        _positive(name, v)
    for e in ns.timestep or []:
        _positive("--timestep", e)
    n_wanted = ns.states if ns.states is not None else (ns.totalstates if ns.totalstates is not None
                                                         else DEFAULTS["states"])
    if ns.totalstates is not None and ns.states is not None and ns.totalstates < ns.ignore_lowest + ns.states:
        raise CLIError("Number of states has to be large enough to include at least all the states wanted to converge!")
    if ns.order % 2:
        raise CLIError("--order has to be even")
    if ns.mineps < 0:
        raise CLIError("--mineps cannot be negative")
    # unsigned in the reference (size_t, unsigned long)
    if ns.ignore_lowest < 0 or (ns.rngseed is not None and ns.rngseed < 0):
        raise CLIError("--ignore-lowest and --rngseed cannot be negative")
    if ns.noise == "impurities":
        raise CLIError("--noise impurities: the reference's random impurity realisations are not reproduced; "
                       "pass realisations through itp_potential_add_impurities")
    if ns.noise != "none":
        raise CLIError(f"unknown noise type '{ns.noise}'")
    if ns.impurity_type is not None or ns.impurity_distribution is not None:
        raise CLIError("--impurity-type and --impurity-distribution have no effect if noise type is not 'impurities'")
    if ns.copy_states is not None:
        raise CLIError("--copy-states is not supported")
    # derived values: N is 1.25 times the wanted count when only that is given
    if ns.states is not None and ns.totalstates is None:
        N = int(ns.states * 1.25)
    else:
        N = ns.totalstates if ns.totalstates is not None else DEFAULTS["N"]
    sx = ns.size if ns.size is not None else (ns.sizex if ns.sizex is not None else DEFAULTS["sizex"])
    sy = ns.size if ns.size is not None else (ns.sizey if ns.sizey is not None else DEFAULTS["sizey"])
    seed = ns.rngseed if ns.rngseed is not None else int.from_bytes(os.urandom(4), "little")
    cfg = dict(sx=sx, sy=sy, lenx=(math.pi if ns.pi else 1.0) * ns.lenx, boundary=1 if ns.dirichlet else 0, N=N,
               n_wanted=n_wanted, ignore_lowest=ns.ignore_lowest, halforder=ns.order // 2, B=ns.magneticfield,
               eps_values=list(ns.timestep or []), eps_divisor=ns.timestep_divisor,
               exhaust_eps=int(ns.exhaust_timestep_list), max_steps=ns.maxsteps, min_time_step=ns.mineps, seed=seed,
               recover=int(ns.recover), timestep_test=parse_convergence_test(ns.timestep_convergence_test),
               final_test=parse_convergence_test(ns.final_convergence_test))
    save = ("nothing" if ns.save_nothing else "energies" if ns.save_only_energies
            else "everything" if ns.save_everything else "final_states")
    pname, pparams = parse_parameter_string(ns.potential)
    from .api import potential_type_params, ITPError
    try:
        potential_type_params(pname, pparams)
    except ITPError as e:
        raise CLIError(f"invalid potential '{ns.potential}': {e}") from None
    opts = dict(save=save, datafile=ns.datafile, clobber=ns.force, verbosity=1 + ns.verbose - ns.quiet,
                potential=(pname, pparams), potential_string=ns.potential, mixed=ns.mixed,
                impurity_constraint=ns.impurity_constraint, threads=ns.threads, wisdomfile=ns.wisdomfile)
    return cfg, opts


def run(cfg, opts, device=0):
    """Builds the potential, runs the driver and saves; returns the exit status."""
    from . import api as I
    from . import datafile
    ctx = None
    try:
        ctx = I.Context(device)
        ctx.set_autotune(1)
        load_wisdom(ctx, opts.get("wisdomfile"), opts["verbosity"])
        dx = cfg["lenx"] / cfg["sx"]
        g = I.Grid(ctx, cfg["sx"], cfg["sy"], dx, cfg["boundary"])
        name, params = opts["potential"]
        V = None if name == "zero" else I.Potential.of_type(g, name, params).values()
        c = dict(cfg, precision=I.FP32 if opts["mixed"] else I.FP64)
        s = I.ITPSystem(ctx, c, V)
        s.set_verbosity(max(0, min(3, opts["verbosity"])))
    except Exception as e:  # set-up failed (itp2d.cpp:98-106)
        print(f"Error during initialization: {e}", file=sys.stderr)
        if ctx is not None:
            ctx.close()
        return 3
    attrs = {"potential": opts["potential_string"], "noise": "none", "impurity_type": "N/A",
             "impurity_distribution": "N/A", "impurity_constraint": opts["impurity_constraint"],
             "num_threads": int(opts["threads"])}
    try:
        if opts["save"] != "nothing":
            d = os.path.dirname(opts["datafile"])
            if d and not os.path.isdir(d):
                raise OSError(f"directory '{d}' does not exist")
        # SIGINT finishes the run at the next step with the error flag, SIGUSR1 saves the states at the
        # next check point (itp2d.cpp:40-66, itpsystem.cpp:235-241, 407-416)
        abort_flag, save_flag = I.install_signal_flags()
        s.set_abort_flag(abort_flag)
        if opts["save"] == "everything":
            datafile.run_saving_every_step(s, opts["datafile"], V=V, clobber=opts["clobber"], attributes=attrs,
                                           save_flag=save_flag)
        elif opts["save"] != "nothing":
            if os.path.exists(opts["datafile"]) and not opts["clobber"]:
                raise OSError(f"datafile '{opts['datafile']}' exists (use --force to overwrite)")
            datafile.run_and_save(s, opts["datafile"], V=V, clobber=opts["clobber"], attributes=attrs,
                                  save="energies" if opts["save"] == "energies" else "final_states",
                                  save_flag=save_flag)
        else:
            s.set_save_flag(save_flag, lambda sys_: None)  # no result file: the flag is only cleared
            s.run()
    except Exception as e:  # the run failed (itp2d.cpp:113-121)
        print(f"Error during the run: {e}", file=sys.stderr)
        ctx.close()
        return 4
    status = 1 if s.error_flag else 0
    save_wisdom(ctx, opts.get("wisdomfile"))
    ctx.close()  # frees the run's device objects before the interpreter exits
    return status


def load_wisdom(ctx, path, verbosity=1):
    """Adds the measured launch shapes of `path` if it exists (itp2d.cpp:86-92);
    a file this engine did not write is left unused, with a note."""
    if not path or not os.path.isfile(path):
        return
    from .api import ITPError
    try:
        with open(path) as f:
            ctx.wisdom_import(f.read())
    except (ITPError, OSError, UnicodeDecodeError) as e:
        if verbosity > 0:
            print(f"itp2d: wisdom file '{path}' not used ({e})", file=sys.stderr)


def save_wisdom(ctx, path):
    """Writes the process's measured launch shapes to `path` (itp2d.cpp:124-126)."""
    if not path:
        return
    try:
        with open(path, "w") as f:
            f.write(ctx.wisdom_export())
    except OSError as e:
        print(f"itp2d: cannot write wisdom file '{path}': {e}", file=sys.stderr)


def main(argv=None):
    try:
        cfg, opts = parse(sys.argv[1:] if argv is None else argv)
    except CLIError as e:
        print(f"itp2d: {e}", file=sys.stderr)
        return 2
    return run(cfg, opts)


if __name__ == "__main__":
    sys.exit(main())
