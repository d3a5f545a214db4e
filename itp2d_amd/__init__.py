"""MI355X (gfx950) imaginary-time-propagation engine for itp2d's hot loop.

Python binding of the C ABI in include/itp2d.h, include/itp2d_system.h and
include/itp2d_batch.h
(library itp2d_amd/lib/libitp2d_amd.so, built by itp2d_amd/build.py). The
classes mirror the reference's hot-path API: Transformer (with the grid),
Potential, StateSet.orthonormalize, MultiProductSplit, Hamiltonian,
EigenSolver, ITPSystem; Batch runs many independent instances together.
There is no CPU fallback: if the library is missing
or no GPU is present, the calls fail.
"""
from .api import (  # noqa: F401
    TRANSFORMS, PERIODIC, DIRICHLET, FP64, FP32, TESTS,
    ITPError, NonPositiveEigenvalue, NonNormalEigenvalue, EigensolverError,
    Context, Grid, Potential, StateSet, MultiProductSplit, Hamiltonian,
    ITPSystem, Batch, init_states, eigh, eigvalsh, install_signal_flags, noalloc_max_bytes, lib, library_path,
)
