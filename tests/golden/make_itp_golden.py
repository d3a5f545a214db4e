"""Runs the CPU oracle's ITP loop on pinned workloads and writes
tests/golden/itp_<name>.npz (sorted energies, deviations, converged flags,
step and converged-state counts) for the GPU parity tests.

Names are workloads.config names, or c4_<i> for instance i of the batched
configuration (workloads.c4_instances).
Run: python tests/golden/make_itp_golden.py [name ...]
(the full-size c2_dirichlet256 takes about 25 minutes on 8 cores)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as O  # noqa: E402
from itp2d_amd import workloads as W  # noqa: E402


def workload(name):
    if name.startswith("c4_"):
        return W.c4_instances(64)[int(name[3:])]
    return W.config(name)


def main(names=("c1_ho64", "c2_small")):
    for name in names:
        cfg, V = workload(name)
        s = O.System(cfg, V)
        s.run()
        E, sd, idx = s.sorted()
        np.savez_compressed(os.path.join(HERE, f"itp_{name}.npz"), E=E, sigma=sd, steps=s.total_steps,
                            converged=s.num_converged, error=int(s.error_flag), eps=s.eps,
                            conv_flags=s.converged().astype(np.int8))
        print(name, s.total_steps, s.num_converged, s.error_flag, flush=True)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]) or ("c1_ho64", "c2_small"))
