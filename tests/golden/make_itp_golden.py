"""Runs the CPU oracle's ITP loop on the small pinned workloads and writes
tests/golden/itp_<name>.npz (sorted energies, deviations, step and
converged-state counts) for the GPU parity tests.
Run: python tests/golden/make_itp_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as O  # noqa: E402
from itp2d_amd import workloads as W  # noqa: E402


def main(names=("c1_ho64", "c2_small")):
    for name in names:
        cfg, V = W.config(name)
        s = O.System(cfg, V)
        s.run()
        E, sd, idx = s.sorted()
        np.savez_compressed(os.path.join(HERE, f"itp_{name}.npz"), E=E, sigma=sd, steps=s.total_steps,
                            converged=s.num_converged, error=int(s.error_flag), eps=s.eps)
        print(name, s.total_steps, s.num_converged, s.error_flag)


if __name__ == "__main__":
    main()
