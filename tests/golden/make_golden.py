"""Writes tests/golden/vacuum_dipole_64.npz: the oracle's (FP64 restatement of the
reference hot loop) DFT plane and probe traces for the 64^3 vacuum-dipole
configuration (8-cell CPML, Ez dipole, Gaussian1 f = 0.1, 1000 steps, Ez plane
at z = 10 with 3 frequencies, Ez probes at (0,0,5), (10,0,0), (0,20,-10)).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ffip_amd import workloads as W  # noqa: E402
from oracle.oracle import OracleSim  # noqa: E402


def main():
    cfg = W.vacuum_dipole()
    sim = OracleSim(cfg, probes=W.VACUUM_DIPOLE_PROBES, trace_steps=1000)
    n = sim.run(cfg["stop condition"])
    d = sim.dft_output(0)
    np.savez_compressed(os.path.join(HERE, "vacuum_dipole_64.npz"), real=d["real"], imag=d["imag"], x=d["x"],
                        y=d["y"], z=d["z"], f=d["f"], trace=sim.trace[:n], steps=n)
    print("steps", n, "dft", d["real"].shape)


if __name__ == "__main__":
    main()
