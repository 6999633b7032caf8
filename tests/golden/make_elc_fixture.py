"""Fixture of the reference's slab test (testsuite/python/
coulomb_mixed_periodicity.py:28-98): 285 charges in a 10 x 10 x 12 box,
all below z = 9, with total forces and the total energy 216.640984711 of an
MMM2D calculation (prefactor 1), which the reference's ELC + P3M test meets
at 1e-5 (forces) and 1e-4 (energy).

    python tests/golden/make_elc_fixture.py

Reads testsuite/python/data/coulomb_mixed_periodicity_system.data in the
reference tree (columns: id, x, y, z, q, fx, fy, fz) and writes
coulomb_mixed_periodicity.npz: pos [285, 3], q [285], force [285, 3],
energy (the value the reference's test asserts).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/testsuite/python/data/coulomb_mixed_periodicity_system.data"
ENERGY = 216.640984711  # coulomb_mixed_periodicity.py:22


def main():
    data = np.genfromtxt(SRC)
    np.savez(os.path.join(HERE, "coulomb_mixed_periodicity.npz"),
             pos=data[:, 1:4], q=data[:, 4], force=data[:, 5:8],
             energy=np.float64(ENERGY))


if __name__ == "__main__":
    main()
