"""Fixture of the reference's dipolar P3M test
(testsuite/python/dipolar_mdlc_p3m_scafacos_p2nfft.py:115-153): 1000 dipoles
in a cubic box of edge cbrt(4 pi 1000 / (3 * 0.09)), with the expected
forces and torques (per unit prefactor) and the dipolar energy 5.570 per
unit prefactor, which the reference meets with DipolarP3M(mesh=32,
accuracy=1e-6, metallic) at vector errors 2e-3 and energy error 1e-3.

    python tests/golden/make_dp3m_fixture.py

Reads testsuite/python/data/p3m_magnetostatics_{system,expected}.data in the
reference tree and writes p3m_magnetostatics.npz: pos [1000, 3], dip
[1000, 3], force [1000, 3], torque [1000, 3], energy, box_l.
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = "/root/reference/testsuite/python/data"


def main():
    sysd = np.genfromtxt(os.path.join(DATA, "p3m_magnetostatics_system.data"))
    exp = np.genfromtxt(os.path.join(DATA, "p3m_magnetostatics_expected.data"))[:, 1:]
    box_l = (4. * 1000 * math.pi / (3. * 0.09)) ** (1. / 3.)
    np.savez(os.path.join(HERE, "p3m_magnetostatics.npz"), pos=sysd[:, 1:4],
             dip=sysd[:, 4:7], force=exp[:, 0:3], torque=exp[:, 3:6],
             energy=np.float64(5.570), box_l=np.float64(box_l))


if __name__ == "__main__":
    main()
