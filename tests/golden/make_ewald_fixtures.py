"""Exact FP64 Coulomb forces on particle subsets of two seeded workloads,
written as fixtures for the accuracy tests and the accuracy table of bench.py.

    python tests/golden/make_ewald_fixtures.py      (a few minutes, all cores)

The total force of a periodic point-charge system (metallic boundary) does not
depend on the Ewald splitting, so each subset is computed once with its own
alpha_e: real space with the exact erfc within r_e (minimum image) plus the
reciprocal sum within k_max, with alpha_e r_e = k_max / (2 alpha_e) = 5.6,
which puts both truncations near exp(-31) relative.

  ewald_c2_subset.npz  N = 10^6, L = 125.99 (seed 1002, 10 % uncharged,
                       shuffled): the first 2000 charged particles
  ewald_c3_subset.npz  N = 250000, L = 79.37 (seed 1003, +1/-1 alternating):
                       the first 2000 particles
Prefactor 1.  Arrays: idx (subset indices), force [2000, 3].
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from tests import p3m_systems as S  # noqa: E402

PRODUCT = 5.6  # alpha_e * r_e and k_max / (2 alpha_e)


def exact_forces(pos, q, L, idx, alpha_e):
    r_e = PRODUCT / alpha_e
    assert r_e < L / 2, "real-space cutoff must stay within the minimum image"
    kmax = 2. * PRODUCT * alpha_e
    order = np.concatenate([idx, np.setdiff1d(np.arange(len(q)), idx)])
    p, qq = pos[order], q[order]
    n = len(idx)
    f_r = O.real_space_subset(p, qq, [L] * 3, 1.0, alpha_e, r_e, n, exact_erfc=True)
    _, f_k = O.ewald_kspace(p, qq, [L] * 3, 1.0, alpha_e, kmax, n_out=n)
    return f_r + f_k, dict(alpha_e=alpha_e, r_e=r_e, kmax=kmax)


def main():
    pos, q, box, _ = S.config_c2()
    idx = np.nonzero(q)[0][:2000]
    t = time.time()
    f, meta = exact_forces(pos, q, box[0], idx, 0.095)
    np.savez(os.path.join(HERE, "ewald_c2_subset.npz"), idx=idx, force=f, **meta)
    print(f"C2 subset done in {time.time() - t:.0f} s, rms |F| {np.sqrt(np.mean(np.sum(f**2, 1))):.4f}")
    pos, q, box = S.config_c3()
    idx = np.arange(2000)
    t = time.time()
    f, meta = exact_forces(pos, q, box[0], idx, 0.145)
    np.savez(os.path.join(HERE, "ewald_c3_subset.npz"), idx=idx, force=f, **meta)
    print(f"C3 subset done in {time.time() - t:.0f} s, rms |F| {np.sqrt(np.mean(np.sum(f**2, 1))):.4f}")


if __name__ == "__main__":
    main()
