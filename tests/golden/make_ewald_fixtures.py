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
  ewald_c4_sample.npz  16 of the 512 batched systems (bench.make_c4: seeds
                       4000 + s, N log-uniform in [500, 20000], density
                       U[0.1, 0.8]) spanning N: the first min(N, 500)
                       particles of each
Prefactor 1.  Arrays: idx (subset indices), force [2000, 3]; for the C4
sample sys (system ids), count (particles per system), force [sum, 3].
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


def c4_sample_ids(ns, k=16):
    """k systems spread over the sorted N (both ends included)."""
    order = np.argsort(ns, kind="stable")
    return sorted({int(order[int(round(j * (len(ns) - 1) / (k - 1)))]) for j in range(k)})


def c4_sample():
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    import bench
    systems = bench.make_c4()
    ids = c4_sample_ids(np.array([s["n"] for s in systems]))
    forces, counts = [], []
    t = time.time()
    for i in ids:
        s = systems[i]
        L = s["L"]
        m = min(s["n"], 500)
        # alpha_e r_e = 5.6 with r_e just inside the minimum image
        alpha_e = PRODUCT / (0.49 * L)
        f, _ = exact_forces(s["pos"], s["q"], L, np.arange(m), alpha_e)
        forces.append(f)
        counts.append(m)
    np.savez(os.path.join(HERE, "ewald_c4_sample.npz"), sys=np.array(ids),
             count=np.array(counts), force=np.concatenate(forces))
    print(f"C4 sample ({len(ids)} systems) done in {time.time() - t:.0f} s")


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "c4":
        c4_sample()
        return
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
