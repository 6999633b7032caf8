"""Seeded synthetic systems and lattice fixtures shared by the tests.

Workloads (all positions uniform in the periodic box, numpy PCG64 seeds):
  C1  N=10,000, L=21.544, q=+1/-1 alternating, prefactor 2, mesh 32, cao 5,
      r_cut 3.0, alpha from the closed-form tuning rule at accuracy 1e-4
  C2  N=1,000,000, L=125.99, 450k +1 / 450k -1 / 100k 0 (shuffled),
      prefactor 1, mesh 128, cao 7, r_cut 2.5, accuracy 1e-5
  C3  N=250,000, L=79.37, +1/-1 alternating, mesh {48,64,96} x cao {3,5,7}
"""
import itertools
import math

import numpy as np


def closed_form_alpha(prefactor, r_cut, box_l, n_q, sum_q2, accuracy):
    """alpha_L of CoulombTuningAlgorithm::calculate_accuracy
    (src/core/electrostatics/p3m.cpp:636-656) at fixed r_cut, unscaled."""
    L = np.asarray(box_l, float)
    V = float(np.prod(L))
    r_cut_iL = r_cut / L[0]
    rs0 = 2. * prefactor * sum_q2 / math.sqrt(n_q * r_cut_iL * L[0] * V)
    if math.sqrt(2.) * rs0 > accuracy:
        alpha_L = math.sqrt(math.log(math.sqrt(2.) * rs0 / accuracy)) / r_cut_iL
    else:
        alpha_L = 0.1
    return alpha_L / L[0]


def uniform_system(n, L, seed, charges="alternating", neutral_zero_frac=0.0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0., L, (n, 3))
    if charges == "alternating":
        q = np.where(np.arange(n) % 2 == 0, 1., -1.)
    else:
        n0 = int(round(neutral_zero_frac * n))
        npos = (n - n0) // 2
        q = np.concatenate([np.ones(npos), -np.ones(n - n0 - npos), np.zeros(n0)])
        q = q[rng.permutation(n)]
    return pos, q


def config_c1():
    L = 21.544
    pos, q = uniform_system(10_000, L, 1001)
    alpha = closed_form_alpha(2.0, 3.0, [L] * 3, len(q), float(np.sum(q * q)), 1e-4)
    params = dict(prefactor=2.0, accuracy=1e-4, mesh=32, cao=5, r_cut=3.0,
                  alpha=alpha)
    return pos, q, [L] * 3, params


def config_c2(n=1_000_000):
    L = 125.99 * (n / 1_000_000) ** (1. / 3.)
    pos, q = uniform_system(n, L, 1002, charges="shuffled",
                            neutral_zero_frac=0.1)
    nq = int(np.count_nonzero(q))
    alpha = closed_form_alpha(1.0, 2.5, [L] * 3, nq, float(np.sum(q * q)), 1e-5)
    params = dict(prefactor=1.0, accuracy=1e-5, mesh=128, cao=7, r_cut=2.5,
                  alpha=alpha)
    return pos, q, [L] * 3, params


def config_c3():
    """N = 250,000 +1/-1 alternating charges in L = 79.37 (seed 1003)."""
    L = 79.37
    pos, q = uniform_system(250_000, L, 1003)
    return pos, q, [L] * 3


def madelung_cube(n=6):
    """NaCl cube of testsuite/python/p3m_madelung.py:279-312."""
    box = [2. * n] * 3
    idx = list(itertools.product(range(2 * n), repeat=3))
    pos = np.array(idx, float)
    q = np.array([(-1.) ** (j + k + l) for j, k, l in idx])
    params = dict(prefactor=1., accuracy=1e-9, mesh=72, r_cut=5.6, cao=7,
                  alpha=0.8163977)
    return pos, q, box, params, -1.74756459463318219, [1., 1., 1.]


def madelung_wire(n_pairs=64):
    """1D ionic wire of p3m_madelung.py:202-234."""
    box = [n_pairs / 2, n_pairs / 2, 2 * n_pairs]
    pos, q = [], []
    for i in range(n_pairs):
        pos.append([0., 0., 2. * i]); q.append(1.)
        pos.append([0., 0., 2. * i + 1.]); q.append(-1.)
    params = dict(prefactor=1., accuracy=1e-8, mesh=32, r_cut=14., cao=7,
                  alpha=0.27362958)
    return np.array(pos), np.array(q), box, params, -2. * np.log(2.), [0., 0., 1.]


def madelung_sheet(n_pairs=24):
    """2D ionic sheet of p3m_madelung.py:236-277."""
    box = [2 * n_pairs, 2 * n_pairs, 4 * n_pairs]
    pos, q = [], []
    for j, k in itertools.product(range(2 * n_pairs), repeat=2):
        pos.append([j, k, 0.]); q.append((-1.) ** (j + k))
    params = dict(prefactor=1., accuracy=1e-8, mesh=48, r_cut=22., cao=7,
                  alpha=0.179719)
    return (np.array(pos, float), np.array(q), box, params,
            -1.6155426267128247, [1., 1., 0.])


def two_charges():
    """electrostatic_interactions.py:76-101 (box 20, charges at x=9 and 11)."""
    pos = np.array([[9., 2., 2.], [11., 2., 2.]])
    q = np.array([1., -1.])
    params = dict(accuracy=1e-7, mesh=22, cao=7, r_cut=8.90625,
                  alpha=0.387611049779351)
    return pos, q, [20., 20., 20.], params


def rel_rms(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def energy_scale(e_ref, q, alpha, prefactor):
    """Magnitude of the sums a P3M energy is the difference of.

    The k-space energy (p3m.cpp:551-589) is the mesh sum minus the self term
    prefactor*alpha/sqrt(pi)*sum(q^2); on dense small systems the two are
    100x larger than their difference, so a rounding error of the FP32 mesh
    sum shows in the total energy amplified by that ratio.  FP32 energy
    tolerances are stated relative to |E| + E_self.
    """
    q = np.asarray(q, dtype=np.float64)
    return abs(e_ref) + prefactor * alpha / math.sqrt(math.pi) * float(np.sum(q * q))
