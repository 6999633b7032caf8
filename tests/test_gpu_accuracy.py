"""Absolute rms force error of P3M against exact FP64 Ewald forces.

The fixtures (tests/golden/make_ewald_fixtures.py) hold the exact total
Coulomb force on 2000-particle subsets of the 10^6- and 250,000-particle
workloads.  P3M's total force is the device k-space force plus the real-space
pair force (A&S erfc, as the reference computes it) within r_cut.  r_cut and
alpha come from the reference's closed-form tuning at fixed mesh and cao
(TuningAlgorithm::get_mc_time bisection).  The error must stay within 1.5x
the estimate the tuning targets, for the FP64 solver and the FP32-mesh form.
"""
import os

import numpy as np
import pytest

from tests import p3m_systems as S

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _total_force_error(pos, q, L, idx, f_exact, mesh, cao, accuracy, single):
    import torch
    from espresso_amd.electrostatics import P3M
    from oracle import oracle as O
    n_q = int(np.count_nonzero(q))
    t = O.tune_r_cut(1.0, [mesh] * 3, cao, n_q, float(np.sum(q * q)), accuracy, [L] * 3)
    assert t is not None
    r_cut, alpha = t[0] * L, t[1] / L
    p3m = P3M(prefactor=1.0, accuracy=accuracy, mesh=mesh, cao=cao, r_cut=r_cut,
              alpha=alpha, tune=False, single_precision=single)
    p3m.activate(pos[:1000], np.where(np.arange(1000) % 2 == 0, 1., -1.), box_l=[L] * 3)
    tp = torch.tensor(pos, device="cuda")
    tq = torch.tensor(q, device="cuda")
    f = torch.zeros_like(tp)
    p3m.add_long_range_forces(tp, tq, f)
    f_k = f.cpu().numpy()[idx]
    order = np.concatenate([idx, np.setdiff1d(np.arange(len(q)), idx)])
    f_r = O.real_space_subset(pos[order], q[order], [L] * 3, 1.0, alpha, r_cut, len(idx))
    d = f_k + f_r - f_exact
    return float(np.sqrt(np.mean(np.sum(d * d, axis=1)))), t[2]


@pytest.mark.parametrize("single", [False, True])
def test_c2_rms_force_error_against_ewald(single):
    pos, q, box, _ = S.config_c2()
    fx = np.load(os.path.join(GOLDEN, "ewald_c2_subset.npz"))
    err, est = _total_force_error(pos, q, box[0], fx["idx"], fx["force"], 128, 7, 1e-5,
                                  single)
    print(f"C2 single={single}: rms force error {err:.3e} (estimate {est:.3e})")
    assert err <= 1.5 * 1e-5


@pytest.mark.parametrize("mesh,cao,acc", [(64, 5, 1e-4), (48, 3, 1e-3), (96, 7, 1e-5)])
def test_c3_rms_force_error_against_ewald(mesh, cao, acc):
    pos, q, box = S.config_c3()
    fx = np.load(os.path.join(GOLDEN, "ewald_c3_subset.npz"))
    for single in (False, True):
        err, est = _total_force_error(pos, q, box[0], fx["idx"], fx["force"], mesh, cao,
                                      acc, single)
        print(f"C3 mesh {mesh} cao {cao} acc {acc} single={single}: {err:.3e} (est {est:.3e})")
        assert err <= 1.5 * acc
