"""ELC on the device (ElectrostaticLayerCorrection without dielectric
contrasts, elc.cpp:184-940) against the ELC oracle (oracle/elc.py, pinned in
test_oracle_elc.py against the reference's MMM2D slab fixture) and against
that fixture itself.

Tolerances: the ELC forces and energy (P3M+ELC minus P3M, both FP64 on the
device) rel-rms <= 1e-10 and rel <= 1e-10 against the oracle; the total
device force and energy on the fixture within the reference test's 1e-5 and
1e-4.
"""
import os

import numpy as np
import pytest

from tests import p3m_systems as S

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _slab(n, box, gap, seed, net=0.):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0., 1., (n, 3)) * np.asarray(box)
    pos[:, 2] = rng.uniform(0.05, box[2] - gap - 0.05, n)
    q = np.where(np.arange(n) % 2 == 0, 1., -1.) * rng.uniform(0.5, 1.5, n)
    q -= q.mean()
    q += net / n
    q[::23] = 0.
    return pos, q


def _p3m(box, mesh, cao, alpha, r_cut, prefactor=1., **kw):
    from espresso_amd.electrostatics import P3M
    return P3M(prefactor=prefactor, accuracy=1e-4, mesh=mesh, cao=cao, alpha=alpha,
               r_cut=r_cut, tune=False, **kw)


def test_elc_fixture_total_force_and_energy():
    import torch
    from espresso_amd.electrostatics import ELC
    from oracle import oracle as O
    d = np.load(os.path.join(GOLDEN, "coulomb_mixed_periodicity.npz"))
    pos, q = d["pos"], d["q"]
    box = [10., 10., 12.]
    mesh, cao, r_cut = [42, 42, 50], 7, 3.
    acc = O.calculate_accuracy(1., mesh, cao, r_cut / box[0], len(q), float(np.sum(q * q)),
                               1e-6, box)
    p3m = _p3m(box, mesh, cao, acc[3] / box[0], r_cut)
    elc = ELC(actor=p3m, maxPWerror=1e-6, gap_size=3.)
    elc.activate(pos, q, box_l=box)
    tp = torch.tensor(pos, device="cuda")
    tq = torch.tensor(q, device="cuda")
    f = torch.zeros_like(tp)
    e_lr, _ = elc.solve(tp, tq, force=f, energy=True)
    e_sr = elc.add_short_range_forces(tp, tq, f, energy=True)
    f = f.cpu().numpy()
    err_f = np.max(np.abs(f - d["force"]))
    err_e = abs(e_lr + e_sr - float(d["energy"]))
    print(f"far_cut {elc.get_params()['far_cut']:.4f}: max force error {err_f:.2e}, "
          f"energy error {err_e:.2e}")
    assert err_f <= 1e-5
    assert err_e <= 1e-4
    assert elc.get_params()["far_cut"] == pytest.approx(0.8)


@pytest.mark.parametrize("neutralize,net,box,gap", [
    (True, 0., (12., 10., 16.), 4.),
    (False, 4., (12., 10., 16.), 4.),
    (True, 3., (9., 13., 20.), 6.),
])
def test_elc_terms_match_oracle(neutralize, net, box, gap):
    import torch
    from espresso_amd.electrostatics import ELC
    from oracle import elc as E
    n = 5000
    pos, q = _slab(n, box, gap, 17, net=net)
    # positions outside the box along x and y (folded by the solver)
    pos[:7, 0] += 3 * box[0]
    pos[7:14, 1] -= 2 * box[1]
    plain = _p3m(box, [24, 20, 32], 5, 1.0, 3.0, prefactor=1.7, check_neutrality=False)
    with_elc = _p3m(box, [24, 20, 32], 5, 1.0, 3.0, prefactor=1.7, check_neutrality=False)
    elc = ELC(actor=with_elc, maxPWerror=1e-5, gap_size=gap, neutralize=neutralize)
    plain.activate(pos, q, box_l=list(box))
    elc.activate(pos, q, box_l=list(box))
    fc = elc.get_params()["far_cut"]
    f0, f1 = np.zeros((n, 3)), np.zeros((n, 3))
    plain.add_long_range_forces(pos, q, f0)
    elc.add_long_range_forces(pos, q, f1)
    e0 = plain.long_range_energy(pos, q)
    e1 = elc.long_range_energy(pos, q)
    f_ref, e_ref = E.elc_forces_energy(pos, q, box, 1.7, gap, fc, neutralize=neutralize)
    err = S.rel_rms(f1 - f0, f_ref)
    print(f"ELC neutralize={neutralize} net={net}: far_cut {fc:.3f}, "
          f"forces rel-rms {err:.2e}, energy {e1 - e0:.6f} vs {e_ref:.6f}")
    assert err <= 1e-10
    assert abs((e1 - e0) - e_ref) <= 1e-10 * max(abs(e_ref), 1.)
    # device tensors, asynchronous forces
    tp = torch.tensor(pos, device="cuda")
    tq = torch.tensor(q, device="cuda")
    tf = torch.zeros_like(tp)
    elc.add_long_range_forces(tp, tq, tf, sync=False)
    elc.actor.synchronize()
    assert S.rel_rms(tf.cpu().numpy(), f1) <= 1e-12
    # no ELC pressure in the reference (zero virial)
    _, vir = elc.solve(pos, q, force=None, energy=False, pressure=True)
    assert np.all(vir == 0.)


def test_elc_gap_particle_is_reported():
    from espresso_amd.electrostatics import ELC
    box = (8., 8., 12.)
    pos, q = _slab(400, box, 3., 4)
    p3m = _p3m(box, [16] * 3, 4, 1.2, 2.5, check_neutrality=False)
    elc = ELC(actor=p3m, maxPWerror=1e-4, gap_size=3.)
    elc.activate(pos, q, box_l=list(box))
    pos[10, 2] = 10.5
    with pytest.raises(RuntimeError, match="gap region"):
        elc.add_long_range_forces(pos, q, np.zeros((400, 3)))
    # uncharged particles may sit in the gap
    q2 = q.copy()
    q2[10] = 0.
    elc.add_long_range_forces(pos, q2, np.zeros((400, 3)))
    with pytest.raises(RuntimeError, match="larger than box length"):
        ELC(actor=_p3m(box, [16] * 3, 4, 1.2, 2.5, check_neutrality=False), maxPWerror=1e-4,
            gap_size=13.).activate(pos, q2, box_l=list(box))


def test_dielectric_contrast_analytic_image_charges():
    """The reference's analytic test (elc_vs_analytic.py): two opposite
    charges 1 apart above a dielectric boundary (delta_mid_bot = 39/41) in a
    200 x 200 x 275 box (gap 75), P3M mesh 58 x 58 x 70, cao 4, accuracy
    1e-7, tuned once; total force on the lower charge and total energy
    against the image-charge solution at atol 1e-4, rtol 1e-7."""
    from espresso_amd.electrostatics import ELC, P3M
    L, gap, dist = 200., 75., 1.
    box = [L, L, L + gap]
    db = 39. / 41.
    pref = 2.0
    p3m = P3M(prefactor=pref, accuracy=1e-7, mesh=[58, 58, 70], cao=4, verbose=False,
              check_neutrality=False)
    elc = ELC(actor=p3m, gap_size=gap, maxPWerror=1e-7, delta_mid_bot=db, delta_mid_top=0.)
    pos = np.array([box[0] / 2, box[1] / 2, box[2] / 2]) * np.ones((2, 3))
    pos[1, 2] += dist
    elc.activate(pos, np.array([-5., 5.]), box_l=box)
    zs = np.linspace(0.1, L - 0.1 - dist, 12)
    worst_f = worst_e = 0.
    for charge in np.arange(-5.0, 5.1, 2.5):
        q = np.array([charge, -charge])
        for z in zs:
            pos = np.array([[L / 2, L / 2, z], [L / 2, L / 2, z + dist]])
            f = np.zeros((2, 3))
            e_lr, _ = elc.solve(pos, q, force=f, energy=True)
            e_sr = elc.add_short_range_forces(pos, q, f, energy=True)
            c2 = pref * charge * charge
            f_ref = c2 * (1. / dist ** 2 + db * (1. / (2 * z) ** 2 - 1. / (2 * z + dist) ** 2))
            e_ref = c2 * (-1. / dist + db * (1. / (4 * z) - 1. / (2 * z + dist) +
                                             1. / (4 * (z + dist))))
            np.testing.assert_allclose(f[0, 2], f_ref, atol=1e-4, rtol=1e-7)
            assert abs(e_lr + e_sr - e_ref) <= 1e-4, (charge, z, e_lr + e_sr, e_ref)
            worst_f = max(worst_f, abs(f[0, 2] - f_ref))
            worst_e = max(worst_e, abs(e_lr + e_sr - e_ref))
    print(f"r_cut {p3m.get_params()['r_cut']:.3f}: worst force error {worst_f:.2e}, "
          f"energy {worst_e:.2e}")


@pytest.mark.parametrize("dt,db,const_pot,pot_diff", [
    (0.3, -0.5, False, 0.),
    (-1., -1., True, 2.5),
    (0., 0.8, False, 0.),
])
def test_dielectric_contrast_total_matches_oracle(dt, db, const_pot, pot_diff):
    """P3M + ELC with contrasts, short range included, against the oracle's
    composition (image charges in the mesh, layer pair and self terms, image
    series): total forces rel-rms <= 1e-9, energy <= 1e-9 relative."""
    from espresso_amd.electrostatics import ELC
    from oracle import elc as E
    from oracle import oracle as O
    box = (9., 8., 14.)
    gap = 4.
    n = 300
    pos, q = _slab(n, box, gap, 29)
    q[::23] = 0.
    q -= np.sum(q) / np.count_nonzero(q) * (q != 0)
    # crowd the layers: some particles close to both boundaries
    pos[:20, 2] = np.linspace(0.05, 1.0, 20)
    pos[20:40, 2] = np.linspace(box[2] - gap - 1.0, box[2] - gap - 0.05, 20)
    alpha, r_cut = 1.4, 2.5
    p3m = _p3m(box, [24, 24, 36], 5, alpha, r_cut, prefactor=1.3)
    elc = ELC(actor=p3m, maxPWerror=1e-6, gap_size=gap, delta_mid_top=dt, delta_mid_bot=db,
              const_pot=const_pot, pot_diff=pot_diff)
    elc.activate(pos, q, box_l=list(box))
    f = np.zeros((n, 3))
    e_lr, _ = elc.solve(pos, q, force=f, energy=True)
    e_sr = elc.add_short_range_forces(pos, q, f, energy=True)
    ep = elc.get_params()
    ora = O.OracleP3M(list(box), [24, 24, 36], 5, alpha, r_cut, 1.3)
    sl = E.space_layer(gap, box[2] - gap, r_cut)
    f_ref, e_ref = E.elc_p3m_total(
        pos, q, box, ora, 1.3, alpha, r_cut, gap, ep["far_cut"], delta_top=dt, delta_bot=db,
        const_pot=const_pot, pot_diff=pot_diff, space_layer=sl,
        real_space=lambda p, qq: O.real_space(p, qq, box, 1.3, alpha, r_cut))
    err = S.rel_rms(f, f_ref)
    print(f"dt {dt} db {db} const_pot {const_pot}: far_cut {ep['far_cut']:.3f}, "
          f"forces rel-rms {err:.2e}, energy {e_lr + e_sr:.8f} vs {e_ref:.8f}")
    assert err <= 1e-9
    assert abs(e_lr + e_sr - e_ref) <= 1e-9 * abs(e_ref)
