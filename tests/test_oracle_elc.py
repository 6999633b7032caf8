"""The ELC oracle (oracle/elc.py) pinned without the reference code:

- ELC + P3M + real space (all oracle) on the reference's slab fixture
  (coulomb_mixed_periodicity.py:76-90) against its MMM2D forces and energy,
  at the reference test's tolerances (1e-5 forces, 1e-4 energy);
- the ELC forces are minus the gradient of the ELC energy (central
  differences), with and without neutralisation, for a charged system;
- tune_far_cut returns the smallest grid value meeting maxPWerror;
- with a dielectric contrast (image charges in the mesh and pair terms
  inside the space layer, the image series of the far field outside it),
  the reference's analytic image-charge test (elc_vs_analytic.py).
"""
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _slab(n, box, gap, seed, net=0.):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0., 1., (n, 3)) * np.asarray(box)
    pos[:, 2] = rng.uniform(0.1, box[2] - gap - 0.1, n)
    q = np.where(np.arange(n) % 2 == 0, 1., -1.) * rng.uniform(0.5, 1.5, n)
    q -= q.mean()
    q += net / n
    return pos, q


def test_elc_p3m_matches_mmm2d_fixture():
    from oracle import elc as E
    from oracle import oracle as O
    d = np.load(os.path.join(GOLDEN, "coulomb_mixed_periodicity.npz"))
    pos, q = d["pos"], d["q"]
    box = [10., 10., 12.]
    mesh, cao, r_cut = [42, 42, 50], 7, 3.
    acc = O.calculate_accuracy(1., mesh, cao, r_cut / box[0], len(q),
                               float(np.sum(q * q)), 1e-6, box)
    alpha = acc[3] / box[0]
    assert acc[0] <= 1e-6
    far_cut = E.tune_far_cut(box, 3., 1e-6)
    p3m = O.OracleP3M(box, mesh, cao, alpha, r_cut, 1.)
    e_rs, f_rs = O.real_space(pos, q, box, 1., alpha, r_cut)
    f_elc, e_elc = E.elc_forces_energy(pos, q, box, 1., 3., far_cut)
    f = f_rs + p3m.forces(pos, q) + f_elc
    e = e_rs + p3m.energy(pos, q) + e_elc
    print(f"far_cut {far_cut:.4f}: max force error {np.max(np.abs(f - d['force'])):.2e}, "
          f"energy error {abs(e - float(d['energy'])):.2e}")
    np.testing.assert_allclose(f, d["force"], atol=1e-5)
    assert abs(e - float(d["energy"])) <= 1e-4
    # without the correction the slab forces are off by far more
    f_no = f_rs + p3m.forces(pos, q)
    assert np.max(np.abs(f_no - d["force"])) > 1e-2


@pytest.mark.parametrize("neutralize,net", [(True, 0.), (False, 0.), (False, 3.),
                                            (True, 2.)])
def test_elc_forces_are_minus_energy_gradient(neutralize, net):
    from oracle import elc as E
    box = [6., 7., 9.]
    gap = 2.
    pos, q = _slab(12, box, gap, 5, net=net)
    far_cut = 0.9
    f, _ = E.elc_forces_energy(pos, q, box, 1.7, gap, far_cut, neutralize=neutralize)
    h = 1e-5
    for i in (0, 3, 7):
        for dim in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, dim] += h
            pm[i, dim] -= h
            ep = E.elc_forces_energy(pp, q, box, 1.7, gap, far_cut, neutralize)[1]
            em = E.elc_forces_energy(pm, q, box, 1.7, gap, far_cut, neutralize)[1]
            g = -(ep - em) / (2. * h)
            assert abs(g - f[i, dim]) <= 1e-6 * max(1., np.max(np.abs(f))), (i, dim, g, f[i, dim])


def test_tune_far_cut_is_minimal():
    from oracle import elc as E
    box = [10., 10., 12.]
    for gap, err in ((3., 1e-6), (2., 1e-4), (5., 1e-3)):
        fc = E.tune_far_cut(box, gap, err)
        step = 0.1
        # the pairwise bound of elc.cpp:942-973 at far_cut fc, evaluated
        # from the slab geometry (h = L_z - gap)

        def bound(c):
            lz, h = box[2], box[2] - gap
            pref = 2. * math.pi * c
            s = pref + 2. * (0.1 + 0.1)
            return 0.5 / (-math.expm1(-pref * lz)) * (
                math.exp(pref * (h - lz)) * (s + 1. / (lz - h)) / (lz - h) +
                math.exp(-pref * (h + lz)) * (s + 1. / (lz + h)) / (lz + h))
        assert bound(fc) <= err
        assert fc - step < step / 2 or bound(fc - step) > err
    with pytest.raises(RuntimeError, match="maxPWerror too small"):
        E.tune_far_cut(box, 0.01, 1e-12)


def test_gap_particles_are_rejected():
    from oracle import elc as E
    pos, q = _slab(10, [5., 5., 8.], 2., 1)
    pos[3, 2] = 6.5
    with pytest.raises(RuntimeError, match="gap"):
        E.elc_forces_energy(pos, q, [5., 5., 8.], 1., 2., 1.)


@pytest.mark.parametrize("charge,zpos", [(-5., 0.1), (2.5, 3.3), (5., 17.), (-2.5, 120.)])
def test_dielectric_contrast_image_charge_analytic(charge, zpos):
    """Two opposite charges 1 apart above a dielectric boundary
    (delta_mid_bot = 39/41) in a 200 x 200 x 275 box with a 75 gap: the
    reference's analytic image-charge solution (elc_vs_analytic.py:35-76)
    for the force on the lower charge and the total energy, at its
    tolerances (atol 1e-4, rtol 1e-7)."""
    from oracle import elc as E
    from oracle import oracle as O
    L, gap, dist = 200., 75., 1.
    box = [L, L, L + gap]
    db = 39. / 41.
    pref = 2.0
    pos = np.array([[L / 2, L / 2, zpos], [L / 2, L / 2, zpos + dist]])
    q = np.array([charge, -charge])
    mesh, cao = [58, 58, 70], 4
    t = O.tune_r_cut(pref, mesh, cao, 2, float(np.sum(q * q)), 1e-7, box)
    r_cut, alpha = t[0] * box[0], t[1] / box[0]
    p3m = O.OracleP3M(box, mesh, cao, alpha, r_cut, pref)
    sl = E.space_layer(gap, L, r_cut)
    fc = E.tune_far_cut(box, gap, 1e-7, space_layer=sl, dielectric=True)
    f, e = E.elc_p3m_total(pos, q, box, p3m, pref, alpha, r_cut, gap, fc, delta_bot=db,
                           real_space=lambda p, qq: O.real_space(p, qq, box, pref, alpha, r_cut))
    c2 = pref * charge * charge
    f_ref = c2 * (1. / dist ** 2 + db * (1. / (2 * zpos) ** 2 - 1. / (2 * zpos + dist) ** 2))
    e_ref = c2 * (-1. / dist + db * (1. / (4 * zpos) - 1. / (2 * zpos + dist) +
                                     1. / (4 * (zpos + dist))))
    print(f"q {charge} z {zpos}: force {f[0, 2]:.8f} vs {f_ref:.8f}, energy {e:.8f} vs {e_ref:.8f}")
    np.testing.assert_allclose(f[0, 2], f_ref, atol=1e-4, rtol=1e-7)
    assert abs(e - e_ref) <= 1e-4
