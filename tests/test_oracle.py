"""Pins the FP64 CPU oracle (oracle/p3m_oracle.c) before it is trusted as the
parity reference: against the reference's own fixtures and known answers
(testsuite/python/p3m_madelung.py, electrostatic_interactions.py,
coulomb_cloud_wall.py, coulomb_tuning.py, src/core/unit_tests/p3m_test.cpp,
EspressoSystemStandAlone_test.cpp) and against what the mathematics fixes
(DFT definition, B-spline recursion and moments, closed-form influence
function, Ewald sum, sinc series)."""
import itertools
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from tests import p3m_systems as S


# --------------------------------------------------------------------------
# small functions
# --------------------------------------------------------------------------
def test_meshift_reference_vectors():
    # src/core/unit_tests/p3m_test.cpp:32-60
    assert list(O.meshift(1)) == [0]
    assert list(O.meshift(4)) == [0, 1, -2, -1]
    assert list(O.meshift(7)) == [0, 1, 2, 3, -3, -2, -1]
    assert list(O.meshift(1, True)) == [0]
    assert list(O.meshift(4, True)) == [0, 1, 0, -1]
    assert list(O.meshift(7, True)) == [0, 1, 2, 0, -3, -2, -1]


def _cardinal_bspline(p, u):
    """M_p(u) by the Cox-de Boor recursion in exact rationals."""
    if p == 1:
        return Fraction(1) if 0 <= u < 1 else Fraction(0)
    return (u * _cardinal_bspline(p - 1, u)
            + (p - u) * _cardinal_bspline(p - 1, u - 1)) / (p - 1)


@pytest.mark.parametrize("p", range(1, 8))
def test_bspline_table_against_recursion(p):
    xs = [Fraction(k, 16) for k in range(-7, 8)]  # open interval: pieces are closed
    for i in range(p):
        num, den = O.bspline_coeffs(p, i)
        for x in xs:
            poly = sum(Fraction(int(c)) * x ** m for m, c in enumerate(num)) / den
            ref = _cardinal_bspline(p, Fraction(i) + Fraction(1, 2) - x)
            assert poly == ref, (p, i, x)


@pytest.mark.parametrize("p", range(1, 8))
def test_bspline_partition_of_unity_and_symmetry(p):
    for x in np.linspace(-0.5, 0.5, 41):
        w = [O.bspline(p, i, x) for i in range(p)]
        assert abs(sum(w) - 1.) < 1e-14
        for i in range(p):
            assert abs(w[i] - O.bspline(p, p - 1 - i, -x)) < 1e-15


def test_sinc_against_numpy():
    # |x| <= 0.1 uses the 8th-order Taylor series, accurate to 2.35e-13
    # (math.hpp:44-51); elsewhere sin(pi x)/(pi x)
    for x in np.concatenate([np.linspace(-2, 2, 401), [1e-9, -0.1, 0.1, 0.0999]]):
        tol = 2.5e-13 if abs(x) <= 0.1 else 1e-15
        assert abs(O.sinc(x) - np.sinc(x)) < tol


@pytest.mark.parametrize("cao", range(1, 8))
def test_cotangent_sum_brute_force(cao):
    M = 16
    m = np.arange(-200000, 200001)
    for n in range(-8, 8):
        if cao == 1:
            ref = 1.0  # sum_m sinc^2((n + m M)/M) = 1 identically
        else:
            ref = float(np.sum(np.sinc((n + m * M) / M) ** (2 * cao)))
        assert abs(O.cotangent_sum(cao, n, 1. / M) - ref) < 1e-10


@pytest.mark.parametrize("shape", [(4, 6, 8), (22, 12, 10), (7, 11, 13), (42, 3, 5),
                                   (1, 8, 9)])
def test_fft3d_is_the_dft(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    np.testing.assert_allclose(O.fft3d(a, -1), np.fft.fftn(a), rtol=0,
                               atol=1e-12 * np.abs(a).sum())
    np.testing.assert_allclose(O.fft3d(a, +1), np.fft.ifftn(a) * a.size, rtol=0,
                               atol=1e-12 * np.abs(a).sum())


def test_as_erfc_part_accuracy():
    # Abramowitz & Stegun 7.1.26: |error of erfc| <= 1.5e-7
    from scipy.special import erfc
    for d in np.linspace(0., 6., 61):
        approx = O.lib().p3mo_as_erfc_part(d) * math.exp(-d * d)
        assert abs(approx - erfc(d)) <= 1.5e-7


# --------------------------------------------------------------------------
# geometry, indices and weights
# --------------------------------------------------------------------------
def test_local_mesh_margins():
    # single-rank CA mesh with skin 0.4 (C1: margin 3, side 38; C2: 4, 136)
    g1 = O.OracleP3M([21.544] * 3, 32, 5, 1.07, 3.0, 2.0).geometry()
    assert list(g1["dim"]) == [38] * 3 and list(g1["margin"]) == [3] * 6
    assert list(g1["inner"]) == [32] * 3
    g2 = O.OracleP3M([125.99] * 3, 128, 7, 1.37, 2.5, 1.0).geometry()
    assert list(g2["dim"]) == [136] * 3 and list(g2["margin"]) == [4] * 6


@pytest.mark.parametrize("cao", [1, 2, 3, 4, 5, 6, 7])
def test_weights_reproduce_moments(cao):
    """sum_m w_m = 1 and, for cao >= 2, sum_m w_m r_m = r with mesh points at
    (ld_ind + nmp + i + mesh_off) a: pins the stencil index and the offsets."""
    L = 10.
    M = 20
    ora = O.OracleP3M([L] * 3, M, cao, 1.0, 2.0, 1.0)
    g = ora.geometry()
    a = L / M
    rng = np.random.default_rng(cao)
    pts = np.concatenate([rng.uniform(-0.2, L + 0.2, (200, 3)),
                          np.stack([np.arange(M + 1) * a] * 3, 1)])
    for r in pts:
        _, nmp, w = ora.weights(r)
        for d in range(3):
            assert abs(w[d].sum() - 1.) < 1e-14
            if cao >= 2:
                xs = (g["ld_ind"][d] + nmp[d] + np.arange(cao) + 0.5) * a
                assert abs(np.dot(w[d], xs) - r[d]) < 1e-11


# --------------------------------------------------------------------------
# influence function
# --------------------------------------------------------------------------
@pytest.mark.parametrize("mesh,cao", [([16, 16, 16], 5), ([12, 20, 10], 3),
                                      ([22, 22, 22], 7)])
def test_influence_function_closed_form(mesh, cao):
    box = [10., 12., 9.]
    alpha = 0.9
    ora = O.OracleP3M(box, mesh, cao, alpha, 2.0, 1.0)
    gf, ge = ora.influence()
    sh = [O.meshift(m) for m in mesh]
    kx, ky, kz = np.meshgrid(*[2 * np.pi * sh[d] / box[d] for d in range(3)],
                             indexing="ij")
    U2 = (np.sinc(sh[0] / mesh[0])[:, None, None] * np.sinc(sh[1] / mesh[1])[None, :, None]
          * np.sinc(sh[2] / mesh[2])[None, None, :]) ** (2 * cao)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = 4 * np.pi * np.exp(-k2 / (4 * alpha ** 2)) / (k2 * U2)
    ref[k2 / (4 * alpha ** 2) >= 30.] = 0.
    corners = np.zeros(mesh, bool)
    for c in itertools.product(*[[0, m // 2] for m in mesh]):
        corners[c] = True
    ref[corners] = 0.
    # sinc's Taylor branch (2.35e-13) enters U^2 to the power 2 cao
    np.testing.assert_allclose(gf, ref, rtol=5e-12, atol=0)
    np.testing.assert_allclose(ge, gf, rtol=1e-14, atol=0)
    assert np.all(gf[corners] == 0.)


# --------------------------------------------------------------------------
# reference fixtures
# --------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["cube", "wire", "sheet"])
def test_madelung_energy_and_pressure(case):
    # p3m_madelung.py:202-312
    pos, q, box, params, mc, base = {"cube": S.madelung_cube,
                                     "wire": S.madelung_wire,
                                     "sheet": S.madelung_sheet}[case]()
    rtol = {"cube": 1e-6, "wire": 5e-7, "sheet": 1e-7}[case]
    ora = O.OracleP3M(box, params["mesh"], params["cao"], params["alpha"],
                      params["r_cut"], 1.0)
    ek = ora.energy(pos, q)
    vk = ora.pressure(pos, q)
    er, _, vr = O.real_space(pos, q, box, 1.0, params["alpha"], params["r_cut"],
                             virial=True)
    n = len(q)
    np.testing.assert_allclose(2. * (ek + er) / n, mc, atol=0., rtol=rtol)
    ref = mc * np.diag(base) / np.sum(base)
    np.testing.assert_allclose(2. * (vk + vr) / n, ref, atol=5e-9, rtol=5e-6)


def test_two_charges_energy_force():
    # electrostatic_interactions.py:76-101
    pos, q, box, params = S.two_charges()
    pref = 1.1
    ora = O.OracleP3M(box, params["mesh"], params["cao"], params["alpha"],
                      params["r_cut"], pref)
    e = ora.energy(pos, q)
    f = ora.forces(pos, q)
    er, fr = O.real_space(pos, q, box, pref, params["alpha"], params["r_cut"])
    np.testing.assert_allclose(e + er, -0.501062398379 * pref, atol=1e-5)
    np.testing.assert_allclose((f + fr)[0], [0.248921612 * pref, 0, 0], atol=1e-5)
    np.testing.assert_allclose((f + fr)[1], [-0.248921612 * pref, 0, 0], atol=1e-5)


def test_dipole_correction_closed_form():
    # electrostatic_interactions.py:103-125
    pos, q, box, params = S.two_charges()
    pref = 1.1
    V = float(np.prod(box))
    dip = q[0] * pos[0] + q[1] * pos[1]
    base = O.OracleP3M(box, params["mesh"], params["cao"], params["alpha"],
                       params["r_cut"], pref)
    e0 = base.energy(pos, q)
    f0 = base.forces(pos, q)
    for eps in np.power(10., np.arange(-4, 5)):
        ora = O.OracleP3M(box, params["mesh"], params["cao"], params["alpha"],
                          params["r_cut"], pref, epsilon=eps)
        corr = 4 * np.pi / V / (1 + 2 * eps)
        np.testing.assert_allclose(ora.energy(pos, q),
                                   e0 + pref * corr * np.dot(dip, dip), atol=1e-10)
        np.testing.assert_allclose(ora.forces(pos, q),
                                   f0 - pref * np.outer(q, corr * dip), atol=1e-10)


def test_standalone_pair():
    # EspressoSystemStandAlone_test.cpp:222-289 (box 12, mesh 12, cao 5)
    box = [12.] * 3
    pref = 2.
    alpha, r_cut = 0.615, 3.5
    ora = O.OracleP3M(box, 12, 5, alpha, r_cut, pref)
    pos1 = np.array([5.9, 5.9, 1.0])
    pos2 = np.array([6.0, 5.9, 1.0])
    pos3 = np.array([6.1, 6.1, 1.0])
    for _ in range(10):
        pos2[0] += 0.02
        pos = np.array([pos1, pos2, pos3])
        q = np.array([0.5, -0.5, 0.])
        r = np.linalg.norm(pos2 - pos1)
        e_ref = -pref * 0.25 / r
        er, fr = O.real_space(pos, q, box, pref, alpha, r_cut)
        e = ora.energy(pos, q) + er
        f = ora.forces(pos, q) + fr
        assert abs(e - e_ref) <= 0.002e-2 * abs(e_ref)
        assert abs(f[0, 0] - (-e_ref / r)) <= 0.02e-2 * abs(e_ref / r)
        assert abs(f[0, 1]) <= 1e-12 and abs(f[0, 2]) <= 1e-12


def test_cloud_wall_fixture(reference_data):
    # coulomb_cloud_wall.py:37-97 (fixture read in place)
    data = np.genfromtxt(os.path.join(reference_data,
                                      "coulomb_cloud_wall_system.data"))
    pos, q, f_ref = data[:, 1:4], data[:, 4], data[:, 5:8]
    pref = 3.
    box = [10.] * 3
    ora = O.OracleP3M(box, 64, 7, 2.70746, 1.001, pref)
    er, fr = O.real_space(pos, q, box, pref, 2.70746, 1.001)
    f = (ora.forces(pos, q) + fr) / pref
    np.testing.assert_allclose(f, f_ref, atol=2e-3)
    assert abs((ora.energy(pos, q) + er) / pref - 148.94229549) <= 1e-3


def test_tuning_fixture(reference_data):
    # coulomb_tuning.py:32-60: any parameter set reaching accuracy 5e-4
    d = np.load(os.path.join(reference_data, "coulomb_tuning_system.npz"))
    box = [10.] * 3
    # the fixture stores unfolded positions; particles are folded into the box
    pos, q, f_ref = np.mod(d["pos"], 10.), d["charges"], d["forces"]
    nq = int(np.count_nonzero(q))
    for mesh, cao in ((24, 5), (32, 7)):
        tuned = O.tune_r_cut(1., [mesh] * 3, cao, nq, float(np.sum(q * q)), 5e-4, box)
        assert tuned is not None
        r_cut = tuned[0] * box[0]
        alpha = tuned[1] / box[0]
        ora = O.OracleP3M(box, mesh, cao, alpha, r_cut, 1.)
        _, fr = O.real_space(pos, q, box, 1., alpha, r_cut)
        np.testing.assert_allclose(ora.forces(pos, q) + fr, f_ref, atol=2e-3)


# --------------------------------------------------------------------------
# against an independent Ewald sum
# --------------------------------------------------------------------------
def test_p3m_converges_to_ewald():
    pos, q = S.uniform_system(300, 8., 11)
    box = [8.] * 3
    alpha = 1.2
    kmax = 2 * alpha * math.sqrt(40.)
    e_ew, f_ew = O.ewald_kspace(pos, q, box, 1.0, alpha, kmax)
    self_e = -np.sum(q * q) * alpha / math.sqrt(math.pi)
    errs = []
    for mesh in (16, 32, 64):
        ora = O.OracleP3M(box, mesh, 7, alpha, 3.0, 1.0)
        f = ora.forces(pos, q)
        errs.append(S.rel_rms(f, f_ew))
        if mesh == 64:
            assert abs(ora.energy(pos, q) - (e_ew + self_e)) < 1e-6 * abs(e_ew + self_e)
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-6


def test_error_estimates_known_values():
    # closed forms of p3m.cpp:145-243 at the C1 / C2 fixed and freed settings
    L1 = 21.544
    acc = O.calculate_accuracy(2.0, [32] * 3, 5, 3.0 / L1, 10000, 10000., 1e-4, [L1] * 3)
    assert abs(acc[3] - 23.1523) < 1e-3
    assert abs(acc[1] - 7.07e-5) < 0.01e-5 and abs(acc[2] - 2.98e-2) < 0.01e-2
    t = O.tune_r_cut(2.0, [32] * 3, 5, 10000, 10000., 1e-4, [L1] * 3)
    assert abs(t[0] * L1 - 6.827) < 0.03 and abs(t[1] / L1 - 0.46281) < 2e-3
    L2 = 125.99
    t = O.tune_r_cut(1.0, [128] * 3, 7, 900000, 900000., 1e-5, [L2] * 3)
    assert abs(t[0] * L2 - 8.558) < 0.13 and abs(t[1] / L2 - 0.38896) < 4e-3
    assert t[2] <= 1e-5


def test_exact_ewald_reference_is_independent_of_splitting():
    # the total force of the exact Ewald reference (real space with libm erfc
    # + reciprocal sum) must not depend on alpha: this pins both parts, their
    # signs and prefactors, as used by tests/golden/make_ewald_fixtures.py
    pos, q = S.uniform_system(1200, 10., 99)
    box = [10.] * 3
    forces = []
    for alpha in (1.2, 1.6):
        r_e = 5.6 / alpha
        f_r = O.real_space_subset(pos, q, box, 1.0, alpha, r_e, 40, exact_erfc=True)
        _, f_k = O.ewald_kspace(pos, q, box, 1.0, alpha, 11.2 * alpha, n_out=40)
        forces.append(f_r + f_k)
    assert S.rel_rms(forces[0], forces[1]) <= 1e-11
    # the A&S pair force of the P3M real-space part stays within 1e-6 of erfc
    f_as = O.real_space_subset(pos, q, box, 1.0, 1.2, 3.0, 40)
    f_ex = O.real_space_subset(pos, q, box, 1.0, 1.2, 3.0, 40, exact_erfc=True)
    assert S.rel_rms(f_as, f_ex) <= 1e-6


def test_ewald_fixtures_present():
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for name, n_sys in (("ewald_c2_subset.npz", 1_000_000), ("ewald_c3_subset.npz", 250_000)):
        d = np.load(os.path.join(here, name))
        assert d["force"].shape == (2000, 3) and np.all(np.isfinite(d["force"]))
        assert d["idx"].shape == (2000,) and d["idx"].max() < n_sys
        assert float(d["alpha_e"]) * float(d["r_e"]) == pytest.approx(5.6)


def test_c4_ewald_sample_fixture_is_pinned():
    """The batched-workload fixture: 16 systems spanning N, exact forces on
    the first min(N, 500) particles.  The smallest system is recomputed with
    another splitting (real-space cutoff 0.35 L instead of 0.49 L): the
    exact total force must not change."""
    import os
    import sys
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, os.path.dirname(here))
    sys.path.insert(0, os.path.dirname(os.path.dirname(here)))
    import bench
    d = np.load(os.path.join(here, "ewald_c4_sample.npz"))
    assert d["sys"].shape == (16,) and d["count"].sum() == d["force"].shape[0]
    assert np.all(np.isfinite(d["force"]))
    systems = bench.make_c4()
    ns = np.array([s["n"] for s in systems])
    assert ns[d["sys"]].min() == ns.min() and ns[d["sys"]].max() == ns.max()
    k = int(np.argmin(ns[d["sys"]]))
    s = systems[int(d["sys"][k])]
    cnt = int(d["count"][k])
    start = int(d["count"][:k].sum())
    alpha = 5.6 / (0.35 * s["L"])
    f_r = O.real_space_subset(s["pos"], s["q"], [s["L"]] * 3, 1.0, alpha, 0.35 * s["L"],
                              cnt, exact_erfc=True)
    _, f_k = O.ewald_kspace(s["pos"], s["q"], [s["L"]] * 3, 1.0, alpha, 11.2 * alpha,
                            n_out=cnt)
    assert S.rel_rms(f_r + f_k, d["force"][start:start + cnt]) <= 1e-9


@pytest.mark.parametrize("cao", [5, 7])
def test_gpu_convention_equals_cpu_convention_with_zero_mesh_offset(cao):
    """For odd cao the reference GPU index rule (nmp = floor(x M/L -
    (cao-1)/2 + 1/2)) is the CPU rule with mesh_off = 0, so the P3MGPU
    oracle must agree with the CPU oracle at mesh_off 0 up to its float
    positions (Nyquist terms are negligible at this alpha)."""
    L, mesh, alpha = 10., 24, 0.8
    pos, q = S.uniform_system(1500, L, 71)
    f_gpu = O.gpu_kspace_forces(pos, q, [L] * 3, [mesh] * 3, cao, alpha, 1.3)
    ora = O.OracleP3M([L] * 3, [mesh] * 3, cao, alpha, 3.0, 1.3, mesh_off=(0., 0., 0.))
    f_cpu = ora.forces(pos, q)
    agree = S.rel_rms(f_gpu, f_cpu)
    assert agree <= 2e-6
    # the default CPU mesh (points at cell centres) is a different, equally
    # accurate discretisation: the two differ several times more
    f_mid = O.OracleP3M([L] * 3, [mesh] * 3, cao, alpha, 3.0, 1.3).forces(pos, q)
    assert S.rel_rms(f_gpu, f_mid) > 5 * agree


def test_gpu_convention_even_cao_against_ewald():
    """Even cao (index rule differs from the CPU one): the P3MGPU oracle's
    total force on the C3 fixture subset is within 1.5x the tuned error
    estimate of exact Ewald."""
    import os
    pos, q, box = S.config_c3()
    L = box[0]
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                              "ewald_c3_subset.npz"))
    acc = 1e-3
    t = O.tune_r_cut(1.0, [48] * 3, 4, len(q), float(np.sum(q * q)), acc, [L] * 3)
    r_cut, alpha = t[0] * L, t[1] / L
    f_k = O.gpu_kspace_forces(pos, q, box, [48] * 3, 4, alpha, 1.0)[fx["idx"]]
    order = np.concatenate([fx["idx"], np.setdiff1d(np.arange(len(q)), fx["idx"])])
    f_r = O.real_space_subset(pos[order], q[order], box, 1.0, alpha, r_cut, len(fx["idx"]))
    d = f_k + f_r - fx["force"]
    err = float(np.sqrt(np.mean(np.sum(d * d, axis=1))))
    assert err <= 1.5 * acc, (err, t[2])


def test_threaded_solve_matches_serial():
    """The threaded oracle (all-cores CPU baseline) changes only the order of
    floating-point additions: forces and energy within 1e-12 of the serial
    solve, for cao 7 on a non-cubic mesh with uncharged particles."""
    from oracle import oracle as O
    rng = np.random.default_rng(77)
    n = 6000
    box = [14., 15., 16.]
    pos = rng.uniform(0., 1., (n, 3)) * box
    q = np.where(np.arange(n) % 2 == 0, 1., -1.)
    q[::11] = 0.
    ora = O.OracleP3M(box, [20, 24, 22], 7, 0.9, 4.0, 1.1)
    f1 = ora.forces(pos, q)
    e1 = ora.energy(pos, q)
    try:
        O.set_threads(4)
        f4 = ora.forces(pos, q)
        e4 = ora.energy(pos, q)
    finally:
        O.set_threads(1)
    assert np.sqrt(np.sum((f4 - f1) ** 2) / np.sum(f1 ** 2)) <= 1e-12
    assert abs(e4 - e1) <= 1e-12 * abs(e1)
    assert O.lib().p3mo_get_threads() == 1
