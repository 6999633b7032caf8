"""The charge-assignment protocol of CoulombP3M through the C ABI.

The reference's ElectrostaticLayerCorrection fills the P3M charge mesh
itself (prepare_fft_mesh, assign_charge with and without skip_cache), sets
the charge sums (count_charged_particles_elc) and runs the k-space kernel on
that mesh (elc.cpp:1110-1256, p3m.cpp:463-468).  Driven through
p3m_prepare_fft_mesh / p3m_assign_charges / p3m_count_charged_particles_elc /
p3m_solve_assigned exactly as elc.cpp drives them, the FP64 device solver
must give the oracle's P3M forces and energies of the same charge sets
(forces rel-rms <= 1e-6, energy rel <= 1e-8): real charges only, real plus
image charges of the boundary layers, images only.  With charge_assign of
all particles, p3m_solve_assigned equals p3m_solve (1e-12).
"""
import ctypes as C

import numpy as np
import pytest

from tests import p3m_systems as S

pytestmark = pytest.mark.gpu

D = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _context(box, n, pos, q, mesh, cao, alpha, r_cut, epsilon=0.):
    from espresso_amd import _lib
    L = _lib.load()
    ctx = C.c_void_p()
    _lib.raise_for(L.p3m_create(0, C.byref(ctx)))
    p = _lib.p3m_params()
    L.p3m_params_default(C.byref(p))
    for d in range(3):
        p.mesh[d] = mesh
        p.box_l[d] = box[d]
    p.cao, p.alpha, p.r_cut, p.accuracy = cao, alpha, r_cut, 1e-3
    p.prefactor, p.epsilon, p.tune, p.verbose = 1.3, epsilon, 0, 0
    _lib.raise_for(L.p3m_set_params(ctx, C.byref(p)), ctx)
    _lib.raise_for(L.p3m_activate(ctx, n, _ptr(pos), _ptr(q), 0), ctx)
    return L, ctx


def _slab(n=600, seed=61):
    rng = np.random.default_rng(seed)
    box = [12., 12., 14.]          # gap 2: charges in 0 <= z <= h = 12
    pos = rng.uniform(0., 1., (n, 3)) * np.array([12., 12., 12.])
    q = np.where(np.arange(n) % 2 == 0, 1., -1.)
    return box, np.ascontiguousarray(pos), np.ascontiguousarray(q)


def _check(status, L, ctx):
    from espresso_amd import _lib
    _lib.raise_for(status, ctx)


@pytest.mark.parametrize("epsilon", [0., 2.])
def test_charge_assign_then_solve_assigned_equals_solve(epsilon):
    from espresso_amd import _lib
    rng = np.random.default_rng(5)
    n, Lb = 800, 11.
    box = [Lb] * 3
    pos = np.ascontiguousarray(rng.uniform(0., Lb, (n, 3)))
    q = np.where(np.arange(n) % 2 == 0, 1., -1.)
    upos = np.ascontiguousarray(pos + Lb * rng.integers(-1, 2, (n, 3)))
    L, ctx = _context(box, n, pos, q, 16, 5, 0.9, 3.0, epsilon)
    try:
        f_ref = np.zeros((n, 3))
        e_ref = C.c_double()
        _check(L.p3m_solve(ctx, n, _ptr(pos), _ptr(q), _ptr(upos), _ptr(f_ref),
                           _lib.P3M_FORCES | _lib.P3M_ENERGY, C.byref(e_ref), None), L, ctx)
        _check(L.p3m_charge_assign(ctx, n, _ptr(pos), _ptr(q), 0), L, ctx)
        cnt = C.c_int64()
        L.p3m_assigned_count(ctx, C.byref(cnt))
        assert cnt.value == n
        _check(L.p3m_count_charged_particles(ctx, n, _ptr(q), 0), L, ctx)
        f = np.zeros((n, 3))
        e = C.c_double()
        _check(L.p3m_solve_assigned(ctx, _ptr(upos), _ptr(f), _lib.P3M_FORCES | _lib.P3M_ENERGY,
                                    C.byref(e)), L, ctx)
        assert S.rel_rms(f, f_ref) <= 1e-12
        assert abs(e.value - e_ref.value) <= 1e-12 * abs(e_ref.value)
        # the mesh is consumed by the solve
        assert L.p3m_solve_assigned(ctx, None, _ptr(f), _lib.P3M_FORCES, None) != 0
        assert b"no charges assigned" in L.p3m_last_error(ctx)
    finally:
        L.p3m_destroy(ctx)


def test_image_charge_mesh_matches_oracle_composition():
    """elc.cpp add_long_range_forces / long_range_energy with dielectric
    contrasts, step by step, against the oracle's P3M on the same sets."""
    from espresso_amd import _lib
    from oracle import elc as E
    from oracle.oracle import OracleP3M
    box, pos, q = _slab()
    n = len(q)
    mesh, cao, alpha, r_cut = 24, 5, 1.0, 3.0
    d_top, d_bot, space_layer = 0.5, -0.3, 1.5
    h = box[2] - 2.
    L, ctx = _context(box, n, pos, q, mesh, cao, alpha, r_cut)
    ora = OracleP3M(box, [mesh] * 3, cao, alpha, r_cut, 1.3)
    try:
        _check(L.p3m_adapt_epsilon_elc(ctx), L, ctx)
        # image charges as charge_assign<IMAGE> places them (unfolded z)
        bot = (q != 0.) & (pos[:, 2] < space_layer)
        top = (q != 0.) & (pos[:, 2] > h - space_layer)
        ip = np.concatenate([pos[bot] * [1, 1, -1], pos[top] * [1, 1, -1] + [0, 0, 2 * h]])
        iq = np.concatenate([d_bot * q[bot], d_top * q[top]])
        ip, iq = np.ascontiguousarray(ip), np.ascontiguousarray(iq)
        assert len(iq) > 20
        oip, oiq = E.image_charges(pos, q, box, 2., space_layer, d_top, d_bot)
        assert np.allclose(np.sort(oiq), np.sort(iq))
        ipf = ip.copy()
        ipf[:, 2] = E.fold_z(ip, box[2])
        both_p, both_q = np.concatenate([pos, ipf]), np.concatenate([q, iq])

        def sums(qq):
            qq = qq[qq != 0.]
            return len(qq), float(np.sum(qq * qq)), float(np.sum(qq)) ** 2

        # forces: modify_p3m_sums<BOTH>, charge_assign<BOTH>, add_long_range_forces
        _check(L.p3m_count_charged_particles_elc(ctx, *sums(both_q)), L, ctx)
        _check(L.p3m_prepare_fft_mesh(ctx, 1), L, ctx)
        _check(L.p3m_assign_charges(ctx, n, _ptr(pos), _ptr(q), 0, 0), L, ctx)
        _check(L.p3m_assign_charges(ctx, len(iq), _ptr(ip), _ptr(iq), 1, 0), L, ctx)
        cnt = C.c_int64()
        L.p3m_assigned_count(ctx, C.byref(cnt))
        assert cnt.value == n
        f = np.zeros((n, 3))
        _check(L.p3m_solve_assigned(ctx, None, _ptr(f), _lib.P3M_FORCES, None), L, ctx)
        f_ref = ora.forces(both_p, both_q)[:n]
        assert S.rel_rms(f, f_ref) <= 1e-6, S.rel_rms(f, f_ref)

        # energy: E(real), E(real + images), E(images)
        e = C.c_double()
        _check(L.p3m_charge_assign(ctx, n, _ptr(pos), _ptr(q), 0), L, ctx)
        _check(L.p3m_count_charged_particles(ctx, n, _ptr(q), 0), L, ctx)
        _check(L.p3m_solve_assigned(ctx, None, None, _lib.P3M_ENERGY, C.byref(e)), L, ctx)
        e_real = e.value
        _check(L.p3m_prepare_fft_mesh(ctx, 1), L, ctx)
        _check(L.p3m_assign_charges(ctx, n, _ptr(pos), _ptr(q), 0, 0), L, ctx)
        _check(L.p3m_assign_charges(ctx, len(iq), _ptr(ip), _ptr(iq), 1, 0), L, ctx)
        _check(L.p3m_count_charged_particles_elc(ctx, *sums(both_q)), L, ctx)
        _check(L.p3m_solve_assigned(ctx, None, None, _lib.P3M_ENERGY, C.byref(e)), L, ctx)
        e_both = e.value
        _check(L.p3m_prepare_fft_mesh(ctx, 1), L, ctx)
        _check(L.p3m_assign_charges(ctx, len(iq), _ptr(ip), _ptr(iq), 1, 0), L, ctx)
        _check(L.p3m_count_charged_particles_elc(ctx, *sums(iq)), L, ctx)
        _check(L.p3m_solve_assigned(ctx, None, None, _lib.P3M_ENERGY, C.byref(e)), L, ctx)
        e_img = e.value
        for got, ref in ((e_real, ora.energy(pos, q)), (e_both, ora.energy(both_p, both_q)),
                         (e_img, ora.energy(ipf, iq))):
            assert abs(got - ref) <= 1e-8 * abs(ref), (got, ref)
        # device pointers give the same forces
        import torch
        tp, tq = torch.tensor(pos, device="cuda"), torch.tensor(q, device="cuda")
        tip, tiq = torch.tensor(ip, device="cuda"), torch.tensor(iq, device="cuda")
        tf = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        dev = _lib.P3M_DEVICE_POINTERS
        _check(L.p3m_count_charged_particles_elc(ctx, *sums(both_q)), L, ctx)
        _check(L.p3m_prepare_fft_mesh(ctx, 1), L, ctx)
        _check(L.p3m_assign_charges(ctx, n, tp.data_ptr(), tq.data_ptr(), 0, dev), L, ctx)
        _check(L.p3m_assign_charges(ctx, len(iq), tip.data_ptr(), tiq.data_ptr(), 1, dev),
               L, ctx)
        _check(L.p3m_solve_assigned(ctx, None, tf.data_ptr(), _lib.P3M_FORCES | dev, None),
               L, ctx)
        assert S.rel_rms(tf.cpu().numpy(), f) <= 1e-12
    finally:
        L.p3m_destroy(ctx)
