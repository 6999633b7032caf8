"""Parity of the MI355X P3M solver with the FP64 oracle (HIP device 0).

Tolerances: FP64 path forces rel-rms <= 1e-6, energy rel <= 1e-8 (the bar
set for the plain solver); mesh indices bit-exact.  FP32-mesh path: forces
rel-rms <= 2e-5 and energy rel <= 2e-6 against the FP64 oracle.
"""
import numpy as np
import pytest

from tests import p3m_systems as S

pytestmark = pytest.mark.gpu


def _p3m(box, params, single_precision=False, epsilon=0., check_neutrality=True,
         **extra):
    from espresso_amd.electrostatics import P3M
    kw = dict(params)
    kw.update(extra)
    kw.setdefault("prefactor", 1.0)
    return P3M(tune=False, single_precision=single_precision, epsilon=epsilon,
               check_neutrality=check_neutrality, **kw)


def _oracle(box, params, epsilon=0., skin=0.4):
    from oracle.oracle import OracleP3M
    p = dict(params)
    return OracleP3M(box, p["mesh"], p["cao"], p["alpha"], p["r_cut"],
                     p.get("prefactor", 1.0), epsilon=epsilon,
                     mesh_off=p.get("mesh_off", (-1., -1., -1.)), skin=skin)


def _compare(pos, q, box, params, epsilon=0., single=False, f_tol=1e-6,
             e_tol=1e-8, pressure=True, upos=None, check_neutrality=True):
    p3m = _p3m(box, params, single_precision=single, epsilon=epsilon,
               check_neutrality=check_neutrality)
    p3m.activate(pos, q, box_l=box)
    ora = _oracle(box, params, epsilon=epsilon)
    f_gpu = np.zeros((len(q), 3))
    p3m.add_long_range_forces(pos, q, f_gpu, pos_unfolded=upos)
    f_ref = ora.forces(pos, q, pos_unfolded=upos)
    assert S.rel_rms(f_gpu, f_ref) <= f_tol, S.rel_rms(f_gpu, f_ref)
    e_gpu = p3m.long_range_energy(pos, q, pos_unfolded=upos)
    e_ref = ora.energy(pos, q, pos_unfolded=upos)
    assert abs(e_gpu - e_ref) <= e_tol * abs(e_ref), (e_gpu, e_ref)
    if pressure:
        p_gpu = p3m.long_range_pressure(pos, q)
        p_ref = ora.pressure(pos, q)
        scale = np.max(np.abs(p_ref))
        assert np.max(np.abs(p_gpu - p_ref)) <= max(e_tol, f_tol) * scale + 1e-300
    return f_gpu, e_gpu


def test_c1_indices_bit_exact():
    pos, q, box, params = S.config_c1()
    L = box[0]
    rng = np.random.default_rng(7)
    # mesh planes, box faces and positions inside the skin outside the box
    a = L / params["mesh"]
    edge = np.concatenate([
        np.arange(0, 33) * a, np.arange(0, 33) * a + 0.5 * a,
        [0., 1e-15, L - 1e-13, np.nextafter(L, 0.), -0.1, -1e-12, L + 0.1, L]])
    epos = np.stack([edge, rng.permutation(edge), rng.permutation(edge)], 1)
    allpos = np.concatenate([pos, epos])
    allq = np.concatenate([q, np.where(np.arange(len(epos)) % 2, 1., -1.)])
    allq[5] = 0.  # uncharged particles report -1
    p3m = _p3m(box, params, check_neutrality=False)
    p3m.activate(allpos, allq, box_l=box)
    nmp = p3m.mesh_indices(allpos, allq)
    ora = _oracle(box, params)
    for i in range(len(allq)):
        if allq[i] == 0.:
            assert tuple(nmp[i]) == (-1, -1, -1)
            continue
        _, ref, _ = ora.weights(allpos[i])
        assert tuple(nmp[i]) == tuple(ref), (i, allpos[i], nmp[i], ref)


def test_c1_charge_mesh():
    pos, q, box, params = S.config_c1()
    p3m = _p3m(box, params)
    p3m.activate(pos, q, box_l=box)
    m_gpu = p3m.charge_mesh(pos, q)
    m_ref = _oracle(box, params).charge_mesh(pos, q)
    assert np.max(np.abs(m_gpu - m_ref)) <= 1e-12 * np.max(np.abs(m_ref))


def test_c1_forces_energy_pressure():
    pos, q, box, params = S.config_c1()
    _compare(pos, q, box, params)


def test_c1_fp32_mesh():
    pos, q, box, params = S.config_c1()
    _compare(pos, q, box, params, single=True, f_tol=2e-5, e_tol=2e-6)


@pytest.mark.parametrize("cao", [2, 5, 7])
def test_fp32_tile_scale_worst_case(cao):
    """FP32 meshes assign charges into 32-bit fixed-point brick tiles whose
    scale bounds a tile point by n max|q| (peak of the order-cao B-spline)^3
    for a brick of n particles.  2000 equal charges on one point where the
    weights peak reach that bound (the oracle's mesh value there is
    2000 peak^3); the mesh and forces must still match the oracle."""
    L, M = 10., 16
    h = L / M
    rng = np.random.default_rng(400 + cao)
    x0 = 4.5 * h  # all three weights at their peak (checked on the oracle)
    pos = np.concatenate([np.full((2000, 3), x0), rng.uniform(0., L, (2000, 3))])
    q = np.concatenate([np.ones(2000), -np.ones(2000)])
    params = dict(prefactor=1.0, mesh=M, cao=cao, r_cut=2.0, alpha=1.0, accuracy=1e-3)
    p3m = _p3m([L] * 3, params, single_precision=True)
    p3m.activate(pos, q, box_l=[L] * 3)
    m_gpu = p3m.charge_mesh(pos, q)
    m_ref = _oracle([L] * 3, params).charge_mesh(pos, q)
    peak3 = {2: 1., 5: 0.21487695199471937, 7: 0.13345187197856614}[cao]
    assert np.max(m_ref) >= 0.99 * 2000 * peak3  # the bound is (nearly) reached
    # coincident particles round identically, so their errors add: at most
    # n 2^-(s+1) at a point of value ~n peak^3 with 2^(30-s) < 2 n_tile peak^3,
    # i.e. a relative error below n_tile 2^-30 (about 2.1e-6 for the ~2250
    # particles of this tile)
    assert np.max(np.abs(m_gpu - m_ref)) <= 4e-6 * np.max(np.abs(m_ref))
    _compare(pos, q, [L] * 3, params, single=True, f_tol=2e-5, e_tol=2e-6)


@pytest.mark.parametrize("cao", [1, 2, 3, 4, 5, 6, 7])
def test_all_orders(cao):
    pos, q = S.uniform_system(3000, 10., 2000 + cao)
    params = dict(prefactor=1.3, mesh=24, cao=cao, r_cut=2.5, alpha=1.1,
                  accuracy=1e-3)
    _compare(pos, q, [10.] * 3, params)


@pytest.mark.parametrize("mesh,box", [
    ([22, 22, 22], [20., 20., 20.]),      # radix 2 x 11 (generic radix)
    ([42, 42, 42], [12., 12., 12.]),      # 2 x 3 x 7
    ([12, 12, 12], [12., 12., 12.]),
    ([10, 14, 26], [9., 11., 13.]),       # 5, 7, 13; non-cubic box and mesh
    ([16, 20, 24], [10., 12., 14.]),
    ([64, 32, 48], [20., 10., 15.]),
    ([46, 34, 38], [15., 11., 12.]),      # 2 x 23, 2 x 17, 2 x 19 (paired first stage)
    ([40, 36, 44], [13., 12., 14.]),      # 8 x 5, 4 x 3 x 3, 11 x 4
    # two radices above 8 in one direction (a generic radix after the first
    # stage), in x (column transforms of the fused pass), y and z
    ([242, 8, 10], [40., 8., 7.]),        # 11 x 11 x 2
    ([8, 242, 10], [8., 40., 7.]),
    ([10, 8, 242], [7., 8., 40.]),
])
def test_meshes(mesh, box):
    pos, q = S.uniform_system(2000, 1., 31)
    pos = pos * np.asarray(box)
    params = dict(prefactor=0.7, mesh=mesh, cao=5, r_cut=2.0, alpha=1.2,
                  accuracy=1e-3)
    _compare(pos, q, box, params)


def test_uncharged_and_nonneutral():
    pos, q = S.uniform_system(4000, 12., 55, charges="shuffled",
                              neutral_zero_frac=0.3)
    q[:7] = 2.0  # net charge: exercises the net-charge energy term
    params = dict(prefactor=1.0, mesh=32, cao=6, r_cut=3.0, alpha=1.0,
                  accuracy=1e-3)
    _compare(pos, q, [12.] * 3, params, check_neutrality=False)


def test_dipole_correction():
    # closed form of electrostatic_interactions.py:103-125 via the oracle
    pos, q = S.uniform_system(1500, 10., 77)
    upos = pos + 10. * np.random.default_rng(3).integers(-2, 3, pos.shape)
    params = dict(prefactor=1.1, mesh=20, cao=7, r_cut=3.0, alpha=1.0,
                  accuracy=1e-3)
    for eps in (0.1, 1.0, 80.):
        _compare(pos, q, [10.] * 3, params, epsilon=eps, upos=upos)


def test_two_charges_reference_values():
    pos, q, box, params = S.two_charges()
    pref = 1.1
    params = dict(params, prefactor=pref)
    p3m = _p3m(box, params)
    p3m.activate(pos, q, box_l=box)
    f = np.zeros((2, 3))
    p3m.add_long_range_forces(pos, q, f)
    e = p3m.long_range_energy(pos, q)
    from oracle.oracle import real_space
    er, fr = real_space(pos, q, box, pref, params["alpha"], params["r_cut"])
    np.testing.assert_allclose(e + er, -0.501062398379 * pref, atol=1e-5)
    np.testing.assert_allclose((f + fr)[0], [0.248921612 * pref, 0, 0], atol=1e-5)


@pytest.mark.parametrize("case", ["cube", "wire", "sheet"])
def test_madelung(case):
    pos, q, box, params, mc, base = {"cube": S.madelung_cube,
                                     "wire": S.madelung_wire,
                                     "sheet": S.madelung_sheet}[case]()
    rtol = {"cube": 1e-6, "wire": 5e-7, "sheet": 1e-7}[case]
    p3m = _p3m(box, params)
    p3m.activate(pos, q, box_l=box)
    from oracle.oracle import real_space
    ek, vk = p3m.solve(pos, q, energy=True, pressure=True)
    er, _, vr = real_space(pos, q, box, 1.0, params["alpha"], params["r_cut"],
                           virial=True)
    n = len(q)
    V = float(np.prod(box))
    np.testing.assert_allclose(2. * (ek + er) / n, mc, atol=0., rtol=rtol)
    p_tensor = 2. * (vk + vr) / n
    ref = mc * np.diag(base) / np.sum(base)
    np.testing.assert_allclose(p_tensor, ref, atol=5e-9, rtol=5e-6)
    _ = V


def test_empty_and_all_zero():
    box = [10.] * 3
    params = dict(prefactor=1.0, mesh=16, cao=5, r_cut=2.0, alpha=1.0,
                  accuracy=1e-3)
    p3m = _p3m(box, params, check_neutrality=False)
    pos = np.zeros((0, 3))
    q = np.zeros(0)
    p3m.activate(pos, q, box_l=box)
    assert p3m.long_range_energy(pos, q) == 0.
    pos2 = np.random.default_rng(0).uniform(0, 10, (50, 3))
    q2 = np.zeros(50)
    f = np.ones((50, 3))
    p3m.add_long_range_forces(pos2, q2, f)
    assert np.all(f == 1.)
    assert p3m.long_range_energy(pos2, q2) == 0.


def test_forces_accumulate_in_place():
    pos, q = S.uniform_system(1000, 8., 9)
    params = dict(prefactor=1.0, mesh=16, cao=4, r_cut=2.0, alpha=1.3,
                  accuracy=1e-3)
    p3m = _p3m([8.] * 3, params)
    p3m.activate(pos, q, box_l=[8.] * 3)
    f0 = np.random.default_rng(1).normal(size=(1000, 3))
    f = f0.copy()
    p3m.add_long_range_forces(pos, q, f)
    df = _oracle([8.] * 3, params).forces(pos, q)
    assert S.rel_rms(f - f0, df) <= 1e-6


def test_device_pointers_match_host():
    torch = pytest.importorskip("torch")
    pos, q, box, params = S.config_c1()
    p3m = _p3m(box, params)
    p3m.activate(pos, q, box_l=box)
    f_h = np.zeros((len(q), 3))
    p3m.add_long_range_forces(pos, q, f_h)
    tp = torch.tensor(pos, device="cuda")
    tq = torch.tensor(q, device="cuda")
    tf = torch.zeros((len(q), 3), dtype=torch.float64, device="cuda")
    p3m.add_long_range_forces(tp, tq, tf)
    torch.cuda.synchronize()
    assert S.rel_rms(tf.cpu().numpy(), f_h) <= 1e-12


def test_c2_subset_parity_fp64():
    """C2 geometry (mesh 128, cao 7) on a 100k-particle sample at the same
    density, FP64 path against the oracle."""
    pos, q, box, params = S.config_c2(n=100_000)
    params = dict(params, mesh=64)
    _compare(pos, q, box, params, pressure=False)


def test_pair_forces_on_targets_match_oracle():
    # device real-space pair forces (A&S erfc) on a subset from all sources
    import torch
    from oracle import oracle as O
    pos, q, box, params = S.config_c1()
    q = q.copy()
    q[5] = 0.  # uncharged target and source
    p3m = _p3m(box, params, check_neutrality=False)
    p3m.activate(pos, q, box_l=box)
    idx = np.array([0, 5, 17, 123, 9999, 4321], dtype=np.int32)
    out = p3m.pair_forces_on(torch.tensor(pos, device="cuda"),
                             torch.tensor(q, device="cuda"), idx).cpu().numpy()
    order = np.concatenate([idx, np.setdiff1d(np.arange(len(q)), idx)])
    ref = O.real_space_subset(pos[order], q[order], box, params["prefactor"], params["alpha"],
                              params["r_cut"], len(idx))
    assert np.all(out[1] == 0.)
    assert S.rel_rms(out, ref) <= 1e-12


def test_actor_autotune_keeps_forces():
    import torch
    pos, q, box, params = S.config_c1()
    for single in (False, True):
        p3m = _p3m(box, params, single_precision=single)
        p3m.activate(pos, q, box_l=box)
        tp, tq = torch.tensor(pos, device="cuda"), torch.tensor(q, device="cuda")
        ref = torch.zeros_like(tp)
        p3m.add_long_range_forces(tp, tq, ref)
        shape = p3m.autotune(tp, tq, reps=1)
        assert shape["gather"] in ("layer", "plane", "plane16", "plane8")
        assert 0. < shape["best_ms"] <= shape["default_ms"]
        f = torch.zeros_like(tp)
        p3m.add_long_range_forces(tp, tq, f)
        tol = 1e-5 if single else 1e-12
        assert S.rel_rms(f.cpu().numpy(), ref.cpu().numpy()) <= tol
