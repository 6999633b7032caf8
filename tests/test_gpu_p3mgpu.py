"""The P3MGPU force convention (index rule and ik of the reference's GPU
kernels, p3m_gpu_cuda.cu:265-360 and 424-540) on the device against the
oracle's restatement of it (oracle.gpu_kspace_forces, pinned in
test_oracle.py against the CPU mesh with zero offset and exact Ewald).

Tolerances: first stencil points bit-exact; FP64 mesh forces rel-rms
<= 1e-10 (same FP32 index, FP64 weights, mesh and transforms on both sides);
the P3MGPU actor (FP32 mesh) forces rel-rms <= 2e-5.  Energy and virial of
the P3MGPU actor come from the CPU kernels' mesh, as in the reference
(p3m.cpp:889-892), so they equal those of an FP32-mesh P3M actor.
"""
import numpy as np
import pytest

from tests import p3m_systems as S

pytestmark = pytest.mark.gpu

PARAMS = dict(prefactor=1.3, accuracy=1e-3, r_cut=2.5, alpha=1.1)


def _actor(cls, mesh, cao, **extra):
    kw = dict(PARAMS)
    kw.update(extra)
    return cls(tune=False, mesh=mesh, cao=cao, **kw)


def _system(n, box, seed):
    pos, q = S.uniform_system(n, box[0], seed)
    pos = pos * (np.asarray(box) / box[0])
    return pos, q


def test_p3mgpu_indices_bit_exact():
    from espresso_amd import _lib
    from espresso_amd.electrostatics import P3M
    from oracle import oracle as O
    box = [10., 11., 12.]
    mesh, cao = [24, 20, 28], 5
    pos, q = _system(2000, box, 3)
    a = np.asarray(box) / np.asarray(mesh)
    # mesh planes and half planes, box faces, positions outside the box
    edge = np.concatenate([np.arange(0, 21) * a[0], np.arange(0, 21) * a[0] + 0.5 * a[0],
                           [0., 1e-15, 10. - 1e-13, np.nextafter(10., 0.), -0.1, -1e-12,
                            10.1, 10., -23.7, 31.3]])
    rng = np.random.default_rng(1)
    epos = np.stack([edge, rng.permutation(edge) * 1.1, rng.permutation(edge) * 1.2], 1)
    allpos = np.concatenate([pos, epos])
    allq = np.concatenate([q, np.where(np.arange(len(epos)) % 2, 1., -1.)])
    p3m = _actor(P3M, mesh, cao, check_neutrality=False)
    _lib.raise_for(_lib.load().p3m_set_force_convention(p3m._ctx, _lib.P3M_CONV_GPU),
                   p3m._ctx)
    p3m.activate(allpos, allq, box_l=box)
    gi = p3m.mesh_indices(allpos, allq)
    nmp, _ = O.gpu_mesh_coordinates(allpos, box, mesh, cao)
    ref = np.mod(nmp, np.asarray(mesh))
    bad = np.nonzero(np.any(gi != ref, axis=1))[0]
    assert len(bad) == 0, (bad[:5], allpos[bad[:5]], gi[bad[:5]], ref[bad[:5]])


@pytest.mark.parametrize("cao,mesh,box", [
    (5, [24, 24, 24], [10., 10., 10.]),
    (7, [32, 24, 28], [10., 9., 11.]),
    (4, [20, 20, 20], [9., 9., 9.]),       # even order: the stencil starts at floor(m + 1/2)
    (1, [16, 16, 16], [8., 8., 8.]),
    (6, [30, 18, 22], [12., 8., 9.]),      # radices 5, 3, 11
])
def test_fp64_mesh_gpu_convention_matches_oracle(cao, mesh, box):
    import torch
    from espresso_amd import _lib
    from espresso_amd.electrostatics import P3M
    from oracle import oracle as O
    pos, q = _system(3000, box, 40 + cao)
    q[::17] = 0.
    pos[:4, 0] -= box[0]
    pos[4:8, 2] += 2 * box[2]
    p3m = _actor(P3M, mesh, cao, check_neutrality=False)
    _lib.raise_for(_lib.load().p3m_set_force_convention(p3m._ctx, _lib.P3M_CONV_GPU),
                   p3m._ctx)
    assert p3m.force_convention() == "gpu"
    p3m.activate(pos, q, box_l=box)
    f = np.zeros((len(q), 3))
    p3m.add_long_range_forces(pos, q, f)
    f_ref = O.gpu_kspace_forces(pos, q, box, mesh, cao, PARAMS["alpha"], PARAMS["prefactor"])
    err = S.rel_rms(f, f_ref)
    print(f"cao {cao} mesh {mesh}: FP64 mesh GPU-convention forces rel-rms {err:.3e}")
    assert err <= 1e-10
    # device tensors, asynchronous enqueue: same forces
    tf = torch.zeros((len(q), 3), dtype=torch.float64, device="cuda")
    p3m.add_long_range_forces(torch.tensor(pos, device="cuda"), torch.tensor(q, device="cuda"),
                              tf, sync=False)
    p3m.synchronize()
    assert S.rel_rms(tf.cpu().numpy(), f) <= 1e-12


def test_p3mgpu_actor_forces_energy_pressure():
    from espresso_amd.electrostatics import P3M, P3MGPU
    from oracle import oracle as O
    box = [10., 10., 10.]
    mesh, cao = [24, 24, 24], 5
    pos, q = _system(4000, box, 77)
    g = _actor(P3MGPU, mesh, cao)
    assert g.force_convention() == "gpu"
    g.activate(pos, q, box_l=box)
    f = np.zeros((len(q), 3))
    g.add_long_range_forces(pos, q, f)
    f_ref = O.gpu_kspace_forces(pos, q, box, mesh, cao, PARAMS["alpha"], PARAMS["prefactor"])
    err = S.rel_rms(f, f_ref)
    print(f"P3MGPU actor forces rel-rms {err:.3e}")
    assert err <= 2e-5
    # the CPU-convention FP32 mesh actor gives different forces (mesh offset)
    c = _actor(P3M, mesh, cao, single_precision=True)
    assert c.force_convention() == "cpu"
    c.activate(pos, q, box_l=box)
    fc = np.zeros((len(q), 3))
    c.add_long_range_forces(pos, q, fc)
    assert S.rel_rms(fc, f) > 5 * err
    # energy and virial: the CPU kernels' mesh, identical to the P3M actor's
    # (spreading flushes with atomics in any order, so repeated FP32 solves
    # agree to a few float ulps, not bit for bit: 1e-6 is about 8 ulps)
    e_g, e_c = g.long_range_energy(pos, q), c.long_range_energy(pos, q)
    assert abs(e_g - e_c) <= 1e-6 * abs(e_c)
    v_c = c.long_range_pressure(pos, q)
    assert np.max(np.abs(g.long_range_pressure(pos, q) - v_c)) <= 1e-6 * np.max(np.abs(v_c))
    ora = O.OracleP3M(box, mesh, cao, PARAMS["alpha"], PARAMS["r_cut"], PARAMS["prefactor"])
    assert abs(e_g - ora.energy(pos, q)) <= 2e-6 * abs(e_g)
    # one call with forces, energy and pressure: GPU-rule forces, CPU-mesh
    # energy and virial
    f2 = np.zeros((len(q), 3))
    e2, v2 = g.solve(pos, q, force=f2, energy=True, pressure=True)
    assert S.rel_rms(f2, f) <= 1e-6
    assert abs(e2 - e_g) <= 1e-6 * abs(e_g)
    assert np.max(np.abs(v2 - v_c)) <= 1e-6 * np.max(np.abs(v_c))
    # after the energy pass the force convention is back in force
    f3 = np.zeros((len(q), 3))
    g.add_long_range_forces(pos, q, f3)
    assert S.rel_rms(f3, f_ref) <= 2e-5


def test_p3mgpu_has_no_dipole_force_term():
    """The reference's GPU force kernel has no epsilon (dipole) term
    (p3m.cpp:867-888); the energy keeps it (CPU kernels)."""
    from espresso_amd.electrostatics import P3MGPU
    box = [10., 10., 10.]
    pos, q = _system(2000, box, 5)
    from espresso_amd.electrostatics import P3M
    upos = pos.copy()
    upos[:100:2, 1] += 10.   # 50 positive charges one image further: dipole moment 500
    out = []
    for cls in (P3MGPU, P3M):
        for eps in (0., 2.0):
            a = _actor(cls, [16] * 3, 4, epsilon=eps, single_precision=True)
            a.activate(pos, q, box_l=box)
            f = np.zeros((len(q), 3))
            a.add_long_range_forces(pos, q, f, pos_unfolded=upos)
            out.append((f, a.long_range_energy(pos, q, pos_unfolded=upos)))
    (fg0, eg0), (fg2, eg2), (fc0, ec0), (fc2, ec2) = out
    # the CPU kernels' dipole force term is far above the FP32 noise ...
    assert S.rel_rms(fc2, fc0) > 1e-3
    # ... and absent from the GPU rule's forces; both energies keep it
    assert S.rel_rms(fg2, fg0) <= 1e-6
    assert abs((eg2 - eg0) - (ec2 - ec0)) <= 1e-5 * abs(ec2 - ec0) + 1e-6 * abs(ec0)


def test_force_convention_errors():
    from espresso_amd import _lib
    from espresso_amd.electrostatics import P3M
    p3m = _actor(P3M, [16] * 3, 3)
    rc = _lib.load().p3m_set_force_convention(p3m._ctx, 2)
    with pytest.raises(ValueError, match="force convention"):
        _lib.raise_for(rc, p3m._ctx)
