"""Deterministic solves on HIP device 0.

FP64 solvers are deterministic by default: the charge assignment rounds each contribution once
to a 64-bit fixed-point value and sums with integer atomics; the gather, the
FFTs and the reductions have a fixed order.  Repeated solves of the same
input must then give identical bits (np.array_equal), also when the
particles arrive in another order (forces, metallic boundary) and when a
system is solved alone or in a mix of systems.  The rounding of the mesh
values (2^-s, s = min(50 - a, 62 - b) for max|q| < 2^a and sum|q| < 2^b)
keeps the results at the FP64 bar against the oracle (forces rel-rms <= 1e-6,
energy rel <= 1e-8).
"""
import numpy as np
import pytest

from tests import p3m_systems as S

pytestmark = pytest.mark.gpu


def _system(n=60_000, L=30., seed=77):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0., L, (n, 3))
    q = np.where(np.arange(n) % 2 == 0, 1., -1.)
    q[0::22] = 0.  # uncharged pairs keep the system neutral
    q[1::22] = 0.
    return pos, q, L


def _actor(L, mesh=40, cao=7, epsilon=0.):
    from espresso_amd.electrostatics import P3M
    p = P3M(prefactor=1.1, accuracy=1e-4, mesh=mesh, cao=cao, r_cut=3.0, alpha=1.0,
            epsilon=epsilon, tune=False, verbose=False)
    return p


def test_repeated_solves_are_bit_identical_and_match_oracle():
    import torch
    from oracle.oracle import OracleP3M
    pos, q, L = _system()
    tp = torch.tensor(pos, device="cuda")
    tq = torch.tensor(q, device="cuda")
    p = _actor(L)
    p.activate(tp, tq, box_l=[L] * 3)
    p.set_deterministic(True)
    runs = []
    for rep in range(3):
        print(f"[det] repeat {rep}", flush=True)
        f = torch.zeros_like(tp)
        p.add_long_range_forces(tp, tq, f)
        e = p.long_range_energy(tp, tq)
        runs.append((f.cpu().numpy(), e))
    for f, e in runs[1:]:
        assert np.array_equal(f, runs[0][0])
        assert e == runs[0][1]
    ora = OracleP3M([L] * 3, [40] * 3, 7, 1.0, 3.0, 1.1)
    f_ref = ora.forces(pos, q)
    e_ref = ora.energy(pos, q)
    assert S.rel_rms(runs[0][0], f_ref) <= 1e-6
    assert abs(runs[0][1] - e_ref) <= 1e-8 * abs(e_ref)
    # grouped spread tiles: the same integer sums, so the same bits
    p.set_launch_shape(group=4, spread="plane")
    f = torch.zeros_like(tp)
    p.add_long_range_forces(tp, tq, f)
    assert np.array_equal(f.cpu().numpy(), runs[0][0])
    p.set_launch_shape(group=1)
    # the floating-point spread agrees to rounding
    p.set_deterministic(False)
    f = torch.zeros_like(tp)
    p.add_long_range_forces(tp, tq, f)
    assert S.rel_rms(f.cpu().numpy(), runs[0][0]) <= 1e-12


def test_forces_do_not_depend_on_particle_order():
    import torch
    pos, q, L = _system(seed=78)
    perm = np.random.default_rng(5).permutation(len(q))
    out = []
    for order in (np.arange(len(q)), perm):
        tp = torch.tensor(pos[order], device="cuda")
        tq = torch.tensor(q[order], device="cuda")
        p = _actor(L, mesh=36, cao=5)
        p.activate(tp, tq, box_l=[L] * 3)
        p.set_deterministic(True)
        f = torch.zeros_like(tp)
        p.add_long_range_forces(tp, tq, f)
        fo = np.empty_like(pos)
        fo[order] = f.cpu().numpy()
        out.append(fo)
        del p
    assert np.array_equal(out[0], out[1])


def test_batch_alone_equals_in_mix_bit_for_bit():
    import torch
    from espresso_amd import P3MBatch
    specs = [(4000, 18., 24, 5), (9000, 25., 32, 7), (2500, 14., 20, 3)]
    systems = []
    for k, (n, L, mesh, cao) in enumerate(specs):
        rng = np.random.default_rng(900 + k)
        systems.append((rng.uniform(0., L, (n, 3)),
                        np.where(np.arange(n) % 2 == 0, 1., -1.), L, mesh, cao))
    b = P3MBatch(single_precision=False)
    for pos, q, L, mesh, cao in systems:
        b.add_system(box_l=[L] * 3, n=len(q), prefactor=1.0, mesh=mesh, cao=cao,
                     alpha=0.9, r_cut=3.0)
    b.set_deterministic(True)
    tp = torch.tensor(np.concatenate([s[0] for s in systems]), device="cuda")
    tq = torch.tensor(np.concatenate([s[1] for s in systems]), device="cuda")
    f_mix = torch.zeros_like(tp)
    e_mix, _ = b.solve(tp, tq, f_mix, energy=True)
    f_mix = f_mix.cpu().numpy()
    offs = np.cumsum([0] + [len(s[1]) for s in systems])
    for k in range(len(systems)):
        f = torch.zeros_like(tp)
        e, _ = b.solve(tp, tq, f, systems=[k], energy=True)
        sl = slice(offs[k], offs[k + 1])
        assert np.array_equal(f.cpu().numpy()[sl], f_mix[sl]), k
        assert e[0] == e_mix[k], k
    # graph replays too
    f1 = torch.zeros_like(tp)
    for _ in range(3):
        f1.zero_()
        b.graph_solve(tp, tq, f1)
        b.synchronize()
        assert np.array_equal(f1.cpu().numpy(), f_mix)


def test_deterministic_argument_errors():
    import torch
    from espresso_amd import P3MBatch
    b32 = P3MBatch(single_precision=True)
    with pytest.raises(ValueError, match="FP64"):
        b32.set_deterministic(True)
    b = P3MBatch(single_precision=False)  # deterministic by default
    with pytest.raises(ValueError, match="8\\^3 bricks"):
        b.set_launch_shape(column=8)
    b.set_deterministic(True)
    with pytest.raises(ValueError, match="8\\^3 bricks"):
        b.set_launch_shape(column=8)
    b.set_deterministic(False)
    b.set_launch_shape(column=8)
    with pytest.raises(ValueError, match="8\\^3 bricks"):
        b.set_deterministic(True)
    with pytest.raises(ValueError, match="brick groups"):
        b.set_launch_shape(column=8, group=4)
    with pytest.raises(ValueError, match="brick groups"):
        b.set_launch_shape(group=4, spread="plane16")
    with pytest.raises(ValueError, match="group must be"):
        b.set_launch_shape(group=8)
