"""The solver inside an MD loop with resident particles (device tensors,
asynchronous solves, energy and virial written on the device): velocity
Verlet over k-space + device real space + a WCA repulsion conserves the total
energy, and the device-side energy and virial equal the host-returned ones.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lattice(n_side, L, seed):
    rng = np.random.default_rng(seed)
    g = (np.arange(n_side) + 0.5) * (L / n_side)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos += rng.normal(scale=0.05, size=pos.shape)
    q = np.where((np.arange(len(pos)) + np.arange(len(pos)) // n_side +
                  np.arange(len(pos)) // n_side ** 2) % 2 == 0, 1., -1.)
    return pos, q


def test_device_results_match_host():
    import torch
    from espresso_amd.electrostatics import P3M
    pos, q = _lattice(10, 12., 1)
    p3m = P3M(prefactor=1.0, accuracy=1e-4, mesh=32, cao=5, r_cut=3.0, alpha=1.0, tune=False)
    p3m.activate(pos, q, box_l=[12.] * 3)
    e_host, v_host = p3m.solve(pos, q, force=np.zeros_like(pos), energy=True, pressure=True)
    tp, tq = torch.tensor(pos, device="cuda"), torch.tensor(q, device="cuda")
    e_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    v_dev = torch.zeros(9, dtype=torch.float64, device="cuda")
    f = torch.zeros_like(tp)
    p3m.solve_to_device(tp, tq, force=f, energy=e_dev, virial=v_dev)
    p3m.synchronize()
    assert abs(e_dev.item() - e_host) <= 1e-12 * abs(e_host)
    assert np.max(np.abs(v_dev.cpu().numpy().reshape(3, 3) - v_host)) <= 1e-12 * np.max(np.abs(v_host))
    with pytest.raises(ValueError, match="P3M_DEVICE_RESULTS requires"):
        from espresso_amd import _lib
        _lib.raise_for(_lib.load().p3m_solve(p3m._ctx, len(q), pos.ctypes.data, q.ctypes.data,
                                             None, None, _lib.P3M_ENERGY | _lib.P3M_DEVICE_RESULTS,
                                             None, None), p3m._ctx)


def test_nve_energy_conservation_with_resident_particles():
    """1000 +-1 charges with a WCA core in a periodic box (density 0.58),
    velocity Verlet dt = 0.002 for 400 steps, all forces on the device each
    step (k-space and real-space P3M asynchronously, WCA in torch); the total
    energy drifts by less than 1e-3 of the kinetic energy."""
    import torch
    from espresso_amd.electrostatics import P3M
    L, n_side = 12., 10
    pos, q = _lattice(n_side, L, 2)
    n = len(q)
    dev = "cuda"
    x = torch.tensor(pos, device=dev)
    tq = torch.tensor(q, device=dev)
    rng = np.random.default_rng(3)
    v = torch.tensor(rng.normal(size=(n, 3)), device=dev)
    v -= v.mean(0)
    p3m = P3M(prefactor=1.0, accuracy=1e-5, mesh=32, cao=7, r_cut=4.0, verbose=False)
    p3m.activate(x, tq, box_l=[L] * 3)
    e_k = torch.zeros(1, dtype=torch.float64, device=dev)

    def wca(xx):
        d = xx[:, None, :] - xx[None, :, :]
        d -= L * torch.round(d / L)
        r2 = (d * d).sum(-1)
        r2.fill_diagonal_(1e6)
        rc2 = 2. ** (1. / 3.)
        m = r2 < rc2
        ir6 = torch.where(m, 1. / r2 ** 3, torch.zeros_like(r2))
        e = torch.where(m, 4. * (ir6 * ir6 - ir6) + 1., torch.zeros_like(r2)).sum() / 2.
        fmag = torch.where(m, 24. * (2. * ir6 * ir6 - ir6) / r2, torch.zeros_like(r2))
        return (fmag[:, :, None] * d).sum(1), e

    def forces():
        f = torch.zeros_like(x)
        p3m.solve_to_device(x, tq, force=f, energy=e_k)
        p3m.add_short_range_forces(x, tq, f, sync=False)
        fw, ew = wca(x)
        p3m.synchronize()
        return f + fw, ew

    def total(f_ew):
        e_sr = p3m.add_short_range_forces(x, tq, torch.zeros_like(x), energy=True)
        return 0.5 * float((v * v).sum()) + e_k.item() + e_sr + float(f_ew)

    dt = 0.002
    f, ew = forces()
    e0 = total(ew)
    ke0 = 0.5 * float((v * v).sum())
    drift = 0.
    for step in range(1, 401):
        v += 0.5 * dt * f
        x += dt * v
        f, ew = forces()
        v += 0.5 * dt * f
        if step % 50 == 0:
            drift = max(drift, abs(total(ew) - e0))
    print(f"NVE 400 steps: max |E - E0| = {drift:.3e} (kinetic {ke0:.3f}, E0 {e0:.3f})")
    assert drift <= 1e-3 * ke0


def test_checkpointed_tuned_actor_gives_the_same_forces():
    """A tuned actor restored from a pickle keeps its tuned parameters, is
    not tuned again on activation, and gives bit-identical forces and energy
    (the reference restores actors from their parameters and is_tuned,
    CoulombP3M.hpp:102-128, test_checkpoint.py:846-866)."""
    import pickle
    from espresso_amd.electrostatics import P3M
    pos, q = _lattice(12, 14., 7)
    p3m = P3M(prefactor=1.0, accuracy=1e-4, timings=2, verbose=False)
    p3m.activate(pos, q, box_l=[14.] * 3)
    g = p3m.get_params()
    assert g["is_tuned"] and g["tune"]
    f1 = np.zeros_like(pos)
    p3m.add_long_range_forces(pos, q, f1)
    e1 = p3m.long_range_energy(pos, q)
    blob = pickle.dumps(p3m)
    del p3m
    r = pickle.loads(blob)
    r.activate(pos, q, box_l=[14.] * 3)
    g2 = r.get_params()
    for k in ("mesh", "cao", "alpha", "r_cut", "accuracy", "is_tuned", "tune"):
        assert g2[k] == g[k], k
    f2 = np.zeros_like(pos)
    r.add_long_range_forces(pos, q, f2)
    assert np.array_equal(f1, f2)
    assert r.long_range_energy(pos, q) == e1
