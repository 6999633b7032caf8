"""Every workload at its real size on HIP device 0.

C2  N = 10^6 (900k charged), L = 125.99, mesh 128, cao 7, r_cut 2.5 and the
    closed-form alpha at accuracy 1e-5 (1.36793), tune=False, with 8^3 bricks
    and with the column march (FP64 8 x 8, FP32 16 x 16 columns): FP64 forces
    rel-rms <= 1e-6 and energy rel <= 1e-8 against the FP64 oracle
    (restatement of CoulombP3MImpl<double, CPU>, p3m.cpp:452-592); the FP32
    mesh within 2e-5 / 2e-6 (FP32 energy relative to |E| + E_self, the size
    of the sums the energy is the difference of).
C4  the 512 batched systems (seeds 4000 + s, N log-uniform in [500, 20000],
    density U[0.1, 0.8]), tuned per system at 1e-4 with mesh 16-48 and cao
    3-7, solved as one FP32 batch and one FP64 batch: 16 systems spanning the
    N and mesh range against the oracle (FP64 1e-6 / 1e-8, FP32 2e-5 / 2e-6,
    the FP32 energy relative to |E| + E_self),
    and all 512 FP32 batch results against the plain FP64 solver (one P3M
    context per system) within 2e-5, the FP64 batch within 1e-12.
C5  the 2048 resident systems (seeds 5000 + s, N log-uniform in
    [1000, 200000], density U[0.2, 0.8]) tuned at 1e-4 behind the serving
    loop: a burst of requests over overlapping batches; 8 served requests
    against the oracle on their displaced positions (2e-5), and 3 requests
    served alone equal to the same requests in the mix (1e-5: the mix only
    changes the order of floating-point atomic additions into the mesh).
"""
import math

import numpy as np
import pytest

from tests import p3m_systems as S

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[configs] {msg}", flush=True)


def test_c2_full_size_parity():
    import torch
    from espresso_amd.electrostatics import P3M
    from oracle.oracle import OracleP3M
    pos, q, box, params = S.config_c2()
    assert abs(params["alpha"] - 1.36793) < 1e-5
    ora = OracleP3M(box, [128] * 3, 7, params["alpha"], 2.5, 1.0)
    f_ref = ora.forces(pos, q)
    e_ref = ora.energy(pos, q)
    tp = torch.tensor(pos, device="cuda")
    tq = torch.tensor(q, device="cuda")
    for single, f_tol, e_tol, column in ((False, 1e-6, 1e-8, 0), (True, 2e-5, 2e-6, 0),
                                         (False, 1e-6, 1e-8, 8), (True, 2e-5, 2e-6, 16)):
        _say(f"C2 N=1e6 mesh 128 cao 7 single={single} column={column}")
        p3m = P3M(tune=False, single_precision=single, **params)
        p3m.activate(tp, tq, box_l=box)
        if column:
            p3m.set_deterministic(False)
            p3m.set_launch_shape(column=column, seg=32, col_lanes=8)
        f = torch.zeros_like(tp)
        p3m.add_long_range_forces(tp, tq, f)
        e = p3m.long_range_energy(tp, tq)
        err_f = S.rel_rms(f.cpu().numpy(), f_ref)
        scale = S.energy_scale(e_ref, q, params["alpha"], 1.0) if single else abs(e_ref)
        err_e = abs(e - e_ref) / scale
        _say(f"C2 single={single} column={column}: force rel-rms {err_f:.3e}, "
             f"energy rel {err_e:.3e}")
        assert err_f <= f_tol and err_e <= e_tol


@pytest.fixture(scope="module")
def c4():
    import bench
    systems = bench.make_c4()
    tuned = bench.tune_systems(systems, 1e-4, (16, 48), (3, 7))
    return systems, tuned


def _c4_batch(systems, tuned, single):
    from espresso_amd import P3MBatch
    b = P3MBatch(single_precision=single)
    for s, t in zip(systems, tuned):
        b.add_system(box_l=[s["L"]] * 3, n=s["n"], prefactor=1.0, mesh=t["mesh"],
                     cao=t["cao"], alpha=t["alpha"], r_cut=t["r_cut"],
                     accuracy=t["accuracy"])
    return b


def _spanning(keys, k):
    """k indices spread over the sorted order of keys (both ends included)."""
    order = np.argsort(keys, kind="stable")
    return sorted({int(order[int(round(j * (len(keys) - 1) / (k - 1)))]) for j in range(k)})


def test_c4_batch_512_systems(c4):
    import torch
    from espresso_amd.electrostatics import P3M
    from oracle.oracle import OracleP3M
    systems, tuned = c4
    assert len(systems) == 512
    assert all(t["accuracy"] <= 1e-4 for t in tuned)
    n_tot = sum(s["n"] for s in systems)
    tp = torch.tensor(np.concatenate([s["pos"] for s in systems]), device="cuda")
    tq = torch.tensor(np.concatenate([s["q"] for s in systems]), device="cuda")
    offs = np.cumsum([0] + [s["n"] for s in systems])
    res = {}
    for single in (True, False):
        _say(f"C4 512 systems, {n_tot} particles, single={single}")
        b = _c4_batch(systems, tuned, single)
        f = torch.zeros((n_tot, 3), dtype=torch.float64, device="cuda")
        e, _ = b.solve(tp, tq, f, energy=True)
        res[single] = (f.cpu().numpy(), np.asarray(e))
        del b
    # 16 systems spanning N, then mesh, against the oracle
    ns = np.array([s["n"] for s in systems])
    ms = np.array([t["mesh"][0] for t in tuned])
    pick = sorted(set(_spanning(ns, 10)) | set(_spanning(ms, 6)))
    k = 0
    while len(pick) < 16:
        if k not in pick:
            pick.append(k)
        k += 1
    for i in pick:
        s, t = systems[i], tuned[i]
        ora = OracleP3M([s["L"]] * 3, t["mesh"], t["cao"], t["alpha"], t["r_cut"], 1.0)
        f_ref = ora.forces(s["pos"], s["q"])
        e_ref = ora.energy(s["pos"], s["q"])
        for single, f_tol, e_tol in ((False, 1e-6, 1e-8), (True, 2e-5, 2e-6)):
            f, e = res[single]
            err = S.rel_rms(f[offs[i]:offs[i + 1]], f_ref)
            assert err <= f_tol, (i, single, err)
            scale = S.energy_scale(e_ref, s["q"], t["alpha"], 1.0) if single else abs(e_ref)
            assert abs(e[i] - e_ref) <= e_tol * scale, (i, single, e[i], e_ref)
    _say(f"C4 oracle checks on systems {pick} passed")
    # all 512: the plain FP64 solver, one context per system
    f_plain = torch.zeros((n_tot, 3), dtype=torch.float64, device="cuda")
    for i, (s, t) in enumerate(zip(systems, tuned)):
        p = P3M(prefactor=1.0, accuracy=t["accuracy"], mesh=t["mesh"], cao=t["cao"],
                r_cut=t["r_cut"], alpha=t["alpha"], tune=False, verbose=False)
        p.activate(s["pos"], s["q"], box_l=[s["L"]] * 3)
        p.add_long_range_forces(tp[offs[i]:offs[i + 1]], tq[offs[i]:offs[i + 1]],
                                f_plain[offs[i]:offs[i + 1]])
        del p
    fp = f_plain.cpu().numpy()
    worst32 = worst64 = 0.
    for i in range(512):
        sl = slice(offs[i], offs[i + 1])
        worst32 = max(worst32, S.rel_rms(res[True][0][sl], fp[sl]))
        worst64 = max(worst64, S.rel_rms(res[False][0][sl], fp[sl]))
    _say(f"C4 all 512 vs plain FP64: worst FP32 {worst32:.3e}, worst FP64 {worst64:.3e}")
    assert worst32 <= 2e-5 and worst64 <= 1e-12


def test_c5_serving_2048_resident_systems():
    import torch
    import bench
    from espresso_amd import P3MServer
    from oracle.oracle import OracleP3M
    specs = bench.make_c5_specs()
    tuned = bench.tune_systems([{"L": s["L"], "n": s["n"],
                                 "q": np.ones(1) * math.sqrt(s["n"])} for s in specs],
                               1e-4, (8, 128), (1, 7), workers=16)
    n_tot = sum(s["n"] for s in specs)
    offs = np.cumsum([0] + [s["n"] for s in specs])
    _say(f"C5 2048 systems, {n_tot} resident particles")
    pos0 = torch.empty((n_tot, 3), dtype=torch.float64, device="cuda")
    q = torch.empty((n_tot,), dtype=torch.float64, device="cuda")
    host_q = []
    for i, s in enumerate(specs):
        rng = np.random.default_rng(s["seed"])
        p = rng.uniform(0., s["L"], (s["n"], 3))
        qq = np.where(np.arange(s["n"]) % 2 == 0, 1., -1.)
        pos0[offs[i]:offs[i + 1]] = torch.from_numpy(p).to("cuda")
        q[offs[i]:offs[i + 1]] = torch.from_numpy(qq).to("cuda")
        host_q.append(qq)
    from espresso_amd import P3MBatch
    b = P3MBatch(single_precision=True)
    for s, t in zip(specs, tuned):
        b.add_system(box_l=[s["L"]] * 3, n=s["n"], prefactor=1.0, mesh=t["mesh"],
                     cao=t["cao"], alpha=t["alpha"], r_cut=t["r_cut"],
                     accuracy=t["accuracy"])
    srv = P3MServer(b, lanes=4, budget_bytes=512 << 20, max_batch=64, sigma=0.05)
    # the checked systems span N: smallest, largest and between
    ns = np.array([s["n"] for s in specs])
    checked = _spanning(ns, 8)
    rng = np.random.default_rng(5999)
    others = [int(i) for i in rng.choice(np.setdiff1d(np.arange(2048), checked), 120,
                                         replace=False)]
    # a burst: the checked requests spread through 128 requests on distinct systems
    order = others[:]
    for j, c in enumerate(checked):
        order.insert(8 + 15 * j, c)
    seeds = [900_000 + i for i in order]
    pos = pos0.clone()
    force = torch.zeros_like(pos)
    cap = torch.zeros((int(ns[order].sum()), 3), dtype=torch.float64, device="cuda")
    _say(f"C5 burst of {len(order)} requests, checked systems {checked}")
    res = srv.run(np.zeros(len(order)), order, seeds, pos, q, force, capture=cap)
    assert res["peak_lanes"] >= 2 and res["batches"] >= 2
    assert res["peak_bytes"] <= 512 << 20
    p_new = pos.cpu().numpy()
    f_mix = cap.cpu().numpy()
    starts = np.concatenate([[0], np.cumsum(ns[order])])
    for c in checked:
        s, t = specs[c], tuned[c]
        sl = slice(offs[c], offs[c + 1])
        ora = OracleP3M([s["L"]] * 3, t["mesh"], t["cao"], t["alpha"], t["r_cut"], 1.0)
        f_ref = ora.forces(p_new[sl], host_q[c])
        k = order.index(c)
        err = S.rel_rms(f_mix[starts[k]:starts[k + 1]], f_ref)
        _say(f"C5 system {c} (N={s['n']}, mesh {t['mesh'][0]}): rel-rms {err:.3e}")
        assert err <= 2e-5, (c, err)
    # three requests served alone equal the same requests in the mix
    for c in (checked[0], checked[4], checked[7]):
        n = int(ns[c])
        pos = pos0.clone()
        force = torch.zeros_like(pos)
        alone = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        srv.run([0.], [c], [seeds[order.index(c)]], pos, q, force, capture=alone)
        k = order.index(c)
        err = S.rel_rms(alone.cpu().numpy(), f_mix[starts[k]:starts[k + 1]])
        _say(f"C5 system {c} alone vs in mix: rel-rms {err:.3e}")
        assert err <= 1e-5, (c, err)
