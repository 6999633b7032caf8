"""Serving loop over resident systems (continuous batching) on HIP device 0.

Served forces are checked against the FP64 oracle on the displaced positions
(FP32 mesh: forces rel-rms <= 2e-5).  The displacement is checked against a
host restatement of its published generator (splitmix64, Box-Muller) to
1e-12.  A request served alone and the same request served in a mix of
overlapping batches agree to rel-rms <= 1e-5: they differ only in the order
of floating-point atomic additions into the mesh.
"""
import numpy as np
import pytest

from tests import p3m_systems as S
from tests.test_gpu_batch import _batch, _device_arrays, _oracle, _systems

pytestmark = pytest.mark.gpu

MASK = (1 << 64) - 1


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK
    return x ^ (x >> 31)


def _displace(pos, box, seed, sigma):
    """Host restatement of the request displacement: for particle i and
    component c, key = seed ^ ((3 i + c) * 0xD1B54A32D192ED03), two
    splitmix64 draws to (0, 1), Box-Muller, fold into the box."""
    out = pos.copy()
    for i in range(len(pos)):
        for c in range(3):
            key = seed ^ (((3 * i + c) * 0xD1B54A32D192ED03) & MASK)
            u1 = ((_splitmix64(key) >> 11) + 0.5) / 2.0 ** 53
            u2 = ((_splitmix64(key ^ 0x8CB92BA72F3D8DD7) >> 11) + 0.5) / 2.0 ** 53
            z = np.sqrt(-2. * np.log(u1)) * np.cos(2. * np.pi * u2)
            L = box[c]
            x = out[i, c] + sigma * z
            x -= L * np.floor(x / L)
            if x >= L:
                x -= L
            out[i, c] = x
    return out


def _server(b, **kw):
    from espresso_amd import P3MServer
    return P3MServer(b, **kw)


def test_served_forces_match_oracle_and_displacement():
    import torch
    systems = _systems(seed=41)
    b = _batch(systems, single=True)
    pos, q = _device_arrays(systems)
    force = torch.zeros_like(pos)
    ids = [s["id"] for s in systems]
    seeds = [1000 + 7 * k for k in range(len(ids))]
    n_out = sum(len(s["q"]) for s in systems)
    cap = torch.zeros((n_out, 3), dtype=torch.float64, device="cuda")
    srv = _server(b, lanes=3, max_batch=2, sigma=0.05)
    res = srv.run(np.zeros(len(ids)), ids, seeds, pos, q, force, capture=cap)
    assert res["batches"] >= 4 and np.all(res["latency_s"] > 0.)
    p_new = pos.cpu().numpy()
    f_cap = cap.cpu().numpy()
    f_res = force.cpu().numpy()
    out = 0
    for s, seed in zip(systems, seeds):
        off, n = b.offset(s["id"]), len(s["q"])
        if n == 0:
            continue
        # displacement reproduces its generator on the host (small systems)
        if n <= 600:
            ref_pos = _displace(s["pos"], s["box"], seed, 0.05)
            assert np.max(np.abs(p_new[off:off + n] - ref_pos)) <= 1e-12
        # served forces (captured and resident) against the oracle on the
        # displaced positions
        f_ref = _oracle(s).forces(p_new[off:off + n], s["q"])
        assert S.rel_rms(f_cap[out:out + n], f_ref) <= 2e-5
        assert np.array_equal(f_cap[out:out + n], f_res[off:off + n])
        out += n


def test_request_alone_equals_in_mix():
    import torch
    systems = _systems(seed=43)
    b = _batch(systems, single=True)
    pos0, q = _device_arrays(systems)
    srv = _server(b, lanes=3, max_batch=2, sigma=0.05)
    ids = [s["id"] for s in systems]
    seeds = [77 + k for k in range(len(ids))]
    n_tot = sum(len(s["q"]) for s in systems)
    for k in (0, 3, 5):
        s = systems[k]
        n = len(s["q"])
        pos = pos0.clone()
        force = torch.zeros_like(pos)
        alone = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        srv.run([0.], [ids[k]], [seeds[k]], pos, q, force, capture=alone)
        pos = pos0.clone()
        force = torch.zeros_like(pos)
        mix = torch.zeros((n_tot, 3), dtype=torch.float64, device="cuda")
        # the request sits in the middle of a burst served by overlapping batches
        order = [j for j in range(len(ids)) if j != k]
        order = order[:3] + [k] + order[3:]
        res = srv.run(np.zeros(len(order)), [ids[j] for j in order],
                      [seeds[j] for j in order], pos, q, force, capture=mix)
        assert res["peak_lanes"] >= 2
        start = sum(len(systems[j]["q"]) for j in order[:order.index(k)])
        a = alone.cpu().numpy()
        m = mix.cpu().numpy()[start:start + n]
        assert S.rel_rms(m, a) <= 1e-5, k


def test_admission_by_working_memory():
    import torch
    systems = _systems(seed=47)
    b = _batch(systems, single=True)
    pos, q = _device_arrays(systems)
    force = torch.zeros_like(pos)
    ids = [s["id"] for s in systems]
    biggest = max(b.working_bytes_of(i) for i in ids)
    srv = _server(b, lanes=4, budget_bytes=biggest + 4096, max_batch=8, sigma=0.)
    reqs = ids * 3
    res = srv.run(np.linspace(0., 2e-3, len(reqs)), reqs, list(range(len(reqs))), pos, q, force)
    assert res["peak_bytes"] <= biggest + 4096
    assert res["batches"] >= 3
    # sigma = 0: forces equal a plain batch solve of the same positions
    ref = torch.zeros_like(pos)
    b.solve(pos, q, ref)
    assert S.rel_rms(force.cpu().numpy(), ref.cpu().numpy()) <= 1e-5
    small = _server(b, lanes=1, budget_bytes=4096, max_batch=8, sigma=0.)
    with pytest.raises(ValueError, match="exceeds the budget"):
        small.run([0.], [ids[3]], [1], pos, q, force)
