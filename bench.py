#!/usr/bin/env python3
"""P3M k-space force-solve throughput on one MI355X.

Workload (single large system): N = 1,000,000 charges uniform in a cubic box
L = 125.99 (density 0.5), 450k +1 / 450k -1 / 100k uncharged, prefactor 1,
mesh 128^3, cao 7, accuracy target 1e-5 with r_cut and alpha from the
solver's own tuning (the k-space cost does not depend on r_cut or alpha).
A step is one k-space force solve (bin -> spread -> r2c FFT -> influence /
ik convolution -> c2r FFTs -> gather) with positions, charges and forces
resident in HBM.  value = N / t_step (Mparticles/s).

Also reported: the FP32-mesh solver on the same input (with its force error
against the FP64 solve), per-stage HIP-event times over the timed region,
the roofline of the dominant kernel and the FP64 CPU oracle timed on this
host on one full solve.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles",
                            "hbm_traffic_c2_n1e6_mesh128_cao7.json")
SEGMENT = {"plane": 64, "plane16": 16, "plane8": 8}


def stage_kernels(shape):
    """Kernel behind each single-kernel stage of the FP64 solve at cao 7 and
    mesh 128 (split FFT passes) for the chosen launch shape."""
    t = shape["threads"]
    tg = shape.get("gather_threads") or t
    g = shape["gather"]
    if shape.get("column"):
        w, lanes = shape["column"], shape["col_lanes"]
        return {"spread": f"k_spread_col<7, double, {w}, {lanes}>",
                "convolve": "k_fft_x_conv<double>",
                "gather": f"k_gather_col<7, double, {w}, {lanes}>"}
    gather = (f"k_gather_layer<7, double, {tg}>" if g == "layer"
              else f"k_gather_plane<7, double, {SEGMENT[g]}, {tg}>")
    grp = "2, 2, 1" if shape.get("group", 1) == 4 else "1, 1, 1"
    return {"spread": f"k_spread<7, double, {SEGMENT[shape['spread']]}, {t}, {grp}>",
            "convolve": "k_fft_x_conv<double>", "gather": gather}


# LDS of one MI355X: 256 CUs x 256 bytes per clock x 2.4 GHz (a stream of
# conflict-free ds_read_b128 measured 96 TB/s on the device, one of
# ds_read_b64 67.5 TB/s)
LDS_PEAK_GBS = 256 * 256 * 2.4


def lds_bytes_gather(n_q, cao, s_m):
    """LDS bytes the gather reads per launch: the three field values of each
    of the cao^3 stencil points of every charged particle."""
    return 3 * cao ** 3 * s_m * n_q


def make_system(n=1_000_000, seed=1002):
    L = 125.99 * (n / 1_000_000) ** (1. / 3.)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0., L, (n, 3))
    n0 = int(round(0.1 * n))
    npos = (n - n0) // 2
    q = np.concatenate([np.ones(npos), -np.ones(n - n0 - npos), np.zeros(n0)])
    q = q[rng.permutation(n)]
    return pos, q, L


def algorithmic_bytes(n_q, mesh, s_m, s_g=None):
    """Bytes per launch of the three single-kernel stages and of one whole
    solve (particles: 3 position + 1 charge value of 8 bytes)."""
    s_g = s_m if s_g is None else s_g
    R = mesh[0] * mesh[1] * mesh[2]
    K = mesh[0] * mesh[1] * (mesh[2] // 2 + 1)
    part = 32 * n_q
    return {
        "spread": part + R * s_m,
        "convolve": K * 2 * s_m + K * s_g + 2 * K * 2 * s_m,
        "gather": part + 3 * R * s_m + 3 * n_q * 8,
        "solve": 2 * part + 3 * n_q * 8 + 8 * R * s_m + 16 * K * s_m + K * s_g,
    }


def make_c4(n_sys=512, seed0=4000):
    """512 independent systems: N log-uniform in [500, 20000] (even), density
    uniform in [0.1, 0.8], +-1 alternating charges, uniform positions."""
    out = []
    for s in range(n_sys):
        rng = np.random.default_rng(seed0 + s)
        n = int(2 * round(math.exp(rng.uniform(math.log(500), math.log(20000))) / 2))
        rho = rng.uniform(0.1, 0.8)
        L = (n / rho) ** (1. / 3.)
        pos = rng.uniform(0., L, (n, 3))
        q = np.where(np.arange(n) % 2 == 0, 1., -1.)
        out.append({"pos": pos, "q": q, "L": L, "n": n})
    return out


# device real-space cost per pair inside r_cut: p3m_short_range on the
# N = 10^6 workload (r_cut 8.558, 5.3e8 pairs) took 9.6 ms on one MI355X;
# bench.py measures it again in its short-range leg and tunes with that rate
DEVICE_PAIR_NS = 0.018
# the 512-system batch is tuned with a margin below its 1e-4 rms force error
# target: the P3M error estimate runs 3-8% below the error measured against
# exact Ewald sums on these systems
C4_TUNE_ACCURACY = 0.9e-4


def model_pairs(n_q, L, r_cut):
    """Pairs within r_cut of a uniform system (the tuning model's count)."""
    return 0.5 * n_q * (n_q / L ** 3) * (4. / 3.) * math.pi * r_cut ** 3


def tune_systems(systems, accuracy, mesh_range, cao_range, workers=8, pair_ns=None,
                 single=True):
    """Per-system tuning with the reference algorithm and the device cost
    model: k-space stage rates of the solver (FP32 mesh if single) plus the
    device real-space pair loop (p3m_short_range) at pair_ns per pair
    (DEVICE_PAIR_NS by default)."""
    from concurrent.futures import ThreadPoolExecutor
    from espresso_amd import tune_model
    pns = DEVICE_PAIR_NS if pair_ns is None else pair_ns

    def one(s):
        return tune_model([s["L"]] * 3, s["n"], float(np.sum(s["q"] ** 2)), 1.0,
                          accuracy, mesh_range=mesh_range, cao_range=cao_range,
                          pair_ns=pns, single_precision=single)
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(one, systems))


def bench_batched(steps, warmup, torch, dev, pair_ns=None):
    """512 systems per launch (FP32 mesh, FP64 reductions and forces),
    autotuned launch shape, hipGraph replay; against the plain FP64 solver
    applied to the same systems one context per system.  The total force
    (batched k-space + device real space) of 16 systems spanning N is
    checked against exact Ewald forces (tests/golden/ewald_c4_sample.npz)."""
    from espresso_amd import P3MBatch
    from espresso_amd.electrostatics import P3M
    systems = make_c4()
    t0 = time.perf_counter()
    tuned = tune_systems(systems, C4_TUNE_ACCURACY, (16, 48), (3, 7), pair_ns=pair_ns)
    t_tune = time.perf_counter() - t0
    n_tot = sum(s["n"] for s in systems)
    print(f"[bench] batched: {len(systems)} systems, {n_tot} particles, tuned "
          f"in {t_tune:.1f} s", flush=True)
    tp = torch.tensor(np.concatenate([s["pos"] for s in systems]), device=dev)
    tq = torch.tensor(np.concatenate([s["q"] for s in systems]), device=dev)
    b = P3MBatch(single_precision=True)
    for s, t in zip(systems, tuned):
        b.add_system(box_l=[s["L"]] * 3, n=s["n"], prefactor=1.0, mesh=t["mesh"],
                     cao=t["cao"], alpha=t["alpha"], r_cut=t["r_cut"],
                     accuracy=t["accuracy"])
    scratch = torch.zeros((n_tot, 3), dtype=torch.float64, device=dev)
    shape = b.autotune(tp, tq, scratch, reps=5)
    print(f"[bench] batched autotune: {shape}", flush=True)
    tf = torch.zeros_like(scratch)
    for _ in range(warmup):
        b.graph_solve(tp, tq, tf)
    b.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.graph_solve(tp, tq, tf)
    b.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    # per-batch latency: each replay bracketed by events on the batch stream
    ext = torch.cuda.ExternalStream(b.stream())
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    for e0, e1 in evs:
        e0.record(ext)
        b.graph_solve(tp, tq, tf)
        e1.record(ext)
    b.synchronize()
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    # stage split of the same selection from eager solves
    b.profile_begin(steps)
    for _ in range(steps):
        b.solve(tp, tq, scratch, sync=False)
    nprof, st = b.profile_end()
    st = {k: round(v / max(nprof, 1), 4) for k, v in st.items()}

    # plain FP64 solver, one context per system, solves issued back to back
    plain = []
    for s, t in zip(systems, tuned):
        p = P3M(prefactor=1.0, accuracy=t["accuracy"], mesh=t["mesh"], cao=t["cao"],
                r_cut=t["r_cut"], alpha=t["alpha"], tune=False, verbose=False)
        p.activate(s["pos"], s["q"], box_l=[s["L"]] * 3)
        plain.append(p)
    offs = np.cumsum([0] + [s["n"] for s in systems])
    f64 = torch.zeros_like(scratch)
    views = [(tp[offs[i]:offs[i + 1]], tq[offs[i]:offs[i + 1]],
              f64[offs[i]:offs[i + 1]]) for i in range(len(systems))]
    # the timed passes call the C ABI directly (what a C++ caller pays):
    # asynchronous device-pointer solves, each context on its own stream
    from espresso_amd import _lib
    L = _lib.load()
    flags = _lib.P3M_FORCES | _lib.P3M_DEVICE_POINTERS | _lib.P3M_ASYNC
    calls = [(p._ctx, vq.shape[0], vp.data_ptr(), vq.data_ptr(), vf.data_ptr())
             for p, (vp, vq, vf) in zip(plain, views)]
    torch.cuda.synchronize()

    def plain_pass():
        for ctx, n, pp, qp, fp in calls:
            _lib.raise_for(L.p3m_solve(ctx, n, pp, qp, None, fp, flags, None, None), ctx)

    plain_pass()  # warm-up
    for p in plain:
        p.synchronize()
    reps = max(1, min(steps, 3))
    t0 = time.perf_counter()
    for _ in range(reps):
        plain_pass()
    for p in plain:
        p.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    # FP32 batch vs FP64 plain forces (one solve each)
    f32 = torch.zeros_like(scratch)
    b.solve(tp, tq, f32)
    f64.zero_()
    torch.cuda.synchronize()
    for p, (vp, vq, vf) in zip(plain, views):
        p.add_long_range_forces(vp, vq, vf)
    d = (f32 - f64).cpu().numpy()
    ref = f64.cpu().numpy()
    rel = float(np.sqrt(np.sum(d * d) / np.sum(ref * ref)))
    # total force error against exact Ewald on the fixture's systems:
    # batched FP32 k-space + the device real-space pair forces
    ewald = None
    fx_path = os.path.join(GOLDEN, "ewald_c4_sample.npz")
    if os.path.exists(fx_path):
        fx = np.load(fx_path)
        f32n = f32.cpu().numpy()
        errs, start = [], 0
        for sid, cnt in zip(fx["sys"], fx["count"]):
            sid, cnt = int(sid), int(cnt)
            vp, vq, _ = views[sid]
            f_r = plain[sid].pair_forces_on(vp, vq, np.arange(cnt)).cpu().numpy()
            dd = f32n[offs[sid]:offs[sid] + cnt] + f_r - fx["force"][start:start + cnt]
            errs.append(float(np.sqrt(np.mean(np.sum(dd * dd, axis=1)))))
            start += cnt
        ewald = {"systems": [int(x) for x in fx["sys"]],
                 "rms_force_error": [round(e, 8) for e in errs],
                 "max_rms_force_error": max(errs),
                 "estimate_max": max(tuned[int(i)]["accuracy"] for i in fx["sys"]),
                 "reference": "exact FP64 Ewald, first min(N, 500) particles of 16 "
                              "systems spanning N (tests/golden/ewald_c4_sample.npz); "
                              "total force = batched FP32 k-space + device real space"}
    meshes = [t["mesh"][0] for t in tuned]
    caos = [t["cao"] for t in tuned]
    res = {
        "workload": "512 systems, N log-uniform [500, 20000], density U[0.1, 0.8], "
                    "+-1 alternating, per-system tuning at 0.9e-4 for a measured rms force error <= 1e-4 (mesh 16-48, cao 3-7)",
        "n_systems": len(systems), "n_particles": n_tot,
        "mesh_range": [min(meshes), max(meshes)], "cao_range": [min(caos), max(caos)],
        "max_estimated_rms_force_error": max(t["accuracy"] for t in tuned),
        "ms_per_batch": round(t_batch * 1e3, 4),
        "batch_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                             "p99": round(float(np.percentile(lat, 99)), 4),
                             "replays": len(lat)},
        "mparticles_per_s": round(n_tot / t_batch / 1e6, 2),
        "solves_per_s": round(len(systems) / t_batch, 1),
        "stages_ms_eager": st,
        "launch_shape": shape,
        "plain_fp64_ms_per_pass": round(t_plain * 1e3, 3),
        "plain_fp64_mparticles_per_s": round(n_tot / t_plain / 1e6, 2),
        "speedup_vs_plain_fp64": round(t_plain / t_batch, 2),
        "force_rel_rms_fp32_batch_vs_fp64_plain": rel,
        "ewald_check": ewald,
        "tuning_pair_ns": DEVICE_PAIR_NS if pair_ns is None else round(pair_ns, 5),
        "tuning_s": round(t_tune, 2),
    }
    print(f"[bench] batched: {res['ms_per_batch']} ms/batch, "
          f"{res['mparticles_per_s']} Mparticles/s, {res['solves_per_s']} solves/s; "
          f"plain FP64 {res['plain_fp64_mparticles_per_s']} Mparticles/s "
          f"(x{res['speedup_vs_plain_fp64']}); rel-rms {rel:.2e}; latency "
          f"{res['batch_latency_ms']}; Ewald max rms err "
          f"{ewald['max_rms_force_error'] if ewald else None}", flush=True)
    del plain
    return res


def bench_c3_batched(steps, warmup, torch, dev, pair_ns=None, n_sys=8):
    """The batched form on systems of the precision-sweep shape: n_sys
    systems of N = 250,000 charges +-1 uniform in L = 79.37 (density 0.5),
    tuned at 1e-4 within mesh 48-96 and cao 3-7 (all systems share N and L,
    so one tuning serves all).  The FP32 batch (hipGraph replay, autotuned
    launch shape) against the plain FP64 solver applied per system, with the
    batch latency p50/p99."""
    from espresso_amd import P3MBatch, _lib
    from espresso_amd.electrostatics import P3M
    L, n = 79.37, 250_000
    systems = []
    for k in range(n_sys):
        rng = np.random.default_rng(3000 + k)
        systems.append({"pos": rng.uniform(0., L, (n, 3)), "L": L, "n": n,
                        "q": np.where(np.arange(n) % 2 == 0, 1., -1.)})
    t = tune_systems(systems[:1], 1e-4, (48, 96), (3, 7), pair_ns=pair_ns)[0]
    tp = torch.tensor(np.concatenate([s["pos"] for s in systems]), device=dev)
    tq = torch.tensor(np.concatenate([s["q"] for s in systems]), device=dev)
    n_tot = n * n_sys
    b = P3MBatch(single_precision=True)
    for _ in systems:
        b.add_system(box_l=[L] * 3, n=n, prefactor=1.0, mesh=t["mesh"], cao=t["cao"],
                     alpha=t["alpha"], r_cut=t["r_cut"], accuracy=t["accuracy"])
    f32 = torch.zeros((n_tot, 3), dtype=torch.float64, device=dev)
    shape = b.autotune(tp, tq, f32, reps=5)
    for _ in range(warmup):
        b.graph_solve(tp, tq, f32)
    b.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.graph_solve(tp, tq, f32)
    b.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    ext = torch.cuda.ExternalStream(b.stream())
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    for e0, e1 in evs:
        e0.record(ext)
        b.graph_solve(tp, tq, f32)
        e1.record(ext)
    b.synchronize()
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    plain = []
    for s in systems:
        p = P3M(prefactor=1.0, accuracy=t["accuracy"], mesh=t["mesh"], cao=t["cao"],
                r_cut=t["r_cut"], alpha=t["alpha"], tune=False, verbose=False)
        p.activate(s["pos"], s["q"], box_l=[L] * 3)
        plain.append(p)
    f64 = torch.zeros_like(f32)
    views = [(tp[k * n:(k + 1) * n], tq[k * n:(k + 1) * n], f64[k * n:(k + 1) * n])
             for k in range(n_sys)]
    lib = _lib.load()
    flags = _lib.P3M_FORCES | _lib.P3M_DEVICE_POINTERS | _lib.P3M_ASYNC
    calls = [(p._ctx, n, vp.data_ptr(), vq.data_ptr(), vf.data_ptr())
             for p, (vp, vq, vf) in zip(plain, views)]
    torch.cuda.synchronize()

    def plain_pass():
        for ctx, nn, pp, qp, fp in calls:
            _lib.raise_for(lib.p3m_solve(ctx, nn, pp, qp, None, fp, flags, None, None), ctx)

    plain_pass()
    for p in plain:
        p.synchronize()
    reps = max(3, steps // 2)
    t0 = time.perf_counter()
    for _ in range(reps):
        plain_pass()
    for p in plain:
        p.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    f32.zero_()
    b.solve(tp, tq, f32)
    f64.zero_()
    torch.cuda.synchronize()
    for p, (vp, vq, vf) in zip(plain, views):
        p.add_long_range_forces(vp, vq, vf)
    d = (f32 - f64).cpu().numpy()
    ref = f64.cpu().numpy()
    res = {"workload": f"{n_sys} systems of N=250000 +-1 in L=79.37 (density 0.5), tuned at "
                       "1e-4 within mesh 48-96, cao 3-7; FP32 batch vs plain FP64 per system",
           "mesh": t["mesh"][0], "cao": t["cao"], "r_cut": round(t["r_cut"], 4),
           "estimated_rms_force_error": t["accuracy"],
           "ms_per_batch": round(t_batch * 1e3, 4),
           "batch_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                                "p99": round(float(np.percentile(lat, 99)), 4),
                                "replays": len(lat)},
           "mparticles_per_s": round(n_tot / t_batch / 1e6, 2),
           "plain_fp64_ms_per_pass": round(t_plain * 1e3, 3),
           "plain_fp64_mparticles_per_s": round(n_tot / t_plain / 1e6, 2),
           "speedup_vs_plain_fp64": round(t_plain / t_batch, 2),
           "force_rel_rms_fp32_batch_vs_fp64_plain":
               float(np.sqrt(np.sum(d * d) / np.sum(ref * ref))),
           "launch_shape": shape}
    print(f"[bench] C3-shaped batch: {res['ms_per_batch']} ms, {res['mparticles_per_s']} "
          f"Mparticles/s; plain FP64 {res['plain_fp64_mparticles_per_s']} (x"
          f"{res['speedup_vs_plain_fp64']}); latency {res['batch_latency_ms']}", flush=True)
    del plain
    return res


def make_c5_specs(n_sys=2048, seed0=5000):
    """2048 resident systems: N log-uniform in [1000, 200000] (even), density
    uniform in [0.2, 0.8]."""
    out = []
    for s in range(n_sys):
        rng = np.random.default_rng(seed0 + s)
        n = int(2 * round(math.exp(rng.uniform(math.log(1000), math.log(200000))) / 2))
        rho = rng.uniform(0.2, 0.8)
        out.append({"n": n, "L": (n / rho) ** (1. / 3.), "seed": seed0 + s})
    return out


def bench_serving(torch, dev, rates=(50, 100, 200, 500, 1000, 2000, 5000, 10000, 15000),
                  seconds=1.5,
                  n_sys=2048, lanes=4, budget_bytes=512 << 20, max_batch=64):
    """Poisson request streams over resident systems: continuous batching,
    admission by working memory, p50/p99 latency and solves/s per offered
    load, and the saturation rate of a burst."""
    from espresso_amd import P3MBatch, P3MServer
    specs = make_c5_specs(n_sys)
    t0 = time.perf_counter()
    # tuning needs only N, sum q^2 and L
    tuned = tune_systems([{"L": s["L"], "n": s["n"], "q": np.ones(1) * math.sqrt(s["n"])}
                          for s in specs], 1e-4, (8, 128), (1, 7), workers=16)
    t_tune = time.perf_counter() - t0
    n_tot = sum(s["n"] for s in specs)
    print(f"[bench] serving: {n_sys} systems, {n_tot} resident particles, tuned in "
          f"{t_tune:.1f} s", flush=True)
    pos = torch.empty((n_tot, 3), dtype=torch.float64, device=dev)
    q = torch.empty((n_tot,), dtype=torch.float64, device=dev)
    off = 0
    for s in specs:
        rng = np.random.default_rng(s["seed"])
        n = s["n"]
        pos[off:off + n] = torch.from_numpy(rng.uniform(0., s["L"], (n, 3))).to(dev)
        q[off:off + n] = torch.from_numpy(np.where(np.arange(n) % 2 == 0, 1., -1.)).to(dev)
        off += n
    force = torch.zeros_like(pos)
    b = P3MBatch(single_precision=True)
    for s, t in zip(specs, tuned):
        b.add_system(box_l=[s["L"]] * 3, n=s["n"], prefactor=1.0, mesh=t["mesh"],
                     cao=t["cao"], alpha=t["alpha"], r_cut=t["r_cut"], accuracy=t["accuracy"])
    resident = sum(b.system_bytes(i) for i in range(n_sys))
    t_setup = time.perf_counter() - t0
    print(f"[bench] serving: resident meshes {resident / 2**30:.2f} GiB, set up in "
          f"{t_setup:.1f} s", flush=True)
    srv = P3MServer(b, lanes=lanes, budget_bytes=budget_bytes, max_batch=max_batch,
                    sigma=0.05)
    stream_rng = np.random.default_rng(5999)
    counts = np.array([s["n"] for s in specs])
    # warm-up: every system once
    srv.run(np.zeros(n_sys), np.arange(n_sys), np.arange(n_sys) + 10**9, pos, q, force)
    points = []
    seed_base = 0
    for rate in rates:
        n_req = max(50, int(rate * seconds))
        arrival = np.cumsum(stream_rng.exponential(1. / rate, n_req))
        ids = stream_rng.integers(0, n_sys, n_req)
        seeds = seed_base + np.arange(n_req)
        seed_base += n_req
        r = srv.run(arrival, ids, seeds, pos, q, force)
        lat = r["latency_s"] * 1e3
        points.append({"offered_req_s": rate, "requests": n_req,
                       "solves_per_s": round(n_req / (arrival[-1] + lat[-1] / 1e3), 1),
                       "p50_ms": round(float(np.percentile(lat, 50)), 3),
                       "p99_ms": round(float(np.percentile(lat, 99)), 3),
                       "mean_batch": round(n_req / r["batches"], 2),
                       "peak_lanes": r["peak_lanes"],
                       "peak_working_mib": round(r["peak_bytes"] / 2**20, 1),
                       "mparticles_per_s": round(float(counts[ids].sum()) / (
                           arrival[-1] + lat[-1] / 1e3) / 1e6, 1)})
        print(f"[bench] serving {rate} req/s: p50 {points[-1]['p50_ms']} ms, p99 "
              f"{points[-1]['p99_ms']} ms, batch {points[-1]['mean_batch']}", flush=True)
    # saturation: a burst of requests all present at t = 0
    n_b = 4 * n_sys
    ids = stream_rng.integers(0, n_sys, n_b)
    r = srv.run(np.zeros(n_b), ids, seed_base + np.arange(n_b), pos, q, force)
    sat = {"requests": n_b, "wall_s": round(r["wall_s"], 4),
           "solves_per_s": round(n_b / r["wall_s"], 1),
           "mparticles_per_s": round(float(counts[ids].sum()) / r["wall_s"] / 1e6, 1),
           "p50_ms": round(float(np.percentile(r["latency_s"], 50)) * 1e3, 3),
           "p99_ms": round(float(np.percentile(r["latency_s"], 99)) * 1e3, 3),
           "mean_batch": round(n_b / r["batches"], 2), "peak_lanes": r["peak_lanes"]}
    print(f"[bench] serving saturation: {sat['solves_per_s']} solves/s, "
          f"{sat['mparticles_per_s']} Mparticles/s", flush=True)
    meshes = [t["mesh"][0] for t in tuned]
    caos = [t["cao"] for t in tuned]
    return {"workload": f"{n_sys} resident systems, N log-uniform [1000, 200000], density "
                        "U[0.2, 0.8], +-1 alternating, tuned at 1e-4; Poisson arrivals, "
                        "uniform system choice, Gaussian displacement sigma=0.05 per request",
            "n_systems": n_sys, "resident_particles": n_tot,
            "resident_mesh_gib": round(resident / 2**30, 3),
            "mesh_range": [min(meshes), max(meshes)], "cao_range": [min(caos), max(caos)],
            "lanes": lanes, "working_budget_mib": budget_bytes >> 20, "max_batch": max_batch,
            "precision": "FP32 mesh, FP64 weights/reductions/forces",
            "sweep": points, "saturation": sat, "setup_s": round(t_setup, 1)}


GOLDEN = os.path.join(ROOT, "tests", "golden")


def total_force_error(p3m, tp, tq, idx, f_exact, torch):
    """Absolute rms error of the total P3M force (device k-space + device
    real-space pair forces) on the fixture subset against exact Ewald."""
    f = torch.zeros_like(tp)
    p3m.add_long_range_forces(tp, tq, f)
    f_r = p3m.pair_forces_on(tp, tq, idx)
    d = (f[torch.as_tensor(idx, device=tp.device)] + f_r).cpu().numpy() - f_exact
    return float(np.sqrt(np.mean(np.sum(d * d, axis=1))))


def accuracy_table_c3(torch, dev):
    """Precision sweep: N=250,000 +-1 in L=79.37, mesh {48,64,96} x cao
    {3,5,7} x accuracy {1e-3,1e-4,1e-5}; r_cut/alpha from the solver's tuning
    at fixed mesh and cao; absolute rms force error of the FP64 solver and the
    FP32-mesh form against exact Ewald on a 2000-particle subset."""
    from espresso_amd.electrostatics import P3M
    L = 79.37
    rng = np.random.default_rng(1003)
    pos = rng.uniform(0., L, (250_000, 3))
    q = np.where(np.arange(250_000) % 2 == 0, 1., -1.)
    fx = np.load(os.path.join(GOLDEN, "ewald_c3_subset.npz"))
    tp, tq = torch.tensor(pos, device=dev), torch.tensor(q, device=dev)
    rows = []
    for mesh in (48, 64, 96):
        for cao in (3, 5, 7):
            for acc in (1e-3, 1e-4, 1e-5):
                row = {"mesh": mesh, "cao": cao, "accuracy": acc}
                try:
                    p64 = P3M(prefactor=1.0, accuracy=acc, mesh=mesh, cao=cao, tune=True,
                              timings=1, verbose=False)
                    p64.activate(pos, q, box_l=[L] * 3)
                except RuntimeError as e:
                    row["infeasible"] = str(e)
                    rows.append(row)
                    continue
                prm = p64.get_params()
                p32 = P3M(prefactor=1.0, accuracy=acc, mesh=mesh, cao=cao, tune=False,
                          r_cut=prm["r_cut"], alpha=prm["alpha"], single_precision=True)
                p32.activate(pos, q, box_l=[L] * 3)
                row.update({"r_cut": round(prm["r_cut"], 4), "alpha": round(prm["alpha"], 5),
                            "estimate": prm["accuracy"],
                            "rms_err_fp64": total_force_error(p64, tp, tq, fx["idx"],
                                                              fx["force"], torch),
                            "rms_err_fp32": total_force_error(p32, tp, tq, fx["idx"],
                                                              fx["force"], torch)})
                print(f"[bench] C3 mesh {mesh} cao {cao} acc {acc:g}: est "
                      f"{row['estimate']:.2e} fp64 {row['rms_err_fp64']:.2e} fp32 "
                      f"{row['rms_err_fp32']:.2e}", flush=True)
                rows.append(row)
    return rows


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_c2(pos, q, box, mesh, prm, f_gpu):
    """The C oracle (restatement of CoulombP3MImpl<double, CPU>: charge
    assignment, c2c FFTs, ik convolution, force interpolation; in-repo FFT,
    no FFTW on this host) timed on this host: T = all CPUs of this process
    (threads over particles and FFT lines, in place of the reference's T MPI
    ranks; median of 5 solves after 1 warm-up) and 1 thread (median of 2)."""
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0))
    n = len(q)
    print(f"[bench] CPU baseline: FP64 oracle, {threads} threads and 1 thread", flush=True)
    ora = O.OracleP3M(box, mesh, 7, prm["alpha"], prm["r_cut"], 1.0)

    def timed(t, reps, warm):
        O.set_threads(t)
        try:
            f = None
            for _ in range(warm):
                f = ora.forces(pos, q)
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                f = ora.forces(pos, q)
                ts.append(time.perf_counter() - t0)
        finally:
            O.set_threads(1)
        return float(np.median(ts)), f

    t_all, f_cpu = timed(threads, 5, 1)
    t_one, _ = timed(1, 2, 0)
    parity = float(np.sqrt(np.sum((f_gpu - f_cpu) ** 2) / np.sum(f_cpu ** 2)))
    cpu = {"value": round(n / t_all / 1e6, 4), "unit": "Mparticles/s", "cores": threads,
           "kind": "port",
           "sample": f"FP64 k-space force solves of the full workload (N=1e6, mesh 128^3, "
                     f"cao 7) by the C oracle (restatement of CoulombP3MImpl<double,CPU>, "
                     f"in-repo FFT, no FFTW on the host) on {threads} threads (particles "
                     f"and FFT lines; the reference would run {threads} MPI ranks): median "
                     f"of 5 solves after 1 warm-up",
           "cpu_model": cpu_model(), "seconds": round(t_all, 4),
           "single_thread": {"value": round(n / t_one / 1e6, 4), "cores": 1,
                             "seconds": round(t_one, 3), "sample": "median of 2 solves"},
           "gpu_vs_oracle_force_rel_rms": parity}
    print(f"[bench] CPU: {t_all:.3f} s/solve on {threads} threads, {t_one:.2f} s on 1; "
          f"GPU-vs-oracle rel-rms {parity:.2e}", flush=True)
    return cpu


def timed_loop(p3m, tp, tq, tf, steps, warmup, torch):
    for _ in range(warmup):
        p3m.add_long_range_forces(tp, tq, tf, sync=False)
    p3m.synchronize()
    torch.cuda.synchronize()
    p3m.profile_begin(steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        p3m.add_long_range_forces(tp, tq, tf, sync=False)
    p3m.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nprof, stages = p3m.profile_end()
    return (t1 - t0) / steps, {k: v / max(nprof, 1) for k, v in stages.items()}


def bench_elc(steps, warmup, torch, dev, n=100_000, L=60., gap=10., seed=7001, cpu=True):
    """Slab workload: n charges (+-1 alternating) in an L x L x (L + gap) box
    with the top gap empty; P3M (FP64, timed tuning at 1e-4) with and
    without ELC (maxPWerror 1e-4).  The ELC share is the difference of the
    two force solves."""
    from espresso_amd.electrostatics import ELC, P3M
    rng = np.random.default_rng(seed)
    box = [L, L, L + gap]
    pos = rng.uniform(0., 1., (n, 3)) * np.array([L, L, L])
    q = np.where(np.arange(n) % 2 == 0, 1., -1.)
    tp = torch.tensor(pos, device=dev)
    tq = torch.tensor(q, device=dev)
    tf = torch.zeros_like(tp)
    plain = P3M(prefactor=1.0, accuracy=1e-4, timings=3, verbose=False)
    plain.activate(tp, tq, box_l=box)
    prm = plain.get_params()
    fixed = dict(prefactor=1.0, accuracy=1e-4, mesh=prm["mesh"], cao=prm["cao"],
                 r_cut=prm["r_cut"], alpha=prm["alpha"], tune=False)
    elc = ELC(actor=P3M(**fixed), maxPWerror=1e-4, gap_size=gap)
    elc.activate(tp, tq, box_l=box)
    t_p3m, _ = timed_loop(plain, tp, tq, tf, steps, warmup, torch)
    t_all, _ = timed_loop(elc.actor, tp, tq, tf, steps, warmup, torch)
    from espresso_amd import _lib
    import ctypes as C
    ep = _lib.p3m_elc_params()
    on = C.c_int(0)
    _lib.load().p3m_get_elc(elc.actor._ctx, C.byref(ep), C.byref(on))
    fc = ep.far_cut
    nx, ny = int(np.ceil(fc * L) + 1), int(np.ceil(fc * L) + 1)
    modes = 0
    for p_ in range(1, nx + 1):
        if (p_ - 1) / L >= fc:
            break
        modes += 1
    modes *= 2
    for p_ in range(1, nx + 1):
        if (p_ - 1) / L >= fc:
            break
        for q_ in range(1, ny + 1):
            if ((p_ - 1) / L) ** 2 + ((q_ - 1) / L) ** 2 >= fc * fc:
                break
            modes += 1
    t_elc = max(t_all - t_p3m, 1e-9)
    rec = {"workload": f"slab N={n}, box {L}x{L}x{L + gap} (gap {gap}), +-1 "
                       f"alternating, P3M FP64 accuracy 1e-4 (timed tuning), "
                       f"ELC maxPWerror 1e-4",
           "mesh": prm["mesh"], "cao": prm["cao"], "r_cut": round(prm["r_cut"], 4),
           "far_cut": round(fc, 5), "modes": modes,
           "p3m_ms": round(t_p3m * 1e3, 4), "p3m_elc_ms": round(t_all * 1e3, 4),
           "elc_ms": round(t_elc * 1e3, 4),
           "elc_particle_modes_per_s": round(n * modes / t_elc, 1)}
    if cpu:
        # CPU baseline: the oracle's numpy restatement of the ELC terms
        # (elc.cpp:830-940) on the same slab, one thread
        from threadpoolctl import threadpool_limits
        from oracle import elc as E
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            E.elc_forces_energy(pos, q, box, 1.0, gap, fc)
            t_cpu = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(n / t_cpu / 1e6, 4), "unit": "Mparticles/s",
                               "cores": 1, "kind": "port",
                               "sample": "ELC forces and energy of the full slab by the "
                                         "oracle's numpy restatement, single thread",
                               "seconds": round(t_cpu, 3),
                               "gpu_vs_cpu": round(t_cpu / t_elc, 1)}
    print(f"[bench] elc: {rec}", flush=True)
    return rec


def bench_dp3m(steps, warmup, torch, dev, n=100_000, seed=7101, cpu=True, cpu_n=20_000):
    """Dipolar P3M at the density of the reference's dipolar test
    (0.09 per unit volume, dipolar_mdlc_p3m_scafacos_p2nfft.py): n random
    unit dipoles in a cubic box, mesh 64, cao 5, accuracy 1e-4 (alpha and
    r_cut from the dipolar error estimates).  Times the k-space solve
    (forces and torques) and the real-space pair terms (cell lists)."""
    from espresso_amd.magnetostatics import DipolarP3M
    rng = np.random.default_rng(seed)
    L = (4. * np.pi * n / (3. * 0.09)) ** (1. / 3.)
    pos = rng.uniform(0., L, (n, 3))
    dip = rng.normal(size=(n, 3))
    dip /= np.linalg.norm(dip, axis=1)[:, None]
    tp, td = torch.tensor(pos, device=dev), torch.tensor(dip, device=dev)
    tf, tt = torch.zeros_like(tp), torch.zeros_like(tp)
    dp = DipolarP3M(prefactor=1.0, accuracy=1e-4, mesh=64, cao=5, verbose=False)
    dp.activate(tp, td, box_l=L)
    prm = dp.get_params()
    for _ in range(max(1, warmup)):
        dp.add_long_range_forces(tp, td, tf, tt, sync=False)
    dp.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dp.add_long_range_forces(tp, td, tf, tt, sync=False)
    dp.synchronize()
    t_k = (time.perf_counter() - t0) / steps
    dp.add_short_range_forces(tp, td, tf, tt, sync=False)
    dp.synchronize()
    t0 = time.perf_counter()
    reps = max(3, steps // 4)
    for _ in range(reps):
        dp.add_short_range_forces(tp, td, tf, tt, sync=False)
    dp.synchronize()
    t_sr = (time.perf_counter() - t0) / reps
    rec = {"workload": f"dipolar P3M: N={n} unit dipoles, density 0.09, box {L:.2f}, "
                       f"mesh 64, cao 5, accuracy 1e-4",
           "alpha": round(prm["alpha"], 5), "r_cut": round(prm["r_cut"], 4),
           "kspace_ms": round(t_k * 1e3, 4),
           "kspace_mparticles_per_s": round(n / t_k / 1e6, 2),
           "real_space_ms": round(t_sr * 1e3, 4),
           "total_mparticles_per_s": round(n / (t_k + t_sr) / 1e6, 2)}
    if cpu:
        # CPU baseline: the oracle's numpy restatement of the dipolar k-space
        # (dp3m.cpp:246-517) for the first cpu_n dipoles of the same box and
        # mesh (the particle work dominates), one thread
        from threadpoolctl import threadpool_limits
        from oracle import dp3m as D
        ora = D.OracleDP3M(L, 64, 5, prm["alpha"], prm["r_cut"], 1.0)
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            ora.solve(pos[:cpu_n], dip[:cpu_n])
            t_cpu = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(cpu_n / t_cpu / 1e6, 4),
                               "unit": "Mparticles/s", "cores": 1, "kind": "port",
                               "sample": f"k-space forces, torques and energy of the first "
                                         f"{cpu_n} dipoles (same box, mesh 64, cao 5) by the "
                                         f"oracle's numpy restatement, single thread",
                               "seconds": round(t_cpu, 3)}
    print(f"[bench] dp3m: {rec}", flush=True)
    return rec


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 512-system batched workload")
    ap.add_argument("--no-c3-batched", action="store_true",
                    help="skip the batched comparison on precision-sweep-shaped systems")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the 2048-system serving workload")
    ap.add_argument("--no-short-range", action="store_true",
                    help="skip the device real-space (short-range) timing")
    ap.add_argument("--no-elc", action="store_true",
                    help="skip the ELC slab workload")
    ap.add_argument("--no-dp3m", action="store_true",
                    help="skip the dipolar P3M workload")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the rms force error against exact Ewald forces")
    ap.add_argument("--no-accuracy-table", action="store_true",
                    help="skip the 27-cell precision sweep (N=250,000) against "
                         "exact Ewald forces")
    ap.add_argument("--only", choices=("batched", "c3batched", "serving", "elc", "dp3m"),
                    help="run only this workload and print its record "
                         "(profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device copy first (reference "
                         "dispatch for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "independent systems shard one per GPU and a single large system "
              "would need mesh halos and FFT transposes that this build does "
              "not implement", file=sys.stderr)
        sys.exit(2)

    import torch
    from espresso_amd.electrostatics import P3M

    if args.only:
        dev = torch.device("cuda:0")
        runs = {"batched": lambda: bench_batched(args.steps, args.warmup, torch, dev),
                "c3batched": lambda: bench_c3_batched(args.steps, args.warmup, torch, dev),
                "serving": lambda: bench_serving(torch, dev),
                "elc": lambda: bench_elc(args.steps, args.warmup, torch, dev,
                                         cpu=not args.no_cpu_baseline),
                "dp3m": lambda: bench_dp3m(args.steps, args.warmup, torch, dev,
                                           cpu=not args.no_cpu_baseline)}
        res = runs[args.only]()
        print(json.dumps({"workload": args.only, "record": res}), flush=True)
        return

    pos, q, L = make_system(args.n)
    box = [L, L, L]
    n = len(q)
    n_q = int(np.count_nonzero(q))
    mesh = [128, 128, 128]
    print(f"[bench] system N={n} N_q={n_q} L={L:.4f}; tuning r_cut/alpha",
          flush=True)

    p3m = P3M(prefactor=1.0, accuracy=1e-5, mesh=mesh, cao=7, tune=True,
              timings=2, verbose=False)
    p3m.activate(pos, q, box_l=box)
    prm = p3m.get_params()
    print(f"[bench] tuned: r_cut={prm['r_cut']:.4f} alpha={prm['alpha']:.5f} "
          f"accuracy={prm['accuracy']:.3e}", flush=True)

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(1 << 25, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    tp = torch.tensor(pos, device=dev)
    tq = torch.tensor(q, device=dev)
    tf = torch.zeros((n, 3), dtype=torch.float64, device=dev)

    shape64 = p3m.autotune(tp, tq, reps=5)
    print(f"[bench] FP64 launch shape: {shape64}", flush=True)
    t_step, stages = timed_loop(p3m, tp, tq, tf, args.steps, args.warmup, torch)
    value = n / t_step / 1e6
    print(f"[bench] FP64: {t_step * 1e3:.3f} ms/step, {value:.1f} Mparticles/s; "
          f"stages {json.dumps({k: round(v, 4) for k, v in stages.items()})}",
          flush=True)

    # the same particles stored in spatial cell order, as ESPResSo's cell
    # system stores them (cells of 8 mesh spacings, x-major): the gather's
    # per-particle force updates and the binning's reads then touch
    # neighbouring memory; reported beside value, never as value
    cell = np.floor(pos / (8. * L / mesh[0])).astype(np.int64) % (mesh[0] // 8)
    order = np.lexsort((cell[:, 2], cell[:, 1], cell[:, 0]))
    tp_c = torch.tensor(pos[order], device=dev)
    tq_c = torch.tensor(q[order], device=dev)
    tf_c = torch.zeros_like(tp_c)
    t_cell, st_cell = timed_loop(p3m, tp_c, tq_c, tf_c, args.steps, args.warmup, torch)
    cell_ordered = {"ms_per_step": round(t_cell * 1e3, 4),
                    "mparticles_per_s": round(n / t_cell / 1e6, 2),
                    "stages_ms": {k: round(v, 4) for k, v in st_cell.items()},
                    "note": "same system and launch shape, particles sorted by cells of "
                            "8 mesh spacings (ESPResSo stores particles cell by cell)"}
    print(f"[bench] FP64 cell-ordered particles: {t_cell * 1e3:.3f} ms/step; stages "
          f"{json.dumps(cell_ordered['stages_ms'])}", flush=True)
    del tp_c, tq_c, tf_c

    # host-pointer (PCIe-inclusive) rate for the record
    f_host = np.zeros((n, 3))
    t0 = time.perf_counter()
    p3m.add_long_range_forces(pos, q, f_host)
    t_host = time.perf_counter() - t0

    ab = algorithmic_bytes(n_q, mesh, 8)
    single = {k: stages[k] for k in ("spread", "convolve", "gather")}
    dom = max(single, key=single.get)
    t_dom = single[dom] * 1e-3
    achieved = ab[dom] / t_dom / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and args.n == 1_000_000:
        with open(TRAFFIC_FILE) as fh:
            traffic = json.load(fh)["bytes_per_dispatch"].get(stage_kernels(shape64)[dom])
    roofline = {"bound": "hbm", "kernel": stage_kernels(shape64)[dom],
                "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": ("stored counter figure: rocprofv3 FETCH_SIZE + WRITE_SIZE "
                                   "per dispatch, calibrated by a device copy "
                                   "(profiles/hbm_traffic_c2_n1e6_mesh128_cao7.json)"
                                   if traffic is not None else None),
                "algorithmic_bytes": ab[dom],
                "avg_launch_ms": round(single[dom], 5)}
    if dom == "gather":
        # the gather's binding resource is LDS, not HBM (SQ counters in
        # profiles/sq_counters_c2_n1e6_mesh128_cao7.md)
        lb = lds_bytes_gather(n_q, 7, 8)
        roofline["lds"] = {"achieved": round(lb / t_dom / 1e9, 1), "peak": LDS_PEAK_GBS,
                           "unit": "GB/s", "frac": round(lb / t_dom / 1e9 / LDS_PEAK_GBS, 4),
                           "bytes": lb,
                           "note": "algorithmic LDS reads: 3 fields x cao^3 stencil points "
                                   "x 8 bytes per charged particle; peak 256 CUs x 256 B/clk "
                                   "x 2.4 GHz (measured ceilings: ds_read_b128 stream 96 TB/s, "
                                   "ds_read_b64 stream 67.5 TB/s)"}
    solve_achieved = ab["solve"] / t_step / 1e9

    # FP32 mesh form on the same input
    fp32 = None
    if not args.no_fp32:
        p32 = P3M(prefactor=1.0, accuracy=1e-5, mesh=mesh, cao=7, tune=False,
                  r_cut=prm["r_cut"], alpha=prm["alpha"], single_precision=True)
        p32.activate(pos, q, box_l=box)
        tf32 = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        shape32 = p32.autotune(tp, tq, reps=5)
        print(f"[bench] FP32 launch shape: {shape32}", flush=True)
        t32, st32 = timed_loop(p32, tp, tq, tf32, args.steps, args.warmup, torch)
        f64 = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        f32 = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        p3m.add_long_range_forces(tp, tq, f64)
        p32.add_long_range_forces(tp, tq, f32)
        d = (f32 - f64).cpu().numpy()
        ref = f64.cpu().numpy()
        rms_abs = float(np.sqrt(np.mean(np.sum(d * d, axis=1)[q != 0])))
        fp32 = {"value": round(n / t32 / 1e6, 2), "ms_per_step": round(t32 * 1e3, 4),
                "launch_shape": shape32,
                "stages_ms": {k: round(v, 4) for k, v in st32.items()},
                "force_rel_rms_vs_fp64": float(np.sqrt(np.sum(d * d) / np.sum(ref * ref))),
                "force_rms_abs_vs_fp64": rms_abs,
                "solve_gbs_algorithmic": round(
                    algorithmic_bytes(n_q, mesh, 4)["solve"] / t32 / 1e9, 1)}
        print(f"[bench] FP32 mesh: {t32 * 1e3:.3f} ms/step, {fp32['value']} "
              f"Mparticles/s, rel-rms vs FP64 {fp32['force_rel_rms_vs_fp64']:.2e}",
              flush=True)

    # absolute rms force error of the measured solvers against exact Ewald
    # forces on the first 2000 charged particles (full workload only)
    accuracy = None
    if args.n == 1_000_000 and not args.no_accuracy:
        fx = np.load(os.path.join(GOLDEN, "ewald_c2_subset.npz"))
        accuracy = {"reference": "exact FP64 Ewald (tests/golden/ewald_c2_subset.npz), "
                                 "2000 charged particles, total force = device k-space + "
                                 "device real-space pair forces",
                    "estimate": prm["accuracy"],
                    "rms_err_fp64": total_force_error(p3m, tp, tq, fx["idx"], fx["force"],
                                                      torch)}
        if fp32 is not None:
            accuracy["rms_err_fp32"] = total_force_error(p32, tp, tq, fx["idx"], fx["force"],
                                                         torch)
        print(f"[bench] rms force error vs Ewald: {accuracy}", flush=True)
    table = None if (args.no_accuracy_table or args.no_accuracy) else accuracy_table_c3(torch, dev)

    # device real-space pair forces of the same system (cell lists), and the
    # total force (k-space + real space) per step
    short_range = None
    pair_ns = None
    if not args.no_short_range:
        fs = torch.zeros_like(tp)
        for _ in range(max(1, args.warmup)):
            p3m.add_short_range_forces(tp, tq, fs, sync=False)
        p3m.synchronize()
        reps = max(3, args.steps // 4)
        t0 = time.perf_counter()
        for _ in range(reps):
            p3m.add_short_range_forces(tp, tq, fs, sync=False)
        p3m.synchronize()
        t_sr = (time.perf_counter() - t0) / reps
        pair_ns = t_sr * 1e9 / model_pairs(n_q, L, prm["r_cut"])
        short_range = {
            "ms": round(t_sr * 1e3, 3),
            "pair_ns": round(pair_ns, 5),
            "cells": p3m.short_range_cells(),
            "r_cut": round(prm["r_cut"], 5),
            "total_force_ms_fp64": round((t_sr + t_step) * 1e3, 3),
            "total_force_mparticles_per_s_fp64": round(n / (t_sr + t_step) / 1e6, 2)}
        print(f"[bench] short range: {short_range}", flush=True)

    elc = None
    if not args.no_elc:
        elc = bench_elc(args.steps, args.warmup, torch, dev, cpu=not args.no_cpu_baseline)

    dp3m = None
    if not args.no_dp3m:
        dp3m = bench_dp3m(args.steps, args.warmup, torch, dev, cpu=not args.no_cpu_baseline)

    batched = None
    if not args.no_batched:
        batched = bench_batched(args.steps, args.warmup, torch, dev, pair_ns=pair_ns)
    c3_batched = None
    if not args.no_c3_batched and not args.no_batched:
        c3_batched = bench_c3_batched(args.steps, args.warmup, torch, dev, pair_ns=pair_ns)
    serving = None
    if not args.no_serving:
        serving = bench_serving(torch, dev)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_c2(pos, q, box, mesh, prm, tf.cpu().numpy() /
                              (args.steps + args.warmup))

    line = {
        "metric": "P3M Mparticles/s at <=1e-4 rms force err",
        "value": round(value, 2),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single system N=1e6, L=125.99, 90% +-1 10% "
                               "uncharged, mesh 128^3, cao 7, accuracy 1e-5 "
                               "(r_cut/alpha tuned), FP64 plain solver",
                   "n_particles": n, "n_charged": n_q, "mesh": mesh, "cao": 7,
                   "r_cut": round(prm["r_cut"], 5), "alpha": round(prm["alpha"], 6),
                   "estimated_rms_force_error": prm["accuracy"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages_ms": {k: round(v, 5) for k, v in stages.items()},
        "launch_shape": shape64,
        "solve_gbs_algorithmic": round(solve_achieved, 1),
        "host_pointer_ms": round(t_host * 1e3, 3),
        "cell_ordered_input": cell_ordered,
        "fp32_mesh": fp32,
        "accuracy": accuracy,
        "short_range": short_range,
        "elc": elc,
        "dp3m": dp3m,
        "batched": batched,
        "batched_c3_shape": c3_batched,
        "serving": serving,
    }
    if table is not None:
        line["accuracy_table_c3"] = table
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
