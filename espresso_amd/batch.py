"""Many independent P3M systems solved per launch.

The reference solves one system per process (one ``CoulombP3M`` actor per
``System``).  ``P3MBatch`` applies the same k-space algorithm to many systems
at once: each system has its own box, mesh, charge-assignment order and
splitting parameter, keeps its meshes and influence function resident in
device memory, and is solved together with any other selected systems in one
launch per pipeline stage.  Particle data of all systems is concatenated in
the order the systems were added and must live on the device (torch tensors)::

    b = P3MBatch(single_precision=True)
    ids = [b.add_system(box_l=[L, L, L], n=n, prefactor=1., mesh=[32] * 3,
                        cao=5, alpha=a, r_cut=rc) for ...]
    pos = torch.empty((b.total_particles, 3), dtype=torch.float64,
                      device="cuda")
    ...
    energies, virials = b.solve(pos, q, force, energy=True)

``tune_model`` picks (mesh, cao, r_cut, alpha) for one system with the
reference's tuning algorithm (src/core/electrostatics/p3m.cpp:594-791) using
a cost model instead of timed integration steps, so that hundreds of systems
can be tuned without device runs.
"""
import ctypes as C

import numpy as np

from . import _lib


def _device_ptr(t, name, numel, device):
    """Data pointer of a contiguous float64 device tensor of ``numel`` values
    on HIP device ``device`` (checked before anything is launched)."""
    from .electrostatics import device_output
    return device_output(t, name, numel, device).data_ptr()


def shape_dict(shape, best_ms=None, default_ms=None):
    """Launch shape (a p3m_launch_shape) as a dict with mapping names, plus
    the autotuner's mean solve times when given."""
    gnames = {v: k for k, v in P3MBatch.GATHER.items()}
    snames = {v: k for k, v in P3MBatch.SPREAD.items()}
    d = {"gather": gnames.get(shape.gather, shape.gather), "ncol": shape.ncol,
         "zlines": shape.zlines, "fused_lds": shape.fused_lds,
         "spread": snames.get(shape.spread, shape.spread), "threads": shape.threads,
         "chunk": shape.chunk, "column": shape.column, "seg": shape.seg,
         "col_lanes": shape.col_lanes, "group": shape.group,
         "gather_threads": shape.gather_threads}
    if best_ms is not None:
        d.update(best_ms=best_ms, default_ms=default_ms)
    return d


class P3MBatch:
    """Batched k-space solver (FP32 mesh by default, FP64 with
    ``single_precision=False``)."""

    def __init__(self, device=0, single_precision=True):
        L = _lib.load()
        self._b = C.c_void_p()
        prec = _lib.P3M_PREC_FP32 if single_precision else _lib.P3M_PREC_FP64
        _lib.raise_for(L.p3m_batch_create(int(device), prec, C.byref(self._b)))
        self.device = int(device)
        self._counts = []
        self.single_precision = bool(single_precision)

    def _check(self, status):
        _lib.raise_for(status, batch=self._b)

    def add_system(self, box_l, n, prefactor, mesh, cao, alpha, r_cut,
                   accuracy=1e-4, epsilon=0., mesh_off=(-1., -1., -1.),
                   skin=0.4):
        """Add a system with fixed parameters and ``n`` particles; returns
        its id.  Validation follows P3MParameters (common.hpp:94-150) and the
        post-tuning sanity checks (p3m.cpp:710-750)."""
        L = _lib.load()
        cp = _lib.p3m_params()
        L.p3m_params_default(C.byref(cp))
        if isinstance(mesh, (int, np.integer)):
            mesh = [int(mesh)] * 3
        for d in range(3):
            cp.mesh[d] = int(mesh[d])
            cp.box_l[d] = float(box_l[d])
            cp.mesh_off[d] = float(mesh_off[d])
        cp.cao = int(cao)
        cp.alpha = float(alpha)
        cp.r_cut = float(r_cut)
        cp.accuracy = float(accuracy)
        cp.epsilon = 0. if epsilon == "metallic" else float(epsilon)
        cp.prefactor = float(prefactor)
        cp.skin = float(skin)
        cp.tune = 0
        sid = C.c_int(-1)
        self._check(L.p3m_batch_add_system(self._b, C.byref(cp), int(n),
                                           C.byref(sid)))
        self._counts.append(int(n))
        return sid.value

    @property
    def num_systems(self):
        return len(self._counts)

    @property
    def total_particles(self):
        return int(sum(self._counts))

    def particle_count(self, sys_id):
        return self._counts[sys_id]

    def offset(self, sys_id):
        off = C.c_int64(0)
        self._check(_lib.load().p3m_batch_particle_offset(self._b, int(sys_id),
                                                          C.byref(off)))
        return off.value

    def system_bytes(self, sys_id):
        """Resident device bytes of one system (meshes, influence function)."""
        out = C.c_int64(0)
        self._check(_lib.load().p3m_batch_system_bytes(self._b, int(sys_id),
                                                       C.byref(out)))
        return out.value

    def working_bytes_of(self, *systems):
        """Device working memory of a solve of the given systems beyond
        their resident data (the particle scratch)."""
        nsel, ids = self._selection(systems)
        out = C.c_int64(0)
        self._check(_lib.load().p3m_batch_working_bytes(self._b, nsel, ids,
                                                        C.byref(out)))
        return out.value

    def _selection(self, systems):
        if systems is None:
            systems = range(self.num_systems)
        ids = [int(s) for s in systems]
        return len(ids), (C.c_int * max(len(ids), 1))(*ids)

    def _arrays(self, pos, q, force):
        n = self.total_particles
        p = _device_ptr(pos, "pos", 3 * n, self.device)
        qq = _device_ptr(q, "q", n, self.device)
        f = None if force is None else _device_ptr(force, "force", 3 * n, self.device)
        _lib.wait_for_torch(self.stream())
        return p, qq, f

    def solve(self, pos, q, force=None, systems=None, energy=False,
              pressure=False, sync=True):
        """Solve the selected systems (all by default).  ``force`` is
        accumulated in place (F -= k-space force).  Returns per-system
        energies [nsel] and virials [nsel, 3, 3] (before division by the
        volume), zeros for quantities not requested; ``sync=False`` only
        enqueues and returns None."""
        p, qq, f = self._arrays(pos, q, force)
        nsel, ids = self._selection(systems)
        flags = 0
        if force is not None:
            flags |= _lib.P3M_FORCES
        if energy:
            flags |= _lib.P3M_ENERGY
        if pressure:
            flags |= _lib.P3M_PRESSURE
        if not sync:
            flags |= _lib.P3M_ASYNC
            self._check(_lib.load().p3m_batch_solve(
                self._b, nsel, ids, p, qq, f, flags, None, None))
            return None
        e = (C.c_double * max(nsel, 1))()
        v = (C.c_double * max(9 * nsel, 9))()
        self._check(_lib.load().p3m_batch_solve(self._b, nsel, ids, p, qq, f,
                                                flags, e, v))
        return (np.array(e[:nsel]),
                np.array(v[:9 * nsel]).reshape(nsel, 3, 3))

    def graph_solve(self, pos, q, force, systems=None):
        """Force solve replayed from a captured hipGraph (enqueue only)."""
        p, qq, f = self._arrays(pos, q, force)
        nsel, ids = self._selection(systems)
        self._check(_lib.load().p3m_batch_graph_solve(self._b, nsel, ids, p,
                                                      qq, f))

    GATHER = {"layer": 0, "plane": 1, "plane16": 2, "plane8": 3}
    SPREAD = {"plane": 0, "plane16": 1, "plane8": 2}

    def set_launch_shape(self, gather="plane16", ncol=8, zlines=16,
                         fused_lds=96 * 1024, spread="plane", threads=512, chunk=64,
                         column=0, seg=32, col_lanes=8, group=1, gather_threads=0):
        """Gather mapping ("layer", "plane", "plane16", "plane8" or 0-3),
        FFT columns per block (split y/x passes), z-line pairs per block, the
        LDS bytes up to which a system's x-y planes are transformed by the
        fused y/x pass (0: never), the spread mapping ("plane", "plane16",
        "plane8" or 0-2), and the block size (256 or 512) and particles per
        staged chunk (1-64, at most threads / 3) of the spread and gather.
        ``column`` 8 or 16 switches spread and gather to the column march
        (W x W columns, ``seg`` x layers per block, ``col_lanes`` lanes per
        particle); 0 keeps 8^3 bricks.  ``group`` 4 lets one spread block
        assign 2 x 2 x 1 bricks into one shared tile (for sparse bricks;
        needs column 0 and spread "plane").  ``gather_threads`` 256 or 512
        gives the brick gather its own block size (0: ``threads``)."""
        g = self.GATHER[gather] if isinstance(gather, str) else int(gather)
        sp = self.SPREAD[spread] if isinstance(spread, str) else int(spread)
        shape = _lib.p3m_launch_shape(g, int(ncol), int(zlines), int(fused_lds), sp,
                                      int(threads), int(chunk), int(column), int(seg),
                                      int(col_lanes), int(group), int(gather_threads))
        self._check(_lib.load().p3m_batch_set_launch_shape(self._b, C.byref(shape)))

    def set_deterministic(self, on=True):
        """Bit-reproducible solves of an FP64 batch (fixed-point charge
        assignment with integer atomics; the default for FP64 batches).
        Column-march launch shapes need it off."""
        self._check(_lib.load().p3m_batch_set_deterministic(self._b, 1 if on else 0))

    def launch_shape(self):
        """The launch shape of later solves (dict with mapping names)."""
        shape = _lib.p3m_launch_shape()
        self._check(_lib.load().p3m_batch_get_launch_shape(self._b, C.byref(shape)))
        return shape_dict(shape)

    def autotune(self, pos, q, scratch_force, systems=None, reps=3):
        """Pick the fastest launch shape (fused pass, particle-mesh blocks,
        gather and spread mappings, FFT columns, z lines) for force solves of
        the selection by timing every candidate; forces accumulate into
        ``scratch_force``.  Returns the chosen shape and the mean solve times
        before and after."""
        p, qq, f = self._arrays(pos, q, scratch_force)
        nsel, ids = self._selection(systems)
        shape = _lib.p3m_launch_shape()
        best, default = C.c_double(), C.c_double()
        self._check(_lib.load().p3m_batch_autotune(
            self._b, nsel, ids, p, qq, f, int(reps), C.byref(shape), C.byref(best),
            C.byref(default)))
        return shape_dict(shape, best.value, default.value)

    def synchronize(self):
        self._check(_lib.load().p3m_batch_synchronize(self._b))

    def stream(self):
        s = C.c_void_p()
        self._check(_lib.load().p3m_batch_get_stream(self._b, C.byref(s)))
        return s.value

    def profile_begin(self, max_solves):
        self._check(_lib.load().p3m_batch_profile_begin(self._b,
                                                        int(max_solves)))

    def profile_end(self):
        """(number of profiled solves, summed stage times in ms)."""
        out = (C.c_double * 7)()
        n = C.c_int(0)
        self._check(_lib.load().p3m_batch_profile_end(self._b, out,
                                                      C.byref(n)))
        keys = ("total", "bin", "spread", "fft_fwd", "convolve", "fft_bwd",
                "gather")
        return n.value, dict(zip(keys, out[:]))

    def __del__(self):
        b = getattr(self, "_b", None)
        if b:
            try:
                _lib.load().p3m_batch_destroy(b)
            except Exception:  # pragma: no cover
                pass
            self._b = None


def tune_model(box_l, n_charged, sum_q2, prefactor, accuracy,
               mesh_range=(1, 1 << 20), cao_range=(1, 7), skin=0.4,
               epsilon=0., verbose=False, pair_ns=1.0, single_precision=False):
    """Tuned parameters of one system from the reference's tuning algorithm
    with a modelled cost per trial (device k-space stage rates of an FP64
    mesh, or an FP32 mesh if ``single_precision``, plus a real-space loop at
    ``pair_ns`` ns per pair).  Returns a dict with mesh,
    cao, alpha, r_cut and the estimated accuracy; raises RuntimeError
    ("CoulombP3M: failed to reach requested accuracy") like the reference."""
    L = _lib.load()
    cp = _lib.p3m_params()
    L.p3m_params_default(C.byref(cp))
    for d in range(3):
        cp.box_l[d] = float(box_l[d])
    cp.prefactor = float(prefactor)
    cp.accuracy = float(accuracy)
    cp.epsilon = float(epsilon)
    cp.skin = float(skin)
    cp.verbose = 1 if verbose else 0
    _lib.raise_for(L.p3m_tune_model(C.byref(cp), int(n_charged), float(sum_q2),
                                    int(mesh_range[0]), int(mesh_range[1]),
                                    int(cao_range[0]), int(cao_range[1]),
                                    float(pair_ns), 1 if single_precision else 0))
    return {"mesh": [cp.mesh[d] for d in range(3)], "cao": cp.cao,
            "alpha": cp.alpha, "r_cut": cp.r_cut, "accuracy": cp.accuracy}


class P3MServer:
    """Continuous-batching request loop over the resident systems of a
    ``P3MBatch``: requests are admitted into batches as lanes (device
    streams) and working memory free up; the particle scratch of all batches
    in flight comes from one device arena of ``budget_bytes``.  Each request
    displaces its system's positions by ``sigma`` N(0, 1) (reproducible from
    its seed), clears its forces and recomputes them."""

    def __init__(self, batch, lanes=4, budget_bytes=256 << 20, max_batch=64,
                 sigma=0.05):
        L = _lib.load()
        self._batch = batch  # keeps the batch alive
        self._s = C.c_void_p()
        _lib.raise_for(L.p3m_serve_create(batch._b, int(lanes), int(budget_bytes),
                                          int(max_batch), float(sigma),
                                          C.byref(self._s)), batch=batch._b)

    def run(self, arrival_s, sys_ids, seeds, pos, q, force, capture=None):
        """Serve the requests (sorted by arrival, seconds after the call) to
        completion.  ``capture`` (device float64 tensor of shape
        (sum of the requests' particle counts, 3)) receives each request's
        forces in request order.  Returns a dict with per-request latencies
        (s), wall time, number of batches, peak working memory and the most
        batches in flight at once."""
        b = self._batch
        arrival = np.ascontiguousarray(arrival_s, dtype=np.float64)
        ids = np.ascontiguousarray(sys_ids, dtype=np.int32)
        sd = np.ascontiguousarray(seeds, dtype=np.uint64)
        n = len(arrival)
        if len(ids) != n or len(sd) != n:
            raise ValueError("arrival_s, sys_ids and seeds must have the same length")
        p, qq, f = b._arrays(pos, q, force)
        out_ptr, off = None, None
        if capture is not None:
            counts = np.array([b.particle_count(int(s)) for s in ids], dtype=np.int64)
            off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(counts)[:-1]])
                                       .astype(np.int64))
            out_ptr = _device_ptr(capture, "capture", 3 * int(counts.sum()), b.device)
        lat = np.zeros(n)
        wall = C.c_double()
        nb, pl = C.c_int(), C.c_int()
        peak = C.c_int64()
        _lib.raise_for(_lib.load().p3m_serve_run(
            self._s, n, arrival.ctypes.data_as(C.POINTER(C.c_double)),
            ids.ctypes.data_as(C.POINTER(C.c_int)),
            sd.ctypes.data_as(C.POINTER(C.c_uint64)), p, qq, f, out_ptr,
            None if off is None else off.ctypes.data_as(C.POINTER(C.c_int64)),
            lat.ctypes.data_as(C.POINTER(C.c_double)), C.byref(wall),
            C.byref(nb), C.byref(peak), C.byref(pl)), server=self._s)
        return {"latency_s": lat, "wall_s": wall.value, "batches": nb.value,
                "peak_bytes": peak.value, "peak_lanes": pl.value}

    def __del__(self):
        s = getattr(self, "_s", None)
        if s:
            try:
                _lib.load().p3m_serve_destroy(s)
            except Exception:  # pragma: no cover
                pass
            self._s = None
