"""P3M Coulomb actors backed by the MI355X solver.

Mirrors ``espressomd.electrostatics.P3M`` / ``P3MGPU``
(src/python/espressomd/electrostatics.py:129-284): the same keyword
arguments, defaults, validation messages and read-only parameters
(src/script_interface/electrostatics/CoulombP3M.hpp:66-128).  The actor
computes the long-range (k-space) part only; particles are passed as arrays:

    p3m = P3M(prefactor=1., accuracy=1e-4, mesh=32, cao=5, r_cut=3.,
              alpha=1.07465, tune=False)
    p3m.activate(pos, q, box_l=[L, L, L])
    f = np.zeros_like(pos); p3m.add_long_range_forces(pos, q, f)
    e = p3m.long_range_energy(pos, q)

Arrays may be numpy arrays (host, copied per call) or torch tensors on a
HIP device (used in place, no PCIe traffic).
"""
import ctypes as C
import numbers

import numpy as np

from . import _lib

_KEYS = ("prefactor", "accuracy", "cao", "r_cut", "alpha", "mesh", "epsilon",
         "mesh_off", "check_neutrality", "charge_neutrality_tolerance",
         "check_complex_residuals", "tune", "timings", "verbose",
         "single_precision", "is_tuned")

# default of Actor::charge_neutrality_tolerance (electrostatics/actor.hpp:38)
NEUTRALITY_TOLERANCE = 2e-12


def neutrality_tolerance(params, given):
    """Charge-neutrality tolerance of an actor, -1 when unchecked, as
    Actor::set_charge_neutrality_tolerance applies the two parameters
    (script_interface/electrostatics/Actor.hpp:71-79, Actor.impl.hpp:44-80):
    an explicit tolerance first (None switches the check off), then
    check_neutrality restores the default or switches it off."""
    tol = NEUTRALITY_TOLERANCE
    if given:
        v = params["charge_neutrality_tolerance"]
        if v is None:
            tol = -1.
        else:
            if not (_is_float(v) or _is_int(v)):
                raise ValueError(
                    "Parameter 'charge_neutrality_tolerance' has to be a float")
            if float(v) < 0.:
                raise ValueError(
                    "Parameter 'charge_neutrality_tolerance' must be >= 0")
            tol = float(v)
    if params["check_neutrality"]:
        if tol == -1.:
            tol = NEUTRALITY_TOLERANCE
    else:
        tol = -1.
    return tol


def _is_int(x):
    return isinstance(x, (numbers.Integral, np.integer)) and not isinstance(x, bool)


def _is_float(x):
    return isinstance(x, (numbers.Real, np.floating)) and not isinstance(x, bool)


def check_device_tensor(t, name, numel, device, minimum=False):
    """Validate a device tensor handed to the solver before anything is
    launched: on HIP device ``device`` and holding ``numel`` elements (at
    least ``numel`` with ``minimum``).  Returns True when it is a contiguous
    float64 tensor that the kernels can use in place."""
    if not getattr(t, "is_cuda", False):
        raise TypeError(f"'{name}' must be a device tensor too")
    dev = t.device.index if t.device.index is not None else 0
    if dev != device:
        raise ValueError(f"'{name}' is on device {dev}, the solver on device {device}")
    size = int(t.numel())
    if (size < numel) if minimum else (size != numel):
        want = f"at least {numel}" if minimum else str(numel)
        raise ValueError(f"'{name}' has {size} elements, expected {want}")
    return str(t.dtype) == "torch.float64" and bool(t.is_contiguous())


def device_input(t, name, numel, device, asynchronous):
    """Float64 contiguous version of a device input tensor.  A converted copy
    is a temporary that the caching allocator may hand out again as soon as
    this call returns, so an asynchronous call (the solver stream still
    reading it) refuses inputs that need one."""
    if check_device_tensor(t, name, numel, device):
        return t
    if asynchronous:
        raise TypeError(f"'{name}' must be a contiguous float64 tensor for an "
                        f"asynchronous call")
    import torch
    return t.contiguous().to(torch.float64)


def device_output(t, name, numel, device, minimum=False):
    """Device output tensor, written in place: it must already be a
    contiguous float64 tensor of the right size on the solver's device."""
    if not check_device_tensor(t, name, numel, device, minimum):
        raise TypeError(f"'{name}' must be a contiguous float64 tensor")
    return t


class _Arrays:
    """Device or host view of pos/q/force for one call.  Device tensors are
    checked against the solver's device and the particle count before any
    launch; with ``asynchronous`` set they must already be contiguous
    float64 tensors."""

    def __init__(self, pos, q, force=None, pos_unfolded=None, device=0,
                 asynchronous=False):
        self.keep = []
        self.device = _is_device_tensor(pos)
        if self.device:
            if not getattr(q, "is_cuda", False):
                raise TypeError("'q' must be a device tensor too")
            self.n = int(q.numel())
            q = device_input(q, "q", self.n, device, asynchronous)
            pos = device_input(pos, "pos", 3 * self.n, device, asynchronous)
            self.keep += [pos, q]
            self.pos = pos.data_ptr()
            self.q = q.data_ptr()
            self.force = None
            if force is not None:
                self.force = device_output(force, "force", 3 * self.n, device).data_ptr()
            self.upos = None
            if pos_unfolded is not None:
                up = device_input(pos_unfolded, "pos_unfolded", 3 * self.n, device,
                                  asynchronous)
                self.keep.append(up)
                self.upos = up.data_ptr()
        else:
            pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64)
                                       .reshape(-1, 3))
            q = np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(-1))
            if pos.shape[0] != q.shape[0]:
                raise ValueError("pos and q must describe the same particles")
            self.keep += [pos, q]
            self.n = int(q.shape[0])
            self.pos = pos.ctypes.data
            self.q = q.ctypes.data
            self.force = None
            if force is not None:
                if not (isinstance(force, np.ndarray) and force.dtype == np.float64
                        and force.flags.c_contiguous and force.size == 3 * self.n):
                    raise TypeError(
                        "'force' must be a C-contiguous float64 array of shape (n, 3)")
                self.force = force.ctypes.data
            self.upos = None
            if pos_unfolded is not None:
                up = np.ascontiguousarray(np.asarray(pos_unfolded, np.float64)
                                          .reshape(-1, 3))
                self.keep.append(up)
                self.upos = up.ctypes.data

    @property
    def flags(self):
        return _lib.P3M_DEVICE_POINTERS if self.device else 0

    def order_after_torch(self, ctx):
        """Device tensors: the solver stream waits for torch's stream."""
        if self.device:
            s = C.c_void_p()
            _lib.raise_for(_lib.load().p3m_get_stream(ctx, C.byref(s)), ctx)
            _lib.wait_for_torch(s.value)


def _is_device_tensor(x):
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda


class _P3MBase:
    """Common parameter handling of the P3M actors
    (electrostatics.py:129-176, CoulombP3M.hpp:102-128)."""

    _name = "P3M"
    _gpu = False
    _force_convention = _lib.P3M_CONV_CPU

    def required_keys(self):
        return {"prefactor", "accuracy"}

    def default_params(self):
        return {"cao": -1,
                "r_cut": -1.,
                "alpha": -1.,
                "mesh": [-1, -1, -1],
                "epsilon": 0.,
                "mesh_off": [-1., -1., -1.],
                "prefactor": 0.,
                "check_neutrality": True,
                "check_complex_residuals": True,
                "tune": True,
                "timings": 10,
                "verbose": True}

    def valid_keys(self):
        return set(_KEYS)

    def __init__(self, device=0, **kwargs):
        L = _lib.load()
        unknown = set(kwargs) - self.valid_keys()
        if unknown:
            raise ValueError(
                f"Only the following keys can be given as keyword arguments: "
                f"{sorted(self.valid_keys())}, got {sorted(kwargs)} "
                f"(unknown {sorted(unknown)})")
        missing = self.required_keys() - set(kwargs)
        if missing:
            raise ValueError(
                f"At least the following keys have to be given as keyword "
                f"arguments: {sorted(self.required_keys())}, got "
                f"{sorted(kwargs)} (missing {sorted(missing)})")
        params = self.default_params()
        params.update(kwargs)
        self.validate_params(params)
        params["charge_neutrality_tolerance"] = neutrality_tolerance(
            params, "charge_neutrality_tolerance" in kwargs)
        params["is_tuned"] = bool(params.get("is_tuned", not params["tune"]))
        self._params = params
        self._ctx = C.c_void_p()
        _lib.raise_for(L.p3m_create(int(device), C.byref(self._ctx)))
        _lib.raise_for(L.p3m_set_force_convention(self._ctx, self._force_convention),
                       self._ctx)
        self._device = int(device)
        self._activated = False
        self._box_l = None
        self._build_cparams(box_l=(1., 1., 1.), skin=0.4)
        _lib.raise_for(L.p3m_set_params(self._ctx, C.byref(self._cp)),
                       self._ctx)

    def validate_params(self, params):
        """electrostatics.py:148-176 plus the script-interface GPU check
        (CoulombP3M.hpp:110-113)."""
        if _is_int(params["mesh"]):
            params["mesh"] = 3 * [params["mesh"]]
        mesh = params["mesh"]
        if not (hasattr(mesh, "__len__") and len(mesh) == 3
                and all(_is_int(m) for m in mesh)):
            raise ValueError(
                "Parameter 'mesh' has to be an integer or integer list of length 3")
        params["mesh"] = [int(m) for m in mesh]
        if any(m % 2 != 0 and m != -1 for m in params["mesh"]):
            raise ValueError(
                "P3M requires an even number of mesh points in all directions")
        if isinstance(params["epsilon"], str) and params["epsilon"] == "metallic":
            params["epsilon"] = 0.0
        if not _is_float(params["epsilon"]):
            raise ValueError(
                "Parameter 'epsilon' has to be a float or 'metallic'")
        mo = params["mesh_off"]
        if not (hasattr(mo, "__len__") and len(mo) == 3
                and all(_is_float(m) for m in mo)):
            raise ValueError(
                "Parameter 'mesh_off' has to be a (3,) array_like of values "
                "between 0.0 and 1.0")
        if not _is_int(params["timings"]):
            raise TypeError("Parameter 'timings' has to be an integer")
        if not isinstance(params["tune"], (bool, np.bool_)):
            raise TypeError("Parameter 'tune' has to be a boolean")
        if self._gpu and not params["single_precision"]:
            raise ValueError("P3M GPU only implemented in single-precision mode")

    def _build_cparams(self, box_l, skin):
        L = _lib.load()
        p = self._params
        cp = _lib.p3m_params()
        L.p3m_params_default(C.byref(cp))
        for d in range(3):
            cp.mesh[d] = int(p["mesh"][d])
            cp.mesh_off[d] = float(p["mesh_off"][d])
            cp.box_l[d] = float(box_l[d])
        cp.cao = int(p["cao"])
        cp.alpha = float(p["alpha"])
        cp.r_cut = float(p["r_cut"])
        cp.epsilon = float(p["epsilon"])
        cp.accuracy = float(p["accuracy"])
        cp.prefactor = float(p["prefactor"])
        cp.skin = float(skin)
        cp.precision = (_lib.P3M_PREC_FP32 if p["single_precision"]
                        else _lib.P3M_PREC_FP64)
        # a restored tuned actor keeps its parameters and is not tuned again:
        # tuning = !is_tuned (CoulombP3M.hpp:114)
        cp.tune = 0 if p["is_tuned"] else 1
        cp.timings = int(p["timings"])
        cp.verbose = 1 if p["verbose"] else 0
        cp.charge_neutrality_tolerance = (p["charge_neutrality_tolerance"]
                                          if p["check_neutrality"] else -1.)
        self._cp = cp

    # -- activation ---------------------------------------------------------
    def activate(self, pos, q, box_l, skin=0.4):
        """Sanity checks, tuning and init on the given particles
        (CoulombP3MImpl::on_activation, p3m.impl.hpp:124-139)."""
        L = _lib.load()
        if not self._activated or tuple(box_l) != self._box_l:
            self._build_cparams(box_l=box_l, skin=skin)
            _lib.raise_for(L.p3m_set_params(self._ctx, C.byref(self._cp)),
                           self._ctx)
        a = _Arrays(pos, q, device=self._device)
        a.order_after_torch(self._ctx)
        _lib.raise_for(L.p3m_activate(self._ctx, a.n, a.pos, a.q, a.flags),
                       self._ctx)
        self._activated = True
        self._box_l = tuple(float(x) for x in box_l)
        return self

    def _require_active(self):
        if not self._activated:
            raise RuntimeError("P3M actor is not activated")

    # -- CoulombP3M long-range interface ------------------------------------
    def add_long_range_forces(self, pos, q, force, pos_unfolded=None,
                              sync=True):
        """force -= k-space force (p3m.impl.hpp:145-154); in place.

        With device tensors and ``sync=False`` the solve is only enqueued on
        the solver's stream (call ``synchronize()`` before reading)."""
        self._require_active()
        a = _Arrays(pos, q, force, pos_unfolded, device=self._device,
                    asynchronous=not sync)
        a.order_after_torch(self._ctx)
        flags = _lib.P3M_FORCES | a.flags
        if not sync:
            if not a.device:
                raise ValueError("sync=False needs device tensors")
            flags |= _lib.P3M_ASYNC
        _lib.raise_for(_lib.load().p3m_solve(
            self._ctx, a.n, a.pos, a.q, a.upos, a.force, flags, None, None),
            self._ctx)
        return force

    def set_tuning_real_space(self, mode):
        """Real-space cost of timed tuning: "model" (a host pair loop at
        1 ns per pair, the default) or "device" (``add_short_range_forces``
        timed for each trial on the activation particles).  Call before
        ``activate``."""
        modes = {"model": _lib.P3M_TUNE_RS_MODEL, "device": _lib.P3M_TUNE_RS_DEVICE}
        if mode not in modes:
            raise ValueError(f"tuning real space must be one of {sorted(modes)}")
        _lib.raise_for(_lib.load().p3m_set_tuning_real_space(self._ctx, modes[mode]),
                       self._ctx)

    def force_convention(self):
        """"cpu" (the CPU kernels' mesh) or "gpu" (the reference's GPU
        kernels' index rule) for the forces."""
        c = C.c_int(0)
        _lib.raise_for(_lib.load().p3m_get_force_convention(self._ctx, C.byref(c)),
                       self._ctx)
        return "gpu" if c.value == _lib.P3M_CONV_GPU else "cpu"

    def set_deterministic(self, on=True):
        """Bit-reproducible solves (FP64 actors, the default there): the
        charge assignment sums 64-bit fixed-point contributions with integer
        atomics, so repeated solves of the same input give identical forces
        and energies.  Column-march launch shapes need it off."""
        _lib.raise_for(_lib.load().p3m_set_deterministic(self._ctx, 1 if on else 0),
                       self._ctx)

    def set_launch_shape(self, **shape):
        """Launch shape of later solves: the keyword arguments of
        ``P3MBatch.set_launch_shape`` (gather, ncol, zlines, fused_lds,
        spread, threads, chunk, column, seg, col_lanes, group, gather_threads);
        unnamed fields keep their current values."""
        from .batch import P3MBatch
        cur = _lib.p3m_launch_shape()
        _lib.raise_for(_lib.load().p3m_get_launch_shape(self._ctx, C.byref(cur)), self._ctx)
        for name, _ in _lib.p3m_launch_shape._fields_:
            if name not in shape:
                continue
            v = shape.pop(name)
            if name == "gather" and isinstance(v, str):
                v = P3MBatch.GATHER[v]
            elif name == "spread" and isinstance(v, str):
                v = P3MBatch.SPREAD[v]
            setattr(cur, name, int(v))
        if shape:
            raise TypeError(f"unknown launch-shape fields: {sorted(shape)}")
        _lib.raise_for(_lib.load().p3m_set_launch_shape(self._ctx, C.byref(cur)), self._ctx)

    def launch_shape(self):
        """The launch shape of later solves (dict with mapping names)."""
        from .batch import shape_dict
        cur = _lib.p3m_launch_shape()
        _lib.raise_for(_lib.load().p3m_get_launch_shape(self._ctx, C.byref(cur)), self._ctx)
        return shape_dict(cur)

    def autotune(self, pos, q, reps=3):
        """Pick the fastest launch shape of the device kernels (fused y/x
        pass, particle-mesh blocks, gather and spread mappings, FFT columns,
        z lines) for force solves of
        these device tensors and keep it.  Returns the shape and the mean
        solve times before and after."""
        self._require_active()
        a = _Arrays(pos, q, device=self._device)
        if not a.device:
            raise TypeError("autotune takes device tensors")
        a.order_after_torch(self._ctx)
        from .batch import shape_dict
        shape = _lib.p3m_launch_shape()
        best, default = C.c_double(), C.c_double()
        _lib.raise_for(_lib.load().p3m_autotune(
            self._ctx, a.n, a.pos, a.q, int(reps), C.byref(shape), C.byref(best),
            C.byref(default)), self._ctx)
        return shape_dict(shape, best.value, default.value)

    def synchronize(self):
        _lib.raise_for(_lib.load().p3m_synchronize(self._ctx), self._ctx)

    def profile_begin(self, max_solves):
        _lib.raise_for(_lib.load().p3m_profile_begin(self._ctx, int(max_solves)),
                       self._ctx)

    def profile_end(self):
        """(number of profiled solves, summed stage times in ms)."""
        out = (C.c_double * 7)()
        n = C.c_int(0)
        _lib.raise_for(_lib.load().p3m_profile_end(self._ctx, out, C.byref(n)),
                       self._ctx)
        keys = ("total", "bin", "spread", "fft_fwd", "convolve", "fft_bwd",
                "gather")
        return n.value, dict(zip(keys, out[:]))

    def long_range_energy(self, pos, q, pos_unfolded=None):
        """k-space energy incl. self/net-charge/dipole terms (p3m.cpp:551-589)."""
        self._require_active()
        a = _Arrays(pos, q, None, pos_unfolded, device=self._device)
        a.order_after_torch(self._ctx)
        e = C.c_double(0.)
        _lib.raise_for(_lib.load().p3m_solve(
            self._ctx, a.n, a.pos, a.q, a.upos, None,
            _lib.P3M_ENERGY | a.flags, C.byref(e), None), self._ctx)
        return e.value

    def long_range_pressure(self, pos, q):
        """k-space virial tensor as returned by long_range_pressure
        (p3m.cpp:393-449), i.e. before the division by the volume."""
        self._require_active()
        a = _Arrays(pos, q, device=self._device)
        a.order_after_torch(self._ctx)
        v = (C.c_double * 9)()
        _lib.raise_for(_lib.load().p3m_solve(
            self._ctx, a.n, a.pos, a.q, None, None,
            _lib.P3M_PRESSURE | a.flags, None, v), self._ctx)
        return np.array(v[:]).reshape(3, 3)

    def solve(self, pos, q, force=None, energy=True, pressure=False,
              pos_unfolded=None):
        """Forces, energy and virial from one charge assignment and FFT."""
        self._require_active()
        a = _Arrays(pos, q, force, pos_unfolded, device=self._device)
        a.order_after_torch(self._ctx)
        flags = a.flags
        if force is not None:
            flags |= _lib.P3M_FORCES
        if energy:
            flags |= _lib.P3M_ENERGY
        if pressure:
            flags |= _lib.P3M_PRESSURE
        e = C.c_double(0.)
        v = (C.c_double * 9)()
        _lib.raise_for(_lib.load().p3m_solve(
            self._ctx, a.n, a.pos, a.q, a.upos, a.force, flags, C.byref(e), v),
            self._ctx)
        return e.value, np.array(v[:]).reshape(3, 3)

    def solve_to_device(self, pos, q, force=None, energy=None, virial=None, sync=False):
        """One solve on device tensors whose energy (1-element float64
        tensor) and virial (9 elements) are written on the device in stream
        order, without a host round trip (P3M_DEVICE_RESULTS): an MD loop's
        per-step energy or NpT virial (p3m.cpp:580-583).  Without sync the
        call only enqueues; ``synchronize()`` waits."""
        self._require_active()
        if not _is_device_tensor(pos):
            raise TypeError("solve_to_device takes device tensors")
        a = _Arrays(pos, q, force, device=self._device, asynchronous=not sync)
        if energy is not None:
            device_output(energy, "energy", 1, self._device, minimum=True)
        if virial is not None:
            device_output(virial, "virial", 9, self._device, minimum=True)
        a.order_after_torch(self._ctx)
        flags = a.flags | _lib.P3M_DEVICE_RESULTS
        if force is not None:
            flags |= _lib.P3M_FORCES
        if energy is not None:
            flags |= _lib.P3M_ENERGY
        if virial is not None:
            flags |= _lib.P3M_PRESSURE
        if not sync:
            flags |= _lib.P3M_ASYNC
        ep = C.c_void_p(energy.data_ptr()) if energy is not None else None
        vp = C.c_void_p(virial.data_ptr()) if virial is not None else None
        _lib.raise_for(_lib.load().p3m_solve(
            self._ctx, a.n, a.pos, a.q, a.upos, a.force, flags,
            C.cast(ep, C.POINTER(C.c_double)) if ep else None,
            C.cast(vp, C.POINTER(C.c_double)) if vp else None), self._ctx)

    def add_short_range_forces(self, pos, q, force, energy=False, pressure=False,
                               sync=True):
        """force += real-space pair forces of all charged pairs closer than
        r_cut (CoulombP3M::pair_force, p3m.hpp:146-165, over cell lists as the
        short-range loop of forces.cpp:178-208 applies it); in place.  Returns
        the pair energy with ``energy=True`` and the virial sum F (x) d with
        ``pressure=True`` (a (energy, virial) pair with both).  With device
        tensors and ``sync=False`` the work is only enqueued."""
        self._require_active()
        a = _Arrays(pos, q, force, device=self._device, asynchronous=not sync)
        a.order_after_torch(self._ctx)
        flags = a.flags
        if energy:
            flags |= _lib.P3M_ENERGY
        if pressure:
            flags |= _lib.P3M_PRESSURE
        if not sync:
            if not a.device:
                raise ValueError("sync=False needs device tensors")
            flags |= _lib.P3M_ASYNC
        e = C.c_double(0.)
        v = (C.c_double * 9)()
        _lib.raise_for(_lib.load().p3m_short_range(
            self._ctx, a.n, a.pos, a.q, flags, a.force, C.byref(e), v), self._ctx)
        vir = np.array(v[:]).reshape(3, 3)
        if energy and pressure:
            return e.value, vir
        if energy:
            return e.value
        if pressure:
            return vir
        return force

    def set_short_range_subdivision(self, cells_per_cutoff):
        """Cells per r_cut of the short-range cell lists (0: automatic)."""
        _lib.raise_for(_lib.load().p3m_set_short_range_subdivision(
            self._ctx, int(cells_per_cutoff)), self._ctx)

    def short_range_cells(self):
        """Cells per dimension of the last short-range call."""
        c = (C.c_int * 3)()
        _lib.raise_for(_lib.load().p3m_short_range_cells(self._ctx, c), self._ctx)
        return list(c)

    def pair_forces_on(self, pos, q, targets):
        """Real-space pair forces (A&S erfc, p3m.hpp:146-165) on the listed
        target particles from all particles within r_cut, minimum image;
        device tensors in, a (len(targets), 3) device tensor out."""
        import torch
        self._require_active()
        a = _Arrays(pos, q, device=self._device)
        if not a.device:
            raise TypeError("pair_forces_on takes device tensors")
        a.order_after_torch(self._ctx)
        t = torch.as_tensor(targets, dtype=torch.int32).to(pos.device).contiguous()
        out = torch.empty((t.numel(), 3), dtype=torch.float64, device=pos.device)
        torch.cuda.synchronize()
        _lib.raise_for(_lib.load().p3m_pair_forces_targets(
            self._ctx, a.n, a.pos, a.q, t.numel(), t.data_ptr(), out.data_ptr()),
            self._ctx)
        return out

    def pair_force(self, q1q2, d, dist):
        out = (C.c_double * 3)()
        dd = (C.c_double * 3)(*[float(x) for x in d])
        _lib.raise_for(_lib.load().p3m_pair_force(self._ctx, q1q2, dd, dist, out),
                       self._ctx)
        return np.array(out[:])

    def pair_energy(self, q1q2, dist):
        out = C.c_double(0.)
        _lib.raise_for(_lib.load().p3m_pair_energy(self._ctx, q1q2, dist,
                                                   C.byref(out)), self._ctx)
        return out.value

    # -- stage outputs for parity tests -------------------------------------
    def charge_mesh(self, pos, q):
        self._require_active()
        a = _Arrays(pos, q, device=self._device)
        a.order_after_torch(self._ctx)
        if a.device:
            raise TypeError("charge_mesh takes host arrays")
        m = self.get_params()["mesh"]
        out = np.zeros(int(np.prod(m)))
        _lib.raise_for(_lib.load().p3m_debug_charge_mesh(
            self._ctx, a.n, a.pos, a.q, out.ctypes.data), self._ctx)
        return out.reshape(tuple(m))

    def mesh_indices(self, pos, q):
        self._require_active()
        a = _Arrays(pos, q, device=self._device)
        a.order_after_torch(self._ctx)
        out = np.zeros((a.n, 3), np.int32)
        _lib.raise_for(_lib.load().p3m_debug_indices(
            self._ctx, a.n, a.pos, a.q, out.ctypes.data), self._ctx)
        return out

    def last_timings(self):
        out = (C.c_double * 8)()
        _lib.load().p3m_last_timings(self._ctx, out, 8)
        keys = ("total", "bin", "spread", "fft_fwd", "convolve", "fft_bwd",
                "gather")
        return dict(zip(keys, out[:7]))

    # -- read-only parameters (CoulombP3M.hpp:66-100) -----------------------
    def get_params(self):
        cp = _lib.p3m_params()
        _lib.raise_for(_lib.load().p3m_get_params(self._ctx, C.byref(cp)),
                       self._ctx)
        tuned = C.c_int(0)
        _lib.load().p3m_is_tuned(self._ctx, C.byref(tuned))
        box = self._box_l or tuple(cp.box_l)
        mesh = [cp.mesh[d] for d in range(3)]
        a = [box[d] / mesh[d] if mesh[d] > 0 else 0. for d in range(3)]
        return {
            "prefactor": cp.prefactor,
            "accuracy": cp.accuracy,
            "mesh": mesh,
            "mesh_off": [cp.mesh_off[d] for d in range(3)],
            "cao": cp.cao,
            "alpha": cp.alpha,
            "r_cut": cp.r_cut,
            "alpha_L": cp.alpha * box[0],
            "r_cut_iL": cp.r_cut / box[0],
            "a": a,
            "epsilon": cp.epsilon,
            "is_tuned": bool(tuned.value),
            "single_precision": bool(self._params["single_precision"]),
            "timings": self._params["timings"],
            "tune": self._params["tune"],
            "verbose": self._params["verbose"],
            "check_neutrality": self._params["check_neutrality"],
            "charge_neutrality_tolerance": (
                self._params["charge_neutrality_tolerance"]
                if self._params["check_neutrality"] else None),
            "check_complex_residuals": self._params["check_complex_residuals"],
        }

    # -- checkpointing (test_checkpoint.py:846-866) -------------------------
    _CHECKPOINT_KEYS = ("prefactor", "accuracy", "mesh", "mesh_off", "cao", "alpha",
                        "r_cut", "epsilon", "tune", "timings", "verbose",
                        "check_neutrality", "charge_neutrality_tolerance",
                        "check_complex_residuals", "single_precision", "is_tuned")

    def __reduce__(self):
        """Pickling keeps the constructor parameters and the tuned ones: a
        restored actor built from them with ``is_tuned`` set is not tuned
        again on activation (CoulombP3M.hpp:102-128 constructs it with
        tuning = !is_tuned).  Particles are not part of the state; the
        restored actor is activated on the caller's particles."""
        p = self.get_params()
        if not p["is_tuned"]:  # still to be tuned: the parameters as given
            p = dict(self._params, is_tuned=False)
        kw = {k: p[k] for k in self._CHECKPOINT_KEYS}
        return (_restore_actor, (type(self), self._device, kw))

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        params = self.get_params()
        if name in params:
            return params[name]
        raise AttributeError(name)

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            try:
                _lib.load().p3m_destroy(ctx)
            except Exception:  # pragma: no cover
                pass
            self._ctx = None


def _restore_actor(cls, device, kwargs):
    return cls(device=device, **kwargs)


def _restore_elc(kwargs):
    return ELC(**kwargs)


class P3M(_P3MBase):
    """P3M electrostatics solver (FP64 by default, FP32 mesh with
    ``single_precision=True``)."""

    def default_params(self):
        return {"single_precision": False, **super().default_params()}


class P3MGPU(_P3MBase):
    """P3M electrostatics solver with GPU support (FP32 mesh, FP64
    reductions and force accumulation).  Forces follow the reference's GPU
    kernels (FP32 mesh index without mesh offset, Nyquist wave number kept,
    no dipole force term; p3m_gpu_cuda.cu:265-360, p3m.cpp:867-888); energy
    and pressure the CPU kernels' mesh, as the reference's GPU actor
    computes them (p3m.cpp:889-892)."""

    _gpu = True
    _force_convention = _lib.P3M_CONV_GPU

    def default_params(self):
        return {"single_precision": True, **super().default_params()}


class ELC:
    """Electrostatic layer correction on top of a P3M or P3MGPU actor, for
    systems with two periodic dimensions (``espressomd.electrostatics.ELC``,
    electrostatics.py class ELC; core ElectrostaticLayerCorrection,
    elc.cpp).  Charges stay in 0 <= z <= L_z - gap_size.  With dielectric
    contrasts (delta_mid_top / delta_mid_bot) the mesh carries the image
    charges of the boundary layers, and ``add_short_range_forces`` adds the
    pair terms with the images.

    The forces, energy and virial calls of the wrapped actor include the
    ELC terms once ELC is constructed on it."""

    _KEYS = ("actor", "maxPWerror", "gap_size", "far_cut", "delta_mid_top",
             "delta_mid_bot", "const_pot", "pot_diff", "neutralize",
             "check_neutrality", "charge_neutrality_tolerance")

    def required_keys(self):
        return {"actor", "maxPWerror", "gap_size"}

    def default_params(self):
        return {"far_cut": -1.,
                "delta_mid_top": 0.,
                "delta_mid_bot": 0.,
                "const_pot": False,
                "pot_diff": 0.,
                "neutralize": True,
                "check_neutrality": True}

    def __init__(self, **kwargs):
        unknown = set(kwargs) - set(self._KEYS)
        if unknown:
            raise ValueError(
                f"Only the following keys can be given as keyword arguments: "
                f"{sorted(self._KEYS)}, got {sorted(kwargs)} (unknown {sorted(unknown)})")
        missing = self.required_keys() - set(kwargs)
        if missing:
            raise ValueError(
                f"At least the following keys have to be given as keyword "
                f"arguments: {sorted(self.required_keys())}, got {sorted(kwargs)} "
                f"(missing {sorted(missing)})")
        params = self.default_params()
        params.update(kwargs)
        self.validate_params(params)
        params["charge_neutrality_tolerance"] = neutrality_tolerance(
            params, "charge_neutrality_tolerance" in kwargs)
        actor = params["actor"]
        if not isinstance(actor, _P3MBase):
            raise ValueError(f"Parameter 'actor' of type {type(actor).__name__} "
                             f"isn't supported by ELC")
        self._params = params
        self.actor = actor
        L = _lib.load()
        ep = _lib.p3m_elc_params()
        L.p3m_elc_params_default(C.byref(ep))
        ep.maxPWerror = float(params["maxPWerror"])
        ep.gap_size = float(params["gap_size"])
        ep.far_cut = float(params["far_cut"])
        ep.neutralize = 1 if params["neutralize"] else 0
        ep.delta_mid_top = float(params["delta_mid_top"])
        ep.delta_mid_bot = float(params["delta_mid_bot"])
        ep.const_pot = 1 if params["const_pot"] else 0
        ep.pot_diff = float(params["pot_diff"])
        _lib.raise_for(L.p3m_set_elc(actor._ctx, C.byref(ep)), actor._ctx)
        # the ELC actor's neutrality check governs the solve it wraps
        if (not params["check_neutrality"] or "charge_neutrality_tolerance" in kwargs):
            actor._params["check_neutrality"] = bool(params["check_neutrality"])
            actor._params["charge_neutrality_tolerance"] = params["charge_neutrality_tolerance"]
            if actor._activated:
                actor._build_cparams(box_l=actor._box_l, skin=actor._cp.skin)
                _lib.raise_for(L.p3m_set_params(actor._ctx, C.byref(actor._cp)),
                               actor._ctx)

    @staticmethod
    def validate_params(params):
        """electrostatics.py ELC.validate_params (type checks with the
        reference's messages)."""
        def is_float(v):
            return _is_float(v) or _is_int(v)
        if not is_float(params["maxPWerror"]):
            raise ValueError("maxPWerror has to be a float -- Got an "
                             f"{type(params['maxPWerror']).__name__}")
        if not is_float(params["gap_size"]):
            raise ValueError("gap_size has to be a float -- Got an "
                             f"{type(params['gap_size']).__name__}")
        if not is_float(params["far_cut"]):
            raise ValueError("far_cut has to be a float -- Got an "
                             f"{type(params['far_cut']).__name__}")
        if not isinstance(params["neutralize"], (bool, np.bool_)):
            raise ValueError("neutralize has to be a bool -- Got an "
                             f"{type(params['neutralize']).__name__}")

    def get_params(self):
        """ELC parameters as the script interface reports them, far_cut as
        tuned by the last activation or solve."""
        ep = _lib.p3m_elc_params()
        on = C.c_int(0)
        _lib.raise_for(_lib.load().p3m_get_elc(self.actor._ctx, C.byref(ep), C.byref(on)),
                       self.actor._ctx)
        return {"actor": self.actor, "maxPWerror": ep.maxPWerror, "gap_size": ep.gap_size,
                "far_cut": ep.far_cut, "neutralize": bool(ep.neutralize),
                "delta_mid_top": ep.delta_mid_top, "delta_mid_bot": ep.delta_mid_bot,
                "const_pot": bool(ep.const_pot), "pot_diff": ep.pot_diff,
                "check_neutrality": bool(self._params["check_neutrality"]),
                "charge_neutrality_tolerance": (
                    self._params["charge_neutrality_tolerance"]
                    if self._params["check_neutrality"] else None)}

    def __reduce__(self):
        """Pickling (test_checkpoint.py:868-888): the ELC parameters as given
        (far_cut -1 is tuned again on activation) and the wrapped actor with
        its own checkpoint state."""
        kw = {k: self._params[k] for k in self._KEYS if k in self._params}
        return (_restore_elc, (kw,))

    def activate(self, pos, q, box_l, skin=0.4):
        """ElectrostaticLayerCorrection::on_activation (elc.hpp:181-200):
        gap and far_cut checks, then the P3M actor's activation."""
        self.actor.activate(pos, q, box_l=box_l, skin=skin)
        return self

    def add_long_range_forces(self, pos, q, force, pos_unfolded=None, sync=True):
        return self.actor.add_long_range_forces(pos, q, force, pos_unfolded, sync)

    def long_range_energy(self, pos, q, pos_unfolded=None):
        return self.actor.long_range_energy(pos, q, pos_unfolded)

    def solve(self, pos, q, force=None, energy=True, pressure=False, pos_unfolded=None):
        return self.actor.solve(pos, q, force, energy, pressure, pos_unfolded)

    def add_short_range_forces(self, *args, **kwargs):
        return self.actor.add_short_range_forces(*args, **kwargs)

    def remove(self):
        """Take ELC off the actor (its epsilon stays metallic)."""
        _lib.raise_for(_lib.load().p3m_set_elc(self.actor._ctx, None), self.actor._ctx)
