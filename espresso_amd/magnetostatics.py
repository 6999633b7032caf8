"""Dipolar P3M actor on the MI355X solver.

Mirrors ``espressomd.magnetostatics.DipolarP3M`` (magnetostatics.py, class
DipolarP3M): the same keyword arguments, defaults and validation messages;
the core is DipolarP3MImpl (src/core/magnetostatics/dp3m.cpp).  Particles
are passed explicitly: positions and dipole moments [n][3], as numpy arrays
(host) or float64 device tensors.
"""
import ctypes as C

import numpy as np

from . import _lib
from .electrostatics import (_is_device_tensor, _is_float, _is_int, device_input,
                              device_output)

_KEYS = ("prefactor", "accuracy", "cao", "r_cut", "alpha", "mesh", "epsilon",
         "mesh_off", "single_precision", "tune", "timings", "verbose", "is_tuned")


class _Vec:
    """Host or device view of [n][3] arrays for one call.  Device tensors are
    checked against the solver's device and the particle count before any
    launch; with ``asynchronous`` set they must already be contiguous float64
    tensors (a converted copy could be reused while the stream reads it)."""

    def __init__(self, pos, dip, force=None, torque=None, device=0, asynchronous=False):
        self.keep = []
        self.device = _is_device_tensor(pos)
        outs = []
        if self.device:
            if int(pos.numel()) % 3:
                raise ValueError("'pos' must hold [n][3] values")
            self.n = int(pos.numel()) // 3
            pos = device_input(pos, "pos", 3 * self.n, device, asynchronous)
            dip = device_input(dip, "dip", 3 * self.n, device, asynchronous)
            self.keep += [pos, dip]
            self.pos, self.dip = pos.data_ptr(), dip.data_ptr()
            for t, name in ((force, "force"), (torque, "torque")):
                outs.append(None if t is None else
                            device_output(t, name, 3 * self.n, device).data_ptr())
        else:
            pos = np.ascontiguousarray(np.asarray(pos, np.float64).reshape(-1, 3))
            dip = np.ascontiguousarray(np.asarray(dip, np.float64).reshape(-1, 3))
            if pos.shape != dip.shape:
                raise ValueError("pos and dip must describe the same particles")
            self.keep += [pos, dip]
            self.n = int(pos.shape[0])
            self.pos, self.dip = pos.ctypes.data, dip.ctypes.data
            for t, name in ((force, "force"), (torque, "torque")):
                if t is None:
                    outs.append(None)
                    continue
                if not (isinstance(t, np.ndarray) and t.dtype == np.float64
                        and t.flags.c_contiguous and t.size == 3 * self.n):
                    raise TypeError(
                        f"'{name}' must be a C-contiguous float64 array of shape (n, 3)")
                outs.append(t.ctypes.data)
        self.force, self.torque = outs

    @property
    def flags(self):
        return _lib.P3M_DEVICE_POINTERS if self.device else 0

    def order_after_torch(self, ctx):
        """Device tensors: the solver stream waits for torch's stream."""
        if self.device:
            s = C.c_void_p()
            _lib.raise_for(_lib.load().p3m_dipolar_get_stream(ctx, C.byref(s)), dipolar=ctx)
            _lib.wait_for_torch(s.value)


class DipolarP3M:
    """Dipolar P3M (magnetostatics) for a cubic box."""

    def required_keys(self):
        return {"accuracy"}

    def default_params(self):
        return {"cao": -1,
                "r_cut": -1,
                "alpha": -1,
                "accuracy": -1,
                "mesh": [-1, -1, -1],
                "epsilon": 0.0,
                "mesh_off": [0.5, 0.5, 0.5],
                "prefactor": 0.,
                "single_precision": False,
                "tune": True,
                "timings": 10,
                "verbose": True}

    def __init__(self, device=0, **kwargs):
        L = _lib.load()
        unknown = set(kwargs) - set(_KEYS)
        if unknown:
            raise ValueError(
                f"Only the following keys can be given as keyword arguments: "
                f"{sorted(_KEYS)}, got {sorted(kwargs)} (unknown {sorted(unknown)})")
        missing = self.required_keys() - set(kwargs)
        if missing:
            raise ValueError(
                f"At least the following keys have to be given as keyword arguments: "
                f"{sorted(self.required_keys())}, got {sorted(kwargs)} "
                f"(missing {sorted(missing)})")
        params = self.default_params()
        params.update(kwargs)
        self.validate_params(params)
        params["is_tuned"] = bool(params.get("is_tuned", not params["tune"]))
        self._params = params
        self._ctx = C.c_void_p()
        _lib.raise_for(L.p3m_dipolar_create(int(device), C.byref(self._ctx)))
        self._device = int(device)
        self._activated = False
        self._build(box_l=1.0, skin=0.4)

    def validate_params(self, params):
        """magnetostatics.py DipolarP3M.validate_params and the cubic-mesh
        check of DipolarP3MImpl (dp3m.impl.hpp:88-91)."""
        if params["prefactor"] is None:
            raise ValueError("Parameter 'prefactor' is missing")
        if _is_int(params["mesh"]):
            params["mesh"] = 3 * [params["mesh"]]
        else:
            m = params["mesh"]
            if not (hasattr(m, "__len__") and len(m) == 3 and all(_is_int(x) for x in m)):
                raise ValueError(
                    "Parameter 'mesh' has to be an integer or integer list of length 3")
        if params["epsilon"] == "metallic":
            params["epsilon"] = 0.0
        if not (_is_float(params["epsilon"]) or _is_int(params["epsilon"])):
            raise ValueError("Parameter 'epsilon' has to be a float or 'metallic'")
        mo = params["mesh_off"]
        if not (hasattr(mo, "__len__") and len(mo) == 3
                and all(_is_float(x) or _is_int(x) for x in mo)):
            raise ValueError("Parameter 'mesh_off' has to be a (3,) array_like of values "
                             "between 0.0 and 1.0")
        if not _is_int(params["timings"]):
            raise TypeError("Parameter 'timings' has to be an integer")
        if not isinstance(params["tune"], (bool, np.bool_)):
            raise TypeError("Parameter 'tune' has to be a boolean")
        m = params["mesh"]
        if not (m[0] == m[1] == m[2]):
            raise ValueError("DipolarP3M requires a cubic mesh")

    def _build(self, box_l, skin):
        L = _lib.load()
        p = self._params
        cp = _lib.p3m_dipolar_params()
        L.p3m_dipolar_params_default(C.byref(cp))
        cp.mesh = int(p["mesh"][0])
        cp.cao = int(p["cao"])
        cp.alpha = float(p["alpha"])
        cp.r_cut = float(p["r_cut"])
        cp.accuracy = float(p["accuracy"])
        cp.prefactor = float(p["prefactor"])
        cp.epsilon = float(p["epsilon"])
        for d in range(3):
            cp.mesh_off[d] = float(p["mesh_off"][d])
        cp.box_l = float(box_l)
        cp.skin = float(skin)
        # tuning = !is_tuned (DipolarP3M.hpp:98-104): a tuned actor keeps its
        # parameters when it is activated again or restored from a checkpoint
        cp.tune = 0 if p["is_tuned"] else 1
        cp.timings = int(p["timings"])
        cp.verbose = 1 if p["verbose"] else 0
        _lib.raise_for(L.p3m_dipolar_set_params(self._ctx, C.byref(cp)), dipolar=self._ctx)

    def activate(self, pos, dip, box_l, skin=0.4):
        """Sanity checks, tuning and init (DipolarP3MImpl::on_activation /
        tune, dp3m.cpp:646-682); box_l is the cubic box edge (a scalar or
        three equal values)."""
        box = np.atleast_1d(np.asarray(box_l, np.float64))
        if box.size == 3 and not (box[0] == box[1] == box[2]):
            raise RuntimeError("DipolarP3M: requires a cubic box")
        self._build(box_l=float(box[0]), skin=skin)
        a = _Vec(pos, dip, device=self._device)
        a.order_after_torch(self._ctx)
        _lib.raise_for(_lib.load().p3m_dipolar_activate(self._ctx, a.n, a.pos, a.dip, a.flags),
                       dipolar=self._ctx)
        self._activated = True
        return self

    def get_params(self):
        cp = _lib.p3m_dipolar_params()
        _lib.raise_for(_lib.load().p3m_dipolar_get_params(self._ctx, C.byref(cp)),
                       dipolar=self._ctx)
        return {"prefactor": cp.prefactor, "accuracy": cp.accuracy, "cao": cp.cao,
                "r_cut": cp.r_cut, "alpha": cp.alpha, "mesh": [cp.mesh] * 3,
                "epsilon": cp.epsilon, "mesh_off": list(cp.mesh_off),
                "tune": bool(self._params["tune"]), "is_tuned": not bool(cp.tune),
                "timings": cp.timings, "single_precision": False,
                "verbose": bool(cp.verbose)}

    def __reduce__(self):
        """Pickling keeps the constructor and tuned parameters; the restored
        actor is built with ``is_tuned`` and is not tuned again."""
        p = self.get_params()
        if not p["is_tuned"]:  # still to be tuned: the parameters as given
            p = dict(self._params, is_tuned=False)
        kw = {k: p[k] for k in _KEYS}
        return (_restore_dipolar, (self._device, kw))

    def _require(self):
        if not self._activated:
            raise RuntimeError("DipolarP3M actor is not activated")

    def add_long_range_forces(self, pos, dip, force, torque, sync=True):
        """k-space forces and torques, added in place."""
        self._require()
        a = _Vec(pos, dip, force, torque, device=self._device, asynchronous=not sync)
        a.order_after_torch(self._ctx)
        flags = _lib.P3M_FORCES | a.flags
        if not sync:
            if not a.device:
                raise ValueError("sync=False needs device tensors")
            flags |= _lib.P3M_ASYNC
        _lib.raise_for(_lib.load().p3m_dipolar_solve(
            self._ctx, a.n, a.pos, a.dip, a.force, a.torque, flags, None), dipolar=self._ctx)

    def long_range_energy(self, pos, dip):
        self._require()
        a = _Vec(pos, dip, device=self._device)
        a.order_after_torch(self._ctx)
        e = C.c_double(0.)
        _lib.raise_for(_lib.load().p3m_dipolar_solve(
            self._ctx, a.n, a.pos, a.dip, None, None, _lib.P3M_ENERGY | a.flags, C.byref(e)),
            dipolar=self._ctx)
        return e.value

    def add_short_range_forces(self, pos, dip, force, torque, energy=False, sync=True):
        """Real-space pair forces and torques (DipolarP3M::pair_force,
        dp3m.hpp:160-211), added in place; returns the pair energy with
        energy=True."""
        self._require()
        a = _Vec(pos, dip, force, torque, device=self._device, asynchronous=not sync)
        a.order_after_torch(self._ctx)
        flags = a.flags | (_lib.P3M_ENERGY if energy else 0)
        if not sync:
            if not a.device or energy:
                raise ValueError("sync=False needs device tensors and no energy")
            flags |= _lib.P3M_ASYNC
        e = C.c_double(0.)
        _lib.raise_for(_lib.load().p3m_dipolar_short_range(
            self._ctx, a.n, a.pos, a.dip, a.force, a.torque, flags, C.byref(e)),
            dipolar=self._ctx)
        return e.value if energy else None

    def set_short_range_subdivision(self, cells_per_cutoff):
        """Cells per r_cut of the real space: 0 (by occupancy), 1 or 2."""
        _lib.raise_for(_lib.load().p3m_dipolar_set_short_range_subdivision(
            self._ctx, int(cells_per_cutoff)), dipolar=self._ctx)

    def synchronize(self):
        _lib.raise_for(_lib.load().p3m_dipolar_synchronize(self._ctx), dipolar=self._ctx)

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            try:
                _lib.load().p3m_dipolar_destroy(ctx)
            except Exception:  # pragma: no cover
                pass
            self._ctx = None


def _restore_dipolar(device, kwargs):
    return DipolarP3M(device=device, **kwargs)
