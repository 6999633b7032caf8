"""ctypes binding of the C ABI in include/espresso_amd_p3m.h.

The shared library is built in-tree by ``espresso_amd/build.py``.  There is no
fallback: when the library is missing, importing the solver raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# ESPRESSO_AMD_LIB selects another build of the same library (e.g. a variant
# built for a measurement); the default is the in-tree build.
LIB_PATH = os.environ.get("ESPRESSO_AMD_LIB",
                          os.path.join(HERE, "lib", "libespresso_amd_p3m.so"))

P3M_OK = 0
P3M_ERR_DOMAIN = 1
P3M_ERR_RUNTIME = 2
P3M_ERR_INVALID = 3
P3M_ERR_DEVICE = 4
P3M_ERR_MEMORY = 5

P3M_FORCES = 1
P3M_ENERGY = 2
P3M_PRESSURE = 4
P3M_DEVICE_POINTERS = 8
P3M_ASYNC = 16
P3M_DEVICE_RESULTS = 32

P3M_PREC_FP64 = 0
P3M_PREC_FP32 = 1

P3M_CONV_CPU = 0
P3M_CONV_GPU = 1

P3M_TUNE_RS_MODEL = 0
P3M_TUNE_RS_DEVICE = 1

# every symbol declared in include/espresso_amd_p3m.h
EXPORTED_SYMBOLS = (
    "p3m_params_default", "p3m_create", "p3m_destroy", "p3m_set_params",
    "p3m_get_params", "p3m_is_tuned", "p3m_activate", "p3m_solve",
    "p3m_solve_noalloc", "p3m_noalloc_scratch_bytes", "p3m_set_launch_shape",
    "p3m_get_launch_shape", "p3m_set_deterministic", "p3m_batch_set_deterministic",
    "p3m_prepare_fft_mesh", "p3m_assign_charges", "p3m_charge_assign",
    "p3m_count_charged_particles", "p3m_count_charged_particles_elc",
    "p3m_adapt_epsilon_elc", "p3m_assigned_count", "p3m_solve_assigned",
    "p3m_debug_charge_mesh", "p3m_debug_indices", "p3m_k_space_error",
    "p3m_real_space_error", "p3m_pair_forces_targets", "p3m_pair_force",
    "p3m_pair_energy",
    "p3m_last_timings", "p3m_profile_begin", "p3m_profile_end", "p3m_autotune",
    "p3m_short_range", "p3m_short_range_cells", "p3m_set_short_range_subdivision",
    "p3m_set_force_convention", "p3m_get_force_convention", "p3m_set_tuning_real_space",
    "p3m_elc_params_default", "p3m_set_elc", "p3m_get_elc",
    "p3m_dipolar_params_default", "p3m_dipolar_create", "p3m_dipolar_destroy",
    "p3m_dipolar_last_error", "p3m_dipolar_set_params", "p3m_dipolar_get_params",
    "p3m_dipolar_activate", "p3m_dipolar_solve", "p3m_dipolar_short_range",
    "p3m_dipolar_get_stream", "p3m_dipolar_synchronize",
    "p3m_dipolar_set_short_range_subdivision",
    "p3m_synchronize", "p3m_get_stream", "p3m_last_error",
    "p3m_last_global_error",
    "p3m_batch_create", "p3m_batch_destroy", "p3m_batch_add_system",
    "p3m_batch_particle_offset", "p3m_batch_system_bytes",
    "p3m_batch_working_bytes", "p3m_batch_solve",
    "p3m_batch_graph_solve", "p3m_batch_set_launch_shape", "p3m_batch_get_launch_shape",
    "p3m_batch_autotune",
    "p3m_batch_synchronize", "p3m_batch_get_stream", "p3m_batch_profile_begin",
    "p3m_batch_profile_end", "p3m_batch_last_error", "p3m_tune_model",
    "p3m_serve_create", "p3m_serve_destroy", "p3m_serve_run",
    "p3m_serve_last_error",
)

_STRING_RESULT = ("p3m_last_error", "p3m_last_global_error",
                  "p3m_batch_last_error", "p3m_serve_last_error",
                  "p3m_dipolar_last_error")


class p3m_params(C.Structure):
    _fields_ = [
        ("mesh", C.c_int * 3),
        ("cao", C.c_int),
        ("alpha", C.c_double),
        ("r_cut", C.c_double),
        ("mesh_off", C.c_double * 3),
        ("epsilon", C.c_double),
        ("accuracy", C.c_double),
        ("prefactor", C.c_double),
        ("box_l", C.c_double * 3),
        ("skin", C.c_double),
        ("precision", C.c_int),
        ("tune", C.c_int),
        ("timings", C.c_int),
        ("verbose", C.c_int),
        ("charge_neutrality_tolerance", C.c_double),
    ]


class p3m_launch_shape(C.Structure):
    _fields_ = [(name, C.c_int) for name in
                ("gather", "ncol", "zlines", "fused_lds", "spread", "threads", "chunk",
                 "column", "seg", "col_lanes", "group", "gather_threads")]


class p3m_elc_params(C.Structure):
    _fields_ = [("maxPWerror", C.c_double), ("gap_size", C.c_double),
                ("far_cut", C.c_double), ("neutralize", C.c_int),
                ("delta_mid_top", C.c_double), ("delta_mid_bot", C.c_double),
                ("const_pot", C.c_int), ("pot_diff", C.c_double)]


class p3m_dipolar_params(C.Structure):
    _fields_ = [("mesh", C.c_int), ("cao", C.c_int), ("alpha", C.c_double),
                ("r_cut", C.c_double), ("accuracy", C.c_double),
                ("prefactor", C.c_double), ("epsilon", C.c_double),
                ("mesh_off", C.c_double * 3), ("box_l", C.c_double),
                ("skin", C.c_double), ("tune", C.c_int), ("timings", C.c_int),
                ("verbose", C.c_int)]


_lib = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load the HIP library (raises LibraryMissing when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if os.environ.get("ESPRESSO_AMD_NO_TORCH") != "1":
        # torch wheels bundle their own libamdhip64.so.7; loading it before
        # this library makes the process share one HIP runtime (the same
        # SONAME would otherwise bind torch to the system runtime)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} is not built; run `python espresso_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.p3m_params_default.argtypes = [C.POINTER(p3m_params)]
    L.p3m_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.p3m_destroy.argtypes = [vp]
    L.p3m_set_params.argtypes = [vp, C.POINTER(p3m_params)]
    L.p3m_get_params.argtypes = [vp, C.POINTER(p3m_params)]
    L.p3m_is_tuned.argtypes = [vp, C.POINTER(C.c_int)]
    L.p3m_activate.argtypes = [vp, C.c_int64, vp, vp, C.c_uint]
    L.p3m_solve.argtypes = [vp, C.c_int64, vp, vp, vp, vp, C.c_uint, dp, dp]
    L.p3m_solve_noalloc.argtypes = [vp, C.c_int64, vp, vp, vp, vp, C.c_uint, dp, dp]
    L.p3m_noalloc_scratch_bytes.argtypes = [C.POINTER(C.c_int64)]
    L.p3m_prepare_fft_mesh.argtypes = [vp, C.c_int]
    L.p3m_assign_charges.argtypes = [vp, C.c_int64, vp, vp, C.c_int, C.c_uint]
    L.p3m_charge_assign.argtypes = [vp, C.c_int64, vp, vp, C.c_uint]
    L.p3m_count_charged_particles.argtypes = [vp, C.c_int64, vp, C.c_uint]
    L.p3m_count_charged_particles_elc.argtypes = [vp, C.c_int, C.c_double, C.c_double]
    L.p3m_adapt_epsilon_elc.argtypes = [vp]
    L.p3m_assigned_count.argtypes = [vp, C.POINTER(C.c_int64)]
    L.p3m_solve_assigned.argtypes = [vp, vp, vp, C.c_uint, dp]
    L.p3m_debug_charge_mesh.argtypes = [vp, C.c_int64, vp, vp, vp]
    L.p3m_debug_indices.argtypes = [vp, C.c_int64, vp, vp, vp]
    L.p3m_k_space_error.argtypes = [C.c_double, C.POINTER(C.c_int), C.c_int,
                                    C.c_int, C.c_double, C.c_double, dp, dp]
    L.p3m_real_space_error.argtypes = [C.c_double, C.c_double, C.c_int,
                                       C.c_double, C.c_double, dp, dp]
    L.p3m_pair_forces_targets.argtypes = [vp, C.c_int64, vp, vp, C.c_int, vp, vp]
    L.p3m_pair_force.argtypes = [vp, C.c_double, dp, C.c_double, dp]
    L.p3m_pair_energy.argtypes = [vp, C.c_double, C.c_double, dp]
    L.p3m_last_timings.argtypes = [vp, dp, C.c_int]
    L.p3m_profile_begin.argtypes = [vp, C.c_int]
    L.p3m_profile_end.argtypes = [vp, dp, C.POINTER(C.c_int)]
    L.p3m_synchronize.argtypes = [vp]
    shp = C.POINTER(p3m_launch_shape)
    L.p3m_autotune.argtypes = [vp, C.c_int64, vp, vp, C.c_int, shp, dp, dp]
    L.p3m_get_stream.argtypes = [vp, C.POINTER(vp)]
    L.p3m_short_range.argtypes = [vp, C.c_int64, vp, vp, C.c_uint, vp, dp, dp]
    L.p3m_short_range_cells.argtypes = [vp, C.POINTER(C.c_int)]
    L.p3m_set_short_range_subdivision.argtypes = [vp, C.c_int]
    L.p3m_set_force_convention.argtypes = [vp, C.c_int]
    L.p3m_get_force_convention.argtypes = [vp, C.POINTER(C.c_int)]
    L.p3m_set_tuning_real_space.argtypes = [vp, C.c_int]
    ep = C.POINTER(p3m_elc_params)
    L.p3m_elc_params_default.argtypes = [ep]
    L.p3m_set_elc.argtypes = [vp, ep]
    L.p3m_get_elc.argtypes = [vp, ep, C.POINTER(C.c_int)]
    dpp = C.POINTER(p3m_dipolar_params)
    L.p3m_dipolar_params_default.argtypes = [dpp]
    L.p3m_dipolar_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.p3m_dipolar_destroy.argtypes = [vp]
    L.p3m_dipolar_last_error.argtypes = [vp]
    L.p3m_dipolar_last_error.restype = C.c_char_p
    L.p3m_dipolar_set_params.argtypes = [vp, dpp]
    L.p3m_dipolar_get_params.argtypes = [vp, dpp]
    L.p3m_dipolar_activate.argtypes = [vp, C.c_int64, vp, vp, C.c_uint]
    L.p3m_dipolar_solve.argtypes = [vp, C.c_int64, vp, vp, vp, vp, C.c_uint, dp]
    L.p3m_dipolar_short_range.argtypes = [vp, C.c_int64, vp, vp, vp, vp, C.c_uint, dp]
    L.p3m_dipolar_get_stream.argtypes = [vp, C.POINTER(vp)]
    L.p3m_dipolar_synchronize.argtypes = [vp]
    L.p3m_dipolar_set_short_range_subdivision.argtypes = [vp, C.c_int]
    L.p3m_last_error.argtypes = [vp]
    L.p3m_last_error.restype = C.c_char_p
    L.p3m_last_global_error.argtypes = []
    L.p3m_last_global_error.restype = C.c_char_p
    ip = C.POINTER(C.c_int)
    L.p3m_batch_create.argtypes = [C.c_int, C.c_int, C.POINTER(vp)]
    L.p3m_batch_destroy.argtypes = [vp]
    L.p3m_batch_add_system.argtypes = [vp, C.POINTER(p3m_params), C.c_int64, ip]
    L.p3m_batch_particle_offset.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
    L.p3m_batch_system_bytes.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
    L.p3m_batch_working_bytes.argtypes = [vp, C.c_int, ip, C.POINTER(C.c_int64)]
    L.p3m_batch_solve.argtypes = [vp, C.c_int, ip, vp, vp, vp, C.c_uint, dp, dp]
    L.p3m_batch_graph_solve.argtypes = [vp, C.c_int, ip, vp, vp, vp]
    L.p3m_set_deterministic.argtypes = [vp, C.c_int]
    L.p3m_batch_set_deterministic.argtypes = [vp, C.c_int]
    L.p3m_set_launch_shape.argtypes = [vp, shp]
    L.p3m_get_launch_shape.argtypes = [vp, shp]
    L.p3m_batch_set_launch_shape.argtypes = [vp, shp]
    L.p3m_batch_get_launch_shape.argtypes = [vp, shp]
    L.p3m_batch_autotune.argtypes = [vp, C.c_int, ip, vp, vp, vp, C.c_int, shp, dp, dp]
    L.p3m_batch_synchronize.argtypes = [vp]
    L.p3m_batch_get_stream.argtypes = [vp, C.POINTER(vp)]
    L.p3m_batch_profile_begin.argtypes = [vp, C.c_int]
    L.p3m_batch_profile_end.argtypes = [vp, dp, ip]
    L.p3m_batch_last_error.argtypes = [vp]
    L.p3m_batch_last_error.restype = C.c_char_p
    L.p3m_tune_model.argtypes = [C.POINTER(p3m_params), C.c_int, C.c_double,
                                 C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_double, C.c_int]
    i64p = C.POINTER(C.c_int64)
    L.p3m_serve_create.argtypes = [vp, C.c_int, C.c_int64, C.c_int, C.c_double,
                                   C.POINTER(vp)]
    L.p3m_serve_destroy.argtypes = [vp]
    L.p3m_serve_run.argtypes = [vp, C.c_int, dp, ip, C.POINTER(C.c_uint64), vp,
                                vp, vp, vp, i64p, dp, dp, ip, i64p, ip]
    L.p3m_serve_last_error.argtypes = [vp]
    L.p3m_serve_last_error.restype = C.c_char_p
    for name in EXPORTED_SYMBOLS:
        if name not in _STRING_RESULT:
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def wait_for_torch(stream_handle):
    """Order the solver stream after the work torch has queued so far on its
    current stream (the producer of the device tensors handed over)."""
    import torch
    if stream_handle:
        torch.cuda.ExternalStream(stream_handle).wait_stream(
            torch.cuda.current_stream())


def raise_for(status, ctx=None, batch=None, server=None, dipolar=None):
    """Map a C status code to the Python exception the reference raises."""
    if status == P3M_OK:
        return
    L = load()
    if dipolar:
        msg = L.p3m_dipolar_last_error(dipolar)
    elif server:
        msg = L.p3m_serve_last_error(server)
    elif batch:
        msg = L.p3m_batch_last_error(batch)
    elif ctx:
        msg = L.p3m_last_error(ctx)
    else:
        msg = L.p3m_last_global_error()
    msg = msg.decode() if msg else ""
    if status in (P3M_ERR_DOMAIN, P3M_ERR_INVALID):
        raise ValueError(msg)
    if status == P3M_ERR_MEMORY:
        raise MemoryError(msg)
    raise RuntimeError(msg)
