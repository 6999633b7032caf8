"""TEST INFRASTRUCTURE ONLY: ctypes view of the FP64 CPU oracle.

The oracle (``p3m_oracle.c``) restates ESPResSo's ``CoulombP3MImpl<double,
Arch::CPU>`` on one rank (src/core/electrostatics/p3m.cpp:97-591) plus the
closed-form error estimates (p3m.cpp:145-243), the real-space pair terms
(p3m.hpp:146-180) and an independent FP64 Ewald sum.  Only tests/,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may use
this module, and only to check or time the reference algorithm; the product
path never calls it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libp3m_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle in place (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "libp3m_oracle.so"],
                   check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    L.p3mo_sinc.restype = C.c_double
    L.p3mo_sinc.argtypes = [C.c_double]
    L.p3mo_cotangent_sum.restype = C.c_double
    L.p3mo_cotangent_sum.argtypes = [C.c_int, C.c_int, C.c_double]
    L.p3mo_as_erfc_part.restype = C.c_double
    L.p3mo_as_erfc_part.argtypes = [C.c_double]
    L.p3mo_bspline.restype = C.c_double
    L.p3mo_bspline.argtypes = [C.c_int, C.c_int, C.c_double]
    L.p3mo_bspline_coeffs.restype = None
    L.p3mo_bspline_coeffs.argtypes = [C.c_int, C.c_int, _lp,
                                      C.POINTER(C.c_int64)]
    L.p3mo_meshift.restype = None
    L.p3mo_meshift.argtypes = [C.c_int, C.c_int, _ip]
    L.p3mo_fft3d.restype = None
    L.p3mo_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int]
    L.p3mo_new.restype = C.c_void_p
    L.p3mo_new.argtypes = [_ip, C.c_int, C.c_double, C.c_double, _dp,
                           C.c_double, _dp, C.c_double, C.c_double]
    L.p3mo_free.restype = None
    L.p3mo_free.argtypes = [C.c_void_p]
    L.p3mo_geometry.restype = None
    L.p3mo_geometry.argtypes = [C.c_void_p, _ip, _dp]
    L.p3mo_weights.restype = C.c_int
    L.p3mo_weights.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int), _ip,
                               _dp]
    L.p3mo_count_charged.restype = None
    L.p3mo_count_charged.argtypes = [C.c_void_p, C.c_int, _dp]
    L.p3mo_add_long_range_forces.restype = C.c_int
    L.p3mo_add_long_range_forces.argtypes = [C.c_void_p, C.c_int, _dp, _dp,
                                             C.c_void_p, _dp]
    L.p3mo_long_range_energy.restype = C.c_double
    L.p3mo_long_range_energy.argtypes = [C.c_void_p, C.c_int, _dp, _dp,
                                         C.c_void_p]
    L.p3mo_long_range_pressure.restype = C.c_int
    L.p3mo_long_range_pressure.argtypes = [C.c_void_p, C.c_int, _dp, _dp,
                                           _dp]
    L.p3mo_charge_mesh.restype = C.c_int
    L.p3mo_charge_mesh.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
    L.p3mo_fields.restype = C.c_int
    L.p3mo_fields.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.p3mo_influence.restype = None
    L.p3mo_influence.argtypes = [C.c_void_p, _dp, _dp]
    L.p3mo_real_space_error.restype = C.c_double
    L.p3mo_real_space_error.argtypes = [C.c_double, C.c_double, C.c_int,
                                        C.c_double, C.c_double, _dp]
    L.p3mo_k_space_error.restype = C.c_double
    L.p3mo_k_space_error.argtypes = [C.c_double, _ip, C.c_int, C.c_int,
                                     C.c_double, C.c_double, _dp]
    L.p3mo_calculate_accuracy.restype = None
    L.p3mo_calculate_accuracy.argtypes = [C.c_double, _ip, C.c_int,
                                          C.c_double, C.c_int, C.c_double,
                                          C.c_double, _dp, _dp]
    L.p3mo_tune_r_cut.restype = C.c_int
    L.p3mo_tune_r_cut.argtypes = [C.c_double, _ip, C.c_int, C.c_int,
                                  C.c_double, C.c_double, _dp, C.c_double,
                                  C.c_double, C.c_double, _dp]
    L.p3mo_real_space.restype = C.c_double
    L.p3mo_real_space.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double,
                                  C.c_double, C.c_double, _dp, _dp]
    L.p3mo_ewald_kspace.restype = C.c_double
    L.p3mo_ewald_kspace.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double,
                                    C.c_double, C.c_double, C.c_int, _dp]
    L.p3mo_real_space_subset.restype = None
    L.p3mo_real_space_subset.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double,
                                         C.c_double, C.c_double, C.c_int,
                                         C.c_int, _dp]
    L.p3mo_set_threads.restype = None
    L.p3mo_set_threads.argtypes = [C.c_int]
    L.p3mo_get_threads.restype = C.c_int
    L.p3mo_get_threads.argtypes = []
    _lib = L
    return L


def set_threads(n):
    """Threads of the oracle's solve: 1 (default) is the serial restatement
    the parity checks use; more threads time the all-cores CPU baseline
    (the charge mesh is then summed from per-thread meshes)."""
    lib().p3mo_set_threads(int(n))


def _v3(x, dtype=np.float64):
    a = np.asarray(x, dtype=dtype)
    if a.ndim == 0:
        a = np.full(3, a, dtype=dtype)
    return np.ascontiguousarray(a)


def _pos(pos):
    return np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))


class OracleP3M:
    """FP64 CPU P3M (k-space part) at fixed parameters."""

    def __init__(self, box_l, mesh, cao, alpha, r_cut, prefactor,
                 epsilon=0.0, mesh_off=(-1., -1., -1.), skin=0.4):
        L = lib()
        self.box_l = _v3(box_l)
        self.mesh = _v3(mesh, np.int32)
        self.cao = int(cao)
        self.alpha = float(alpha)
        self.r_cut = float(r_cut)
        self.prefactor = float(prefactor)
        self.epsilon = float(epsilon)
        self.skin = float(skin)
        self._h = L.p3mo_new(self.mesh, self.cao, self.alpha, self.r_cut,
                             _v3(mesh_off), self.epsilon, self.box_l,
                             self.prefactor, self.skin)
        if not self._h:
            raise RuntimeError("oracle: invalid CA-mesh geometry")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().p3mo_free(self._h)
            self._h = None

    def geometry(self):
        ints = np.zeros(32, np.int32)
        dbls = np.zeros(16, np.float64)
        lib().p3mo_geometry(self._h, ints, dbls)
        return dict(dim=ints[0:3].copy(), size=int(ints[3]),
                    ld_ind=ints[4:7].copy(), inner=ints[7:10].copy(),
                    in_ld=ints[10:13].copy(), in_ur=ints[13:16].copy(),
                    margin=ints[16:22].copy(), q_2_off=int(ints[22]),
                    q_21_off=int(ints[23]), ld_pos=dbls[0:3].copy(),
                    a=dbls[3:6].copy(), ai=dbls[6:9].copy(),
                    cao_cut=dbls[9:12].copy())

    def weights(self, pos):
        """CA-mesh linear index, nearest mesh point and weights (3, cao)."""
        ind = C.c_int(0)
        nmp = np.zeros(3, np.int32)
        w = np.zeros(3 * self.cao)
        rc = lib().p3mo_weights(self._h, _v3(pos), C.byref(ind), nmp, w)
        if rc:
            raise ValueError("position outside the charge-assignment mesh")
        return ind.value, nmp, w.reshape(3, self.cao)

    def forces(self, pos, q, pos_unfolded=None):
        pos = _pos(pos)
        q = np.ascontiguousarray(q, dtype=np.float64)
        f = np.zeros_like(pos)
        up = None
        if pos_unfolded is not None:
            up_arr = _pos(pos_unfolded)
            up = up_arr.ctypes.data
        rc = lib().p3mo_add_long_range_forces(self._h, len(q), pos, q, up, f)
        if rc:
            raise ValueError("position outside the charge-assignment mesh")
        return f

    def energy(self, pos, q, pos_unfolded=None):
        pos = _pos(pos)
        q = np.ascontiguousarray(q, dtype=np.float64)
        up = None
        if pos_unfolded is not None:
            up_arr = _pos(pos_unfolded)
            up = up_arr.ctypes.data
        return lib().p3mo_long_range_energy(self._h, len(q), pos, q, up)

    def pressure(self, pos, q):
        out = np.zeros(9)
        rc = lib().p3mo_long_range_pressure(
            self._h, len(q), _pos(pos),
            np.ascontiguousarray(q, dtype=np.float64), out)
        if rc:
            raise ValueError("position outside the charge-assignment mesh")
        return out.reshape(3, 3)

    def charge_mesh(self, pos, q):
        out = np.zeros(int(np.prod(self.mesh)))
        rc = lib().p3mo_charge_mesh(self._h, len(q), _pos(pos),
                                    np.ascontiguousarray(q, np.float64), out)
        if rc:
            raise ValueError("position outside the charge-assignment mesh")
        return out.reshape(tuple(self.mesh))

    def fields(self, pos, q):
        m = tuple(self.mesh)
        outs = [np.zeros(int(np.prod(m))) for _ in range(3)]
        rc = lib().p3mo_fields(self._h, len(q), _pos(pos),
                               np.ascontiguousarray(q, np.float64), *outs)
        if rc:
            raise ValueError("position outside the charge-assignment mesh")
        return [o.reshape(m) for o in outs]

    def influence(self):
        m = int(np.prod(self.mesh))
        gf = np.zeros(m)
        ge = np.zeros(m)
        lib().p3mo_influence(self._h, gf, ge)
        return gf.reshape(tuple(self.mesh)), ge.reshape(tuple(self.mesh))


def sinc(x):
    return lib().p3mo_sinc(float(x))


def bspline(order, i, x):
    return lib().p3mo_bspline(int(order), int(i), float(x))


def bspline_coeffs(order, i):
    num = np.zeros(order, np.int64)
    den = C.c_int64(0)
    lib().p3mo_bspline_coeffs(order, i, num, C.byref(den))
    return num, den.value


def _bspline_eval(order, x):
    """(order, len(x)) B-spline weights of the stencil points at offsets x in
    [-1/2, 1/2] from the exact integer tables."""
    x = np.asarray(x, np.float64)
    w = np.empty((order, x.size))
    for i in range(order):
        num, den = bspline_coeffs(order, i)
        acc = np.full(x.size, float(num[order - 1]))
        for m in range(order - 2, -1, -1):
            acc = float(num[m]) + x * acc
        w[i] = acc / den
    return w


def gpu_mesh_coordinates(pos, box_l, mesh, cao):
    """First stencil point nmp (not wrapped) and offset dist of each particle
    under the index rule of the reference's GPU kernels
    (p3m_gpu_cuda.cu:334-345): positions folded into [0, L) in FP64
    (x - L floor(x / L)) and cast to float, m = x hi - pos_shift with
    hi = float(M) / float(L) and pos_shift = float((cao - 1) / 2) (integer
    division) rounded once (an FMA; the FP64 product of two floats is exact
    and the difference needs at most 53 bits, so rounding it to float is
    the FMA's single rounding), nmp = floor(m + 1/2) in float, dist = m - nmp.
    Returns (nmp int64 [n, 3], dist float64 [n, 3])."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    M = np.asarray(mesh, np.int64).reshape(3)
    L = np.asarray(box_l, np.float64).reshape(3)
    hi = (M.astype(np.float32) / L.astype(np.float32)).astype(np.float32)
    ps = np.float32((cao - 1) // 2)
    xf = pos - L * np.floor(pos / L)
    xf = np.where(xf >= L, xf - L, xf).astype(np.float32)
    m = (xf.astype(np.float64) * hi.astype(np.float64) - float(ps)).astype(np.float32)
    nmp = np.floor((m + np.float32(0.5)).astype(np.float32)).astype(np.int64)
    dist = (m - nmp.astype(np.float32)).astype(np.float32).astype(np.float64)
    return nmp, dist


def gpu_kspace_forces(pos, q, box_l, mesh, cao, alpha, prefactor):
    """k-space forces of the reference's GPU P3M (P3MGPU;
    p3m_gpu_cuda.cu:165-723), FP64 arithmetic with the index rule of its
    FP32 kernels (:315-360, 424-470):

    - positions folded into the box and cast to float
      (GpuParticleData.cpp:90); m = x hi - pos_shift with hi = float(M) /
      float(L), pos_shift = float((cao - 1) / 2) (integer division), one
      rounding (the CUDA compiler contracts it to an FMA); nmp =
      floor(m + 1/2), dist = m - nmp, both in float
    - stencil points nmp + k (k < cao), periodic wrap
    - unnormalised r2c FFT; G_hat = 2/pi exp(-pi^2 n_L^2 / alpha^2) /
      (n_L^2 U^2) with n_L = n / L, U^2 = prod_d sinc(n_d / M_d)^(2 cao)
      (P3M_BRILLOUIN = 0), zero at n = 0 and where every index is a
      multiple of M/2 (:213-248)
    - i k = 2 pi i n / L, n = j - M for j > M/2, so the Nyquist index keeps
      +M/2, and n_z = j on the half spectrum (:265-295)
    - unnormalised c2r (imaginary parts of the kz = 0 and Nyquist bins are
      dropped); F = -prefactor / (2 V) q sum_stencil w E (:424-723)."""
    q = np.asarray(q, np.float64)
    M = np.asarray(mesh, np.int64).reshape(3)
    L = np.asarray(box_l, np.float64).reshape(3)
    nmp, dist = gpu_mesh_coordinates(pos, box_l, mesh, cao)
    w = [_bspline_eval(cao, dist[:, d]) for d in range(3)]          # (cao, n)
    idx = [np.mod(nmp[:, d][None, :] + np.arange(cao)[:, None], M[d]) for d in range(3)]
    rho = np.zeros(tuple(M))
    for i in range(cao):
        for j in range(cao):
            for k in range(cao):
                np.add.at(rho, (idx[0][i], idx[1][j], idx[2][k]),
                          q * w[0][i] * w[1][j] * w[2][k])
    rhat = np.fft.rfftn(rho)
    K2 = M[2] // 2 + 1
    n = []
    for d, size in enumerate((M[0], M[1], K2)):
        j = np.arange(size)
        n.append(np.where(j > M[d] // 2, j - M[d], j).astype(np.float64))
    nx, ny, nz = np.meshgrid(n[0], n[1], n[2], indexing="ij")
    nm2 = (nx / L[0]) ** 2 + (ny / L[1]) ** 2 + (nz / L[2]) ** 2
    S = (np.sinc(nx / M[0]) * np.sinc(ny / M[1]) * np.sinc(nz / M[2])) ** (2 * cao)
    with np.errstate(divide="ignore", invalid="ignore"):
        G = 2. / np.pi * np.exp(-(np.pi / alpha) ** 2 * nm2) / (nm2 * S)
    jx, jy, jz = np.meshgrid(np.arange(M[0]), np.arange(M[1]), np.arange(K2), indexing="ij")
    zero = ((jx == 0) & (jy == 0) & (jz == 0)) | (
        (jx % (M[0] // 2) == 0) & (jy % (M[1] // 2) == 0) & (jz % (M[2] // 2) == 0))
    G[zero] = 0.
    phi = rhat * G
    N = float(np.prod(M))
    E = [np.fft.irfftn(2j * np.pi * nd / L[d] * phi, s=tuple(M), axes=(0, 1, 2)) * N
         for d, nd in enumerate((nx, ny, nz))]
    pref = prefactor / (2. * float(np.prod(L.astype(np.float32).astype(np.float64))))
    f = np.zeros((len(q), 3))
    for i in range(cao):
        for j in range(cao):
            for k in range(cao):
                ww = w[0][i] * w[1][j] * w[2][k]
                sel = (idx[0][i], idx[1][j], idx[2][k])
                for d in range(3):
                    f[:, d] += ww * E[d][sel]
    return -pref * q[:, None] * f


def meshift(n, zero_out_midpoint=False):
    out = np.zeros(n, np.int32)
    lib().p3mo_meshift(n, int(bool(zero_out_midpoint)), out)
    return out


def fft3d(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    buf = a.view(np.float64).reshape(-1).copy()
    lib().p3mo_fft3d(a.shape[0], a.shape[1], a.shape[2], buf, int(sign))
    return buf.view(np.complex128).reshape(a.shape)


def cotangent_sum(cao, n, mesh_i):
    return lib().p3mo_cotangent_sum(cao, n, mesh_i)


def real_space_error(pref, r_cut_iL, n_c_part, sum_q2, alpha_L, box_l):
    return lib().p3mo_real_space_error(pref, r_cut_iL, n_c_part, sum_q2,
                                       alpha_L, _v3(box_l))


def k_space_error(pref, mesh, cao, n_c_part, sum_q2, alpha_L, box_l):
    return lib().p3mo_k_space_error(pref, _v3(mesh, np.int32), cao, n_c_part,
                                    sum_q2, alpha_L, _v3(box_l))


def calculate_accuracy(pref, mesh, cao, r_cut_iL, n_c_part, sum_q2,
                       accuracy, box_l):
    out = np.zeros(4)
    lib().p3mo_calculate_accuracy(pref, _v3(mesh, np.int32), cao, r_cut_iL,
                                  n_c_part, sum_q2, accuracy, _v3(box_l), out)
    return tuple(out)


def tune_r_cut(pref, mesh, cao, n_c_part, sum_q2, accuracy, box_l,
               skin=0.4):
    """Bisection of TuningAlgorithm::get_mc_time for fixed (mesh, cao).

    Returns (r_cut_iL, alpha_L, accuracy) or None."""
    box_l = _v3(box_l)
    r_max = (min(box_l) / 2. - skin) / box_l[0]
    out = np.zeros(3)
    rc = lib().p3mo_tune_r_cut(pref, _v3(mesh, np.int32), cao, n_c_part,
                               sum_q2, accuracy, box_l, skin, 0.0, r_max, out)
    if rc:
        return None
    return tuple(out)


def real_space(pos, q, box_l, prefactor, alpha, r_cut, virial=False):
    pos = _pos(pos)
    f = np.zeros_like(pos)
    vir = np.zeros(9)
    e = lib().p3mo_real_space(len(q), pos, np.ascontiguousarray(q, np.float64),
                              _v3(box_l), prefactor, alpha, r_cut, f, vir)
    if virial:
        return e, f, vir.reshape(3, 3)
    return e, f


def real_space_subset(pos, q, box_l, prefactor, alpha, r_cut, n_out,
                      exact_erfc=False):
    """Real-space forces on the first n_out particles from all others:
    the P3M pair force (A&S erfc) or, with exact_erfc, the exact Ewald one."""
    pos = _pos(pos)
    f = np.zeros((int(n_out), 3))
    lib().p3mo_real_space_subset(len(q), pos, np.ascontiguousarray(q, np.float64),
                                 _v3(box_l), prefactor, alpha, r_cut, int(n_out),
                                 1 if exact_erfc else 0, f)
    return f


def ewald_kspace(pos, q, box_l, prefactor, alpha, kmax, n_out=None):
    """FP64 Ewald reciprocal-space energy and forces on the first n_out."""
    pos = _pos(pos)
    n = len(q)
    n_out = n if n_out is None else int(n_out)
    f = np.zeros((n_out, 3))
    e = lib().p3mo_ewald_kspace(n, pos, np.ascontiguousarray(q, np.float64),
                                _v3(box_l), prefactor, alpha, kmax, n_out, f)
    return e, f
