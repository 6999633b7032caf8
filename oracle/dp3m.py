"""Dipolar P3M oracle: a numpy restatement of DipolarP3MImpl<double, CPU>
(src/core/magnetostatics/dp3m.cpp, dp3m.hpp, p3m/influence_function_dipolar.hpp)
for a cubic box and mesh.  Test infrastructure only: tests and the fixture
scripts use it as the checker; the product path never imports it.

Charge assignment reuses the Coulomb oracle's bit-exact CA-mesh weights
(the reference assigns each dipole component with the same weights,
AssignDipole, dp3m.cpp:157-174), so a component mesh is the charge mesh of
"charges" mu_d.  Transforms are numpy's complex FFTs, unnormalised forward
(sign -1) and backward (sign +1) as FFTW in the reference.
"""
import math

import numpy as np

from oracle import oracle as O

SQRT_PI_INV = 1. / math.sqrt(math.pi)


def meshift(M, zero_nyquist):
    j = np.arange(M)
    s = np.where(2 * j < M, j, j - M)
    if zero_nyquist:
        s = np.where(2 * j == M, 0, s)
    return s.astype(np.float64)


def influence_dipolar(M, cao, alpha_L, L, S):
    """grid_influence_function<S> (influence_function_dipolar.hpp:40-140)
    with P3M_BRILLOUIN = 0 on the full mesh [M][M][M]: S = 3 for forces and
    torques, S = 2 for the energy.  Zero where every index is a multiple of
    M/2."""
    shift = meshift(M, False)
    dop = meshift(M, True)
    sx, sy, sz = np.meshgrid(shift, shift, shift, indexing="ij")
    dx, dy, dz = np.meshgrid(dop, dop, dop, indexing="ij")
    nm2 = sx * sx + sy * sy + sz * sz
    U = (np.sinc(sx / M) * np.sinc(sy / M) * np.sinc(sz / M)) ** (2 * cao)
    f2 = (math.pi / alpha_L) ** 2
    dn = dx * sx + dy * sy + dz * sz
    d2 = dx * dx + dy * dy + dz * dz
    with np.errstate(divide="ignore", invalid="ignore"):
        ex = np.where(f2 * nm2 < 30., np.exp(-f2 * nm2), 0.)
        num = U * ex / nm2 * dn ** S
        G = num / (d2 ** S * U * U)
    ix = np.arange(M)
    half = M // 2
    keep = ~((ix[:, None, None] % half == 0) & (ix[None, :, None] % half == 0) &
             (ix[None, None, :] % half == 0))
    G = np.where(keep & np.isfinite(G), G, 0.)
    return G * (M ** 3 * 2. / (L * L))


def self_energy_sum(M, cao, G_energy):
    """grid_influence_function_self_energy (influence_function_dipolar.hpp:
    160-218): sum_k G_e(k) U2(k) |d_op(k)|^2, U2 over aliases m in
    [-1, 1]^3 (P3M_BRILLOUIN + 1)."""
    shift = meshift(M, False)
    dop = meshift(M, True)
    sx, sy, sz = np.meshgrid(shift, shift, shift, indexing="ij")
    dx, dy, dz = np.meshgrid(dop, dop, dop, indexing="ij")
    U2 = np.zeros_like(sx)
    for mx in (-1, 0, 1):
        for my in (-1, 0, 1):
            for mz in (-1, 0, 1):
                U2 += (np.sinc(sx / M + mx * M) * np.sinc(sy / M + my * M) *
                       np.sinc(sz / M + mz * M)) ** (2 * cao)
    ix = np.arange(M)
    half = M // 2
    keep = ~((ix[:, None, None] % half == 0) & (ix[None, :, None] % half == 0) &
             (ix[None, None, :] % half == 0))
    return float(np.sum(np.where(keep, G_energy * U2 * (dx * dx + dy * dy + dz * dz), 0.)))


class OracleDP3M:
    """k-space part of dipolar P3M (long_range_kernel, dp3m.cpp:246-449,
    calc_surface_term :451-517, calc_energy_correction :892-901)."""

    def __init__(self, L, M, cao, alpha, r_cut, prefactor, epsilon=0., accuracy=1e-4,
                 mesh_off=(0.5, 0.5, 0.5), skin=0.4):
        self.L, self.M, self.cao = float(L), int(M), int(cao)
        self.alpha, self.r_cut, self.prefactor = float(alpha), float(r_cut), float(prefactor)
        self.epsilon, self.accuracy = float(epsilon), float(accuracy)
        self.ca = O.OracleP3M([L] * 3, [M] * 3, cao, alpha, r_cut, 1.0, mesh_off=mesh_off,
                              skin=skin)
        alpha_L = self.alpha * self.L
        self.Gf = influence_dipolar(self.M, self.cao, alpha_L, self.L, 3)
        self.Ge = influence_dipolar(self.M, self.cao, alpha_L, self.L, 2)
        geo = self.ca.geometry()
        self.marg = geo["margin"][0::2]

    def _meshes(self, pos, dip):
        return [self.ca.charge_mesh(pos, dip[:, d]) for d in range(3)]

    def _gather(self, fields, pos):
        """interpolate each real field at each particle: [n][len(fields)]"""
        M, cao = self.M, self.cao
        out = np.zeros((len(pos), len(fields)))
        for i, p in enumerate(pos):
            _, nmp, w = self.ca.weights(p)
            ix = [np.mod(nmp[d] - self.marg[d] + np.arange(cao), M) for d in range(3)]
            W = w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]
            sel = np.ix_(ix[0], ix[1], ix[2])
            for c, F in enumerate(fields):
                out[i, c] = np.sum(W * F[sel])
        return out

    def solve(self, pos, dip):
        """(forces [n,3], torques [n,3], energy) of the k-space part."""
        pos = np.asarray(pos, np.float64).reshape(-1, 3)
        dip = np.asarray(dip, np.float64).reshape(-1, 3)
        M, L = self.M, self.L
        mu2 = float(np.sum(dip * dip))
        dipole_prefac = self.prefactor / M ** 3
        dop = meshift(M, True)
        D = np.meshgrid(dop, dop, dop, indexing="ij")
        A = [np.fft.fftn(m) for m in self._meshes(pos, dip)]
        z = D[0] * A[0] + D[1] * A[1] + D[2] * A[2]
        # energy (dp3m.cpp:270-311)
        e = float(np.sum(self.Ge * np.abs(z) ** 2)) * dipole_prefac * math.pi / L
        e -= self.prefactor * mu2 * SQRT_PI_INV * (2. / 3.) * self.alpha ** 3
        V = L ** 3
        alpha_L = self.alpha * L
        ukp3m = math.pi * self_energy_sum(M, self.cao, self.Ge) / (3. * L * M ** 3) * V
        eself = -2. * alpha_L ** 3 * SQRT_PI_INV / 3.
        corr = -mu2 * (ukp3m + eself + 2. * math.pi / 3.)
        e += self.prefactor * corr / V
        # torques (:313-362)
        kt = self.Ge * z
        wn = 2. * math.pi / L
        E = [np.real(np.fft.ifftn(D[d] * kt)) * M ** 3 for d in range(3)]
        Eg = self._gather(E, pos) * (dipole_prefac * wn)
        tq = -np.cross(dip, Eg)
        # forces (:364-412): 6 symmetric second-derivative meshes
        kf = -1j * self.Gf * z
        F = {}
        for a in range(3):
            for b in range(a, 3):
                F[a, b] = np.real(np.fft.ifftn(D[a] * D[b] * kf)) * M ** 3
        order = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
        Fg = self._gather([F[k] for k in order], pos) * (dipole_prefac * wn * wn)
        H = np.zeros((len(pos), 3, 3))
        for c, (a, b) in enumerate(order):
            H[:, a, b] = Fg[:, c]
            H[:, b, a] = Fg[:, c]
        f = np.einsum("nab,nb->na", H, dip)
        # surface term (:451-517)
        if self.epsilon != 0.:
            pref = self.prefactor * 4. * math.pi / V / (2. * self.epsilon + 1.)
            Mt = np.sum(dip, axis=0)
            e += 0.5 * pref * float(np.sum(dip @ Mt))
            tq -= pref * np.cross(dip, Mt)
        return f, tq, e


def pair_terms(pos, dip, L, prefactor, alpha, r_cut, accuracy):
    """Real-space dipolar pair forces, torques and energy over all pairs
    within r_cut (minimum image): DipolarP3M::pair_force / pair_energy
    (dp3m.hpp:160-235), torque on the partner from calc_opposing_force
    (forces_inline.hpp:159-171).  B_r uses the A&S erfc for accuracy > 5e-6,
    libm erfc otherwise."""
    from scipy.special import erfc
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    dip = np.asarray(dip, np.float64).reshape(-1, 3)
    n = len(pos)
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r2 = np.sum(d * d, axis=2)
    ok = (r2 < r_cut * r_cut) & (r2 > 0.)
    ok &= (np.sum(dip * dip, axis=1)[:, None] > 0.) & (np.sum(dip * dip, axis=1)[None, :] > 0.)
    r = np.sqrt(np.where(ok, r2, 1.))
    ad = alpha * r
    coeff = 2. * alpha * SQRT_PI_INV
    ex = np.exp(-ad * ad)
    r2i = 1. / np.where(ok, r2, 1.)
    if accuracy > 5e-6:
        t = 1. / (1. + 0.3275911 * ad)
        ep = t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 +
                  t * (-1.453152027 + t * 1.061405429))))
        B = (ep / r + coeff) * ex * r2i
    else:
        B = (erfc(ad) / r + coeff * ex) * r2i
    common = alpha * alpha * coeff * ex
    Cr = r2i * (3. * B + 2. * common)
    Dr = r2i * (5. * Cr + 4. * common * alpha * alpha)
    B, Cr, Dr = (np.where(ok, x, 0.) for x in (B, Cr, Dr))
    mimj = dip @ dip.T
    mir = np.einsum("ijk,ik->ij", d, dip)
    mjr = np.einsum("ijk,jk->ij", d, dip)
    f = prefactor * ((mimj * Cr)[:, :, None] * d + dip[:, None, :] * (mjr * Cr)[:, :, None] +
                     dip[None, :, :] * (mir * Cr)[:, :, None] - (mir * mjr * Dr)[:, :, None] * d)
    mixmj = np.cross(dip[:, None, :], dip[None, :, :])
    mixr = np.cross(dip[:, None, :], d)
    tq = prefactor * (-mixmj * B[:, :, None] + mixr * (mjr * Cr)[:, :, None])
    e = prefactor * (mimj * B - mir * mjr * Cr)
    return f.sum(axis=1), tq.sum(axis=1), 0.5 * float(np.sum(e))


# ---------------------------------------------------------------------------
# error estimates and tuning (dp3m.cpp:533-820)
# ---------------------------------------------------------------------------
def real_space_error(L, r_cut_iL, n_c_part, sum_mu2, alpha_L):
    """dp3m_real_space_error (dp3m.cpp:757-787), eq. (33) of Wang & Holm."""
    rc = r_cut_iL * L
    rc2 = rc * rc
    rc4 = rc2 * rc2
    a2 = alpha_L * alpha_L / (L * L)
    c = sum_mu2 * math.exp(-a2 * rc2)
    cc = 4. * a2 * a2 * rc2 * rc2 + 6. * a2 * rc2 + 3.
    dc = 8. * a2 ** 3 * rc2 ** 3 + 20. * a2 * a2 * rc4 + 30. * a2 * rc2 + 15.
    con = 1. / math.sqrt(L ** 3 * a2 * a2 * rc * rc4 * rc4 * n_c_part)
    return c * con * math.sqrt((13. / 6.) * cc * cc + (2. / 15.) * dc * dc -
                               (13. / 15.) * cc * dc)


def k_space_error(L, M, cao, n_c_part, sum_mu2, alpha_L):
    """dp3m_k_space_error (dp3m.cpp:714-755) with P3M_BRILLOUIN = 0."""
    mesh_i = 1. / M
    idx = np.arange(-(M // 2), M // 2)
    cs1 = np.array([O.cotangent_sum(cao, int(n), mesh_i) for n in idx])
    nx, ny, nz = np.meshgrid(idx, idx, idx, indexing="ij")
    cs = cs1[:, None, None] * cs1[None, :, None] * cs1[None, None, :]
    n2 = (nx * nx + ny * ny + nz * nz).astype(np.float64)
    nz0 = n2 > 0
    n2s = np.where(nz0, n2, 1.)
    f1 = (math.pi / alpha_L) ** 2
    ex = np.exp(-f1 * n2s)
    U2 = (np.sinc(nx * mesh_i) * np.sinc(ny * mesh_i) * np.sinc(nz * mesh_i)) ** (2 * cao)
    alias1 = ex * ex * n2s
    alias2 = U2 * ex * n2s ** 3 / n2s
    d = alias1 - (alias2 / cs) ** 2 / n2s ** 3
    with np.errstate(divide="ignore", invalid="ignore"):
        good = nz0 & (d > 0.) & (np.abs(d / alias1) > 1e-14)
    he_q = float(np.sum(np.where(good, d, 0.)))
    return 8. * math.pi ** 2 / 3. * sum_mu2 * math.sqrt(he_q / n_c_part) / L ** 4


def rtbisection(L, r_cut_iL, n_c_part, sum_mu2, x1, x2, xacc, accuracy):
    """dp3m_rtbisection (dp3m.cpp:789-820)."""
    const = accuracy / math.sqrt(2.)
    f1 = real_space_error(L, r_cut_iL, n_c_part, sum_mu2, x1) - const
    f2 = real_space_error(L, r_cut_iL, n_c_part, sum_mu2, x2) - const
    if f1 * f2 >= 0.:
        raise RuntimeError("Root must be bracketed for bisection in dp3m_rtbisection")
    if f1 < 0.:
        dx, rtb = x2 - x1, x1
    else:
        dx, rtb = x1 - x2, x2
    for _ in range(40):
        dx *= 0.5
        xmid = rtb + dx
        fmid = real_space_error(L, r_cut_iL, n_c_part, sum_mu2, xmid) - const
        if fmid <= 0.:
            rtb = xmid
        if abs(dx) < xacc or fmid == 0.:
            return rtb
    raise RuntimeError("Too many bisections in dp3m_rtbisection")


def calculate_accuracy(L, M, cao, r_cut_iL, n_c_part, sum_mu2, accuracy):
    """DipolarTuningAlgorithm::calculate_accuracy (dp3m.cpp:566-598):
    (total, real-space, k-space error, alpha_L)."""
    rs = real_space_error(L, r_cut_iL, n_c_part, sum_mu2, 0.001)
    if math.sqrt(2.) * rs > accuracy:
        alpha_L = rtbisection(L, r_cut_iL, n_c_part, sum_mu2, 1e-4 * L, 5. * L, 1e-4, accuracy)
    else:
        alpha_L = 0.1
    rs = real_space_error(L, r_cut_iL, n_c_part, sum_mu2, alpha_L)
    ks = k_space_error(L, M, cao, n_c_part, sum_mu2, alpha_L)
    return math.hypot(rs, ks), rs, ks, alpha_L


def tune_r_cut(L, M, cao, n_c_part, sum_mu2, accuracy, skin=0.4):
    """r_cut bisection of TuningAlgorithm::get_mc_time (TuningAlgorithm.cpp:
    116-172) for fixed (mesh, cao) on one rank: the smallest r_cut_iL (to
    1e-3) meeting the accuracy.  Returns (r_cut_iL, alpha_L, accuracy) or
    None."""
    lo, hi = 0., (L / 2. - skin) / L
    if calculate_accuracy(L, M, cao, hi, n_c_part, sum_mu2, accuracy)[0] > accuracy:
        return None
    while hi - lo >= 1e-3:
        mid = 0.5 * (lo + hi)
        if calculate_accuracy(L, M, cao, mid, n_c_part, sum_mu2, accuracy)[0] > accuracy:
            lo = mid
        else:
            hi = mid
    acc = calculate_accuracy(L, M, cao, hi, n_c_part, sum_mu2, accuracy)
    return hi, acc[3], acc[0]
