"""ELC (electrostatic layer correction) oracle: a numpy restatement of
ElectrostaticLayerCorrection's far-field terms without dielectric
contrasts (src/core/electrostatics/elc.cpp).  Test infrastructure only:
tests, smoke() and bench.py's CPU baseline use it as the checker; the
product path never imports it.

ELC corrects a 3D-periodic P3M solve of a slab (particles in
0 <= z <= h = L_z - gap) for the periodic images along z.  The correction is
a sum over the 2D reciprocal vectors (p/L_x, q/L_y) within far_cut, each
with weight 1 / expm1(omega L_z), plus the Yeh-Berkowitz dipole term
(elc.cpp:201-265 forces, 267-342 energy) and, without neutralisation, the
removal of P3M's homogeneous background.
"""
import math

import numpy as np

TWO_PI = 2. * math.pi


def tune_far_cut(box_l, gap_size, max_pw_error, space_layer=0., dielectric=False):
    """ElectrostaticLayerCorrection::tune_far_cut (elc.cpp:942-973): the
    smallest multiple of min(1/L_x, 1/L_y) whose pairwise error bound is at
    most maxPWerror; raises RuntimeError past far_cut 50."""
    Lx, Ly, Lz = (float(v) for v in box_l)
    h = Lz - gap_size
    ux, uy = 1. / Lx, 1. / Ly
    step = min(ux, uy)
    lz = h + space_layer if dielectric else Lz
    fc = step
    while True:
        pref = TWO_PI * fc
        s = pref + 2. * (ux + uy)
        den = -math.expm1(-pref * lz)
        n1 = math.exp(pref * (h - lz))
        n2 = math.exp(-pref * (h + lz))
        err = 0.5 / den * (n1 * (s + 1. / (lz - h)) / (lz - h) +
                           n2 * (s + 1. / (lz + h)) / (lz + h))
        fc += step
        if not (err > max_pw_error and fc < 50.):
            break
    if fc >= 50.:
        raise RuntimeError("ELC tuning failed: maxPWerror too small")
    return fc - step


def modes(box_l, far_cut):
    """Reciprocal vectors of the far-field sums in the reference's loop
    order (elc.cpp:830-883): p-modes (p, 0), q-modes (0, q), then (p, q)
    with p, q >= 1.  Each entry is (p, q, omega, pref_di factor 4 or 8)."""
    Lx, Ly = float(box_l[0]), float(box_l[1])
    ux, uy = 1. / Lx, 1. / Ly
    nx = int(math.ceil(far_cut * Lx) + 1.)
    ny = int(math.ceil(far_cut * Ly) + 1.)
    out = []
    p = 1
    while ux * (p - 1) < far_cut and p <= nx:
        out.append((p, 0, TWO_PI * ux * p, 4.))
        p += 1
    q = 1
    while uy * (q - 1) < far_cut and q <= ny:
        out.append((0, q, TWO_PI * uy * q, 4.))
        q += 1
    p = 1
    while ux * (p - 1) < far_cut and p <= nx:
        q = 1
        while (ux * (p - 1)) ** 2 + (uy * (q - 1)) ** 2 < far_cut ** 2 and q <= ny:
            out.append((p, q, TWO_PI * math.sqrt((ux * p) ** 2 + (uy * q) ** 2), 8.))
            q += 1
        p += 1
    return out


def fold_z(pos, Lz):
    z = pos[:, 2] - Lz * np.floor(pos[:, 2] / Lz)
    return np.where(z >= Lz, z - Lz, z)


def space_layer(gap_size, box_h, r_cut):
    """ElectrostaticLayerCorrection::recalc_space_layer (elc.cpp:1032-1054):
    gap/3, capped by gap - r_cut and h/2; RuntimeError when no room is left."""
    sl = gap_size / 3.
    mx = min(gap_size - r_cut, box_h / 2.)
    if sl > mx:
        if mx <= 0.:
            raise RuntimeError("P3M real-space cutoff too large for ELC w/ dielectric contrast")
        sl = mx
    return sl


def elc_forces_energy(pos, q, box_l, prefactor, gap_size, far_cut, neutralize=True,
                      delta_top=0., delta_bot=0., const_pot=False, pot_diff=0.,
                      space_layer=0.):
    """Forces [n, 3] and energy of ElectrostaticLayerCorrection::add_force /
    calc_energy (elc.cpp:830-940) for particles folded into the box along z:
    the dipole term (:201-265, 267-342), the z term of dielectric contrasts
    (:365-498) and the far-field sums (:500-828).

    Per reciprocal vector (k_x, k_y) = 2 pi (p/L_x, q/L_y), omega = |k|:
    a particle carries the eight values q sx sy e^{+-omega z} with
    sx in {sin, cos}(k_x x), sy in {sin, cos}(k_y y) (a missing direction
    contributes cos = 1, sin = 0); the global sums G are scaled by
    -c prefactor pi / (L_x L_y expm1(omega L_z)), c = 4 for p- and q-modes,
    8 for pq-modes.  With dielectric contrasts the sums also take the image
    charges of the boundary layers (z < space_layer: delta_bot q at -z;
    z > h - space_layer: delta_top q at 2h - z) and the infinite image
    series (scaled by c prefactor pi / (L_x L_y)).  The energy is half of the
    dipole and z terms plus the sum over modes of
    sum_i (minus-values . plus-sums + plus-values . minus-sums) / omega."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    q = np.asarray(q, np.float64)
    Lx, Ly, Lz = (float(v) for v in box_l)
    x, y = pos[:, 0], pos[:, 1]
    z = fold_z(pos, Lz)
    h = Lz - gap_size
    sl = float(space_layer)
    dt, db = float(delta_top), float(delta_bot)
    diel = (dt != 0. or db != 0.)
    const_pot = bool(const_pot) and diel
    neutralize = bool(neutralize) and not diel
    bad = (q != 0.) & ((z < 0.) | (z > h))
    if np.any(bad):
        raise RuntimeError("particle entered ELC gap region")
    bot = diel & (z < sl) if diel else np.zeros(len(q), bool)
    top = diel & (z > h - sl) if diel else np.zeros(len(q), bool)
    V = Lx * Ly * Lz
    f = np.zeros((len(q), 3))
    shift = 0.5 * Lz
    # dipole force (elc.cpp:201-265)
    pref = prefactor * 4. * math.pi / V
    g0 = np.sum(q * (z - shift)) + np.sum((db * q * (-z - shift))[bot]) \
        + np.sum((dt * q * (2. * h - z - shift))[top])
    g1 = np.sum(q * z)
    g2 = np.sum(q) + np.sum((db * q)[bot]) + np.sum((dt * q)[top])
    g0 *= pref
    g1 *= pref / h * Lz
    g2 *= pref
    field = g0
    if const_pot:
        field -= pot_diff / h + g1
    f[:, 2] -= field * q
    if not neutralize:
        f[:, 2] += g2 * q * (z - shift)
    # dipole energy (elc.cpp:267-342)
    pe = prefactor * 2. * math.pi / V
    b0 = np.sum(q)
    b1 = np.sum((db * q)[bot]) + np.sum((dt * q)[top])
    b2 = np.sum(q * (z - shift))
    b3 = np.sum((db * q * (-z - shift))[bot]) + np.sum((dt * q * (2. * h - z - shift))[top])
    b4 = np.sum(q * (z - shift) ** 2)
    b5 = np.sum((db * q * (-z - shift) ** 2)[bot]) \
        + np.sum((dt * q * (2. * h - z - shift) ** 2)[top])
    b6 = np.sum(q * z)
    energy = 2. * pe * (b2 * b2 + b2 * b3)
    if not neutralize:
        energy += 2. * pe * (-b0 * b4 - (0.25 - 0.5 / 3.) * (b0 * Lz) ** 2)
    if diel:
        if const_pot:
            energy += pe / h * Lz * b6 * b6
            energy -= 2. * pot_diff / h * b6
        energy += pe * (-(b1 * b4 + b0 * b5) - (1. - 2. / 3.) * b0 * b1 * Lz * Lz)
    # z term (elc.cpp:365-498), dielectric contrasts only
    if diel:
        pz = prefactor * 2. * math.pi / (Lx * Ly)
        delta = dt * db
        if not const_pot:  # (delta = 1 needs const_pot, elc.cpp:1094-1101)
            fdb, fdt, fd = db / (1. - delta), dt / (1. - delta), delta / (1. - delta)
        if const_pot:
            c0 = np.sum((-db * q)[bot]) + np.sum((dt * q)[top])
        else:
            c0 = np.sum(np.where(z < sl, fd * (db + 1.) * q, fdb * (dt + 1.) * q)) \
                - np.sum(np.where(z > h - sl, fd * (dt + 1.) * q, fdt * (db + 1.) * q))
        f[:, 2] += pz * c0 * q
        # energy sums
        if const_pot:
            e0 = np.sum(q)
            e1 = np.sum(q * (z - shift))
            e2 = np.sum((-db * q)[bot]) + np.sum((dt * q)[top])
            e3 = np.sum((-db * q * (-z - shift))[bot]) \
                + np.sum((dt * q * (2. * h - z - shift))[top])
        else:
            # ImageSum (elc.cpp:344-363): b for the bottom, t for the top
            dci = 1. / (1. - delta)

            def isb(qq, zz):
                return qq * dci * (zz - 2. * delta * Lz * dci) - qq * dci * shift

            def ist(qq, zz):
                return qq * dci * (zz + 2. * delta * Lz * dci) - qq * dci * shift
            e0 = np.sum(q)
            e1 = np.sum(q * (z - shift))
            low = z < sl
            high = z > h - sl
            e2 = np.sum(np.where(low, fd * (db + 1.) * q, fdb * (1. + dt) * q))
            e3 = np.sum(q * np.where(low,
                                     isb(db * delta, -(2. * h + z)) + isb(delta, -(2. * h - z)),
                                     isb(db, -z) + isb(delta, -(2. * h - z))))
            e2 -= np.sum(np.where(high, fd * (dt + 1.) * q, fdt * (1. + db) * q))
            e3 -= np.sum(q * np.where(high,
                                      ist(dt * delta, 4. * h - z) + ist(delta, 2. * h + z),
                                      ist(dt, 2. * h - z) + ist(delta, 2. * h + z)))
        energy += -pz * (e1 * e2 - e0 * e3)
    # far-field modes (elc.cpp:500-828, 842-883, 902-937)
    e_modes = 0.
    for (p, qq, omega, c) in modes(box_l, far_cut):
        kx, ky = TWO_PI * p / Lx, TWO_PI * qq / Ly
        sx, cx = np.sin(kx * x), np.cos(kx * x)
        sy, cy = np.sin(ky * y), np.cos(ky * y)
        if p == 0:
            sx, cx = np.zeros_like(x), np.ones_like(x)
        if qq == 0:
            sy, cy = np.zeros_like(y), np.ones_like(y)
        ep = np.exp(omega * z)
        em = 1. / ep
        # values [type][sign]: types ss, sc, cs, cc
        tv = {"ss": sx * sy, "sc": sx * cy, "cs": cx * sy, "cc": cx * cy}
        a = {(t, s): q * v * (ep if s == "p" else em)
             for t, v in tv.items() for s in ("p", "m")}
        pref_di = c * prefactor * math.pi / (Lx * Ly)
        g_pref = -pref_di / math.expm1(omega * Lz)
        G = {k: np.sum(v) for k, v in a.items()}
        if diel:
            delta = dt * db
            fac_elc = 1. / (1. - delta * math.exp(-omega * 2. * h))
            fdb, fdt = db * fac_elc, dt * fac_elc
            fd = fdb * dt
            # boundary-layer images: bottom at -z (values T e^{-+omega z}),
            # top at 2h - z
            eb = np.exp(-omega * z)
            et = np.exp(np.where(top, omega * (2. * h - z), 0.))
            extra = {}
            for t, v in tv.items():
                extra[t, "p"] = np.sum((db * q * v * eb)[bot]) + np.sum((dt * q * v * et)[top])
                extra[t, "m"] = np.sum((db * q * v / eb)[bot]) + np.sum((dt * q * v / et)[top])
            ew_p = np.where(z < sl,
                            (np.exp(omega * (-z - 2. * h)) * db + np.exp(omega * (z - 2. * h))) * fd,
                            (np.exp(-omega * z) + np.exp(omega * (z - 2. * h)) * dt) * fdb)
            ew_m = np.where(z > h - sl,
                            (np.exp(omega * (z - 4. * h)) * dt + np.exp(omega * (-z - 2. * h))) * fd,
                            (np.exp(omega * (z - 2. * h)) + np.exp(omega * (-z - 2. * h)) * db) * fdt)
            for t, v in tv.items():
                G[t, "p"] = g_pref * (G[t, "p"] + extra[t, "p"]) + pref_di * np.sum(q * v * ew_p)
                G[t, "m"] = g_pref * (G[t, "m"] + extra[t, "m"]) + pref_di * np.sum(q * v * ew_m)
        else:
            G = {k: g_pref * v for k, v in G.items()}
        # d/dx of sx: ss -> cs, sc -> cc, cs -> -ss, cc -> -sc (times k_x)
        fx = (a["sc", "m"] * G["cc", "p"] + a["ss", "m"] * G["cs", "p"]
              - a["cc", "m"] * G["sc", "p"] - a["cs", "m"] * G["ss", "p"]
              + a["sc", "p"] * G["cc", "m"] + a["ss", "p"] * G["cs", "m"]
              - a["cc", "p"] * G["sc", "m"] - a["cs", "p"] * G["ss", "m"])
        fy = (a["cs", "m"] * G["cc", "p"] + a["ss", "m"] * G["sc", "p"]
              - a["cc", "m"] * G["cs", "p"] - a["sc", "m"] * G["ss", "p"]
              + a["cs", "p"] * G["cc", "m"] + a["ss", "p"] * G["sc", "m"]
              - a["cc", "p"] * G["cs", "m"] - a["sc", "p"] * G["ss", "m"])
        fz = (a["cc", "m"] * G["cc", "p"] + a["cs", "m"] * G["cs", "p"]
              + a["sc", "m"] * G["sc", "p"] + a["ss", "m"] * G["ss", "p"]
              - a["cc", "p"] * G["cc", "m"] - a["cs", "p"] * G["cs", "m"]
              - a["sc", "p"] * G["sc", "m"] - a["ss", "p"] * G["ss", "m"])
        f[:, 0] += kx / omega * fx
        f[:, 1] += ky / omega * fy
        f[:, 2] += fz
        e_modes += np.sum(sum(a[t, "m"] * G[t, "p"] + a[t, "p"] * G[t, "m"]
                              for t in tv)) / omega
    # the reference halves the whole sum, dipole term included
    # (elc.cpp:885-940)
    return f, 0.5 * (energy + e_modes)


def as_erfc_part(d):
    """exp(d^2) erfc(d) by Abramowitz & Stegun 7.1.26 (AS_erfc_part.hpp)."""
    t = 1. / (1. + 0.3275911 * d)
    return t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 +
                t * (-1.453152027 + t * 1.061405429))))


def pair_force(prefactor, alpha, r_cut, q1q2, d):
    """CoulombP3M::pair_force (p3m.hpp:146-165), rows of d."""
    d = np.atleast_2d(d)
    r = np.sqrt(np.sum(d * d, axis=1))
    ok = (q1q2 != 0.) & (r < r_cut) & (r > 0.)
    rs = np.where(ok, r, 1.)
    ad = alpha * rs
    fac = np.exp(-ad * ad) * (as_erfc_part(ad) / rs + 2. * alpha / math.sqrt(math.pi)) / (rs * rs)
    return np.where(ok, fac * prefactor * q1q2, 0.)[:, None] * d


def pair_energy(prefactor, alpha, r_cut, q1q2, r):
    """CoulombP3M::pair_energy (p3m.hpp:167-180)."""
    r = np.asarray(r, np.float64)
    ok = (q1q2 != 0.) & (r < r_cut) & (r > 0.)
    rs = np.where(ok, r, 1.)
    ad = alpha * rs
    return np.where(ok, prefactor * q1q2 * as_erfc_part(ad) / rs * np.exp(-ad * ad), 0.)


def _mi(d, box):
    box = np.asarray(box, np.float64)
    return d - box * np.round(d / box)


def image_charges(pos, q, box_l, gap_size, space_layer, delta_top, delta_bot):
    """Image charges of the boundary layers (charge_assign IMAGE,
    elc.cpp:1110-1141): delta_bot q at (x, y, -z) for z < space_layer,
    delta_top q at (x, y, 2h - z) for z > h - space_layer.  Returns
    (positions, charges) of the images."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    q = np.asarray(q, np.float64)
    h = float(box_l[2]) - gap_size
    z = fold_z(pos, float(box_l[2]))
    b = (q != 0.) & (z < space_layer)
    t = (q != 0.) & (z > h - space_layer)
    pb = pos[b].copy()
    pb[:, 2] = -z[b]
    pt = pos[t].copy()
    pt[:, 2] = 2. * h - z[t]
    return np.concatenate([pb, pt]), np.concatenate([delta_bot * q[b], delta_top * q[t]])


def layer_corrections(pos, q, box_l, prefactor, alpha, r_cut, gap_size, space_layer,
                      delta_top, delta_bot):
    """Short-range image terms of dielectric contrasts, all pairs and self:
    add_pair_force_corrections / pair_energy_correction (elc.hpp:250-296)
    for every pair closer than r_cut, and dielectric_layers_self_forces /
    _self_energy (elc.hpp:140-171).  Returns (pair forces, pair energy,
    self forces, self energy)."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    q = np.asarray(q, np.float64)
    Lz = float(box_l[2])
    h = Lz - gap_size
    n = len(q)
    zf = fold_z(pos, Lz)
    P = pos.copy()
    P[:, 2] = zf

    def contributions(i, j):
        """image terms of particle i against particle j: list of (q_eff, d)"""
        out = []
        if P[i, 2] < space_layer:
            img = np.array([P[i, 0], P[i, 1], -P[i, 2]])
            out.append((delta_bot * q[i] * q[j], _mi(P[j] - img, box_l)))
        if P[i, 2] > h - space_layer:
            img = np.array([P[i, 0], P[i, 1], 2. * h - P[i, 2]])
            out.append((delta_top * q[i] * q[j], _mi(P[j] - img, box_l)))
        return out

    f_pair = np.zeros((n, 3))
    e_pair = 0.
    for i in range(n):
        for j in range(i + 1, n):
            if q[i] == 0. or q[j] == 0.:
                continue
            if np.linalg.norm(_mi(P[i] - P[j], box_l)) >= r_cut:
                continue
            e = 0.
            for qe, d in contributions(i, j):
                f_pair[i] += pair_force(prefactor, alpha, r_cut, qe, d)[0]
                e += pair_energy(prefactor, alpha, r_cut, qe, np.linalg.norm(d))
            for qe, d in contributions(j, i):
                f_pair[j] += pair_force(prefactor, alpha, r_cut, qe, d)[0]
                e += pair_energy(prefactor, alpha, r_cut, qe, np.linalg.norm(d))
            e_pair += e / 2.
    f_self = np.zeros((n, 3))
    e_self = 0.
    for i in range(n):
        for qe, d in contributions(i, i):
            f_self[i] += pair_force(prefactor, alpha, r_cut, qe, d)[0]
            e_self += pair_energy(prefactor, alpha, r_cut, qe, np.linalg.norm(d))
    return f_pair, e_pair, f_self, e_self


def elc_p3m_total(pos, q, box_l, p3m, prefactor, alpha, r_cut, gap_size, far_cut,
                  delta_top=0., delta_bot=0., const_pot=False, pot_diff=0., neutralize=True,
                  space_layer=None, real_space=None):
    """Total Coulomb forces and energy of P3M + ELC as the reference computes
    them (ElectrostaticLayerCorrection::add_long_range_forces /
    long_range_energy, elc.cpp:1189-1256, plus the short-range loop with the
    pair corrections of elc.hpp:250-296).  p3m is an OracleP3M of the box;
    real_space(pos, q) -> (energy, forces) gives the plain pair terms.
    With dielectric contrasts the mesh holds the real and the image charges
    and the energy is
    0.5 E(real) + 0.5 E_self(layers) + 0.5 E(real + images) - 0.5 E(images)."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    q = np.asarray(q, np.float64)
    n = len(q)
    Lz = float(box_l[2])
    h = Lz - gap_size
    diel = (delta_top != 0. or delta_bot != 0.)
    if space_layer is None:
        space_layer = globals()["space_layer"](gap_size, h, r_cut) if diel else 0.
    sl = space_layer
    P = pos.copy()
    P[:, 2] = fold_z(pos, Lz)
    e_sr, f = real_space(P, q)
    f = f.copy()
    energy = e_sr
    if diel:
        ip, iq = image_charges(P, q, box_l, gap_size, sl, delta_top, delta_bot)
        ipf = ip.copy()
        ipf[:, 2] = fold_z(ip, Lz)          # the mesh is periodic
        both_p = np.concatenate([P, ipf])
        both_q = np.concatenate([q, iq])
        f += p3m.forces(both_p, both_q)[:n]
        fp, ep_, fs, es = layer_corrections(P, q, box_l, prefactor, alpha, r_cut, gap_size, sl,
                                            delta_top, delta_bot)
        f += fp + fs
        energy += ep_
        energy += 0.5 * p3m.energy(P, q) + 0.5 * es + 0.5 * p3m.energy(both_p, both_q) \
            - 0.5 * p3m.energy(ipf, iq)
    else:
        f += p3m.forces(P, q)
        energy += p3m.energy(P, q)
    fe, ee = elc_forces_energy(P, q, box_l, prefactor, gap_size, far_cut, neutralize,
                               delta_top, delta_bot, const_pot, pot_diff, sl)
    return f + fe, energy + ee
