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


def elc_forces_energy(pos, q, box_l, prefactor, gap_size, far_cut, neutralize=True):
    """Forces [n, 3] and energy of the ELC far-field correction for a system
    without dielectric contrasts (elc.cpp:201-265, 267-342, 500-640,
    641-828, 830-940).  Positions are folded into the box along z first.

    Per reciprocal vector (k_x, k_y) = 2 pi (p/L_x, q/L_y), omega = |k|:
    a particle carries the eight values q sx sy e^{+-omega z} with
    sx in {sin, cos}(k_x x), sy in {sin, cos}(k_y y) (a missing direction
    contributes cos = 1, sin = 0); the global sums G are scaled by
    -c prefactor pi / (L_x L_y expm1(omega L_z)), c = 4 for p- and q-modes,
    8 for pq-modes.  The energy is half the sum over modes of
    sum_i (minus-values . plus-sums + plus-values . minus-sums) / omega, and
    half the dipole term 2 * 2 pi prefactor / V (sum_i q_i (z_i - L_z/2))^2."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    q = np.asarray(q, np.float64)
    Lx, Ly, Lz = (float(v) for v in box_l)
    x, y = pos[:, 0], pos[:, 1]
    z = fold_z(pos, Lz)
    h = Lz - gap_size
    bad = (q != 0.) & ((z < 0.) | (z > h))
    if np.any(bad):
        raise RuntimeError("particle entered ELC gap region")
    V = Lx * Ly * Lz
    f = np.zeros((len(q), 3))
    # dipole term and background (elc.cpp:201-265, 267-342)
    shift = 0.5 * Lz
    pref = prefactor * 4. * math.pi / V
    m0 = pref * np.sum(q * (z - shift))
    qt = np.sum(q)
    f[:, 2] -= m0 * q
    if not neutralize:
        f[:, 2] += pref * qt * q * (z - shift)
    pe = prefactor * 2. * math.pi / V
    s2 = np.sum(q * (z - shift))
    energy = 2. * pe * s2 * s2
    if not neutralize:
        s4 = np.sum(q * (z - shift) ** 2)
        energy += 2. * pe * (-qt * s4 - (0.25 - 0.5 / 3.) * (qt * Lz) ** 2)
    # far-field modes
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
        g_pref = -c * prefactor * math.pi / (Lx * Ly) / math.expm1(omega * Lz)
        G = {k: g_pref * np.sum(v) for k, v in a.items()}
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
