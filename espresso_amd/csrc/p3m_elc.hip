// ELC far-field terms on the device; see p3m_elc.hpp.
//
// Layout of one call (all FP64, deterministic reductions):
//   k_elc_prep     fold positions into the box, gap check, per-block sums of
//                  the dipole and z terms (moments of the real charges and of
//                  the layer images)
//   k_elc_sums     per (particle block, tile of modes): the eight scaled sums
//                  q {sin,cos}(k_x x) {sin,cos}(k_y y) e^{+-omega z}, with
//                  dielectric contrasts also those of the layer images and of
//                  the image series
//   k_elc_reduce   per mode: sum of the block partials in block order; one
//                  more block reduces the dipole sums
//   k_elc_forces   per particle: all modes against the global sums (force
//                  and energy), plus the dipole, background and z forces
#include "p3m_elc.hpp"
#include "p3m_batch.hpp"
#include "p3m_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>

namespace espresso_amd {
namespace {
using namespace dev;

constexpr int kElcBlock = 256;
constexpr int kModeTile = 4;       // modes per k_elc_sums block
constexpr int kMaxSumBlocks = 512; // particle blocks of k_elc_sums at most
constexpr int kMaxModes = 65536;   // reciprocal vectors at most (grid y, partials)

struct ElcGeo {
  double L[3];
  double h;     // L_z - gap: charged particles must stay in [0, h]
  double shift; // L_z / 2
  double sl;    // space layer (0 without dielectric contrasts)
  double dt, db;
  int diel, const_pot;
};

constexpr int kDipSums = 10; // prep sums per block (see k_elc_prep)

__device__ __forceinline__ double fold_into(double x, double L) {
  double xf = x - L * floor(x / L);
  if (xf >= L) xf -= L;
  return xf;
}

/** sin and cos of kx x and ky y of one particle over a sequence of modes:
 *  p = 1 (q = 1) from one sincos per particle, p = p_prev + 1 by angle
 *  addition, p = p_prev reused, any other p by sincos (kx x); p = 0 (q = 0)
 *  is the missing direction (cos = 1, sin = 0).  Angle addition over p
 *  steps rounds to about p ulps. */
struct ElcTrig {
  double x, y, s1x, c1x, s1y, c1y, sx, cx, sy, cy;
  int px, qy;
  __device__ __forceinline__ ElcTrig(double4 const &p, ElcGeo const &g) : x(p.x), y(p.y) {
    double const two_pi = 2. * 3.14159265358979323846;
    sincos(two_pi * (1. / g.L[0]) * x, &s1x, &c1x);
    sincos(two_pi * (1. / g.L[1]) * y, &s1y, &c1y);
    sx = sy = 0.;
    cx = cy = 1.;
    px = qy = 0;
  }
  __device__ __forceinline__ void get(ElcMode const &md, double &osx, double &ocx, double &osy,
                                      double &ocy) {
    if (md.p == 0) {
      osx = 0.;
      ocx = 1.;
    } else {
      if (md.p != px) {
        if (md.p == 1) {
          sx = s1x;
          cx = c1x;
        } else if (md.p == px + 1) {
          double const s = sx * c1x + cx * s1x, c = cx * c1x - sx * s1x;
          sx = s;
          cx = c;
        } else {
          sincos(md.kx * x, &sx, &cx);
        }
        px = md.p;
      }
      osx = sx;
      ocx = cx;
    }
    if (md.q == 0) {
      osy = 0.;
      ocy = 1.;
    } else {
      if (md.q != qy) {
        if (md.q == 1) {
          sy = s1y;
          cy = c1y;
        } else if (md.q == qy + 1) {
          double const s = sy * c1y + cy * s1y, c = cy * c1y - sy * s1y;
          sy = s;
          cy = c;
        } else {
          sincos(md.ky * y, &sy, &cy);
        }
        qy = md.q;
      }
      osy = sy;
      ocy = cy;
    }
  }
};

/** The eight values of particle (x, y, z, q) for mode md, in the order
 *  [ss, sc, cs, cc] x [e^{+omega z}, e^{-omega z}] (v[0..3] plus, v[4..7]
 *  minus).  kx = 0 or ky = 0 stands for the missing direction (cos = 1,
 *  sin = 0), which turns the pq-mode values into those of the p- and
 *  q-modes (elc.cpp:500-600, 641-770). */
__device__ __forceinline__ void mode_values(ElcMode const &md, double4 const &p, ElcTrig &tr,
                                            double v[8]) {
  double sx, cx, sy, cy;
  tr.get(md, sx, cx, sy, cy);
  double const ep = exp(md.omega * p.z);
  double const qp = p.w * ep, qm = p.w / ep;
  double const ss = sx * sy, sc = sx * cy, cs = cx * sy, cc = cx * cy;
  v[0] = ss * qp;
  v[1] = sc * qp;
  v[2] = cs * qp;
  v[3] = cc * qp;
  v[4] = ss * qm;
  v[5] = sc * qm;
  v[6] = cs * qm;
  v[7] = cc * qm;
}

__global__ void k_elc_clear(int *flag) { *flag = 0; }

/** Per particle: folded position (spos), gap check, and the block sums
 *  v0 = q (z - s), v1 = q, v2 = q (z - s)^2, v3 = q z (s = L_z / 2); with
 *  dielectric contrasts v4 = sum of layer image charges q', v5 = q' (z' - s),
 *  v6 = q' (z' - s)^2 (z' = -z below, 2h - z above; elc.cpp:201-342), v7 the
 *  z-force sum (:441-498), v8 and v9 the z-energy sums (:365-439). */
__global__ __launch_bounds__(kElcBlock) void k_elc_prep(int64_t n, double const *__restrict__ pos,
                                                        double const *__restrict__ q, ElcGeo g,
                                                        double4 *__restrict__ spos,
                                                        double *partials, int *err) {
  __shared__ double red[kDipSums * 16];
  int64_t const i = (int64_t)blockIdx.x * kElcBlock + threadIdx.x;
  double v[kDipSums];
#pragma unroll
  for (int k = 0; k < kDipSums; ++k) v[k] = 0.;
  if (i < n) {
    double const qi = q[i];
    double4 p{fold_into(pos[3 * i], g.L[0]), fold_into(pos[3 * i + 1], g.L[1]),
              fold_into(pos[3 * i + 2], g.L[2]), qi};
    if (qi != 0. && !(p.z >= 0. && p.z <= g.h)) atomicOr(err, 1);
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) {
      atomicOr(err, 2);
      p = double4{0., 0., 0., 0.};
    }
    spos[i] = p;
    double const z = p.z, s = g.shift, h = g.h;
    double const dz = z - s;
    v[0] = qi * dz;
    v[1] = qi;
    v[2] = qi * dz * dz;
    v[3] = qi * z;
    if (g.diel) {
      bool const bot = z < g.sl, top = z > h - g.sl;
      if (bot) {
        double const qe = g.db * qi, zi = -z - s;
        v[4] += qe;
        v[5] += qe * zi;
        v[6] += qe * zi * zi;
      }
      if (top) {
        double const qe = g.dt * qi, zi = 2. * h - z - s;
        v[4] += qe;
        v[5] += qe * zi;
        v[6] += qe * zi * zi;
      }
      double const delta = g.dt * g.db;
      double const fdb = g.db / (1. - delta), fdt = g.dt / (1. - delta),
                   fd = delta / (1. - delta);
      if (g.const_pot) {
        // z force: counter the 2 pi |z| term of P3M (elc.cpp:450-461)
        if (bot) v[7] -= g.db * qi;
        if (top) v[7] += g.dt * qi;
        if (bot) {
          v[8] -= g.db * qi;
          v[9] -= g.db * qi * (-z - s);
        }
        if (top) {
          v[8] += g.dt * qi;
          v[9] += g.dt * qi * (2. * h - z - s);
        }
      } else {
        v[7] += bot ? fd * (g.db + 1.) * qi : fdb * (g.dt + 1.) * qi;
        v[7] -= top ? fd * (g.dt + 1.) * qi : fdt * (g.db + 1.) * qi;
        // ImageSum b / t (elc.cpp:344-363) with delta = dt db
        double const dci = 1. / (1. - delta), Lz = g.L[2];
        auto isb = [&](double qq, double zz) {
          return qq * dci * (zz - 2. * delta * Lz * dci) - qq * dci * s;
        };
        auto ist = [&](double qq, double zz) {
          return qq * dci * (zz + 2. * delta * Lz * dci) - qq * dci * s;
        };
        v[8] += bot ? fd * (g.db + 1.) * qi : fdb * (1. + g.dt) * qi;
        v[9] += qi * (bot ? isb(g.db * delta, -(2. * h + z)) + isb(delta, -(2. * h - z))
                          : isb(g.db, -z) + isb(delta, -(2. * h - z)));
        v[8] -= top ? fd * (g.dt + 1.) * qi : fdt * (1. + g.db) * qi;
        v[9] -= qi * (top ? ist(g.dt * delta, 4. * h - z) + ist(delta, 2. * h + z)
                          : ist(g.dt, 2. * h - z) + ist(delta, 2. * h + z));
      }
    }
  }
  block_sum<kDipSums>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < kDipSums; ++k) partials[blockIdx.x * kDipSums + k] = v[k];
}

/** The per-particle contributions of one mode to its eight sums, scaled:
 *  pref (e^{+-omega z} values of the charge and, with contrasts, of its
 *  layer images) and pref_di (the image series, elc.cpp:520-597). */
template <bool DIEL>
__device__ __forceinline__ void mode_sums(ElcMode const &md, double4 const &p, ElcGeo const &g,
                                          double Lz, ElcTrig &tr, double acc[8]) {
  double sx, cx, sy, cy;
  tr.get(md, sx, cx, sy, cy);
  double const w = md.omega, z = p.z, q = p.w;
  double const ep = exp(w * z);
  double cp = md.pref * q * ep, cm = md.pref * q / ep;
  if constexpr (DIEL) {
    double const h = g.h;
    double const pref_di = md.pref_di;
    bool const bot = z < g.sl, top = z > h - g.sl;
    if (bot && g.db != 0.) {
      // image at -z: e^{+omega z'} = e^{-omega z}
      cp += md.pref * g.db * q / ep;
      cm += md.pref * g.db * q * ep;
    }
    if (top && g.dt != 0.) {
      // image at 2h - z, with the scale folded into the exponent so that
      // e^{omega (2h - z)} cannot overflow: pref e^{omega z'} =
      // pref_di e^{omega (z' - L_z)} / expm1(-omega L_z)
      double const em1 = expm1(-w * Lz);
      cp += pref_di * g.dt * q * exp(w * (2. * h - z - Lz)) / em1;
      cm += md.pref * g.dt * q * exp(-w * (2. * h - z));
    }
    double const delta = g.dt * g.db;
    double const fac = 1. / (1. - delta * exp(-w * 2. * h));
    double const fdb = g.db * fac, fdt = g.dt * fac, fd = fdb * g.dt;
    double const ewp = bot ? (exp(w * (-z - 2. * h)) * g.db + exp(w * (z - 2. * h))) * fd
                           : (exp(-w * z) + exp(w * (z - 2. * h)) * g.dt) * fdb;
    double const ewm = top ? (exp(w * (z - 4. * h)) * g.dt + exp(w * (-z - 2. * h))) * fd
                           : (exp(w * (z - 2. * h)) + exp(w * (-z - 2. * h)) * g.db) * fdt;
    cp += pref_di * q * ewp;
    cm += pref_di * q * ewm;
  }
  double const ss = sx * sy, sc = sx * cy, cs = cx * sy, cc = cx * cy;
  acc[0] += ss * cp;
  acc[1] += sc * cp;
  acc[2] += cs * cp;
  acc[3] += cc * cp;
  acc[4] += ss * cm;
  acc[5] += sc * cm;
  acc[6] += cs * cm;
  acc[7] += cc * cm;
}

/** blockIdx.x: particle block (ppb particles, ppb / kElcBlock per thread);
 *  blockIdx.y: tile of kModeTile modes.  partials[(blk * nmodes + m) * 8 + k]. */
template <bool DIEL>
__global__ __launch_bounds__(kElcBlock) void k_elc_sums(int64_t n, int64_t ppb,
                                                        double4 const *__restrict__ spos,
                                                        ElcMode const *__restrict__ modes,
                                                        int nmodes, ElcGeo g,
                                                        double *partials) {
  __shared__ double red[kModeTile * 8 * (kElcBlock / 64)];
  int const m0 = blockIdx.y * kModeTile;
  int const nm = min(kModeTile, nmodes - m0);
  double acc[kModeTile][8];
#pragma unroll
  for (int t = 0; t < kModeTile; ++t)
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[t][k] = 0.;
  int64_t const beg = (int64_t)blockIdx.x * ppb;
  int64_t const end = min(n, beg + ppb);
  for (int64_t i = beg + threadIdx.x; i < end; i += kElcBlock) {
    double4 const p = spos[i];
    if (p.w == 0.) continue;
    ElcTrig tr(p, g);
#pragma unroll
    for (int t = 0; t < kModeTile; ++t)
      if (t < nm) mode_sums<DIEL>(modes[m0 + t], p, g, g.L[2], tr, acc[t]);
  }
  int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < kModeTile; ++t)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double const s = wave_sum(acc[t][k]);
      if (lane == 0) red[(t * 8 + k) * (kElcBlock / 64) + wave] = s;
    }
  __syncthreads();
  int const t = threadIdx.x >> 3, k = threadIdx.x & 7;
  if (threadIdx.x < kModeTile * 8 && t < nm) {
    double s = 0.;
    for (int w = 0; w < kElcBlock / 64; ++w) s += red[(t * 8 + k) * (kElcBlock / 64) + w];
    partials[((int64_t)blockIdx.x * nmodes + m0 + t) * 8 + k] = s;
  }
}

/** Blocks 0 .. nmodes-1: G[m][k] = sum over the nsb particle blocks (in
 *  block order); block nmodes: dip[0 .. kDipSums) = sums of the prep
 *  partials. */
__global__ __launch_bounds__(kElcBlock) void k_elc_reduce(double const *__restrict__ partials,
                                                          int nsb, int nmodes,
                                                          double const *__restrict__ prep_part,
                                                          int nprep, double *G, double *dip) {
  __shared__ double red[kDipSums * 16];
  int const m = blockIdx.x;
  double v[kDipSums];
#pragma unroll
  for (int k = 0; k < kDipSums; ++k) v[k] = 0.;
  if (m < nmodes) {
    for (int b = threadIdx.x; b < nsb; b += kElcBlock)
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] += partials[((int64_t)b * nmodes + m) * 8 + k];
    block_sum<kDipSums>(v, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 8; ++k) G[m * 8 + k] = v[k];
  } else {
    for (int b = threadIdx.x; b < nprep; b += kElcBlock)
#pragma unroll
      for (int k = 0; k < kDipSums; ++k) v[k] += prep_part[b * kDipSums + k];
    block_sum<kDipSums>(v, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < kDipSums; ++k) dip[k] = v[k];
  }
}

/** Forces and energy of every particle against all modes
 *  (add_PoQ_force / add_PQ_force, PoQ_energy / PQ_energy,
 *  elc.cpp:601-640, 771-828) plus the dipole force and, without
 *  neutralisation, the removal of P3M's background force (:201-265). */
struct ElcForceConst {
  double dip_pref;  // 4 pi prefactor / V
  double z_pref;    // 2 pi prefactor / (L_x L_y)
  double pot_diff;
  int neutralize;
};

template <bool EN>
__global__ __launch_bounds__(kElcBlock) void k_elc_forces(
    int64_t n, double4 const *__restrict__ spos, ElcMode const *__restrict__ modes, int nmodes,
    double const *__restrict__ G, double const *__restrict__ dip, ElcGeo g, ElcForceConst c,
    double *__restrict__ force, double *partials) {
  __shared__ double red[16];
  int64_t const i = (int64_t)blockIdx.x * kElcBlock + threadIdx.x;
  double e = 0.;
  if (i < n) {
    double4 const p = spos[i];
    if (p.w != 0.) {
      double fx = 0., fy = 0., fz = 0.;
      ElcTrig tr(p, g);
      for (int m = 0; m < nmodes; ++m) {
        ElcMode const md = modes[m];
        double a[8];
        mode_values(md, p, tr, a);
        double const *g = G + m * 8;
        // plus values against minus sums and minus values against plus sums
        double const gp0 = g[0], gp1 = g[1], gp2 = g[2], gp3 = g[3];
        double const gm0 = g[4], gm1 = g[5], gm2 = g[6], gm3 = g[7];
        double const inv_w = md.inv_omega;
        if (md.kx != 0.) {
          double const t = a[5] * gp3 + a[4] * gp2 - a[7] * gp1 - a[6] * gp0 +
                           a[1] * gm3 + a[0] * gm2 - a[3] * gm1 - a[2] * gm0;
          fx += md.kx * inv_w * t;
        }
        if (md.ky != 0.) {
          double const t = a[6] * gp3 + a[4] * gp1 - a[7] * gp2 - a[5] * gp0 +
                           a[2] * gm3 + a[0] * gm1 - a[3] * gm2 - a[1] * gm0;
          fy += md.ky * inv_w * t;
        }
        fz += a[7] * gp3 + a[6] * gp2 + a[5] * gp1 + a[4] * gp0 - a[3] * gm3 - a[2] * gm2 -
              a[1] * gm1 - a[0] * gm0;
        if constexpr (EN)
          e += (a[7] * gp3 + a[6] * gp2 + a[5] * gp1 + a[4] * gp0 + a[3] * gm3 + a[2] * gm2 +
                a[1] * gm1 + a[0] * gm0) *
               inv_w;
      }
      // dipole term (field of the total z moment with the layer images),
      // constant potential, background removal (elc.cpp:201-265) and the
      // z term of dielectric contrasts (:441-498)
      double field = c.dip_pref * (dip[0] + dip[5]);
      if (g.const_pot) field -= c.pot_diff / g.h + c.dip_pref / g.h * g.L[2] * dip[3];
      fz -= field * p.w;
      if (!c.neutralize) fz += c.dip_pref * (dip[1] + dip[4]) * p.w * (p.z - g.shift);
      if (g.diel) fz += c.z_pref * dip[7] * p.w;
      if (force) {
        force[3 * i] += fx;
        force[3 * i + 1] += fy;
        force[3 * i + 2] += fz;
      }
    }
  }
  if constexpr (EN) {
    double v[1] = {e};
    block_sum<1>(v, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
  }
}

/** Real particles (folded) and their bottom and top images: rows [0, n),
 *  [n, 2n), [2n, 3n) (charge_assign, ChargeProtocol BOTH / IMAGE,
 *  elc.cpp:1110-1141).  An image outside its layer, of an uncharged
 *  particle or of a zero contrast gets charge 0.  The bottom image at -z is
 *  stored at L_z - z (the mesh is periodic). */
__global__ __launch_bounds__(kElcBlock) void k_elc_images(int64_t n,
                                                          double4 const *__restrict__ spos,
                                                          ElcGeo g, double *__restrict__ pos3,
                                                          double *__restrict__ q_both,
                                                          double *__restrict__ q_img) {
  int64_t const i = (int64_t)blockIdx.x * kElcBlock + threadIdx.x;
  if (i >= n) return;
  double4 const p = spos[i];
  double const z = p.z, h = g.h;
  bool const bot = p.w != 0. && z < g.sl;
  bool const top = p.w != 0. && z > h - g.sl;
  double zb = g.L[2] - z;
  if (zb >= g.L[2]) zb -= g.L[2];
  int64_t const ib = n + i, it = 2 * n + i;
  pos3[3 * i] = p.x;
  pos3[3 * i + 1] = p.y;
  pos3[3 * i + 2] = z;
  pos3[3 * ib] = p.x;
  pos3[3 * ib + 1] = p.y;
  pos3[3 * ib + 2] = zb;
  pos3[3 * it] = p.x;
  pos3[3 * it + 1] = p.y;
  pos3[3 * it + 2] = 2. * h - z;
  double const qb = bot ? g.db * p.w : 0., qt = top ? g.dt * p.w : 0.;
  q_both[i] = p.w;
  q_img[i] = 0.;
  q_both[ib] = q_img[ib] = qb;
  q_both[it] = q_img[it] = qt;
}

__global__ __launch_bounds__(kElcBlock) void k_elc_add_rows(int64_t n3,
                                                            double const *__restrict__ f3,
                                                            double *__restrict__ force) {
  int64_t const i = (int64_t)blockIdx.x * kElcBlock + threadIdx.x;
  if (i < n3) force[i] += f3[i];
}

__global__ __launch_bounds__(kElcBlock) void k_elc_zero(int64_t n, double *__restrict__ a) {
  int64_t const i = (int64_t)blockIdx.x * kElcBlock + threadIdx.x;
  if (i < n) a[i] = 0.;
}

/** CoulombP3M::pair_force / pair_energy (p3m.hpp:146-180, A&S erfc) of a
 *  pair term q1q2 at displacement d. */
__device__ __forceinline__ double elc_pair(double prefactor, double alpha, double r_cut,
                                           double q1q2, double const d[3], double f[3]) {
  double const r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (q1q2 == 0. || !(r < r_cut) || !(r > 0.)) return 0.;
  double const ad = alpha * r;
  double const t = 1. / (1. + 0.3275911 * ad);
  double const erfc_part =
      t * (0.254829592 +
           t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
  double const ex = exp(-ad * ad);
  double const erfc_ri = erfc_part / r;
  double const fac =
      ex * (erfc_ri + 2. * alpha * std::numbers::inv_sqrtpi) / (r * r) * prefactor * q1q2;
  for (int k = 0; k < 3; ++k) f[k] += fac * d[k];
  return prefactor * q1q2 * erfc_ri * ex;
}

/** dielectric_layers_self_forces / _self_energy (elc.hpp:140-171): each
 *  particle in a layer against its own image, d = r - r' (minimum image). */
template <bool EN>
__global__ __launch_bounds__(kElcBlock) void k_elc_self(int64_t n,
                                                        double4 const *__restrict__ spos,
                                                        ElcGeo g, double prefactor, double alpha,
                                                        double r_cut, double *__restrict__ force,
                                                        double *partials) {
  __shared__ double red[16];
  int64_t const i = (int64_t)blockIdx.x * kElcBlock + threadIdx.x;
  double e = 0.;
  if (i < n) {
    double4 const p = spos[i];
    double f[3] = {0., 0., 0.};
    double const Lz = g.L[2];
    if (p.w != 0. && p.z < g.sl) {
      double d[3] = {0., 0., 2. * p.z};
      d[2] -= Lz * rint(d[2] / Lz);
      e += elc_pair(prefactor, alpha, r_cut, g.db * p.w * p.w, d, f);
    }
    if (p.w != 0. && p.z > g.h - g.sl) {
      double d[3] = {0., 0., p.z - (2. * g.h - p.z)};
      d[2] -= Lz * rint(d[2] / Lz);
      e += elc_pair(prefactor, alpha, r_cut, g.dt * p.w * p.w, d, f);
    }
    if (force) {
      force[3 * i] += f[0];
      force[3 * i + 1] += f[1];
      force[3 * i + 2] += f[2];
    }
  }
  if constexpr (EN) {
    double v[1] = {e};
    block_sum<1>(v, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
  }
}

template <typename T> void elc_grow(T *&p, size_t &cap, size_t need, char const *what) {
  if (need <= cap) return;
  if (p) (void)hipFree(p);
  p = nullptr;
  size_t const n = need + need / 4 + 64;
  hip_check(hipMalloc(&p, sizeof(T) * n), what);
  cap = n;
}
} // namespace

std::vector<ElcMode> elc_modes(ElcParams const &p) {
  double const two_pi = 2. * std::numbers::pi;
  double const ux = 1. / p.box[0], uy = 1. / p.box[1];
  double const Lz = p.box[2];
  auto const nx = static_cast<std::size_t>(std::ceil(p.far_cut * p.box[0]) + 1.);
  auto const ny = static_cast<std::size_t>(std::ceil(p.far_cut * p.box[1]) + 1.);
  double const area_pi = p.prefactor * std::numbers::pi * ux * uy;
  std::vector<ElcMode> out;
  auto add = [&](double kx, double ky, double omega, double c, std::size_t i, std::size_t j) {
    out.push_back(ElcMode{kx, ky, omega, -c * area_pi / std::expm1(omega * Lz), c * area_pi,
                          1. / omega, static_cast<int>(i), static_cast<int>(j)});
  };
  for (std::size_t i = 1; ux * static_cast<double>(i - 1) < p.far_cut && i <= nx; ++i) {
    double const om = two_pi * ux * static_cast<double>(i);
    add(om, 0., om, 4., i, 0);
  }
  for (std::size_t j = 1; uy * static_cast<double>(j - 1) < p.far_cut && j <= ny; ++j) {
    double const om = two_pi * uy * static_cast<double>(j);
    add(0., om, om, 4., 0, j);
  }
  double const fc2 = p.far_cut * p.far_cut;
  for (std::size_t i = 1; ux * static_cast<double>(i - 1) < p.far_cut && i <= nx; ++i)
    for (std::size_t j = 1;
         (ux * static_cast<double>(i - 1)) * (ux * static_cast<double>(i - 1)) +
                 (uy * static_cast<double>(j - 1)) * (uy * static_cast<double>(j - 1)) <
             fc2 &&
         j <= ny;
         ++j) {
      double const kx = two_pi * ux * static_cast<double>(i);
      double const ky = two_pi * uy * static_cast<double>(j);
      add(kx, ky,
          two_pi * std::sqrt((ux * static_cast<double>(i)) * (ux * static_cast<double>(i)) +
                             (uy * static_cast<double>(j)) * (uy * static_cast<double>(j))),
          8., i, j);
    }
  return out;
}

double elc_tune_far_cut(double const box[3], double gap_size, double max_pw_error,
                        bool dielectric, double space_layer) {
  constexpr double maximal_far_cut = 50.;
  double const ux = 1. / box[0], uy = 1. / box[1];
  double const step = std::min(ux, uy);
  double const h = box[2] - gap_size;
  double const lz = dielectric ? h + space_layer : box[2];
  double fc = step;
  double err;
  do {
    double const pref = 2. * std::numbers::pi * fc;
    double const sum = pref + 2. * (ux + uy);
    double const den = -std::expm1(-pref * lz);
    double const num1 = std::exp(pref * (h - lz));
    double const num2 = std::exp(-pref * (h + lz));
    err = 0.5 / den *
          (num1 * (sum + 1. / (lz - h)) / (lz - h) + num2 * (sum + 1. / (lz + h)) / (lz + h));
    fc += step;
  } while (err > max_pw_error && fc < maximal_far_cut);
  if (fc >= maximal_far_cut) throw std::runtime_error("ELC tuning failed: maxPWerror too small");
  return fc - step;
}

double elc_space_layer(double gap_size, double box_h, double r_cut) {
  double sl = gap_size / 3.;
  double const mx = std::min(gap_size - r_cut, box_h / 2.);
  if (sl > mx) {
    if (mx <= 0.)
      throw std::runtime_error("P3M real-space cutoff too large for ELC w/ dielectric contrast");
    sl = mx;
  }
  return sl;
}

struct ElcSolver::Impl {
  int device = 0;
  ElcParams cached{};
  bool have_modes = false;
  std::vector<ElcMode> modes;
  ElcMode *d_modes = nullptr;
  double4 *spos = nullptr;
  double *prep_part = nullptr, *sum_part = nullptr, *G = nullptr, *dip = nullptr,
         *e_part = nullptr;
  int *err = nullptr;
  size_t cap_modes = 0, cap_spos = 0, cap_prep = 0, cap_sum = 0, cap_G = 0, cap_e = 0;
  // image arrays of dielectric contrasts
  double4 *ispos = nullptr;
  double *pos3 = nullptr, *q_both = nullptr, *q_img = nullptr, *f3 = nullptr,
         *self_part = nullptr;
  size_t cap_ispos = 0, cap_pos3 = 0, cap_qb = 0, cap_qi = 0, cap_f3 = 0, cap_self = 0;
};

namespace {
ElcGeo make_geo(ElcParams const &p) {
  ElcGeo g;
  for (int d = 0; d < 3; ++d) g.L[d] = p.box[d];
  g.h = p.box[2] - p.gap_size;
  g.shift = 0.5 * p.box[2];
  g.diel = p.dielectric() ? 1 : 0;
  g.sl = g.diel ? p.space_layer : 0.;
  g.dt = p.delta_top;
  g.db = p.delta_bot;
  g.const_pot = (p.const_pot && g.diel) ? 1 : 0;
  return g;
}
} // namespace

ElcSolver::ElcSolver(int device) : m(new Impl) {
  m->device = device;
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipMalloc(&m->dip, sizeof(double) * kDipSums), "hipMalloc elc dipole");
  hip_check(hipMalloc(&m->err, sizeof(int)), "hipMalloc elc flag");
}

ElcSolver::~ElcSolver() {
  (void)hipSetDevice(m->device);
  for (void *p : {static_cast<void *>(m->d_modes), static_cast<void *>(m->spos),
                  static_cast<void *>(m->prep_part), static_cast<void *>(m->sum_part),
                  static_cast<void *>(m->G), static_cast<void *>(m->dip),
                  static_cast<void *>(m->e_part), static_cast<void *>(m->err),
                  static_cast<void *>(m->ispos), static_cast<void *>(m->pos3),
                  static_cast<void *>(m->q_both), static_cast<void *>(m->q_img),
                  static_cast<void *>(m->f3), static_cast<void *>(m->self_part)})
    if (p) (void)hipFree(p);
  delete m;
}

int ElcSolver::num_modes() const { return static_cast<int>(m->modes.size()); }

void ElcSolver::compute(hipStream_t st, ElcParams const &p, int64_t n, double const *pos,
                        double const *q, double *force, double *energy, bool sync) {
  if (energy) *energy = 0.;
  for (int d = 0; d < 3; ++d)
    if (!(p.box[d] > 0.)) throw std::invalid_argument("ELC: box lengths must be positive");
  if (!(p.far_cut > 0.)) throw std::invalid_argument("ELC: far_cut must be positive");
  if (n < 0) throw std::invalid_argument("ELC: negative particle count");
  if (n == 0) return;
  hip_check(hipSetDevice(m->device), "hipSetDevice");
  bool const same = m->have_modes && m->cached.box[0] == p.box[0] &&
                    m->cached.box[1] == p.box[1] && m->cached.box[2] == p.box[2] &&
                    m->cached.far_cut == p.far_cut && m->cached.prefactor == p.prefactor;
  if (!same) {
    m->modes = elc_modes(p);
    elc_grow(m->d_modes, m->cap_modes, m->modes.size(), "hipMalloc elc modes");
    hip_check(hipMemcpyAsync(m->d_modes, m->modes.data(), sizeof(ElcMode) * m->modes.size(),
                             hipMemcpyHostToDevice, st),
              "copy elc modes");
    hip_check(hipStreamSynchronize(st), "sync elc modes");
    m->cached = p;
    m->have_modes = true;
  }
  if (m->modes.size() > static_cast<size_t>(kMaxModes)) {
    size_t const nm = m->modes.size();
    m->have_modes = false;
    throw std::runtime_error("ELC: far_cut " + std::to_string(p.far_cut) + " gives " +
                             std::to_string(nm) + " reciprocal vectors (at most " +
                             std::to_string(kMaxModes) + ")");
  }
  int const nmodes = static_cast<int>(m->modes.size());
  int const nprep = static_cast<int>((n + kElcBlock - 1) / kElcBlock);
  // particle blocks of the sums: at most kMaxSumBlocks, whole thread strides
  int64_t ppb = kElcBlock;
  while ((n + ppb - 1) / ppb > kMaxSumBlocks) ppb += kElcBlock;
  int const nsb = static_cast<int>((n + ppb - 1) / ppb);
  elc_grow(m->spos, m->cap_spos, static_cast<size_t>(n), "hipMalloc elc positions");
  elc_grow(m->prep_part, m->cap_prep, static_cast<size_t>(nprep) * kDipSums,
           "hipMalloc elc partials");
  elc_grow(m->sum_part, m->cap_sum, static_cast<size_t>(nsb) * std::max(nmodes, 1) * 8,
           "hipMalloc elc mode partials");
  elc_grow(m->G, m->cap_G, static_cast<size_t>(std::max(nmodes, 1)) * 8, "hipMalloc elc sums");
  if (energy) elc_grow(m->e_part, m->cap_e, static_cast<size_t>(nprep), "hipMalloc elc energy");
  ElcGeo const g = make_geo(p);
  k_elc_clear<<<1, 1, 0, st>>>(m->err);
  k_elc_prep<<<nprep, kElcBlock, 0, st>>>(n, pos, q, g, m->spos, m->prep_part, m->err);
  if (nmodes > 0) {
    dim3 const grid(nsb, (nmodes + kModeTile - 1) / kModeTile);
    if (g.diel)
      k_elc_sums<true><<<grid, kElcBlock, 0, st>>>(n, ppb, m->spos, m->d_modes, nmodes, g,
                                                   m->sum_part);
    else
      k_elc_sums<false><<<grid, kElcBlock, 0, st>>>(n, ppb, m->spos, m->d_modes, nmodes, g,
                                                    m->sum_part);
  }
  k_elc_reduce<<<nmodes + 1, kElcBlock, 0, st>>>(m->sum_part, nsb, nmodes, m->prep_part, nprep,
                                                 m->G, m->dip);
  double const volume = p.box[0] * p.box[1] * p.box[2];
  ElcForceConst c;
  c.dip_pref = p.prefactor * 4. * std::numbers::pi / volume;
  c.z_pref = p.prefactor * 2. * std::numbers::pi / (p.box[0] * p.box[1]);
  c.pot_diff = p.pot_diff;
  c.neutralize = (p.neutralize && !g.diel) ? 1 : 0;
  if (energy)
    k_elc_forces<true><<<nprep, kElcBlock, 0, st>>>(n, m->spos, m->d_modes, nmodes, m->G, m->dip,
                                                    g, c, force, m->e_part);
  else
    k_elc_forces<false><<<nprep, kElcBlock, 0, st>>>(n, m->spos, m->d_modes, nmodes, m->G,
                                                     m->dip, g, c, force, nullptr);
  hip_check(hipGetLastError(), "ELC kernels");
  if (!energy && !sync) return;
  std::vector<double> ep(energy ? static_cast<size_t>(nprep) : 0);
  double v[kDipSums];
  int err = 0;
  if (energy) {
    hip_check(hipMemcpyAsync(ep.data(), m->e_part, sizeof(double) * nprep,
                             hipMemcpyDeviceToHost, st),
              "copy elc energy");
    hip_check(hipMemcpyAsync(v, m->dip, sizeof(v), hipMemcpyDeviceToHost, st), "copy elc dip");
  }
  hip_check(hipMemcpyAsync(&err, m->err, sizeof(int), hipMemcpyDeviceToHost, st), "copy elc flag");
  hip_check(hipStreamSynchronize(st), "sync elc");
  if (err & 2) throw std::runtime_error("ELC: non-finite particle position");
  if (err & 1) throw std::runtime_error("ELC: a charged particle entered the ELC gap region");
  if (!energy) return;
  double e_modes = 0.;
  for (double e : ep) e_modes += e;
  // dipole energy (elc.cpp:267-342): b0 = v1, b1 = v4, b2 = v0, b3 = v5,
  // b4 = v2, b5 = v6, b6 = v3
  double const pe = p.prefactor * 2. * std::numbers::pi / volume;
  double const Lz = p.box[2], h = g.h;
  double e_dip = 2. * pe * (v[0] * v[0] + v[0] * v[5]);
  if (c.neutralize == 0)
    e_dip += 2. * pe * (-v[1] * v[2] - (0.25 - 0.5 / 3.) * (v[1] * Lz) * (v[1] * Lz));
  double e_z = 0.;
  if (g.diel) {
    if (g.const_pot) {
      e_dip += pe / h * Lz * v[3] * v[3];
      e_dip -= 2. * p.pot_diff / h * v[3];
    }
    e_dip += pe * (-(v[4] * v[2] + v[1] * v[6]) - (1. - 2. / 3.) * v[1] * v[4] * Lz * Lz);
    // z energy (elc.cpp:365-439)
    e_z = -c.z_pref * (v[0] * v[8] - v[1] * v[9]);
  }
  // the reference halves the whole sum (elc.cpp:938)
  *energy = 0.5 * (e_dip + e_z + e_modes);
}

void ElcSolver::images(hipStream_t st, ElcParams const &p, int64_t n, double const *pos,
                       double const *q, double **pos3, double **q_both, double **q_img) {
  hip_check(hipSetDevice(m->device), "hipSetDevice");
  size_t const n3 = 3 * static_cast<size_t>(std::max<int64_t>(n, 1));
  elc_grow(m->ispos, m->cap_ispos, static_cast<size_t>(std::max<int64_t>(n, 1)),
           "hipMalloc elc image positions");
  elc_grow(m->pos3, m->cap_pos3, 3 * n3, "hipMalloc elc image arrays");
  elc_grow(m->q_both, m->cap_qb, n3, "hipMalloc elc image charges");
  elc_grow(m->q_img, m->cap_qi, n3, "hipMalloc elc image charges");
  elc_grow(m->prep_part, m->cap_prep,
           static_cast<size_t>((n + kElcBlock - 1) / kElcBlock + 1) * kDipSums,
           "hipMalloc elc partials");
  ElcGeo const g = make_geo(p);
  int const nb = static_cast<int>((n + kElcBlock - 1) / kElcBlock);
  if (n > 0) {
    k_elc_clear<<<1, 1, 0, st>>>(m->err);
    k_elc_prep<<<nb, kElcBlock, 0, st>>>(n, pos, q, g, m->ispos, m->prep_part, m->err);
    k_elc_images<<<nb, kElcBlock, 0, st>>>(n, m->ispos, g, m->pos3, m->q_both, m->q_img);
    hip_check(hipGetLastError(), "ELC image kernels");
  }
  *pos3 = m->pos3;
  *q_both = m->q_both;
  *q_img = m->q_img;
}

double *ElcSolver::image_forces(hipStream_t st, int64_t n) {
  size_t const need = 9 * static_cast<size_t>(std::max<int64_t>(n, 1));
  elc_grow(m->f3, m->cap_f3, need, "hipMalloc elc image forces");
  int const nb = static_cast<int>((need + kElcBlock - 1) / kElcBlock);
  k_elc_zero<<<nb, kElcBlock, 0, st>>>(static_cast<int64_t>(need), m->f3);
  hip_check(hipGetLastError(), "ELC clear");
  return m->f3;
}

void ElcSolver::add_real_rows(hipStream_t st, int64_t n, double const *f3, double *force) {
  if (n <= 0) return;
  int64_t const n3 = 3 * n;
  k_elc_add_rows<<<static_cast<int>((n3 + kElcBlock - 1) / kElcBlock), kElcBlock, 0, st>>>(
      n3, f3, force);
  hip_check(hipGetLastError(), "ELC add rows");
}

void ElcSolver::layer_self(hipStream_t st, ElcParams const &p, int64_t n, double const *pos,
                           double const *q, double *force, double *energy) {
  if (energy) *energy = 0.;
  if (n <= 0) return;
  hip_check(hipSetDevice(m->device), "hipSetDevice");
  int const nb = static_cast<int>((n + kElcBlock - 1) / kElcBlock);
  elc_grow(m->ispos, m->cap_ispos, static_cast<size_t>(n), "hipMalloc elc image positions");
  elc_grow(m->prep_part, m->cap_prep, static_cast<size_t>(nb) * kDipSums,
           "hipMalloc elc partials");
  elc_grow(m->self_part, m->cap_self, static_cast<size_t>(nb), "hipMalloc elc self energy");
  ElcGeo const g = make_geo(p);
  k_elc_clear<<<1, 1, 0, st>>>(m->err);
  k_elc_prep<<<nb, kElcBlock, 0, st>>>(n, pos, q, g, m->ispos, m->prep_part, m->err);
  if (energy)
    k_elc_self<true><<<nb, kElcBlock, 0, st>>>(n, m->ispos, g, p.prefactor, p.alpha, p.r_cut,
                                               force, m->self_part);
  else
    k_elc_self<false><<<nb, kElcBlock, 0, st>>>(n, m->ispos, g, p.prefactor, p.alpha, p.r_cut,
                                                force, nullptr);
  hip_check(hipGetLastError(), "ELC self kernels");
  if (!energy) return;
  std::vector<double> e(static_cast<size_t>(nb));
  hip_check(hipMemcpyAsync(e.data(), m->self_part, sizeof(double) * nb, hipMemcpyDeviceToHost,
                           st),
            "copy elc self energy");
  hip_check(hipStreamSynchronize(st), "sync elc self");
  for (double v : e) *energy += v;
}

} // namespace espresso_amd
