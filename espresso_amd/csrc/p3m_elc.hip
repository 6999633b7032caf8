// ELC far-field terms on the device; see p3m_elc.hpp.
//
// Layout of one call (all FP64, deterministic reductions):
//   k_elc_prep     fold positions into the box, gap check, per-block sums of
//                  q (z - L_z/2), q and q (z - L_z/2)^2 (dipole terms)
//   k_elc_sums     per (particle block, tile of modes): the eight sums
//                  q {sin,cos}(k_x x) {sin,cos}(k_y y) e^{+-omega z}
//   k_elc_reduce   per mode: sum of the block partials in block order, times
//                  the mode's scale; one more block reduces the dipole sums
//   k_elc_forces   per particle: all modes against the global sums (force
//                  and energy), plus the dipole and background force
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
};

__device__ __forceinline__ double fold_into(double x, double L) {
  double xf = x - L * floor(x / L);
  if (xf >= L) xf -= L;
  return xf;
}

/** The eight values of particle (x, y, z, q) for mode md, in the order
 *  [ss, sc, cs, cc] x [e^{+omega z}, e^{-omega z}] (v[0..3] plus, v[4..7]
 *  minus).  kx = 0 or ky = 0 stands for the missing direction (cos = 1,
 *  sin = 0), which turns the pq-mode values into those of the p- and
 *  q-modes (elc.cpp:500-600, 641-770). */
__device__ __forceinline__ void mode_values(ElcMode const &md, double4 const &p, double v[8]) {
  double sx = 0., cx = 1., sy = 0., cy = 1.;
  if (md.kx != 0.) sincos(md.kx * p.x, &sx, &cx);
  if (md.ky != 0.) sincos(md.ky * p.y, &sy, &cy);
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

__global__ __launch_bounds__(kElcBlock) void k_elc_prep(int64_t n, double const *__restrict__ pos,
                                                        double const *__restrict__ q, ElcGeo g,
                                                        double4 *__restrict__ spos,
                                                        double *partials, int *err) {
  __shared__ double red[4 * 16];
  int64_t const i = (int64_t)blockIdx.x * kElcBlock + threadIdx.x;
  double v[4] = {0., 0., 0., 0.};
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
    double const dz = p.z - g.shift;
    v[0] = p.w * dz;
    v[1] = p.w;
    v[2] = p.w * dz * dz;
  }
  block_sum<4>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) partials[blockIdx.x * 4 + k] = v[k];
}

/** blockIdx.x: particle block (ppb particles, ppb / kElcBlock per thread);
 *  blockIdx.y: tile of kModeTile modes.  partials[(blk * nmodes + m) * 8 + k]. */
__global__ __launch_bounds__(kElcBlock) void k_elc_sums(int64_t n, int64_t ppb,
                                                        double4 const *__restrict__ spos,
                                                        ElcMode const *__restrict__ modes,
                                                        int nmodes, double *partials) {
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
#pragma unroll
    for (int t = 0; t < kModeTile; ++t) {
      if (t < nm) {
        double v[8];
        mode_values(modes[m0 + t], p, v);
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[t][k] += v[k];
      }
    }
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

/** Blocks 0 .. nmodes-1: G[m][k] = pref_m * sum over the nsb particle blocks
 *  (in block order); block nmodes: dip[0..3] = sums of the prep partials. */
__global__ __launch_bounds__(kElcBlock) void k_elc_reduce(double const *__restrict__ partials,
                                                          int nsb, int nmodes,
                                                          ElcMode const *__restrict__ modes,
                                                          double const *__restrict__ prep_part,
                                                          int nprep, double *G, double *dip) {
  __shared__ double red[8 * 16];
  int const m = blockIdx.x;
  double v[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  if (m < nmodes) {
    for (int b = threadIdx.x; b < nsb; b += kElcBlock)
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] += partials[((int64_t)b * nmodes + m) * 8 + k];
    block_sum<8>(v, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 8; ++k) G[m * 8 + k] = modes[m].pref * v[k];
  } else {
    for (int b = threadIdx.x; b < nprep; b += kElcBlock)
#pragma unroll
      for (int k = 0; k < 4; ++k) v[k] += prep_part[b * 4 + k];
    block_sum<8>(v, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 4; ++k) dip[k] = v[k];
  }
}

/** Forces and energy of every particle against all modes
 *  (add_PoQ_force / add_PQ_force, PoQ_energy / PQ_energy,
 *  elc.cpp:601-640, 771-828) plus the dipole force and, without
 *  neutralisation, the removal of P3M's background force (:201-265). */
template <bool EN>
__global__ __launch_bounds__(kElcBlock) void k_elc_forces(
    int64_t n, double4 const *__restrict__ spos, ElcMode const *__restrict__ modes, int nmodes,
    double const *__restrict__ G, double const *__restrict__ dip, double dip_pref, int neutralize,
    double shift, double *__restrict__ force, double *partials) {
  __shared__ double red[16];
  int64_t const i = (int64_t)blockIdx.x * kElcBlock + threadIdx.x;
  double e = 0.;
  if (i < n) {
    double4 const p = spos[i];
    if (p.w != 0.) {
      double fx = 0., fy = 0., fz = 0.;
      for (int m = 0; m < nmodes; ++m) {
        ElcMode const md = modes[m];
        double a[8];
        mode_values(md, p, a);
        double const *g = G + m * 8;
        // plus values against minus sums and minus values against plus sums
        double const gp0 = g[0], gp1 = g[1], gp2 = g[2], gp3 = g[3];
        double const gm0 = g[4], gm1 = g[5], gm2 = g[6], gm3 = g[7];
        double const inv_w = 1. / md.omega;
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
      // dipole term (field of the total z moment) and background removal
      fz -= dip_pref * dip[0] * p.w;
      if (!neutralize) fz += dip_pref * dip[1] * p.w * (p.z - shift);
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
  auto add = [&](double kx, double ky, double omega, double c) {
    out.push_back(ElcMode{kx, ky, omega, -c * area_pi / std::expm1(omega * Lz)});
  };
  for (std::size_t i = 1; ux * static_cast<double>(i - 1) < p.far_cut && i <= nx; ++i) {
    double const om = two_pi * ux * static_cast<double>(i);
    add(om, 0., om, 4.);
  }
  for (std::size_t j = 1; uy * static_cast<double>(j - 1) < p.far_cut && j <= ny; ++j) {
    double const om = two_pi * uy * static_cast<double>(j);
    add(0., om, om, 4.);
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
          8.);
    }
  return out;
}

double elc_tune_far_cut(double const box[3], double gap_size, double max_pw_error) {
  constexpr double maximal_far_cut = 50.;
  double const ux = 1. / box[0], uy = 1. / box[1];
  double const step = std::min(ux, uy);
  double const lz = box[2];
  double const h = box[2] - gap_size;
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
};

ElcSolver::ElcSolver(int device) : m(new Impl) {
  m->device = device;
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipMalloc(&m->dip, sizeof(double) * 4), "hipMalloc elc dipole");
  hip_check(hipMalloc(&m->err, sizeof(int)), "hipMalloc elc flag");
}

ElcSolver::~ElcSolver() {
  (void)hipSetDevice(m->device);
  for (void *p : {static_cast<void *>(m->d_modes), static_cast<void *>(m->spos),
                  static_cast<void *>(m->prep_part), static_cast<void *>(m->sum_part),
                  static_cast<void *>(m->G), static_cast<void *>(m->dip),
                  static_cast<void *>(m->e_part), static_cast<void *>(m->err)})
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
  elc_grow(m->prep_part, m->cap_prep, static_cast<size_t>(nprep) * 4, "hipMalloc elc partials");
  elc_grow(m->sum_part, m->cap_sum, static_cast<size_t>(nsb) * std::max(nmodes, 1) * 8,
           "hipMalloc elc mode partials");
  elc_grow(m->G, m->cap_G, static_cast<size_t>(std::max(nmodes, 1)) * 8, "hipMalloc elc sums");
  if (energy) elc_grow(m->e_part, m->cap_e, static_cast<size_t>(nprep), "hipMalloc elc energy");
  ElcGeo g;
  for (int d = 0; d < 3; ++d) g.L[d] = p.box[d];
  g.h = p.box[2] - p.gap_size;
  g.shift = 0.5 * p.box[2];
  k_elc_clear<<<1, 1, 0, st>>>(m->err);
  k_elc_prep<<<nprep, kElcBlock, 0, st>>>(n, pos, q, g, m->spos, m->prep_part, m->err);
  if (nmodes > 0) {
    dim3 const grid(nsb, (nmodes + kModeTile - 1) / kModeTile);
    k_elc_sums<<<grid, kElcBlock, 0, st>>>(n, ppb, m->spos, m->d_modes, nmodes, m->sum_part);
  }
  k_elc_reduce<<<nmodes + 1, kElcBlock, 0, st>>>(m->sum_part, nsb, nmodes, m->d_modes,
                                                 m->prep_part, nprep, m->G, m->dip);
  double const volume = p.box[0] * p.box[1] * p.box[2];
  double const dip_pref = p.prefactor * 4. * std::numbers::pi / volume;
  if (energy)
    k_elc_forces<true><<<nprep, kElcBlock, 0, st>>>(n, m->spos, m->d_modes, nmodes, m->G, m->dip,
                                                    dip_pref, p.neutralize ? 1 : 0, g.shift,
                                                    force, m->e_part);
  else
    k_elc_forces<false><<<nprep, kElcBlock, 0, st>>>(n, m->spos, m->d_modes, nmodes, m->G,
                                                     m->dip, dip_pref, p.neutralize ? 1 : 0,
                                                     g.shift, force, nullptr);
  hip_check(hipGetLastError(), "ELC kernels");
  if (!energy && !sync) return;
  std::vector<double> ep(energy ? static_cast<size_t>(nprep) : 0);
  double dip[4];
  int err = 0;
  if (energy) {
    hip_check(hipMemcpyAsync(ep.data(), m->e_part, sizeof(double) * nprep,
                             hipMemcpyDeviceToHost, st),
              "copy elc energy");
    hip_check(hipMemcpyAsync(dip, m->dip, sizeof(dip), hipMemcpyDeviceToHost, st),
              "copy elc dip");
  }
  hip_check(hipMemcpyAsync(&err, m->err, sizeof(int), hipMemcpyDeviceToHost, st), "copy elc flag");
  hip_check(hipStreamSynchronize(st), "sync elc");
  if (err & 2) throw std::runtime_error("ELC: non-finite particle position");
  if (err & 1) throw std::runtime_error("ELC: a charged particle entered the ELC gap region");
  if (!energy) return;
  double e_modes = 0.;
  for (double v : ep) e_modes += v;
  // dipole energy (elc.cpp:267-342), halved with the mode sums (elc.cpp:938)
  double const pe = p.prefactor * 2. * std::numbers::pi / volume;
  double e_dip = 2. * pe * dip[0] * dip[0];
  if (!p.neutralize)
    e_dip += 2. * pe * (-dip[1] * dip[2] - (0.25 - 0.5 / 3.) * (dip[1] * p.box[2]) *
                                               (dip[1] * p.box[2]));
  *energy = 0.5 * (e_dip + e_modes);
}

} // namespace espresso_amd
