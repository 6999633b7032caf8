// Device real-space pair forces of P3M Coulomb over cell lists; see
// p3m_short_range.hpp.
#include "p3m_short_range.hpp"
#include "p3m_batch.hpp"
#include "p3m_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <numbers>
#include <stdexcept>
#include <vector>

namespace espresso_amd {
namespace {
using namespace dev;

constexpr int kSRMaxCells = 64; // cells per dimension at most

/** Cell grid: nc cells per dimension of edge L / nc >= r_cut / nsub. */
struct CellGrid {
  int nc[3];
  double L[3];
  double inv_cell[3]; // nc / L
};

/** x folded into [0, L). */
__device__ __forceinline__ double fold(double x, double L) {
  double xf = x - L * floor(x / L);
  if (xf >= L) xf -= L;
  return xf < 0. ? 0. : xf;
}

__global__ void k_sr_fill(int *p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

/** Cell of every charged particle (uncharged ones take no part in pair
 *  terms, p3m.hpp:150) and its rank in the cell. */
__global__ __launch_bounds__(256) void k_sr_bin(int64_t n, double const *__restrict__ pos,
                                                double const *__restrict__ q, CellGrid g,
                                                int *counts, int *p_cell, int *p_rank) {
  int64_t const i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (q[i] == 0.) {
    p_cell[i] = -1;
    return;
  }
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    c[d] = min(g.nc[d] - 1, static_cast<int>(fold(pos[3 * i + d], g.L[d]) * g.inv_cell[d]));
  int const cell = (c[0] * g.nc[1] + c[1]) * g.nc[2] + c[2];
  p_cell[i] = cell;
  p_rank[i] = atomicAdd(&counts[cell], 1);
}

/** Exclusive scan of the cell counts (one block); offsets[ncells] = total. */
__global__ __launch_bounds__(1024) void k_sr_scan(int const *counts, int ncells, int *offsets) {
  __shared__ int wsum[16];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  int const lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int base = 0; base < ncells; base += blockDim.x) {
    int const i = base + threadIdx.x;
    int const c = (i < ncells) ? counts[i] : 0;
    int s = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      int const t = __shfl_up(s, off, 64);
      if (lane >= off) s += t;
    }
    if (lane == 63) wsum[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0;
      for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
        int const t = wsum[w];
        wsum[w] = acc;
        acc += t;
      }
    }
    __syncthreads();
    if (i < ncells) offsets[i] = carry + wsum[wid] + s - c;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += wsum[wid] + s;
    __syncthreads();
  }
  if (threadIdx.x == 0) offsets[ncells] = carry;
}

/** Charged particles in cell order: folded position and charge, and the
 *  original index. */
__global__ __launch_bounds__(256) void k_sr_scatter(int64_t n, double const *__restrict__ pos,
                                                    double const *__restrict__ q, CellGrid g,
                                                    int const *p_cell, int const *p_rank,
                                                    int const *offsets, double4 *spos,
                                                    int *sidx) {
  int64_t const i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int const cell = p_cell[i];
  if (cell < 0) return;
  int const slot = offsets[cell] + p_rank[i];
  spos[slot] = double4{fold(pos[3 * i], g.L[0]), fold(pos[3 * i + 1], g.L[1]),
                       fold(pos[3 * i + 2], g.L[2]), q[i]};
  sidx[slot] = static_cast<int>(i);
}

/** 1/x and 1/sqrt(x) in FP64: the hardware estimate refined by two
 *  Newton-Raphson steps (each doubles the correct bits). */
__device__ __forceinline__ double nr_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.);
  y = fma(y, e, y);
  e = fma(-x, y, 1.);
  return fma(y, e, y);
}
__device__ __forceinline__ double nr_rsq(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x, y * y, 1.);
  y = fma(0.5 * y, e, y);
  e = fma(-x, y * y, 1.);
  return fma(0.5 * y, e, y);
}

/** Pair term of CoulombP3M::pair_force / pair_energy (p3m.hpp:146-180) with
 *  the A&S erfc of AS_erfc_part.hpp:29-40 (the published 7.1.26 constants)
 *  for a pair already known to lie within the cutoff: d = r_i - r_j,
 *  r2 = |d|^2 > 0, qq = prefactor q_i q_j.  Adds the force on i and, with
 *  EV, the pair energy and F (x) d (upper triangle: xx xy xz yy yz zz). */
template <bool EV>
__device__ __forceinline__ void pair_term(double dx, double dy, double dz, double r2, double qq,
                                          double alpha, double two_a_sqrt_pi_i, double &fx,
                                          double &fy, double &fz, double (&ev)[7]) {
  double const rinv = nr_rsq(r2);
  double const adist = alpha * (r2 * rinv);
  double const t = nr_rcp(fma(0.3275911, adist, 1.));
  double const erfc_part =
      t * (0.254829592 +
           t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
  double const ex = exp(-adist * adist);
  double const erfc_ri = erfc_part * rinv;
  double const fac = ex * (erfc_ri + two_a_sqrt_pi_i) * (rinv * rinv) * qq;
  fx += fac * dx;
  fy += fac * dy;
  fz += fac * dz;
  if constexpr (EV) {
    ev[0] += qq * erfc_ri * ex;
    ev[1] += fac * dx * dx;
    ev[2] += fac * dx * dy;
    ev[3] += fac * dx * dz;
    ev[4] += fac * dy * dy;
    ev[5] += fac * dy * dz;
    ev[6] += fac * dz * dz;
  }
}

/** ELC image terms of dielectric contrasts (off when on == 0). */
struct SrImages {
  int on;
  double h, sl, dt, db;
};

/** Pair term of particle i with the image of itself in its layer against
 *  partner j (add_pair_force_corrections / pair_energy_correction,
 *  elc.hpp:250-296; dielectric_layers_contribution, :119-138): q_eff =
 *  delta q_i q_j, d = r_j - r_i' with r_i' = (x_i, y_i, -z_i) below the
 *  layer bound or (x_i, y_i, 2h - z_i) above it, minimum image along z.
 *  dx, dy = r_i - r_j in the plane; zj is the partner's folded z.  Force on
 *  i and, with EV, the pair energy (ev[0]). */
template <bool EV>
__device__ __forceinline__ void image_terms(SrImages const &im, double zi, double zj, double dx,
                                            double dy, double Lz, double qq, double rc2,
                                            double alpha, double two_a, double &fx, double &fy,
                                            double &fz, double (&ev)[7]) {
  auto one = [&](double delta, double dz) {
    dz -= Lz * rint(dz / Lz);
    double const r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < rc2) || !(r2 > 0.) || delta == 0.) return;
    double ax = -dx, ay = -dy, az = dz; // d = r_j - r_i'
    double const rinv = nr_rsq(r2);
    double const adist = alpha * (r2 * rinv);
    double const t = nr_rcp(fma(0.3275911, adist, 1.));
    double const erfc_part =
        t * (0.254829592 +
             t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
    double const ex = exp(-adist * adist);
    double const erfc_ri = erfc_part * rinv;
    double const q = qq * delta;
    double const fac = ex * (erfc_ri + two_a) * (rinv * rinv) * q;
    fx += fac * ax;
    fy += fac * ay;
    fz += fac * az;
    if constexpr (EV) ev[0] += q * erfc_ri * ex;
  };
  if (zi < im.sl) one(im.db, zj + zi);
  if (zi > im.h - im.sl) one(im.dt, zj - (2. * im.h - zi));
}

/** Displacement r_i - r_j, with the minimum image when MI. */
template <bool MI>
__device__ __forceinline__ void displacement(double4 const &pi, double4 const &pj,
                                             CellGrid const &g, double &dx, double &dy,
                                             double &dz) {
  dx = pi.x - pj.x;
  dy = pi.y - pj.y;
  dz = pi.z - pj.z;
  if constexpr (MI) {
    dx -= g.L[0] * rint(dx / g.L[0]);
    dy -= g.L[1] * rint(dy / g.L[1]);
    dz -= g.L[2] * rint(dz / g.L[2]);
  }
}

/** The staged candidates tile[0, cnt) against this thread's particle, in
 *  groups of 64: a cheap pass marks the partners within the cutoff in a
 *  64-bit mask, then each lane evaluates only its own marked pairs (the
 *  expensive pair term runs with the lanes that have work, instead of on
 *  every candidate that any lane of the wave needs). */
template <bool MI, bool EV>
__device__ __forceinline__ void tile_pairs(double4 const *tile, int cnt, double4 const &pi,
                                           double qpi, CellGrid const &g, double rc2,
                                           double alpha, double two_a, double &fx, double &fy,
                                           double &fz, double (&ev)[7], SrImages const &im,
                                           double sz) {
  for (int kb = 0; kb < cnt; kb += 64) {
    int const kn = min(64, cnt - kb);
    unsigned long long mask = 0;
    for (int k = 0; k < kn; ++k) {
      double dx, dy, dz;
      displacement<MI>(pi, tile[kb + k], g, dx, dy, dz);
      double const r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < rc2 && r2 > 0.) mask |= 1ull << k;
    }
    while (mask) {
      int const k = __builtin_ctzll(mask);
      mask &= mask - 1;
      double4 const pj = tile[kb + k];
      double dx, dy, dz;
      displacement<MI>(pi, pj, g, dx, dy, dz);
      double const r2 = dx * dx + dy * dy + dz * dz;
      pair_term<EV>(dx, dy, dz, r2, qpi * pj.w, alpha, two_a, fx, fy, fz, ev);
      if (im.on)
        image_terms<EV>(im, pi.z, pj.z - sz, dx, dy, g.L[2], qpi * pj.w, rc2, alpha, two_a, fx,
                        fy, fz, ev);
    }
  }
}

/** One block of NT threads per cell: the cell's particles (NT at a time)
 *  against the (2 NSUB + 1)^3 neighbour cells, each staged through LDS with
 *  its periodic shift (nc >= 2 NSUB + 1 per dimension: every neighbour is a
 *  distinct cell).  Each thread sums its particle's force over all partners
 *  (both orders of a pair are evaluated), so no atomics are needed. */
template <int NSUB, int NT, bool EV>
__global__ __launch_bounds__(NT) void k_sr_cells(CellGrid g, int const *__restrict__ offsets,
                                                 double4 const *__restrict__ spos,
                                                 int const *__restrict__ sidx, double prefactor,
                                                 double alpha, double rc2,
                                                 double *__restrict__ force, double *partials,
                                                 SrImages im) {
  __shared__ double4 tile[NT];
  __shared__ double red[8 * 16];
  constexpr int W = 2 * NSUB + 1;
  int const cell = blockIdx.x;
  int const cz = cell % g.nc[2], cy = (cell / g.nc[2]) % g.nc[1], cx = cell / (g.nc[2] * g.nc[1]);
  int const beg = offsets[cell], end = offsets[cell + 1];
  double const two_a = 2. * alpha * std::numbers::inv_sqrtpi;
  double ev[7] = {0., 0., 0., 0., 0., 0., 0.};
  for (int hb = beg; hb < end; hb += NT) {
    int const i = hb + static_cast<int>(threadIdx.x);
    bool const has = i < end;
    double4 const pi = has ? spos[i] : double4{0., 0., 0., 0.};
    double const qpi = prefactor * pi.w;
    double fx = 0., fy = 0., fz = 0.;
    for (int o = 0; o < W * W * W; ++o) {
      int nx = cx + o / (W * W) - NSUB, ny = cy + (o / W) % W - NSUB, nz = cz + o % W - NSUB;
      double sx = 0., sy = 0., sz = 0.;
      if (nx < 0) { nx += g.nc[0]; sx = -g.L[0]; } else if (nx >= g.nc[0]) { nx -= g.nc[0]; sx = g.L[0]; }
      if (ny < 0) { ny += g.nc[1]; sy = -g.L[1]; } else if (ny >= g.nc[1]) { ny -= g.nc[1]; sy = g.L[1]; }
      if (nz < 0) { nz += g.nc[2]; sz = -g.L[2]; } else if (nz >= g.nc[2]) { nz -= g.nc[2]; sz = g.L[2]; }
      int const nb = (nx * g.nc[1] + ny) * g.nc[2] + nz;
      int const jb = offsets[nb], je = offsets[nb + 1];
      for (int tb = jb; tb < je; tb += NT) {
        __syncthreads(); // the previous tile is consumed
        int const j = tb + static_cast<int>(threadIdx.x);
        if (j < je) {
          double4 const pj = spos[j];
          tile[threadIdx.x] = double4{pj.x + sx, pj.y + sy, pj.z + sz, pj.w};
        }
        __syncthreads();
        if (has)
          tile_pairs<false, EV>(tile, min(NT, je - tb), pi, qpi, g, rc2, alpha, two_a, fx, fy,
                                fz, ev, im, sz);
      }
    }
    if (has) {
      double *f = force + 3 * static_cast<int64_t>(sidx[i]);
      f[0] += fx;
      f[1] += fy;
      f[2] += fz;
    }
  }
  if constexpr (EV) {
    block_sum<7>(ev, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 7; ++k) partials[blockIdx.x * 8 + k] = ev[k];
  }
}

/** Pair terms over cells with a per-wave pair queue (no ELC images): one
 *  block of 256 threads per home cell.  Wave w takes the candidates w,
 *  w + 4, ... of each neighbour cell (one uniform load per candidate), its
 *  lanes the home particles (up to 256 per pass); the pairs inside r_cut go
 *  to the wave's own LDS queue (ballot, wave-uniform count), which the wave
 *  drains with every lane evaluating one pair.  The pair terms so run
 *  without the idle lanes of per-lane pair loops (lanes without a home
 *  particle, lanes with fewer partners than the wave's busiest); forces on
 *  the home particles accumulate in LDS (FP64 LDS atomics, so the order of
 *  the sums, not their set, depends on scheduling) and are added to the
 *  caller's array once.  Energy and virial are per-lane partials reduced
 *  per block. */
template <int NSUB, bool EV>
__global__ __launch_bounds__(256) void k_sr_wave(CellGrid g, int const *__restrict__ offsets,
                                                 double4 const *__restrict__ spos,
                                                 int const *__restrict__ sidx, double prefactor,
                                                 double alpha, double rc2,
                                                 double *__restrict__ force, double *partials) {
  constexpr int NT = 256, NW = NT / 64, W = 2 * NSUB + 1, NN = W * W * W;
  constexpr int QW = 256; // queue entries per wave
  __shared__ double4 hp[NT];
  __shared__ double acc[NT * 3];
  __shared__ int2 queue[NW][QW]; // (candidate index, home slot | neighbour << 16)
  __shared__ int nbeg[NN], ncnt[NN];
  __shared__ signed char nsc[NN * 3]; // periodic shift of each neighbour in box lengths
  __shared__ double red[8 * 16];
  int const cell = blockIdx.x;
  int const cz = cell % g.nc[2], cy = (cell / g.nc[2]) % g.nc[1], cx = cell / (g.nc[2] * g.nc[1]);
  int const beg = offsets[cell], end = offsets[cell + 1];
  if (beg == end) {
    if (EV && threadIdx.x < 7) partials[blockIdx.x * 8 + threadIdx.x] = 0.;
    return;
  }
  for (int o = threadIdx.x; o < NN; o += NT) {
    int nx = cx + o / (W * W) - NSUB, ny = cy + (o / W) % W - NSUB, nz = cz + o % W - NSUB;
    int sx = 0, sy = 0, sz = 0;
    if (nx < 0) { nx += g.nc[0]; sx = -1; } else if (nx >= g.nc[0]) { nx -= g.nc[0]; sx = 1; }
    if (ny < 0) { ny += g.nc[1]; sy = -1; } else if (ny >= g.nc[1]) { ny -= g.nc[1]; sy = 1; }
    if (nz < 0) { nz += g.nc[2]; sz = -1; } else if (nz >= g.nc[2]) { nz -= g.nc[2]; sz = 1; }
    int const nb = (nx * g.nc[1] + ny) * g.nc[2] + nz;
    nbeg[o] = offsets[nb];
    ncnt[o] = offsets[nb + 1] - offsets[nb];
    nsc[3 * o] = static_cast<signed char>(sx);
    nsc[3 * o + 1] = static_cast<signed char>(sy);
    nsc[3 * o + 2] = static_cast<signed char>(sz);
  }
  int const lane = static_cast<int>(threadIdx.x & 63);
  int const wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  int2 *const wq = queue[wave];
  double const two_a = 2. * alpha * std::numbers::inv_sqrtpi;
  double ev[7] = {0., 0., 0., 0., 0., 0., 0.};
  for (int hb = beg; hb < end; hb += NT) {
    int const ni = min(NT, end - hb);
    __syncthreads(); // neighbour table written; previous pass's readers done
    if (static_cast<int>(threadIdx.x) < ni) hp[threadIdx.x] = spos[hb + threadIdx.x];
    for (int x = threadIdx.x; x < ni * 3; x += NT) acc[x] = 0.;
    __syncthreads();
    int const nr = (ni + 63) >> 6;
    // the lane's home particles stay in registers for the tests
    double hx[NT / 64], hy[NT / 64], hz[NT / 64];
#pragma unroll
    for (int r = 0; r < NT / 64; ++r) {
      double4 const pi = hp[lane + 64 * r];
      hx[r] = pi.x;
      hy[r] = pi.y;
      hz[r] = pi.z;
    }
    int count = 0; // wave-uniform
    auto drain = [&]() {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int q = lane; q < count; q += 64) {
        int2 const en = wq[q];
        int const il = en.y & 0xffff, o = en.y >> 16;
        double4 const pj = spos[en.x];
        double4 const pi = hp[il];
        double const dx = pi.x - (pj.x + nsc[3 * o] * g.L[0]);
        double const dy = pi.y - (pj.y + nsc[3 * o + 1] * g.L[1]);
        double const dz = pi.z - (pj.z + nsc[3 * o + 2] * g.L[2]);
        double const r2 = dx * dx + dy * dy + dz * dz;
        double fx = 0., fy = 0., fz = 0.;
        pair_term<EV>(dx, dy, dz, r2, prefactor * pi.w * pj.w, alpha, two_a, fx, fy, fz, ev);
        atomicAdd(&acc[il * 3], fx);
        atomicAdd(&acc[il * 3 + 1], fy);
        atomicAdd(&acc[il * 3 + 2], fz);
      }
      __builtin_amdgcn_wave_barrier();
      count = 0;
    };
    for (int o = 0; o < NN; ++o) {
      int const jb = __builtin_amdgcn_readfirstlane(nbeg[o]);
      int const cnt = __builtin_amdgcn_readfirstlane(ncnt[o]);
      double const shx = nsc[3 * o] * g.L[0], shy = nsc[3 * o + 1] * g.L[1],
                   shz = nsc[3 * o + 2] * g.L[2];
      // the wave's candidates t = wave + NW u, 64 at a time: lane u loads
      // candidate u, the tests read them back with lane reads (no memory
      // latency per candidate)
      for (int t0 = wave; t0 < cnt; t0 += NW * 64) {
        int const tl = t0 + NW * lane;
        double4 const cj = (tl < cnt) ? spos[jb + tl] : double4{0., 0., 0., 0.};
        int const nu = min(64, (cnt - t0 + NW - 1) / NW);
        for (int u = 0; u < nu; ++u) {
          double const qx = readlane_f64(cj.x, u) + shx;
          double const qy = readlane_f64(cj.y, u) + shy;
          double const qz = readlane_f64(cj.z, u) + shz;
          int const t = t0 + NW * u;
#pragma unroll
          for (int r = 0; r < NT / 64; ++r) {
            if (r < nr) {
              int const il = lane + 64 * r;
              double const dx = hx[r] - qx, dy = hy[r] - qy, dz = hz[r] - qz;
              double const r2 = dx * dx + dy * dy + dz * dz;
              bool const hit = il < ni && r2 < rc2 && r2 > 0.;
              uint64_t const mask = __ballot(hit);
              if (hit)
                wq[count + __popcll(mask & ((1ull << lane) - 1ull))] =
                    make_int2(jb + t, il | (o << 16));
              count += __popcll(mask);
              // at most 64 entries join per test: drain while they still fit
              if (count > QW - 64) drain();
            }
          }
        }
      }
    }
    drain();
    __syncthreads(); // every wave's pairs are in acc
    if (static_cast<int>(threadIdx.x) < ni) {
      double *f = force + 3 * static_cast<int64_t>(sidx[hb + threadIdx.x]);
      double const *a = acc + threadIdx.x * 3;
      f[0] += a[0];
      f[1] += a[1];
      f[2] += a[2];
    }
  }
  if constexpr (EV) {
    block_sum<7>(ev, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 7; ++k) partials[blockIdx.x * 8 + k] = ev[k];
  }
}

/** All charged pairs with the minimum-image displacement (boxes narrower
 *  than three cells of edge r_cut); one block per NT sorted particles. */
template <int NT, bool EV>
__global__ __launch_bounds__(NT) void k_sr_allpairs(int nq, CellGrid g,
                                                    double4 const *__restrict__ spos,
                                                    int const *__restrict__ sidx,
                                                    double prefactor, double alpha, double rc2,
                                                    double *__restrict__ force,
                                                    double *partials, SrImages im) {
  __shared__ double4 tile[NT];
  __shared__ double red[8 * 16];
  int const i = blockIdx.x * NT + static_cast<int>(threadIdx.x);
  bool const has = i < nq;
  double4 const pi = has ? spos[i] : double4{0., 0., 0., 0.};
  double const qpi = prefactor * pi.w;
  double const two_a = 2. * alpha * std::numbers::inv_sqrtpi;
  double fx = 0., fy = 0., fz = 0.;
  double ev[7] = {0., 0., 0., 0., 0., 0., 0.};
  for (int tb = 0; tb < nq; tb += NT) {
    __syncthreads();
    int const j = tb + static_cast<int>(threadIdx.x);
    if (j < nq) tile[threadIdx.x] = spos[j];
    __syncthreads();
    if (has)
      tile_pairs<true, EV>(tile, min(NT, nq - tb), pi, qpi, g, rc2, alpha, two_a, fx, fy, fz,
                           ev, im, 0.);
  }
  if (has) {
    double *f = force + 3 * static_cast<int64_t>(sidx[i]);
    f[0] += fx;
    f[1] += fy;
    f[2] += fz;
  }
  if constexpr (EV) {
    block_sum<7>(ev, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 7; ++k) partials[blockIdx.x * 8 + k] = ev[k];
  }
}

/** |mu|^2 of every particle (the dipolar binning key: zero dipoles take no
 *  part, dp3m.hpp:162). */
__global__ __launch_bounds__(256) void k_sr_mu2(int64_t n, double const *__restrict__ dip,
                                                double *__restrict__ mu2) {
  int64_t const i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double const a = dip[3 * i], b = dip[3 * i + 1], c = dip[3 * i + 2];
  mu2[i] = a * a + b * b + c * c;
}

/** Dipoles in cell order. */
__global__ __launch_bounds__(256) void k_sr_sort_dip(int64_t n, int const *__restrict__ total,
                                                     int const *__restrict__ sidx,
                                                     double const *__restrict__ dip,
                                                     double4 *__restrict__ sdip) {
  int64_t const s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= n || s >= *total) return;
  int64_t const i = sidx[s];
  sdip[s] = double4{dip[3 * i], dip[3 * i + 1], dip[3 * i + 2], 0.};
}

/** Dipolar pair term (DipolarP3M::pair_force / pair_energy,
 *  dp3m.hpp:160-235) of particle i with partner j at d = r_i - r_j, r2 =
 *  |d|^2 within the cutoff: force and torque on i, pair energy with EV. */
template <bool EV>
__device__ __forceinline__ void dip_pair(double dx, double dy, double dz, double r2,
                                         double4 const &mi, double4 const &mj, double prefactor,
                                         double alpha, int exact, double (&f)[3],
                                         double (&t)[3], double &e) {
  double const r = sqrt(r2);
  double const ad = alpha * r;
  double const ex = exp(-ad * ad);
  double const r2i = 1. / r2;
  double const coeff = 2. * alpha * std::numbers::inv_sqrtpi;
  double B;
  if (exact) {
    B = (erfc(ad) / r + coeff * ex) * r2i;
  } else {
    double const tt = 1. / (1. + 0.3275911 * ad);
    double const ep =
        tt * (0.254829592 +
              tt * (-0.284496736 + tt * (1.421413741 + tt * (-1.453152027 + tt * 1.061405429))));
    B = (ep / r + coeff) * ex * r2i;
  }
  double const alpsq = alpha * alpha;
  double const common = alpsq * coeff * ex;
  double const Cr = r2i * (3. * B + 2. * common);
  double const Dr = r2i * (5. * Cr + 4. * common * alpsq);
  double const mimj = mi.x * mj.x + mi.y * mj.y + mi.z * mj.z;
  double const mir = mi.x * dx + mi.y * dy + mi.z * dz;
  double const mjr = mj.x * dx + mj.y * dy + mj.z * dz;
  double const cd = mimj * Cr - mir * mjr * Dr;
  f[0] += prefactor * (cd * dx + (mi.x * mjr + mj.x * mir) * Cr);
  f[1] += prefactor * (cd * dy + (mi.y * mjr + mj.y * mir) * Cr);
  f[2] += prefactor * (cd * dz + (mi.z * mjr + mj.z * mir) * Cr);
  double const c1 = mjr * Cr;
  t[0] += prefactor * (-(mi.y * mj.z - mi.z * mj.y) * B + (mi.y * dz - mi.z * dy) * c1);
  t[1] += prefactor * (-(mi.z * mj.x - mi.x * mj.z) * B + (mi.z * dx - mi.x * dz) * c1);
  t[2] += prefactor * (-(mi.x * mj.y - mi.y * mj.x) * B + (mi.x * dy - mi.y * dx) * c1);
  if constexpr (EV) e += prefactor * (mimj * B - mir * mjr * Cr);
}

/** Dipolar pair terms over cells of edge r_cut/2 (NSUB = 2, dense systems),
 *  dense evaluation per block: one block of 256 threads per home cell.  The block enumerates the (home particle,
 *  candidate) pairs of the (2 NSUB + 1)^3 neighbour cells 256 at a time,
 *  appends the pairs inside r_cut to an LDS queue (wave ballots, one LDS
 *  atomic per wave) and drains the queue with every lane evaluating one
 *  pair, so the pair terms run without idle lanes however sparse the cells.
 *  Force and torque on each home particle accumulate in LDS (FP64 LDS
 *  atomics) and are added to the caller's arrays once. */
template <int NSUB, bool EV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_sr_dip_queue(CellGrid g, int const *__restrict__ offsets,
                                                      double4 const *__restrict__ spos,
                                                      double4 const *__restrict__ sdip,
                                                      int const *__restrict__ sidx,
                                                      double prefactor, double alpha, double rc2,
                                                      int exact, double *__restrict__ force,
                                                      double *__restrict__ torque,
                                                      double *partials) {
  constexpr int NT = 256, W = 2 * NSUB + 1, NN = W * W * W;
  constexpr int TPT = 2;                  // candidate pairs tested per thread and chunk
  constexpr int CH = NT * TPT;            // pairs per chunk
  constexpr int QCAP = 1024, TCAP = 256;  // queue entries; candidates per table window
  __shared__ double4 hp[NT], hm[NT];
  __shared__ double acc[NT * 6];
  __shared__ int2 queue[QCAP]; // (candidate index, home slot | neighbour << 16)
  __shared__ int2 ctab[TCAP];  // (candidate index, neighbour)
  __shared__ int nbeg[NN], npre[NN + 1];
  __shared__ signed char nsc[NN * 3]; // periodic shift of each neighbour in box lengths
  __shared__ int qn;
  __shared__ double red[16];
  int const cell = blockIdx.x;
  int const cz = cell % g.nc[2], cy = (cell / g.nc[2]) % g.nc[1], cx = cell / (g.nc[2] * g.nc[1]);
  int const beg = offsets[cell], end = offsets[cell + 1];
  if (beg == end) {
    if (EV && threadIdx.x == 0) partials[blockIdx.x * 8] = 0.;
    return;
  }
  for (int o = threadIdx.x; o < NN; o += NT) {
    int nx = cx + o / (W * W) - NSUB, ny = cy + (o / W) % W - NSUB, nz = cz + o % W - NSUB;
    int sx = 0, sy = 0, sz = 0;
    if (nx < 0) { nx += g.nc[0]; sx = -1; } else if (nx >= g.nc[0]) { nx -= g.nc[0]; sx = 1; }
    if (ny < 0) { ny += g.nc[1]; sy = -1; } else if (ny >= g.nc[1]) { ny -= g.nc[1]; sy = 1; }
    if (nz < 0) { nz += g.nc[2]; sz = -1; } else if (nz >= g.nc[2]) { nz -= g.nc[2]; sz = 1; }
    int const nb = (nx * g.nc[1] + ny) * g.nc[2] + nz;
    nbeg[o] = offsets[nb];
    npre[o + 1] = offsets[nb + 1] - offsets[nb];
    nsc[3 * o] = static_cast<signed char>(sx);
    nsc[3 * o + 1] = static_cast<signed char>(sy);
    nsc[3 * o + 2] = static_cast<signed char>(sz);
  }
  if (threadIdx.x == 0) qn = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    npre[0] = 0;
    for (int o = 0; o < NN; ++o) npre[o + 1] += npre[o];
  }
  __syncthreads();
  int const ncand = npre[NN];
  int const lane = static_cast<int>(threadIdx.x & 63);
  double e = 0.;
  for (int hb = beg; hb < end; hb += NT) {
    int const ni = min(NT, end - hb);
    if (static_cast<int>(threadIdx.x) < ni) {
      hp[threadIdx.x] = spos[hb + threadIdx.x];
      hm[threadIdx.x] = sdip[hb + threadIdx.x];
    }
    for (int x = threadIdx.x; x < ni * 6; x += NT) acc[x] = 0.;
    for (int kw = 0; kw < ncand; kw += TCAP) {
      int const nw = min(TCAP, ncand - kw);
      __syncthreads(); // table readers of the previous window are done
      for (int o = 0; o < NN; ++o) {
        int const s0 = npre[o];
        int const lo = max(s0, kw), hi = min(npre[o + 1], kw + nw);
        for (int k = lo + static_cast<int>(threadIdx.x); k < hi; k += NT)
          ctab[k - kw] = make_int2(nbeg[o] + (k - s0), o);
      }
      __syncthreads();
      int const total = ni * nw;
      bool const last_window = kw + TCAP >= ncand;
      // pair c = k ni + il (candidate k, home slot il) of this thread:
      // c = tid at the start, advanced by NT per test
      int const dk = NT / ni, dil = NT - dk * ni;
      int k = static_cast<int>(threadIdx.x) / ni;
      int il = static_cast<int>(threadIdx.x) - k * ni;
      for (int c0 = 0; c0 < total; c0 += CH) {
#pragma unroll
        for (int u = 0; u < TPT; ++u) {
          bool hit = false;
          int2 ent = make_int2(0, 0);
          if (k < nw) { // c < total
            int2 const cj = ctab[k];
            double4 const pj = spos[cj.x];
            double4 const pi = hp[il];
            double const dx = pi.x - (pj.x + nsc[3 * cj.y] * g.L[0]);
            double const dy = pi.y - (pj.y + nsc[3 * cj.y + 1] * g.L[1]);
            double const dz = pi.z - (pj.z + nsc[3 * cj.y + 2] * g.L[2]);
            double const r2 = dx * dx + dy * dy + dz * dz;
            hit = r2 < rc2 && r2 > 0.;
            ent = make_int2(cj.x, il | (cj.y << 16));
          }
          il += dil;
          k += dk;
          if (il >= ni) {
            il -= ni;
            ++k;
          }
          uint64_t const mask = __ballot(hit);
          int base = 0;
          if (lane == 0 && mask) base = atomicAdd(&qn, __popcll(mask));
          base = __shfl(base, 0, 64);
          if (hit) queue[base + __popcll(mask & ((1ull << lane) - 1ull))] = ent;
        }
        __syncthreads();
        int const qcount = qn;
        __syncthreads(); // every wave has read the count before it changes
        // at most CH entries join per chunk: drain while they still fit
        if (qcount > QCAP - CH || (last_window && c0 + CH >= total)) {
          for (int q = threadIdx.x; q < qcount; q += NT) {
            int2 const en = queue[q];
            int const il = en.y & 0xffff, o = en.y >> 16;
            double4 const pj = spos[en.x];
            double4 const pi = hp[il];
            double const dx = pi.x - (pj.x + nsc[3 * o] * g.L[0]);
            double const dy = pi.y - (pj.y + nsc[3 * o + 1] * g.L[1]);
            double const dz = pi.z - (pj.z + nsc[3 * o + 2] * g.L[2]);
            double const r2 = dx * dx + dy * dy + dz * dz;
            double f[3] = {0., 0., 0.}, t[3] = {0., 0., 0.};
            dip_pair<EV>(dx, dy, dz, r2, hm[il], sdip[en.x], prefactor, alpha, exact, f, t, e);
            atomicAdd(&acc[il * 6], f[0]);
            atomicAdd(&acc[il * 6 + 1], f[1]);
            atomicAdd(&acc[il * 6 + 2], f[2]);
            atomicAdd(&acc[il * 6 + 3], t[0]);
            atomicAdd(&acc[il * 6 + 4], t[1]);
            atomicAdd(&acc[il * 6 + 5], t[2]);
          }
          __syncthreads();
          if (threadIdx.x == 0) qn = 0;
          __syncthreads();
        }
      }
    }
    if (static_cast<int>(threadIdx.x) < ni) {
      int64_t const o = 3 * static_cast<int64_t>(sidx[hb + threadIdx.x]);
      double const *a = acc + threadIdx.x * 6;
      if (force) {
        force[o] += a[0];
        force[o + 1] += a[1];
        force[o + 2] += a[2];
      }
      if (torque) {
        torque[o] += a[3];
        torque[o + 1] += a[4];
        torque[o + 2] += a[5];
      }
    }
    __syncthreads(); // before the next home chunk restages hp, hm and acc
  }
  if constexpr (EV) {
    double v[1] = {e};
    block_sum<1>(v, red);
    if (threadIdx.x == 0) partials[blockIdx.x * 8] = v[0];
  }
}

/** Dipolar pair terms over cells of edge r_cut (NSUB = 1), dense evaluation
 *  per wave: one block of 256 threads per home cell.  Wave w takes the candidates w, w + 4, ... of
 *  each neighbour cell (one uniform load per candidate), its lanes the home
 *  particles; the pairs inside r_cut go to the wave's own LDS queue (ballot
 *  and a wave-uniform count, no atomics), which the wave drains itself with
 *  every lane evaluating one pair, so the pair terms run without idle
 *  lanes however sparse the cells and the loop needs no block barriers.
 *  Force and torque on each home particle accumulate in LDS (FP64 LDS
 *  atomics) and are added to the caller's arrays once. */
template <int NSUB, bool EV>
__global__ __launch_bounds__(256) void k_sr_dip_wave(
    CellGrid g, int const *__restrict__ offsets, double4 const *__restrict__ spos,
    double4 const *__restrict__ sdip, int const *__restrict__ sidx, double prefactor,
    double alpha, double rc2, int exact, double *__restrict__ force, double *__restrict__ torque,
    double *partials) {
  constexpr int NT = 256, NW = NT / 64, W = 2 * NSUB + 1, NN = W * W * W;
  constexpr int QW = 256;   // queue entries per wave
  __shared__ double4 hp[NT], hm[NT];
  __shared__ double acc[NT * 6];
  __shared__ int2 queue[NW][QW]; // (candidate index, home slot | neighbour << 16)
  __shared__ int nbeg[NN], ncnt[NN];
  __shared__ signed char nsc[NN * 3]; // periodic shift of each neighbour in box lengths
  __shared__ double red[16];
  int const cell = blockIdx.x;
  int const cz = cell % g.nc[2], cy = (cell / g.nc[2]) % g.nc[1], cx = cell / (g.nc[2] * g.nc[1]);
  int const beg = offsets[cell], end = offsets[cell + 1];
  if (beg == end) {
    if (EV && threadIdx.x == 0) partials[blockIdx.x * 8] = 0.;
    return;
  }
  for (int o = threadIdx.x; o < NN; o += NT) {
    int nx = cx + o / (W * W) - NSUB, ny = cy + (o / W) % W - NSUB, nz = cz + o % W - NSUB;
    int sx = 0, sy = 0, sz = 0;
    if (nx < 0) { nx += g.nc[0]; sx = -1; } else if (nx >= g.nc[0]) { nx -= g.nc[0]; sx = 1; }
    if (ny < 0) { ny += g.nc[1]; sy = -1; } else if (ny >= g.nc[1]) { ny -= g.nc[1]; sy = 1; }
    if (nz < 0) { nz += g.nc[2]; sz = -1; } else if (nz >= g.nc[2]) { nz -= g.nc[2]; sz = 1; }
    int const nb = (nx * g.nc[1] + ny) * g.nc[2] + nz;
    nbeg[o] = offsets[nb];
    ncnt[o] = offsets[nb + 1] - offsets[nb];
    nsc[3 * o] = static_cast<signed char>(sx);
    nsc[3 * o + 1] = static_cast<signed char>(sy);
    nsc[3 * o + 2] = static_cast<signed char>(sz);
  }
  int const lane = static_cast<int>(threadIdx.x & 63);
  int const wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  int2 *const wq = queue[wave];
  double e = 0.;
  for (int hb = beg; hb < end; hb += NT) {
    int const ni = min(NT, end - hb);
    __syncthreads(); // neighbour table written; previous chunk's readers done
    if (static_cast<int>(threadIdx.x) < ni) {
      hp[threadIdx.x] = spos[hb + threadIdx.x];
      hm[threadIdx.x] = sdip[hb + threadIdx.x];
    }
    for (int x = threadIdx.x; x < ni * 6; x += NT) acc[x] = 0.;
    __syncthreads();
    int const nr = (ni + 63) >> 6;
    int count = 0; // wave-uniform
    auto drain = [&]() {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int q = lane; q < count; q += 64) {
        int2 const en = wq[q];
        int const il = en.y & 0xffff, o = en.y >> 16;
        double4 const pj = spos[en.x];
        double4 const pi = hp[il];
        double const dx = pi.x - (pj.x + nsc[3 * o] * g.L[0]);
        double const dy = pi.y - (pj.y + nsc[3 * o + 1] * g.L[1]);
        double const dz = pi.z - (pj.z + nsc[3 * o + 2] * g.L[2]);
        double const r2 = dx * dx + dy * dy + dz * dz;
        double f[3] = {0., 0., 0.}, t[3] = {0., 0., 0.};
        dip_pair<EV>(dx, dy, dz, r2, hm[il], sdip[en.x], prefactor, alpha, exact, f, t, e);
        atomicAdd(&acc[il * 6], f[0]);
        atomicAdd(&acc[il * 6 + 1], f[1]);
        atomicAdd(&acc[il * 6 + 2], f[2]);
        atomicAdd(&acc[il * 6 + 3], t[0]);
        atomicAdd(&acc[il * 6 + 4], t[1]);
        atomicAdd(&acc[il * 6 + 5], t[2]);
      }
      __builtin_amdgcn_wave_barrier();
      count = 0;
    };
    for (int o = 0; o < NN; ++o) {
      int const jb = __builtin_amdgcn_readfirstlane(nbeg[o]);
      int const cnt = __builtin_amdgcn_readfirstlane(ncnt[o]);
      double const shx = nsc[3 * o] * g.L[0], shy = nsc[3 * o + 1] * g.L[1],
                   shz = nsc[3 * o + 2] * g.L[2];
      for (int t = wave; t < cnt; t += NW) {
        double4 const pj = spos[jb + t];
        double const qx = pj.x + shx, qy = pj.y + shy, qz = pj.z + shz;
        for (int r = 0; r < nr; ++r) {
          int const il = lane + 64 * r;
          double4 const pi = hp[min(il, NT - 1)];
          double const dx = pi.x - qx, dy = pi.y - qy, dz = pi.z - qz;
          double const r2 = dx * dx + dy * dy + dz * dz;
          bool const hit = il < ni && r2 < rc2 && r2 > 0.;
          uint64_t const mask = __ballot(hit);
          if (hit)
            wq[count + __popcll(mask & ((1ull << lane) - 1ull))] = make_int2(jb + t, il | (o << 16));
          count += __popcll(mask);
          // at most 64 entries join per test: drain while they still fit
          if (count > QW - 64) drain();
        }
      }
    }
    drain();
    __syncthreads(); // every wave's pairs are in acc
    if (static_cast<int>(threadIdx.x) < ni) {
      int64_t const o = 3 * static_cast<int64_t>(sidx[hb + threadIdx.x]);
      double const *a = acc + threadIdx.x * 6;
      if (force) {
        force[o] += a[0];
        force[o + 1] += a[1];
        force[o + 2] += a[2];
      }
      if (torque) {
        torque[o] += a[3];
        torque[o + 1] += a[4];
        torque[o + 2] += a[5];
      }
    }
  }
  if constexpr (EV) {
    double v[1] = {e};
    block_sum<1>(v, red);
    if (threadIdx.x == 0) partials[blockIdx.x * 8] = v[0];
  }
}

template <typename T> void sr_grow(T *&p, size_t &cap, size_t need, char const *what) {
  if (need <= cap) return;
  if (p) (void)hipFree(p);
  p = nullptr;
  size_t const n = need + need / 4 + 64;
  hip_check(hipMalloc(&p, sizeof(T) * n), what);
  cap = n;
}
} // namespace

struct ShortRange::Impl {
  int device = 0;
  int nc[3] = {0, 0, 0};
  double *mu2 = nullptr;
  double4 *sdip = nullptr;
  size_t cap_mu2 = 0, cap_sdip = 0;
  int *counts = nullptr, *offsets = nullptr, *p_cell = nullptr, *p_rank = nullptr,
      *sidx = nullptr;
  double4 *spos = nullptr;
  double *partials = nullptr;
  size_t cap_cells = 0, cap_off = 0, cap_pc = 0, cap_pr = 0, cap_si = 0, cap_sp = 0, cap_part = 0;
};

ShortRange::ShortRange(int device) : m(new Impl) { m->device = device; }

ShortRange::~ShortRange() {
  (void)hipSetDevice(m->device);
  for (void *p : {static_cast<void *>(m->counts), static_cast<void *>(m->offsets),
                  static_cast<void *>(m->p_cell), static_cast<void *>(m->p_rank),
                  static_cast<void *>(m->sidx), static_cast<void *>(m->spos),
                  static_cast<void *>(m->partials), static_cast<void *>(m->mu2),
                  static_cast<void *>(m->sdip)})
    if (p) (void)hipFree(p);
  delete m;
}

void ShortRange::cells(int out[3]) const {
  for (int d = 0; d < 3; ++d) out[d] = m->nc[d];
}

void ShortRange::compute(hipStream_t st, ShortRangeParams const &p, int64_t n, double const *pos,
                         double const *q, double *force, ShortRangeResult *out) {
  if (out) *out = ShortRangeResult{};
  if (!(p.r_cut > 0.)) throw std::invalid_argument("P3M: short range needs r_cut > 0");
  for (int d = 0; d < 3; ++d)
    if (!(p.box[d] > 0.)) throw std::invalid_argument("P3M: short range needs a box");
  if (n > INT32_MAX) throw std::invalid_argument("P3M: too many particles for the short range");
  if (n == 0) return;
  hip_check(hipSetDevice(m->device), "hipSetDevice");
  // cells of edge >= r_cut / nsub: two per r_cut when the box is wide enough
  // for 5 distinct neighbour cells, one when it allows 3, else all pairs
  if (p.subdivision < 0 || p.subdivision > 2)
    throw std::invalid_argument("P3M: short-range subdivision must be 0, 1 or 2");
  int nsub = (p.subdivision == 1) ? 1 : 2;
  // automatic: cells of r_cut when they hold 128 particles or more on
  // average, where the per-wave kernel keeps its lanes busy and tests fewer
  // cells (C2 at density 0.5: 9.7 against 10.6 ms for cells of r_cut/2)
  if (p.subdivision == 0 && !p.elc_images &&
      static_cast<double>(n) * p.r_cut * p.r_cut * p.r_cut / (p.box[0] * p.box[1] * p.box[2]) >=
          128.)
    nsub = 1;
  int nc[3];
  auto fit = [&](int s) {
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
      nc[d] = std::min(kSRMaxCells, static_cast<int>(std::floor(p.box[d] * s / p.r_cut)));
      if (nc[d] < 2 * s + 1) ok = false;
    }
    return ok;
  };
  if (!fit(nsub)) {
    if (nsub == 1 && p.subdivision == 0 && fit(2)) {
      nsub = 2; // too narrow for three cells of r_cut, wide enough for five of r_cut/2
    } else {
      nsub = 1;
      if (!fit(1)) {
        nsub = 0;
        nc[0] = nc[1] = nc[2] = 1;
      }
    }
  }
  CellGrid g;
  for (int d = 0; d < 3; ++d) {
    m->nc[d] = g.nc[d] = nc[d];
    g.L[d] = p.box[d];
    g.inv_cell[d] = nc[d] / p.box[d];
  }
  int const ncells = nc[0] * nc[1] * nc[2];
  sr_grow(m->counts, m->cap_cells, static_cast<size_t>(ncells), "hipMalloc cell counts");
  sr_grow(m->offsets, m->cap_off, static_cast<size_t>(ncells) + 1, "hipMalloc cell offsets");
  sr_grow(m->p_cell, m->cap_pc, static_cast<size_t>(n), "hipMalloc cell ids");
  sr_grow(m->p_rank, m->cap_pr, static_cast<size_t>(n), "hipMalloc cell ranks");
  sr_grow(m->sidx, m->cap_si, static_cast<size_t>(n), "hipMalloc cell order");
  sr_grow(m->spos, m->cap_sp, static_cast<size_t>(n), "hipMalloc cell positions");
  int const nb = static_cast<int>((n + 255) / 256);
  k_sr_fill<<<std::min((ncells + 255) / 256, 1024), 256, 0, st>>>(m->counts, ncells);
  k_sr_bin<<<nb, 256, 0, st>>>(n, pos, q, g, m->counts, m->p_cell, m->p_rank);
  k_sr_scan<<<1, 1024, 0, st>>>(m->counts, ncells, m->offsets);
  k_sr_scatter<<<nb, 256, 0, st>>>(n, pos, q, g, m->p_cell, m->p_rank, m->offsets, m->spos,
                                   m->sidx);
  double const rc2 = p.r_cut * p.r_cut;
  int blocks = ncells;
  int nq = 0;
  if (nsub == 0) {
    // all-pairs: the number of charged particles decides the grid
    hip_check(hipMemcpyAsync(&nq, m->offsets + ncells, sizeof(int), hipMemcpyDeviceToHost, st),
              "copy charged count");
    hip_check(hipStreamSynchronize(st), "sync short range");
    blocks = std::max(1, (nq + 255) / 256);
  }
  if (out) sr_grow(m->partials, m->cap_part, static_cast<size_t>(blocks) * 8, "hipMalloc partials");
  SrImages im{p.elc_images ? 1 : 0, p.elc_h, p.elc_sl, p.elc_dt, p.elc_db};
  auto launch = [&](auto EVc) {
    constexpr bool EV = decltype(EVc)::value;
    if (nsub == 2 && !im.on)
      k_sr_wave<2, EV><<<blocks, 256, 0, st>>>(g, m->offsets, m->spos, m->sidx, p.prefactor,
                                               p.alpha, rc2, force, m->partials);
    else if (nsub == 1 && !im.on)
      k_sr_wave<1, EV><<<blocks, 256, 0, st>>>(g, m->offsets, m->spos, m->sidx, p.prefactor,
                                               p.alpha, rc2, force, m->partials);
    else if (nsub == 2)
      k_sr_cells<2, 64, EV><<<blocks, 64, 0, st>>>(g, m->offsets, m->spos, m->sidx, p.prefactor,
                                                   p.alpha, rc2, force, m->partials, im);
    else if (nsub == 1)
      k_sr_cells<1, 256, EV><<<blocks, 256, 0, st>>>(g, m->offsets, m->spos, m->sidx,
                                                     p.prefactor, p.alpha, rc2, force,
                                                     m->partials, im);
    else
      k_sr_allpairs<256, EV><<<blocks, 256, 0, st>>>(nq, g, m->spos, m->sidx, p.prefactor,
                                                     p.alpha, rc2, force, m->partials, im);
  };
  if (out)
    launch(std::true_type{});
  else
    launch(std::false_type{});
  hip_check(hipGetLastError(), "short-range kernels");
  if (!out) return;
  std::vector<double> part(static_cast<size_t>(blocks) * 8);
  hip_check(hipMemcpyAsync(part.data(), m->partials, sizeof(double) * part.size(),
                           hipMemcpyDeviceToHost, st),
            "copy short-range partials");
  hip_check(hipStreamSynchronize(st), "sync short range");
  double s[7] = {0., 0., 0., 0., 0., 0., 0.};
  for (int b = 0; b < blocks; ++b)
    for (int k = 0; k < 7; ++k) s[k] += part[static_cast<size_t>(b) * 8 + k];
  // every pair was evaluated from both sides
  out->energy = 0.5 * s[0];
  double const v[3][3] = {{s[1], s[2], s[3]}, {s[2], s[4], s[5]}, {s[3], s[5], s[6]}};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) out->virial[3 * a + b] = 0.5 * v[a][b];
}

bool ShortRange::compute_dipolar(hipStream_t st, ShortRangeParams const &p, bool exact_erfc,
                                 int64_t n, double const *pos, double const *dip, double *force,
                                 double *torque, double *energy) {
  if (energy) *energy = 0.;
  if (!(p.r_cut > 0.)) throw std::invalid_argument("P3M: short range needs r_cut > 0");
  if (n > INT32_MAX) throw std::invalid_argument("P3M: too many particles for the short range");
  int nsub = (p.subdivision == 1) ? 1 : 2;
  int nc[3];
  auto fit = [&](int s) {
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
      nc[d] = std::min(kSRMaxCells, static_cast<int>(std::floor(p.box[d] * s / p.r_cut)));
      if (nc[d] < 2 * s + 1) ok = false;
    }
    return ok;
  };
  if (!fit(nsub)) {
    nsub = 1;
    if (!fit(1)) return false;
  }
  if (n == 0) return true;
  hip_check(hipSetDevice(m->device), "hipSetDevice");
  CellGrid g;
  for (int d = 0; d < 3; ++d) {
    m->nc[d] = g.nc[d] = nc[d];
    g.L[d] = p.box[d];
    g.inv_cell[d] = nc[d] / p.box[d];
  }
  int const ncells = nc[0] * nc[1] * nc[2];
  sr_grow(m->counts, m->cap_cells, static_cast<size_t>(ncells), "hipMalloc cell counts");
  sr_grow(m->offsets, m->cap_off, static_cast<size_t>(ncells) + 1, "hipMalloc cell offsets");
  sr_grow(m->p_cell, m->cap_pc, static_cast<size_t>(n), "hipMalloc cell ids");
  sr_grow(m->p_rank, m->cap_pr, static_cast<size_t>(n), "hipMalloc cell ranks");
  sr_grow(m->sidx, m->cap_si, static_cast<size_t>(n), "hipMalloc cell order");
  sr_grow(m->spos, m->cap_sp, static_cast<size_t>(n), "hipMalloc cell positions");
  sr_grow(m->mu2, m->cap_mu2, static_cast<size_t>(n), "hipMalloc dipole norms");
  sr_grow(m->sdip, m->cap_sdip, static_cast<size_t>(n), "hipMalloc cell dipoles");
  if (energy) sr_grow(m->partials, m->cap_part, static_cast<size_t>(ncells) * 8, "hipMalloc partials");
  int const nb = static_cast<int>((n + 255) / 256);
  k_sr_mu2<<<nb, 256, 0, st>>>(n, dip, m->mu2);
  k_sr_fill<<<std::min((ncells + 255) / 256, 1024), 256, 0, st>>>(m->counts, ncells);
  k_sr_bin<<<nb, 256, 0, st>>>(n, pos, m->mu2, g, m->counts, m->p_cell, m->p_rank);
  k_sr_scan<<<1, 1024, 0, st>>>(m->counts, ncells, m->offsets);
  k_sr_scatter<<<nb, 256, 0, st>>>(n, pos, m->mu2, g, m->p_cell, m->p_rank, m->offsets, m->spos,
                                   m->sidx);
  k_sr_sort_dip<<<nb, 256, 0, st>>>(n, m->offsets + ncells, m->sidx, dip, m->sdip);
  double const rc2 = p.r_cut * p.r_cut;
  int const ex = exact_erfc ? 1 : 0;
  auto launch = [&](auto EVc) {
    constexpr bool EV = decltype(EVc)::value;
    if (nsub == 2)
      k_sr_dip_queue<2, EV><<<ncells, 256, 0, st>>>(g, m->offsets, m->spos, m->sdip, m->sidx,
                                                    p.prefactor, p.alpha, rc2, ex, force, torque,
                                                    m->partials);
    // cells of r_cut: the per-wave kernel needs about half a wave of home
    // particles per cell; sparser cells run faster on the block queue
    else if (static_cast<double>(n) >= 32. * static_cast<double>(ncells))
      k_sr_dip_wave<1, EV><<<ncells, 256, 0, st>>>(g, m->offsets, m->spos, m->sdip, m->sidx,
                                                   p.prefactor, p.alpha, rc2, ex, force, torque,
                                                   m->partials);
    else
      k_sr_dip_queue<1, EV><<<ncells, 256, 0, st>>>(g, m->offsets, m->spos, m->sdip, m->sidx,
                                                    p.prefactor, p.alpha, rc2, ex, force, torque,
                                                    m->partials);
  };
  if (energy)
    launch(std::true_type{});
  else
    launch(std::false_type{});
  hip_check(hipGetLastError(), "dipolar short-range kernels");
  if (!energy) return true;
  std::vector<double> part(static_cast<size_t>(ncells) * 8);
  hip_check(hipMemcpyAsync(part.data(), m->partials, sizeof(double) * part.size(),
                           hipMemcpyDeviceToHost, st),
            "copy partials");
  hip_check(hipStreamSynchronize(st), "sync dipolar short range");
  double s = 0.;
  for (int c = 0; c < ncells; ++c) s += part[static_cast<size_t>(c) * 8];
  *energy = 0.5 * s;
  return true;
}

} // namespace espresso_amd
