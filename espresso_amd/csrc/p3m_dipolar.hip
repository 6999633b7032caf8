// Dipolar P3M on the device; see p3m_dipolar.hpp.
#include "p3m_dipolar.hpp"
#include "p3m_batch.hpp"
#include "p3m_kernels.hpp"
#include "p3m_short_range.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>

namespace espresso_amd {
namespace {
using namespace dev;
using C2 = double2;

constexpr int kDpBlock = 256;
constexpr int kDpLines = 8; // FFT lines per block
constexpr int kDpGatherBlock = 512;
constexpr int kDpPlaneBlock = 1024; // threads of the z/y plane transform
// bucket counters per bucket: block b counts into copy b % kDpCopies, which
// spreads the returning atomics on one bucket over several addresses
constexpr int kDpCopies = 8;
// dynamic LDS of the gather: gfx950's 160 KiB per workgroup less a margin
// for the kernel's static LDS
constexpr int kDpGatherLds = 160 * 1024 - 256;



/** Start of a solve: the three dipole meshes (blockIdx.y), the bucket
 *  counters and the error flag set to zero in one launch. */
__global__ __launch_bounds__(kDpBlock) void k_dp_prep(C2 *a0, C2 *a1, C2 *a2, int64_t m3,
                                                      int *counts, int nk, int *err) {
  int64_t const i = (int64_t)blockIdx.x * kDpBlock + threadIdx.x;
  C2 *a = blockIdx.y == 0 ? a0 : blockIdx.y == 1 ? a1 : a2;
  if (i < m3) a[i] = C2{0., 0.};
  if (blockIdx.y == 0 && i < (int64_t)nk * kDpCopies) counts[i] = 0;
  if (blockIdx.y == 0 && i == 0) *err = 0;
}

/** Weights and periodic index of the first stencil point of one particle
 *  (the Coulomb solver's bit-exact CA-mesh rule). */
template <int P>
__device__ __forceinline__ bool dp_weights(DevGeo const &g, double const x[3], int gi[3],
                                           double w[3][P]) {
  int nmp[3];
  double dist[3];
  bool const ok = particle_coordinates_rt(g, P, x, nmp, gi, dist);
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline_weights<P>(dist[d], w[d]);
  return ok;
}

/** dipole_assign (dp3m.cpp:157-193) over particles sorted into buckets of
 *  mesh cells: one
 *  block per bucket accumulates its particles in an LDS tile of E^3 points
 *  per component (LDS atomics) and adds the tile to the meshes with one
 *  device atomic per point.  The tile origin is the lowest stencil base of
 *  the bucket's particles (taken modulo M relative to its first particle);
 *  a particle whose stencil leaves the tile uses device atomics directly,
 *  so every placement is correct, and tile points wrap modulo M. */
template <int P>
__global__ __launch_bounds__(kDpBlock) void k_dp_spread_tile(
    int const *__restrict__ offsets, int const *__restrict__ counts,
    double const *__restrict__ pos, double const *__restrict__ dip, DevGeo g, int E, C2 *A0,
    C2 *A1, C2 *A2, int *err) {
  extern __shared__ double tile[]; // [3][E][E][E]
  __shared__ int org[3];
  int const beg = offsets[blockIdx.x], cnt = counts[blockIdx.x];
  if (cnt == 0) return;
  int const M = g.M[0];
  int const E3 = E * E * E;
  for (int e = threadIdx.x; e < 3 * E3; e += kDpBlock) tile[e] = 0.;
  if (threadIdx.x < 3) org[threadIdx.x] = INT_MAX;
  int gref[3];
  {
    double const x[3] = {pos[3 * beg], pos[3 * beg + 1], pos[3 * beg + 2]};
    int nmp[3];
    double dist[3];
    (void)particle_coordinates_rt(g, P, x, nmp, gref, dist);
  }
  auto rel = [M](int gi, int gr) {
    int d = gi - gr;
    if (d < -(M / 2)) d += M;
    if (d >= M - M / 2) d -= M;
    return d;
  };
  __syncthreads();
  for (int j = threadIdx.x; j < cnt; j += kDpBlock) {
    int64_t const i = beg + j;
    double const x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    int nmp[3], gi[3];
    double dist[3];
    if (!particle_coordinates_rt(g, P, x, nmp, gi, dist)) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d) atomicMin(&org[d], rel(gi[d], gref[d]));
  }
  __syncthreads();
  int const o[3] = {org[0] == INT_MAX ? 0 : org[0], org[1] == INT_MAX ? 0 : org[1],
                    org[2] == INT_MAX ? 0 : org[2]};
  int const a = static_cast<int>(threadIdx.x & 7);
  for (int j = threadIdx.x >> 3; j < cnt; j += kDpBlock / 8) {
    if (a >= P) continue;
    int64_t const i = beg + j;
    double const mu[3] = {dip[3 * i], dip[3 * i + 1], dip[3 * i + 2]};
    if (mu[0] == 0. && mu[1] == 0. && mu[2] == 0.) continue;
    double const x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    int gi[3];
    double w[3][P];
    if (!dp_weights<P>(g, x, gi, w)) {
      if (a == 0) atomicOr(err, 1);
      continue;
    }
    int r[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) r[d] = rel(gi[d], gref[d]) - o[d];
    double const wz = w[2][a];
    if (r[0] + P <= E && r[1] + P <= E && r[2] + P <= E) {
      for (int rr = 0; rr < P; ++rr)
        for (int b = 0; b < P; ++b) {
          int const t = ((r[0] + rr) * E + r[1] + b) * E + r[2] + a;
          double const wt = w[0][rr] * w[1][b] * wz;
          atomicAdd(&tile[t], wt * mu[0]);
          atomicAdd(&tile[E3 + t], wt * mu[1]);
          atomicAdd(&tile[2 * E3 + t], wt * mu[2]);
        }
    } else {
      int const iz = wrap_once(gi[2] + a, M);
      for (int rr = 0; rr < P; ++rr) {
        int const ix = wrap_once(gi[0] + rr, M);
        for (int b = 0; b < P; ++b) {
          int64_t const idx = ((int64_t)ix * M + wrap_once(gi[1] + b, M)) * M + iz;
          double const wt = w[0][rr] * w[1][b] * wz;
          atomicAdd(&A0[idx].x, wt * mu[0]);
          atomicAdd(&A1[idx].x, wt * mu[1]);
          atomicAdd(&A2[idx].x, wt * mu[2]);
        }
      }
    }
  }
  __syncthreads();
  int const b0 = ((gref[0] + o[0]) % M + M) % M, b1 = ((gref[1] + o[1]) % M + M) % M,
            b2 = ((gref[2] + o[2]) % M + M) % M;
  for (int e = threadIdx.x; e < E3; e += kDpBlock) {
    double const v0 = tile[e], v1 = tile[E3 + e], v2 = tile[2 * E3 + e];
    if (v0 == 0. && v1 == 0. && v2 == 0.) continue;
    int const tz = e % E, ty = (e / E) % E, tx = e / (E * E);
    int64_t const idx =
        ((int64_t)((b0 + tx) % M) * M + (b1 + ty) % M) * M + (b2 + tz) % M;
    if (v0 != 0.) atomicAdd(&A0[idx].x, v0);
    if (v1 != 0.) atomicAdd(&A1[idx].x, v1);
    if (v2 != 0.) atomicAdd(&A2[idx].x, v2);
  }
}

/** In-place complex FFT of the lines along `axis` of nmesh cubic meshes
 *  (blockIdx.y: mesh), kDpLines lines per block, unnormalised, forward
 *  sign -1 (inv: +1).  With rout, the real parts go to the real meshes
 *  rout[mesh] instead (the last pass of a back transform). */
__global__ __launch_bounds__(kDpBlock) void k_dp_fft(C2 *const *meshes, int M, int axis,
                                                     FftPlanDev plan, C2 const *__restrict__ tw,
                                                     bool inv, double *const *rout) {
  extern __shared__ double smem_d[];
  C2 *twl = reinterpret_cast<C2 *>(smem_d);
  int const ls = M + 1;
  C2 *X = twl + M;
  C2 *Y = X + kDpLines * ls;
  for (int m = threadIdx.x; m < M; m += blockDim.x) twl[m] = tw[m];
  C2 *data = meshes[blockIdx.y];
  int64_t const M2 = (int64_t)M * M;
  int const l0 = blockIdx.x * kDpLines;
  int const nl = min(kDpLines, static_cast<int>(M2 - l0));
  int64_t const stride = axis == 2 ? 1 : axis == 1 ? M : M2;
  auto base = [&](int l) -> int64_t {
    if (axis == 2) return (int64_t)l * M;
    int const hi = l / M, lo = l - hi * M;
    return axis == 1 ? (int64_t)hi * M2 + lo : (int64_t)hi * M + lo;
  };
  for (int e = threadIdx.x; e < kDpLines * M; e += blockDim.x) {
    int const l = e / M, j = e - l * M;
    X[l * ls + j] = (l < nl) ? data[base(l0 + l) + j * stride] : C2{0., 0.};
  }
  __syncthreads();
  C2 *res = lds_fft(X, Y, kDpLines, ls, plan, twl, inv);
  if (rout) { // last pass of a back transform: real parts only, compact
    double *r = rout[blockIdx.y];
    for (int e = threadIdx.x; e < nl * M; e += blockDim.x) {
      int const l = e / M, j = e - l * M;
      r[base(l0 + l) + j * stride] = res[l * ls + j].x;
    }
    return;
  }
  for (int e = threadIdx.x; e < nl * M; e += blockDim.x) {
    int const l = e / M, j = e - l * M;
    data[base(l0 + l) + j * stride] = res[l * ls + j];
  }
}

/** z and y transforms of one x plane (blockIdx.x) of nmesh cubic meshes
 *  (blockIdx.y) in LDS: z lines, a transpose, then y lines, for meshes whose
 *  plane fits twice in a workgroup's LDS.  Unnormalised, sign as k_dp_fft. */
__global__ __launch_bounds__(kDpPlaneBlock) void k_dp_fft_zy(C2 *const *meshes, int M,
                                                             FftPlanDev plan,
                                                             C2 const *__restrict__ tw, bool inv) {
  extern __shared__ double smem_d[];
  C2 *twl = reinterpret_cast<C2 *>(smem_d);
  int const ls = M + 1;
  int const MM = M * M;
  C2 *X = twl + M;
  C2 *Y = X + M * ls;
  for (int m = threadIdx.x; m < M; m += blockDim.x) twl[m] = tw[m];
  C2 *data = meshes[blockIdx.y] + (int64_t)blockIdx.x * MM;
  for (int e = threadIdx.x; e < MM; e += blockDim.x) {
    int const y = e / M, z = e - y * M;
    X[y * ls + z] = data[e];
  }
  __syncthreads();
  C2 *r1 = lds_fft(X, Y, M, ls, plan, twl, inv); // along z, line y
  C2 *o = (r1 == X) ? Y : X;
  for (int e = threadIdx.x; e < MM; e += blockDim.x) {
    int const y = e / M, z = e - y * M;
    o[z * ls + y] = r1[y * ls + z];
  }
  __syncthreads();
  C2 *r2 = lds_fft(o, r1, M, ls, plan, twl, inv); // along y, line z
  for (int e = threadIdx.x; e < MM; e += blockDim.x) {
    int const y = e / M, z = e - y * M;
    data[e] = r2[z * ls + y];
  }
}

/** k-space step (dp3m.cpp:270-311, 322-343, 369-389) and back-transform
 *  meshes in one pass over the mesh (Nyquist wave number 0): with
 *  the transformed dipole components A_d (table entries 0..2) and the
 *  integer wave numbers D, z = sum_d D_d A_d gives the energy partials
 *  G_e |z|^2.  With back != 0 the field meshes T_d = D_d G_e z overwrite
 *  A_d, and with forces != 0 the six force meshes F_ab = D_a D_b (-i G_f z)
 *  (xx xy xz yy yz zz, table entries 3..8) are written. */
__global__ __launch_bounds__(kDpBlock) void k_dp_kspace(C2 *const *T, C2 *const *F,
                                                        double const *__restrict__ Ge,
                                                        double const *__restrict__ Gf, int M,
                                                        int back, int forces,
                                                        double *partials) {
  __shared__ double red[16];
  int64_t const M3 = (int64_t)M * M * M;
  int64_t const k = (int64_t)blockIdx.x * kDpBlock + threadIdx.x;
  double e = 0.;
  if (k < M3) {
    int const iz = static_cast<int>(k % M);
    int const iy = static_cast<int>((k / M) % M);
    int const ix = static_cast<int>(k / ((int64_t)M * M));
    double const D[3] = {static_cast<double>(dop_shift(ix, M)),
                         static_cast<double>(dop_shift(iy, M)),
                         static_cast<double>(dop_shift(iz, M))};
    C2 const a0 = T[0][k], a1 = T[1][k], a2 = T[2][k];
    double const re = D[0] * a0.x + D[1] * a1.x + D[2] * a2.x;
    double const im = D[0] * a0.y + D[1] * a1.y + D[2] * a2.y;
    double const ge = Ge[k];
    e = ge * (re * re + im * im);
    if (back) {
      C2 const kt = C2{ge * re, ge * im};
#pragma unroll
      for (int d = 0; d < 3; ++d) T[d][k] = C2{D[d] * kt.x, D[d] * kt.y};
      if (forces) {
        double const gf = Gf[k];
        C2 const kf = C2{gf * im, -gf * re};
        int c = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = a; b < 3; ++b, ++c) {
            double const sab = D[a] * D[b];
            F[c][k] = C2{sab * kf.x, sab * kf.y};
          }
      }
    }
  }
  double v[1] = {e};
  block_sum<1>(v, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

/** AssignTorques / AssignForces (dp3m.cpp:195-244): torque -= mu x (pt E),
 *  force_a += pf sum_b mu_b H_ab with E and H interpolated from the real
 *  parts of the back-transformed meshes, over particles sorted into buckets
 *  of mesh cells.  One block per bucket stages the bucket's E^3 tile of the
 *  three field meshes (and the six force meshes when forces are wanted) in
 *  LDS, tile origin as in the spread; eight lanes per particle (lane a < P
 *  sums the stencil points with z offset a), combined by lane shuffles.  A
 *  particle whose stencil leaves the tile reads the meshes directly.
 *  Results go to the caller's row perm[i]. */
template <int P>
__global__ __launch_bounds__(kDpGatherBlock) void k_dp_gather_tile(
    int const *__restrict__ offsets, int const *__restrict__ counts,
    double const *__restrict__ pos, double const *__restrict__ dip, DevGeo g, int E,
    double const *const *R, double pt, double pf, int const *__restrict__ perm,
    double *__restrict__ force, double *__restrict__ torque) {
  extern __shared__ double tile[]; // [nm][E][E][E]: E_x E_y E_z H_xx .. H_zz
  __shared__ int org[3];
  int const beg = offsets[blockIdx.x], cnt = counts[blockIdx.x];
  if (cnt == 0) return;
  bool const wf = force != nullptr;
  int const nm = wf ? 9 : 3;
  int const M = g.M[0];
  int const E3 = E * E * E;
  if (threadIdx.x < 3) org[threadIdx.x] = INT_MAX;
  int gref[3];
  {
    double const x[3] = {pos[3 * beg], pos[3 * beg + 1], pos[3 * beg + 2]};
    int nmp[3];
    double dist[3];
    (void)particle_coordinates_rt(g, P, x, nmp, gref, dist);
  }
  auto rel = [M](int gi, int gr) {
    int d = gi - gr;
    if (d < -(M / 2)) d += M;
    if (d >= M - M / 2) d -= M;
    return d;
  };
  __syncthreads();
  for (int j = threadIdx.x; j < cnt; j += kDpGatherBlock) {
    int64_t const i = beg + j;
    double const x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    int nmp[3], gi[3];
    double dist[3];
    if (!particle_coordinates_rt(g, P, x, nmp, gi, dist)) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d) atomicMin(&org[d], rel(gi[d], gref[d]));
  }
  __syncthreads();
  int const o[3] = {org[0] == INT_MAX ? 0 : org[0], org[1] == INT_MAX ? 0 : org[1],
                    org[2] == INT_MAX ? 0 : org[2]};
  int const b0 = ((gref[0] + o[0]) % M + M) % M, b1 = ((gref[1] + o[1]) % M + M) % M,
            b2 = ((gref[2] + o[2]) % M + M) % M;
  for (int e = threadIdx.x; e < nm * E3; e += kDpGatherBlock) {
    int const c = e / E3, q = e - c * E3;
    int const tz = q % E, ty = (q / E) % E, tx = q / (E * E);
    int64_t const idx =
        ((int64_t)((b0 + tx) % M) * M + (b1 + ty) % M) * M + (b2 + tz) % M;
    tile[e] = R[c][idx];
  }
  __syncthreads();
  int const a = static_cast<int>(threadIdx.x & 7);
  // the eight lanes of a particle share j and run the loop together
  for (int j = threadIdx.x >> 3; j < cnt; j += kDpGatherBlock / 8) {
    int64_t const i = beg + j;
    double v[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.}; // E[3], H[6]
    double const mu[3] = {dip[3 * i], dip[3 * i + 1], dip[3 * i + 2]};
    bool const act = (mu[0] != 0. || mu[1] != 0. || mu[2] != 0.);
    if (act && a < P) {
      double const x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
      int gi[3];
      double w[3][P];
      if (dp_weights<P>(g, x, gi, w)) {
        int r[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) r[d] = rel(gi[d], gref[d]) - o[d];
        double const wz = w[2][a];
        if (r[0] + P <= E && r[1] + P <= E && r[2] + P <= E) {
          for (int rr = 0; rr < P; ++rr)
            for (int b = 0; b < P; ++b) {
              int const t = ((r[0] + rr) * E + r[1] + b) * E + r[2] + a;
              double const wt = w[0][rr] * w[1][b] * wz;
#pragma unroll
              for (int d = 0; d < 3; ++d) v[d] += wt * tile[d * E3 + t];
              if (wf)
#pragma unroll
                for (int d = 0; d < 6; ++d) v[3 + d] += wt * tile[(3 + d) * E3 + t];
            }
        } else {
          int const iz = wrap_once(gi[2] + a, M);
          for (int rr = 0; rr < P; ++rr) {
            int const ix = wrap_once(gi[0] + rr, M);
            for (int b = 0; b < P; ++b) {
              int64_t const idx = ((int64_t)ix * M + wrap_once(gi[1] + b, M)) * M + iz;
              double const wt = w[0][rr] * w[1][b] * wz;
#pragma unroll
              for (int d = 0; d < 3; ++d) v[d] += wt * R[d][idx];
              if (wf)
#pragma unroll
                for (int d = 0; d < 6; ++d) v[3 + d] += wt * R[3 + d][idx];
            }
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      v[k] += __shfl_xor(v[k], 1, 64);
      v[k] += __shfl_xor(v[k], 2, 64);
      v[k] += __shfl_xor(v[k], 4, 64);
    }
    if (!act || a != 0) continue;
    int64_t const ro = 3 * static_cast<int64_t>(perm[i]); // caller's row
    if (torque) {
      double const e0 = pt * v[0], e1 = pt * v[1], e2 = pt * v[2];
      torque[ro] -= mu[1] * e2 - mu[2] * e1;
      torque[ro + 1] -= mu[2] * e0 - mu[0] * e2;
      torque[ro + 2] -= mu[0] * e1 - mu[1] * e0;
    }
    if (wf) {
      double const *H = v + 3; // xx xy xz yy yz zz
      force[ro] += pf * (mu[0] * H[0] + mu[1] * H[1] + mu[2] * H[2]);
      force[ro + 1] += pf * (mu[0] * H[1] + mu[1] * H[3] + mu[2] * H[4]);
      force[ro + 2] += pf * (mu[0] * H[2] + mu[1] * H[4] + mu[2] * H[5]);
    }
  }
}

/** Counting sort of the particles into nbk^3 buckets of mesh cells (the
 *  spread and gather take one bucket per block): bucket key and rank per
 *  particle, and the block partials of sum |mu|^2 and of the total dipole
 *  (energy self term and surface term). */
__global__ __launch_bounds__(kDpBlock) void k_dp_key(int64_t n, double const *__restrict__ pos,
                                                     double const *__restrict__ dip, double L,
                                                     int nbk, int *__restrict__ counts,
                                                     int *__restrict__ key, int *__restrict__ rank,
                                                     double *__restrict__ partials) {
  __shared__ double red[4 * 16];
  int64_t const i = (int64_t)blockIdx.x * kDpBlock + threadIdx.x;
  double v[4] = {0., 0., 0., 0.};
  if (i < n) {
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double x = pos[3 * i + d];
      x -= L * floor(x / L);
      c[d] = isfinite(x) ? min(max(static_cast<int>(x / L * nbk), 0), nbk - 1) : 0;
    }
    int const k = (c[0] * nbk + c[1]) * nbk + c[2];
    int const copy = static_cast<int>(blockIdx.x) & (kDpCopies - 1);
    key[i] = k;
    rank[i] = atomicAdd(&counts[k * kDpCopies + copy], 1) * kDpCopies + copy;
    double const m0 = dip[3 * i], m1 = dip[3 * i + 1], m2 = dip[3 * i + 2];
    v[0] = m0 * m0 + m1 * m1 + m2 * m2;
    v[1] = m0;
    v[2] = m1;
    v[3] = m2;
  }
  // moment partials of the block: sum |mu|^2 and the total dipole
  block_sum<4>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) partials[blockIdx.x * 4 + k] = v[k];
}

/** Exclusive scan of the bucket counts (one block of 1024 threads): bucket
 *  start offsets[k] and size total[k], and each counter copy turned into the
 *  start of its run within the bucket. */
__global__ __launch_bounds__(1024) void k_dp_scan(int *__restrict__ counts, int nk,
                                                  int *__restrict__ offsets,
                                                  int *__restrict__ total) {
  __shared__ int part[1024];
  int const per = (nk + 1023) / 1024;
  int const b = threadIdx.x * per;
  int s = 0;
  for (int k = b; k < min(nk, b + per); ++k)
    for (int c = 0; c < kDpCopies; ++c) s += counts[k * kDpCopies + c];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int t = 0; t < 1024; ++t) {
      int const v = part[t];
      part[t] = acc;
      acc += v;
    }
  }
  __syncthreads();
  int acc = part[threadIdx.x];
  for (int k = b; k < min(nk, b + per); ++k) {
    offsets[k] = acc;
    int const a0 = acc;
    for (int c = 0; c < kDpCopies; ++c) {
      int const v = counts[k * kDpCopies + c];
      counts[k * kDpCopies + c] = acc;
      acc += v;
    }
    total[k] = acc - a0;
  }
}

__global__ __launch_bounds__(kDpBlock) void k_dp_scatter(int64_t n, double const *__restrict__ pos,
                                                         double const *__restrict__ dip,
                                                         int const *__restrict__ key,
                                                         int const *__restrict__ rank,
                                                         int const *__restrict__ starts,
                                                         double *__restrict__ spos,
                                                         double *__restrict__ sdip,
                                                         int *__restrict__ perm) {
  int64_t const i = (int64_t)blockIdx.x * kDpBlock + threadIdx.x;
  if (i >= n) return;
  int const r = rank[i];
  int64_t const s = static_cast<int64_t>(starts[key[i] * kDpCopies + (r & (kDpCopies - 1))]) +
                    r / kDpCopies;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    spos[3 * s + d] = pos[3 * i + d];
    sdip[3 * s + d] = dip[3 * i + d];
  }
  perm[s] = static_cast<int>(i);
}

/** Deterministic sum of the moment partials into out[4]. */
__global__ __launch_bounds__(kDpBlock) void k_dp_reduce4(double const *__restrict__ part, int nb,
                                                         double *__restrict__ out) {
  __shared__ double red[4 * 16];
  double v[4] = {0., 0., 0., 0.};
  for (int b = threadIdx.x; b < nb; b += kDpBlock)
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] += part[b * 4 + k];
  block_sum<4>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) out[k] = v[k];
}

/** Sums of |mu|^2 and of the dipole vector (block partials). */

/** Surface term torques (calc_surface_term, dp3m.cpp:451-517):
 *  torque -= pref mu x M_total. */
__global__ __launch_bounds__(kDpBlock) void k_dp_surface(int64_t n, double const *__restrict__ dip,
                                                         double const *__restrict__ mom,
                                                         double pref, double *torque) {
  int64_t const i = (int64_t)blockIdx.x * kDpBlock + threadIdx.x;
  if (i >= n || !torque) return;
  double const mx = mom[1], my = mom[2], mz = mom[3];
  double const m0 = dip[3 * i], m1 = dip[3 * i + 1], m2 = dip[3 * i + 2];
  torque[3 * i] -= pref * (m1 * mz - m2 * my);
  torque[3 * i + 1] -= pref * (m2 * mx - m0 * mz);
  torque[3 * i + 2] -= pref * (m0 * my - m1 * mx);
}

/** Real-space dipolar pair terms (DipolarP3M::pair_force / pair_energy,
 *  dp3m.hpp:160-235) of all pairs within r_cut with the minimum image; both
 *  orders of a pair, so each thread accumulates its particle's force and
 *  torque (the partner's torque of calc_opposing_force,
 *  forces_inline.hpp:159-171, is the other order's torque). */
template <bool EV>
__global__ __launch_bounds__(kDpBlock) void k_dp_pairs(int64_t n, double const *__restrict__ pos,
                                                       double const *__restrict__ dip, double L,
                                                       double prefactor, double alpha,
                                                       double r_cut, int exact_erfc,
                                                       double *__restrict__ force,
                                                       double *__restrict__ torque,
                                                       double *partials) {
  __shared__ double4 tp[kDpBlock], tm[kDpBlock];
  __shared__ double red[16];
  int64_t const i = (int64_t)blockIdx.x * kDpBlock + threadIdx.x;
  bool const has = i < n;
  double4 pi{0., 0., 0., 0.}, mi{0., 0., 0., 0.};
  auto fold = [&](double v) {
    double f = v - L * floor(v / L);
    return f >= L ? f - L : f;
  };
  if (has) {
    pi = double4{fold(pos[3 * i]), fold(pos[3 * i + 1]), fold(pos[3 * i + 2]), 0.};
    mi = double4{dip[3 * i], dip[3 * i + 1], dip[3 * i + 2], 0.};
  }
  bool const act = has && (mi.x != 0. || mi.y != 0. || mi.z != 0.);
  double const rc2 = r_cut * r_cut;
  double const coeff = 2. * alpha * std::numbers::inv_sqrtpi;
  double const alpsq = alpha * alpha;
  double f[3] = {0., 0., 0.}, t[3] = {0., 0., 0.}, e = 0.;
  for (int64_t jb = 0; jb < n; jb += kDpBlock) {
    __syncthreads();
    int64_t const j = jb + threadIdx.x;
    if (j < n) {
      tp[threadIdx.x] = double4{fold(pos[3 * j]), fold(pos[3 * j + 1]), fold(pos[3 * j + 2]), 0.};
      tm[threadIdx.x] = double4{dip[3 * j], dip[3 * j + 1], dip[3 * j + 2], 0.};
    }
    __syncthreads();
    if (!act) continue;
    int const cnt = static_cast<int>(min<int64_t>(kDpBlock, n - jb));
    for (int k = 0; k < cnt; ++k) {
      double4 const pj = tp[k], mj = tm[k];
      if (mj.x == 0. && mj.y == 0. && mj.z == 0.) continue;
      double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      dx -= L * rint(dx / L);
      dy -= L * rint(dy / L);
      dz -= L * rint(dz / L);
      double const r2 = dx * dx + dy * dy + dz * dz;
      if (!(r2 < rc2) || !(r2 > 0.)) continue;
      double const r = sqrt(r2);
      double const ad = alpha * r;
      double const ex = exp(-ad * ad);
      double const r2i = 1. / r2;
      double B;
      if (exact_erfc) {
        B = (erfc(ad) / r + coeff * ex) * r2i;
      } else {
        double const tt = 1. / (1. + 0.3275911 * ad);
        double const ep =
            tt * (0.254829592 +
                  tt * (-0.284496736 + tt * (1.421413741 + tt * (-1.453152027 + tt * 1.061405429))));
        B = (ep / r + coeff) * ex * r2i;
      }
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
  }
  if (act) {
    if (force) {
      force[3 * i] += f[0];
      force[3 * i + 1] += f[1];
      force[3 * i + 2] += f[2];
    }
    if (torque) {
      torque[3 * i] += t[0];
      torque[3 * i + 1] += t[1];
      torque[3 * i + 2] += t[2];
    }
  }
  if constexpr (EV) {
    double v[1] = {e};
    block_sum<1>(v, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
  }
}

void dp_free(void *p) {
  if (p) (void)hipFree(p);
}

double sinc_pi(double x) {
  if (std::fabs(x) < 1e-12) return 1.;
  double const px = std::numbers::pi * x;
  return std::sin(px) / px;
}
int dshift(int j, int M, bool zero_nyquist) {
  if (zero_nyquist && 2 * j == M) return 0;
  return (2 * j < M) ? j : j - M;
}
} // namespace

// ---------------------------------------------------------------------------
// error estimates (dp3m.cpp:684-820)
// ---------------------------------------------------------------------------
double dp3m_real_space_error(double box, double r_cut_iL, int n_c_part, double sum_mu2,
                             double alpha_L) {
  double const rc = r_cut_iL * box;
  double const rc2 = rc * rc;
  double const rc4 = rc2 * rc2;
  double const a2 = alpha_L * alpha_L / (box * box);
  double const c = sum_mu2 * std::exp(-a2 * rc2);
  double const cc = 4. * a2 * a2 * rc2 * rc2 + 6. * a2 * rc2 + 3.;
  double const dc = 8. * a2 * a2 * a2 * rc2 * rc2 * rc2 + 20. * a2 * a2 * rc4 + 30. * a2 * rc2 + 15.;
  double const con =
      1. / std::sqrt(box * box * box * a2 * a2 * rc * rc4 * rc4 * static_cast<double>(n_c_part));
  return c * con *
         std::sqrt((13. / 6.) * cc * cc + (2. / 15.) * dc * dc - (13. / 15.) * cc * dc);
}

double dp3m_k_space_error(double box, int mesh, int cao, int n_c_part, double sum_mu2,
                          double alpha_L) {
  double const mesh_i = 1. / mesh;
  double const f1 = (std::numbers::pi / alpha_L) * (std::numbers::pi / alpha_L);
  int const h = mesh / 2;
  std::vector<double> cs(static_cast<size_t>(2 * h));
  for (int n = -h; n < h; ++n) cs[static_cast<size_t>(n + h)] = analytic_cotangent_sum(cao, n, mesh_i);
  double he_q = 0.;
  for (int x = -h; x < h; ++x)
    for (int y = -h; y < h; ++y)
      for (int z = -h; z < h; ++z) {
        if (x == 0 && y == 0 && z == 0) continue;
        double const n2 = static_cast<double>(x * x + y * y + z * z);
        double const c = cs[static_cast<size_t>(x + h)] * cs[static_cast<size_t>(y + h)] *
                         cs[static_cast<size_t>(z + h)];
        double const ex = std::exp(-f1 * n2);
        double const U2 = std::pow(sinc_pi(x * mesh_i) * sinc_pi(y * mesh_i) * sinc_pi(z * mesh_i),
                                   2 * cao);
        double const alias1 = ex * ex * n2;
        double const alias2 = U2 * ex * n2 * n2 * n2 / n2;
        double const d = alias1 - (alias2 / c) * (alias2 / c) / (n2 * n2 * n2);
        if (d > 0. && std::fabs(d / alias1) > 1e-14) he_q += d;
      }
  return 8. * std::numbers::pi * std::numbers::pi / 3. * sum_mu2 *
         std::sqrt(he_q / n_c_part) / (box * box * box * box);
}

namespace {
double dp3m_rtbisection(double box, double r_cut_iL, int n, double mu2, double x1, double x2,
                        double xacc, double accuracy) {
  double const cst = accuracy / std::numbers::sqrt2;
  double const f1 = dp3m_real_space_error(box, r_cut_iL, n, mu2, x1) - cst;
  double const f2 = dp3m_real_space_error(box, r_cut_iL, n, mu2, x2) - cst;
  if (f1 * f2 >= 0.)
    throw std::runtime_error("Root must be bracketed for bisection in dp3m_rtbisection");
  double dx, rtb;
  if (f1 < 0.) {
    dx = x2 - x1;
    rtb = x1;
  } else {
    dx = x1 - x2;
    rtb = x2;
  }
  for (int j = 1; j <= 40; ++j) {
    dx *= 0.5;
    double const xmid = rtb + dx;
    double const fmid = dp3m_real_space_error(box, r_cut_iL, n, mu2, xmid) - cst;
    if (fmid <= 0.) rtb = xmid;
    if (std::fabs(dx) < xacc || fmid == 0.) return rtb;
  }
  throw std::runtime_error("Too many bisections in dp3m_rtbisection");
}
} // namespace

void dp3m_calculate_accuracy(double box, int mesh, int cao, double r_cut_iL, int n_c_part,
                             double sum_mu2, double accuracy, double out[4]) {
  double rs = dp3m_real_space_error(box, r_cut_iL, n_c_part, sum_mu2, 0.001);
  double alpha_L;
  if (std::numbers::sqrt2 * rs > accuracy)
    alpha_L = dp3m_rtbisection(box, r_cut_iL, n_c_part, sum_mu2, 0.0001 * box, 5. * box, 0.0001,
                               accuracy);
  else
    alpha_L = 0.1;
  rs = dp3m_real_space_error(box, r_cut_iL, n_c_part, sum_mu2, alpha_L);
  double const ks = dp3m_k_space_error(box, mesh, cao, n_c_part, sum_mu2, alpha_L);
  out[0] = std::hypot(rs, ks);
  out[1] = rs;
  out[2] = ks;
  out[3] = alpha_L;
}

// ---------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------
struct DipolarSolver::Impl {
  int device = 0;
  hipStream_t st = nullptr;
  bool configured = false;
  DipolarParams p;
  DevGeo g{};
  FftPlanDev plan{};
  int M = 0;
  C2 *tw = nullptr, *twi = nullptr;
  C2 *mesh[9] = {};       // A0..A2 (T0..T2), F0..F5
  C2 **d_ptrs = nullptr;  // device pointer table: A[3], F[6]
  double *rmesh[9] = {};  // real parts of the back-transformed meshes
  double **d_rptrs = nullptr; // device pointer table of rmesh
  double *Ge = nullptr, *Gf = nullptr;
  double *part = nullptr;
  size_t cap_part = 0;
  int *err = nullptr;
  double self_sum = -1.; // sum_k G_e U2 |d_op|^2
  std::unique_ptr<ShortRange> sr; // cell lists of the real space
  int subdivision = 0;
  double *mom = nullptr; // device: sum |mu|^2, total dipole
  // particles sorted by blocks of mesh cells
  double *spos = nullptr, *sdip = nullptr;
  int *key = nullptr, *rank = nullptr, *perm = nullptr, *counts = nullptr, *offsets = nullptr,
      *total = nullptr;
  int64_t cap_sort = 0;
  int cap_keys = 0;
};

DipolarSolver::DipolarSolver(int device) : m(std::make_unique<Impl>()) {
  m->device = device;
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipStreamCreate(&m->st), "hipStreamCreate");
  hip_check(hipMalloc(&m->err, sizeof(int)), "hipMalloc dp3m flag");
  hip_check(hipMalloc(&m->d_ptrs, sizeof(C2 *) * 12), "hipMalloc dp3m table");
  hip_check(hipMalloc(&m->d_rptrs, sizeof(double *) * 9), "hipMalloc dp3m table");
  hip_check(hipMalloc(&m->mom, sizeof(double) * 4), "hipMalloc dp3m moments");
}

DipolarSolver::~DipolarSolver() {
  (void)hipSetDevice(m->device);
  if (m->st) (void)hipStreamSynchronize(m->st);
  for (auto *x : m->mesh) dp_free(x);
  for (auto *x : m->rmesh) dp_free(x);
  dp_free(m->tw);
  dp_free(m->twi);
  dp_free(m->d_ptrs);
  dp_free(m->d_rptrs);
  dp_free(m->Ge);
  dp_free(m->Gf);
  dp_free(m->part);
  dp_free(m->err);
  for (void *x : {static_cast<void *>(m->mom), static_cast<void *>(m->spos),
                  static_cast<void *>(m->sdip), static_cast<void *>(m->key),
                  static_cast<void *>(m->rank), static_cast<void *>(m->perm),
                  static_cast<void *>(m->counts), static_cast<void *>(m->offsets),
                  static_cast<void *>(m->total)})
    dp_free(x);
  if (m->st) (void)hipStreamDestroy(m->st);
}

bool DipolarSolver::configured() const { return m->configured; }
DipolarParams const &DipolarSolver::params() const { return m->p; }
hipStream_t DipolarSolver::stream() const { return m->st; }

void DipolarSolver::set_short_range_subdivision(int s) {
  if (s < 0 || s > 2)
    throw std::invalid_argument("DipolarP3M: short-range subdivision must be 0, 1 or 2");
  m->subdivision = s;
}

void DipolarSolver::configure(DipolarParams const &p) {
  if (p.mesh < 2 || p.mesh % 2) throw std::runtime_error("DipolarP3M: mesh must be even and >= 2");
  if (p.cao < 1 || p.cao > 7) throw std::runtime_error("DipolarP3M: cao must be in [1, 7]");
  if (!(p.alpha > 0.) || !(p.r_cut > 0.) || !(p.box > 0.))
    throw std::runtime_error("DipolarP3M: alpha, r_cut and the box must be positive");
  hip_check(hipSetDevice(m->device), "hipSetDevice");
  hip_check(hipStreamSynchronize(m->st), "sync");
  int const M = p.mesh;
  Vec3d const box{p.box, p.box, p.box};
  P3MParams prm(false, p.epsilon, p.r_cut, Vec3i{M, M, M},
                Vec3d{p.mesh_off[0], p.mesh_off[1], p.mesh_off[2]}, p.cao, p.alpha, p.accuracy);
  prm.recalc_a_ai_cao_cut(box);
  LocalMesh lm;
  lm.calc_local_ca_mesh(prm, box, p.skin);
  DevGeo &g = m->g;
  g = DevGeo{};
  for (int k = 0; k < 3; ++k) {
    g.M[k] = M;
    g.dim[k] = lm.dim[k];
    g.marg[k] = lm.margin[2 * k];
    g.ld_pos[k] = lm.ld_pos[k];
    g.ai[k] = prm.ai[k];
    g.box_l[k] = p.box;
  }
  g.pos_shift = std::floor((p.cao - 1) / 2.0) - (p.cao % 2) / 2.0;
  g.conv = 0;
  FftPlan1D const fp = make_fft_plan(M);
  m->plan.n = fp.n;
  m->plan.nstages = fp.nstages;
  for (int s = 0; s < 16; ++s) m->plan.radix[s] = fp.radix[s];
  // buffers
  if (M != m->M) {
    for (auto *&x : m->mesh) {
      dp_free(x);
      x = nullptr;
    }
    for (auto *&x : m->rmesh) {
      dp_free(x);
      x = nullptr;
    }
    dp_free(m->tw);
    dp_free(m->Ge);
    dp_free(m->Gf);
    size_t const M3 = static_cast<size_t>(M) * M * M;
    for (auto *&x : m->mesh) hip_check(hipMalloc(&x, sizeof(C2) * M3), "hipMalloc dp3m mesh");
    for (auto *&x : m->rmesh)
      hip_check(hipMalloc(&x, sizeof(double) * M3), "hipMalloc dp3m real mesh");
    hip_check(hipMemcpy(m->d_rptrs, m->rmesh, sizeof(m->rmesh), hipMemcpyHostToDevice),
              "copy table");
    hip_check(hipMalloc(&m->Ge, sizeof(double) * M3), "hipMalloc dp3m G");
    hip_check(hipMalloc(&m->Gf, sizeof(double) * M3), "hipMalloc dp3m G");
    hip_check(hipMalloc(&m->tw, sizeof(C2) * M), "hipMalloc dp3m twiddles");
    std::vector<C2> tw(static_cast<size_t>(M));
    for (int k = 0; k < M; ++k) {
      long double const ang = -2.0L * std::numbers::pi_v<long double> * k / M;
      tw[static_cast<size_t>(k)] =
          C2{static_cast<double>(std::cos(ang)), static_cast<double>(std::sin(ang))};
    }
    hip_check(hipMemcpy(m->tw, tw.data(), sizeof(C2) * M, hipMemcpyHostToDevice), "copy tw");
    C2 *tab[12] = {m->mesh[0], m->mesh[1], m->mesh[2], m->mesh[3], m->mesh[4], m->mesh[5],
                   m->mesh[6], m->mesh[7], m->mesh[8], nullptr, nullptr, nullptr};
    hip_check(hipMemcpy(m->d_ptrs, tab, sizeof(tab), hipMemcpyHostToDevice), "copy table");
    m->M = M;
  }
  // influence functions (influence_function_dipolar.hpp:40-140), P3M_BRILLOUIN = 0
  {
    size_t const M3 = static_cast<size_t>(M) * M * M;
    std::vector<double> ge(M3, 0.), gf(M3, 0.);
    double const alpha_L = p.alpha * p.box;
    double const f2 = (std::numbers::pi / alpha_L) * (std::numbers::pi / alpha_L);
    double const pref = static_cast<double>(M) * M * M * 2. / (p.box * p.box);
    int const half = M / 2;
    std::vector<double> sn(static_cast<size_t>(M));
    for (int j = 0; j < M; ++j)
      sn[static_cast<size_t>(j)] = sinc_pi(static_cast<double>(dshift(j, M, false)) / M);
    size_t idx = 0;
    for (int x = 0; x < M; ++x)
      for (int y = 0; y < M; ++y)
        for (int z = 0; z < M; ++z, ++idx) {
          if (x % half == 0 && y % half == 0 && z % half == 0) continue;
          double const s[3] = {static_cast<double>(dshift(x, M, false)),
                               static_cast<double>(dshift(y, M, false)),
                               static_cast<double>(dshift(z, M, false))};
          double const d[3] = {static_cast<double>(dshift(x, M, true)),
                               static_cast<double>(dshift(y, M, true)),
                               static_cast<double>(dshift(z, M, true))};
          double const nm2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
          double const sz = std::pow(sn[static_cast<size_t>(x)] * sn[static_cast<size_t>(y)] *
                                         sn[static_cast<size_t>(z)],
                                     2 * p.cao);
          double const dn = d[0] * s[0] + d[1] * s[1] + d[2] * s[2];
          double const d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          double const et = f2 * nm2;
          double num2 = 0., num3 = 0.;
          if (et < 30.) {
            double const f3 = sz * std::exp(-et) / nm2;
            num2 = f3 * dn * dn;
            num3 = f3 * dn * dn * dn;
          }
          double const den = sz * sz;
          ge[idx] = pref * num2 / (d2 * d2 * den);
          gf[idx] = pref * num3 / (d2 * d2 * d2 * den);
          if (!std::isfinite(ge[idx])) ge[idx] = 0.;
          if (!std::isfinite(gf[idx])) gf[idx] = 0.;
        }
    hip_check(hipMemcpy(m->Ge, ge.data(), sizeof(double) * M3, hipMemcpyHostToDevice), "copy G");
    hip_check(hipMemcpy(m->Gf, gf.data(), sizeof(double) * M3, hipMemcpyHostToDevice), "copy G");
    // self-energy sum of the energy correction (influence_function_dipolar.hpp:160-218)
    double se = 0.;
    idx = 0;
    for (int x = 0; x < M; ++x)
      for (int y = 0; y < M; ++y)
        for (int z = 0; z < M; ++z, ++idx) {
          if (x % half == 0 && y % half == 0 && z % half == 0) continue;
          double const o[3] = {static_cast<double>(dshift(x, M, false)) / M,
                               static_cast<double>(dshift(y, M, false)) / M,
                               static_cast<double>(dshift(z, M, false)) / M};
          double u2 = 0.;
          for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b)
              for (int c = -1; c <= 1; ++c)
                u2 += std::pow(sinc_pi(o[0] + a * M) * sinc_pi(o[1] + b * M) *
                                   sinc_pi(o[2] + c * M),
                               2 * p.cao);
          double const d[3] = {static_cast<double>(dshift(x, M, true)),
                               static_cast<double>(dshift(y, M, true)),
                               static_cast<double>(dshift(z, M, true))};
          se += ge[idx] * u2 * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        }
    m->self_sum = se;
  }
  int const lds = static_cast<int>(sizeof(C2) * (M + 2 * kDpLines * (M + 1)));
  if (lds > 64 * 1024)
    hip_check(hipFuncSetAttribute(reinterpret_cast<void const *>(k_dp_fft),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds),
              "hipFuncSetAttribute");
  size_t const lds_plane = sizeof(C2) * (M + 2 * static_cast<size_t>(M) * (M + 1));
  if (lds_plane > 64 * 1024 && lds_plane <= static_cast<size_t>(kDpGatherLds))
    hip_check(hipFuncSetAttribute(reinterpret_cast<void const *>(k_dp_fft_zy),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds_plane)),
              "hipFuncSetAttribute");
  m->p = p;
  m->configured = true;
}

namespace {

template <int P> void launch_spread(hipStream_t st, int nk, int E, int const *offsets,
                                    int const *counts, double const *pos, double const *dip,
                                    DevGeo const &g, C2 *const *mesh, int *err) {
  size_t const lds = sizeof(double) * 3 * static_cast<size_t>(E) * E * E;
  k_dp_spread_tile<P><<<nk, kDpBlock, lds, st>>>(offsets, counts, pos, dip, g, E, mesh[0],
                                                 mesh[1], mesh[2], err);
}
template <int P> void launch_gather(hipStream_t st, int nk, int E, int const *offsets,
                                    int const *counts, double const *pos, double const *dip,
                                    DevGeo const &g, double const *const *R, double pt,
                                    double pf, int const *perm, double *force, double *torque) {
  static bool const raised = [] {
    hip_check(hipFuncSetAttribute(reinterpret_cast<void const *>(k_dp_gather_tile<P>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kDpGatherLds),
              "hipFuncSetAttribute");
    return true;
  }();
  (void)raised;
  size_t const lds = sizeof(double) * (force ? 9 : 3) * static_cast<size_t>(E) * E * E;
  if (lds > static_cast<size_t>(kDpGatherLds))
    throw std::runtime_error("DipolarP3M: gather tile exceeds the LDS of a workgroup");
  k_dp_gather_tile<P><<<nk, kDpGatherBlock, lds, st>>>(offsets, counts, pos, dip, g, E, R,
                                                       pt, pf, perm, force, torque);
}
} // namespace

void DipolarSolver::solve(int64_t n, double const *pos, double const *dip, double *force,
                          double *torque, double *energy, bool sync) {
  if (!m->configured) throw std::runtime_error("DipolarP3M solver used before configure()");
  if (energy) *energy = 0.;
  if (n <= 0) return;
  hip_check(hipSetDevice(m->device), "hipSetDevice");
  hipStream_t const st = m->st;
  DipolarParams const &p = m->p;
  int const M = m->M;
  int64_t const M3 = static_cast<int64_t>(M) * M * M;
  int const nb = static_cast<int>((n + kDpBlock - 1) / kDpBlock);
  int const kb = static_cast<int>((M3 + kDpBlock - 1) / kDpBlock);
  size_t const need = std::max(static_cast<size_t>(kb), static_cast<size_t>(nb) * 4);
  if (need > m->cap_part) {
    hip_check(hipStreamSynchronize(st), "sync");
    dp_free(m->part);
    hip_check(hipMalloc(&m->part, sizeof(double) * need), "hipMalloc dp3m partials");
    m->cap_part = need;
  }
  // particles sorted into buckets of S^3 mesh cells for the spread and
  // gather; the spread's LDS tile of E = ceil(M / nbk) + cao points per edge
  // and component stays within 64 KiB
  int nbk = 1, E = 0;
  for (int S : {8, 4, 2, 1}) {
    nbk = std::max(1, M / S);
    E = (M + nbk - 1) / nbk + p.cao;
    if (3 * E * E * E * static_cast<int>(sizeof(double)) <= 65536) break;
  }
  int const nk = nbk * nbk * nbk;
  if (n > m->cap_sort || nk > m->cap_keys) {
    hip_check(hipStreamSynchronize(st), "sync");
    if (n > m->cap_sort) {
      for (void *x : {static_cast<void *>(m->spos), static_cast<void *>(m->sdip),
                      static_cast<void *>(m->key), static_cast<void *>(m->rank),
                      static_cast<void *>(m->perm)})
        dp_free(x);
      hip_check(hipMalloc(&m->spos, sizeof(double) * 3 * n), "hipMalloc dp3m sort");
      hip_check(hipMalloc(&m->sdip, sizeof(double) * 3 * n), "hipMalloc dp3m sort");
      hip_check(hipMalloc(&m->key, sizeof(int) * n), "hipMalloc dp3m sort");
      hip_check(hipMalloc(&m->rank, sizeof(int) * n), "hipMalloc dp3m sort");
      hip_check(hipMalloc(&m->perm, sizeof(int) * n), "hipMalloc dp3m sort");
      m->cap_sort = n;
    }
    if (nk > m->cap_keys) {
      dp_free(m->counts);
      dp_free(m->offsets);
      dp_free(m->total);
      hip_check(hipMalloc(&m->counts, sizeof(int) * nk * kDpCopies), "hipMalloc dp3m sort");
      hip_check(hipMalloc(&m->offsets, sizeof(int) * nk), "hipMalloc dp3m sort");
      hip_check(hipMalloc(&m->total, sizeof(int) * nk), "hipMalloc dp3m sort");
      m->cap_keys = nk;
    }
  }
  k_dp_prep<<<dim3(static_cast<unsigned>(std::max<int64_t>(
                       kb, (static_cast<int64_t>(nk) * kDpCopies + kDpBlock - 1) / kDpBlock)),
                   3),
              kDpBlock, 0, st>>>(m->mesh[0], m->mesh[1], m->mesh[2], M3, m->counts, nk, m->err);
  // bucket keys, and the moments (sum |mu|^2, total dipole) on the device
  k_dp_key<<<nb, kDpBlock, 0, st>>>(n, pos, dip, p.box, nbk, m->counts, m->key, m->rank,
                                    m->part);
  k_dp_reduce4<<<1, kDpBlock, 0, st>>>(m->part, nb, m->mom);
  k_dp_scan<<<1, 1024, 0, st>>>(m->counts, nk, m->offsets, m->total);
  k_dp_scatter<<<nb, kDpBlock, 0, st>>>(n, pos, dip, m->key, m->rank, m->counts, m->spos,
                                        m->sdip, m->perm);
  double const *sp = m->spos, *sd = m->sdip;
  int const *perm = m->perm;
  double const V = p.box * p.box * p.box;
  double const wn = 2. * std::numbers::pi / p.box;
  double const dipole_prefac = p.prefactor / static_cast<double>(M3);
  {
    switch (p.cao) {
    case 1: launch_spread<1>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->mesh, m->err); break;
    case 2: launch_spread<2>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->mesh, m->err); break;
    case 3: launch_spread<3>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->mesh, m->err); break;
    case 4: launch_spread<4>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->mesh, m->err); break;
    case 5: launch_spread<5>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->mesh, m->err); break;
    case 6: launch_spread<6>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->mesh, m->err); break;
    default: launch_spread<7>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->mesh, m->err); break;
    }
    size_t const lds = sizeof(C2) * (M + 2 * kDpLines * (M + 1));
    int const lines = static_cast<int>((static_cast<int64_t>(M) * M + kDpLines - 1) / kDpLines);
    size_t const lds_plane = sizeof(C2) * (M + 2 * static_cast<size_t>(M) * (M + 1));
    bool const plane = lds_plane <= static_cast<size_t>(kDpGatherLds);
    // rout: real meshes for the real parts of the last (x) pass, or null
    auto fft = [&](C2 *const *table, int nmesh, bool inv, double *const *rout) {
      if (plane) {
        k_dp_fft_zy<<<dim3(M, nmesh), kDpPlaneBlock, lds_plane, st>>>(table, M, m->plan, m->tw,
                                                                       inv);
        k_dp_fft<<<dim3(lines, nmesh), kDpBlock, lds, st>>>(table, M, 0, m->plan, m->tw, inv,
                                                            rout);
      } else {
        for (int axis = 2; axis >= 0; --axis)
          k_dp_fft<<<dim3(lines, nmesh), kDpBlock, lds, st>>>(table, M, axis, m->plan, m->tw,
                                                              inv, axis == 0 ? rout : nullptr);
      }
    };
    // device table: [0..2] A (field meshes after the k-space step), [3..8] F
    fft(m->d_ptrs, 3, false, nullptr);
    bool const wf = force != nullptr, wt = torque != nullptr;
    k_dp_kspace<<<kb, kDpBlock, 0, st>>>(m->d_ptrs, m->d_ptrs + 3, m->Ge, m->Gf, M,
                                         (wf || wt) ? 1 : 0, wf ? 1 : 0, m->part);
    std::vector<double> ep(static_cast<size_t>(kb));
    double mom[4] = {0., 0., 0., 0.};
    if (energy) {
      hip_check(hipMemcpyAsync(ep.data(), m->part, sizeof(double) * kb, hipMemcpyDeviceToHost,
                               st),
                "copy dp3m energy");
      hip_check(hipMemcpyAsync(mom, m->mom, sizeof(mom), hipMemcpyDeviceToHost, st),
                "copy dp3m moments");
    }
    if (wf || wt) {
      // the field meshes [0..2] and force meshes [3..8] are consecutive in
      // the table: one launch per axis transforms all of them
      fft(m->d_ptrs, wf ? 9 : 3, true, m->d_rptrs);
      double const pt = dipole_prefac * wn, pf = dipole_prefac * wn * wn;
      switch (p.cao) {
      case 1: launch_gather<1>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->d_rptrs, pt, pf, perm, force, torque); break;
      case 2: launch_gather<2>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->d_rptrs, pt, pf, perm, force, torque); break;
      case 3: launch_gather<3>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->d_rptrs, pt, pf, perm, force, torque); break;
      case 4: launch_gather<4>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->d_rptrs, pt, pf, perm, force, torque); break;
      case 5: launch_gather<5>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->d_rptrs, pt, pf, perm, force, torque); break;
      case 6: launch_gather<6>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->d_rptrs, pt, pf, perm, force, torque); break;
      default: launch_gather<7>(st, nk, E, m->offsets, m->total, sp, sd, m->g, m->d_rptrs, pt, pf, perm, force, torque); break;
      }
    }
    hip_check(hipGetLastError(), "dipolar P3M kernels");
    if (energy || sync) {
      int err = 0;
      hip_check(hipMemcpyAsync(&err, m->err, sizeof(int), hipMemcpyDeviceToHost, st),
                "copy flag");
      hip_check(hipStreamSynchronize(st), "sync dp3m");
      if (err) throw std::runtime_error("DipolarP3M: particle outside the charge-assignment mesh");
    }
    if (energy) {
      double const mu2 = mom[0];
      double ek = 0.;
      for (double v : ep) ek += v;
      double e = ek * dipole_prefac * std::numbers::pi / p.box;
      e -= p.prefactor * mu2 * std::numbers::inv_sqrtpi * (2. / 3.) * p.alpha * p.alpha * p.alpha;
      // calc_energy_correction (dp3m.cpp:892-901)
      double const alpha_L = p.alpha * p.box;
      double const ukp3m = std::numbers::pi * m->self_sum / (3. * p.box * static_cast<double>(M3)) * V;
      double const eself = -2. * alpha_L * alpha_L * alpha_L * std::numbers::inv_sqrtpi / 3.;
      double const corr = -mu2 * (ukp3m + eself + 2. * std::numbers::pi / 3.);
      e += p.prefactor * corr / V;
      *energy = (mu2 > 0.) ? e : 0.;
      if (p.epsilon != 0.) {
        // calc_surface_term energy (dp3m.cpp:483-491)
        double const pref = p.prefactor * 4. * std::numbers::pi / V / (2. * p.epsilon + 1.);
        *energy += 0.5 * pref * (mom[1] * mom[1] + mom[2] * mom[2] + mom[3] * mom[3]);
      }
    }
  }
  if (p.epsilon != 0. && torque) {
    // calc_surface_term torques (dp3m.cpp:493-514)
    double const pref = p.prefactor * 4. * std::numbers::pi / V / (2. * p.epsilon + 1.);
    k_dp_surface<<<nb, kDpBlock, 0, st>>>(n, dip, m->mom, pref, torque);
  }
  hip_check(hipGetLastError(), "dipolar P3M surface");
  if (sync || energy) hip_check(hipStreamSynchronize(st), "sync dp3m");
}

void DipolarSolver::short_range(int64_t n, double const *pos, double const *dip, double *force,
                                double *torque, double *energy, bool sync) {
  if (!m->configured) throw std::runtime_error("DipolarP3M solver used before configure()");
  if (energy) *energy = 0.;
  if (n <= 0) return;
  hip_check(hipSetDevice(m->device), "hipSetDevice");
  hipStream_t const st = m->st;
  DipolarParams const &p = m->p;
  int const nb = static_cast<int>((n + kDpBlock - 1) / kDpBlock);
  if (static_cast<size_t>(nb) > m->cap_part) {
    hip_check(hipStreamSynchronize(st), "sync");
    dp_free(m->part);
    hip_check(hipMalloc(&m->part, sizeof(double) * nb), "hipMalloc dp3m partials");
    m->cap_part = static_cast<size_t>(nb);
  }
  int const exact = p.accuracy > 5e-6 ? 0 : 1;
  // cell lists when the box holds three cells of edge r_cut, else all pairs
  if (!m->sr) m->sr = std::make_unique<ShortRange>(m->device);
  ShortRangeParams sp;
  for (int d = 0; d < 3; ++d) sp.box[d] = p.box;
  sp.prefactor = p.prefactor;
  sp.alpha = p.alpha;
  sp.r_cut = p.r_cut;
  // cells of r_cut/2 test 42 % fewer candidates and give 8 times more
  // blocks; cells of r_cut cost fewer passes over neighbour tables.
  // Measured (N = 2e5, r_cut 10), r_cut against r_cut/2 cells: 0.76 against
  // 2.41 ms at 2.7 dipoles per r_cut/2 cell, 1.98 against 2.54 at 12.5,
  // 6.6 against 6.1 at 48, 71.5 against 19.8 at 125
  double const half = 0.5 * p.r_cut;
  double const per_cell = static_cast<double>(n) * half * half * half / (p.box * p.box * p.box);
  sp.subdivision = m->subdivision ? m->subdivision : (per_cell >= 40. ? 2 : 1);
  if (m->sr->compute_dipolar(st, sp, exact != 0, n, pos, dip, force, torque, energy)) {
    if (sync && !energy) hip_check(hipStreamSynchronize(st), "sync dp3m pairs");
    return;
  }
  if (energy)
    k_dp_pairs<true><<<nb, kDpBlock, 0, st>>>(n, pos, dip, p.box, p.prefactor, p.alpha, p.r_cut,
                                              exact, force, torque, m->part);
  else
    k_dp_pairs<false><<<nb, kDpBlock, 0, st>>>(n, pos, dip, p.box, p.prefactor, p.alpha, p.r_cut,
                                               exact, force, torque, nullptr);
  hip_check(hipGetLastError(), "dipolar pair kernel");
  if (energy) {
    std::vector<double> e(static_cast<size_t>(nb));
    hip_check(hipMemcpyAsync(e.data(), m->part, sizeof(double) * nb, hipMemcpyDeviceToHost, st),
              "copy dp3m pair energy");
    hip_check(hipStreamSynchronize(st), "sync dp3m pairs");
    double s = 0.;
    for (double v : e) s += v;
    *energy = 0.5 * s;
  } else if (sync) {
    hip_check(hipStreamSynchronize(st), "sync dp3m pairs");
  }
}

} // namespace espresso_amd
