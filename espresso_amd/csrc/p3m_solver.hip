// Device P3M k-space solver for MI355X; see p3m_solver.hpp for the pipeline.
#include "p3m_kernels.hpp"
#include "p3m_solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>

namespace espresso_amd {

void hip_check(hipError_t err, char const *what) {
  if (err != hipSuccess)
    throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(err));
}

FftPlan1D make_fft_plan(int n) {
  if (n < 1) throw std::runtime_error("FFT length must be >= 1");
  FftPlan1D p;
  p.n = n;
  int m = n;
  auto push = [&p](int r) {
    if (p.nstages >= FftPlan1D::max_stages) throw std::runtime_error("FFT plan too long");
    p.radix[p.nstages++] = r;
  };
  while (m % 8 == 0) { push(8); m /= 8; }
  while (m % 4 == 0) { push(4); m /= 4; }
  while (m % 2 == 0) { push(2); m /= 2; }
  for (int f : {3, 5, 7}) while (m % f == 0) { push(f); m /= f; }
  for (int f = 11; m > 1 && f * f <= m; f += 2) while (m % f == 0) { push(f); m /= f; }
  if (m > 1) push(m);
  return p;
}

namespace {
using namespace dev;

constexpr int kBlock = 256;
constexpr int kPG = 512;    // block size of the particle-mesh kernels
constexpr int kChunk = 64;  // particles staged per brick pass

// ---------------------------------------------------------------------------
// small utility kernels
// ---------------------------------------------------------------------------
template <typename T> __global__ void k_fill(T *p, T v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// influence function on the half spectrum [Mx][My][Kz] (FP64 arithmetic):
// grid_influence_function<T, S=1, m=0> (influence_function.hpp:110-150) with
// G_opt (influence_function.hpp:53-89): G = 4 pi exp(-k^2/4a^2) / (k^2 U^2),
// U^2 = prod_d sinc(k_d h_d / 2pi)^(2 cao), zero where every index is a
// multiple of M/2 and where k^2/4a^2 >= 30.
// ---------------------------------------------------------------------------
__device__ double dev_sinc(double x) {
  double const pix = std::numbers::pi * x;
  if (fabs(x) > 0.1) return sin(pix) / pix;
  double const c2 = -1. / 6., c4 = 1. / 120., c6 = -1. / 5040., c8 = 1. / 362880.;
  double const pix2 = pix * pix;
  return 1. + pix2 * (c2 + pix2 * (c4 + pix2 * (c6 + pix2 * c8)));
}

template <typename MT>
__global__ void k_influence(MT *G, int Mx, int My, int Mz, int Kz, double wx, double wy,
                            double wz, double hx, double hy, double hz, double alpha, int cao) {
  int64_t const total = (int64_t)Mx * My * Kz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int const kz = static_cast<int>(e % Kz);
    int const y = static_cast<int>((e / Kz) % My);
    int const x = static_cast<int>(e / ((int64_t)Kz * My));
    double g = 0.;
    int const hmx = Mx / 2, hmy = My / 2, hmz = Mz / 2;
    bool const zero = (hmx == 0 || x % hmx == 0) && (hmy == 0 || y % hmy == 0) &&
                      (hmz == 0 || kz % hmz == 0);
    if (!zero) {
      double const k0 = plain_shift(x, Mx) * wx;
      double const k1 = plain_shift(y, My) * wy;
      double const k2v = plain_shift(kz, Mz) * wz;
      double const k2 = k0 * k0 + k1 * k1 + k2v * k2v;
      if (k2 != 0.) {
        double const f0 = dev_sinc(k0 * (1. / ((2. * std::numbers::pi) / hx)));
        double const f1 = dev_sinc(k1 * (1. / ((2. * std::numbers::pi) / hy)));
        double const f2 = dev_sinc(k2v * (1. / ((2. * std::numbers::pi) / hz)));
        double const U2 = pow(f0 * f1 * f2, 2 * cao);
        double const ex = (1. / (2. * alpha)) * (1. / (2. * alpha)) * k2;
        double numerator = 0.;
        if (ex < 30.) {
          double const f3 = exp(-ex) * (4. * std::numbers::pi / k2);
          numerator = U2 * f3 * k2;
        }
        g = numerator / (k2 * (U2 * U2));
      }
    }
    G[e] = static_cast<MT>(g);
  }
}

// ---------------------------------------------------------------------------
// particle binning: first stencil point, brick id, rank inside the brick,
// block partial sums (n_q, sum q, sum q^2, dipole)
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(kBlock) void k_bin(int64_t n, double const *__restrict__ pos,
                                                double const *__restrict__ q,
                                                double const *__restrict__ upos, DevGeo g,
                                                int *brick_count, int *p_brick, int *p_rank,
                                                int *p_nmp, double *partials, int *err_flag) {
  __shared__ double red[8 * 16];
  int64_t const i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double v[7] = {0., 0., 0., 0., 0., 0., 0.};
  if (i < n) {
    double const qi = q[i];
    int b = -1;
    if (qi != 0.) {
      double const x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
      int nmp[3], gi[3];
      double dist[3];
      if (!particle_coordinates<P>(g, x, nmp, gi, dist)) atomicOr(err_flag, 1);
      b = ((gi[0] / g.B[0]) * g.nb[1] + gi[1] / g.B[1]) * g.nb[2] + gi[2] / g.B[2];
      p_rank[i] = atomicAdd(&brick_count[b], 1);
      if (p_nmp) {
        p_nmp[3 * i] = nmp[0];
        p_nmp[3 * i + 1] = nmp[1];
        p_nmp[3 * i + 2] = nmp[2];
      }
      double const *u = upos ? upos + 3 * i : x;
      v[0] = 1.;
      v[1] = qi;
      v[2] = qi * qi;
      v[3] = qi * u[0];
      v[4] = qi * u[1];
      v[5] = qi * u[2];
    } else if (p_nmp) {
      p_nmp[3 * i] = p_nmp[3 * i + 1] = p_nmp[3 * i + 2] = -1;
    }
    p_brick[i] = b;
  }
  block_sum<7>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 7; ++k) partials[blockIdx.x * 8 + k] = v[k];
}

/** deterministic reduction of per-block partials: out[k] = sum_b part[b*8+k] */
__global__ void k_reduce_partials(double const *part, int nblk, int nk, double *out) {
  __shared__ double red[16 * 8];
  double v[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x)
    for (int k = 0; k < nk; ++k) v[k] += part[b * 8 + k];
  block_sum<8>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < nk; ++k) out[k] = v[k];
}

/** exclusive scan of the brick counts into offset[0..nb] (one block) */
__global__ __launch_bounds__(1024) void k_scan(int const *count, int *offset, int nb) {
  __shared__ int wsum[16];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  int const lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int base = 0; base < nb; base += blockDim.x) {
    int const i = base + threadIdx.x;
    int const c = (i < nb) ? count[i] : 0;
    // inclusive wave scan
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
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        int const t = wsum[w];
        wsum[w] = acc;
        acc += t;
      }
    }
    __syncthreads();
    if (i < nb) offset[i] = carry + wsum[wid] + s - c;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += wsum[wid] + s;
    __syncthreads();
  }
  if (threadIdx.x == 0) offset[nb] = carry;
}

__global__ void k_scatter(int64_t n, double const *__restrict__ pos, double const *__restrict__ q,
                          int const *p_brick, int const *p_rank, int const *offset,
                          double4 *spos, int *sidx) {
  int64_t const i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int const b = p_brick[i];
  if (b < 0) return;
  int const s = offset[b] + p_rank[i];
  spos[s] = double4{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], q[i]};
  sidx[s] = static_cast<int>(i);
}

// ---------------------------------------------------------------------------
// charge assignment into LDS brick tiles (AssignCharge<cao>, p3m.cpp:283-330;
// p3m_interpolate, interpolation.hpp:191-208).  One block per brick: the
// tile holds the brick's B^3 first-stencil points plus cao-1 planes on the
// high side; contributions are added with LDS atomics, the tile is flushed
// into the periodic mesh with device-scope atomics (neighbouring tiles
// overlap by cao-1 planes).
// ---------------------------------------------------------------------------
struct ChunkSlots {
  double *w;  // [kChunk][3P]
  double *qq; // [kChunk]
  int *lo;    // [kChunk] tile index of the first stencil point
};

/** Lane segments for stencil order P: a particle occupies S = next power of
 *  two >= P^2 lanes of a wave64 (ppw particles per wave); lane r < P^2 of a
 *  segment handles stencil column (j, k) = (r / P, r % P). */
template <int P> struct Seg {
  static constexpr int S = (P * P <= 1) ? 1 : (P * P <= 4) ? 4 : (P * P <= 16) ? 16
                           : (P * P <= 32) ? 32 : 64;
  static constexpr int ppw = 64 / S;
};
template <int P> struct LaneSeg {
  int wave, nwaves, sub, r, j, k;
  bool active;
  __device__ LaneSeg() {
    int const lane = threadIdx.x & 63;
    wave = threadIdx.x >> 6;
    nwaves = blockDim.x >> 6;
    sub = lane / Seg<P>::S;
    r = lane - sub * Seg<P>::S;
    active = r < P * P;
    j = active ? r / P : 0;
    k = active ? r - (r / P) * P : 0;
  }
};

template <int P>
__device__ __forceinline__ void stage_chunk(DevGeo const &g, double4 const *spos, int c0, int cnt,
                                            int o0, int o1, int o2, int TY, int TZ,
                                            ChunkSlots s) {
  int const t = threadIdx.x;
  if (t < cnt) {
    double4 const p = spos[c0 + t];
    double const x[3] = {p.x, p.y, p.z};
    int nmp[3], gi[3];
    double dist[3];
    particle_coordinates<P>(g, x, nmp, gi, dist);
    double *w = s.w + t * 3 * P;
    bspline_weights<P>(dist[0], w);
    bspline_weights<P>(dist[1], w + P);
    bspline_weights<P>(dist[2], w + 2 * P);
    s.qq[t] = p.w;
    s.lo[t] = ((gi[0] - o0) * TY + (gi[1] - o1)) * TZ + (gi[2] - o2);
  }
}

template <int P, typename MT>
__global__ __launch_bounds__(kPG) void k_spread(DevGeo g, int const *__restrict__ offset,
                                                double4 const *__restrict__ spos,
                                                MT *__restrict__ rho) {
  // The tile accumulates in FP64 for both mesh precisions: LDS FP64 atomics
  // (ds_add_f64) are several times faster than FP32 ones (ds_add_f32) for
  // this scattered pattern on gfx950, and the sum is rounded once.
  extern __shared__ double smem_d[];
  int const TY = g.T[1], TZ = g.T[2];
  int const TYZ = TY * TZ;
  int const tile_n = g.T[0] * TYZ;
  ChunkSlots s;
  s.w = smem_d;
  s.qq = s.w + kChunk * 3 * P;
  s.lo = reinterpret_cast<int *>(s.qq + kChunk);
  double *tile = reinterpret_cast<double *>(s.lo + kChunk);

  int const b = blockIdx.x;
  int const bz = b % g.nb[2], by = (b / g.nb[2]) % g.nb[1], bx = b / (g.nb[2] * g.nb[1]);
  int const o0 = bx * g.B[0], o1 = by * g.B[1], o2 = bz * g.B[2];
  int const beg = offset[b], end = offset[b + 1];
  if (beg == end) return;
  LaneSeg<P> const seg;
  for (int e = threadIdx.x; e < tile_n; e += blockDim.x) tile[e] = 0.;
  for (int c0 = beg; c0 < end; c0 += kChunk) {
    int const cnt = min(kChunk, end - c0);
    __syncthreads();
    stage_chunk<P>(g, spos, c0, cnt, o0, o1, o2, TY, TZ, s);
    __syncthreads();
    // one lane segment per particle, lanes over the (j, k) plane of the
    // stencil, loop over i: every instruction touches one compact patch
    int const ngroups = (cnt + Seg<P>::ppw - 1) / Seg<P>::ppw;
    for (int grp = seg.wave; grp < ngroups; grp += seg.nwaves) {
      int const p = grp * Seg<P>::ppw + seg.sub;
      if (seg.active && p < cnt) {
        double const *w = s.w + p * 3 * P;
        double const wyzq = (w[P + seg.j] * w[2 * P + seg.k]) * s.qq[p];
        double *dst = tile + s.lo[p] + seg.j * TZ + seg.k;
#pragma unroll
        for (int i = 0; i < P; ++i) atomicAdd(dst + i * TYZ, w[i] * wyzq);
      }
    }
  }
  __syncthreads();
  int const M0 = g.M[0], M1 = g.M[1], M2 = g.M[2];
  for (int e = threadIdx.x; e < tile_n; e += blockDim.x) {
    double const v = tile[e];
    if (v != 0.) {
      int const a = magic_div(e, g.tyz_magic);
      int const rr = e - a * TYZ;
      int const bb = magic_div(rr, g.tz_magic);
      int const c = rr - bb * TZ;
      int const x = wrap_once(o0 + a, M0), y = wrap_once(o1 + bb, M1), z = wrap_once(o2 + c, M2);
      unsafeAtomicAdd(rho + ((int64_t)x * M1 + y) * M2 + z, static_cast<MT>(v));
    }
  }
}

// ---------------------------------------------------------------------------
// FFT passes.  Real mesh [Mx][My][Mz]; half spectrum [Mx][My][Kz].
// ---------------------------------------------------------------------------
template <typename C>
__device__ __forceinline__ void load_twiddles(C *tw_lds, C const *tw, int n) {
  for (int m = threadIdx.x; m < n; m += blockDim.x) tw_lds[m] = tw[m];
}

/** forward r2c along z: two real rows per complex line
 *  (X1 = (Z_k + conj Z_{n-k})/2, X2 = (Z_k - conj Z_{n-k})/2i) */
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_fft_z_r2c(MT const *__restrict__ rho,
                                                      cx_t<MT> *__restrict__ khat, int rows,
                                                      int Kz, FftPlanDev plan,
                                                      cx_t<MT> const *__restrict__ tw, int nlp) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  int const n = plan.n;
  int const ls = n + 1;
  C *twl = reinterpret_cast<C *>(smem_d);
  C *b0 = twl + n;
  C *b1 = b0 + nlp * ls;
  load_twiddles(twl, tw, n);
  int const pair0 = blockIdx.x * nlp;
  for (int e = threadIdx.x; e < nlp * n; e += blockDim.x) {
    int const l = e / n, j = e - l * n;
    int64_t const r0 = 2 * (int64_t)(pair0 + l), r1 = r0 + 1;
    MT const re = (r0 < rows) ? rho[r0 * n + j] : MT(0);
    MT const im = (r1 < rows) ? rho[r1 * n + j] : MT(0);
    b0[l * ls + j] = C{re, im};
  }
  __syncthreads();
  C *res = lds_fft(b0, b1, nlp, ls, plan, twl, false);
  for (int e = threadIdx.x; e < nlp * Kz; e += blockDim.x) {
    int const l = e / Kz, k = e - l * Kz;
    int64_t const r0 = 2 * (int64_t)(pair0 + l), r1 = r0 + 1;
    C const Z = res[l * ls + k];
    C const Zc = cconj(res[l * ls + (n - k) % n]);
    C const s = cadd(Z, Zc), d = csub(Z, Zc);
    if (r0 < rows) khat[r0 * Kz + k] = C{MT(0.5) * s.x, MT(0.5) * s.y};
    if (r1 < rows) khat[r1 * Kz + k] = C{MT(0.5) * d.y, MT(-0.5) * d.x};
  }
}

/** c2c along y for a tile of kz columns at fixed x (in place). */
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_fft_y(cx_t<MT> *__restrict__ khat, int My, int Kz,
                                                  FftPlanDev plan, cx_t<MT> const *__restrict__ tw,
                                                  int ncol, int inv) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  int const n = plan.n;
  int const ls = n + 1;
  C *twl = reinterpret_cast<C *>(smem_d);
  C *b0 = twl + n;
  C *b1 = b0 + ncol * ls;
  load_twiddles(twl, tw, n);
  int const c0 = blockIdx.x * ncol;
  int64_t const x = blockIdx.y;
  int const nc = min(ncol, Kz - c0);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = e / ncol, c = e - j * ncol;
    if (c < nc) b0[c * ls + j] = khat[(x * My + j) * Kz + c0 + c];
  }
  __syncthreads();
  C *res = lds_fft(b0, b1, nc, ls, plan, twl, inv != 0);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = e / ncol, c = e - j * ncol;
    if (c < nc) khat[(x * My + j) * Kz + c0 + c] = res[c * ls + j];
  }
}

/** forward c2c along x, convolution with the influence function, energy and
 *  virial partial sums (p3m.cpp:393-449, 551-567), ik along x and inverse c2c
 *  along x (p3m.cpp:491-522): A = IFFT_x(G rho^), B = IFFT_x(i k_x G rho^). */
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_fft_x_conv(
    cx_t<MT> const *__restrict__ khat, MT const *__restrict__ G, cx_t<MT> *__restrict__ Aout,
    cx_t<MT> *__restrict__ Bout, int Mx, int My, int Mz, int Kz, double wvx, double wvy,
    double wvz, double half_alpha_inv_sq, FftPlanDev plan, cx_t<MT> const *__restrict__ tw,
    int ncol, int want_fields, double *partials) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  __shared__ double red[8 * 16];
  int const n = plan.n; // == Mx
  int const ls = n + 1;
  C *twl = reinterpret_cast<C *>(smem_d);
  C *X = twl + n;          // 2*ncol lines
  C *Y = X + 2 * ncol * ls; // 2*ncol lines
  load_twiddles(twl, tw, n);
  int const c0 = blockIdx.x * ncol;
  int const y = blockIdx.y;
  int const nc = min(ncol, Kz - c0);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = e / ncol, c = e - j * ncol;
    if (c < nc) X[c * ls + j] = khat[((int64_t)j * My + y) * Kz + c0 + c];
  }
  __syncthreads();
  C *res = lds_fft(X, Y, nc, ls, plan, twl, false);
  C *oth = (res == X) ? Y : X;
  double acc[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  double const ky = dop_shift(y, My) * wvy;
  MT const kyf = static_cast<MT>(dop_shift(y, My)) * static_cast<MT>(wvy);
  (void)kyf;
  for (int e = threadIdx.x; e < n * nc; e += blockDim.x) {
    int const c = e / n, j = e - c * n;
    int const kz = c0 + c;
    C const rh = res[c * ls + j];
    MT const gv = G[((int64_t)j * My + y) * Kz + kz];
    double const wgt = (kz == 0 || 2 * kz == Mz) ? 1. : 2.;
    double const ek = static_cast<double>(gv) *
                      (static_cast<double>(rh.x) * rh.x + static_cast<double>(rh.y) * rh.y);
    double const kx = dop_shift(j, Mx) * wvx;
    double const kzv = dop_shift(kz, Mz) * wvz;
    double const norm_sq = kx * kx + ky * ky + kzv * kzv;
    acc[0] += wgt * ek;
    if (norm_sq != 0.) {
      double const pref = wgt * ek * (-2. * (1. / norm_sq + half_alpha_inv_sq));
      acc[1] += pref * kx * kx;
      acc[2] += pref * kx * ky;
      acc[3] += pref * kx * kzv;
      acc[4] += pref * ky * ky;
      acc[5] += pref * ky * kzv;
      acc[6] += pref * kzv * kzv;
      acc[7] += wgt * ek;
    }
    if (want_fields) {
      C const phi = C{gv * rh.x, gv * rh.y};
      MT const kxf = static_cast<MT>(dop_shift(j, Mx)) * static_cast<MT>(wvx);
      oth[c * ls + j] = phi;                                   // A
      oth[(ncol + c) * ls + j] = C{-kxf * phi.y, kxf * phi.x}; // B = i kx phi
    }
  }
  block_sum<8>(acc, red);
  if (threadIdx.x == 0) {
    int const blk = blockIdx.y * gridDim.x + blockIdx.x;
    for (int k = 0; k < 8; ++k) partials[blk * 8 + k] = acc[k];
  }
  if (!want_fields) return;
  // inverse along x of A (lines 0..nc) and B (lines ncol..ncol+nc): run on
  // all 2*ncol lines (unused lines hold stale data and are not stored)
  C *ires = lds_fft(oth, res, 2 * ncol, ls, plan, twl, true);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = e / ncol, c = e - j * ncol;
    if (c < nc) {
      int64_t const gidx = ((int64_t)j * My + y) * Kz + c0 + c;
      Aout[gidx] = ires[c * ls + j];
      Bout[gidx] = ires[(ncol + c) * ls + j];
    }
  }
}

/** inverse c2c along y of the x-transformed fields, in place:
 *  A <- IFFT_y(A) (z field before i k_z), B <- IFFT_y(B) (x field),
 *  Fy <- IFFT_y(i k_y A) (y field). */
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_fft_y_inv(cx_t<MT> *__restrict__ A,
                                                      cx_t<MT> *__restrict__ B,
                                                      cx_t<MT> *__restrict__ Fy, int My, int Kz,
                                                      double wvy, FftPlanDev plan,
                                                      cx_t<MT> const *__restrict__ tw, int ncol) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  int const n = plan.n; // == My
  int const ls = n + 1;
  C *twl = reinterpret_cast<C *>(smem_d);
  C *S = twl + n;
  C *W0 = S + ncol * ls;
  C *W1 = W0 + ncol * ls;
  load_twiddles(twl, tw, n);
  int const c0 = blockIdx.x * ncol;
  int64_t const x = blockIdx.y;
  int const nc = min(ncol, Kz - c0);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = e / ncol, c = e - j * ncol;
    if (c < nc) {
      C const v = A[(x * My + j) * Kz + c0 + c];
      S[c * ls + j] = v;
      W0[c * ls + j] = v;
    }
  }
  __syncthreads();
  C *res = lds_fft(W0, W1, nc, ls, plan, twl, true);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = e / ncol, c = e - j * ncol;
    if (c < nc) A[(x * My + j) * Kz + c0 + c] = res[c * ls + j];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * nc; e += blockDim.x) {
    int const c = e / n, j = e - c * n;
    MT const kyf = static_cast<MT>(dop_shift(j, My)) * static_cast<MT>(wvy);
    C const v = S[c * ls + j];
    W0[c * ls + j] = C{-kyf * v.y, kyf * v.x};
  }
  __syncthreads();
  res = lds_fft(W0, W1, nc, ls, plan, twl, true);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = e / ncol, c = e - j * ncol;
    if (c < nc) Fy[(x * My + j) * Kz + c0 + c] = res[c * ls + j];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = e / ncol, c = e - j * ncol;
    if (c < nc) W0[c * ls + j] = B[(x * My + j) * Kz + c0 + c];
  }
  __syncthreads();
  res = lds_fft(W0, W1, nc, ls, plan, twl, true);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = e / ncol, c = e - j * ncol;
    if (c < nc) B[(x * My + j) * Kz + c0 + c] = res[c * ls + j];
  }
}

/** inverse c2r along z: two half-spectrum rows per complex line,
 *  optionally multiplied by i k_z (z component of the field). */
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_fft_z_c2r(cx_t<MT> const *__restrict__ S,
                                                      MT *__restrict__ E, int rows, int Kz,
                                                      double wvz, int mul_ikz, FftPlanDev plan,
                                                      cx_t<MT> const *__restrict__ tw, int nlp) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  int const n = plan.n; // == Mz
  int const ls = n + 1;
  C *twl = reinterpret_cast<C *>(smem_d);
  C *b0 = twl + n;
  C *b1 = b0 + nlp * ls;
  load_twiddles(twl, tw, n);
  int const pair0 = blockIdx.x * nlp;
  for (int e = threadIdx.x; e < nlp * n; e += blockDim.x) {
    int const l = e / n, k = e - l * n;
    int64_t const r0 = 2 * (int64_t)(pair0 + l), r1 = r0 + 1;
    bool const upper = (k >= Kz);
    int const ks = upper ? n - k : k;
    C x1 = (r0 < rows) ? S[r0 * Kz + ks] : C{MT(0), MT(0)};
    C x2 = (r1 < rows) ? S[r1 * Kz + ks] : C{MT(0), MT(0)};
    if (mul_ikz) {
      MT const kz = static_cast<MT>(dop_shift(ks, n)) * static_cast<MT>(wvz);
      x1 = C{-kz * x1.y, kz * x1.x};
      x2 = C{-kz * x2.y, kz * x2.x};
    }
    if (ks == 0 || 2 * ks == n) {
      x1.y = MT(0);
      x2.y = MT(0);
    }
    if (upper) {
      x1 = cconj(x1);
      x2 = cconj(x2);
    }
    b0[l * ls + k] = C{x1.x - x2.y, x1.y + x2.x};
  }
  __syncthreads();
  C *res = lds_fft(b0, b1, nlp, ls, plan, twl, true);
  for (int e = threadIdx.x; e < nlp * n; e += blockDim.x) {
    int const l = e / n, j = e - l * n;
    int64_t const r0 = 2 * (int64_t)(pair0 + l), r1 = r0 + 1;
    C const v = res[l * ls + j];
    if (r0 < rows) E[r0 * n + j] = v.x;
    if (r1 < rows) E[r1 * n + j] = v.y;
  }
}

// ---------------------------------------------------------------------------
// force back-interpolation (AssignForces<cao>, p3m.cpp:344-373): one block
// per brick with the brick's field tiles in LDS.  A lane segment per particle
// covers the (j, k) plane of the stencil and loops over i; the segment is
// reduced with a fixed xor butterfly (deterministic, no atomics) and its
// first lane applies F -= q * prefactor/V * E in FP64.
// ---------------------------------------------------------------------------
/** Lane segments of the gather: S = next power of two >= P lanes per
 *  particle, lane i < P of a segment sweeps the (j, k) plane of stencil
 *  layer i. */
template <int P> struct GSeg {
  static constexpr int S = (P <= 1) ? 1 : (P <= 2) ? 2 : (P <= 4) ? 4 : 8;
  static constexpr int ppw = 64 / S;
};

/** Sum over aligned segments of S in {1, 2, 4, 8} lanes with DPP; valid in
 *  every lane of the segment.  Executed by all lanes of the wave. */
template <int S> __device__ __forceinline__ double small_segment_total(double v) {
  if constexpr (S >= 2) v += dpp_f64<0xB1>(v);  // quad_perm [1,0,3,2]
  if constexpr (S >= 4) v += dpp_f64<0x4E>(v);  // quad_perm [2,3,0,1]
  if constexpr (S >= 8) v += dpp_f64<0x141>(v); // row_half_mirror
  return v;
}

template <int P, int BB, typename MT>
__global__ __launch_bounds__(kPG) void k_gather(
    DevGeo g, int const *__restrict__ offset, double4 const *__restrict__ spos,
    int const *__restrict__ sidx, MT const *__restrict__ Ex, MT const *__restrict__ Ey,
    MT const *__restrict__ Ez, double *__restrict__ force, double force_prefac,
    double const *__restrict__ sums, double dip_pref) {
  constexpr int T = BB + P - 1;
  constexpr int TYZ = T * T;
  constexpr int tile_n = T * TYZ;
  constexpr int S = GSeg<P>::S;
  constexpr int ppw = GSeg<P>::ppw;
  extern __shared__ double smem_d[];
  ChunkSlots s;
  s.w = smem_d;
  s.qq = s.w + kChunk * 3 * P;
  s.lo = reinterpret_cast<int *>(s.qq + kChunk);
  MT *tx = reinterpret_cast<MT *>(s.lo + kChunk);
  MT *ty = tx + tile_n;
  MT *tz = ty + tile_n;

  int const b = blockIdx.x;
  int const beg = offset[b], end = offset[b + 1];
  if (beg == end) return;
  int const bz = b % g.nb[2], by = (b / g.nb[2]) % g.nb[1], bx = b / (g.nb[2] * g.nb[1]);
  int const o0 = bx * BB, o1 = by * BB, o2 = bz * BB;
  int const M0 = g.M[0], M1 = g.M[1], M2 = g.M[2];
  for (int e = threadIdx.x; e < tile_n; e += blockDim.x) {
    int const a = e / TYZ;
    int const rr = e - a * TYZ;
    int const bb = rr / T;
    int const c = rr - bb * T;
    int64_t const gidx =
        ((int64_t)wrap_once(o0 + a, M0) * M1 + wrap_once(o1 + bb, M1)) * M2 + wrap_once(o2 + c, M2);
    tx[e] = Ex[gidx];
    ty[e] = Ey[gidx];
    tz[e] = Ez[gidx];
  }
  double dm[3] = {0., 0., 0.};
  if (dip_pref != 0.) {
    dm[0] = dip_pref * sums[3];
    dm[1] = dip_pref * sums[4];
    dm[2] = dip_pref * sums[5];
  }
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const nwaves = blockDim.x >> 6;
  int const sub = lane / S;
  int const li = lane - sub * S;
  bool const active = li < P;
  for (int c0 = beg; c0 < end; c0 += kChunk) {
    int const cnt = min(kChunk, end - c0);
    __syncthreads();
    stage_chunk<P>(g, spos, c0, cnt, o0, o1, o2, T, T, s);
    __syncthreads();
    int const ngroups = (cnt + ppw - 1) / ppw;
    for (int grp = wave; grp < ngroups; grp += nwaves) {
      int const p = grp * ppw + sub;
      bool const valid = active && p < cnt;
      double fx = 0., fy = 0., fz = 0.;
      if (valid) {
        double const *w = s.w + p * 3 * P;
        double wz[P];
#pragma unroll
        for (int k = 0; k < P; ++k) wz[k] = w[2 * P + k];
        double const wxi = w[li];
        int const base = s.lo[p] + li * TYZ;
#pragma unroll
        for (int j = 0; j < P; ++j) {
          double const wj = wxi * w[P + j];
          int const row = base + j * T;
#pragma unroll
          for (int k = 0; k < P; ++k) {
            double const ww = wj * wz[k];
            fx += ww * static_cast<double>(tx[row + k]);
            fy += ww * static_cast<double>(ty[row + k]);
            fz += ww * static_cast<double>(tz[row + k]);
          }
        }
      }
      fx = small_segment_total<S>(fx);
      fy = small_segment_total<S>(fy);
      fz = small_segment_total<S>(fz);
      if (li == 0 && p < cnt) {
        double const qv = s.qq[p];
        double const pf = qv * force_prefac;
        int64_t const idx = sidx[c0 + p];
        double *f = force + 3 * idx;
        f[0] = (f[0] - pf * fx) - qv * dm[0];
        f[1] = (f[1] - pf * fy) - qv * dm[1];
        f[2] = (f[2] - pf * fz) - qv * dm[2];
      }
    }
  }
}

/** Gather with one lane segment per particle over the (j, k) plane of the
 *  stencil (every LDS instruction serves one particle, so reads are free of
 *  inter-particle bank conflicts; the segment is reduced with DPP and lane
 *  reads).  Preferred for FP64 fields. */
template <int P, int BB, typename MT>
__global__ __launch_bounds__(kPG) void k_gather_plane(
    DevGeo g, int const *__restrict__ offset, double4 const *__restrict__ spos,
    int const *__restrict__ sidx, MT const *__restrict__ Ex, MT const *__restrict__ Ey,
    MT const *__restrict__ Ez, double *__restrict__ force, double force_prefac,
    double const *__restrict__ sums, double dip_pref) {
  constexpr int T = BB + P - 1;
  constexpr int TYZ = T * T;
  constexpr int tile_n = T * TYZ;
  extern __shared__ double smem_d[];
  ChunkSlots s;
  s.w = smem_d;
  s.qq = s.w + kChunk * 3 * P;
  s.lo = reinterpret_cast<int *>(s.qq + kChunk);
  MT *tx = reinterpret_cast<MT *>(s.lo + kChunk);
  MT *ty = tx + tile_n;
  MT *tz = ty + tile_n;

  int const b = blockIdx.x;
  int const beg = offset[b], end = offset[b + 1];
  if (beg == end) return;
  int const bz = b % g.nb[2], by = (b / g.nb[2]) % g.nb[1], bx = b / (g.nb[2] * g.nb[1]);
  int const o0 = bx * BB, o1 = by * BB, o2 = bz * BB;
  int const M0 = g.M[0], M1 = g.M[1], M2 = g.M[2];
  for (int e = threadIdx.x; e < tile_n; e += blockDim.x) {
    int const a = e / TYZ;
    int const rr = e - a * TYZ;
    int const bb = rr / T;
    int const c = rr - bb * T;
    int64_t const gidx =
        ((int64_t)wrap_once(o0 + a, M0) * M1 + wrap_once(o1 + bb, M1)) * M2 + wrap_once(o2 + c, M2);
    tx[e] = Ex[gidx];
    ty[e] = Ey[gidx];
    tz[e] = Ez[gidx];
  }
  double dm[3] = {0., 0., 0.};
  if (dip_pref != 0.) {
    dm[0] = dip_pref * sums[3];
    dm[1] = dip_pref * sums[4];
    dm[2] = dip_pref * sums[5];
  }
  LaneSeg<P> const seg;
  for (int c0 = beg; c0 < end; c0 += kChunk) {
    int const cnt = min(kChunk, end - c0);
    __syncthreads();
    stage_chunk<P>(g, spos, c0, cnt, o0, o1, o2, T, T, s);
    __syncthreads();
    int const ngroups = (cnt + Seg<P>::ppw - 1) / Seg<P>::ppw;
    for (int grp = seg.wave; grp < ngroups; grp += seg.nwaves) {
      int const p = grp * Seg<P>::ppw + seg.sub;
      bool const valid = seg.active && p < cnt;
      double fx = 0., fy = 0., fz = 0.;
      if (valid) {
        double const *w = s.w + p * 3 * P;
        double const wyz = w[P + seg.j] * w[2 * P + seg.k];
        int const base = s.lo[p] + seg.j * T + seg.k;
#pragma unroll
        for (int i = 0; i < P; ++i) {
          double const ww = w[i] * wyz;
          fx += ww * static_cast<double>(tx[base + i * TYZ]);
          fy += ww * static_cast<double>(ty[base + i * TYZ]);
          fz += ww * static_cast<double>(tz[base + i * TYZ]);
        }
      }
      fx = segment_total<Seg<P>::S>(fx);
      fy = segment_total<Seg<P>::S>(fy);
      fz = segment_total<Seg<P>::S>(fz);
      if (seg.r == 0 && p < cnt) {
        double const qv = s.qq[p];
        double const pf = qv * force_prefac;
        int64_t const idx = sidx[c0 + p];
        double *f = force + 3 * idx;
        f[0] = (f[0] - pf * fx) - qv * dm[0];
        f[1] = (f[1] - pf * fy) - qv * dm[1];
        f[2] = (f[2] - pf * fz) - qv * dm[2];
      }
    }
  }
}

template <typename F> void launch_cao(int cao, F &&f) {
  switch (cao) {
  case 1: f(std::integral_constant<int, 1>{}); break;
  case 2: f(std::integral_constant<int, 2>{}); break;
  case 3: f(std::integral_constant<int, 3>{}); break;
  case 4: f(std::integral_constant<int, 4>{}); break;
  case 5: f(std::integral_constant<int, 5>{}); break;
  case 6: f(std::integral_constant<int, 6>{}); break;
  case 7: f(std::integral_constant<int, 7>{}); break;
  default: throw std::runtime_error("cao must be in [1, 7]");
  }
}

int grid_for(int64_t n, int block = kBlock, int64_t cap = 65536) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<int>(g);
}

template <typename T> void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

} // namespace

// ---------------------------------------------------------------------------
// Solver implementation
// ---------------------------------------------------------------------------
struct Solver::Impl {
  P3MParams params;
  Vec3d box_l{};
  double skin = 0.;
  double prefactor = 0.;
  MeshPrecision prec = MeshPrecision::FP64;
  DevGeo geo{};
  FftPlan1D plan[3];
  FftPlanDev dplan[3];
  int nbricks = 0;
  // mesh buffers (element size depends on precision)
  void *rho = nullptr, *khat = nullptr, *A = nullptr, *B = nullptr, *G = nullptr;
  void *Ey = nullptr, *Ez = nullptr;
  void *tw[3] = {nullptr, nullptr, nullptr};
  // particle buffers
  int64_t cap_n = 0;
  int *p_brick = nullptr, *p_rank = nullptr, *p_nmp = nullptr, *sidx = nullptr;
  double4 *spos = nullptr;
  double *bin_part = nullptr;
  int *count = nullptr, *offset = nullptr;
  double *conv_part = nullptr;
  int conv_blocks = 0;
  double *sums = nullptr; // [0..7] bin sums, [8..15] conv sums
  int *err_flag = nullptr;
  double *h_sums = nullptr; // pinned
  hipEvent_t ev[8] = {};
  // event ring for profiling runs of many asynchronous solves
  std::vector<hipEvent_t> ring;
  int ring_used = 0;
  bool profiling = false;
  // launch shapes
  int nlp_z = 1, ncol_y = 1, ncol_x = 1, ncol_yi = 1;
  // gather mapping: plane segments (FP64 fields) or layer segments (FP32)
  bool gather_plane = true;
  size_t lds_z = 0, lds_y = 0, lds_x = 0, lds_yi = 0, lds_spread = 0, lds_gather = 0;

  void free_mesh() {
    dfree(rho); dfree(khat); dfree(A); dfree(B); dfree(G); dfree(Ey); dfree(Ez);
    for (auto &t : tw) dfree(t);
    dfree(count); dfree(offset); dfree(conv_part);
  }
  void free_particles() {
    dfree(p_brick); dfree(p_rank); dfree(p_nmp); dfree(sidx); dfree(spos); dfree(bin_part);
    cap_n = 0;
  }
};

Solver::Solver(int device) : m(new Impl), m_device(device) {
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking), "hipStreamCreate");
  for (auto &e : m->ev) hip_check(hipEventCreate(&e), "hipEventCreate");
  hip_check(hipMalloc(&m->sums, 32 * sizeof(double)), "hipMalloc sums");
  hip_check(hipMalloc(&m->err_flag, sizeof(int)), "hipMalloc err");
  hip_check(hipHostMalloc(&m->h_sums, 40 * sizeof(double)), "hipHostMalloc");
}

Solver::~Solver() {
  (void)hipSetDevice(m_device);
  if (m_stream) (void)hipStreamSynchronize(m_stream);
  m->free_mesh();
  m->free_particles();
  dfree(m->sums);
  dfree(m->err_flag);
  if (m->h_sums) (void)hipHostFree(m->h_sums);
  for (auto &e : m->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : m->ring) (void)hipEventDestroy(e);
  if (m_stream) (void)hipStreamDestroy(m_stream);
  delete m;
}

void Solver::set_brick(int bx, int by, int bz) {
  if (bx != 8 || by != 8 || bz != 8)
    throw std::invalid_argument("P3M: only 8x8x8 bricks are built into this library");
  m_brick = {bx, by, bz};
  if (m_configured) configure(m->params, m->box_l, m->skin, m->prefactor, m->prec);
}

namespace {
template <typename MT> void upload_twiddles(void *&dst, int n, hipStream_t st) {
  using C = cx_t<MT>;
  std::vector<C> h(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) {
    long double const ang = -2.0L * std::numbers::pi_v<long double> * k / n;
    h[k] = C{static_cast<MT>(std::cos(ang)), static_cast<MT>(std::sin(ang))};
  }
  hip_check(hipMalloc(&dst, sizeof(C) * n), "hipMalloc twiddles");
  hip_check(hipMemcpyAsync(dst, h.data(), sizeof(C) * n, hipMemcpyHostToDevice, st),
            "hipMemcpy twiddles");
  hip_check(hipStreamSynchronize(st), "sync twiddles");
}

size_t lds_limit() { return 160 * 1024; }

template <typename Kern> void allow_lds(Kern k, size_t bytes) {
  hip_check(hipFuncSetAttribute(reinterpret_cast<void const *>(k),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(bytes)),
            "hipFuncSetAttribute");
}

/** lines per block so that `lines_per_unit * units * line_bytes + tw` fits */
int pick_lines(int n, size_t csz, int lines_per_unit, int want, size_t budget) {
  size_t const line = static_cast<size_t>(n + 1) * csz;
  size_t const tw = static_cast<size_t>(n) * csz;
  int u = want;
  while (u > 1 && tw + static_cast<size_t>(lines_per_unit) * u * line > budget) --u;
  return u;
}
} // namespace

void Solver::configure(P3MParams const &params, Vec3d const &box_l, double skin,
                       double prefactor, MeshPrecision precision) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  hip_check(hipStreamSynchronize(m_stream), "sync");
  m->free_mesh();
  m->params = params;
  m->box_l = box_l;
  m->skin = skin;
  m->prefactor = prefactor;
  m->prec = precision;
  m->params.recalc_a_ai_cao_cut(box_l);
  m_lm.calc_local_ca_mesh(m->params, box_l, skin);
  for (int d = 0; d < 3; ++d)
    if (m_lm.inner[d] != params.mesh[d])
      throw std::runtime_error("P3M: charge-assignment mesh does not match the FFT mesh");
  int const cao = params.cao;
  auto &g = m->geo;
  for (int d = 0; d < 3; ++d) {
    g.M[d] = params.mesh[d];
    g.dim[d] = m_lm.dim[d];
    g.marg[d] = m_lm.margin[2 * d];
    g.B[d] = m_brick[d];
    g.nb[d] = (params.mesh[d] + g.B[d] - 1) / g.B[d];
    g.ld_pos[d] = m_lm.ld_pos[d];
    g.ai[d] = m->params.ai[d];
    g.box_l[d] = box_l[d];
  }
  g.Kz = params.mesh[2] / 2 + 1;
  for (int d = 0; d < 3; ++d) g.T[d] = g.B[d] + cao - 1;
  if (g.T[0] * g.T[1] * g.T[2] >= 65536)
    throw std::runtime_error("P3M: brick tile too large");
  g.tz_magic = (1ull << 32) / static_cast<unsigned long long>(g.T[2]) + 1;
  g.tyz_magic = (1ull << 32) / static_cast<unsigned long long>(g.T[1] * g.T[2]) + 1;
  g.pos_shift = std::floor((cao - 1) / 2.0) - (cao % 2) / 2.0;
  m->nbricks = g.nb[0] * g.nb[1] * g.nb[2];

  size_t const msz = (precision == MeshPrecision::FP64) ? 8 : 4;
  size_t const csz = 2 * msz;
  int64_t const R = (int64_t)g.M[0] * g.M[1] * g.M[2];
  int64_t const K = (int64_t)g.M[0] * g.M[1] * g.Kz;
  hip_check(hipMalloc(&m->rho, R * msz), "hipMalloc rho");
  hip_check(hipMalloc(&m->Ey, R * msz), "hipMalloc Ey");
  hip_check(hipMalloc(&m->Ez, R * msz), "hipMalloc Ez");
  hip_check(hipMalloc(&m->khat, K * csz), "hipMalloc khat");
  hip_check(hipMalloc(&m->A, K * csz), "hipMalloc A");
  hip_check(hipMalloc(&m->B, K * csz), "hipMalloc B");
  hip_check(hipMalloc(&m->G, K * msz), "hipMalloc G");
  hip_check(hipMalloc(&m->count, sizeof(int) * m->nbricks), "hipMalloc count");
  hip_check(hipMalloc(&m->offset, sizeof(int) * (m->nbricks + 1)), "hipMalloc offset");

  for (int d = 0; d < 3; ++d) {
    m->plan[d] = make_fft_plan(g.M[d]);
    m->dplan[d].n = m->plan[d].n;
    m->dplan[d].nstages = m->plan[d].nstages;
    for (int s = 0; s < 16; ++s) m->dplan[d].radix[s] = m->plan[d].radix[s];
    if (precision == MeshPrecision::FP64)
      upload_twiddles<double>(m->tw[d], g.M[d], m_stream);
    else
      upload_twiddles<float>(m->tw[d], g.M[d], m_stream);
  }

  // launch shapes and LDS budgets
  size_t const budget = 96 * 1024;
  m->nlp_z = pick_lines(std::max(g.M[2], 1), csz, 2, 16, budget);
  m->ncol_y = pick_lines(g.M[1], csz, 2, 16, budget);
  m->ncol_x = pick_lines(g.M[0], csz, 4, 16, budget);
  m->ncol_yi = pick_lines(g.M[1], csz, 3, 16, budget);
  auto lds_lines = [csz](int n, int lines) {
    return static_cast<size_t>(n) * csz + static_cast<size_t>(lines) * (n + 1) * csz;
  };
  m->lds_z = lds_lines(g.M[2], 2 * m->nlp_z);
  m->lds_y = lds_lines(g.M[1], 2 * m->ncol_y);
  m->lds_x = lds_lines(g.M[0], 4 * m->ncol_x);
  m->lds_yi = lds_lines(g.M[1], 3 * m->ncol_yi);
  int const tile_n = (g.B[0] + cao - 1) * (g.B[1] + cao - 1) * (g.B[2] + cao - 1);
  size_t const chunk_bytes = kChunk * (3 * cao + 1) * sizeof(double) + kChunk * sizeof(int);
  m->lds_spread = chunk_bytes + tile_n * sizeof(double);
  m->lds_gather = chunk_bytes + 3 * tile_n * msz;
  for (size_t b : {m->lds_z, m->lds_y, m->lds_x, m->lds_yi, m->lds_spread, m->lds_gather})
    if (b > lds_limit())
      throw std::runtime_error("P3M: mesh or brick too large for the LDS budget");
  m->conv_blocks = ((g.Kz + m->ncol_x - 1) / m->ncol_x) * g.M[1];
  m->gather_plane = (precision == MeshPrecision::FP64);
  hip_check(hipMalloc(&m->conv_part, sizeof(double) * 8 * m->conv_blocks), "hipMalloc conv");

  auto set_attrs = [&](auto tag) {
    using MT = decltype(tag);
    allow_lds(k_fft_z_r2c<MT>, m->lds_z);
    allow_lds(k_fft_z_c2r<MT>, m->lds_z);
    allow_lds(k_fft_y<MT>, m->lds_y);
    allow_lds(k_fft_x_conv<MT>, m->lds_x);
    allow_lds(k_fft_y_inv<MT>, m->lds_yi);
    launch_cao(cao, [&](auto P) {
      allow_lds(k_spread<P.value, MT>, m->lds_spread);
      allow_lds(k_gather<P.value, 8, MT>, m->lds_gather);
      allow_lds(k_gather_plane<P.value, 8, MT>, m->lds_gather);
    });
  };
  if (precision == MeshPrecision::FP64)
    set_attrs(double{});
  else
    set_attrs(float{});

  // influence function (FP64 arithmetic, stored in mesh precision)
  double const two_pi = 2. * std::numbers::pi;
  double const wv[3] = {two_pi * (1. / box_l[0]), two_pi * (1. / box_l[1]),
                        two_pi * (1. / box_l[2])};
  int const gg = grid_for(K, 256, 16384);
  if (precision == MeshPrecision::FP64)
    k_influence<double><<<gg, 256, 0, m_stream>>>(
        static_cast<double *>(m->G), g.M[0], g.M[1], g.M[2], g.Kz, wv[0], wv[1], wv[2],
        m->params.a[0], m->params.a[1], m->params.a[2], m->params.alpha, cao);
  else
    k_influence<float><<<gg, 256, 0, m_stream>>>(
        static_cast<float *>(m->G), g.M[0], g.M[1], g.M[2], g.Kz, wv[0], wv[1], wv[2],
        m->params.a[0], m->params.a[1], m->params.a[2], m->params.alpha, cao);
  hip_check(hipGetLastError(), "k_influence");
  hip_check(hipStreamSynchronize(m_stream), "sync influence");
  m_configured = true;
}

namespace {
void ensure_particles(Solver::Impl *m, int64_t n) {
  if (n <= m->cap_n) return;
  m->free_particles();
  int64_t const cap = std::max<int64_t>(n, 1024);
  hip_check(hipMalloc(&m->p_brick, sizeof(int) * cap), "hipMalloc p_brick");
  hip_check(hipMalloc(&m->p_rank, sizeof(int) * cap), "hipMalloc p_rank");
  hip_check(hipMalloc(&m->sidx, sizeof(int) * cap), "hipMalloc sidx");
  hip_check(hipMalloc(&m->spos, sizeof(double4) * cap), "hipMalloc spos");
  hip_check(hipMalloc(&m->bin_part, sizeof(double) * 8 * ((cap + kBlock - 1) / kBlock)),
            "hipMalloc bin_part");
  m->cap_n = cap;
}

template <typename MT>
void run_solve(Solver::Impl *m, hipStream_t st, int64_t n, double const *pos, double const *q,
               double const *upos, double *force, unsigned flags, int *p_nmp, bool stop_after_spread,
               hipEvent_t *ev) {
  auto const &g = m->geo;
  using C = cx_t<MT>;
  int const cao = m->params.cao;
  bool const want_force = (flags & 1u) && force;
  bool const want_energy = (flags & 6u) != 0;
  int const nblk = static_cast<int>((n + kBlock - 1) / kBlock);
  int64_t const R = (int64_t)g.M[0] * g.M[1] * g.M[2];
  int const rows = g.M[0] * g.M[1];

  (void)hipEventRecord(ev[0], st);
  k_fill<int><<<grid_for(m->nbricks), kBlock, 0, st>>>(m->count, 0, m->nbricks);
  k_fill<int><<<1, 64, 0, st>>>(m->err_flag, 0, 1);
  k_fill<double><<<1, 64, 0, st>>>(m->sums, 0., 32);
  if (n > 0) {
    launch_cao(cao, [&](auto P) {
      k_bin<P.value><<<nblk, kBlock, 0, st>>>(n, pos, q, upos, g, m->count, m->p_brick, m->p_rank,
                                              p_nmp, m->bin_part, m->err_flag);
    });
    k_reduce_partials<<<1, 1024, 0, st>>>(m->bin_part, nblk, 7, m->sums);
  }
  k_scan<<<1, 1024, 0, st>>>(m->count, m->offset, m->nbricks);
  if (n > 0)
    k_scatter<<<nblk, kBlock, 0, st>>>(n, pos, q, m->p_brick, m->p_rank, m->offset, m->spos,
                                       m->sidx);
  k_fill<MT><<<grid_for(R, kBlock, 8192), kBlock, 0, st>>>(static_cast<MT *>(m->rho), MT(0), R);
  (void)hipEventRecord(ev[1], st);
  launch_cao(cao, [&](auto P) {
    k_spread<P.value, MT><<<m->nbricks, kPG, m->lds_spread, st>>>(g, m->offset, m->spos,
                                                                     static_cast<MT *>(m->rho));
  });
  (void)hipEventRecord(ev[2], st);
  if (stop_after_spread) return;

  double const two_pi = 2. * std::numbers::pi;
  double const wv[3] = {two_pi * (1. / m->box_l[0]), two_pi * (1. / m->box_l[1]),
                        two_pi * (1. / m->box_l[2])};
  C *khat = static_cast<C *>(m->khat);
  C *A = static_cast<C *>(m->A);
  C *B = static_cast<C *>(m->B);
  k_fft_z_r2c<MT><<<((rows + 1) / 2 + m->nlp_z - 1) / m->nlp_z, kBlock, m->lds_z, st>>>(
      static_cast<MT const *>(m->rho), khat, rows, g.Kz, m->dplan[2],
      static_cast<C const *>(m->tw[2]), m->nlp_z);
  {
    dim3 grid((g.Kz + m->ncol_y - 1) / m->ncol_y, g.M[0]);
    k_fft_y<MT><<<grid, kBlock, m->lds_y, st>>>(khat, g.M[1], g.Kz, m->dplan[1],
                                                static_cast<C const *>(m->tw[1]), m->ncol_y, 0);
  }
  (void)hipEventRecord(ev[3], st);
  {
    dim3 grid((g.Kz + m->ncol_x - 1) / m->ncol_x, g.M[1]);
    double const hais = (1. / 2. / m->params.alpha) * (1. / 2. / m->params.alpha);
    k_fft_x_conv<MT><<<grid, kBlock, m->lds_x, st>>>(
        khat, static_cast<MT const *>(m->G), A, B, g.M[0], g.M[1], g.M[2], g.Kz, wv[0], wv[1],
        wv[2], hais, m->dplan[0], static_cast<C const *>(m->tw[0]), m->ncol_x, want_force ? 1 : 0,
        m->conv_part);
  }
  (void)hipEventRecord(ev[4], st);
  k_reduce_partials<<<1, 1024, 0, st>>>(m->conv_part, m->conv_blocks, 8, m->sums + 8);
  if (want_force) {
    dim3 grid((g.Kz + m->ncol_yi - 1) / m->ncol_yi, g.M[0]);
    k_fft_y_inv<MT><<<grid, kBlock, m->lds_yi, st>>>(A, B, khat, g.M[1], g.Kz, wv[1],
                                                     m->dplan[1],
                                                     static_cast<C const *>(m->tw[1]),
                                                     m->ncol_yi);
    int const gz = ((rows + 1) / 2 + m->nlp_z - 1) / m->nlp_z;
    MT *Ex = static_cast<MT *>(m->rho);
    k_fft_z_c2r<MT><<<gz, kBlock, m->lds_z, st>>>(B, Ex, rows, g.Kz, wv[2], 0, m->dplan[2],
                                                  static_cast<C const *>(m->tw[2]), m->nlp_z);
    k_fft_z_c2r<MT><<<gz, kBlock, m->lds_z, st>>>(khat, static_cast<MT *>(m->Ey), rows, g.Kz,
                                                  wv[2], 0, m->dplan[2],
                                                  static_cast<C const *>(m->tw[2]), m->nlp_z);
    k_fft_z_c2r<MT><<<gz, kBlock, m->lds_z, st>>>(A, static_cast<MT *>(m->Ez), rows, g.Kz,
                                                  wv[2], 1, m->dplan[2],
                                                  static_cast<C const *>(m->tw[2]), m->nlp_z);
  }
  (void)hipEventRecord(ev[5], st);
  if (want_force && n > 0) {
    double const volume = m->box_l[0] * m->box_l[1] * m->box_l[2];
    double const force_prefac = m->prefactor / volume;
    double dip_pref = 0.;
    if (m->params.epsilon != P3M_EPSILON_METALLIC)
      dip_pref = m->prefactor *
                 (4. * std::numbers::pi / volume / (2. * m->params.epsilon + 1.));
    launch_cao(cao, [&](auto P) {
      if (m->gather_plane)
        k_gather_plane<P.value, 8, MT><<<m->nbricks, kPG, m->lds_gather, st>>>(
            g, m->offset, m->spos, m->sidx, static_cast<MT const *>(m->rho),
            static_cast<MT const *>(m->Ey), static_cast<MT const *>(m->Ez), force, force_prefac,
            m->sums, dip_pref);
      else
        k_gather<P.value, 8, MT><<<m->nbricks, kPG, m->lds_gather, st>>>(
            g, m->offset, m->spos, m->sidx, static_cast<MT const *>(m->rho),
            static_cast<MT const *>(m->Ey), static_cast<MT const *>(m->Ez), force, force_prefac,
            m->sums, dip_pref);
    });
  }
  (void)hipEventRecord(ev[6], st);
  (void)want_energy;
}
} // namespace

void Solver::profile_begin(int max_solves) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  hip_check(hipStreamSynchronize(m_stream), "sync");
  size_t const need = static_cast<size_t>(std::max(max_solves, 1)) * 8;
  while (m->ring.size() < need) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    m->ring.push_back(e);
  }
  m->ring_used = 0;
  m->profiling = true;
}

int Solver::profile_end(double *stage_ms) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  hip_check(hipStreamSynchronize(m_stream), "sync");
  for (int k = 0; k < 7; ++k) stage_ms[k] = 0.;
  for (int i = 0; i < m->ring_used; ++i) {
    hipEvent_t *ev = &m->ring[8 * static_cast<size_t>(i)];
    float t = 0.f;
    (void)hipEventElapsedTime(&t, ev[0], ev[6]);
    stage_ms[0] += t;
    for (int s = 1; s <= 6; ++s) {
      (void)hipEventElapsedTime(&t, ev[s - 1], ev[s]);
      stage_ms[s] += t;
    }
  }
  m->profiling = false;
  int const n = m->ring_used;
  m->ring_used = 0;
  return n;
}

void Solver::solve(int64_t n, double const *d_pos, double const *d_q, double const *d_upos,
                   double *d_force, unsigned flags, SolveResult *res, bool sync) {
  if (!m_configured) throw std::runtime_error("P3M solver used before configure()");
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  ensure_particles(m, n);
  hipEvent_t *ev = m->ev;
  if (m->profiling && static_cast<size_t>(m->ring_used + 1) * 8 <= m->ring.size())
    ev = &m->ring[8 * static_cast<size_t>(m->ring_used++)];
  if (m->prec == MeshPrecision::FP64)
    run_solve<double>(m, m_stream, n, d_pos, d_q, d_upos, d_force, flags, nullptr, false, ev);
  else
    run_solve<float>(m, m_stream, n, d_pos, d_q, d_upos, d_force, flags, nullptr, false, ev);
  hip_check(hipGetLastError(), "P3M solve launch");
  if (!sync) return;
  hip_check(hipMemcpyAsync(m->h_sums, m->sums, 16 * sizeof(double), hipMemcpyDeviceToHost,
                           m_stream),
            "copy sums");
  hip_check(hipMemcpyAsync(m->h_sums + 16, m->err_flag, sizeof(int), hipMemcpyDeviceToHost,
                           m_stream),
            "copy err");
  hip_check(hipStreamSynchronize(m_stream), "P3M solve");
  int err = 0;
  std::memcpy(&err, m->h_sums + 16, sizeof(int));
  if (err) throw std::runtime_error("P3M: non-finite particle position");
  float t = 0.f;
  (void)hipEventElapsedTime(&t, ev[0], ev[6]);
  last_ms[0] = t;
  for (int s = 1; s <= 6; ++s) {
    (void)hipEventElapsedTime(&t, ev[s - 1], ev[s]);
    last_ms[s] = t;
  }
  if (!res) return;
  double const *h = m->h_sums;
  res->n_q = h[0];
  res->sum_q = h[1];
  res->sum_q2 = h[2];
  for (int d = 0; d < 3; ++d) res->dipole[d] = h[3 + d];
  // energy (p3m.cpp:551-589)
  double const volume = m->box_l[0] * m->box_l[1] * m->box_l[2];
  double const alpha = m->params.alpha;
  double energy = h[8] / (2. * volume);
  energy -= res->sum_q2 * alpha * std::numbers::inv_sqrtpi;
  energy -= (res->sum_q * res->sum_q) * std::numbers::pi / (2. * volume * alpha * alpha);
  if (m->params.epsilon != P3M_EPSILON_METALLIC) {
    double const pref = 4. * std::numbers::pi / volume / (2. * m->params.epsilon + 1.);
    energy += pref * (res->dipole[0] * res->dipole[0] + res->dipole[1] * res->dipole[1] +
                      res->dipole[2] * res->dipole[2]);
  }
  res->energy = (res->n_q > 0.) ? energy * m->prefactor : 0.;
  // virial tensor (p3m.cpp:393-449)
  double v[9] = {h[9] + h[15], h[10], h[11], h[10], h[12] + h[15], h[13], h[11], h[13],
                 h[14] + h[15]};
  for (int k = 0; k < 9; ++k) res->virial[k] = v[k] * m->prefactor / (2. * volume);
}

void Solver::debug_charge_mesh(int64_t n, double const *d_pos, double const *d_q,
                               std::vector<double> &out) {
  if (!m_configured) throw std::runtime_error("P3M solver used before configure()");
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  ensure_particles(m, n);
  if (m->prec == MeshPrecision::FP64)
    run_solve<double>(m, m_stream, n, d_pos, d_q, nullptr, nullptr, 0, nullptr, true, m->ev);
  else
    run_solve<float>(m, m_stream, n, d_pos, d_q, nullptr, nullptr, 0, nullptr, true, m->ev);
  hip_check(hipGetLastError(), "P3M spread launch");
  int64_t const R = (int64_t)m->geo.M[0] * m->geo.M[1] * m->geo.M[2];
  out.resize(R);
  if (m->prec == MeshPrecision::FP64) {
    hip_check(hipMemcpyAsync(out.data(), m->rho, R * sizeof(double), hipMemcpyDeviceToHost,
                             m_stream),
              "copy rho");
    hip_check(hipStreamSynchronize(m_stream), "sync");
  } else {
    std::vector<float> tmp(R);
    hip_check(hipMemcpyAsync(tmp.data(), m->rho, R * sizeof(float), hipMemcpyDeviceToHost,
                             m_stream),
              "copy rho");
    hip_check(hipStreamSynchronize(m_stream), "sync");
    for (int64_t i = 0; i < R; ++i) out[i] = tmp[i];
  }
}

void Solver::debug_indices(int64_t n, double const *d_pos, double const *d_q,
                           std::vector<int32_t> &out) {
  if (!m_configured) throw std::runtime_error("P3M solver used before configure()");
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  ensure_particles(m, n);
  int *d_nmp = nullptr;
  hip_check(hipMalloc(&d_nmp, sizeof(int) * 3 * std::max<int64_t>(n, 1)), "hipMalloc nmp");
  try {
    if (m->prec == MeshPrecision::FP64)
      run_solve<double>(m, m_stream, n, d_pos, d_q, nullptr, nullptr, 0, d_nmp, true, m->ev);
    else
      run_solve<float>(m, m_stream, n, d_pos, d_q, nullptr, nullptr, 0, d_nmp, true, m->ev);
    hip_check(hipGetLastError(), "P3M bin launch");
    out.resize(3 * n);
    hip_check(hipMemcpyAsync(out.data(), d_nmp, sizeof(int) * 3 * n, hipMemcpyDeviceToHost,
                             m_stream),
              "copy nmp");
    hip_check(hipStreamSynchronize(m_stream), "sync");
  } catch (...) {
    (void)hipFree(d_nmp);
    throw;
  }
  (void)hipFree(d_nmp);
}

} // namespace espresso_amd
