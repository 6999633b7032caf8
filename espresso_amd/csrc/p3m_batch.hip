// Descriptor-driven batched P3M solver; see p3m_batch.hpp.
#include "p3m_batch.hpp"
#include "p3m_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numbers>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>

namespace {
/** Working memory of the allocation-free solves (BatchSolver::solve_fixed):
 *  one array per device, held until the process ends.  Host code never
 *  passes `scratch` itself to a kernel: it uses the device address that
 *  hipGetSymbolAddress returns for the current device. */
constexpr size_t kFixedScratchBytes = size_t{48} << 20;
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];
} // namespace

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
  // prime factors above 8 first: the first stage has no inter-stage
  // twiddles, which lets a generic radix compute its outputs in symmetric
  // pairs.  Then few stages: radix 8, a leftover factor 2 merged with a 3
  // into a radix-6 register DFT (each stage is one LDS round trip and barrier)
  int small = 1;
  for (int f : {2, 3, 5, 7})
    while (m % f == 0) { small *= f; m /= f; }
  for (int f = 11; m > 1 && f * f <= m; f += 2) while (m % f == 0) { push(f); m /= f; }
  if (m > 1) push(m);
  m = small;
  while (m % 8 == 0) { push(8); m /= 8; }
  if (m % 4 == 0) { push(4); m /= 4; }
  if (m % 2 == 0) {
    if (m % 3 == 0) { push(6); m /= 6; } else { push(2); m /= 2; }
  }
  for (int f : {3, 5, 7}) while (m % f == 0) { push(f); m /= f; }
  return p;
}

namespace {
using namespace dev;

constexpr int kBlock = 256;  // block size of the mesh and particle-list kernels
constexpr int kChunkMax = 64; // most particles staged per brick pass (3 x chunk <= threads)
constexpr int kCountCopies = 8; // brick counter copies (one per XCD-sized share of the blocks;
                                // k_scan reads a brick's copies as two 16-byte words)
constexpr int kBrick = 8;    // first-stencil points per brick edge
constexpr int kFill = 4096;  // mesh elements zeroed per block

/** Per-system state read by every kernel. */
struct SysDesc {
  DevGeo g;
  int cao;
  int n;
  long long part_off;
  int brick_off; // into the brick counters
  int off_off;   // into the brick offsets (brick_off + system index)
  int nbricks;
  void *rho, *Ey, *Ez, *khat, *A, *B, *G;
  int tw_off[3];
  int lz, nseg; // column blocks: x layers per block and blocks per column
  int kzmajor; // 1: half spectra stored [Kz][Mx][My] and the fused y/x pass
  FftPlanDev plan[3];
  double wv[3];
  double hais;
  double force_prefac;
  double dip_pref;
};

/** slot i of the selection with wt[i] <= b < wt[i+1] */
__device__ __forceinline__ int find_slot(int const *__restrict__ wt, int nsel, int b) {
  int lo = 0, hi = nsel;
  while (hi - lo > 1) {
    int const mid = (lo + hi) >> 1;
    if (wt[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

/** Work-table index of this block.  Workgroups are dealt to the 8 XCDs
 *  round-robin by blockIdx.x; the index is remapped so that each XCD takes a
 *  contiguous range of the table (neighbouring bricks and lines, whose halo
 *  tiles overlap, then share that XCD's L2).  Bijective on [0, gridDim.x). */
__device__ __forceinline__ int xcd_block() {
  int const n = gridDim.x, b = blockIdx.x;
  int const x = b & 7, k = b >> 3;
  int const q = n >> 3, r = n & 7;
  return x * q + min(x, r) + k;
}

struct Slot {
  int slot, s, lb, vb; // table slot, system, block within the system, table index
};
template <bool XCD = true>
__device__ __forceinline__ Slot locate(int const *wt, int const *sel, int nsel) {
  Slot r;
  r.vb = XCD ? xcd_block() : static_cast<int>(blockIdx.x);
  r.slot = find_slot(wt, nsel, r.vb);
  r.s = sel[r.slot];
  r.lb = r.vb - wt[r.slot];
  return r;
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
                            double wz, double hx, double hy, double hz, double alpha, int cao,
                            int kzmajor) {
  int64_t const total = (int64_t)Mx * My * Kz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int kz, y, x;
    if (kzmajor) { // [Kz][Mx][My]
      y = static_cast<int>(e % My);
      x = static_cast<int>((e / My) % Mx);
      kz = static_cast<int>(e / ((int64_t)My * Mx));
    } else { // [Mx][My][Kz]
      kz = static_cast<int>(e % Kz);
      y = static_cast<int>((e / Kz) % My);
      x = static_cast<int>(e / ((int64_t)Kz * My));
    }
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
// particle binning (one block per 256 particles of a system)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_bin(SysDesc const *__restrict__ sys,
                                                int const *__restrict__ sel,
                                                int const *__restrict__ wt, int nsel,
                                                int const *__restrict__ soff, ParticleArrays pa,
                                                int *counts, int *p_brick, int *p_rank,
                                                int *p_nmp, double *partials, int *err_flag) {
  __shared__ double red[8 * 16];
  Slot const L = locate<false>(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  int64_t const li = (int64_t)L.lb * blockDim.x + threadIdx.x;
  double v[7] = {0., 0., 0., 0., 0., 0., 0.};
  if (li < d.n) {
    int64_t const i = d.part_off + li;     // particle arrays
    int64_t const si = soff[L.s] + li;     // lane scratch
    double const qi = pa.q[i];
    int b = -1;
    if (qi != 0.) {
      double const x[3] = {pa.pos[3 * i], pa.pos[3 * i + 1], pa.pos[3 * i + 2]};
      int nmp[3], gi[3];
      double dist[3];
      if (!particle_coordinates_rt(d.g, d.cao, x, nmp, gi, dist)) atomicOr(err_flag, 1);
      b = d.g.col ? ((gi[1] / d.g.col) * d.g.nb[2] + gi[2] / d.g.col) * d.g.M[0] + gi[0]
                  : ((gi[0] / kBrick) * d.g.nb[1] + gi[1] / kBrick) * d.g.nb[2] + gi[2] / kBrick;
      // brick counters [brick][copy]: blocks take copy blockIdx.x % 8, which
      // spreads the returning atomics of one brick over eight addresses
      int const copy = static_cast<int>(blockIdx.x) & (kCountCopies - 1);
      int const rank = atomicAdd(&counts[(d.brick_off + b) * kCountCopies + copy], 1);
      p_rank[si] = rank * kCountCopies + copy;
      if (p_nmp) {
        p_nmp[3 * li] = nmp[0];
        p_nmp[3 * li + 1] = nmp[1];
        p_nmp[3 * li + 2] = nmp[2];
      }
      double const *u = pa.upos ? pa.upos + 3 * i : x;
      v[0] = 1.;
      v[1] = qi;
      v[2] = qi * qi;
      v[3] = qi * u[0];
      v[4] = qi * u[1];
      v[5] = qi * u[2];
      v[6] = fabs(qi); // bounds the fixed-point charge mesh of deterministic solves
    } else if (p_nmp) {
      p_nmp[3 * li] = p_nmp[3 * li + 1] = p_nmp[3 * li + 2] = -1;
    }
    p_brick[si] = b;
  }
  double qmax = v[6]; // |q| of this particle (0 if uncharged)
  block_sum<7>(v, red);
  qmax = block_max(qmax, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 7; ++k) partials[L.vb * 8 + k] = v[k];
    partials[L.vb * 8 + 7] = qmax;
  }
}

/** per-system deterministic reduction of block partials: one block per
 *  selected system, sums[s*16 + base + k] = sum over the system's blocks */
__global__ void k_reduce_sys(int const *__restrict__ sel, int const *__restrict__ wt, int nsel,
                             double const *part, int nk, int base, double *sums,
                             int skip_empty) {
  __shared__ double red[16 * 8];
  int const slot = blockIdx.x;
  int const s = sel[slot];
  if (skip_empty && wt[slot + 1] == wt[slot]) return; // system handled by another pass
  double v[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  for (int b = wt[slot] + threadIdx.x; b < wt[slot + 1]; b += blockDim.x)
    for (int k = 0; k < nk; ++k) v[k] += part[b * 8 + k];
  block_sum<8>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < nk; ++k) sums[s * 16 + base + k] = v[k];
  (void)nsel;
}

/** One block per selected system: the per-system sums of the binning
 *  partials (n_q, sum q, sum q^2, dipole, sum |q|; block order,
 *  deterministic) and max |q| (slot 7), then
 *  the exclusive scan of the system's brick counts, kScanItems bricks per
 *  thread and pass.  The counters become the start of each copy's run
 *  within its brick; offsets[b] is the first particle of brick b and
 *  offsets[nb] the system's charged-particle count. */
constexpr int kScanItems = 4;
__global__ __launch_bounds__(1024) void k_scan(SysDesc const *__restrict__ sys,
                                               int const *__restrict__ sel,
                                               int const *__restrict__ wt_bin,
                                               double const *__restrict__ bin_part,
                                               double *sums, int *counts, int *offsets) {
  __shared__ int wsum[16];
  __shared__ int carry;
  __shared__ double red[16 * 8];
  int const slot = blockIdx.x;
  int const s = sel[slot];
  SysDesc const &d = sys[s];
  {
    double v[7] = {0., 0., 0., 0., 0., 0., 0.};
    double qmax = 0.;
    // four partials per thread and pass, loaded before any is summed (the
    // sums keep the order b, b + blockDim.x, ...)
    int const bend = wt_bin[slot + 1];
    for (int b = wt_bin[slot] + threadIdx.x; b < bend; b += 4 * blockDim.x) {
      double4 pv[4][2];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        int const bb = b + u * static_cast<int>(blockDim.x);
        double4 const *src = reinterpret_cast<double4 const *>(bin_part + bb * 8);
        pv[u][0] = bb < bend ? src[0] : double4{0., 0., 0., 0.};
        pv[u][1] = bb < bend ? src[1] : double4{0., 0., 0., 0.};
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        v[0] += pv[u][0].x;
        v[1] += pv[u][0].y;
        v[2] += pv[u][0].z;
        v[3] += pv[u][0].w;
        v[4] += pv[u][1].x;
        v[5] += pv[u][1].y;
        v[6] += pv[u][1].z;
        qmax = fmax(qmax, pv[u][1].w);
      }
    }
    block_sum<7>(v, red);
    qmax = block_max(qmax, red);
    if (threadIdx.x == 0) {
      for (int k = 0; k < 7; ++k) sums[s * 16 + k] = v[k];
      sums[s * 16 + 7] = qmax;
    }
  }
  int const nb = d.nbricks;
  int *count = counts + d.brick_off * kCountCopies;
  int *offset = offsets + d.off_off;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  int const lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int base = 0; base < nb; base += blockDim.x * kScanItems) {
    int const i0 = base + threadIdx.x * kScanItems;
    // a brick's kCountCopies = 8 counters are 32 contiguous, 32-byte
    // aligned bytes: two 16-byte loads
    int cc[kScanItems][kCountCopies];
    int c = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
      int4 lo4 = {0, 0, 0, 0}, hi4 = {0, 0, 0, 0};
      if (i0 + it < nb) {
        int4 const *src = reinterpret_cast<int4 const *>(count + (i0 + it) * kCountCopies);
        lo4 = src[0];
        hi4 = src[1];
      }
      cc[it][0] = lo4.x;
      cc[it][1] = lo4.y;
      cc[it][2] = lo4.z;
      cc[it][3] = lo4.w;
      cc[it][4] = hi4.x;
      cc[it][5] = hi4.y;
      cc[it][6] = hi4.z;
      cc[it][7] = hi4.w;
#pragma unroll
      for (int k = 0; k < kCountCopies; ++k) c += cc[it][k];
    }
    int sc = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      int const t = __shfl_up(sc, off, 64);
      if (lane >= off) sc += t;
    }
    if (lane == 63) wsum[wid] = sc;
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
    int o = carry + wsum[wid] + sc - c;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
      if (i0 + it < nb) {
        offset[i0 + it] = o;
        int st[kCountCopies];
#pragma unroll
        for (int k = 0; k < kCountCopies; ++k) {
          st[k] = o;
          o += cc[it][k];
        }
        int4 *dst = reinterpret_cast<int4 *>(count + (i0 + it) * kCountCopies);
        dst[0] = int4{st[0], st[1], st[2], st[3]};
        dst[1] = int4{st[4], st[5], st[6], st[7]};
      }
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += wsum[wid] + sc;
    __syncthreads();
  }
  if (threadIdx.x == 0) offset[nb] = carry;
}

__global__ __launch_bounds__(kBlock) void k_scatter(SysDesc const *__restrict__ sys,
                                                    int const *__restrict__ sel,
                                                    int const *__restrict__ wt, int nsel,
                                                    int const *__restrict__ soff,
                                                    ParticleArrays pa, int const *p_brick,
                                                    int const *p_rank, int const *counts,
                                                    double4 *spos, int *sidx) {
  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  int64_t const li = (int64_t)L.lb * blockDim.x + threadIdx.x;
  if (li >= d.n) return;
  int64_t const i = d.part_off + li;
  int64_t const si = soff[L.s] + li;
  int const b = p_brick[si];
  if (b < 0) return;
  int const rc = p_rank[si];
  int64_t const slot =
      soff[L.s] + counts[(d.brick_off + b) * kCountCopies + (rc & (kCountCopies - 1))] +
      (rc / kCountCopies);
  spos[slot] = double4{pa.pos[3 * i], pa.pos[3 * i + 1], pa.pos[3 * i + 2], pa.q[i]};
  sidx[slot] = static_cast<int>(li);
}

/** Start of a solve, one launch over the systems' mesh blocks: zero the
 *  charge mesh (unless it is kept), each block its share of the system's
 *  brick counters, and the first block the error flag. */
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_clear(SysDesc const *__restrict__ sys,
                                                  int const *__restrict__ sel,
                                                  int const *__restrict__ wt, int nsel,
                                                  int *counts, int *err_flag, int zero_rho) {
  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  if (L.vb == 0 && threadIdx.x == 0) *err_flag = 0;
  int const nblk = wt[L.slot + 1] - wt[L.slot];
  int const nc = d.nbricks * kCountCopies;
  int const per = (nc + nblk - 1) / nblk;
  int *c = counts + d.brick_off * kCountCopies;
  for (int e = L.lb * per + threadIdx.x; e < min(nc, (L.lb + 1) * per); e += blockDim.x) c[e] = 0;
  if (!zero_rho) return;
  int64_t const R = (int64_t)d.g.M[0] * d.g.M[1] * d.g.M[2];
  MT *rho = static_cast<MT *>(d.rho);
  int64_t const beg = (int64_t)L.lb * kFill;
  int64_t const end = min(beg + kFill, R);
  // 16-byte stores over the block's whole words, element stores for the
  // rest (and for a mesh that is not 16-byte aligned)
  constexpr int kPerWord = 16 / static_cast<int>(sizeof(MT));
  int64_t const words =
      (reinterpret_cast<uintptr_t>(rho + beg) & 15) ? 0 : (end - beg) / kPerWord;
  int4 *w = reinterpret_cast<int4 *>(rho + beg);
  for (int64_t e = threadIdx.x; e < words; e += blockDim.x) w[e] = int4{0, 0, 0, 0};
  for (int64_t e = beg + words * kPerWord + threadIdx.x; e < end; e += blockDim.x) rho[e] = MT(0);
}

// ---------------------------------------------------------------------------
// charge assignment into LDS brick tiles (AssignCharge<cao>, p3m.cpp:283-330;
// p3m_interpolate, interpolation.hpp:191-208).  One block per brick of
// 8x8x8 first-stencil points: the tile holds the brick plus cao-1 planes on
// the high side; contributions are added in fixed point with LDS integer
// atomics and the tile is flushed into the periodic mesh with device-scope
// atomics.
// ---------------------------------------------------------------------------
/** Bytes of the chunk slots ([chunk][3P] weights, [chunk] charges, [chunk]
 *  tile indices) rounded up to 16, so that the brick tile after them is
 *  16-byte aligned for any chunk (LDS b64 accesses must be aligned). */
__host__ __device__ constexpr int chunk_slot_bytes(int P, int chunk) {
  return (chunk * (3 * P + 1) * 8 + chunk * 4 + 15) & ~15;
}

struct ChunkSlots {
  double *w;  // [chunk][3P]
  double *qq; // [chunk]
  int *lo;    // [chunk] tile index of the first stencil point
};

/** Weights of a chunk of cnt particles, one thread per (particle, dimension):
 *  thread 3 p + d writes the P weights of dimension d of particle p; thread
 *  3 p also writes the charge and the tile index of the first stencil point
 *  (it evaluates the integer coordinates of all three dimensions). */
template <int P>
__device__ __forceinline__ void stage_chunk(DevGeo const &g, double4 const *spos, int64_t c0,
                                            int cnt, int o0, int o1, int o2, int TY, int TZ,
                                            ChunkSlots s) {
  int const t = threadIdx.x;
  if (t < 3 * cnt) {
    int const p = t / 3;
    int const dd = t - 3 * p;
    double4 const pp = spos[c0 + p];
    double const x[3] = {pp.x, pp.y, pp.z};
    int nmp, gi;
    double dist;
    coordinate_1d(g, P, x[dd], dd, nmp, gi, dist);
    bspline_weights<P>(dist, s.w + p * 3 * P + dd * P);
    if (dd == 0) {
      int n1, g1, n2, g2;
      double d1, d2;
      coordinate_1d(g, P, x[1], 1, n1, g1, d1);
      coordinate_1d(g, P, x[2], 2, n2, g2, d2);
      s.qq[p] = pp.w;
      s.lo[p] = ((gi - o0) * TY + (g1 - o1)) * TZ + (g2 - o2);
    }
  }
}

/** Row mapping of a brick tile of T^3 points (T = 8 + P - 1 <= 16) over a
 *  block of NTH threads: lane c = threadIdx.x % 16 covers z offset c (active
 *  for c < T), thread row threadIdx.x / 16 + k NTH / 16 covers tile row
 *  r = a T + b (tile index r T + c).  One periodic wrap per coordinate. */
template <int P, int NTH> struct TileRows {
  static constexpr int T = kBrick + P - 1;
  static constexpr int ROWS = T * T;
  static constexpr int RPP = NTH / 16;
  static constexpr int NPASS = (ROWS + RPP - 1) / RPP;
  static_assert(T <= 16, "tile rows wider than 16 lanes");
  int c, row0, zz;
  bool lane_ok;
  __device__ TileRows(int o2, int M2) {
    c = threadIdx.x & 15;
    row0 = threadIdx.x >> 4;
    lane_ok = c < T;
    zz = wrap_once(o2 + c, M2);
  }
  __device__ bool valid(int k) const { return lane_ok && row0 + k * RPP < ROWS; }
  __device__ int row(int k) const { return row0 + k * RPP; }
  __device__ int64_t mesh_index(int k, int o0, int o1, int M0, int M1, int M2) const {
    int const r = row(k);
    int const a = r / T;
    int const b = r - a * T;
    return ((int64_t)wrap_once(o0 + a, M0) * M1 + wrap_once(o1 + b, M1)) * M2 + zz;
  }
};

/** Lane segments over the (j, k) stencil plane: S = next power of two >= P^2
 *  lanes per particle, ppw particles per wave64. */
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

template <int P> __device__ __forceinline__ void brick_origin(SysDesc const &d, int b, int &o0,
                                                              int &o1, int &o2) {
  int const bz = b % d.g.nb[2], by = (b / d.g.nb[2]) % d.g.nb[1], bx = b / (d.g.nb[2] * d.g.nb[1]);
  o0 = bx * kBrick;
  o1 = by * kBrick;
  o2 = bz * kBrick;
}

/** Scale 2^s of the 64-bit fixed-point charge mesh of a system with charges
 *  of magnitude below 2^a (max |q|) summing to below 2^b (sum |q|).  A
 *  contribution q w_x w_y w_z 2^s is below 2^(a+s) (weights <= 1) and must
 *  stay below 2^51 for fixed_round; a mesh point or partial sum is below
 *  2^(b+s), kept below 2^62.  s = min(50 - a, 62 - b). */
__host__ __device__ inline double fixed_scale(double sum_abs_q, double max_abs_q) {
  int a = 0, b = 0;
  (void)frexp(max_abs_q, &a); // |q| < 2^a (a = 0 for no charges)
  (void)frexp(sum_abs_q, &b);
  return ldexp(1., (50 - a < 62 - b) ? 50 - a : 62 - b);
}

/** a b rounded to the nearest integer (ties to even) for |a b| < 2^51, as a
 *  two's-complement 64-bit value: a b + 1.5 2^52 lies in [2^52, 2^53), where
 *  the spacing of doubles is 1, so the fused multiply-add rounds the exact
 *  product once, and the bits of the sum are those of 1.5 2^52 plus the
 *  rounded integer. */
template <typename TT = unsigned long long>
__device__ __forceinline__ TT fixed_round(double a, double b) {
  double const y = fma(a, b, 6755399441055744.0); // 1.5 2^52
  // the low 32 bits of 1.5 2^52 are zero: a 32-bit result (|a b| < 2^31) is
  // the low word of the sum
  return static_cast<TT>(static_cast<unsigned long long>(__double_as_longlong(y)) -
                         0x4338000000000000ull);
}

/** LDS tile word of the charge assignment: 64-bit fixed point for FP64
 *  meshes; 32-bit for FP32 meshes, at a per-brick scale (brick_scale32),
 *  which halves the tile and its atomic traffic (ds_add_u32 runs at about
 *  1.6x the rate of ds_add_u64 on the stencil pattern, ds_add_f32 at a
 *  twentieth). */
template <typename MT> struct SpreadWord {
  using type = unsigned long long;
};
template <> struct SpreadWord<float> {
  using type = unsigned int;
};

/** Upper bounds of w^3 for the weights w of an order-P charge assignment:
 *  the cube of the peak of the cardinal B-spline of order P (1, 1, 3/4, 2/3,
 *  115/192, 11/20, 5887/11520 for P = 1..7), rounded up. */
__constant__ double kPeakCube[8] = {1., 1., 1., 0.4219, 0.2963, 0.2149, 0.1664, 0.1335};

/** Scale 2^s of a brick's 32-bit fixed-point tile: each of the brick's n
 *  particles (|q| <= max_abs_q) adds at most |q| kPeakCube[P] to a tile
 *  point, so every tile point and contribution stays below 2^(e+s) = 2^30
 *  for n max_abs_q kPeakCube[P] < 2^e.  The resolution 2^-s (2^-27 for a
 *  brick of 49 unit charges at P = 7) is below the FP32 mesh's own rounding
 *  once the tile is flushed. */
__device__ __forceinline__ double brick_scale32(int n, double max_abs_q, int P) {
  int e = 0;
  (void)frexp(static_cast<double>(n) * max_abs_q * kPeakCube[P], &e);
  return ldexp(1., 30 - e);
}

/** Tile rows of an SX x SY x SZ group of bricks over a block of NTH threads:
 *  LPR lanes per row of TZ points (z), rows r = a TY + b.  One periodic
 *  wrap per coordinate. */
template <int P, int NTH, int SX, int SY, int SZ> struct GroupRows {
  static constexpr int TX = SX * kBrick + P - 1, TY = SY * kBrick + P - 1,
                       TZ = SZ * kBrick + P - 1;
  static constexpr int LPR = (TZ <= 16) ? 16 : 32;
  static_assert(TZ <= 32, "tile rows wider than 32 lanes");
  static constexpr int RPP = NTH / LPR;
  static constexpr int ROWS = TX * TY;
  static constexpr int NPASS = (ROWS + RPP - 1) / RPP;
  int c, row0, zz;
  bool lane_ok;
  __device__ GroupRows(int o2, int M2) {
    c = threadIdx.x % LPR;
    row0 = threadIdx.x / LPR;
    lane_ok = c < TZ;
    zz = wrap_once(o2 + c, M2);
  }
  __device__ bool valid(int k) const { return lane_ok && row0 + k * RPP < ROWS; }
  __device__ int row(int k) const { return row0 + k * RPP; }
  __device__ int64_t mesh_index(int k, int o0, int o1, int M0, int M1, int M2) const {
    int const r = row(k);
    int const a = r / TY;
    int const b = r - a * TY;
    return ((int64_t)wrap_once(o0 + a, M0) * M1 + wrap_once(o1 + b, M1)) * M2 + zz;
  }
};

/** Blocks of the spread of a system with bricks nb[3] grouped SX x SY x SZ. */
__host__ __device__ inline int brick_groups(int const nb[3], int sx, int sy, int sz) {
  return ((nb[0] + sx - 1) / sx) * ((nb[1] + sy - 1) / sy) * ((nb[2] + sz - 1) / sz);
}

/** Charge assignment of one group of SX x SY x SZ bricks (one brick by
 *  default; a group shares one tile and its halo, which pays for sparse
 *  bricks; a 32-bit tile's resolution falls with the group's particle
 *  count):
 *  contributions are rounded once to fixed point (fixed_round; 64-bit at
 *  fixed_scale of the system's charges for FP64 meshes, 32-bit at
 *  brick_scale32 for FP32 meshes) and summed in the LDS tile with integer
 *  atomics (faster than FP atomics on this pattern, and integer sums do not
 *  depend on the order).  The tile is flushed into the periodic mesh with
 *  device atomics: as 64-bit integers into a fixed-point mesh (int_mesh,
 *  FP64 meshes: the whole charge assignment is then independent of the order
 *  of particles, bricks and atomics), or converted to the mesh type.  Lanes
 *  over the (j, k) stencil plane in segments of up to SEGMAX lanes per
 *  particle. */
template <int P, typename MT, int SEGMAX, int NTH, int SX = 1, int SY = 1, int SZ = 1>
__global__ __launch_bounds__(NTH, SEGMAX == 64 ? 8 : 1) void k_spread(SysDesc const *__restrict__ sys,
                                               int const *__restrict__ sel,
                                               int const *__restrict__ wt, int nsel,
                                               int const *__restrict__ soff,
                                               int const *__restrict__ offsets,
                                               double4 const *__restrict__ spos, int chunk,
                                               double const *__restrict__ sums, int int_mesh) {
  using GR = GroupRows<P, NTH, SX, SY, SZ>;
  constexpr int TY = GR::TY, TZ = GR::TZ;
  constexpr int TYZ = TY * TZ;
  constexpr int tile_n = GR::TX * TYZ;
  constexpr int NSUB = SX * SY * SZ;
  extern __shared__ double smem_d[];
  ChunkSlots s;
  s.w = smem_d;
  s.qq = s.w + chunk * 3 * P;
  s.lo = reinterpret_cast<int *>(s.qq + chunk);
  using TT = typename SpreadWord<MT>::type;
  TT *tile = reinterpret_cast<TT *>(reinterpret_cast<char *>(smem_d) + chunk_slot_bytes(P, chunk));

  // the block's group of SX x SY x SZ bricks and their particle runs
  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  int const nb0 = d.g.nb[0], nb1 = d.g.nb[1], nb2 = d.g.nb[2];
  int const ng2 = (nb2 + SZ - 1) / SZ, ng1 = (nb1 + SY - 1) / SY;
  int const gz = L.lb % ng2, gy = (L.lb / ng2) % ng1, gx = L.lb / (ng2 * ng1);
  int const o0 = gx * SX * kBrick, o1 = gy * SY * kBrick, o2 = gz * SZ * kBrick;
  int rb[NSUB], re[NSUB];
  int ntot = 0;
#pragma unroll
  for (int u = 0; u < NSUB; ++u) {
    int const bx = gx * SX + u / (SY * SZ), by = gy * SY + (u / SZ) % SY, bz = gz * SZ + u % SZ;
    rb[u] = re[u] = 0;
    if (bx < nb0 && by < nb1 && bz < nb2) {
      int const b = (bx * nb1 + by) * nb2 + bz;
      rb[u] = offsets[d.off_off + b];
      re[u] = offsets[d.off_off + b + 1];
    }
    ntot += re[u] - rb[u];
  }
  if (ntot == 0) return;
  double const scale = (sizeof(TT) == 8)
                           ? fixed_scale(sums[L.s * 16 + 6], sums[L.s * 16 + 7])
                           : brick_scale32(ntot, sums[L.s * 16 + 7], P);
  double4 const *sp = spos + soff[L.s];
  constexpr int S = (Seg<P>::S < SEGMAX) ? Seg<P>::S : SEGMAX;
  constexpr int ppw = 64 / S;
  constexpr int NPAIR = P * P;
  constexpr int NT = (NPAIR + S - 1) / S;
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const nwaves = blockDim.x >> 6;
  int const sub = lane / S;
  int const r0 = lane - sub * S;
  for (int e = threadIdx.x; e < tile_n; e += blockDim.x) tile[e] = TT(0);
  bool first = true;
#pragma unroll 1
  for (int u = 0; u < NSUB; ++u) {
    for (int c0 = rb[u]; c0 < re[u]; c0 += chunk) {
      int const cnt = min(chunk, re[u] - c0);
      if (!first) __syncthreads(); // the tile zeroing and the slots are disjoint
      first = false;
      stage_chunk<P>(d.g, sp, c0, cnt, o0, o1, o2, TY, TZ, s);
      __syncthreads();
      int const ngroups = (cnt + ppw - 1) / ppw;
      for (int grp = wave; grp < ngroups; grp += nwaves) {
        int const p = grp * ppw + sub;
        if (p < cnt) {
          double const *w = s.w + p * 3 * P;
          double const qs = s.qq[p] * scale;
          int const lo = s.lo[p];
          // x weights in registers: the LDS atomics below would otherwise
          // make the compiler re-read them from LDS after every atomic
          double wx[P];
#pragma unroll
          for (int i = 0; i < P; ++i) wx[i] = w[i];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            int const r = r0 + t * S;
            if (r < NPAIR) {
              int const jj = r / P, kk = r - (r / P) * P;
              double const wyzq = (w[P + jj] * w[2 * P + kk]) * qs;
              TT *dst = tile + lo + jj * TZ + kk;
#pragma unroll
              for (int i = 0; i < P; ++i) atomicAdd(dst + i * TYZ, fixed_round<TT>(wx[i], wyzq));
            }
          }
        }
      }
    }
  }
  __syncthreads();
  int const M0 = d.g.M[0], M1 = d.g.M[1], M2 = d.g.M[2];
  double const inv_scale = 1. / scale;
  GR const tr(o2, M2);
#pragma unroll
  for (int k = 0; k < GR::NPASS; ++k) {
    if (tr.valid(k)) {
      TT const v = tile[tr.row(k) * TZ + tr.c];
      if (v == TT(0)) continue;
      int64_t const g = tr.mesh_index(k, o0, o1, M0, M1, M2);
      double const sv = (sizeof(TT) == 8) ? static_cast<double>(static_cast<long long>(v))
                                          : static_cast<double>(static_cast<int>(v));
      if (sizeof(TT) == 8 && int_mesh)
        atomicAdd(static_cast<unsigned long long *>(d.rho) + g,
                  static_cast<unsigned long long>(v));
      else
        unsafeAtomicAdd(static_cast<MT *>(d.rho) + g, static_cast<MT>(sv * inv_scale));
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
__global__ __launch_bounds__(kBlock) void k_fft_z_r2c(SysDesc const *__restrict__ sys,
                                                      int const *__restrict__ sel,
                                                      int const *__restrict__ wt, int nsel,
                                                      cx_t<MT> const *__restrict__ tw, int nlp,
                                                      double const *__restrict__ sums, int det) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  // deterministic solves: the charge mesh holds 64-bit fixed-point sums
  double const inv_scale =
      det ? 1. / fixed_scale(sums[L.s * 16 + 6], sums[L.s * 16 + 7]) : 0.;
  long long const *rho_fx = static_cast<long long const *>(d.rho);
  FftPlanDev const &plan = d.plan[2];
  int const n = plan.n;
  int const ls = n + 1;
  int const rows = d.g.M[0] * d.g.M[1];
  int const Kz = d.g.Kz;
  FastDiv const fn = make_fastdiv(n), fnl = make_fastdiv(nlp), fkz = make_fastdiv(Kz);
  MT const *rho = static_cast<MT const *>(d.rho);
  C *khat = static_cast<C *>(d.khat);
  C *twl = reinterpret_cast<C *>(smem_d);
  C *b0 = twl + n;
  C *b1 = b0 + nlp * ls;
  load_twiddles(twl, tw + d.tw_off[2], n);
  int const pair0 = L.lb * nlp;
  for (int e = threadIdx.x; e < nlp * n; e += blockDim.x) {
    int const l = fdiv(e, fn), j = e - l * n;
    int64_t const r0 = 2 * (int64_t)(pair0 + l), r1 = r0 + 1;
    MT re, im;
    if (det) {
      re = (r0 < rows) ? static_cast<MT>(static_cast<double>(rho_fx[r0 * n + j]) * inv_scale) : MT(0);
      im = (r1 < rows) ? static_cast<MT>(static_cast<double>(rho_fx[r1 * n + j]) * inv_scale) : MT(0);
    } else {
      re = (r0 < rows) ? rho[r0 * n + j] : MT(0);
      im = (r1 < rows) ? rho[r1 * n + j] : MT(0);
    }
    b0[l * ls + j] = C{re, im};
  }
  __syncthreads();
  C *res = lds_fft(b0, b1, nlp, ls, plan, twl, false);
  bool const kzm = d.kzmajor != 0;
  for (int e = threadIdx.x; e < nlp * Kz; e += blockDim.x) {
    // kz-major spectra: consecutive threads take consecutive rows
    int const q0 = fdiv(e, kzm ? fnl : fkz);
    int const l = kzm ? e - q0 * nlp : q0;
    int const k = kzm ? q0 : e - q0 * Kz;
    int64_t const r0 = 2 * (int64_t)(pair0 + l), r1 = r0 + 1;
    C const Z = res[l * ls + k];
    C const Zc = cconj(res[l * ls + (k == 0 ? 0 : n - k)]);
    C const sm = cadd(Z, Zc), df = csub(Z, Zc);
    int64_t const i0 = kzm ? (int64_t)k * rows + r0 : r0 * Kz + k;
    int64_t const i1 = kzm ? i0 + 1 : i0 + Kz;
    if (r0 < rows) khat[i0] = C{MT(0.5) * sm.x, MT(0.5) * sm.y};
    if (r1 < rows) khat[i1] = C{MT(0.5) * df.y, MT(-0.5) * df.x};
  }
}

/** forward c2c along y for a tile of kz columns at fixed x (in place) */
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_fft_y(SysDesc const *__restrict__ sys,
                                                  int const *__restrict__ sel,
                                                  int const *__restrict__ wt, int nsel,
                                                  cx_t<MT> const *__restrict__ tw, int ncol) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  FftPlanDev const &plan = d.plan[1];
  int const n = plan.n;
  int const ls = n + 1;
  int const Kz = d.g.Kz;
  int const ncb = (Kz + ncol - 1) / ncol;
  FastDiv const fnc = make_fastdiv(ncol);
  int const cb = L.lb % ncb;
  int64_t const x = L.lb / ncb;
  int const My = d.g.M[1];
  C *khat = static_cast<C *>(d.khat);
  C *twl = reinterpret_cast<C *>(smem_d);
  C *b0 = twl + n;
  C *b1 = b0 + ncol * ls;
  load_twiddles(twl, tw + d.tw_off[1], n);
  int const c0 = cb * ncol;
  int const nc = min(ncol, Kz - c0);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = fdiv(e, fnc), c = e - j * ncol;
    if (c < nc) b0[c * ls + j] = khat[(x * My + j) * Kz + c0 + c];
  }
  __syncthreads();
  C *res = lds_fft(b0, b1, nc, ls, plan, twl, false);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = fdiv(e, fnc), c = e - j * ncol;
    if (c < nc) khat[(x * My + j) * Kz + c0 + c] = res[c * ls + j];
  }
}

/** forward c2c along x, convolution with the influence function, energy and
 *  virial partial sums (p3m.cpp:393-449, 551-567), ik along x and inverse c2c
 *  along x (p3m.cpp:491-522): A = IFFT_x(G rho^), B = IFFT_x(i k_x G rho^). */
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_fft_x_conv(SysDesc const *__restrict__ sys,
                                                       int const *__restrict__ sel,
                                                       int const *__restrict__ wt, int nsel,
                                                       cx_t<MT> const *__restrict__ tw, int ncol,
                                                       int want_fields, int want_energy,
                                                       double *partials) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  __shared__ double red[8 * 16];
  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  FftPlanDev const &plan = d.plan[0];
  int const n = plan.n; // == Mx
  int const ls = n + 1;
  int const Kz = d.g.Kz;
  int const Mx = d.g.M[0], My = d.g.M[1], Mz = d.g.M[2];
  int const ncb = (Kz + ncol - 1) / ncol;
  FastDiv const fnc = make_fastdiv(ncol), fn = make_fastdiv(n);
  int const cb = L.lb % ncb;
  int const y = L.lb / ncb;
  C const *khat = static_cast<C const *>(d.khat);
  MT const *G = static_cast<MT const *>(d.G);
  C *Aout = static_cast<C *>(d.A);
  C *Bout = static_cast<C *>(d.B);
  C *twl = reinterpret_cast<C *>(smem_d);
  C *X = twl + n;            // 2*ncol lines
  C *Y = X + 2 * ncol * ls;  // 2*ncol lines
  load_twiddles(twl, tw + d.tw_off[0], n);
  int const c0 = cb * ncol;
  int const nc = min(ncol, Kz - c0);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = fdiv(e, fnc), c = e - j * ncol;
    if (c < nc) X[c * ls + j] = khat[((int64_t)j * My + y) * Kz + c0 + c];
  }
  __syncthreads();
  C *res = lds_fft(X, Y, nc, ls, plan, twl, false);
  C *oth = (res == X) ? Y : X;
  double acc[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  double const wvx = d.wv[0], wvy = d.wv[1], wvz = d.wv[2];
  double const ky = dop_shift(y, My) * wvy;
  for (int e = threadIdx.x; e < n * nc; e += blockDim.x) {
    int const c = fdiv(e, fn), j = e - c * n;
    int const kz = c0 + c;
    C const rh = res[c * ls + j];
    MT const gv = G[((int64_t)j * My + y) * Kz + kz];
    double const wgt = (kz == 0 || 2 * kz == Mz) ? 1. : 2.;
    double const ek = static_cast<double>(gv) *
                      (static_cast<double>(rh.x) * rh.x + static_cast<double>(rh.y) * rh.y);
    if (want_energy) {
      double const kx = dop_shift(j, Mx) * wvx;
      double const kzv = dop_shift(kz, Mz) * wvz;
      double const norm_sq = kx * kx + ky * ky + kzv * kzv;
      acc[0] += wgt * ek;
      if (norm_sq != 0.) {
        double const pref = wgt * ek * (-2. * (1. / norm_sq + d.hais));
        acc[1] += pref * kx * kx;
        acc[2] += pref * kx * ky;
        acc[3] += pref * kx * kzv;
        acc[4] += pref * ky * ky;
        acc[5] += pref * ky * kzv;
        acc[6] += pref * kzv * kzv;
        acc[7] += wgt * ek;
      }
    }
    if (want_fields) {
      C const phi = C{gv * rh.x, gv * rh.y};
      MT const kxf = static_cast<MT>(ik_shift(j, Mx, d.g.conv)) * static_cast<MT>(wvx);
      oth[c * ls + j] = phi;                                   // A
      oth[(ncol + c) * ls + j] = C{-kxf * phi.y, kxf * phi.x}; // B = i kx phi
    }
  }
  if (want_energy) {
    block_sum<8>(acc, red); // ends with a barrier
    if (threadIdx.x == 0)
      for (int k = 0; k < 8; ++k) partials[L.vb * 8 + k] = acc[k];
  } else {
    __syncthreads();
  }
  if (!want_fields) return;
  C *ires = lds_fft(oth, res, 2 * ncol, ls, plan, twl, true);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = fdiv(e, fnc), c = e - j * ncol;
    if (c < nc) {
      int64_t const gidx = ((int64_t)j * My + y) * Kz + c0 + c;
      Aout[gidx] = ires[c * ls + j];
      Bout[gidx] = ires[(ncol + c) * ls + j];
    }
  }
}

/** inverse c2c along y of the x-transformed fields, in place:
 *  A <- IFFT_y(A) (z field before i k_z), B <- IFFT_y(B) (x field),
 *  khat <- IFFT_y(i k_y A) (y field). */
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_fft_y_inv(SysDesc const *__restrict__ sys,
                                                      int const *__restrict__ sel,
                                                      int const *__restrict__ wt, int nsel,
                                                      cx_t<MT> const *__restrict__ tw, int ncol) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  FftPlanDev const &plan = d.plan[1];
  int const n = plan.n; // == My
  int const ls = n + 1;
  int const Kz = d.g.Kz;
  int const My = d.g.M[1];
  int const ncb = (Kz + ncol - 1) / ncol;
  FastDiv const fnc = make_fastdiv(ncol), fn = make_fastdiv(n);
  int const cb = L.lb % ncb;
  int64_t const x = L.lb / ncb;
  C *A = static_cast<C *>(d.A);
  C *B = static_cast<C *>(d.B);
  C *Fy = static_cast<C *>(d.khat);
  double const wvy = d.wv[1];
  C *twl = reinterpret_cast<C *>(smem_d);
  C *S = twl + n;
  C *W0 = S + ncol * ls;
  C *W1 = W0 + ncol * ls;
  load_twiddles(twl, tw + d.tw_off[1], n);
  int const c0 = cb * ncol;
  int const nc = min(ncol, Kz - c0);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = fdiv(e, fnc), c = e - j * ncol;
    if (c < nc) {
      C const v = A[(x * My + j) * Kz + c0 + c];
      S[c * ls + j] = v;
      W0[c * ls + j] = v;
    }
  }
  __syncthreads();
  C *res = lds_fft(W0, W1, nc, ls, plan, twl, true);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = fdiv(e, fnc), c = e - j * ncol;
    if (c < nc) A[(x * My + j) * Kz + c0 + c] = res[c * ls + j];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * nc; e += blockDim.x) {
    int const c = fdiv(e, fn), j = e - c * n;
    MT const kyf = static_cast<MT>(ik_shift(j, My, d.g.conv)) * static_cast<MT>(wvy);
    C const v = S[c * ls + j];
    W0[c * ls + j] = C{-kyf * v.y, kyf * v.x};
  }
  __syncthreads();
  res = lds_fft(W0, W1, nc, ls, plan, twl, true);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = fdiv(e, fnc), c = e - j * ncol;
    if (c < nc) Fy[(x * My + j) * Kz + c0 + c] = res[c * ls + j];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = fdiv(e, fnc), c = e - j * ncol;
    if (c < nc) W0[c * ls + j] = B[(x * My + j) * Kz + c0 + c];
  }
  __syncthreads();
  res = lds_fft(W0, W1, nc, ls, plan, twl, true);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    int const j = fdiv(e, fnc), c = e - j * ncol;
    if (c < nc) B[(x * My + j) * Kz + c0 + c] = res[c * ls + j];
  }
}

/** Fused y/x pass for meshes whose x-y plane fits in LDS twice (kz-major
 *  spectra), one block per kz column: forward c2c along y and x, convolution
 *  with the influence function with energy and virial partial sums
 *  (p3m.cpp:393-449, 551-567), then the inverse transforms of the three field
 *  components (i k_x phi, i k_y phi, phi) along x and y.  k_y and k_z are
 *  constant along an x line, so the y and z components share one inverse x
 *  transform a = IFFT_x(phi): five inverse line passes instead of six.  The
 *  block's own khat column keeps phi, then a, between the passes (two LDS
 *  planes instead of three).  The plane stays in [x][y] order throughout:
 *  y transforms run along its rows, x transforms down its columns
 *  (lds_fft_cols), so no pass transposes it.  Outputs: B <- x field, khat <- y field,
 *  A <- z field before i k_z (which the c2r z pass applies), all still
 *  transformed along z. */
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_fft_yx_conv(SysDesc const *__restrict__ sys,
                                                        int const *__restrict__ sel,
                                                        int const *__restrict__ wt, int nsel,
                                                        cx_t<MT> const *__restrict__ tw,
                                                        int want_fields, int want_energy,
                                                        double *partials) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  __shared__ double red[8 * 16];
  Slot const L = locate<false>(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  int const kz = L.lb;
  int const Mx = d.g.M[0], My = d.g.M[1], Mz = d.g.M[2];
  int const lsy = My + 1, lsx = Mx + 1; // line strides of [x][y] and [y][x] planes
  int const plane = max(Mx * lsy, My * lsx);
  int const pl = Mx * My;
  FastDiv const fMy = make_fastdiv(My);
  C *twx = reinterpret_cast<C *>(smem_d);
  C *twy = twx + Mx;
  C *P0 = twy + My;
  C *P1 = P0 + plane;
  load_twiddles(twx, tw + d.tw_off[0], Mx);
  load_twiddles(twy, tw + d.tw_off[1], My);
  C *const col = static_cast<C *>(d.khat) + (int64_t)kz * pl; // this block's column
  for (int e = threadIdx.x; e < pl; e += blockDim.x) {
    int const x = fdiv(e, fMy), y = e - x * My;
    P0[x * lsy + y] = col[e];
  }
  __syncthreads();
  C *r = lds_fft(P0, P1, Mx, lsy, d.plan[1], twy, false); // along y
  C *o = (r == P0) ? P1 : P0;
  C *q = lds_fft_cols(r, o, My, lsy, d.plan[0], twx, false); // along x
  C *w = (q == r) ? o : r;

  double acc[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  double const wvx = d.wv[0], wvy = d.wv[1], wvz = d.wv[2];
  double const wgt = (kz == 0 || 2 * kz == Mz) ? 1. : 2.;
  double const kzv = dop_shift(kz, Mz) * wvz;
  MT const *G = static_cast<MT const *>(d.G) + (int64_t)kz * pl;
  for (int e = threadIdx.x; e < pl; e += blockDim.x) {
    int const x = fdiv(e, fMy), y = e - x * My;
    C const rh = q[x * lsy + y];
    MT const gv = G[e];
    if (want_energy) {
      double const ek = static_cast<double>(gv) *
                        (static_cast<double>(rh.x) * rh.x + static_cast<double>(rh.y) * rh.y);
      double const kx = dop_shift(x, Mx) * wvx;
      double const ky = dop_shift(y, My) * wvy;
      double const norm_sq = kx * kx + ky * ky + kzv * kzv;
      acc[0] += wgt * ek;
      if (norm_sq != 0.) {
        double const pref = wgt * ek * (-2. * (1. / norm_sq + d.hais));
        acc[1] += pref * kx * kx;
        acc[2] += pref * kx * ky;
        acc[3] += pref * kx * kzv;
        acc[4] += pref * ky * ky;
        acc[5] += pref * ky * kzv;
        acc[6] += pref * kzv * kzv;
        acc[7] += wgt * ek;
      }
    }
    q[x * lsy + y] = C{gv * rh.x, gv * rh.y};
  }
  if (want_energy) {
    block_sum<8>(acc, red); // ends with a barrier: q is complete
    if (threadIdx.x == 0)
      for (int k = 0; k < 8; ++k) partials[L.vb * 8 + k] = acc[k];
  } else {
    __syncthreads(); // q is complete
  }
  if (!want_fields) return;
  C *const Bp = static_cast<C *>(d.B) + (int64_t)kz * pl;
  C *const Ap = static_cast<C *>(d.A) + (int64_t)kz * pl;
  // phi = G rho^ to the column and i k_x phi into w
  for (int e = threadIdx.x; e < pl; e += blockDim.x) {
    int const x = fdiv(e, fMy), y = e - x * My;
    C const v = q[x * lsy + y];
    col[e] = v;
    MT const kf = static_cast<MT>(ik_shift(x, Mx, d.g.conv)) * static_cast<MT>(wvx);
    w[x * lsy + y] = C{-kf * v.y, kf * v.x};
  }
  __syncthreads();
  // x component: inverse along x, then along y, into B
  C *a = lds_fft_cols(w, q, My, lsy, d.plan[0], twx, true);
  C *b = (a == w) ? q : w;
  C *c = lds_fft(a, b, Mx, lsy, d.plan[1], twy, true);
  for (int e = threadIdx.x; e < pl; e += blockDim.x) {
    int const x = fdiv(e, fMy), y = e - x * My;
    Bp[e] = c[x * lsy + y];
  }
  __syncthreads(); // c may be P0
  // a = IFFT_x(phi), shared by the y and z components
  for (int e = threadIdx.x; e < pl; e += blockDim.x) {
    int const x = fdiv(e, fMy), y = e - x * My;
    P0[x * lsy + y] = col[e];
  }
  __syncthreads();
  a = lds_fft_cols(P0, P1, My, lsy, d.plan[0], twx, true);
  b = (a == P0) ? P1 : P0;
  // a to the column (the first stage below reads a and writes b only;
  // its barrier precedes any write to a)
  for (int e = threadIdx.x; e < pl; e += blockDim.x) {
    int const x = fdiv(e, fMy), y = e - x * My;
    col[e] = a[x * lsy + y];
  }
  // z component (i k_z is applied by the c2r pass): inverse along y into A
  c = lds_fft(a, b, Mx, lsy, d.plan[1], twy, true);
  for (int e = threadIdx.x; e < pl; e += blockDim.x) {
    int const x = fdiv(e, fMy), y = e - x * My;
    Ap[e] = c[x * lsy + y];
  }
  __syncthreads(); // c may be P0
  // y component: i k_y a, inverse along y into this column (khat)
  for (int e = threadIdx.x; e < pl; e += blockDim.x) {
    int const x = fdiv(e, fMy), y = e - x * My;
    C const v = col[e];
    MT const kf = static_cast<MT>(ik_shift(y, My, d.g.conv)) * static_cast<MT>(wvy);
    P0[x * lsy + y] = C{-kf * v.y, kf * v.x};
  }
  __syncthreads();
  c = lds_fft(P0, P1, Mx, lsy, d.plan[1], twy, true);
  for (int e = threadIdx.x; e < pl; e += blockDim.x) {
    int const x = fdiv(e, fMy), y = e - x * My;
    col[e] = c[x * lsy + y];
  }
}

/** inverse c2r along z of the three field components (x from B into rho,
 *  y from khat into Ey, z from A times i k_z into Ez), two half-spectrum
 *  rows per complex line.  NC = 3: a block takes the same nlp line pairs of
 *  all three components (shared index arithmetic, six loads per element in
 *  flight, one batched LDS transform of 3 nlp lines); NC = 1: blockIdx.y
 *  selects the component (long lines, whose three-component blocks would
 *  not fit the LDS budget with enough lines). */
template <typename MT, int NC>
__global__ __launch_bounds__(kBlock) void k_fft_z_c2r(SysDesc const *__restrict__ sys,
                                                      int const *__restrict__ sel,
                                                      int const *__restrict__ wt, int nsel,
                                                      cx_t<MT> const *__restrict__ tw, int nlp) {
  using C = cx_t<MT>;
  extern __shared__ double smem_d[];
  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  FftPlanDev const &plan = d.plan[2];
  int const n = plan.n; // == Mz
  int const ls = n + 1;
  int const rows = d.g.M[0] * d.g.M[1];
  int const Kz = d.g.Kz;
  FastDiv const fn = make_fastdiv(n), fnl = make_fastdiv(nlp);
  int const c0 = (NC == 3) ? 0 : static_cast<int>(blockIdx.y);
  C const *S[3] = {static_cast<C const *>(d.B), static_cast<C const *>(d.khat),
                   static_cast<C const *>(d.A)};
  MT *E[3] = {static_cast<MT *>(d.rho), static_cast<MT *>(d.Ey), static_cast<MT *>(d.Ez)};
  double const wvz = d.wv[2];
  C *twl = reinterpret_cast<C *>(smem_d);
  C *b0 = twl + n;
  C *b1 = b0 + NC * nlp * ls;
  load_twiddles(twl, tw + d.tw_off[2], n);
  int const pair0 = L.lb * nlp;
  bool const kzm = d.kzmajor != 0;
  int const cstride = nlp * ls; // component c0 + c occupies lines [c nlp, (c + 1) nlp)
  for (int e = threadIdx.x; e < nlp * n; e += blockDim.x) {
    int const q0 = fdiv(e, kzm ? fnl : fn);
    int const l = kzm ? e - q0 * nlp : q0;
    int const k = kzm ? q0 : e - q0 * n;
    int64_t const r0 = 2 * (int64_t)(pair0 + l), r1 = r0 + 1;
    bool const upper = (k >= Kz);
    int const ks = upper ? n - k : k;
    int64_t const i0 = kzm ? (int64_t)ks * rows + r0 : r0 * Kz + ks;
    int64_t const i1 = kzm ? i0 + 1 : i0 + Kz;
    bool const v0 = r0 < rows, v1 = r1 < rows;
    C const zero{MT(0), MT(0)};
    C x[NC][2];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      x[c][0] = v0 ? S[c0 + c][i0] : zero;
      x[c][1] = v1 ? S[c0 + c][i1] : zero;
    }
    MT const kz = static_cast<MT>(ik_shift(ks, n, d.g.conv)) * static_cast<MT>(wvz);
    bool const real_row = (ks == 0 || 2 * ks == n);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      C x1 = x[c][0], x2 = x[c][1];
      if (c0 + c == 2) { // z component: times i k_z
        x1 = C{-kz * x1.y, kz * x1.x};
        x2 = C{-kz * x2.y, kz * x2.x};
      }
      if (real_row) {
        x1.y = MT(0);
        x2.y = MT(0);
      }
      if (upper) {
        x1 = cconj(x1);
        x2 = cconj(x2);
      }
      b0[c * cstride + l * ls + k] = C{x1.x - x2.y, x1.y + x2.x};
    }
  }
  __syncthreads();
  C *res = lds_fft(b0, b1, NC * nlp, ls, plan, twl, true);
  for (int e = threadIdx.x; e < nlp * n; e += blockDim.x) {
    int const l = fdiv(e, fn), j = e - l * n;
    int64_t const r0 = 2 * (int64_t)(pair0 + l), r1 = r0 + 1;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      C const v = res[c * cstride + l * ls + j];
      if (r0 < rows) E[c0 + c][r0 * n + j] = v.x;
      if (r1 < rows) E[c0 + c][r1 * n + j] = v.y;
    }
  }
}

// ---------------------------------------------------------------------------
// force back-interpolation (AssignForces<cao>, p3m.cpp:344-373): one block
// per brick with the brick's three field tiles in LDS; F -= q prefactor/V E
// accumulated in FP64 per particle (plus the dipole term for non-metallic
// epsilon, p3m.cpp:538-547).
// ---------------------------------------------------------------------------
/** Field tiles of a brick fetched into registers with the row mapping, the
 *  first chunk of particles staged while those loads are in flight, then the
 *  tiles stored to LDS; ends with a barrier. */
template <int P, typename MT, int NTH>
__device__ __forceinline__ void fetch_tiles_and_stage(SysDesc const &d, double4 const *sp, int beg,
                                                      int cnt, int o0, int o1, int o2, MT *tx,
                                                      MT *ty, MT *tz, ChunkSlots s) {
  constexpr int T = kBrick + P - 1;
  constexpr int NP = TileRows<P, NTH>::NPASS;
  int const M0 = d.g.M[0], M1 = d.g.M[1], M2 = d.g.M[2];
  MT const *Ex = static_cast<MT const *>(d.rho);
  MT const *Ey = static_cast<MT const *>(d.Ey);
  MT const *Ez = static_cast<MT const *>(d.Ez);
  TileRows<P, NTH> const tr(o2, M2);
  MT vx[NP], vy[NP], vz[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    vx[k] = vy[k] = vz[k] = MT(0);
    if (tr.valid(k)) {
      int64_t const g = tr.mesh_index(k, o0, o1, M0, M1, M2);
      vx[k] = Ex[g];
      vy[k] = Ey[g];
      vz[k] = Ez[g];
    }
  }
  stage_chunk<P>(d.g, sp, beg, cnt, o0, o1, o2, T, T, s);
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    if (tr.valid(k)) {
      int const e = tr.row(k) * T + tr.c;
      tx[e] = vx[k];
      ty[e] = vy[k];
      tz[e] = vz[k];
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void apply_force(double *force, int64_t idx, double qv,
                                            double force_prefac, double fx, double fy, double fz,
                                            double const dm[3]) {
  double const pf = qv * force_prefac;
  double *f = force + 3 * idx;
  f[0] = (f[0] - pf * fx) - qv * dm[0];
  f[1] = (f[1] - pf * fy) - qv * dm[1];
  f[2] = (f[2] - pf * fz) - qv * dm[2];
}

/** Plane mapping: one lane segment per particle over the (j, k) plane, loop
 *  over i.  SEGMAX = 64: a segment of the next power of two >= P^2 lanes, so
 *  the LDS reads of an instruction serve one particle (no inter-particle bank
 *  conflicts), reduced with 64-lane DPP and lane reads.  SEGMAX = 16: at most
 *  16 lanes per particle, each sweeping (j, k) pairs r0, r0 + 16, ...; four
 *  particles per wave and a 4-step DPP reduction. */
template <int P, typename MT, int SEGMAX, int NTH>
__global__ __launch_bounds__(NTH) void k_gather_plane(
    SysDesc const *__restrict__ sys, int const *__restrict__ sel, int const *__restrict__ wt,
    int nsel, int const *__restrict__ soff, int const *__restrict__ offsets,
    double4 const *__restrict__ spos, int const *__restrict__ sidx, double *__restrict__ force,
    double const *__restrict__ sums, int chunk) {
  constexpr int T = kBrick + P - 1;
  constexpr int TYZ = T * T;
  constexpr int tile_n = T * TYZ;
  extern __shared__ double smem_d[];
  ChunkSlots s;
  s.w = smem_d;
  s.qq = s.w + chunk * 3 * P;
  s.lo = reinterpret_cast<int *>(s.qq + chunk);
  MT *tx = reinterpret_cast<MT *>(reinterpret_cast<char *>(smem_d) + chunk_slot_bytes(P, chunk));
  MT *ty = tx + tile_n;
  MT *tz = ty + tile_n;

  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  int const tyz = d.g.T[1] * d.g.T[2]; // == TYZ (bricks), read at run time
  int const b = L.lb;
  int const beg = offsets[d.off_off + b], end = offsets[d.off_off + b + 1];
  if (beg == end) return;
  int o0, o1, o2;
  brick_origin<P>(d, b, o0, o1, o2);
  double dm[3] = {0., 0., 0.};
  // the GPU index rule follows the reference's GPU force kernel, which has
  // no dipole term (p3m.cpp:867-888)
  if (d.dip_pref != 0. && !d.g.conv)
    for (int k = 0; k < 3; ++k) dm[k] = d.dip_pref * sums[L.s * 16 + 3 + k];
  double4 const *sp = spos + soff[L.s];
  fetch_tiles_and_stage<P, MT, NTH>(d, sp, beg, min(chunk, end - beg), o0, o1, o2, tx, ty, tz, s);
  int const *si = sidx + soff[L.s];
  double *fbase = force + 3 * d.part_off;
  constexpr int S = (Seg<P>::S < SEGMAX) ? Seg<P>::S : SEGMAX;
  constexpr int ppw = 64 / S;
  constexpr int NPAIR = P * P;
  constexpr int NT = (NPAIR + S - 1) / S;
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const nwaves = blockDim.x >> 6;
  int const sub = lane / S;
  int const r0 = lane - sub * S;
  for (int c0 = beg; c0 < end; c0 += chunk) {
    int const cnt = min(chunk, end - c0);
    if (c0 != beg) { // the first chunk was staged with the tiles
      __syncthreads();
      stage_chunk<P>(d.g, sp, c0, cnt, o0, o1, o2, T, T, s);
      __syncthreads();
    }
    int const ngroups = (cnt + ppw - 1) / ppw;
    for (int grp = wave; grp < ngroups; grp += nwaves) {
      int const p = grp * ppw + sub;
      double fx = 0., fy = 0., fz = 0.;
      if (p < cnt) {
        double const *w = s.w + p * 3 * P;
        int const lo = s.lo[p];
        double wx[P];
#pragma unroll
        for (int i = 0; i < P; ++i) wx[i] = w[i];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          int const r = r0 + t * S;
          if (r < NPAIR) {
            int const j = r / P, k = r - (r / P) * P;
            double const wyz = w[P + j] * w[2 * P + k];
            int const base = lo + j * T + k;
            // the stencil column of (j, k) summed over i, then weighted; the
            // layer pitch comes from the descriptor, so the compiler keeps
            // single 64-bit LDS reads (paired ds_read2 reads at this pitch
            // measured about 30% slower)
            double ax = 0., ay = 0., az = 0.;
#pragma unroll
            for (int i = 0; i < P; ++i) {
              ax += wx[i] * static_cast<double>(tx[base + i * tyz]);
              ay += wx[i] * static_cast<double>(ty[base + i * tyz]);
              az += wx[i] * static_cast<double>(tz[base + i * tyz]);
            }
            fx += wyz * ax;
            fy += wyz * ay;
            fz += wyz * az;
          }
        }
      }
      fx = segment_total<S>(fx);
      fy = segment_total<S>(fy);
      fz = segment_total<S>(fz);
      if (r0 == 0 && p < cnt)
        apply_force(fbase, si[c0 + p], s.qq[p], d.force_prefac, fx, fy, fz, dm);
    }
  }
}

/** Layer mapping: S = next power of two >= P lanes per particle, lane i sweeps
 *  the (j, k) plane of stencil layer i in the mesh precision; the lanes of a
 *  particle are summed in FP64 with a 3-step DPP reduction.  Several particles
 *  per instruction.  Used for FP32 fields. */
template <int P> struct GSeg {
  static constexpr int S = (P <= 1) ? 1 : (P <= 2) ? 2 : (P <= 4) ? 4 : 8;
  static constexpr int ppw = 64 / S;
};
template <int S> __device__ __forceinline__ double small_segment_total(double v) {
  if constexpr (S >= 2) v += dpp_f64<0xB1>(v);  // quad_perm [1,0,3,2]
  if constexpr (S >= 4) v += dpp_f64<0x4E>(v);  // quad_perm [2,3,0,1]
  if constexpr (S >= 8) v += dpp_f64<0x141>(v); // row_half_mirror
  return v;
}

// FP32 tiles in 512-thread blocks: capped at 64 VGPRs (8 waves per SIMD),
// so that four blocks share a CU (the LDS holds four such tiles) instead of
// two; the cap spills up to 20 bytes per lane and is still faster (C4 batch
// gather 0.91 -> 0.87 ms, single-system FP32 step 0.514 -> 0.502 ms)
template <int P, typename MT, int NTH>
__global__ __launch_bounds__(NTH, (NTH == 512 && sizeof(MT) == 4) ? 8 : 1) void k_gather_layer(
    SysDesc const *__restrict__ sys, int const *__restrict__ sel, int const *__restrict__ wt,
    int nsel, int const *__restrict__ soff, int const *__restrict__ offsets,
    double4 const *__restrict__ spos, int const *__restrict__ sidx, double *__restrict__ force,
    double const *__restrict__ sums, int chunk) {
  constexpr int T = kBrick + P - 1;
  constexpr int TYZ = T * T;
  constexpr int tile_n = T * TYZ;
  constexpr int S = GSeg<P>::S;
  constexpr int ppw = GSeg<P>::ppw;
  extern __shared__ double smem_d[];
  ChunkSlots s;
  s.w = smem_d;
  s.qq = s.w + chunk * 3 * P;
  s.lo = reinterpret_cast<int *>(s.qq + chunk);
  MT *tx = reinterpret_cast<MT *>(reinterpret_cast<char *>(smem_d) + chunk_slot_bytes(P, chunk));
  MT *ty = tx + tile_n;
  MT *tz = ty + tile_n;

  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  int const b = L.lb;
  int const beg = offsets[d.off_off + b], end = offsets[d.off_off + b + 1];
  if (beg == end) return;
  int o0, o1, o2;
  brick_origin<P>(d, b, o0, o1, o2);
  double dm[3] = {0., 0., 0.};
  // the GPU index rule follows the reference's GPU force kernel, which has
  // no dipole term (p3m.cpp:867-888)
  if (d.dip_pref != 0. && !d.g.conv)
    for (int k = 0; k < 3; ++k) dm[k] = d.dip_pref * sums[L.s * 16 + 3 + k];
  double4 const *sp = spos + soff[L.s];
  fetch_tiles_and_stage<P, MT, NTH>(d, sp, beg, min(chunk, end - beg), o0, o1, o2, tx, ty, tz, s);
  int const *si = sidx + soff[L.s];
  double *fbase = force + 3 * d.part_off;
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const nwaves = blockDim.x >> 6;
  int const sub = lane / S;
  int const li = lane - sub * S;
  bool const active = li < P;
  for (int c0 = beg; c0 < end; c0 += chunk) {
    int const cnt = min(chunk, end - c0);
    if (c0 != beg) { // the first chunk was staged with the tiles
      __syncthreads();
      stage_chunk<P>(d.g, sp, c0, cnt, o0, o1, o2, T, T, s);
      __syncthreads();
    }
    int const ngroups = (cnt + ppw - 1) / ppw;
    for (int grp = wave; grp < ngroups; grp += nwaves) {
      int const p = grp * ppw + sub;
      bool const valid = active && p < cnt;
      // per-lane stencil layer in the mesh precision MT; the cross-lane sum
      // and the force update are FP64
      MT ax = MT(0), ay = MT(0), az = MT(0);
      if (valid) {
        double const *w = s.w + p * 3 * P;
        MT wz[P];
#pragma unroll
        for (int k = 0; k < P; ++k) wz[k] = static_cast<MT>(w[2 * P + k]);
        double const wxi = w[li];
        int const base = s.lo[p] + li * TYZ;
        // each stencil row summed with the z weights, then weighted by
        // w_x w_y: three FMAs per stencil point
#pragma unroll
        for (int j = 0; j < P; ++j) {
          int const row = base + j * T;
          MT bx = wz[0] * tx[row], by = wz[0] * ty[row], bz = wz[0] * tz[row];
#pragma unroll
          for (int k = 1; k < P; ++k) {
            bx += wz[k] * tx[row + k];
            by += wz[k] * ty[row + k];
            bz += wz[k] * tz[row + k];
          }
          MT const wj = static_cast<MT>(wxi * w[P + j]);
          ax += wj * bx;
          ay += wj * by;
          az += wj * bz;
        }
      }
      double const fx = small_segment_total<S>(static_cast<double>(ax));
      double const fy = small_segment_total<S>(static_cast<double>(ay));
      double const fz = small_segment_total<S>(static_cast<double>(az));
      if (li == 0 && p < cnt)
        apply_force(fbase, si[c0 + p], s.qq[p], d.force_prefac, fx, fy, fz, dm);
    }
  }
}

// ---------------------------------------------------------------------------
// Column march (spread and gather).  Particles are binned into columns of
// W x W first stencil points in (y, z) and, within a column, into x layers.
// A block takes a run of x layers of one column and keeps the P mesh planes
// x = X .. X+P-1 that layer X's stencils touch in an LDS ring (slot X mod P),
// T x T points per plane (T = W + P - 1, the column plus the stencil reach).
// Ring layout of a field: point (jy, kz), slot q at jy IP + kz P + q with
// IP >= T P and IP = P^2 (mod 64): the P^2 values (kz0..kz0+P-1, every slot)
// of one stencil row jy are consecutive, and row jy + 1 continues the bank
// sequence, so the lanes of a particle read or add consecutive LDS words.
// Moving to the next layer brings in one plane (gather) or flushes one
// completed plane into the mesh (spread) instead of a whole brick tile.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int ring_pitch(int T, int P) {
  return T * P + (((P * P) - T * P) & 63);
}

template <int P, int W> struct ColGeo {
  static constexpr int T = W + P - 1;
  static constexpr int IP = ring_pitch(T, P);
  static constexpr int RING = T * IP; // elements of one field's ring
  static constexpr int RL = (T <= 16) ? 16 : 32; // lanes per tile row in plane copies
  static constexpr int RPP = 256 / RL;           // rows per pass of a 256-thread block
  static constexpr int NPASS = (T + RPP - 1) / RPP;
  static_assert(T <= RL, "tile row wider than a plane-copy row");
};

/** Column block: column c of the system, x layers [x0, x1), tile origin
 *  (oy, oz) in y and z. */
struct ColBlock {
  int c, x0, x1, oy, oz;
};
__device__ __forceinline__ ColBlock col_block(SysDesc const &d, int lb) {
  ColBlock b;
  b.c = lb / d.nseg;
  int const sg = lb - b.c * d.nseg;
  b.x0 = sg * d.lz;
  b.x1 = min(b.x0 + d.lz, d.g.M[0]);
  int const c1 = b.c / d.g.nb[2], c2 = b.c - c1 * d.g.nb[2];
  b.oy = c1 * d.g.col;
  b.oz = c2 * d.g.col;
  return b;
}

/** Weights of a chunk of one layer's particles, one thread per (particle,
 *  dimension), as stage_chunk; the particle's ring offset is
 *  (y - oy) IP + (z - oz) P. */
template <int P, int IP>
__device__ __forceinline__ void stage_layer(DevGeo const &g, double4 const *spos, int64_t c0,
                                            int cnt, int oy, int oz, ChunkSlots s) {
  int const t = threadIdx.x;
  if (t < 3 * cnt) {
    int const p = t / 3;
    int const dd = t - 3 * p;
    double4 const pp = spos[c0 + p];
    double const x[3] = {pp.x, pp.y, pp.z};
    int nmp, gi;
    double dist;
    coordinate_1d(g, P, x[dd], dd, nmp, gi, dist);
    bspline_weights<P>(dist, s.w + p * 3 * P + dd * P);
    if (dd == 0) {
      int n1, g1, n2, g2;
      double d1, d2;
      coordinate_1d(g, P, x[1], 1, n1, g1, d1);
      coordinate_1d(g, P, x[2], 2, n2, g2, d2);
      s.qq[p] = pp.w;
      s.lo[p] = (g1 - oy) * IP + (g2 - oz) * P;
    }
  }
}

/** Row mapping of one ring plane over a 256-thread block: lane c of a row
 *  covers kz = c (active for c < T), row jy = threadIdx.x / RL + k RPP. */
template <int P, int W> struct PlaneRows {
  using G = ColGeo<P, W>;
  int kz, row0, zz;
  bool lane_ok;
  __device__ PlaneRows(int oz, int M2) {
    kz = threadIdx.x % G::RL;
    row0 = threadIdx.x / G::RL;
    lane_ok = kz < G::T;
    zz = wrap_once(oz + kz, M2);
  }
  __device__ bool valid(int k) const { return lane_ok && row0 + k * G::RPP < G::T; }
  __device__ int row(int k) const { return row0 + k * G::RPP; }
  __device__ int64_t mesh_index(int k, int x, int oy, int M1, int M2) const {
    return ((int64_t)x * M1 + wrap_once(oy + row(k), M1)) * M2 + zz;
  }
  __device__ int lds_index(int k, int q) const { return row(k) * G::IP + kz * P + q; }
};

template <int P, typename MT, int W, int S>
__global__ __launch_bounds__(256) void k_gather_col(
    SysDesc const *__restrict__ sys, int const *__restrict__ sel, int const *__restrict__ wt,
    int nsel, int const *__restrict__ soff, int const *__restrict__ offsets,
    double4 const *__restrict__ spos, int const *__restrict__ sidx, double *__restrict__ force,
    double const *__restrict__ sums, int chunk) {
  using G = ColGeo<P, W>;
  constexpr int IP = G::IP;
  constexpr int PPW = 64 / S;                // particles per wave
  constexpr int NM = (P * P + S - 1) / S;    // stencil-row elements per lane
  extern __shared__ double smem_d[];
  ChunkSlots s;
  s.w = smem_d;
  s.qq = s.w + chunk * 3 * P;
  s.lo = reinterpret_cast<int *>(s.qq + chunk);
  MT *rx = reinterpret_cast<MT *>(reinterpret_cast<char *>(smem_d) + chunk_slot_bytes(P, chunk));
  MT *ry = rx + G::RING;
  MT *rz = ry + G::RING;

  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  ColBlock const cb = col_block(d, L.lb);
  int const M0 = d.g.M[0], M1 = d.g.M[1], M2 = d.g.M[2];
  int const *off = offsets + d.off_off + cb.c * M0;
  if (off[cb.x0] == off[cb.x1]) return; // no particles in these layers
  double dm[3] = {0., 0., 0.};
  // the GPU index rule follows the reference's GPU force kernel, which has
  // no dipole term (p3m.cpp:867-888)
  if (d.dip_pref != 0. && !d.g.conv)
    for (int k = 0; k < 3; ++k) dm[k] = d.dip_pref * sums[L.s * 16 + 3 + k];
  double4 const *sp = spos + soff[L.s];
  int const *si = sidx + soff[L.s];
  double *fbase = force + 3 * d.part_off;
  MT const *Ex = static_cast<MT const *>(d.rho);
  MT const *Ey = static_cast<MT const *>(d.Ey);
  MT const *Ez = static_cast<MT const *>(d.Ez);
  PlaneRows<P, W> const pr(cb.oz, M2);
  // prologue: planes x0 .. x0 + P - 2 straight into their slots, plane
  // x0 + P - 1 into registers
  for (int X = cb.x0; X < cb.x0 + P - 1; ++X) {
    int const xm = wrap_once(X, M0), q = X % P;
#pragma unroll
    for (int k = 0; k < G::NPASS; ++k)
      if (pr.valid(k)) {
        int64_t const gidx = pr.mesh_index(k, xm, cb.oy, M1, M2);
        int const e = pr.lds_index(k, q);
        rx[e] = Ex[gidx];
        ry[e] = Ey[gidx];
        rz[e] = Ez[gidx];
      }
  }
  MT vx[G::NPASS], vy[G::NPASS], vz[G::NPASS];
  auto fetch = [&](int X) {
    int const xm = wrap_once(X, M0);
#pragma unroll
    for (int k = 0; k < G::NPASS; ++k) {
      vx[k] = vy[k] = vz[k] = MT(0);
      if (pr.valid(k)) {
        int64_t const gidx = pr.mesh_index(k, xm, cb.oy, M1, M2);
        vx[k] = Ex[gidx];
        vy[k] = Ey[gidx];
        vz[k] = Ez[gidx];
      }
    }
  };
  fetch(cb.x0 + P - 1);
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const nwaves = blockDim.x >> 6;
  int const sub = lane / S;
  int const hl = lane - sub * S;
  for (int X = cb.x0; X < cb.x1; ++X) {
    {
      // plane X + P - 1 into the slot of plane X - 1 (last read by an
      // earlier layer, before that layer's closing barrier)
      int const q = (X + P - 1) % P;
#pragma unroll
      for (int k = 0; k < G::NPASS; ++k)
        if (pr.valid(k)) {
          int const e = pr.lds_index(k, q);
          rx[e] = vx[k];
          ry[e] = vy[k];
          rz[e] = vz[k];
        }
    }
    if (X + 1 < cb.x1) fetch(X + P); // in flight during this layer
    int const beg = off[X], end = off[X + 1];
    if (beg == end) continue;
    int const lq = X % P; // slot of plane X: weight index of slot q is (q - lq) mod P
    for (int c0 = beg; c0 < end; c0 += chunk) {
      int const cnt = min(chunk, end - c0);
      if (c0 != beg) __syncthreads(); // the previous chunk's slots are read
      stage_layer<P, IP>(d.g, sp, c0, cnt, cb.oy, cb.oz, s);
      __syncthreads();
      int const ngroups = (cnt + PPW - 1) / PPW;
      for (int grp = wave; grp < ngroups; grp += nwaves) {
        int const p = grp * PPW + sub;
        bool const valid = p < cnt;
        double fx = 0., fy = 0., fz = 0.;
        if (valid) {
          double const *w = s.w + p * 3 * P;
          double wy[P];
#pragma unroll
          for (int r = 0; r < P; ++r) wy[r] = w[P + r];
          int const lo = s.lo[p];
          // one stencil-row element at a time: P loads per field in flight
          // per wave, and registers for more waves
#pragma unroll 1
          for (int mi = 0; mi < NM; ++mi) {
            int const m = hl + mi * S;
            if (m < P * P) {
              int const kk = m / P, q = m - (m / P) * P;
              int k = q - lq;
              if (k < 0) k += P;
              double const wzx = w[2 * P + kk] * w[k];
              int const base = lo + m;
              MT ax = MT(0), ay = MT(0), az = MT(0);
#pragma unroll
              for (int r = 0; r < P; ++r) {
                MT const wr = static_cast<MT>(wy[r]);
                ax += wr * rx[base + r * IP];
                ay += wr * ry[base + r * IP];
                az += wr * rz[base + r * IP];
              }
              fx += wzx * static_cast<double>(ax);
              fy += wzx * static_cast<double>(ay);
              fz += wzx * static_cast<double>(az);
            }
          }
        }
        fx = segment_total<S>(fx);
        fy = segment_total<S>(fy);
        fz = segment_total<S>(fz);
        if (hl == 0 && valid)
          apply_force(fbase, si[c0 + p], s.qq[p], d.force_prefac, fx, fy, fz, dm);
      }
    }
    __syncthreads(); // this layer's reads of the ring and of the slots are done
  }
}

template <int P, typename MT, int W, int S>
__global__ __launch_bounds__(256) void k_spread_col(SysDesc const *__restrict__ sys,
                                                   int const *__restrict__ sel,
                                                   int const *__restrict__ wt, int nsel,
                                                   int const *__restrict__ soff,
                                                   int const *__restrict__ offsets,
                                                   double4 const *__restrict__ spos, int chunk) {
  using G = ColGeo<P, W>;
  constexpr int IP = G::IP;
  constexpr int PPW = 64 / S;
  constexpr int NM = (P * P + S - 1) / S;
  extern __shared__ double smem_d[];
  ChunkSlots s;
  s.w = smem_d;
  s.qq = s.w + chunk * 3 * P;
  s.lo = reinterpret_cast<int *>(s.qq + chunk);
  double *ring = reinterpret_cast<double *>(reinterpret_cast<char *>(smem_d) +
                                            chunk_slot_bytes(P, chunk));

  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  ColBlock const cb = col_block(d, L.lb);
  int const M0 = d.g.M[0], M1 = d.g.M[1], M2 = d.g.M[2];
  int const *off = offsets + d.off_off + cb.c * M0;
  if (off[cb.x0] == off[cb.x1]) return;
  double4 const *sp = spos + soff[L.s];
  MT *rho = static_cast<MT *>(d.rho);
  PlaneRows<P, W> const pr(cb.oz, M2);
  for (int e = threadIdx.x; e < G::RING; e += blockDim.x) ring[e] = 0.;
  __syncthreads(); // flushes of leading empty layers read other threads' zeros
  // flush the completed plane X (slot X mod P) into the mesh and clear it
  auto flush = [&](int X) {
    int const xm = wrap_once(X, M0), q = X % P;
#pragma unroll
    for (int k = 0; k < G::NPASS; ++k)
      if (pr.valid(k)) {
        int const e = pr.lds_index(k, q);
        double const v = ring[e];
        if (v != 0.) {
          unsafeAtomicAdd(rho + pr.mesh_index(k, xm, cb.oy, M1, M2), static_cast<MT>(v));
          ring[e] = 0.;
        }
      }
  };
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const nwaves = blockDim.x >> 6;
  int const sub = lane / S;
  int const hl = lane - sub * S;
  for (int X = cb.x0; X < cb.x1; ++X) {
    int const beg = off[X], end = off[X + 1];
    if (beg != end) {
      int const lq = X % P;
      for (int c0 = beg; c0 < end; c0 += chunk) {
        int const cnt = min(chunk, end - c0);
        if (c0 != beg) __syncthreads();
        // the barrier after staging also orders the clearing of the slot
        // flushed last (plane X - 1, now the slot of plane X + P - 1) before
        // the additions
        stage_layer<P, IP>(d.g, sp, c0, cnt, cb.oy, cb.oz, s);
        __syncthreads();
        int const ngroups = (cnt + PPW - 1) / PPW;
        for (int grp = wave; grp < ngroups; grp += nwaves) {
          int const p = grp * PPW + sub;
          if (p < cnt) {
            double const *w = s.w + p * 3 * P;
            double const qp = s.qq[p];
            double wy[P];
#pragma unroll
            for (int r = 0; r < P; ++r) wy[r] = w[P + r];
            int const lo = s.lo[p];
#pragma unroll
            for (int mi = 0; mi < NM; ++mi) {
              int const m = hl + mi * S;
              if (m < P * P) {
                int const kk = m / P, q = m - (m / P) * P;
                int k = q - lq;
                if (k < 0) k += P;
                double const wzxq = (w[2 * P + kk] * w[k]) * qp;
                double *dst = ring + lo + m;
#pragma unroll
                for (int r = 0; r < P; ++r) atomicAdd(dst + r * IP, wy[r] * wzxq);
              }
            }
          }
        }
      }
      __syncthreads(); // every addition of layer X is in the ring
    }
    flush(X); // plane X receives nothing from later layers of this block
  }
  // planes x1 .. x1 + P - 2 (the last additions are behind the last barrier)
  for (int X = cb.x1; X < cb.x1 + P - 1; ++X) flush(X);
}

// ---------------------------------------------------------------------------
// request preparation for served solves: positions of the selected systems
// get a Gaussian displacement (x += sigma N(0,1), folded into the box) and
// their forces are cleared.  Normal draws are Box-Muller over a counter-based
// splitmix64 hash of (request seed, particle, component), so a request is
// reproducible whatever else runs with it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double unit_open(unsigned long long h) {
  return (static_cast<double>(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); // (0, 1)
}

__global__ __launch_bounds__(kBlock) void k_prepare(SysDesc const *__restrict__ sys,
                                                    int const *__restrict__ sel,
                                                    int const *__restrict__ wt, int nsel,
                                                    double *pos, double *force, double sigma,
                                                    unsigned long long const *seeds) {
  Slot const L = locate(wt, sel, nsel);
  SysDesc const &d = sys[L.s];
  int64_t const li = (int64_t)L.lb * blockDim.x + threadIdx.x;
  if (li >= d.n) return;
  int64_t const i = d.part_off + li;
  if (sigma > 0.) {
    unsigned long long const seed = seeds[L.slot];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      unsigned long long const key = seed ^ (static_cast<unsigned long long>(3 * li + c) *
                                             0xD1B54A32D192ED03ull);
      double const u1 = unit_open(splitmix64(key));
      double const u2 = unit_open(splitmix64(key ^ 0x8CB92BA72F3D8DD7ull));
      double const z = sqrt(-2. * log(u1)) * cospi(2. * u2);
      double const L0 = d.g.box_l[c];
      double x = pos[3 * i + c] + sigma * z;
      x -= L0 * floor(x / L0);
      if (x >= L0) x -= L0;
      pos[3 * i + c] = x;
    }
  }
  force[3 * i] = 0.;
  force[3 * i + 1] = 0.;
  force[3 * i + 2] = 0.;
}

// ---------------------------------------------------------------------------
// real-space pair forces (CoulombP3M::pair_force, p3m.hpp:146-165, with the
// Abramowitz-Stegun erfc of AS_erfc_part.hpp:29-40) on listed target
// particles from all particles within r_cut, minimum image; one block per
// target scans all sources.  Used to complete the total force of selected
// particles (error measurements); O(targets x N).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_pair_forces_targets(
    int64_t n, double const *__restrict__ pos, double const *__restrict__ q,
    int const *__restrict__ targets, double Lx, double Ly, double Lz, double prefactor,
    double alpha, double r_cut, double *__restrict__ out) {
  __shared__ double red[3 * 16];
  int const t = targets[blockIdx.x];
  double const qi = q[t];
  double const xi = pos[3 * (int64_t)t], yi = pos[3 * (int64_t)t + 1], zi = pos[3 * (int64_t)t + 2];
  double const rc2 = r_cut * r_cut;
  double const two_a_sqrt_pi_i = 2. * alpha * (1. / sqrt(std::numbers::pi));
  double v[3] = {0., 0., 0.};
  if (qi != 0.) {
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
      double const qj = q[j];
      if (j == t || qj == 0.) continue;
      double dx = xi - pos[3 * j], dy = yi - pos[3 * j + 1], dz = zi - pos[3 * j + 2];
      dx -= Lx * rint(dx / Lx);
      dy -= Ly * rint(dy / Ly);
      dz -= Lz * rint(dz / Lz);
      double const r2 = dx * dx + dy * dy + dz * dz;
      if (r2 >= rc2 || r2 == 0.) continue;
      double const dist = sqrt(r2);
      double const adist = alpha * dist;
      double const tt = 1. / (1. + 0.3275911 * adist);
      double const erfc_part = tt * (0.254829592 + tt * (-0.284496736 + tt * (1.421413741 +
                                     tt * (-1.453152027 + tt * 1.061405429))));
      double const ex = exp(-adist * adist);
      double const fac = ex * (erfc_part / dist + two_a_sqrt_pi_i) / r2 * prefactor * qi * qj;
      v[0] += fac * dx;
      v[1] += fac * dy;
      v[2] += fac * dz;
    }
  }
  block_sum<3>(v, red);
  if (threadIdx.x == 0) {
    out[3 * (int64_t)blockIdx.x] = v[0];
    out[3 * (int64_t)blockIdx.x + 1] = v[1];
    out[3 * (int64_t)blockIdx.x + 2] = v[2];
  }
}

template <typename F> void dispatch_cao(int cao, F &&f) {
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

/** Block sizes of the particle-mesh kernels (LaunchShape::threads). */
template <typename F> void with_threads(int threads, F &&f) {
  if (threads == 256)
    f(std::integral_constant<int, 256>{});
  else if (threads == 512)
    f(std::integral_constant<int, 512>{});
  else
    throw std::runtime_error("particle-mesh block size must be 256 or 512");
}
template <typename F> void for_threads(F &&f) {
  f(std::integral_constant<int, 256>{});
  f(std::integral_constant<int, 512>{});
}
/** Column width and lanes per particle of the column-march kernels. */
template <typename F> void with_column(int w, int lanes, F &&f) {
  auto by_lanes = [&](auto W) {
    if (lanes == 8)
      f(W, std::integral_constant<int, 8>{});
    else if (lanes == 16)
      f(W, std::integral_constant<int, 16>{});
    else
      throw std::runtime_error("column march: lanes per particle must be 8 or 16");
  };
  if (w == 8)
    by_lanes(std::integral_constant<int, 8>{});
  else if (w == 16)
    by_lanes(std::integral_constant<int, 16>{});
  else
    throw std::runtime_error("column march: column width must be 8 or 16");
}
template <typename F> void for_columns(F &&f) {
  for (int w : {8, 16})
    for (int l : {8, 16}) with_column(w, l, f);
}


template <typename T> void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

constexpr int kLdsBytes = 160 * 1024; // LDS per workgroup on gfx950

/** Raise the dynamic LDS limit of a kernel to what its static LDS leaves. */
template <typename Kern> void allow_lds(Kern k) {
  hipFuncAttributes attr{};
  hip_check(hipFuncGetAttributes(&attr, reinterpret_cast<void const *>(k)),
            "hipFuncGetAttributes");
  int const dyn = kLdsBytes - static_cast<int>(attr.sharedSizeBytes);
  hip_check(hipFuncSetAttribute(reinterpret_cast<void const *>(k),
                                hipFuncAttributeMaxDynamicSharedMemorySize, dyn),
            "hipFuncSetAttribute");
}

size_t chunk_bytes(int cao, int chunk) { return chunk_slot_bytes(cao, chunk); }
/** Ring elements of one field of a column block. */
int ring_points(int cao, int w) { return (w + cao - 1) * ring_pitch(w + cao - 1, cao); }
int tile_points(int cao, int group = 1) {
  int const t = kBrick + cao - 1, t2 = 2 * kBrick + cao - 1;
  return group == 4 ? t2 * t2 * t : t * t * t;
}

/** LDS of the fused y/x pass: both twiddle tables and two x-y planes. */
size_t fused_lds_bytes(int mx, int my, size_t csz) {
  size_t const plane = static_cast<size_t>(std::max(mx * (my + 1), my * (mx + 1)));
  return (static_cast<size_t>(mx + my) + 2 * plane) * csz;
}
bool uses_fused_pass(int mx, int my, MeshPrecision prec, LaunchShape const &shape) {
  size_t const csz = (prec == MeshPrecision::FP64) ? 16 : 8;
  return fused_lds_bytes(mx, my, csz) <= static_cast<size_t>(shape.fused_lds);
}

/** lines per block such that twiddles + lines_per_unit * units lines fit */
int pick_lines(int n, size_t csz, int lines_per_unit, int want, size_t budget) {
  size_t const line = static_cast<size_t>(n + 1) * csz;
  size_t const twb = static_cast<size_t>(n) * csz;
  int u = want;
  while (u > 1 && twb + static_cast<size_t>(lines_per_unit) * u * line > budget) --u;
  return u;
}
} // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct HostSys {
  P3MParams params;
  Vec3d box_l{};
  double skin = 0.;
  double prefactor = 0.;
  LocalMesh lm;
  int64_t n = 0;
  int64_t part_off = 0;
  int brick_off = 0;
  int nbricks = 0;
  size_t bytes = 0;
  SysDesc desc{};
};

/** Work tables of one selection of systems. */
struct Plan {
  std::vector<int> sel;
  int nsel = 0;
  int *dbuf = nullptr; // device copy of all tables below (owned by the lane)
  int *d_sel = nullptr, *wt_bin = nullptr, *wt_fill = nullptr, *wt_z = nullptr, *wt_zc = nullptr,
      *wt_y = nullptr, *wt_x = nullptr, *wt_yi = nullptr, *wt_f = nullptr;
  int blocks_bin = 0, blocks_fill = 0, blocks_z = 0, blocks_zc = 0, blocks_y = 0, blocks_x = 0,
      blocks_yi = 0, blocks_f = 0;
  int *soff = nullptr; // per system id: first particle of the system in the lane scratch
  int64_t npart = 0;   // particles of the selection
  struct Group {
    int cao;
    int nsel;
    int *d_sel, *wt;
    int blocks;
    int *wt_sp;    // spread work table (brick groups of LaunchShape::group)
    int blocks_sp;
  };
  std::vector<Group> groups;
  int nlp_z = 1, nlp_zc = 1, c2r_comps = 3, ncol_y = 1, ncol_x = 1, ncol_yi = 1;
  size_t lds_z = 0, lds_zc = 0, lds_y = 0, lds_x = 0, lds_yi = 0, lds_f = 0;
  bool any_dipole = false;
  // graph
  hipGraphExec_t graph = nullptr;
  ParticleArrays graph_pa{};
  unsigned graph_flags = 0;
  int64_t version = -1;
};

/** Working state of one stream: selection plan, particle scratch sized to
 *  the selection, partial sums and events.  Systems (meshes, G,
 *  descriptors) are shared by all lanes; a system is never in two lanes'
 *  in-flight solves at once (the caller's contract). */
struct Lane {
  hipStream_t stream = nullptr;
  int *counts = nullptr, *offsets = nullptr;
  int counts_cap = 0, offsets_cap = 0;
  // particle scratch of the current solve: carved from the lane's own block
  // or from memory the caller passes in
  int *p_brick = nullptr, *p_rank = nullptr, *sidx = nullptr;
  double4 *spos = nullptr;
  void *own = nullptr;
  size_t own_bytes = 0;
  double *bin_part = nullptr, *conv_part = nullptr;
  int bin_cap = 0, conv_cap = 0;
  double *sums = nullptr;
  int sums_cap = 0;
  int *err_flag = nullptr;
  double *h_sums = nullptr;
  unsigned long long *seeds = nullptr, *h_seeds = nullptr;
  int seeds_cap = 0;
  int *tables = nullptr, *h_tables = nullptr; // plan tables (device, pinned staging)
  size_t tables_cap = 0;
  hipEvent_t ev[8] = {};
  std::vector<hipEvent_t> ring;
  int ring_used = 0;
  bool profiling = false;
  Plan plan;
  // the fixed lane of allocation-free solves: lane 0's stream, every device
  // buffer carved from the fixed scratch array, sums read back into h_fixed
  bool fixed = false;
  std::vector<double> h_fixed;
};

struct BatchSolver::Impl {
  std::vector<HostSys> sys;
  SysDesc *d_sys = nullptr;
  size_t d_sys_cap = 0;
  bool descs_dirty = true;
  int64_t version = 0; // bumped when systems or particle counts change
  // twiddle tables by (size, precision)
  std::map<int, int> tw_offset; // size -> offset (complex elements)
  void *d_tw = nullptr;
  size_t tw_len = 0;
  std::vector<std::unique_ptr<Lane>> lanes;
  bool kernels_ready = false;
  std::unique_ptr<Lane> fixed; // see enable_fixed()
  char *fixed_base = nullptr;  // this device's address of the fixed scratch array
  size_t fixed_nsys = 0;       // systems when the fixed lane was set up
};

namespace {
void free_plan(Plan &p) {
  if (p.graph) (void)hipGraphExecDestroy(p.graph);
  p = Plan{};
}

void create_lane(Lane &ln) {
  hip_check(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking), "hipStreamCreate");
  for (auto &e : ln.ev) hip_check(hipEventCreate(&e), "hipEventCreate");
  hip_check(hipMalloc(&ln.err_flag, sizeof(int)), "hipMalloc err");
}

void destroy_lane(Lane &ln) {
  if (ln.stream) (void)hipStreamSynchronize(ln.stream);
  free_plan(ln.plan);
  dfree(ln.counts);
  dfree(ln.offsets);
  if (ln.own) (void)hipFree(ln.own);
  ln.own = nullptr;
  dfree(ln.bin_part);
  dfree(ln.conv_part);
  dfree(ln.sums);
  dfree(ln.err_flag);
  dfree(ln.seeds);
  dfree(ln.tables);
  if (ln.h_tables) (void)hipHostFree(ln.h_tables);
  if (ln.h_sums) (void)hipHostFree(ln.h_sums);
  if (ln.h_seeds) (void)hipHostFree(ln.h_seeds);
  for (auto &e : ln.ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : ln.ring) (void)hipEventDestroy(e);
  if (ln.stream) (void)hipStreamDestroy(ln.stream);
  ln.stream = nullptr;
}

/** Before shared state changes: wait for every lane and drop their plans. */
void quiesce(BatchSolver::Impl *m) {
  for (auto &ln : m->lanes) {
    hip_check(hipStreamSynchronize(ln->stream), "sync lane");
    free_plan(ln->plan);
  }
  if (m->fixed) free_plan(m->fixed->plan); // on lane 0's stream, drained above
}
} // namespace

BatchSolver::BatchSolver(int device, MeshPrecision precision)
    : m(new Impl), m_device(device), m_prec(precision),
      m_det(precision == MeshPrecision::FP64) {
  hip_check(hipSetDevice(device), "hipSetDevice");
  m->lanes.push_back(std::make_unique<Lane>());
  create_lane(*m->lanes[0]);
  m_stream = m->lanes[0]->stream;
  m_shape.gather = (precision == MeshPrecision::FP64) ? kGatherPlane16 : kGatherLayer;
  m_shape.ncol = 8;
}

BatchSolver::~BatchSolver() {
  (void)hipSetDevice(m_device);
  if (m->fixed) { // borrowed stream and carved buffers: only the events are its own
    free_plan(m->fixed->plan);
    for (auto &e : m->fixed->ev)
      if (e) (void)hipEventDestroy(e);
  }
  for (auto &ln : m->lanes) destroy_lane(*ln);
  for (auto &h : m->sys) {
    void *ptrs[] = {h.desc.rho, h.desc.Ey, h.desc.Ez, h.desc.khat, h.desc.A, h.desc.B, h.desc.G};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
  }
  dfree(m->d_sys);
  dfree(m->d_tw);
  delete m;
}

int BatchSolver::num_systems() const { return static_cast<int>(m->sys.size()); }
int64_t BatchSolver::particle_offset(int s) const { return m->sys.at(s).part_off; }
int64_t BatchSolver::particle_count(int s) const { return m->sys.at(s).n; }
int64_t BatchSolver::total_particles() const {
  return m->sys.empty() ? 0 : m->sys.back().part_off + m->sys.back().n;
}
size_t BatchSolver::system_bytes(int s) const { return m->sys.at(s).bytes; }
LocalMesh const &BatchSolver::local_mesh(int s) const { return m->sys.at(s).lm; }

int BatchSolver::add_lane() {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  auto ln = std::make_unique<Lane>();
  create_lane(*ln);
  m->lanes.push_back(std::move(ln));
  return static_cast<int>(m->lanes.size()) - 1;
}
int BatchSolver::num_lanes() const { return static_cast<int>(m->lanes.size()); }
hipStream_t BatchSolver::lane_stream(int lane) const { return m->lanes.at(lane)->stream; }

namespace {
size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }
/** Particle scratch of a solve over n particles: sorted (x, y, z, q) as
 *  double4, brick id, rank in brick and original index. */
size_t scratch_bytes(int64_t n) {
  size_t const u = static_cast<size_t>(std::max<int64_t>(n, 1));
  return align256(u * sizeof(double4)) + 3 * align256(u * sizeof(int));
}
void carve(Lane &ln, void *base, int64_t n) {
  size_t const u = static_cast<size_t>(std::max<int64_t>(n, 1));
  char *p = static_cast<char *>(base);
  ln.spos = reinterpret_cast<double4 *>(p);
  p += align256(u * sizeof(double4));
  ln.p_brick = reinterpret_cast<int *>(p);
  p += align256(u * sizeof(int));
  ln.p_rank = reinterpret_cast<int *>(p);
  p += align256(u * sizeof(int));
  ln.sidx = reinterpret_cast<int *>(p);
}
} // namespace

size_t BatchSolver::working_bytes(std::vector<int> const &sel) const {
  int64_t n = 0;
  for (int s : sel) n += m->sys.at(s).n;
  return scratch_bytes(n);
}

namespace {
void compute_influence(HostSys const &h, MeshPrecision prec, hipStream_t st);
void set_bins(HostSys &h, LaunchShape const &shape);
void recompute_offsets(BatchSolver::Impl *m);
} // namespace

void BatchSolver::set_launch_shape(LaunchShape const &s) {
  // validate before anything changes
  if (s.gather < kGatherLayer || s.gather > kGatherPlane8)
    throw std::invalid_argument("P3M: unknown gather mapping");
  if (s.spread < kSpreadSeg64 || s.spread > kSpreadSeg8)
    throw std::invalid_argument("P3M: unknown spread mapping");
  if ((s.threads != 256 && s.threads != 512) || s.chunk < 1 || s.chunk > kChunkMax ||
      3 * s.chunk > s.threads ||
      (s.gather_threads != 0 && s.gather_threads != 256 && s.gather_threads != 512) ||
      3 * s.chunk > s.gather_block())
    throw std::invalid_argument("P3M: invalid particle-mesh block shape");
  if ((s.column != 0 && s.column != 8 && s.column != 16) || s.seg < 1 || s.seg > 1024 ||
      (s.col_lanes != 8 && s.col_lanes != 16))
    throw std::invalid_argument("P3M: invalid column-march shape");
  if (m_det && s.column)
    throw std::invalid_argument("P3M: deterministic solves need 8^3 bricks (column = 0)");
  if ((s.group != 1 && s.group != 4) ||
      (s.group == 4 && (s.column != 0 || s.spread != kSpreadSeg64)))
    throw std::invalid_argument(
        "P3M: spread brick groups are 1 or 4, the latter with 8^3 bricks and the plane spread "
        "mapping");
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  bool const rebin = s.column != m_shape.column || s.seg != m_shape.seg;
  bool const regroup = s.group != m_shape.group;
  m_shape = s;
  m_shape.ncol = std::max(1, std::min(64, s.ncol));
  m_shape.zlines = std::max(1, std::min(64, s.zlines));
  m_shape.fused_lds = std::max(0, std::min(128 * 1024, s.fused_lds));
  quiesce(m);
  if (rebin) { // particle bins and block counts of every system
    for (auto &h : m->sys) set_bins(h, m_shape);
    recompute_offsets(m);
  } else if (regroup) {
    m->version++; // spread work tables change
  }
  // systems whose y/x pass changes switch spectrum layout: recompute G
  bool changed = false;
  for (auto &h : m->sys) {
    int const kzm = uses_fused_pass(h.desc.g.M[0], h.desc.g.M[1], m_prec, m_shape) ? 1 : 0;
    if (kzm != h.desc.kzmajor) {
      h.desc.kzmajor = kzm;
      compute_influence(h, m_prec, m_stream);
      changed = true;
    }
  }
  if (changed) {
    hip_check(hipStreamSynchronize(m_stream), "sync influence");
    m->descs_dirty = true;
  }
}

void BatchSolver::set_deterministic(bool on) {
  if (on && m_prec != MeshPrecision::FP64)
    throw std::invalid_argument("P3M: deterministic solves need the FP64 mesh");
  if (on && m_shape.column)
    throw std::invalid_argument("P3M: deterministic solves need 8^3 bricks (column = 0)");
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  quiesce(m); // captured graphs hold the previous spread
  m_det = on;
}

AutotuneResult BatchSolver::autotune(std::vector<int> const &sel, ParticleArrays const &pa,
                                     int reps) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  reps = std::max(reps, 1);
  hipEvent_t t0, t1;
  hip_check(hipEventCreate(&t0), "hipEventCreate");
  hip_check(hipEventCreate(&t1), "hipEventCreate");
  auto time_shape = [&](LaunchShape const &shape) {
    set_launch_shape(shape);
    solve(sel, pa, 1u, false, nullptr); // builds the plan, warms caches
    hip_check(hipEventRecord(t0, m_stream), "hipEventRecord");
    for (int r = 0; r < reps; ++r) solve(sel, pa, 1u, false, nullptr);
    hip_check(hipEventRecord(t1, m_stream), "hipEventRecord");
    hip_check(hipEventSynchronize(t1), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
    return static_cast<double>(ms) / reps;
  };
  AutotuneResult res;
  LaunchShape const start = m_shape;
  res.default_ms = time_shape(start);
  res.best = start;
  res.best_ms = res.default_ms;
  res.candidates = 1;
  auto consider = [&](LaunchShape const &c) {
    double const ms = time_shape(c);
    ++res.candidates;
    if (ms < res.best_ms) {
      res.best_ms = ms;
      res.best = c;
    }
  };
  // coordinate descent: fused y/x pass, particle-mesh blocks, gather and
  // spread mappings, blocks again, FFT columns, z lines
  {
    LaunchShape c = res.best;
    c.fused_lds = (start.fused_lds > 0) ? 0 : LaunchShape{}.fused_lds;
    consider(c);
  }
  // particle-mesh block size and chunk: large blocks for dense bricks, small
  // ones (more bricks in flight per CU) for sparse bricks
  // (the gather keeps its block size, so that a move of the spread's is
  // judged on its own; gather_threads is a coordinate of its own below)
  auto try_blocks = [&] {
    LaunchShape const base = res.best;
    for (auto [t, c] : {std::pair{512, 64}, std::pair{256, 64}, std::pair{256, 32}}) {
      if (t == base.threads && c == base.chunk) continue;
      LaunchShape cand = base;
      cand.threads = t;
      cand.chunk = c;
      cand.gather_threads = base.gather_block();
      consider(cand);
    }
  };
  try_blocks();
  for (int v : {kGatherLayer, kGatherPlane, kGatherPlane16, kGatherPlane8}) {
    if (v == res.best.gather) continue;
    LaunchShape c = res.best;
    c.gather = v;
    consider(c);
  }
  // spread mappings, and brick groups sharing one tile (sparse bricks; the
  // 64-lane mapping only) as candidates of the same coordinate
  for (auto [v, gr] : {std::pair{kSpreadSeg64, 1}, std::pair{kSpreadSeg64, 4},
                       std::pair{kSpreadSeg16, 1}, std::pair{kSpreadSeg8, 1}}) {
    if (v == res.best.spread && gr == res.best.group) continue;
    if (gr != 1 && res.best.column != 0) continue;
    LaunchShape c = res.best;
    c.spread = v;
    c.group = gr;
    consider(c);
  }
  try_blocks();
  // gather block size apart from the spread's (brick gather only)
  if (res.best.column == 0)
    for (int v : {0, 256, 512}) {
      if (v == res.best.gather_threads || (v != 0 && 3 * res.best.chunk > v)) continue;
      LaunchShape c = res.best;
      c.gather_threads = v;
      consider(c);
    }
  // column march: width and lanes per particle, then x layers per block
  if (!m_det) {
    LaunchShape const base = res.best;
    for (int w : {8, 16})
      for (int l : {8, 16}) {
        if (w == base.column && l == base.col_lanes) continue;
        LaunchShape c = base;
        c.column = w;
        c.col_lanes = l;
        c.group = 1;
        consider(c);
      }
    if (res.best.column)
      for (int sg : {16, 64}) {
        LaunchShape c = res.best;
        c.seg = sg;
        consider(c);
      }
  }
  for (int v : {2, 4, 8, 16}) {
    if (v == res.best.ncol) continue;
    LaunchShape c = res.best;
    c.ncol = v;
    consider(c);
  }
  for (int v : {4, 8, 16}) {
    if (v == res.best.zlines) continue;
    LaunchShape c = res.best;
    c.zlines = v;
    consider(c);
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  set_launch_shape(res.best);
  return res;
}

namespace {
template <typename MT> void append_twiddles(BatchSolver::Impl *m, int n, hipStream_t st) {
  using C = cx_t<MT>;
  if (m->tw_offset.count(n)) return;
  size_t const old = m->tw_len;
  std::vector<C> h(old + n);
  if (old) {
    hip_check(hipMemcpy(h.data(), m->d_tw, sizeof(C) * old, hipMemcpyDeviceToHost), "copy tw");
  }
  for (int k = 0; k < n; ++k) {
    long double const ang = -2.0L * std::numbers::pi_v<long double> * k / n;
    h[old + k] = C{static_cast<MT>(std::cos(ang)), static_cast<MT>(std::sin(ang))};
  }
  hip_check(hipStreamSynchronize(st), "sync");
  void *nt = nullptr;
  hip_check(hipMalloc(&nt, sizeof(C) * h.size()), "hipMalloc twiddles");
  hip_check(hipMemcpy(nt, h.data(), sizeof(C) * h.size(), hipMemcpyHostToDevice), "copy tw");
  dfree(m->d_tw);
  m->d_tw = nt;
  m->tw_offset[n] = static_cast<int>(old);
  m->tw_len = h.size();
}

template <typename MT> void set_kernel_attrs() {
  allow_lds(k_fft_z_r2c<MT>);
  allow_lds(k_fft_z_c2r<MT, 1>);
  allow_lds(k_fft_z_c2r<MT, 3>);
  allow_lds(k_fft_y<MT>);
  allow_lds(k_fft_x_conv<MT>);
  allow_lds(k_fft_y_inv<MT>);
  allow_lds(k_fft_yx_conv<MT>);
  for (int cao = 1; cao <= 7; ++cao)
    dispatch_cao(cao, [&](auto P) {
      for_threads([&](auto NT) {
        allow_lds(k_spread<P.value, MT, 64, NT.value>);
        allow_lds(k_spread<P.value, MT, 64, NT.value, 2, 2, 1>);
        allow_lds(k_spread<P.value, MT, 16, NT.value>);
        allow_lds(k_spread<P.value, MT, 8, NT.value>);
        allow_lds(k_gather_plane<P.value, MT, 64, NT.value>);
        allow_lds(k_gather_plane<P.value, MT, 16, NT.value>);
        allow_lds(k_gather_plane<P.value, MT, 8, NT.value>);
        allow_lds(k_gather_layer<P.value, MT, NT.value>);
      });
      for_columns([&](auto W, auto S) {
        allow_lds(k_spread_col<P.value, MT, W.value, S.value>);
        allow_lds(k_gather_col<P.value, MT, W.value, S.value>);
      });
    });
}

/** G on the half spectrum in the system's layout (FP64 arithmetic). */
void compute_influence(HostSys const &h, MeshPrecision prec, hipStream_t st) {
  SysDesc const &d = h.desc;
  DevGeo const &g = d.g;
  int64_t const K = (int64_t)g.M[0] * g.M[1] * g.Kz;
  int const gg = static_cast<int>(std::min<int64_t>((K + 255) / 256, 16384));
  if (prec == MeshPrecision::FP64)
    k_influence<double><<<gg, 256, 0, st>>>(
        static_cast<double *>(d.G), g.M[0], g.M[1], g.M[2], g.Kz, d.wv[0], d.wv[1], d.wv[2],
        h.params.a[0], h.params.a[1], h.params.a[2], h.params.alpha, d.cao, d.kzmajor);
  else
    k_influence<float><<<gg, 256, 0, st>>>(
        static_cast<float *>(d.G), g.M[0], g.M[1], g.M[2], g.Kz, d.wv[0], d.wv[1], d.wv[2],
        h.params.a[0], h.params.a[1], h.params.a[2], h.params.alpha, d.cao, d.kzmajor);
  hip_check(hipGetLastError(), "k_influence");
}

/** Particle bins of a system for the launch shape: 8^3 bricks, or W x W
 *  columns in (y, z) with one bin per x layer and blocks of `seg` layers. */
void set_bins(HostSys &h, LaunchShape const &shape) {
  DevGeo &g = h.desc.g;
  int const cao = h.desc.cao;
  g.col = shape.column;
  if (shape.column == 0) {
    for (int k = 0; k < 3; ++k) {
      g.B[k] = kBrick;
      g.nb[k] = (g.M[k] + kBrick - 1) / kBrick;
      g.T[k] = kBrick + cao - 1;
    }
    h.desc.lz = 0;
    h.desc.nseg = 0;
  } else {
    int const w = shape.column;
    g.B[0] = 1;
    g.nb[0] = g.M[0];
    g.T[0] = cao;
    for (int k = 1; k < 3; ++k) {
      g.B[k] = w;
      g.nb[k] = (g.M[k] + w - 1) / w;
      g.T[k] = w + cao - 1;
    }
    h.desc.lz = std::max(1, std::min(shape.seg, g.M[0]));
    h.desc.nseg = (g.M[0] + h.desc.lz - 1) / h.desc.lz;
  }
  g.tz_magic = (1ull << 32) / static_cast<unsigned long long>(g.T[2]) + 1;
  g.tyz_magic = (1ull << 32) / static_cast<unsigned long long>(g.T[1] * g.T[2]) + 1;
  h.nbricks = g.nb[0] * g.nb[1] * g.nb[2];
}

/** Blocks of a system in the spread and gather launches. */
int particle_mesh_blocks(HostSys const &h) {
  SysDesc const &d = h.desc;
  return d.g.col ? d.g.nb[1] * d.g.nb[2] * d.nseg : h.nbricks;
}

void recompute_offsets(BatchSolver::Impl *m) {
  int64_t off = 0;
  int boff = 0;
  for (size_t s = 0; s < m->sys.size(); ++s) {
    auto &h = m->sys[s];
    h.part_off = off;
    off += h.n;
    h.brick_off = boff;
    boff += h.nbricks;
    h.desc.n = static_cast<int>(h.n);
    h.desc.part_off = h.part_off;
    h.desc.brick_off = h.brick_off;
    h.desc.off_off = h.brick_off + static_cast<int>(s);
    h.desc.nbricks = h.nbricks;
  }
  m->descs_dirty = true;
  m->version++;
}
} // namespace

int BatchSolver::add_system(P3MParams const &params_in, Vec3d const &box_l, double skin,
                            double prefactor, int64_t n_particles) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  quiesce(m);
  if (!m->kernels_ready) {
    if (m_prec == MeshPrecision::FP64)
      set_kernel_attrs<double>();
    else
      set_kernel_attrs<float>();
    m->kernels_ready = true;
  }
  HostSys h;
  h.params = params_in;
  h.box_l = box_l;
  h.skin = skin;
  h.prefactor = prefactor;
  h.n = n_particles;
  h.params.recalc_a_ai_cao_cut(box_l);
  h.lm.calc_local_ca_mesh(h.params, box_l, skin);
  for (int d = 0; d < 3; ++d)
    if (h.lm.inner[d] != h.params.mesh[d])
      throw std::runtime_error("P3M: charge-assignment mesh does not match the FFT mesh");
  int const cao = h.params.cao;
  if (cao < 1 || cao > 7) throw std::runtime_error("cao must be in [1, 7]");
  SysDesc &d = h.desc;
  DevGeo &g = d.g;
  for (int k = 0; k < 3; ++k) {
    g.M[k] = h.params.mesh[k];
    g.dim[k] = h.lm.dim[k];
    g.marg[k] = h.lm.margin[2 * k];
    g.ld_pos[k] = h.lm.ld_pos[k];
    g.ai[k] = h.params.ai[k];
    g.box_l[k] = box_l[k];
  }
  g.Kz = g.M[2] / 2 + 1;
  g.pos_shift = std::floor((cao - 1) / 2.0) - (cao % 2) / 2.0;
  g.conv = 0;
  for (int k = 0; k < 3; ++k)
    g.hi_f[k] = static_cast<float>(g.M[k]) / static_cast<float>(box_l[k]);
  g.ps_f = static_cast<float>((cao - 1) / 2);
  d.cao = cao;
  set_bins(h, m_shape);
  double const two_pi = 2. * std::numbers::pi;
  double const volume = box_l[0] * box_l[1] * box_l[2];
  for (int k = 0; k < 3; ++k) d.wv[k] = two_pi * (1. / box_l[k]);
  d.hais = (1. / 2. / h.params.alpha) * (1. / 2. / h.params.alpha);
  d.force_prefac = prefactor / volume;
  d.dip_pref = (h.params.epsilon != P3M_EPSILON_METALLIC)
                   ? prefactor * (4. * std::numbers::pi / volume / (2. * h.params.epsilon + 1.))
                   : 0.;
  for (int k = 0; k < 3; ++k) {
    FftPlan1D const p = make_fft_plan(g.M[k]);
    d.plan[k].n = p.n;
    d.plan[k].nstages = p.nstages;
    for (int st = 0; st < 16; ++st) d.plan[k].radix[st] = p.radix[st];
    if (m_prec == MeshPrecision::FP64)
      append_twiddles<double>(m, g.M[k], m_stream);
    else
      append_twiddles<float>(m, g.M[k], m_stream);
  }
  size_t const msz = (m_prec == MeshPrecision::FP64) ? 8 : 4;
  int64_t const R = (int64_t)g.M[0] * g.M[1] * g.M[2];
  int64_t const K = (int64_t)g.M[0] * g.M[1] * g.Kz;
  hip_check(hipMalloc(&d.rho, R * msz), "hipMalloc rho");
  hip_check(hipMalloc(&d.Ey, R * msz), "hipMalloc Ey");
  hip_check(hipMalloc(&d.Ez, R * msz), "hipMalloc Ez");
  hip_check(hipMalloc(&d.khat, K * 2 * msz), "hipMalloc khat");
  hip_check(hipMalloc(&d.A, K * 2 * msz), "hipMalloc A");
  hip_check(hipMalloc(&d.B, K * 2 * msz), "hipMalloc B");
  hip_check(hipMalloc(&d.G, K * msz), "hipMalloc G");
  h.bytes = static_cast<size_t>(3 * R * msz + 3 * K * 2 * msz + K * msz);
  d.kzmajor = uses_fused_pass(g.M[0], g.M[1], m_prec, m_shape) ? 1 : 0;
  compute_influence(h, m_prec, m_stream);
  hip_check(hipStreamSynchronize(m_stream), "sync influence");
  m->sys.push_back(h);
  recompute_offsets(m);
  // twiddle offsets of all systems (the table may have been reallocated)
  for (auto &hs : m->sys)
    for (int k = 0; k < 3; ++k) hs.desc.tw_off[k] = m->tw_offset.at(hs.desc.g.M[k]);
  return static_cast<int>(m->sys.size()) - 1;
}

void BatchSolver::set_particle_count(int s, int64_t n) {
  if (m->sys.at(s).n == n) return;
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  quiesce(m);
  m->sys.at(s).n = n;
  recompute_offsets(m);
}

void BatchSolver::set_convention(int s, int conv) {
  if (conv != 0 && conv != 1) throw std::invalid_argument("P3M: convention must be 0 or 1");
  DevGeo &g = m->sys.at(s).desc.g;
  if (g.conv == conv) return;
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  quiesce(m);
  g.conv = conv;
  m->descs_dirty = true;
}

int BatchSolver::convention(int s) const { return m->sys.at(s).desc.g.conv; }

namespace {
void upload_descs(BatchSolver::Impl *m, hipStream_t st) {
  if (!m->descs_dirty) return;
  size_t const need = m->sys.size();
  if (need > m->d_sys_cap) {
    dfree(m->d_sys);
    hip_check(hipMalloc(&m->d_sys, sizeof(SysDesc) * need), "hipMalloc sys");
    m->d_sys_cap = need;
  }
  std::vector<SysDesc> h(need);
  for (size_t s = 0; s < need; ++s) h[s] = m->sys[s].desc;
  hip_check(hipMemcpyAsync(m->d_sys, h.data(), sizeof(SysDesc) * need, hipMemcpyHostToDevice, st),
            "copy sys");
  hip_check(hipStreamSynchronize(st), "sync sys");
  m->descs_dirty = false;
}

/** Particle scratch for `npart` particles of a selection, brick counters of
 *  all systems, per-system sums; grows only, after draining the lane. */
void ensure_scratch(BatchSolver::Impl *m, Lane &ln, int64_t npart, void *external,
                    size_t external_bytes) {
  int nb_need = 0;
  for (auto const &h : m->sys) nb_need += h.nbricks;
  int const nsys = static_cast<int>(m->sys.size());
  size_t const need = scratch_bytes(npart);
  bool const grow_own = !external && need > ln.own_bytes;
  bool const grow = grow_own || nb_need > ln.counts_cap || nb_need + nsys > ln.offsets_cap ||
                    nsys * 16 > ln.sums_cap;
  if (grow) hip_check(hipStreamSynchronize(ln.stream), "sync before realloc"); // queued work
  if (external) {
    if (external_bytes < need) throw std::invalid_argument("P3M: scratch block too small");
    carve(ln, external, npart);
  } else {
    if (grow_own) {
      if (ln.own) (void)hipFree(ln.own);
      ln.own = nullptr;
      ln.own_bytes = 0;
      size_t const cap = scratch_bytes(std::max<int64_t>(npart + npart / 4, 1024));
      hip_check(hipMalloc(&ln.own, cap), "hipMalloc particle scratch");
      ln.own_bytes = cap;
    }
    carve(ln, ln.own, npart);
  }
  if (nb_need > ln.counts_cap) {
    dfree(ln.counts);
    hip_check(hipMalloc(&ln.counts, sizeof(int) * kCountCopies * std::max(nb_need, 1)),
              "hipMalloc counts");
    ln.counts_cap = nb_need;
  }
  if (nb_need + nsys > ln.offsets_cap) {
    dfree(ln.offsets);
    hip_check(hipMalloc(&ln.offsets, sizeof(int) * std::max(nb_need + nsys, 1)),
              "hipMalloc offsets");
    ln.offsets_cap = nb_need + nsys;
  }
  if (nsys * 16 > ln.sums_cap) {
    dfree(ln.sums);
    hip_check(hipMalloc(&ln.sums, sizeof(double) * std::max(nsys * 16, 16)), "hipMalloc sums");
    if (ln.h_sums) (void)hipHostFree(ln.h_sums);
    hip_check(hipHostMalloc(&ln.h_sums, sizeof(double) * (std::max(nsys * 16, 16) + 8)),
              "hipHostMalloc sums");
    ln.sums_cap = nsys * 16;
  }
}

/** Every device buffer of a fixed-lane solve, carved from the fixed scratch
 *  array in 256-byte aligned pieces: error flag, per-system sums, brick
 *  counters and offsets of all systems, binning and convolution partials,
 *  work tables, then the particle scratch of the selection. */
void carve_fixed(BatchSolver::Impl *m, Lane &ln, size_t table_ints, int64_t npart,
                 int blocks_bin, int conv_blocks) {
  int nb = 0;
  for (auto const &h : m->sys) nb += h.nbricks;
  size_t const nsys = m->sys.size();
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *const p = m->fixed_base + off;
    off += align256(std::max<size_t>(bytes, 1));
    return p;
  };
  ln.err_flag = reinterpret_cast<int *>(take(sizeof(int)));
  ln.sums = reinterpret_cast<double *>(take(sizeof(double) * 16 * nsys));
  ln.counts = reinterpret_cast<int *>(take(sizeof(int) * kCountCopies * nb));
  ln.offsets = reinterpret_cast<int *>(take(sizeof(int) * (nb + nsys)));
  ln.bin_part = reinterpret_cast<double *>(take(sizeof(double) * 8 * blocks_bin));
  ln.conv_part = reinterpret_cast<double *>(take(sizeof(double) * 8 * conv_blocks));
  ln.tables = reinterpret_cast<int *>(take(sizeof(int) * table_ints));
  char *const parts = take(scratch_bytes(npart));
  if (off > kFixedScratchBytes)
    throw std::invalid_argument(
        "P3M: an allocation-free solve of " + std::to_string(npart) + " particles needs " +
        std::to_string(off) + " bytes of scratch, more than the fixed " +
        std::to_string(kFixedScratchBytes) + " (about 44 bytes per particle)");
  carve(ln, parts, npart);
  ln.counts_cap = nb;
  ln.offsets_cap = nb + static_cast<int>(nsys);
  ln.sums_cap = static_cast<int>(16 * nsys);
  ln.bin_cap = blocks_bin;
  ln.conv_cap = conv_blocks;
  ln.tables_cap = table_ints;
}

void build_plan(BatchSolver::Impl *m, Lane &ln, std::vector<int> const &sel, MeshPrecision prec,
                LaunchShape const &shape) {
  Plan &p = ln.plan;
  hipStream_t const st = ln.stream;
  // the fixed lane's tables live in the device-wide fixed scratch, which
  // other contexts' fixed solves overwrite: rebuilt on every solve
  if (!ln.fixed && p.version == m->version && p.sel == sel && p.dbuf) return;
  hip_check(hipStreamSynchronize(st), "sync before plan rebuild"); // queued work reads the tables
  free_plan(p);
  p.sel = sel;
  p.nsel = static_cast<int>(sel.size());
  size_t const csz = (prec == MeshPrecision::FP64) ? 16 : 8;
  // split y/x passes: sizes over the systems that use them; z passes: all
  int max_mx = 1, max_my = 1, max_mz = 1;
  for (int s : sel) {
    auto const &d = m->sys.at(s).desc;
    max_mz = std::max(max_mz, d.g.M[2]);
    if (d.kzmajor) {
      p.lds_f = std::max(p.lds_f, fused_lds_bytes(d.g.M[0], d.g.M[1], csz));
    } else {
      max_mx = std::max(max_mx, d.g.M[0]);
      max_my = std::max(max_my, d.g.M[1]);
    }
    if (d.dip_pref != 0.) p.any_dipole = true;
  }
  size_t const budget = 64 * 1024;
  p.nlp_z = pick_lines(max_mz, csz, 2, shape.zlines, budget);
  // c2r: three components per block when as many lines fit as the other
  // z pass takes, else one component per block (blockIdx.y)
  p.nlp_zc = pick_lines(max_mz, csz, 6, shape.zlines, budget);
  p.c2r_comps = (p.nlp_zc >= p.nlp_z) ? 3 : 1;
  if (p.c2r_comps == 1) p.nlp_zc = p.nlp_z;
  p.ncol_y = pick_lines(max_my, csz, 2, shape.ncol, budget);
  p.ncol_x = pick_lines(max_mx, csz, 4, shape.ncol, budget);
  p.ncol_yi = pick_lines(max_my, csz, 3, shape.ncol, budget);
  auto lds_lines = [csz](int n, int lines) {
    return static_cast<size_t>(n) * csz + static_cast<size_t>(lines) * (n + 1) * csz;
  };
  p.lds_z = lds_lines(max_mz, 2 * p.nlp_z);
  p.lds_zc = lds_lines(max_mz, 2 * p.c2r_comps * p.nlp_zc);
  p.lds_y = lds_lines(max_my, 2 * p.ncol_y);
  p.lds_x = lds_lines(max_mx, 4 * p.ncol_x);
  p.lds_yi = lds_lines(max_my, 3 * p.ncol_yi);
  for (size_t b : {p.lds_z, p.lds_zc, p.lds_y, p.lds_x, p.lds_yi, p.lds_f})
    if (b > 128 * 1024) throw std::runtime_error("P3M: mesh too large for the LDS budget");

  // tables: sel, then prefix arrays (each nsel+1), then per-cao groups
  std::vector<int> host;
  auto push_prefix = [&](auto blocks_of) {
    int const start = static_cast<int>(host.size());
    int acc = 0;
    host.push_back(0);
    for (int s : sel) {
      acc += blocks_of(m->sys[s]);
      host.push_back(acc);
    }
    return std::make_pair(start, acc);
  };
  host.insert(host.end(), sel.begin(), sel.end());
  auto const bin = push_prefix([](HostSys const &h) {
    return static_cast<int>((h.n + kBlock - 1) / kBlock);
  });
  auto const fill = push_prefix([](HostSys const &h) {
    int64_t const R = (int64_t)h.desc.g.M[0] * h.desc.g.M[1] * h.desc.g.M[2];
    return static_cast<int>((R + kFill - 1) / kFill);
  });
  int const nlp = p.nlp_z, nlpc = p.nlp_zc, ncy = p.ncol_y, ncx = p.ncol_x, ncyi = p.ncol_yi;
  auto const fz = push_prefix([nlp](HostSys const &h) {
    int const pairs = (h.desc.g.M[0] * h.desc.g.M[1] + 1) / 2;
    return (pairs + nlp - 1) / nlp;
  });
  auto const fzc = push_prefix([nlpc](HostSys const &h) {
    int const pairs = (h.desc.g.M[0] * h.desc.g.M[1] + 1) / 2;
    return (pairs + nlpc - 1) / nlpc;
  });
  auto const fy = push_prefix([ncy](HostSys const &h) {
    return h.desc.kzmajor ? 0 : ((h.desc.g.Kz + ncy - 1) / ncy) * h.desc.g.M[0];
  });
  auto const fx = push_prefix([ncx](HostSys const &h) {
    return h.desc.kzmajor ? 0 : ((h.desc.g.Kz + ncx - 1) / ncx) * h.desc.g.M[1];
  });
  auto const fyi = push_prefix([ncyi](HostSys const &h) {
    return h.desc.kzmajor ? 0 : ((h.desc.g.Kz + ncyi - 1) / ncyi) * h.desc.g.M[0];
  });
  auto const ff = push_prefix([](HostSys const &h) { return h.desc.kzmajor ? h.desc.g.Kz : 0; });
  struct GH {
    int cao, sel_start, nsel, wt_start, blocks, wsp_start, blocks_sp;
  };
  int const gsx = 2, gsy = 2, gsz = 1; // brick groups of the spread (LaunchShape::group 4)
  std::vector<GH> gh;
  for (int cao = 1; cao <= 7; ++cao) {
    std::vector<int> gs;
    for (int s : sel)
      if (m->sys[s].desc.cao == cao) gs.push_back(s);
    if (gs.empty()) continue;
    GH g{cao, static_cast<int>(host.size()), static_cast<int>(gs.size()), 0, 0, 0, 0};
    host.insert(host.end(), gs.begin(), gs.end());
    g.wt_start = static_cast<int>(host.size());
    int acc = 0;
    host.push_back(0);
    for (int s : gs) {
      acc += particle_mesh_blocks(m->sys[s]);
      host.push_back(acc);
    }
    g.blocks = acc;
    g.wsp_start = g.wt_start;
    g.blocks_sp = g.blocks;
    if (shape.group > 1 && shape.column == 0) {
      g.wsp_start = static_cast<int>(host.size());
      acc = 0;
      host.push_back(0);
      for (int s : gs) {
        acc += brick_groups(m->sys[s].desc.g.nb, gsx, gsy, gsz);
        host.push_back(acc);
      }
      g.blocks_sp = acc;
    }
    gh.push_back(g);
  }
  // per-system offsets of the selection's particles in the lane scratch
  int const soff_start = static_cast<int>(host.size());
  host.resize(host.size() + m->sys.size(), 0);
  int64_t npart = 0;
  for (int s : sel) {
    host[soff_start + s] = static_cast<int>(npart);
    npart += m->sys[s].n;
  }
  if (npart > INT32_MAX) throw std::runtime_error("P3M: too many particles in one selection");
  if (ln.fixed) {
    carve_fixed(m, ln, host.size(), npart, bin.second, fx.second + ff.second);
    hip_check(hipMemcpyAsync(ln.tables, host.data(), sizeof(int) * host.size(),
                             hipMemcpyHostToDevice, st),
              "copy plan");
    hip_check(hipStreamSynchronize(st), "copy plan"); // `host` is a local
  }
  // the lane is idle (synchronised above): its table buffers can be reused
  if (!ln.fixed && host.size() > ln.tables_cap) {
    dfree(ln.tables);
    if (ln.h_tables) (void)hipHostFree(ln.h_tables);
    ln.h_tables = nullptr;
    size_t const cap = host.size() + host.size() / 2 + 64;
    hip_check(hipMalloc(&ln.tables, sizeof(int) * cap), "hipMalloc plan");
    hip_check(hipHostMalloc(&ln.h_tables, sizeof(int) * cap), "hipHostMalloc plan");
    ln.tables_cap = cap;
  }
  p.dbuf = ln.tables;
  if (!ln.fixed) {
    std::memcpy(ln.h_tables, host.data(), sizeof(int) * host.size());
    hip_check(hipMemcpyAsync(p.dbuf, ln.h_tables, sizeof(int) * host.size(),
                             hipMemcpyHostToDevice, st),
              "copy plan");
  }
  p.d_sel = p.dbuf;
  p.wt_bin = p.dbuf + bin.first;
  p.blocks_bin = bin.second;
  p.wt_fill = p.dbuf + fill.first;
  p.blocks_fill = fill.second;
  p.wt_z = p.dbuf + fz.first;
  p.blocks_z = fz.second;
  p.wt_zc = p.dbuf + fzc.first;
  p.blocks_zc = fzc.second;
  p.wt_y = p.dbuf + fy.first;
  p.blocks_y = fy.second;
  p.wt_x = p.dbuf + fx.first;
  p.blocks_x = fx.second;
  p.wt_yi = p.dbuf + fyi.first;
  p.blocks_yi = fyi.second;
  p.wt_f = p.dbuf + ff.first;
  p.blocks_f = ff.second;
  p.soff = p.dbuf + soff_start;
  p.npart = npart;
  for (auto const &g : gh)
    p.groups.push_back({g.cao, g.nsel, p.dbuf + g.sel_start, p.dbuf + g.wt_start, g.blocks,
                        p.dbuf + g.wsp_start, g.blocks_sp});
  p.version = m->version;
  if (ln.fixed) return; // partial-sum buffers carved above
  if (p.blocks_bin > ln.bin_cap) {
    dfree(ln.bin_part); // the stream is idle (synchronised above)
    hip_check(hipMalloc(&ln.bin_part, sizeof(double) * 8 * std::max(p.blocks_bin, 1)),
              "hipMalloc bin partials");
    ln.bin_cap = p.blocks_bin;
  }
  int const conv_blocks = p.blocks_x + p.blocks_f; // split partials, then fused
  if (conv_blocks > ln.conv_cap) {
    dfree(ln.conv_part);
    hip_check(hipMalloc(&ln.conv_part, sizeof(double) * 8 * std::max(conv_blocks, 1)),
              "hipMalloc conv partials");
    ln.conv_cap = conv_blocks;
  }
}

template <typename MT>
void enqueue(BatchSolver::Impl *m, Lane &ln, ParticleArrays const &pa, unsigned flags,
             LaunchShape const &shape, hipEvent_t *ev, int *p_nmp, bool stop_after_spread) {
  using C = cx_t<MT>;
  Plan &p = ln.plan;
  hipStream_t const st = ln.stream;
  bool const want_force = (flags & 1u) && pa.force;
  bool const want_sums = (flags & 6u) != 0; // energy or virial (P3M_ENERGY | P3M_PRESSURE)
  int const ns = p.nsel;
  C const *tw = static_cast<C const *>(m->d_tw);
  SysDesc const *sys = m->d_sys;
  if (ev) (void)hipEventRecord(ev[0], st);
  bool const no_spread = (flags & kSolveNoSpread) != 0;
  bool const zero_rho = !(flags & kSolveAccumulate) && !no_spread;
  k_clear<MT><<<p.blocks_fill, kBlock, 0, st>>>(sys, p.d_sel, p.wt_fill, ns, ln.counts,
                                                ln.err_flag, zero_rho ? 1 : 0);
  if (p.blocks_bin > 0) {
    k_bin<<<p.blocks_bin, kBlock, 0, st>>>(sys, p.d_sel, p.wt_bin, ns, p.soff, pa, ln.counts,
                                            ln.p_brick, ln.p_rank, p_nmp, ln.bin_part,
                                            ln.err_flag);
  }
  k_scan<<<ns, 1024, 0, st>>>(sys, p.d_sel, p.wt_bin, ln.bin_part, ln.sums, ln.counts,
                              ln.offsets);
  if (p.blocks_bin > 0)
    k_scatter<<<p.blocks_bin, kBlock, 0, st>>>(sys, p.d_sel, p.wt_bin, ns, p.soff, pa,
                                                ln.p_brick, ln.p_rank, ln.counts, ln.spos,
                                                ln.sidx);
  if (ev) (void)hipEventRecord(ev[1], st);
  bool const det = (flags & kSolveDeterministic) && !no_spread && !stop_after_spread &&
                   !(flags & (kSolveAccumulate | kSolveMeshOnly));
  if (det && (shape.column || !std::is_same_v<MT, double>))
    throw std::runtime_error("P3M: deterministic solves need FP64 meshes and 8^3 bricks");
  // the fixed-point mesh of deterministic solves: the brick spread flushes
  // integers, the r2c pass converts them
  for (auto const &g : p.groups) {
    if (no_spread) break;
    if (shape.column) {
      dispatch_cao(g.cao, [&](auto P) {
        with_column(shape.column, shape.col_lanes, [&](auto W, auto S) {
          size_t const lds = chunk_bytes(P.value, shape.chunk) +
                             ring_points(P.value, W.value) * sizeof(double);
          k_spread_col<P.value, MT, W.value, S.value><<<g.blocks, 256, lds, st>>>(
              sys, g.d_sel, g.wt, g.nsel, p.soff, ln.offsets,
              static_cast<double4 const *>(ln.spos), shape.chunk);
        });
      });
      continue;
    }
    dispatch_cao(g.cao, [&](auto P) {
      with_threads(shape.threads, [&](auto NT) {
        size_t const lds = chunk_bytes(P.value, shape.chunk) +
                           tile_points(P.value, shape.group) *
                               sizeof(typename SpreadWord<MT>::type);
        auto const args = std::make_tuple(sys, g.d_sel, g.wt_sp, g.nsel, p.soff, ln.offsets,
                                          static_cast<double4 const *>(ln.spos), shape.chunk,
                                          static_cast<double const *>(ln.sums), det ? 1 : 0);
        auto launch = [&](auto kern) {
          std::apply([&](auto... a) { kern<<<g.blocks_sp, NT.value, lds, st>>>(a...); }, args);
        };
        if (shape.group == 4)
          launch(k_spread<P.value, MT, 64, NT.value, 2, 2, 1>);
        else if (shape.spread == kSpreadSeg8)
          launch(k_spread<P.value, MT, 8, NT.value>);
        else if (shape.spread == kSpreadSeg16)
          launch(k_spread<P.value, MT, 16, NT.value>);
        else
          launch(k_spread<P.value, MT, 64, NT.value>);
      });
    });
  }
  if (ev) (void)hipEventRecord(ev[2], st);
  if (stop_after_spread || (flags & kSolveMeshOnly)) {
    // the stage events of a mesh-only solve still bracket it: an event that
    // was never recorded makes hipEventElapsedTime fail
    if (ev)
      for (int k = 3; k <= 6; ++k) (void)hipEventRecord(ev[k], st);
    return;
  }
  k_fft_z_r2c<MT><<<p.blocks_z, kBlock, p.lds_z, st>>>(sys, p.d_sel, p.wt_z, ns, tw, p.nlp_z,
                                                       ln.sums, det ? 1 : 0);
  if (p.blocks_y > 0)
    k_fft_y<MT><<<p.blocks_y, kBlock, p.lds_y, st>>>(sys, p.d_sel, p.wt_y, ns, tw, p.ncol_y);
  if (ev) (void)hipEventRecord(ev[3], st);
  if (p.blocks_x > 0) {
    k_fft_x_conv<MT><<<p.blocks_x, kBlock, p.lds_x, st>>>(sys, p.d_sel, p.wt_x, ns, tw, p.ncol_x,
                                                          want_force ? 1 : 0, want_sums ? 1 : 0,
                                                          ln.conv_part);
    if (want_sums)
      k_reduce_sys<<<ns, 1024, 0, st>>>(p.d_sel, p.wt_x, ns, ln.conv_part, 8, 8, ln.sums, 1);
  }
  if (p.blocks_f > 0) {
    double *fpart = ln.conv_part + 8 * static_cast<size_t>(p.blocks_x);
    k_fft_yx_conv<MT><<<p.blocks_f, kBlock, p.lds_f, st>>>(sys, p.d_sel, p.wt_f, ns, tw,
                                                           want_force ? 1 : 0, want_sums ? 1 : 0,
                                                           fpart);
    if (want_sums)
      k_reduce_sys<<<ns, 1024, 0, st>>>(p.d_sel, p.wt_f, ns, fpart, 8, 8, ln.sums, 1);
  }
  if (ev) (void)hipEventRecord(ev[4], st);
  if (want_force) {
    if (p.blocks_yi > 0)
      k_fft_y_inv<MT><<<p.blocks_yi, kBlock, p.lds_yi, st>>>(sys, p.d_sel, p.wt_yi, ns, tw,
                                                              p.ncol_yi);
    if (p.c2r_comps == 3) {
      k_fft_z_c2r<MT, 3><<<p.blocks_zc, kBlock, p.lds_zc, st>>>(sys, p.d_sel, p.wt_zc, ns, tw,
                                                                p.nlp_zc);
    } else {
      dim3 const gz(p.blocks_zc, 3);
      k_fft_z_c2r<MT, 1><<<gz, kBlock, p.lds_zc, st>>>(sys, p.d_sel, p.wt_zc, ns, tw, p.nlp_zc);
    }
  }
  if (ev) (void)hipEventRecord(ev[5], st);
  if (want_force) {
    for (auto const &g : p.groups) {
      if (shape.column) {
        dispatch_cao(g.cao, [&](auto P) {
          with_column(shape.column, shape.col_lanes, [&](auto W, auto S) {
            size_t const lds = chunk_bytes(P.value, shape.chunk) +
                               3 * ring_points(P.value, W.value) * sizeof(MT);
            k_gather_col<P.value, MT, W.value, S.value><<<g.blocks, 256, lds, st>>>(
                sys, g.d_sel, g.wt, g.nsel, p.soff, ln.offsets,
                static_cast<double4 const *>(ln.spos), static_cast<int const *>(ln.sidx),
                pa.force, static_cast<double const *>(ln.sums), shape.chunk);
          });
        });
        continue;
      }
      dispatch_cao(g.cao, [&](auto P) {
        with_threads(shape.gather_block(), [&](auto NT) {
          size_t const lds =
              chunk_bytes(P.value, shape.chunk) + 3 * tile_points(P.value) * sizeof(MT);
          auto const args =
              std::make_tuple(sys, g.d_sel, g.wt, g.nsel, p.soff, ln.offsets,
                              static_cast<double4 const *>(ln.spos),
                              static_cast<int const *>(ln.sidx), pa.force,
                              static_cast<double const *>(ln.sums), shape.chunk);
          auto launch = [&](auto kern) {
            std::apply([&](auto... a) { kern<<<g.blocks, NT.value, lds, st>>>(a...); }, args);
          };
          if (shape.gather == kGatherPlane8)
            launch(k_gather_plane<P.value, MT, 8, NT.value>);
          else if (shape.gather == kGatherPlane16)
            launch(k_gather_plane<P.value, MT, 16, NT.value>);
          else if (shape.gather == kGatherPlane)
            launch(k_gather_plane<P.value, MT, 64, NT.value>);
          else
            launch(k_gather_layer<P.value, MT, NT.value>);
        });
      });
    }
  }
  if (ev) (void)hipEventRecord(ev[6], st);
}

/** finalize_results on the device for one system: the k-space energy
 *  (p3m.cpp:551-589) and virial (:393-449) from the per-system sums, written
 *  in stream order (no host synchronisation). */
__global__ void k_finalize_device(double const *__restrict__ h, double volume, double alpha,
                                  double epsilon, double prefactor, double *energy,
                                  double *virial) {
  if (threadIdx.x != 0) return;
  if (energy) {
    double e = h[8] / (2. * volume);
    e -= h[2] * alpha * std::numbers::inv_sqrtpi;
    e -= (h[1] * h[1]) * std::numbers::pi / (2. * volume * alpha * alpha);
    if (epsilon != 0.) {
      double const pref = 4. * std::numbers::pi / volume / (2. * epsilon + 1.);
      e += pref * (h[3] * h[3] + h[4] * h[4] + h[5] * h[5]);
    }
    *energy = (h[0] > 0.) ? e * prefactor : 0.;
  }
  if (virial) {
    double const v[9] = {h[9] + h[15], h[10], h[11], h[10], h[12] + h[15], h[13], h[11], h[13],
                         h[14] + h[15]};
    for (int k = 0; k < 9; ++k) virial[k] = v[k] * prefactor / (2. * volume);
  }
}

void finalize_results(BatchSolver::Impl *m, Lane const &ln, std::vector<int> const &sel,
                      std::vector<SolveResult> *out) {
  out->assign(sel.size(), SolveResult{});
  for (size_t i = 0; i < sel.size(); ++i) {
    int const s = sel[i];
    HostSys const &hs = m->sys[s];
    double const *h = ln.h_sums + 16 * s;
    SolveResult &r = (*out)[i];
    r.n_q = h[0];
    r.sum_q = h[1];
    r.sum_q2 = h[2];
    for (int d = 0; d < 3; ++d) r.dipole[d] = h[3 + d];
    // energy (p3m.cpp:551-589)
    double const volume = hs.box_l[0] * hs.box_l[1] * hs.box_l[2];
    double const alpha = hs.params.alpha;
    r.kspace = h[8] / (2. * volume);
    double energy = r.kspace;
    energy -= r.sum_q2 * alpha * std::numbers::inv_sqrtpi;
    energy -= (r.sum_q * r.sum_q) * std::numbers::pi / (2. * volume * alpha * alpha);
    if (hs.params.epsilon != P3M_EPSILON_METALLIC) {
      double const pref = 4. * std::numbers::pi / volume / (2. * hs.params.epsilon + 1.);
      energy += pref * (r.dipole[0] * r.dipole[0] + r.dipole[1] * r.dipole[1] +
                        r.dipole[2] * r.dipole[2]);
    }
    r.energy = (r.n_q > 0.) ? energy * hs.prefactor : 0.;
    // virial tensor (p3m.cpp:393-449)
    double const v[9] = {h[9] + h[15], h[10], h[11], h[10], h[12] + h[15], h[13], h[11], h[13],
                         h[14] + h[15]};
    for (int k = 0; k < 9; ++k) r.virial[k] = v[k] * hs.prefactor / (2. * volume);
  }
}
} // namespace

void BatchSolver::solve(std::vector<int> const &sel, ParticleArrays const &pa, unsigned flags,
                        bool sync, std::vector<SolveResult> *out) {
  solve_on(0, sel, pa, flags, sync, out);
}

void BatchSolver::solve_on(int lane, std::vector<int> const &sel, ParticleArrays const &pa,
                           unsigned flags, bool sync, std::vector<SolveResult> *out,
                           void *scratch, size_t scratch_size) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  if (sel.empty()) {
    if (out) out->clear();
    return;
  }
  for (int s : sel)
    if (s < 0 || s >= num_systems()) throw std::invalid_argument("P3M: invalid system id");
  Lane &ln = *m->lanes.at(lane);
  upload_descs(m, ln.stream);
  build_plan(m, ln, sel, m_prec, m_shape);
  ensure_scratch(m, ln, ln.plan.npart, scratch, scratch_size);
  hipEvent_t *ev = ln.ev;
  if (ln.profiling && static_cast<size_t>(ln.ring_used + 1) * 8 <= ln.ring.size())
    ev = &ln.ring[8 * static_cast<size_t>(ln.ring_used++)];
  if (m_det) flags |= kSolveDeterministic;
  if (m_prec == MeshPrecision::FP64)
    enqueue<double>(m, ln, pa, flags, m_shape, ev, nullptr, false);
  else
    enqueue<float>(m, ln, pa, flags, m_shape, ev, nullptr, false);
  hip_check(hipGetLastError(), "P3M solve launch");
  if (!sync) return;
  int const ns = num_systems() * 16;
  hip_check(hipMemcpyAsync(ln.h_sums, ln.sums, sizeof(double) * ns, hipMemcpyDeviceToHost,
                           ln.stream),
            "copy sums");
  hip_check(hipMemcpyAsync(ln.h_sums + ns, ln.err_flag, sizeof(int), hipMemcpyDeviceToHost,
                           ln.stream),
            "copy err");
  hip_check(hipStreamSynchronize(ln.stream), "P3M solve");
  int err = 0;
  std::memcpy(&err, ln.h_sums + ns, sizeof(int));
  if (err) throw std::runtime_error("P3M: non-finite particle position");
  float t = 0.f;
  (void)hipEventElapsedTime(&t, ev[0], ev[6]);
  last_ms[0] = t;
  for (int st = 1; st <= 6; ++st) {
    (void)hipEventElapsedTime(&t, ev[st - 1], ev[st]);
    last_ms[st] = t;
  }
  if (out) finalize_results(m, ln, sel, out);
}

namespace {
/** This device's address of the fixed scratch array, looked up once per
 *  device (the device must be current). */
char *fixed_scratch_address(int device) {
  static std::mutex mu;
  static std::map<int, char *> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto const it = addr.find(device);
  if (it != addr.end()) return it->second;
  void *p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!p) throw std::runtime_error("P3M: no device address for the fixed scratch array");
  addr[device] = static_cast<char *>(p);
  return static_cast<char *>(p);
}
} // namespace

size_t BatchSolver::fixed_scratch_bytes() { return kFixedScratchBytes; }

void BatchSolver::enable_fixed() {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  char *const base = fixed_scratch_address(m_device);
  if (!m->fixed) {
    auto ln = std::make_unique<Lane>();
    ln->fixed = true;
    ln->stream = m_stream;
    for (auto &e : ln->ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    m->fixed = std::move(ln);
  }
  m->fixed_base = base;
  m->fixed->h_fixed.assign(16 * m->sys.size() + 8, 0.);
  m->fixed->h_sums = m->fixed->h_fixed.data();
  m->fixed_nsys = m->sys.size();
  upload_descs(m, m_stream);
}

void BatchSolver::solve_fixed(std::vector<int> const &sel, ParticleArrays const &pa,
                              unsigned flags, std::vector<SolveResult> *out) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  if (!m->fixed) throw std::runtime_error("P3M: allocation-free solve before its set-up");
  if (m->sys.size() != m->fixed_nsys || m->sys.size() > m->d_sys_cap)
    throw std::runtime_error("P3M: systems were added after the allocation-free set-up");
  if (sel.empty()) {
    if (out) out->clear();
    return;
  }
  for (int s : sel)
    if (s < 0 || s >= num_systems()) throw std::invalid_argument("P3M: invalid system id");
  Lane &ln = *m->fixed;
  upload_descs(m, ln.stream); // copies into the existing descriptor array
  build_plan(m, ln, sel, m_prec, m_shape);
  if (m_det) flags |= kSolveDeterministic;
  if (m_prec == MeshPrecision::FP64)
    enqueue<double>(m, ln, pa, flags, m_shape, ln.ev, nullptr, false);
  else
    enqueue<float>(m, ln, pa, flags, m_shape, ln.ev, nullptr, false);
  hip_check(hipGetLastError(), "P3M solve launch");
  int const ns = num_systems() * 16;
  hip_check(hipMemcpyAsync(ln.h_sums, ln.sums, sizeof(double) * ns, hipMemcpyDeviceToHost,
                           ln.stream),
            "copy sums");
  hip_check(hipMemcpyAsync(ln.h_sums + ns, ln.err_flag, sizeof(int), hipMemcpyDeviceToHost,
                           ln.stream),
            "copy err");
  hip_check(hipStreamSynchronize(ln.stream), "P3M solve");
  int err = 0;
  std::memcpy(&err, ln.h_sums + ns, sizeof(int));
  if (err) throw std::runtime_error("P3M: non-finite particle position");
  float t = 0.f;
  (void)hipEventElapsedTime(&t, ln.ev[0], ln.ev[6]);
  last_ms[0] = t;
  for (int st = 1; st <= 6; ++st) {
    (void)hipEventElapsedTime(&t, ln.ev[st - 1], ln.ev[st]);
    last_ms[st] = t;
  }
  if (out) finalize_results(m, ln, sel, out);
}

void BatchSolver::device_results(int lane, int s, double *d_energy, double *d_virial) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  Lane &ln = *m->lanes.at(lane);
  HostSys const &hs = m->sys.at(s);
  if (!ln.sums || 16 * (s + 1) > ln.sums_cap)
    throw std::runtime_error("P3M: device results before a solve of this system");
  double const volume = hs.box_l[0] * hs.box_l[1] * hs.box_l[2];
  k_finalize_device<<<1, 64, 0, ln.stream>>>(ln.sums + 16 * s, volume, hs.params.alpha,
                                             hs.params.epsilon, hs.prefactor, d_energy, d_virial);
  hip_check(hipGetLastError(), "P3M device results");
}

void BatchSolver::prepare_on(int lane, std::vector<int> const &sel, double *pos, double *force,
                             double sigma, std::vector<unsigned long long> const &seeds) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  if (sel.empty()) return;
  if (seeds.size() != sel.size()) throw std::invalid_argument("P3M: one seed per system");
  Lane &ln = *m->lanes.at(lane);
  upload_descs(m, ln.stream);
  build_plan(m, ln, sel, m_prec, m_shape);
  int const ns = static_cast<int>(sel.size());
  if (ns > ln.seeds_cap) {
    hip_check(hipStreamSynchronize(ln.stream), "sync before realloc");
    dfree(ln.seeds);
    if (ln.h_seeds) (void)hipHostFree(ln.h_seeds);
    ln.h_seeds = nullptr;
    hip_check(hipMalloc(&ln.seeds, sizeof(unsigned long long) * ns), "hipMalloc seeds");
    hip_check(hipHostMalloc(&ln.h_seeds, sizeof(unsigned long long) * ns), "hipHostMalloc");
    ln.seeds_cap = ns;
  }
  // the previous copy out of h_seeds has completed: the lane's earlier work
  // was drained by the caller before it formed this batch
  std::memcpy(ln.h_seeds, seeds.data(), sizeof(unsigned long long) * ns);
  hip_check(hipMemcpyAsync(ln.seeds, ln.h_seeds, sizeof(unsigned long long) * ns,
                           hipMemcpyHostToDevice, ln.stream),
            "copy seeds");
  Plan const &p = ln.plan;
  if (p.blocks_bin > 0)
    k_prepare<<<p.blocks_bin, kBlock, 0, ln.stream>>>(m->d_sys, p.d_sel, p.wt_bin, p.nsel, pos,
                                                       force, sigma, ln.seeds);
  hip_check(hipGetLastError(), "P3M prepare launch");
}

void BatchSolver::graph_solve(std::vector<int> const &sel, ParticleArrays const &pa,
                              unsigned flags) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  if (sel.empty()) return;
  Lane &ln = *m->lanes[0];
  upload_descs(m, ln.stream);
  build_plan(m, ln, sel, m_prec, m_shape);
  ensure_scratch(m, ln, ln.plan.npart, nullptr, 0);
  if (m_det) flags |= kSolveDeterministic;
  Plan &p = ln.plan;
  bool const same = p.graph && p.graph_flags == flags && p.graph_pa.pos == pa.pos &&
                    p.graph_pa.q == pa.q && p.graph_pa.upos == pa.upos &&
                    p.graph_pa.force == pa.force;
  if (!same) {
    if (p.graph) (void)hipGraphExecDestroy(p.graph);
    p.graph = nullptr;
    hipGraph_t graph;
    hip_check(hipStreamBeginCapture(ln.stream, hipStreamCaptureModeRelaxed), "begin capture");
      if (m_prec == MeshPrecision::FP64)
      enqueue<double>(m, ln, pa, flags, m_shape, nullptr, nullptr, false);
    else
      enqueue<float>(m, ln, pa, flags, m_shape, nullptr, nullptr, false);
    hip_check(hipStreamEndCapture(ln.stream, &graph), "end capture");
    hip_check(hipGraphInstantiate(&p.graph, graph, nullptr, nullptr, 0), "graph instantiate");
    (void)hipGraphDestroy(graph);
    p.graph_pa = pa;
    p.graph_flags = flags;
  }
  hip_check(hipGraphLaunch(p.graph, ln.stream), "graph launch");
}

void BatchSolver::profile_begin(int max_solves) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  Lane &ln = *m->lanes[0];
  hip_check(hipStreamSynchronize(ln.stream), "sync");
  size_t const need = static_cast<size_t>(std::max(max_solves, 1)) * 8;
  while (ln.ring.size() < need) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    ln.ring.push_back(e);
  }
  ln.ring_used = 0;
  ln.profiling = true;
}

int BatchSolver::profile_end(double *stage_ms) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  Lane &ln = *m->lanes[0];
  hip_check(hipStreamSynchronize(ln.stream), "sync");
  for (int k = 0; k < 7; ++k) stage_ms[k] = 0.;
  for (int i = 0; i < ln.ring_used; ++i) {
    hipEvent_t *ev = &ln.ring[8 * static_cast<size_t>(i)];
    float t = 0.f;
    (void)hipEventElapsedTime(&t, ev[0], ev[6]);
    stage_ms[0] += t;
    for (int s = 1; s <= 6; ++s) {
      (void)hipEventElapsedTime(&t, ev[s - 1], ev[s]);
      stage_ms[s] += t;
    }
  }
  ln.profiling = false;
  int const n = ln.ring_used;
  ln.ring_used = 0;
  return n;
}

void BatchSolver::debug_charge_mesh(int s, ParticleArrays const &pa, std::vector<double> &out) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  Lane &ln = *m->lanes[0];
  upload_descs(m, ln.stream);
  build_plan(m, ln, {s}, m_prec, m_shape);
  ensure_scratch(m, ln, ln.plan.npart, nullptr, 0);
  if (m_prec == MeshPrecision::FP64)
    enqueue<double>(m, ln, pa, 0, m_shape, ln.ev, nullptr, true);
  else
    enqueue<float>(m, ln, pa, 0, m_shape, ln.ev, nullptr, true);
  hip_check(hipGetLastError(), "P3M spread launch");
  auto const &g = m->sys[s].desc.g;
  int64_t const R = (int64_t)g.M[0] * g.M[1] * g.M[2];
  out.resize(R);
  if (m_prec == MeshPrecision::FP64) {
    hip_check(hipMemcpyAsync(out.data(), m->sys[s].desc.rho, R * sizeof(double),
                             hipMemcpyDeviceToHost, m_stream),
              "copy rho");
    hip_check(hipStreamSynchronize(m_stream), "sync");
  } else {
    std::vector<float> tmp(R);
    hip_check(hipMemcpyAsync(tmp.data(), m->sys[s].desc.rho, R * sizeof(float),
                             hipMemcpyDeviceToHost, m_stream),
              "copy rho");
    hip_check(hipStreamSynchronize(m_stream), "sync");
    for (int64_t i = 0; i < R; ++i) out[i] = tmp[i];
  }
}

void BatchSolver::debug_indices(int s, ParticleArrays const &pa, std::vector<int32_t> &out) {
  hip_check(hipSetDevice(m_device), "hipSetDevice");
  Lane &ln = *m->lanes[0];
  upload_descs(m, ln.stream);
  build_plan(m, ln, {s}, m_prec, m_shape);
  ensure_scratch(m, ln, ln.plan.npart, nullptr, 0);
  int64_t const n = m->sys[s].n;
  int *d_nmp = nullptr;
  hip_check(hipMalloc(&d_nmp, sizeof(int) * 3 * std::max<int64_t>(n, 1)), "hipMalloc nmp");
  try {
      if (m_prec == MeshPrecision::FP64)
      enqueue<double>(m, ln, pa, 0, m_shape, ln.ev, d_nmp, true);
    else
      enqueue<float>(m, ln, pa, 0, m_shape, ln.ev, d_nmp, true);
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

void pair_forces_targets(int device, hipStream_t stream, int64_t n, double const *pos,
                         double const *q, int ntargets, int const *targets, Vec3d const &box,
                         double prefactor, double alpha, double r_cut, double *out) {
  hip_check(hipSetDevice(device), "hipSetDevice");
  if (ntargets <= 0) return;
  k_pair_forces_targets<<<ntargets, kBlock, 0, stream>>>(n, pos, q, targets, box[0], box[1],
                                                         box[2], prefactor, alpha, r_cut, out);
  hip_check(hipGetLastError(), "pair force launch");
  hip_check(hipStreamSynchronize(stream), "pair forces");
}

} // namespace espresso_amd
