// Device building blocks shared by the P3M kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace espresso_amd::dev {

template <typename T> struct CxOf;
template <> struct CxOf<double> { using type = double2; };
template <> struct CxOf<float> { using type = float2; };
template <typename T> using cx_t = typename CxOf<T>::type;

template <typename C> __host__ __device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C> __host__ __device__ __forceinline__ C cadd(C a, C b) {
  return C{a.x + b.x, a.y + b.y};
}
template <typename C> __host__ __device__ __forceinline__ C csub(C a, C b) {
  return C{a.x - b.x, a.y - b.y};
}
template <typename C> __host__ __device__ __forceinline__ C cconj(C a) { return C{a.x, -a.y}; }
template <typename C, typename T> __host__ __device__ __forceinline__ C cscale(C a, T s) {
  return C{s * a.x, s * a.y};
}
/** a + s b */
template <typename C, typename T> __host__ __device__ __forceinline__ C caxpy(C a, T s, C b) {
  return C{a.x + s * b.x, a.y + s * b.y};
}
/** multiply by -i (forward) or +i (inverse) */
template <typename C> __host__ __device__ __forceinline__ C cmul_mi(C a, bool inv) {
  return inv ? C{-a.y, a.x} : C{a.y, -a.x};
}

// ---------------------------------------------------------------------------
// Cardinal B-spline weights (Hockney & Eastwood 5-189).  For order P and
// stencil point i the weight at offset x in [-1/2, 1/2] is M_P(i + 1/2 - x),
// M_P(u) = 1/(P-1)! sum_{j<=floor(u)} (-1)^j C(P,j) (u-j)^(P-1).  The integer
// numerators and denominators are derived here at compile time and reduced by
// their gcd, which yields the tabulated polynomials of
// src/utils/include/utils/math/bspline.hpp:30-177.
// ---------------------------------------------------------------------------
template <int P> struct BsplineTable {
  double num[P][P];
  double den[P];
  double inv_den[P]; // 1 / den, correctly rounded at compile time
};

constexpr long long cx_binom(int n, int k) {
  if (k < 0 or k > n) return 0;
  long long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}
constexpr long long cx_ipow(long long b, int e) {
  long long r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}
constexpr long long cx_gcd(long long a, long long b) {
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b != 0) {
    long long const t = a % b;
    a = b;
    b = t;
  }
  return a;
}

template <int P> constexpr BsplineTable<P> make_bspline_table() {
  BsplineTable<P> t{};
  long long fact = 1;
  for (int k = 2; k <= P - 1; ++k) fact *= k;
  for (int i = 0; i < P; ++i) {
    long long c[P] = {};
    for (int j = 0; j <= i; ++j) {
      long long const sgn = (j % 2) ? -1 : 1;
      long long const a = 2 * (i - j) + 1;
      for (int m = 0; m < P; ++m)
        c[m] += sgn * cx_binom(P, j) * cx_binom(P - 1, m) * cx_ipow(a, P - 1 - m) *
                cx_ipow(-2, m);
    }
    long long const den = cx_ipow(2, P - 1) * fact;
    long long g = den;
    for (int m = 0; m < P; ++m) g = cx_gcd(g, c[m]);
    for (int m = 0; m < P; ++m) t.num[i][m] = static_cast<double>(c[m] / g);
    t.den[i] = static_cast<double>(den / g);
    t.inv_den[i] = 1.0 / t.den[i];
  }
  return t;
}

/** w[i] = B-spline weight of stencil point i at offset x (FP64 Horner,
 *  integer-coefficient numerator times the reciprocal of its denominator). */
template <int P> __device__ __forceinline__ void bspline_weights(double x, double *w) {
  constexpr BsplineTable<P> T = make_bspline_table<P>();
#pragma unroll
  for (int i = 0; i < P; ++i) {
    double acc = T.num[i][P - 1];
#pragma unroll
    for (int m = P - 2; m >= 0; --m) acc = T.num[i][m] + x * acc;
    w[i] = acc * T.inv_den[i];
  }
}

// ---------------------------------------------------------------------------
// Bit-exact mesh coordinate of a particle (interpolation.hpp:144-160):
//   t = ((x - ld_pos) * ai) - pos_shift,  nmp = (int) t,  dist = (t - nmp) - 0.5
// evaluated with every operation rounded separately, as on the reference's
// x86-64 build.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ca_coordinate(double x, double ld_pos, double ai,
                                                double pos_shift) {
#pragma clang fp contract(off)
  double const u = x - ld_pos;
  double const v = u * ai;
  return v - pos_shift;
}
__device__ __forceinline__ double ca_distance(double t, int nmp) {
#pragma clang fp contract(off)
  double const u = t - static_cast<double>(nmp);
  return u - 0.5;
}

/** Geometry of one system shared by the particle kernels. */
struct DevGeo {
  int M[3];      // mesh
  int Kz;        // Mz/2 + 1
  int dim[3];    // CA mesh dims
  int marg[3];   // CA-local index of the FFT origin (left margin)
  int B[3];      // brick dims (first stencil points per brick)
  int nb[3];     // bricks per dim
  int T[3];      // tile dims B + cao - 1
  unsigned long long tz_magic, tyz_magic; // floor(2^32 / d) + 1 for d = T[2], T[1]*T[2]
  double ld_pos[3];
  double ai[3];
  double box_l[3];
  double pos_shift;
  // index rule: 0 = the CPU kernels' CA mesh (above); 1 = the FP32 rule of
  // the reference's GPU kernels, m = fma(x, hi_f, -ps_f), nmp = floor(m + 1/2)
  int conv;
  float hi_f[3]; // float(M) / float(L)
  float ps_f;    // float((cao - 1) / 2), integer division
  // particle bins: 0 = bricks of B[0] x B[1] x B[2] first stencil points;
  // W > 0 = columns of W x W points in (y, z), one bin per x layer (the bins
  // of a column are consecutive, x fastest)
  int col;
};

/** floor(e / d) for 0 <= e < 2^16 and 1 <= d < 2^16 from the magic number
 *  floor(2^32 / d) + 1 (exact in that range). */
__device__ __forceinline__ int magic_div(int e, unsigned long long magic) {
  return static_cast<int>((static_cast<unsigned long long>(e) * magic) >> 32);
}

/** Division by a block-uniform divisor d >= 1 as a multiply-high with the
 *  magic number m = floor((2^32 - 1) / d) + 1; exact for 0 <= e with
 *  e * d < 2^32 (d = 1 is passed through).  Build it once per block or per
 *  FFT stage from a uniform d. */
struct FastDiv {
  unsigned m;
  int d;
};
__device__ __forceinline__ FastDiv make_fastdiv(int d) {
  return FastDiv{d > 1 ? 0xFFFFFFFFu / static_cast<unsigned>(d) + 1u : 0u, d};
}
__device__ __forceinline__ int fdiv(int e, FastDiv f) {
  return f.d == 1 ? e : static_cast<int>(__umulhi(static_cast<unsigned>(e), f.m));
}

/** Periodic mesh index of tile row/column v = o + a >= 0: one conditional
 *  subtraction (v < 2M, every mesh at least as wide as a tile), a remainder
 *  for the rare lanes of tiny meshes whose tiles wrap more than once. */
__device__ __forceinline__ int wrap_once(int v, int M) {
  int const w = v - M;
  v = (w < 0) ? v : w;
  if (v >= M) v %= M;
  return v;
}

// ---------------------------------------------------------------------------
// DPP lane exchanges on FP64 values (two 32-bit moves per value).
// ---------------------------------------------------------------------------
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v) {
  int const lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  int const hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  int const lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  int const hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

/** Sum over aligned segments of S lanes (S in {1, 4, 8, 16, 32, 64}); the
 *  result is valid in the first lane of each segment.  Must be executed by
 *  all 64 lanes (lanes outside the work contribute zeros). */
template <int S> __device__ __forceinline__ double segment_total(double v) {
  static_assert(S == 1 || S == 4 || S == 8 || S == 16 || S == 32 || S == 64,
                "unsupported segment size");
  if constexpr (S >= 4) {
    v += dpp_f64<0xB1>(v); // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v); // quad_perm [2,3,0,1]: every lane holds its quad sum
  }
  if constexpr (S >= 8) v += dpp_f64<0x141>(v);  // row_half_mirror: 8-lane sums
  if constexpr (S >= 16) v += dpp_f64<0x140>(v); // row_mirror: 16-lane row sums
  if constexpr (S == 32) {
    double const a = readlane_f64(v, 0) + readlane_f64(v, 16);
    double const b = readlane_f64(v, 32) + readlane_f64(v, 48);
    v = ((threadIdx.x & 63) < 32) ? a : b;
  } else if constexpr (S == 64) {
    v = (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
  }
  return v;
}

/** One dimension of the index rule of the reference's GPU kernels
 *  (p3m_gpu_cuda.cu:334-345): the position folded into [0, L) in FP64
 *  (x - L floor(x / L), every operation rounded; the reference's particles
 *  arrive folded, GpuParticleData.cpp:90) and cast to float, m = x hi -
 *  pos_shift as one FMA, nmp = floor(m + 1/2), dist = m - nmp in float.
 *  gi is the periodic index of the first stencil point nmp; the CA-local
 *  index is reported as gi as well.  Returns false for non-finite x. */
__device__ __forceinline__ bool gpu_coordinate_1d(DevGeo const &g, double x, int d, int &nmp,
                                                  int &gi, double &dist) {
#pragma clang fp contract(off)
  bool const ok = isfinite(x);
  double const L = g.box_l[d];
  double xf = ok ? x : 0.;
  double const k = floor(xf / L);
  double const sh = L * k;
  xf = xf - sh;
  if (xf >= L) xf = xf - L;
  float const xs = static_cast<float>(xf);
  float const m = __fmaf_rn(xs, g.hi_f[d], -g.ps_f);
  float const nf = floorf(m + 0.5f);
  int const n = static_cast<int>(nf);
  dist = static_cast<double>(m - nf);
  int gg = n % g.M[d];
  if (gg < 0) gg += g.M[d];
  gi = gg;
  nmp = gg;
  return ok;
}

/** Mesh coordinates of one particle for stencil order P: CA-local first
 *  stencil point nmp, its periodic FFT-mesh index gi and the offsets dist.
 *  Positions whose stencil would leave the CA mesh (further outside the box
 *  than the Verlet skin, which the reference never sees) are folded into
 *  the box first.  Returns false for non-finite positions. */
__device__ __forceinline__ bool particle_coordinates_rt(DevGeo const &g, int P, double const x[3],
                                                        int nmp[3], int gi[3], double dist[3]) {
  bool ok = true;
  if (g.conv) {
#pragma unroll
    for (int d = 0; d < 3; ++d) ok &= gpu_coordinate_1d(g, x[d], d, nmp[d], gi[d], dist[d]);
    return ok;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double t = ca_coordinate(x[d], g.ld_pos[d], g.ai[d], g.pos_shift);
    if (!(t >= 0. && t < static_cast<double>(g.dim[d] - P + 1))) {
      double xf = x[d] - g.box_l[d] * floor(x[d] / g.box_l[d]);
      if (xf >= g.box_l[d]) xf -= g.box_l[d];
      t = ca_coordinate(xf, g.ld_pos[d], g.ai[d], g.pos_shift);
      if (!(t >= 0. && t < static_cast<double>(g.dim[d] - P + 1))) {
        ok = false;
        t = 0.;
      }
    }
    int const n = static_cast<int>(t);
    nmp[d] = n;
    dist[d] = ca_distance(t, n);
    int gg = n - g.marg[d];
    gg %= g.M[d];
    if (gg < 0) gg += g.M[d];
    gi[d] = gg;
  }
  return ok;
}
/** One dimension d of particle_coordinates_rt (same operations, so the
 *  result is identical). */
__device__ __forceinline__ bool coordinate_1d(DevGeo const &g, int P, double x, int d, int &nmp,
                                              int &gi, double &dist) {
  if (g.conv) return gpu_coordinate_1d(g, x, d, nmp, gi, dist);
  bool ok = true;
  double t = ca_coordinate(x, g.ld_pos[d], g.ai[d], g.pos_shift);
  if (!(t >= 0. && t < static_cast<double>(g.dim[d] - P + 1))) {
    double xf = x - g.box_l[d] * floor(x / g.box_l[d]);
    if (xf >= g.box_l[d]) xf -= g.box_l[d];
    t = ca_coordinate(xf, g.ld_pos[d], g.ai[d], g.pos_shift);
    if (!(t >= 0. && t < static_cast<double>(g.dim[d] - P + 1))) {
      ok = false;
      t = 0.;
    }
  }
  int const n = static_cast<int>(t);
  nmp = n;
  dist = ca_distance(t, n);
  int gg = n - g.marg[d];
  gg %= g.M[d];
  if (gg < 0) gg += g.M[d];
  gi = gg;
  return ok;
}

template <int P>
__device__ __forceinline__ bool particle_coordinates(DevGeo const &g, double const x[3],
                                                     int nmp[3], int gi[3], double dist[3]) {
  return particle_coordinates_rt(g, P, x, nmp, gi, dist);
}

/** i k_d for FFT index j of a dimension of size M with the Nyquist entry set
 *  to zero (d_op = calc_meshift(mesh, true), data_struct.hpp:68-70). */
__device__ __forceinline__ int dop_shift(int j, int M) {
  if (2 * j == M) return 0;
  return (2 * j < M) ? j : j - M;
}
/** Wave-number index of the ik-differentiation of the field components:
 *  the CPU kernels' d_op (Nyquist zero, dop_shift) or, for the GPU index
 *  rule, n = j - M for j > M/2 with the Nyquist entry kept at +M/2
 *  (apply_diff_op, p3m_gpu_cuda.cu:265-295). */
__device__ __forceinline__ int ik_shift(int j, int M, int conv) {
  if (conv) return (2 * j > M) ? j - M : j;
  return dop_shift(j, M);
}
/** calc_meshift(mesh, false) entry (common.hpp:248-264). */
__device__ __forceinline__ int plain_shift(int j, int M) { return (2 * j < M) ? j : j - M; }

// ---------------------------------------------------------------------------
// wave64 / block reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

/** Sum K values over the block (blockDim.x multiple of 64, <= 1024); the
 *  result is valid in thread 0.  `scratch` needs K * 16 doubles. */
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *scratch) {
  int const lane = threadIdx.x & 63;
  int const wid = threadIdx.x >> 6;
  int const nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) scratch[k * 16 + wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.;
      for (int w = 0; w < nw; ++w) s += scratch[k * 16 + w];
      v[k] = s;
    }
  }
  __syncthreads();
}

/** Maximum of v over the block (blockDim.x multiple of 64, <= 1024); the
 *  result is valid in thread 0.  `scratch` needs 16 doubles. */
__device__ __forceinline__ double block_max(double v, double *scratch) {
  int const lane = threadIdx.x & 63;
  int const wid = threadIdx.x >> 6;
  int const nw = (blockDim.x + 63) >> 6;
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 0; w < nw; ++w) v = fmax(v, scratch[w]);
  __syncthreads();
  return v;
}

// ---------------------------------------------------------------------------
// LDS-resident batched 1D FFT (Stockham autosort, mixed radix).
//   stage with radix R, current span Ns, butterfly j in [0, n/R):
//     v_r   = in[j + r n/R] * w_n^(r (j mod Ns) n/(Ns R))
//     V     = DFT_R(v)
//     out[(j / Ns) Ns R + (j mod Ns) + r Ns] = V_r
// Twiddles come from an LDS table tw[m] = exp(-2 pi i m / n) (conjugated for
// the inverse).  Radices 2, 4, 8 have closed butterflies, 3, 5, 6, 7 use a
// register DFT with table roots, larger primes one output per work item.
// Unnormalised in both directions, as FFTW in the reference (fft.cpp:732-769).
// ---------------------------------------------------------------------------
struct FftPlanDev {
  int n;
  int nstages;
  int radix[16];
};

template <typename C> __host__ __device__ __forceinline__ C tw_get(C const *tw, int m, bool inv) {
  C w = tw[m];
  if (inv) w.y = -w.y;
  return w;
}

/** 3-point DFT of (a, b, c) with constant twiddles: with t = b + c,
 *  X0 = a + t, X1,2 = a - t/2 -+ i sin(pi/3) (b - c) (signs of i flipped for
 *  the inverse). */
template <typename C> __host__ __device__ __forceinline__ void dft3(C &a, C &b, C &c, bool inv) {
  using T = decltype(a.x);
  T const s3 = T(0.86602540378443864676372317075293618347); // sin(pi/3)
  C const t = cadd(b, c);
  C const m = caxpy(a, T(-0.5), t);
  C const d = cscale(cmul_mi(csub(b, c), inv), s3);
  a = cadd(a, t);
  b = cadd(m, d);
  c = csub(m, d);
}

/** In-register R-point DFT, X_k = sum_r v_r w^(rk), w = exp(-+2 pi i / R).
 *  Radices 2-8 use constant twiddles (radix 3, 5, 7: the symmetric
 *  a_m = v_m + v_(R-m), b_m = v_m - v_(R-m) split; radix 6: two radix-3 DFTs
 *  of the even and odd points merged with w_6^k); other radices use the
 *  length-n twiddle table tw (stride n / R). */
template <typename C, int R>
__host__ __device__ __forceinline__ void dft_reg(C (&v)[R], C const *tw, int n, bool inv) {
  using T = decltype(v[0].x);
  if constexpr (R == 1) {
  } else if constexpr (R == 3) {
    dft3(v[0], v[1], v[2], inv);
  } else if constexpr (R == 5) {
    T const c1 = T(0.30901699437494742410229341718281905886);  // cos(2 pi / 5)
    T const c2 = T(-0.80901699437494742410229341718281905886); // cos(4 pi / 5)
    T const s1 = T(0.95105651629515357211643933337938214340);  // sin(2 pi / 5)
    T const s2 = T(0.58778525229247312916870595463907276860);  // sin(4 pi / 5)
    C const a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    C const a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    C const p1 = caxpy(caxpy(v[0], c1, a1), c2, a2);
    C const p2 = caxpy(caxpy(v[0], c2, a1), c1, a2);
    C const r1 = cmul_mi(caxpy(cscale(b1, s1), s2, b2), inv);
    C const r2 = cmul_mi(caxpy(cscale(b1, s2), -s1, b2), inv);
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(p1, r1);
    v[4] = csub(p1, r1);
    v[2] = cadd(p2, r2);
    v[3] = csub(p2, r2);
  } else if constexpr (R == 6) {
    T const s3 = T(0.86602540378443864676372317075293618347); // sin(pi/3)
    C e0 = v[0], e1 = v[2], e2 = v[4];
    C o0 = v[1], o1 = v[3], o2 = v[5];
    dft3(e0, e1, e2, inv);
    dft3(o0, o1, o2, inv);
    // w_6 = 1/2 -+ i sin(pi/3), w_6^2 = -1/2 -+ i sin(pi/3)
    C const t1 = caxpy(cscale(o1, T(0.5)), s3, cmul_mi(o1, inv));
    C const t2 = caxpy(cscale(o2, T(-0.5)), s3, cmul_mi(o2, inv));
    v[0] = cadd(e0, o0);
    v[3] = csub(e0, o0);
    v[1] = cadd(e1, t1);
    v[4] = csub(e1, t1);
    v[2] = cadd(e2, t2);
    v[5] = csub(e2, t2);
  } else if constexpr (R == 7) {
    T const c1 = T(0.62348980185873353052500488400423981063);  // cos(2 pi / 7)
    T const c2 = T(-0.22252093395631440428890256449679475947); // cos(4 pi / 7)
    T const c3 = T(-0.90096886790241912623610231950744505117); // cos(6 pi / 7)
    T const s1 = T(0.78183148246802980870844452667405775023);  // sin(2 pi / 7)
    T const s2 = T(0.97492791218182360701813168299393121723);  // sin(4 pi / 7)
    T const s3 = T(0.43388373911755812047576833284835875461);  // sin(6 pi / 7)
    C const a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
    C const a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
    C const a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
    // X_k = v0 + sum_m cos(2 pi mk/7) a_m -+ i sum_m sin(2 pi mk/7) b_m, k = 1..3
    C const p1 = caxpy(caxpy(caxpy(v[0], c1, a1), c2, a2), c3, a3);
    C const p2 = caxpy(caxpy(caxpy(v[0], c2, a1), c3, a2), c1, a3);
    C const p3 = caxpy(caxpy(caxpy(v[0], c3, a1), c1, a2), c2, a3);
    C const r1 = cmul_mi(caxpy(caxpy(cscale(b1, s1), s2, b2), s3, b3), inv);
    C const r2 = cmul_mi(caxpy(caxpy(cscale(b1, s2), -s3, b2), -s1, b3), inv);
    C const r3 = cmul_mi(caxpy(caxpy(cscale(b1, s3), -s1, b2), s2, b3), inv);
    v[0] = cadd(v[0], cadd(a1, cadd(a2, a3)));
    v[1] = cadd(p1, r1);
    v[6] = csub(p1, r1);
    v[2] = cadd(p2, r2);
    v[5] = csub(p2, r2);
    v[3] = cadd(p3, r3);
    v[4] = csub(p3, r3);
  } else if constexpr (R == 2) {
    C const a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    C const s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    C const s13 = cadd(v[1], v[3]), d13 = cmul_mi(csub(v[1], v[3]), inv);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  } else if constexpr (R == 8) {
    T const h = T(0.70710678118654752440084436210484903928);
    C a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a[r] = cadd(v[r], v[r + 4]);
      b[r] = csub(v[r], v[r + 4]);
    }
    // b_r *= w_8^r  (forward w_8 = exp(-i pi/4))
    b[1] = inv ? C{h * (b[1].x - b[1].y), h * (b[1].x + b[1].y)}
               : C{h * (b[1].x + b[1].y), h * (b[1].y - b[1].x)};
    b[2] = cmul_mi(b[2], inv);
    b[3] = inv ? C{-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y)}
               : C{h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y)};
    C const as02 = cadd(a[0], a[2]), ad02 = csub(a[0], a[2]);
    C const as13 = cadd(a[1], a[3]), ad13 = cmul_mi(csub(a[1], a[3]), inv);
    C const bs02 = cadd(b[0], b[2]), bd02 = csub(b[0], b[2]);
    C const bs13 = cadd(b[1], b[3]), bd13 = cmul_mi(csub(b[1], b[3]), inv);
    v[0] = cadd(as02, as13);
    v[4] = csub(as02, as13);
    v[2] = cadd(ad02, ad13);
    v[6] = csub(ad02, ad13);
    v[1] = cadd(bs02, bs13);
    v[5] = csub(bs02, bs13);
    v[3] = cadd(bd02, bd13);
    v[7] = csub(bd02, bd13);
  } else {
    int const step = n / R;
    C out[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      C s = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) s = cadd(s, cmul(v[r], tw_get(tw, ((r * k) % R) * step, inv)));
      out[k] = s;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
}

/** One radix-R Stockham butterfly j of a stage with Ns = product of the
 *  radices already applied; `fNs` divides by Ns and `tstride` = n / (Ns R). */
template <typename C, int R>
__device__ __forceinline__ void stockham_bfly(C const *src, C *dst, int j, int nb, int Ns, int n,
                                              FastDiv fNs, int tstride, C const *tw, bool inv) {
  C v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = src[j + r * nb];
  int const jq = fdiv(j, fNs);
  int const jm = j - jq * Ns;
  if (Ns > 1) {
    int const step = tstride * jm;
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw_get(tw, r * step, inv));
  }
  dft_reg<C, R>(v, tw, n, inv);
  int const idxD = jq * Ns * R + jm;
#pragma unroll
  for (int r = 0; r < R; ++r) dst[idxD + r * Ns] = v[r];
}

/** stockham_bfly for lines whose consecutive elements lie `es` apart */
template <typename C, int R>
__device__ __forceinline__ void stockham_bfly_es(C const *src, C *dst, int j, int nb, int Ns, int n,
                                                 FastDiv fNs, int tstride, C const *tw, bool inv,
                                                 int es) {
  C v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = src[(j + r * nb) * es];
  int const jq = fdiv(j, fNs);
  int const jm = j - jq * Ns;
  if (Ns > 1) {
    int const step = tstride * jm;
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw_get(tw, r * step, inv));
  }
  dft_reg<C, R>(v, tw, n, inv);
  int const idxD = jq * Ns * R + jm;
#pragma unroll
  for (int r = 0; r < R; ++r) dst[(idxD + r * Ns) * es] = v[r];
}

/** Transform the `nlines` columns of an LDS plane: column l holds its
 *  plan.n elements at l + m * es (m = 0 .. n-1), i.e. the first index of a
 *  [n][es] layout, so a plane FFTed along its second index is transformed
 *  along its first without a transpose.  Consecutive threads take
 *  consecutive columns (unit-stride LDS accesses).  Returns the buffer
 *  holding the result (buf0 or buf1); called by the whole block; ends with
 *  a barrier. */
template <typename C>
__device__ __forceinline__ C *lds_fft_cols(C *buf0, C *buf1, int nlines, int es,
                                           FftPlanDev const &plan, C const *tw, bool inv) {
  int const n = __builtin_amdgcn_readfirstlane(plan.n);
  int const nstages = __builtin_amdgcn_readfirstlane(plan.nstages);
  FastDiv const fnl = make_fastdiv(nlines);
  C *in = buf0;
  C *out = buf1;
  int Ns = 1;
  for (int s = 0; s < nstages; ++s) {
    int const R = __builtin_amdgcn_readfirstlane(plan.radix[s]);
    int const nb = __builtin_amdgcn_readfirstlane(n / R);
    int const tstride = __builtin_amdgcn_readfirstlane(nb / Ns); // n / (Ns R)
    FastDiv const fNs = make_fastdiv(Ns);
    if (R <= 8) {
      int const items = nlines * nb;
      for (int it = threadIdx.x; it < items; it += blockDim.x) {
        int const j = fdiv(it, fnl);
        int const l = it - j * nlines;
        C const *src = in + l;
        C *dst = out + l;
        switch (R) {
        case 2: stockham_bfly_es<C, 2>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv, es); break;
        case 3: stockham_bfly_es<C, 3>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv, es); break;
        case 4: stockham_bfly_es<C, 4>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv, es); break;
        case 5: stockham_bfly_es<C, 5>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv, es); break;
        case 6: stockham_bfly_es<C, 6>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv, es); break;
        case 7: stockham_bfly_es<C, 7>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv, es); break;
        default: stockham_bfly_es<C, 8>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv, es); break;
        }
      }
    } else if (Ns == 1) {
      // generic radix as the first stage, as in lds_fft
      int const half = (R + 1) >> 1;
      int const items = nlines * nb * half;
      for (int it = threadIdx.x; it < items; it += blockDim.x) {
        int const jk = fdiv(it, fnl); // (j, kp) pair
        int const l = it - jk * nlines;
        int const j = jk / half;
        int const kp = jk - j * half;
        C const *src = in + l;
        C *dst = out + l;
        C const v0 = src[j * es];
        if (kp == 0) {
          C sum = v0;
          for (int r = 1; r < R; ++r) sum = cadd(sum, src[(j + r * nb) * es]);
          dst[j * R * es] = sum;
        } else {
          C sa = v0, sb = C{0, 0};
          int idx = 0;
          for (int mm = 1; mm < half; ++mm) {
            idx += kp; // (mm kp) mod R
            if (idx >= R) idx -= R;
            C const w = tw[idx * nb]; // exp(-2 pi i idx / R) = cos - i sin
            C const u = src[(j + mm * nb) * es], z = src[(j + (R - mm) * nb) * es];
            sa = caxpy(sa, w.x, cadd(u, z));
            sb = caxpy(sb, -w.y, csub(u, z));
          }
          C const isb = cmul_mi(sb, inv);
          dst[(j * R + kp) * es] = cadd(sa, isb);
          dst[(j * R + R - kp) * es] = csub(sa, isb);
        }
      }
    } else {
      // generic radix: one output element per item, as in lds_fft
      int const items = nlines * n;
      for (int it = threadIdx.x; it < items; it += blockDim.x) {
        int const e = fdiv(it, fnl);
        int const l = it - e * nlines;
        int const j = e / R; // butterfly
        int const k = e - j * R;
        int const jq = fdiv(j, fNs);
        int const jm = j - jq * Ns;
        C const *src = in + l;
        C *dst = out + l;
        int const base = tstride * (jm + k * Ns); // < n
        C sum = src[j * es];
        int m = 0;
        for (int r = 1; r < R; ++r) {
          m += base; // (r * base) mod n
          if (m >= n) m -= n;
          sum = cadd(sum, cmul(src[(j + r * nb) * es], tw_get(tw, m, inv)));
        }
        dst[(jq * Ns * R + jm + k * Ns) * es] = sum;
      }
    }
    __syncthreads();
    C *t = in;
    in = out;
    out = t;
    Ns *= R;
  }
  return in;
}

/** Transform `nlines` lines of length plan.n held in LDS (line l starts at
 *  l * ls).  Returns the buffer holding the result (buf0 or buf1).  Must be
 *  called by the whole block; ends with a barrier. */
template <typename C>
__device__ __forceinline__ C *lds_fft(C *buf0, C *buf1, int nlines, int ls, FftPlanDev const &plan,
                                      C const *tw,
                      bool inv) {
  // plan fields are block-uniform: keep them in scalar registers so that the
  // radix dispatch below is a scalar branch
  int const n = __builtin_amdgcn_readfirstlane(plan.n);
  int const nstages = __builtin_amdgcn_readfirstlane(plan.nstages);
  C *in = buf0;
  C *out = buf1;
  int Ns = 1;
  for (int s = 0; s < nstages; ++s) {
    int const R = __builtin_amdgcn_readfirstlane(plan.radix[s]);
    int const nb = __builtin_amdgcn_readfirstlane(n / R);
    int const tstride = __builtin_amdgcn_readfirstlane(nb / Ns); // n / (Ns R)
    FastDiv const fNs = make_fastdiv(Ns);
    if (R <= 8) {
      FastDiv const fnb = make_fastdiv(nb);
      int const items = nlines * nb;
      for (int it = threadIdx.x; it < items; it += blockDim.x) {
        int const l = fdiv(it, fnb);
        int const j = it - l * nb;
        C const *src = in + l * ls;
        C *dst = out + l * ls;
        switch (R) {
        case 2: stockham_bfly<C, 2>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv); break;
        case 3: stockham_bfly<C, 3>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv); break;
        case 4: stockham_bfly<C, 4>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv); break;
        case 5: stockham_bfly<C, 5>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv); break;
        case 6: stockham_bfly<C, 6>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv); break;
        case 7: stockham_bfly<C, 7>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv); break;
        default: stockham_bfly<C, 8>(src, dst, j, nb, Ns, n, fNs, tstride, tw, inv); break;
        }
      }
    } else if (Ns == 1) {
      // generic radix as the first stage (no inter-stage twiddles): item kp
      // of butterfly j gives X_kp and X_(R-kp) from a_m = v_m + v_(R-m),
      // b_m = v_m - v_(R-m): X_(kp, R-kp) = v_0 + sum_m cos(2 pi m kp / R) a_m
      // -+ i sum_m sin(2 pi m kp / R) b_m; kp = 0 gives X_0 = sum_r v_r
      int const half = (R + 1) >> 1;
      FastDiv const fh = make_fastdiv(half), fnb = make_fastdiv(nb);
      int const items = nlines * nb * half;
      for (int it = threadIdx.x; it < items; it += blockDim.x) {
        int const lj = fdiv(it, fh);
        int const kp = it - lj * half;
        int const l = fdiv(lj, fnb);
        int const j = lj - l * nb;
        C const *src = in + l * ls;
        C *dst = out + l * ls;
        C const v0 = src[j];
        if (kp == 0) {
          C s = v0;
          for (int r = 1; r < R; ++r) s = cadd(s, src[j + r * nb]);
          dst[j * R] = s;
        } else {
          C sa = v0, sb = C{0, 0};
          int idx = 0;
          for (int mm = 1; mm < half; ++mm) {
            idx += kp; // (mm kp) mod R
            if (idx >= R) idx -= R;
            C const w = tw[idx * nb]; // exp(-2 pi i idx / R) = cos - i sin
            C const u = src[j + mm * nb], z = src[j + (R - mm) * nb];
            sa = caxpy(sa, w.x, cadd(u, z));
            sb = caxpy(sb, -w.y, csub(u, z));
          }
          C const isb = cmul_mi(sb, inv);
          dst[j * R + kp] = cadd(sa, isb);
          dst[j * R + R - kp] = csub(sa, isb);
        }
      }
    } else {
      // generic radix: one output element per item
      FastDiv const fn = make_fastdiv(n), fR = make_fastdiv(R);
      int const items = nlines * n;
      for (int it = threadIdx.x; it < items; it += blockDim.x) {
        int const l = fdiv(it, fn);
        int const e = it - l * n;
        int const j = fdiv(e, fR); // butterfly
        int const k = e - j * R;
        int const jq = fdiv(j, fNs);
        int const jm = j - jq * Ns;
        C const *src = in + l * ls;
        C *dst = out + l * ls;
        int const base = tstride * (jm + k * Ns); // < n
        C sum = src[j];
        int m = 0;
        for (int r = 1; r < R; ++r) {
          m += base; // (r * base) mod n
          if (m >= n) m -= n;
          sum = cadd(sum, cmul(src[j + r * nb], tw_get(tw, m, inv)));
        }
        dst[jq * Ns * R + jm + k * Ns] = sum;
      }
    }
    __syncthreads();
    C *t = in;
    in = out;
    out = t;
    Ns *= R;
  }
  return in;
}

} // namespace espresso_amd::dev
