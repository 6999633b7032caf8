/*
 * p3m_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Plain-C FP64 restatement of ESPResSo's CoulombP3MImpl<double, Arch::CPU>
 * on a single MPI rank, plus the closed-form error estimates, the
 * deterministic part of the tuning algorithm, the real-space pair terms and
 * an independent FP64 Ewald sum.  Only tests/, __graft_entry__.smoke() and
 * the CPU-baseline leg of bench.py may load this library, and only to check
 * or time the reference algorithm.
 *
 * Citations are relative to the ESPResSo source tree (version 4.3-dev):
 *   p3m.cpp               = src/core/electrostatics/p3m.cpp
 *   p3m.hpp               = src/core/electrostatics/p3m.hpp
 *   common.hpp/.cpp       = src/core/p3m/common.hpp / common.cpp
 *   interpolation.hpp     = src/core/p3m/interpolation.hpp
 *   influence_function.hpp= src/core/p3m/influence_function.hpp
 *   math.hpp              = src/core/p3m/math.hpp
 *   send_mesh.cpp         = src/core/p3m/send_mesh.cpp
 *   fft.cpp               = src/core/fft/fft.cpp
 *   bspline.hpp           = src/utils/include/utils/math/bspline.hpp
 *   AS_erfc_part.hpp      = src/utils/include/utils/math/AS_erfc_part.hpp
 *   TuningAlgorithm.cpp   = src/core/p3m/TuningAlgorithm.cpp
 *
 * Layout note: the reference leaves the k-space mesh in the permuted YZX
 * order of its pencil FFT (FFTBackendLegacy.hpp:75-80).  The permutation only
 * changes the storage order, so this restatement keeps k-space in natural
 * [kx][ky][kz] order; sums over k therefore run in a different order (the
 * results differ only by FP64 rounding).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define P3MO_PI 3.14159265358979323846264338327950288
#define P3MO_INV_SQRTPI 0.564189583547756286948079451560772586
#define P3MO_SQRT2 1.41421356237309504880168872420969808
#define ROUND_ERROR_PREC 1.0e-14 /* config.hpp:64-67 */
#define P3M_MESHOFF 0.5          /* config.hpp:45-47 */

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */
static double sqr(double x) { return x * x; }

/* Threads of the solve (charge assignment, FFT lines, k-space loop, force
 * interpolation).  1 (the default) keeps the serial loops that the parity
 * checks use; more threads are for timing the CPU baseline: the charge mesh
 * is then summed from per-thread meshes, which changes the order of the
 * floating-point additions and nothing else. */
static int g_threads = 1;
void p3mo_set_threads(int t) { g_threads = t < 1 ? 1 : t; }
int p3mo_get_threads(void) { return g_threads; }

/* math::sinc, math.hpp:53-81: sin(pi x)/(pi x), 8th-order Taylor series for
 * |x| <= 0.1 (coefficients 1/1!, -1/3!, 1/5!, -1/7!, 1/9!). */
double p3mo_sinc(double x) {
  double const pix = P3MO_PI * x;
  if (fabs(x) > 0.1) return sin(pix) / pix;
  double const c0 = 1.0, c2 = -1.0 / 6.0, c4 = 1.0 / 120.0,
               c6 = -1.0 / 5040.0, c8 = 1.0 / 362880.0;
  double const pix2 = pix * pix;
  return c0 + pix2 * (c2 + pix2 * (c4 + pix2 * (c6 + pix2 * c8)));
}

/* math::analytic_cotangent_sum<cao>, math.hpp:92-126: closed form of
 * sum_m sinc^(2 cao)((n + m M)/M) as an even trigonometric polynomial in
 * c = cos^2(pi n / M) (Hockney & Eastwood eq. 7-66). */
double p3mo_cotangent_sum(int cao, int n, double mesh_i) {
  double const theta = (double)n * mesh_i * P3MO_PI;
  double const c = sqr(cos(theta));
  switch (cao) {
  case 1: return 1.;
  case 2: return (1. + c * 2.) / 3.;
  case 3: return (2. + c * (11. + c * 2.)) / 15.;
  case 4: return (17. + c * (180. + c * (114. + c * 4.))) / 315.;
  case 5: return (62. + c * (1072. + c * (1452. + c * (247. + c * 2.)))) / 2835.;
  case 6:
    return (1382. + c * (35396. + c * (83021. + c * (34096. + c * (2026. + c * 4.))))) /
           155925.;
  default:
    return (21844. +
            c * (776661. +
                 c * (2801040. + c * (2123860. + c * (349500. + c * (8166. + c * 4.)))))) /
           6081075.;
  }
}

/* Utils::AS_erfc_part, AS_erfc_part.hpp:29-40: exp(d^2) erfc(d) via
 * Abramowitz & Stegun 7.1.26. */
double p3mo_as_erfc_part(double d) {
  double const a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741,
               a4 = -1.453152027, a5 = 1.061405429, p = 0.3275911;
  double const t = 1.0 / (1.0 + p * d);
  return t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5))));
}

/* ------------------------------------------------------------------------ */
/* B-spline charge assignment weights, bspline.hpp:30-177.                  */
/* The reference tabulates, for order p and stencil point i, the polynomial */
/* w_i(x) = M_p(i + 1/2 - x) with x in [-1/2, 1/2] and M_p the cardinal     */
/* B-spline of order p, written as (integer-coefficient polynomial)/D and   */
/* evaluated by Horner's rule.  This restatement derives the same integer   */
/* coefficients at start-up from the explicit piecewise form                */
/*   M_p(u) = 1/(p-1)! sum_{j<=floor(u)} (-1)^j C(p,j) (u-j)^(p-1)          */
/* reduced by their common gcd, and evaluates them the same way.            */
/* ------------------------------------------------------------------------ */
static int64_t bs_num[8][7][7]; /* [order][i][power] */
static int64_t bs_den[8][7];
static int bs_ready = 0;

static int64_t binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}
static int64_t ipow(int64_t b, int e) {
  int64_t r = 1;
  while (e-- > 0) r *= b;
  return r;
}
static int64_t gcd64(int64_t a, int64_t b) {
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b) {
    int64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

static void bspline_init(void) {
  if (bs_ready) return;
  for (int p = 1; p <= 7; ++p) {
    int64_t fact = 1;
    for (int k = 2; k <= p - 1; ++k) fact *= k;
    for (int i = 0; i < p; ++i) {
      /* (u - j) with u = i + 1/2 - x  ->  ((2(i-j)+1) - 2x)/2 */
      int64_t c[7] = {0};
      for (int j = 0; j <= i; ++j) {
        int64_t const sgn = (j % 2) ? -1 : 1;
        int64_t const bj = binom(p, j);
        int64_t const a = 2 * (i - j) + 1;
        for (int m = 0; m <= p - 1; ++m)
          c[m] += sgn * bj * binom(p - 1, m) * ipow(a, p - 1 - m) * ipow(-2, m);
      }
      int64_t den = ipow(2, p - 1) * fact;
      int64_t g = den;
      for (int m = 0; m <= p - 1; ++m) g = gcd64(g, c[m]);
      for (int m = 0; m <= p - 1; ++m) bs_num[p][i][m] = c[m] / g;
      bs_den[p][i] = den / g;
    }
  }
  bs_ready = 1;
}

double p3mo_bspline(int order, int i, double x) {
  bspline_init();
  if (order < 1 || order > 7 || i < 0 || i >= order) return NAN;
  double acc = (double)bs_num[order][i][order - 1];
  for (int m = order - 2; m >= 0; --m) acc = (double)bs_num[order][i][m] + x * acc;
  return acc / (double)bs_den[order][i];
}

/* integer numerators/denominators, exported for the tests */
void p3mo_bspline_coeffs(int order, int i, int64_t *num, int64_t *den) {
  bspline_init();
  for (int m = 0; m < order; ++m) num[m] = bs_num[order][i][m];
  *den = bs_den[order][i];
}

/* detail::calc_meshift, common.hpp:248-264. */
void p3mo_meshift(int n, int zero_out_midpoint, int *out) {
  for (int j = 0; j < n; ++j) out[j] = 0;
  for (int j = 1; j <= n / 2; ++j) {
    out[j] = j;
    out[n - j] = -j;
  }
  if (zero_out_midpoint) out[n / 2] = 0;
}

/* ------------------------------------------------------------------------ */
/* complex FFT (any length), the unnormalised DFT FFTW computes at          */
/* fft.cpp:732-769 (forward sign -1, backward sign +1).                     */
/* Recursive mixed-radix decimation in time; prime lengths by direct DFT.   */
/* ------------------------------------------------------------------------ */
typedef struct {
  double re, im;
} cplx;

static void dft_rec(cplx *x, cplx *tmp, int n, int stride_in, cplx const *tw,
                    int tw_stride, cplx *out) {
  /* out[k] = sum_j x[j*stride_in] w^(jk), w = tw[tw_stride] (root of order n) */
  if (n == 1) {
    out[0] = x[0];
    return;
  }
  int p = 0;
  for (int f = 2; f * f <= n; ++f)
    if (n % f == 0) {
      p = f;
      break;
    }
  if (p == 0) { /* prime: direct DFT */
    int const N = n * tw_stride; /* length of the root table */
    for (int k = 0; k < n; ++k) {
      double sr = 0., si = 0.;
      for (int j = 0; j < n; ++j) {
        int const e = (int)(((int64_t)j * k % n) * tw_stride % N);
        cplx const a = x[j * stride_in];
        sr += a.re * tw[e].re - a.im * tw[e].im;
        si += a.re * tw[e].im + a.im * tw[e].re;
      }
      tmp[k].re = sr;
      tmp[k].im = si;
    }
    memcpy(out, tmp, sizeof(cplx) * (size_t)n);
    return;
  }
  int const m = n / p;
  /* p sub-transforms of length m on the decimated sequences */
  for (int r = 0; r < p; ++r)
    dft_rec(x + r * stride_in, tmp + r * m, m, stride_in * p, tw, tw_stride * p,
            out + r * m);
  /* combine: X[k + q m] = sum_r w_n^(r (k + q m)) Y_r[k] */
  cplx *y = tmp;
  memcpy(y, out, sizeof(cplx) * (size_t)n);
  int const N = n * tw_stride;
  for (int k = 0; k < m; ++k)
    for (int q = 0; q < p; ++q) {
      int const kk = k + q * m;
      double sr = 0., si = 0.;
      for (int r = 0; r < p; ++r) {
        int const e = (int)(((int64_t)r * kk % n) * tw_stride % N);
        cplx const a = y[r * m + k];
        sr += a.re * tw[e].re - a.im * tw[e].im;
        si += a.re * tw[e].im + a.im * tw[e].re;
      }
      out[kk].re = sr;
      out[kk].im = si;
    }
}

static void fft1d_strided(cplx *data, int n, int stride, int sign, cplx *line,
                          cplx *tmp, cplx *res, cplx const *tw) {
  (void)sign;
  for (int j = 0; j < n; ++j) line[j] = data[(size_t)j * stride];
  dft_rec(line, tmp, n, 1, tw, 1, res);
  for (int j = 0; j < n; ++j) data[(size_t)j * stride] = res[j];
}

/* 3D unnormalised c2c DFT of an interleaved complex [n0][n1][n2] array. */
void p3mo_fft3d(int n0, int n1, int n2, double *data_interleaved, int sign) {
  cplx *d = (cplx *)data_interleaved;
  int const n[3] = {n0, n1, n2};
  int nmax = n0 > n1 ? n0 : n1;
  if (n2 > nmax) nmax = n2;
  cplx *line = malloc(sizeof(cplx) * (size_t)nmax);
  cplx *tmp = malloc(sizeof(cplx) * (size_t)nmax);
  cplx *res = malloc(sizeof(cplx) * (size_t)nmax);
  cplx *tw = malloc(sizeof(cplx) * (size_t)nmax);
  size_t const strides[3] = {(size_t)n1 * n2, (size_t)n2, 1};
  for (int ax = 0; ax < 3; ++ax) {
    int const len = n[ax];
    for (int j = 0; j < len; ++j) {
      double const ang = sign * 2. * P3MO_PI * (double)j / (double)len;
      tw[j].re = cos(ang);
      tw[j].im = sin(ang);
    }
    int const o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
    if (g_threads > 1) {
      int const nlines = n[o1] * n[o2];
#pragma omp parallel num_threads(g_threads)
      {
        cplx *tl = malloc(sizeof(cplx) * 3 * (size_t)nmax);
#pragma omp for schedule(static)
        for (int ab = 0; ab < nlines; ++ab) {
          int const a = ab / n[o2], b = ab % n[o2];
          size_t const base = (size_t)a * strides[o1] + (size_t)b * strides[o2];
          fft1d_strided(d + base, len, (int)strides[ax], sign, tl, tl + nmax, tl + 2 * nmax, tw);
        }
        free(tl);
      }
      continue;
    }
    for (int a = 0; a < n[o1]; ++a)
      for (int b = 0; b < n[o2]; ++b) {
        size_t const base = (size_t)a * strides[o1] + (size_t)b * strides[o2];
        fft1d_strided(d + base, len, (int)strides[ax], sign, line, tmp, res, tw);
      }
  }
  free(line);
  free(tmp);
  free(res);
  free(tw);
}

/* ------------------------------------------------------------------------ */
/* solver state                                                             */
/* ------------------------------------------------------------------------ */
typedef struct p3mo {
  /* P3MParameters, common.hpp:59-160 */
  int mesh[3];
  int cao;
  double alpha, r_cut, alpha_L, r_cut_iL;
  double mesh_off[3];
  double epsilon;
  double a[3], ai[3], cao_cut[3];
  /* box and actor */
  double box_l[3];
  double prefactor;
  double skin;
  /* P3MLocalMesh, common.hpp:175-214 */
  int dim[3], size, ld_ind[3], inner[3], in_ld[3], in_ur[3];
  int margin[6], q_2_off, q_21_off;
  double ld_pos[3];
  /* halo blocks, send_mesh.cpp:89-160 */
  int s_ld[6][3], s_ur[6][3], s_dim[6][3], s_size[6];
  int r_ld[6][3], r_ur[6][3], r_dim[6][3], r_size[6];
  /* meshes */
  int *d_op[3];
  double *g_force, *g_energy; /* natural [kx][ky][kz] order */
  double *rs_scalar;          /* CA mesh */
  double *rs_fields[3];       /* CA meshes */
  cplx *kmesh;                /* M^3 complex */
  cplx *kfield;               /* M^3 complex scratch */
  /* interpolation cache, interpolation.hpp:55-130 */
  int *ca_fmp;
  double *ca_frac;
  int n_cache, cap_cache;
  /* count_charged_particles, p3m.cpp:97-115 */
  int sum_qpart;
  double sum_q2, square_sum_q;
} p3mo;

/* P3MParameters::recalc_a_ai_cao_cut, common.hpp:158-163 */
static void recalc_a_ai_cao_cut(p3mo *s) {
  for (int d = 0; d < 3; ++d) {
    s->ai[d] = (double)s->mesh[d] / s->box_l[d];
    s->a[d] = 1. / s->ai[d];
    s->cao_cut[d] = ((double)s->cao / 2.) * s->a[d];
  }
}

/* P3MLocalMesh::calc_local_ca_mesh, common.cpp:34-106, single rank:
 * the local box is the whole box [0, L), no ELC space layer. */
static void calc_local_ca_mesh(p3mo *s) {
  int ind[3];
  double full_skin[3], inner_ld_gp[3], inner_ur_gp[3], outer_ld_gp[3], outer_ur_gp[3];
  for (int i = 0; i < 3; ++i) {
    full_skin[i] = (s->cao_cut[i] + s->skin) + 0.;
    double const inner_ld_pos = 0., inner_ur_pos = s->box_l[i];
    double const outer_ur_pos = inner_ur_pos + full_skin[i];
    double const outer_ld_pos = inner_ld_pos - full_skin[i];
    /* calc_grid_pos, common.hpp:169-171 */
    inner_ld_gp[i] = inner_ld_pos * s->ai[i] - s->mesh_off[i];
    inner_ur_gp[i] = inner_ur_pos * s->ai[i] - s->mesh_off[i];
    outer_ld_gp[i] = outer_ld_pos * s->ai[i] - s->mesh_off[i];
    outer_ur_gp[i] = outer_ur_pos * s->ai[i] - s->mesh_off[i];
  }
  for (int i = 0; i < 3; ++i) s->in_ld[i] = (int)ceil(inner_ld_gp[i]);
  for (int i = 0; i < 3; ++i) s->in_ur[i] = (int)floor(inner_ur_gp[i]);
  for (int i = 0; i < 3; ++i) {
    if (inner_ur_gp[i] - s->in_ur[i] < ROUND_ERROR_PREC) s->in_ur[i]--;
    if (inner_ld_gp[i] - s->in_ld[i] + 1. < ROUND_ERROR_PREC) s->in_ld[i]--;
  }
  for (int i = 0; i < 3; ++i) s->inner[i] = s->in_ur[i] - s->in_ld[i] + 1;
  for (int i = 0; i < 3; ++i) s->ld_ind[i] = (int)ceil(outer_ld_gp[i]);
  for (int i = 0; i < 3; ++i) s->margin[i * 2] = s->in_ld[i] - s->ld_ind[i];
  for (int i = 0; i < 3; ++i) ind[i] = (int)floor(outer_ur_gp[i]);
  for (int i = 0; i < 3; ++i)
    if (outer_ur_gp[i] - ind[i] == 0.) ind[i]--;
  for (int i = 0; i < 3; ++i) s->margin[i * 2 + 1] = ind[i] - s->in_ur[i];
  s->size = 1;
  for (int i = 0; i < 3; ++i) {
    s->dim[i] = ind[i] - s->ld_ind[i] + 1;
    s->size *= s->dim[i];
  }
  for (int i = 0; i < 3; ++i) s->in_ld[i] = s->margin[i * 2];
  for (int i = 0; i < 3; ++i) s->in_ur[i] = s->margin[i * 2] + s->inner[i];
  s->q_2_off = s->dim[2] - s->cao;
  s->q_21_off = s->dim[2] * (s->dim[1] - s->cao);
}

/* P3MLocalMesh::recalc_ld_pos, common.hpp:204-209 */
static void recalc_ld_pos(p3mo *s) {
  for (int i = 0; i < 3; ++i) s->ld_pos[i] = (s->ld_ind[i] + s->mesh_off[i]) * s->a[i];
}

/* p3m_send_mesh::resize, send_mesh.cpp:89-160, single rank: the neighbour in
 * every direction is the rank itself, so r_margin[j] = margin[opposite]. */
static void halo_resize(p3mo *s) {
  int done[3] = {0, 0, 0};
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) {
      s->s_ld[i * 2][j] = 0 + done[j] * s->margin[j * 2];
      if (j == i)
        s->s_ur[i * 2][j] = s->margin[j * 2];
      else
        s->s_ur[i * 2][j] = s->dim[j] - done[j] * s->margin[(j * 2) + 1];
      if (j == i)
        s->s_ld[(i * 2) + 1][j] = s->in_ur[j];
      else
        s->s_ld[(i * 2) + 1][j] = 0 + done[j] * s->margin[j * 2];
      s->s_ur[(i * 2) + 1][j] = s->dim[j] - done[j] * s->margin[(j * 2) + 1];
    }
    done[i] = 1;
  }
  for (int i = 0; i < 6; i++) {
    s->s_size[i] = 1;
    for (int j = 0; j < 3; j++) {
      s->s_dim[i][j] = s->s_ur[i][j] - s->s_ld[i][j];
      s->s_size[i] *= s->s_dim[i][j];
    }
  }
  int r_margin[6];
  for (int i = 0; i < 6; i++) {
    int const j = (i % 2 == 0) ? i + 1 : i - 1;
    r_margin[j] = s->margin[i];
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      if (j == i) {
        s->r_ld[i * 2][j] = s->s_ld[i * 2][j] + s->margin[2 * j];
        s->r_ur[i * 2][j] = s->s_ur[i * 2][j] + r_margin[2 * j];
        s->r_ld[(i * 2) + 1][j] = s->s_ld[(i * 2) + 1][j] - r_margin[(2 * j) + 1];
        s->r_ur[(i * 2) + 1][j] = s->s_ur[(i * 2) + 1][j] - s->margin[(2 * j) + 1];
      } else {
        s->r_ld[i * 2][j] = s->s_ld[i * 2][j];
        s->r_ur[i * 2][j] = s->s_ur[i * 2][j];
        s->r_ld[(i * 2) + 1][j] = s->s_ld[(i * 2) + 1][j];
        s->r_ur[(i * 2) + 1][j] = s->s_ur[(i * 2) + 1][j];
      }
    }
  for (int i = 0; i < 6; i++) {
    s->r_size[i] = 1;
    for (int j = 0; j < 3; j++) {
      s->r_dim[i][j] = s->r_ur[i][j] - s->r_ld[i][j];
      s->r_size[i] *= s->r_dim[i][j];
    }
  }
}

/* fft_pack_block / fft_unpack_block, packing.hpp:43-106 (element = 1) and
 * p3m_add_block, send_mesh.cpp:63-87. */
static void pack_block(double const *in, double *out, int const start[3], int const size[3],
                       int const dim[3]) {
  int li_out = 0;
  for (int s0 = 0; s0 < size[0]; s0++)
    for (int m = 0; m < size[1]; m++) {
      int const li_in = start[2] + dim[2] * ((start[1] + m) + dim[1] * (start[0] + s0));
      memmove(out + li_out, in + li_in, sizeof(double) * (size_t)size[2]);
      li_out += size[2];
    }
}
static void unpack_block(double const *in, double *out, int const start[3], int const size[3],
                         int const dim[3]) {
  int li_in = 0;
  for (int s0 = 0; s0 < size[0]; s0++)
    for (int m = 0; m < size[1]; m++) {
      int const li_out = start[2] + dim[2] * ((start[1] + m) + dim[1] * (start[0] + s0));
      memmove(out + li_out, in + li_in, sizeof(double) * (size_t)size[2]);
      li_in += size[2];
    }
}
static void add_block(double const *in, double *out, int const start[3], int const size[3],
                      int const dim[3]) {
  int li_in = 0;
  for (int s0 = 0; s0 < size[0]; s0++)
    for (int m = 0; m < size[1]; m++) {
      int li_out = start[2] + dim[2] * ((start[1] + m) + dim[1] * (start[0] + s0));
      for (int f = 0; f < size[2]; f++) out[li_out++] += in[li_in++];
    }
}

/* p3m_send_mesh::gather_grid, send_mesh.cpp:163-201, single rank */
static void halo_gather(p3mo *s, double *mesh) {
  int maxsz = 0;
  for (int i = 0; i < 6; ++i) {
    if (s->s_size[i] > maxsz) maxsz = s->s_size[i];
    if (s->r_size[i] > maxsz) maxsz = s->r_size[i];
  }
  double *buf = malloc(sizeof(double) * (size_t)(maxsz > 0 ? maxsz : 1));
  for (int s_dir = 0; s_dir < 6; s_dir++) {
    int const r_dir = (s_dir % 2 == 0) ? s_dir + 1 : s_dir - 1;
    if (s->s_size[s_dir] > 0) pack_block(mesh, buf, s->s_ld[s_dir], s->s_dim[s_dir], s->dim);
    if (s->r_size[r_dir] > 0) add_block(buf, mesh, s->r_ld[r_dir], s->r_dim[r_dir], s->dim);
  }
  free(buf);
}

/* p3m_send_mesh::spread_grid, send_mesh.cpp:203-240, single rank */
static void halo_spread(p3mo *s, double *mesh) {
  int maxsz = 0;
  for (int i = 0; i < 6; ++i) {
    if (s->s_size[i] > maxsz) maxsz = s->s_size[i];
    if (s->r_size[i] > maxsz) maxsz = s->r_size[i];
  }
  double *buf = malloc(sizeof(double) * (size_t)(maxsz > 0 ? maxsz : 1));
  for (int s_dir = 5; s_dir >= 0; s_dir--) {
    int const r_dir = (s_dir % 2 == 0) ? s_dir + 1 : s_dir - 1;
    if (s->r_size[r_dir] > 0) pack_block(mesh, buf, s->r_ld[r_dir], s->r_dim[r_dir], s->dim);
    if (s->s_size[s_dir] > 0) unpack_block(buf, mesh, s->s_ld[s_dir], s->s_dim[s_dir], s->dim);
  }
  free(buf);
}

/* G_opt<S, m=0>, influence_function.hpp:53-89: with no aliasing images the
 * sum has the single term km = k. */
static double g_opt(int S, int cao, double alpha, double const k[3], double const h[3]) {
  double const k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
  if (k2 == 0.) return 0.;
  double const limit = 30.;
  double const exponent_prefactor = sqr(1. / (2. * alpha));
  double numerator = 0., denominator = 0.;
  double fnm[3];
  for (int d = 0; d < 3; ++d) {
    double const wavevector = (2. * P3MO_PI) / h[d];
    double const wavevector_i = 1. / wavevector;
    fnm[d] = p3mo_sinc(k[d] * wavevector_i);
  }
  double const U2 = pow(fnm[0] * fnm[1] * fnm[2], 2 * cao);
  double const km2 = k2;
  double const exponent = exponent_prefactor * km2;
  if (exponent < limit) {
    double const f3 = exp(-exponent) * (4. * P3MO_PI / km2);
    double const kkm = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    numerator += U2 * f3 * (S == 1 ? kkm : 1.);
  }
  denominator += U2;
  return numerator / ((S == 1 ? k2 : 1.) * sqr(denominator));
}

/* grid_influence_function<double, S, 0>, influence_function.hpp:110-150,
 * evaluated in natural [kx][ky][kz] order. */
static void grid_influence_function(p3mo *s, int S, double *g) {
  int *shifts[3];
  for (int d = 0; d < 3; ++d) {
    shifts[d] = malloc(sizeof(int) * (size_t)s->mesh[d]);
    p3mo_meshift(s->mesh[d], 0, shifts[d]);
  }
  double wavevector[3];
  for (int d = 0; d < 3; ++d) wavevector[d] = (2. * P3MO_PI) * (1. / s->box_l[d]);
  int const hm[3] = {s->mesh[0] / 2, s->mesh[1] / 2, s->mesh[2] / 2};
  size_t index = 0;
  for (int x = 0; x < s->mesh[0]; ++x)
    for (int y = 0; y < s->mesh[1]; ++y)
      for (int z = 0; z < s->mesh[2]; ++z) {
        g[index] = 0.;
        if ((x % hm[0] != 0) || (y % hm[1] != 0) || (z % hm[2] != 0)) {
          double const k[3] = {shifts[0][x] * wavevector[0], shifts[1][y] * wavevector[1],
                               shifts[2][z] * wavevector[2]};
          g[index] = g_opt(S, s->cao, s->alpha, k, s->a);
        }
        ++index;
      }
  for (int d = 0; d < 3; ++d) free(shifts[d]);
}

static void free_meshes(p3mo *s) {
  for (int d = 0; d < 3; ++d) {
    free(s->d_op[d]);
    s->d_op[d] = NULL;
    free(s->rs_fields[d]);
    s->rs_fields[d] = NULL;
  }
  free(s->g_force);
  free(s->g_energy);
  free(s->rs_scalar);
  free(s->kmesh);
  free(s->kfield);
  s->g_force = s->g_energy = s->rs_scalar = NULL;
  s->kmesh = s->kfield = NULL;
}

/* init_cpu_kernels + scaleby_box_l, p3m.cpp:245-280 and 856-866, for fixed
 * (mesh, cao, alpha, r_cut). Returns NULL on invalid geometry. */
p3mo *p3mo_new(int const mesh[3], int cao, double alpha, double r_cut, double const mesh_off[3],
               double epsilon, double const box_l[3], double prefactor, double skin) {
  p3mo *s = calloc(1, sizeof(p3mo));
  for (int d = 0; d < 3; ++d) {
    s->mesh[d] = mesh[d];
    s->box_l[d] = box_l[d];
    s->mesh_off[d] = mesh_off[d];
    /* P3MParameters ctor, common.hpp:136-143: -1 selects P3M_MESHOFF */
    if (mesh_off[0] == -1. && mesh_off[1] == -1. && mesh_off[2] == -1.)
      s->mesh_off[d] = P3M_MESHOFF;
  }
  s->cao = cao;
  s->alpha = alpha;
  s->r_cut = r_cut;
  s->epsilon = epsilon;
  s->prefactor = prefactor;
  s->skin = skin;
  s->alpha_L = alpha * box_l[0];
  s->r_cut_iL = r_cut * (1. / box_l[0]);
  recalc_a_ai_cao_cut(s);
  calc_local_ca_mesh(s);
  recalc_ld_pos(s);
  for (int d = 0; d < 3; ++d)
    if (s->inner[d] != s->mesh[d]) {
      free(s);
      return NULL;
    }
  halo_resize(s);
  size_t const M3 = (size_t)mesh[0] * mesh[1] * mesh[2];
  for (int d = 0; d < 3; ++d) {
    s->d_op[d] = malloc(sizeof(int) * (size_t)mesh[d]);
    p3mo_meshift(mesh[d], 1, s->d_op[d]); /* data_struct.hpp:68-70 */
    s->rs_fields[d] = calloc((size_t)s->size, sizeof(double));
  }
  s->rs_scalar = calloc((size_t)s->size, sizeof(double));
  s->kmesh = calloc(M3, sizeof(cplx));
  s->kfield = calloc(M3, sizeof(cplx));
  s->g_force = malloc(sizeof(double) * M3);
  s->g_energy = malloc(sizeof(double) * M3);
  grid_influence_function(s, 1, s->g_force);  /* p3m.cpp:126-132 */
  grid_influence_function(s, 0, s->g_energy); /* p3m.cpp:137-143 */
  return s;
}

void p3mo_free(p3mo *s) {
  if (!s) return;
  free_meshes(s);
  free(s->ca_fmp);
  free(s->ca_frac);
  free(s);
}

/* geometry export: ints[0..] = dim[3], size, ld_ind[3], inner[3], in_ld[3],
 * in_ur[3], margin[6], q_2_off, q_21_off ; dbls = ld_pos[3], a[3], ai[3], cao_cut[3] */
void p3mo_geometry(p3mo const *s, int *ints, double *dbls) {
  int k = 0;
  for (int d = 0; d < 3; ++d) ints[k++] = s->dim[d];
  ints[k++] = s->size;
  for (int d = 0; d < 3; ++d) ints[k++] = s->ld_ind[d];
  for (int d = 0; d < 3; ++d) ints[k++] = s->inner[d];
  for (int d = 0; d < 3; ++d) ints[k++] = s->in_ld[d];
  for (int d = 0; d < 3; ++d) ints[k++] = s->in_ur[d];
  for (int d = 0; d < 6; ++d) ints[k++] = s->margin[d];
  ints[k++] = s->q_2_off;
  ints[k++] = s->q_21_off;
  k = 0;
  for (int d = 0; d < 3; ++d) dbls[k++] = s->ld_pos[d];
  for (int d = 0; d < 3; ++d) dbls[k++] = s->a[d];
  for (int d = 0; d < 3; ++d) dbls[k++] = s->ai[d];
  for (int d = 0; d < 3; ++d) dbls[k++] = s->cao_cut[d];
}

/* p3m_calculate_interpolation_weights<cao>, interpolation.hpp:138-178.
 * Writes the linear CA-mesh index, the nearest mesh point and 3*cao weights. */
int p3mo_weights(p3mo const *s, double const pos[3], int *ind, int nmp[3], double *w) {
  int const cao = s->cao;
  double const pos_shift = floor((cao - 1) / 2.0) - (cao % 2) / 2.0;
  double dist[3];
  for (int d = 0; d < 3; d++) {
    double const t = ((pos[d] - s->ld_pos[d]) * s->ai[d]) - pos_shift;
    nmp[d] = (int)t;
    dist[d] = (t - nmp[d]) - 0.5;
  }
  *ind = s->dim[1] * s->dim[2] * nmp[0] + s->dim[2] * nmp[1] + nmp[2]; /* index.hpp:52 */
  for (int d = 0; d < 3; ++d)
    if (nmp[d] < 0 || nmp[d] + cao > s->dim[d]) return -1;
  for (int i = 0; i < cao; i++) {
    w[0 * cao + i] = p3mo_bspline(cao, i, dist[0]);
    w[1 * cao + i] = p3mo_bspline(cao, i, dist[1]);
    w[2 * cao + i] = p3mo_bspline(cao, i, dist[2]);
  }
  return 0;
}

/* count_charged_particles, p3m.cpp:97-115 */
void p3mo_count_charged(p3mo *s, int n, double const *q) {
  int local_n = 0;
  double local_q2 = 0., local_q = 0.;
  for (int i = 0; i < n; ++i)
    if (q[i] != 0.0) {
      local_n++;
      local_q2 += sqr(q[i]);
      local_q += q[i];
    }
  s->sum_qpart = local_n;
  s->sum_q2 = local_q2;
  s->square_sum_q = sqr(local_q);
}

/* prepare_fft_mesh + charge_assign (AssignCharge<cao>), p3m.impl.hpp:161-168,
 * p3m.cpp:283-330 and the p3m_interpolate stencil walk, interpolation.hpp:191-208 */
static int charge_assign_threaded(p3mo *s, int n, double const *pos, double const *q);

static int charge_assign(p3mo *s, int n, double const *pos, double const *q) {
  int const cao = s->cao;
  for (int i = 0; i < s->size; i++) s->rs_scalar[i] = 0.;
  s->n_cache = 0;
  if (s->cap_cache < n) {
    s->cap_cache = n;
    s->ca_fmp = realloc(s->ca_fmp, sizeof(int) * (size_t)n);
    s->ca_frac = realloc(s->ca_frac, sizeof(double) * 3 * (size_t)cao * (size_t)n);
  }
  if (g_threads > 1) return charge_assign_threaded(s, n, pos, q);
  for (int p = 0; p < n; ++p) {
    if (q[p] == 0.0) continue;
    int ind, nmp[3];
    double *w = s->ca_frac + 3 * (size_t)cao * (size_t)s->n_cache;
    if (p3mo_weights(s, pos + 3 * (size_t)p, &ind, nmp, w) != 0) return -1;
    s->ca_fmp[s->n_cache++] = ind;
    int q_ind = ind;
    for (int i0 = 0; i0 < cao; i0++) {
      double const tmp0 = w[i0];
      for (int i1 = 0; i1 < cao; i1++) {
        double const tmp1 = tmp0 * w[cao + i1];
        for (int i2 = 0; i2 < cao; i2++) {
          s->rs_scalar[q_ind] += (tmp1 * w[2 * cao + i2]) * q[p];
          q_ind++;
        }
        q_ind += s->q_2_off;
      }
      q_ind += s->q_21_off;
    }
  }
  return 0;
}

/* charge_assign on g_threads threads: the cache slot of each charged
 * particle from a serial prefix count, weights in parallel, each thread
 * spreading its share of the particles into a private CA mesh, then the
 * private meshes summed point by point. */
static int charge_assign_threaded(p3mo *s, int n, double const *pos, double const *q) {
  int const cao = s->cao;
  int *slot = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int nc = 0;
  for (int p = 0; p < n; ++p) slot[p] = (q[p] != 0.0) ? nc++ : -1;
  s->n_cache = nc;
  int const T = g_threads;
  size_t const size = (size_t)s->size;
  double *priv = calloc((size_t)T * size, sizeof(double));
  int bad = 0;
#pragma omp parallel num_threads(T) reduction(| : bad)
  {
    double *mesh = priv + (size_t)omp_get_thread_num() * size;
#pragma omp for schedule(static)
    for (int p = 0; p < n; ++p) {
      if (slot[p] < 0) continue;
      int ind, nmp[3];
      double *w = s->ca_frac + 3 * (size_t)cao * (size_t)slot[p];
      if (p3mo_weights(s, pos + 3 * (size_t)p, &ind, nmp, w) != 0) {
        bad = 1;
        continue;
      }
      s->ca_fmp[slot[p]] = ind;
      int q_ind = ind;
      for (int i0 = 0; i0 < cao; i0++) {
        double const tmp0 = w[i0];
        for (int i1 = 0; i1 < cao; i1++) {
          double const tmp1 = tmp0 * w[cao + i1];
          for (int i2 = 0; i2 < cao; i2++) {
            mesh[q_ind] += (tmp1 * w[2 * cao + i2]) * q[p];
            q_ind++;
          }
          q_ind += s->q_2_off;
        }
        q_ind += s->q_21_off;
      }
    }
#pragma omp for schedule(static)
    for (size_t i = 0; i < size; ++i) {
      double acc = 0.;
      for (int t = 0; t < T; ++t) acc += priv[(size_t)t * size + i];
      s->rs_scalar[i] = acc;
    }
  }
  free(priv);
  free(slot);
  return bad ? -1 : 0;
}

/* inner block of the CA mesh -> complex M^3 array and forward DFT
 * (FFTBackendLegacy::forward_fft -> fft.cpp:715-745) */
static void forward_fft(p3mo *s) {
  int const *M = s->mesh;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int x = 0; x < M[0]; ++x) {
    size_t idx = (size_t)x * M[1] * M[2];
    for (int y = 0; y < M[1]; ++y)
      for (int z = 0; z < M[2]; ++z) {
        int const li = (s->in_ld[2] + z) + s->dim[2] * ((s->in_ld[1] + y) + s->dim[1] * (s->in_ld[0] + x));
        s->kmesh[idx].re = s->rs_scalar[li];
        s->kmesh[idx].im = 0.;
        ++idx;
      }
  }
  p3mo_fft3d(M[0], M[1], M[2], (double *)s->kmesh, -1);
}

/* backward DFT of kfield, real part into the inner block of a CA mesh
 * (fft.cpp:747-785) */
static void backward_fft(p3mo *s, double *rs) {
  int const *M = s->mesh;
  p3mo_fft3d(M[0], M[1], M[2], (double *)s->kfield, +1);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int x = 0; x < M[0]; ++x) {
    size_t idx = (size_t)x * M[1] * M[2];
    for (int y = 0; y < M[1]; ++y)
      for (int z = 0; z < M[2]; ++z) {
        int const li = (s->in_ld[2] + z) + s->dim[2] * ((s->in_ld[1] + y) + s->dim[1] * (s->in_ld[0] + x));
        rs[li] = s->kfield[idx].re;
        ++idx;
      }
  }
}

static void calc_dipole(int n, double const *q, double const *upos, double dip[3]) {
  dip[0] = dip[1] = dip[2] = 0.;
  for (int p = 0; p < n; ++p)
    for (int d = 0; d < 3; ++d) dip[d] = dip[d] + q[p] * upos[3 * (size_t)p + d];
}

/* long_range_kernel, p3m.cpp:452-592: force and/or energy.  `force` is
 * accumulated in place (may be NULL), the energy is returned. `upos` are the
 * unfolded positions used by the dipole term (NULL: use pos). */
static double long_range_kernel(p3mo *s, int force_flag, int energy_flag, int n, double const *pos,
                                double const *q, double const *upos, double *force) {
  int const *M = s->mesh;
  size_t const M3 = (size_t)M[0] * M[1] * M[2];
  p3mo_count_charged(s, n, q);
  /* no charges: an empty mesh (the reference clears it in charge_assign /
   * prepare_fft_mesh before the solve, p3m.impl.hpp:161-168), so forces and
   * energy are zero */
  if (!(s->sum_q2 > 0.)) return 0.;
  if (charge_assign(s, n, pos, q) != 0) return NAN;
  halo_gather(s, s->rs_scalar);
  forward_fft(s);
  double const volume = s->box_l[0] * s->box_l[1] * s->box_l[2];
  int const has_dipole = (s->epsilon != 0.);
  double box_dipole[3] = {0., 0., 0.};
  if (has_dipole) calc_dipole(n, q, upos ? upos : pos, box_dipole);
  double const pref = 4. * P3MO_PI / volume / (2. * s->epsilon + 1.);

  if (force_flag) {
    double wavevector[3];
    for (int d = 0; d < 3; ++d) wavevector[d] = (2. * P3MO_PI) * (1. / s->box_l[d]);
    for (int d_rs = 0; d_rs < 3; ++d_rs) {
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
      for (int x = 0; x < M[0]; ++x) {
        size_t idx = (size_t)x * M[1] * M[2];
        for (int y = 0; y < M[1]; ++y)
          for (int z = 0; z < M[2]; ++z) {
            int const nn[3] = {x, y, z};
            double const phr = s->g_force[idx] * s->kmesh[idx].re;
            double const phi = s->g_force[idx] * s->kmesh[idx].im;
            double const k = (double)s->d_op[d_rs][nn[d_rs]] * wavevector[d_rs];
            s->kfield[idx].re = -k * phi;
            s->kfield[idx].im = +k * phr;
            ++idx;
          }
      }
      backward_fft(s, s->rs_fields[d_rs]);
      (void)M3;
    }
    for (int d = 0; d < 3; ++d) halo_spread(s, s->rs_fields[d]);

    /* AssignForces<cao>, p3m.cpp:344-373 */
    double const force_prefac = s->prefactor / volume;
    int const cao = s->cao;
    /* cache slot of each charged particle: its rank among the charged ones */
    int *slot = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int nc = 0;
    for (int p = 0; p < n; ++p) slot[p] = (q[p] != 0.0) ? nc++ : -1;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int p = 0; p < n; ++p) {
      if (q[p] == 0.0) continue;
      int const p_index = slot[p];
      double const pf = q[p] * force_prefac;
      double const *w = s->ca_frac + 3 * (size_t)cao * (size_t)p_index;
      int q_ind = s->ca_fmp[p_index];
      double f[3] = {0., 0., 0.};
      for (int i0 = 0; i0 < cao; i0++) {
        double const tmp0 = w[i0];
        for (int i1 = 0; i1 < cao; i1++) {
          double const tmp1 = tmp0 * w[cao + i1];
          for (int i2 = 0; i2 < cao; i2++) {
            double const ww = tmp1 * w[2 * cao + i2];
            f[0] += ww * s->rs_fields[0][q_ind];
            f[1] += ww * s->rs_fields[1][q_ind];
            f[2] += ww * s->rs_fields[2][q_ind];
            q_ind++;
          }
          q_ind += s->q_2_off;
        }
        q_ind += s->q_21_off;
      }
      if (force)
        for (int d = 0; d < 3; ++d) force[3 * (size_t)p + d] -= pf * f[d];
    }
    free(slot);
    /* dipole correction, p3m.cpp:538-547 */
    if (has_dipole && force) {
      double dm[3];
      for (int d = 0; d < 3; ++d) dm[d] = s->prefactor * pref * box_dipole[d];
      for (int p = 0; p < n; ++p)
        for (int d = 0; d < 3; ++d) force[3 * (size_t)p + d] -= q[p] * dm[d];
    }
  }

  if (energy_flag) {
    double node_energy = 0.;
    if (g_threads > 1) {
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : node_energy)
      for (size_t i = 0; i < M3; i++)
        node_energy += s->g_energy[i] * (sqr(s->kmesh[i].re) + sqr(s->kmesh[i].im));
    } else {
      for (size_t i = 0; i < M3; i++)
        node_energy += s->g_energy[i] * (sqr(s->kmesh[i].re) + sqr(s->kmesh[i].im));
    }
    node_energy /= 2. * volume;
    double energy = node_energy;
    energy -= s->sum_q2 * s->alpha * P3MO_INV_SQRTPI;
    energy -= s->square_sum_q * P3MO_PI / (2. * volume * sqr(s->alpha));
    if (has_dipole)
      energy += pref * (sqr(box_dipole[0]) + sqr(box_dipole[1]) + sqr(box_dipole[2]));
    return energy * s->prefactor;
  }
  return 0.;
}

/* add_long_range_forces (CPU), p3m.impl.hpp:145-154 */
int p3mo_add_long_range_forces(p3mo *s, int n, double const *pos, double const *q,
                               double const *upos, double *force) {
  double const e = long_range_kernel(s, 1, 0, n, pos, q, upos, force);
  return isnan(e) ? -1 : 0;
}

/* long_range_energy, p3m.impl.hpp:141-143 */
double p3mo_long_range_energy(p3mo *s, int n, double const *pos, double const *q,
                              double const *upos) {
  return long_range_kernel(s, 0, 1, n, pos, q, upos, NULL);
}

/* long_range_pressure, p3m.cpp:393-449 (row-major 3x3 tensor) */
int p3mo_long_range_pressure(p3mo *s, int n, double const *pos, double const *q, double *out9) {
  int const *M = s->mesh;
  for (int i = 0; i < 9; ++i) out9[i] = 0.;
  p3mo_count_charged(s, n, q);
  double const volume = s->box_l[0] * s->box_l[1] * s->box_l[2];
  if (s->sum_q2 > 0.) {
    if (charge_assign(s, n, pos, q) != 0) return -1;
    halo_gather(s, s->rs_scalar);
    forward_fft(s);
    double const half_alpha_inv_sq = sqr(1. / 2. / s->alpha);
    double wavevector[3];
    for (int d = 0; d < 3; ++d) wavevector[d] = (2. * P3MO_PI) * (1. / s->box_l[d]);
    double diagonal = 0.;
    size_t index = 0;
    for (int x = 0; x < M[0]; ++x)
      for (int y = 0; y < M[1]; ++y)
        for (int z = 0; z < M[2]; ++z) {
          double const kx = s->d_op[0][x] * wavevector[0];
          double const ky = s->d_op[1][y] * wavevector[1];
          double const kz = s->d_op[2][z] * wavevector[2];
          double const norm_sq = sqr(kx) + sqr(ky) + sqr(kz);
          if (norm_sq != 0.) {
            double const e = s->g_energy[index] * (sqr(s->kmesh[index].re) + sqr(s->kmesh[index].im));
            double const vterm = -2. * (1. / norm_sq + half_alpha_inv_sq);
            double const pref = e * vterm;
            out9[0] += pref * kx * kx;
            out9[1] += pref * kx * ky;
            out9[2] += pref * kx * kz;
            out9[4] += pref * ky * ky;
            out9[5] += pref * ky * kz;
            out9[8] += pref * kz * kz;
            diagonal += e;
          }
          ++index;
        }
    out9[0] += diagonal;
    out9[4] += diagonal;
    out9[8] += diagonal;
    out9[3] = out9[1];
    out9[6] = out9[2];
    out9[7] = out9[5];
  }
  for (int i = 0; i < 9; ++i) out9[i] = out9[i] * s->prefactor / (2. * volume);
  return 0;
}

/* Stage exports for stage-by-stage parity of the GPU kernels:
 * charge mesh after the halo wrap (inner M^3 block, natural order), the
 * forward transform, and the three field meshes. */
int p3mo_charge_mesh(p3mo *s, int n, double const *pos, double const *q, double *out) {
  if (charge_assign(s, n, pos, q) != 0) return -1;
  halo_gather(s, s->rs_scalar);
  int const *M = s->mesh;
  size_t idx = 0;
  for (int x = 0; x < M[0]; ++x)
    for (int y = 0; y < M[1]; ++y)
      for (int z = 0; z < M[2]; ++z)
        out[idx++] = s->rs_scalar[(s->in_ld[2] + z) +
                                  s->dim[2] * ((s->in_ld[1] + y) + s->dim[1] * (s->in_ld[0] + x))];
  return 0;
}

int p3mo_fields(p3mo *s, int n, double const *pos, double const *q, double *ex, double *ey,
                double *ez) {
  double *tmp = calloc(3 * (size_t)n, sizeof(double));
  int const rc = p3mo_add_long_range_forces(s, n, pos, q, NULL, tmp);
  free(tmp);
  if (rc) return rc;
  double *outs[3] = {ex, ey, ez};
  int const *M = s->mesh;
  for (int d = 0; d < 3; ++d) {
    size_t idx = 0;
    for (int x = 0; x < M[0]; ++x)
      for (int y = 0; y < M[1]; ++y)
        for (int z = 0; z < M[2]; ++z)
          outs[d][idx++] = s->rs_fields[d][(s->in_ld[2] + z) +
                                           s->dim[2] * ((s->in_ld[1] + y) +
                                                        s->dim[1] * (s->in_ld[0] + x))];
  }
  return 0;
}

void p3mo_influence(p3mo const *s, double *g_force, double *g_energy) {
  size_t const M3 = (size_t)s->mesh[0] * s->mesh[1] * s->mesh[2];
  if (g_force) memcpy(g_force, s->g_force, sizeof(double) * M3);
  if (g_energy) memcpy(g_energy, s->g_energy, sizeof(double) * M3);
}

/* ------------------------------------------------------------------------ */
/* error estimates, p3m.cpp:145-243                                         */
/* ------------------------------------------------------------------------ */
static void tune_aliasing_sums(int const shift[3], int const mesh[3], double const mesh_i[3],
                               int cao, double alpha_L_i, double *alias1, double *alias2) {
  double const factor1 = sqr(P3MO_PI * alpha_L_i);
  *alias1 = 0.;
  *alias2 = 0.;
  /* P3M_BRILLOUIN = 0: the single image nm = shift */
  int nm[3];
  double fnm[3];
  for (int d = 0; d < 3; ++d) {
    nm[d] = shift[d] + 0 * mesh[d];
    fnm[d] = p3mo_sinc(nm[d] * mesh_i[d]);
  }
  int const norm_sq = nm[0] * nm[0] + nm[1] * nm[1] + nm[2] * nm[2];
  double const ex = exp(-factor1 * norm_sq);
  double const energy = pow(fnm[0] * fnm[1] * fnm[2], 2 * cao);
  *alias1 += sqr(ex) / norm_sq;
  *alias2 += energy * ex * (shift[0] * nm[0] + shift[1] * nm[1] + shift[2] * nm[2]) / norm_sq;
}

double p3mo_real_space_error(double pref, double r_cut_iL, int n_c_part, double sum_q2,
                             double alpha_L, double const box_l[3]) {
  double const volume = box_l[0] * box_l[1] * box_l[2];
  return (2. * pref * sum_q2 * exp(-sqr(r_cut_iL * alpha_L))) /
         sqrt((double)n_c_part * r_cut_iL * box_l[0] * volume);
}

double p3mo_k_space_error(double pref, int const mesh[3], int cao, int n_c_part, double sum_q2,
                          double alpha_L, double const box_l[3]) {
  double const mesh_i[3] = {1. / mesh[0], 1. / mesh[1], 1. / mesh[2]};
  double const alpha_L_i = 1. / alpha_L;
  int const stop[3] = {mesh[0] / 2, mesh[1] / 2, mesh[2] / 2};
  double he_q = 0.;
  double values[3];
  int ind[3];
  for (ind[0] = -stop[0]; ind[0] < stop[0]; ++ind[0]) {
    values[0] = p3mo_cotangent_sum(cao, ind[0], mesh_i[0]);
    for (ind[1] = -stop[1]; ind[1] < stop[1]; ++ind[1]) {
      values[1] = p3mo_cotangent_sum(cao, ind[1], mesh_i[1]);
      for (ind[2] = -stop[2]; ind[2] < stop[2]; ++ind[2]) {
        values[2] = p3mo_cotangent_sum(cao, ind[2], mesh_i[2]);
        if (ind[0] != 0 || ind[1] != 0 || ind[2] != 0) {
          int const n2 = ind[0] * ind[0] + ind[1] * ind[1] + ind[2] * ind[2];
          double const cs = values[0] * values[1] * values[2];
          double alias1, alias2;
          tune_aliasing_sums(ind, mesh, mesh_i, cao, alpha_L_i, &alias1, &alias2);
          double const d = alias1 - sqr(alias2 / cs) / n2;
          if (d > 0. && fabs(d / alias1) > ROUND_ERROR_PREC) he_q += d;
        }
      }
    }
  }
  return 2. * pref * sum_q2 * sqrt(he_q / (double)n_c_part) / (box_l[1] * box_l[2]);
}

/* CoulombTuningAlgorithm::calculate_accuracy, p3m.cpp:636-674.
 * out[0..3] = total error, real-space error, k-space error, alpha_L */
void p3mo_calculate_accuracy(double pref, int const mesh[3], int cao, double r_cut_iL,
                             int n_c_part, double sum_q2, double accuracy,
                             double const box_l[3], double *out) {
  double alpha_L;
  double rs_err = p3mo_real_space_error(pref, r_cut_iL, n_c_part, sum_q2, 0., box_l);
  if (P3MO_SQRT2 * rs_err > accuracy)
    alpha_L = sqrt(log(P3MO_SQRT2 * rs_err / accuracy)) / r_cut_iL;
  else
    alpha_L = 0.1;
  rs_err = p3mo_real_space_error(pref, r_cut_iL, n_c_part, sum_q2, alpha_L, box_l);
  double const ks_err = p3mo_k_space_error(pref, mesh, cao, n_c_part, sum_q2, alpha_L, box_l);
  out[0] = sqrt(rs_err * rs_err + ks_err * ks_err);
  out[1] = rs_err;
  out[2] = ks_err;
  out[3] = alpha_L;
}

/* r_cut bisection of TuningAlgorithm::get_mc_time, TuningAlgorithm.cpp:116-172,
 * for fixed (mesh, cao).  Returns 0 on success, -1 when cao is too large,
 * -3 when the accuracy cannot be reached.  out = {r_cut_iL, alpha_L, accuracy}.
 * r_cut_iL_min/max follow TuningAlgorithm::determine_r_cut_limits on one rank. */
int p3mo_tune_r_cut(double pref, int const mesh[3], int cao, int n_c_part, double sum_q2,
                    double accuracy, double const box_l[3], double skin, double r_cut_iL_min,
                    double r_cut_iL_max, double *out) {
  double min_box_l = box_l[0];
  for (int d = 1; d < 3; ++d)
    if (box_l[d] < min_box_l) min_box_l = box_l[d];
  double k_cut = 1e300;
  int min_mesh = mesh[0];
  for (int d = 0; d < 3; ++d) {
    double const kc = ((double)cao / 2.) * (box_l[d] / mesh[d]);
    if (kc < k_cut) k_cut = kc;
    if (mesh[d] < min_mesh) min_mesh = mesh[d];
  }
  double const k_cut_max = min_box_l - skin; /* min(box, local box) on one rank */
  if (cao >= min_mesh || k_cut >= k_cut_max) return -1;
  double res[4];
  p3mo_calculate_accuracy(pref, mesh, cao, r_cut_iL_max, n_c_part, sum_q2, accuracy, box_l, res);
  if (res[0] > accuracy) return -3;
  double lo = r_cut_iL_min, hi = r_cut_iL_max, r_cut_iL;
  for (;;) {
    r_cut_iL = 0.5 * (lo + hi);
    if (hi - lo < 1e-3) break; /* P3M_RCUT_PREC */
    p3mo_calculate_accuracy(pref, mesh, cao, r_cut_iL, n_c_part, sum_q2, accuracy, box_l, res);
    if (res[0] > accuracy)
      lo = r_cut_iL;
    else
      hi = r_cut_iL;
  }
  r_cut_iL = hi;
  p3mo_calculate_accuracy(pref, mesh, cao, r_cut_iL, n_c_part, sum_q2, accuracy, box_l, res);
  out[0] = r_cut_iL;
  out[1] = res[3];
  out[2] = res[0];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* real-space pair terms, p3m.hpp:146-180 (USE_ERFC_APPROXIMATION = 1)      */
/* ------------------------------------------------------------------------ */
static void pair_force(double prefactor, double alpha, double r_cut, double q1q2,
                       double const d[3], double dist, double out[3]) {
  out[0] = out[1] = out[2] = 0.;
  if (q1q2 == 0. || dist >= r_cut || dist <= 0.) return;
  double const adist = alpha * dist;
  double const exp_adist_sq = exp(-adist * adist);
  double const dist_sq = dist * dist;
  double const two_a_sqrt_pi_i = 2. * alpha * P3MO_INV_SQRTPI;
  double const erfc_part_ri = p3mo_as_erfc_part(adist) / dist;
  double const fac = exp_adist_sq * (erfc_part_ri + two_a_sqrt_pi_i) / dist_sq;
  for (int k = 0; k < 3; ++k) out[k] = (fac * prefactor * q1q2) * d[k];
}

static double pair_energy(double prefactor, double alpha, double r_cut, double q1q2,
                          double dist) {
  if (q1q2 == 0. || dist >= r_cut || dist <= 0.) return 0.;
  double const adist = alpha * dist;
  double const erfc_part_ri = p3mo_as_erfc_part(adist) / dist;
  return prefactor * q1q2 * erfc_part_ri * exp(-adist * adist);
}

/* Brute-force short-range loop over all pairs with the minimum image
 * convention (valid for r_cut < L/2).  Adds forces, returns the energy and
 * the real-space virial tensor (sum over pairs of F_ij (x) d_ij). */
double p3mo_real_space(int n, double const *pos, double const *q, double const box_l[3],
                       double prefactor, double alpha, double r_cut, double *force,
                       double *virial9) {
  double energy = 0.;
  if (virial9)
    for (int i = 0; i < 9; ++i) virial9[i] = 0.;
  for (int i = 0; i < n; ++i) {
    if (q[i] == 0.) continue;
    for (int j = i + 1; j < n; ++j) {
      if (q[j] == 0.) continue;
      double d[3], r2 = 0.;
      for (int k = 0; k < 3; ++k) {
        double x = pos[3 * (size_t)i + k] - pos[3 * (size_t)j + k];
        x -= box_l[k] * round(x / box_l[k]);
        d[k] = x;
        r2 += x * x;
      }
      double const dist = sqrt(r2);
      if (dist >= r_cut) continue;
      double f[3];
      pair_force(prefactor, alpha, r_cut, q[i] * q[j], d, dist, f);
      if (force)
        for (int k = 0; k < 3; ++k) {
          force[3 * (size_t)i + k] += f[k];
          force[3 * (size_t)j + k] -= f[k];
        }
      if (virial9)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) virial9[3 * a + b] += f[a] * d[b];
      energy += pair_energy(prefactor, alpha, r_cut, q[i] * q[j], dist);
    }
  }
  return energy;
}

/* ------------------------------------------------------------------------ */
/* FP64 Ewald reciprocal sum (independent of P3M), metallic boundaries.     */
/*   E_rec = pref/(2V) sum_{k!=0} 4 pi exp(-k^2/4a^2)/k^2 |S(k)|^2          */
/*   F_i   = pref q_i/V sum_{k!=0} 4 pi exp(-k^2/4a^2)/k^2 k Im(S*(k) e^{ikr_i}) */
/* The k-vectors n with |k| <= kmax are visited in the half space n > 0     */
/* (lexicographic) and doubled.  Phases use exp(ik.r) recursion per axis.  */
/* Forces are written for the first n_out particles (all sources used).     */
/* Returns E_rec (without self term).                                        */
/* ------------------------------------------------------------------------ */
double p3mo_ewald_kspace(int n, double const *pos, double const *q, double const box_l[3],
                         double prefactor, double alpha, double kmax, int n_out, double *force) {
  double const volume = box_l[0] * box_l[1] * box_l[2];
  double const two_pi = 2. * P3MO_PI;
  int nmax[3];
  for (int d = 0; d < 3; ++d) nmax[d] = (int)ceil(kmax * box_l[d] / two_pi);
  /* per-particle phase tables e^{i 2pi m x/L} for m in [-nmax, nmax] */
  size_t const w0 = 2 * (size_t)nmax[0] + 1, w1 = 2 * (size_t)nmax[1] + 1,
               w2 = 2 * (size_t)nmax[2] + 1;
  cplx *ex = malloc(sizeof(cplx) * (size_t)n * w0);
  cplx *ey = malloc(sizeof(cplx) * (size_t)n * w1);
  cplx *ez = malloc(sizeof(cplx) * (size_t)n * w2);
  cplx *tabs[3] = {ex, ey, ez};
  size_t const ws[3] = {w0, w1, w2};
  for (int d = 0; d < 3; ++d) {
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; ++p) {
      cplx *t = tabs[d] + (size_t)p * ws[d];
      int const c = nmax[d];
      double const ang = two_pi * pos[3 * (size_t)p + d] / box_l[d];
      for (int m = -c; m <= c; ++m) {
        t[m + c].re = cos(m * ang);
        t[m + c].im = sin(m * ang);
      }
    }
  }
  /* half space: (a>0) or (a==0,b>0) or (a==0,b==0,c>0), |k| <= kmax */
  double const kmax2 = kmax * kmax;
  size_t nk = 0;
  int *kv = malloc(sizeof(int) * 3 * w0 * w1 * w2);
  for (int a = 0; a <= nmax[0]; ++a)
    for (int b = -nmax[1]; b <= nmax[1]; ++b)
      for (int c = -nmax[2]; c <= nmax[2]; ++c) {
        if (a == 0 && (b < 0 || (b == 0 && c <= 0))) continue;
        double const k[3] = {two_pi * a / box_l[0], two_pi * b / box_l[1], two_pi * c / box_l[2]};
        if (k[0] * k[0] + k[1] * k[1] + k[2] * k[2] > kmax2) continue;
        kv[3 * nk] = a;
        kv[3 * nk + 1] = b;
        kv[3 * nk + 2] = c;
        ++nk;
      }
  for (int i = 0; i < 3 * n_out; ++i) force[i] = 0.;
  double energy = 0.;
#pragma omp parallel reduction(+ : energy)
  {
    double *fl = calloc(3 * (size_t)(n_out > 0 ? n_out : 1), sizeof(double));
#pragma omp for schedule(dynamic, 4)
    for (size_t kk = 0; kk < nk; ++kk) {
      int const a = kv[3 * kk], b = kv[3 * kk + 1], c = kv[3 * kk + 2];
      double const k[3] = {two_pi * a / box_l[0], two_pi * b / box_l[1], two_pi * c / box_l[2]};
      double const k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
      double const g = 4. * P3MO_PI * exp(-k2 / (4. * alpha * alpha)) / k2;
      double sr = 0., si = 0.;
      for (int p = 0; p < n; ++p) {
        if (q[p] == 0.) continue;
        cplx const A = ex[(size_t)p * w0 + (size_t)(a + nmax[0])];
        cplx const B = ey[(size_t)p * w1 + (size_t)(b + nmax[1])];
        cplx const C = ez[(size_t)p * w2 + (size_t)(c + nmax[2])];
        double const abr = A.re * B.re - A.im * B.im, abi = A.re * B.im + A.im * B.re;
        double const er = abr * C.re - abi * C.im, ei = abr * C.im + abi * C.re;
        sr += q[p] * er;
        si += q[p] * ei;
      }
      energy += 2. * g * (sr * sr + si * si); /* doubled half space */
      for (int p = 0; p < n_out; ++p) {
        cplx const A = ex[(size_t)p * w0 + (size_t)(a + nmax[0])];
        cplx const B = ey[(size_t)p * w1 + (size_t)(b + nmax[1])];
        cplx const C = ez[(size_t)p * w2 + (size_t)(c + nmax[2])];
        double const abr = A.re * B.re - A.im * B.im, abi = A.re * B.im + A.im * B.re;
        double const er = abr * C.re - abi * C.im, ei = abr * C.im + abi * C.re;
        /* Im(S* e^{ikr}) = sr*ei - si*er ; doubled half space */
        double const im = sr * ei - si * er;
        double const fac = 2. * g * q[p] * im * prefactor / volume;
        for (int d = 0; d < 3; ++d) fl[3 * (size_t)p + d] += fac * k[d];
      }
    }
#pragma omp critical
    for (int i = 0; i < 3 * n_out; ++i) force[i] += fl[i];
    free(fl);
  }
  free(kv);
  free(ex);
  free(ey);
  free(ez);
  return prefactor * energy / (2. * volume);
}

/* Real-space forces on the first n_out particles from all others within
 * r_cut (minimum image, r_cut < half the smallest box side): the P3M pair
 * force with the Abramowitz-Stegun erfc (p3m.hpp:146-165) when exact_erfc is
 * 0, or the exact Ewald real-space force with libm erfc when it is 1. */
void p3mo_real_space_subset(int n, double const *pos, double const *q, double const box_l[3],
                            double prefactor, double alpha, double r_cut, int n_out,
                            int exact_erfc, double *force) {
  double const two_a_sqrt_pi_i = 2. * alpha / sqrt(P3MO_PI);
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < n_out; ++i) {
    double f[3] = {0., 0., 0.};
    if (q[i] != 0.) {
      for (int j = 0; j < n; ++j) {
        if (j == i || q[j] == 0.) continue;
        double d[3], r2 = 0.;
        for (int k = 0; k < 3; ++k) {
          double x = pos[3 * (size_t)i + k] - pos[3 * (size_t)j + k];
          x -= box_l[k] * round(x / box_l[k]);
          d[k] = x;
          r2 += x * x;
        }
        if (r2 >= r_cut * r_cut || r2 == 0.) continue;
        double const dist = sqrt(r2);
        double fp[3];
        if (exact_erfc) {
          double const ad = alpha * dist;
          double const fac = prefactor * q[i] * q[j] *
                             (erfc(ad) / dist + two_a_sqrt_pi_i * exp(-ad * ad)) / r2;
          for (int k = 0; k < 3; ++k) fp[k] = fac * d[k];
        } else {
          pair_force(prefactor, alpha, r_cut, q[i] * q[j], d, dist, fp);
        }
        for (int k = 0; k < 3; ++k) f[k] += fp[k];
      }
    }
    for (int k = 0; k < 3; ++k) force[3 * (size_t)i + k] = f[k];
  }
}
