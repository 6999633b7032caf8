// C ABI of the MI355X P3M solver (include/espresso_amd_p3m.h).
//
// p3m_context plays the role of ESPResSo's CoulombP3MImpl<FloatType, Arch>
// (src/core/electrostatics/p3m.impl.hpp:62-184) for one system: it holds the
// validated parameters, the tuned state and the device solver.  Exceptions
// never cross the ABI; they become status codes plus a message.
#include "espresso_amd_p3m.h"

#include "p3m_dipolar.hpp"
#include "p3m_elc.hpp"
#include "p3m_host.hpp"
#include "p3m_batch.hpp"
#include "p3m_serve.hpp"
#include "p3m_short_range.hpp"
#include "p3m_solver.hpp"
#include "p3m_tuning.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sstream>
#include <memory>
#include <mutex>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>

using namespace espresso_amd;

struct p3m_context {
  int device = 0;
  std::unique_ptr<Solver> solver;
  std::unique_ptr<ShortRange> short_range; // created on first use
  int short_range_subdivision = 0;
  int force_conv = P3M_CONV_CPU;
  int tune_real_space = P3M_TUNE_RS_MODEL;
  // ELC on top of this solver (ElectrostaticLayerCorrection wrapping the
  // P3M actor): parameters as given, far_cut as last used
  bool elc_on = false;
  p3m_elc_params elc{};
  double elc_far_cut = -1.;
  std::unique_ptr<ElcSolver> elc_solver;
  p3m_params user{};
  P3MParams params{};
  bool have_params = false;
  bool is_tuned = false;
  bool initialized = false;
  std::string last_error;
  // staging for host pointers
  int64_t cap = 0;
  double *d_pos = nullptr, *d_q = nullptr, *d_upos = nullptr, *d_force = nullptr;
  // charge-assignment protocol of CoulombP3M (prepare_fft_mesh,
  // assign_charge, count_charged_particles_elc; p3m.hpp:91-143) as ELC drives
  // it: the interpolation cache keeps the positions and charges of the
  // assignments made without skip_cache, in order; the charge sums feed the
  // energy of p3m_solve_assigned (sum_qpart, sum_q2, square_sum_q)
  bool mesh_assigned = false;
  int64_t cache_n = 0, cache_cap = 0;
  double *d_cache_pos = nullptr, *d_cache_q = nullptr;
  double sum_qpart = 0., sum_q2 = 0., square_sum_q = 0.;
};

struct p3m_dipolar {
  int device = 0;
  std::unique_ptr<DipolarSolver> solver;
  p3m_dipolar_params user{};
  bool have_params = false;
  bool is_tuned = false;
  std::string last_error;
  int64_t cap = 0;
  double *d_pos = nullptr, *d_dip = nullptr, *d_f = nullptr, *d_t = nullptr;
};

struct p3m_batch {
  int device = 0;
  // shared with the servers built over the batch, which keep the solver
  // alive when the batch handle is destroyed before them
  std::shared_ptr<BatchSolver> solver;
  std::string last_error;
};

struct p3m_server {
  int device = 0;
  std::shared_ptr<BatchSolver> solver; // declared first: outlives `server`
  std::unique_ptr<Server> server;
  std::string last_error;
};

namespace {

p3m_launch_shape to_abi(LaunchShape const &s) {
  return p3m_launch_shape{s.gather, s.ncol,    s.zlines, s.fused_lds, s.spread,
                          s.threads, s.chunk, s.column, s.seg,       s.col_lanes,
                          s.group,   s.gather_threads};
}

LaunchShape from_abi(p3m_launch_shape const &a) {
  for (int v : {a.ncol, a.zlines})
    if (v < 1 || v > 64 || (v & (v - 1)) != 0)
      throw std::invalid_argument("P3M batch: ncol and zlines must be powers of two in [1, 64]");
  if (a.fused_lds < 0 || a.fused_lds > 128 * 1024)
    throw std::invalid_argument("P3M batch: fused_lds must be in [0, 131072]");
  if (a.gather < kGatherLayer || a.gather > kGatherPlane8)
    throw std::invalid_argument("P3M batch: gather must be 0 (layer), 1, 2 or 3 (plane)");
  if (a.spread < kSpreadSeg64 || a.spread > kSpreadSeg8)
    throw std::invalid_argument("P3M batch: spread must be 0, 1 or 2");
  if (a.threads != 256 && a.threads != 512)
    throw std::invalid_argument("P3M batch: threads must be 256 or 512");
  if (a.chunk < 1 || a.chunk > 64 || 3 * a.chunk > a.threads)
    throw std::invalid_argument("P3M batch: chunk must be in [1, 64] and at most threads / 3");
  if (a.column != 0 && a.column != 8 && a.column != 16)
    throw std::invalid_argument("P3M batch: column must be 0 (bricks), 8 or 16");
  if (a.seg < 1 || a.seg > 1024)
    throw std::invalid_argument("P3M batch: seg must be in [1, 1024]");
  if (a.col_lanes != 8 && a.col_lanes != 16)
    throw std::invalid_argument("P3M batch: col_lanes must be 8 or 16");
  if (a.group != 1 && a.group != 4)
    throw std::invalid_argument("P3M batch: group must be 1 or 4");
  if (a.gather_threads != 0 && a.gather_threads != 256 && a.gather_threads != 512)
    throw std::invalid_argument("P3M batch: gather_threads must be 0, 256 or 512");
  if (a.gather_threads != 0 && 3 * a.chunk > a.gather_threads)
    throw std::invalid_argument("P3M batch: chunk must be at most gather_threads / 3");
  LaunchShape s;
  s.gather = a.gather;
  s.ncol = a.ncol;
  s.zlines = a.zlines;
  s.fused_lds = a.fused_lds;
  s.spread = a.spread;
  s.threads = a.threads;
  s.chunk = a.chunk;
  s.column = a.column;
  s.seg = a.seg;
  s.col_lanes = a.col_lanes;
  s.group = a.group;
  s.gather_threads = a.gather_threads;
  return s;
}
std::mutex g_err_mutex;
std::string g_last_error;

void set_global_error(std::string const &msg) {
  std::lock_guard<std::mutex> lock(g_err_mutex);
  g_last_error = msg;
}

template <typename F> int guarded(std::string *err, F &&f) {
  auto fail = [&](std::string const &msg, int code) {
    if (err) *err = msg;
    set_global_error(msg);
    return code;
  };
  try {
    f();
    return P3M_OK;
  } catch (std::domain_error const &e) {
    return fail(e.what(), P3M_ERR_DOMAIN);
  } catch (std::invalid_argument const &e) {
    return fail(e.what(), P3M_ERR_INVALID);
  } catch (std::bad_alloc const &e) {
    return fail(e.what(), P3M_ERR_MEMORY);
  } catch (std::runtime_error const &e) {
    std::string const msg = e.what();
    return fail(msg, msg.rfind("HIP error", 0) == 0 ? P3M_ERR_DEVICE : P3M_ERR_RUNTIME);
  } catch (std::exception const &e) {
    return fail(e.what(), P3M_ERR_RUNTIME);
  } catch (...) {
    return fail("unknown error", P3M_ERR_RUNTIME);
  }
}

template <typename F> int guarded(p3m_context *ctx, F &&f) {
  return guarded(ctx ? &ctx->last_error : nullptr, std::forward<F>(f));
}

Vec3d v3(double const *a) { return {a[0], a[1], a[2]}; }
Vec3i v3i(int const *a) { return {a[0], a[1], a[2]}; }

void free_staging(p3m_context *ctx) {
  for (double **p : {&ctx->d_pos, &ctx->d_q, &ctx->d_upos, &ctx->d_force}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  ctx->cap = 0;
}

void free_cache(p3m_context *ctx) {
  for (double **p : {&ctx->d_cache_pos, &ctx->d_cache_q}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  ctx->cache_cap = 0;
  ctx->cache_n = 0;
}

/** Charge sums of count_charged_particles (p3m.cpp:97-115). */
void count_charges(p3m_context *ctx, std::vector<double> const &charges) {
  ctx->sum_qpart = 0.;
  ctx->sum_q2 = 0.;
  double sum_q = 0.;
  for (double const qq : charges)
    if (qq != 0.) {
      ctx->sum_qpart += 1.;
      ctx->sum_q2 += qq * qq;
      sum_q += qq;
    }
  ctx->square_sum_q = sum_q * sum_q;
}

void ensure_staging(p3m_context *ctx, int64_t n) {
  if (n <= ctx->cap) return;
  free_staging(ctx);
  int64_t const cap = std::max<int64_t>(n, 1024);
  hip_check(hipMalloc(&ctx->d_pos, sizeof(double) * 3 * cap), "hipMalloc pos");
  hip_check(hipMalloc(&ctx->d_q, sizeof(double) * cap), "hipMalloc q");
  hip_check(hipMalloc(&ctx->d_upos, sizeof(double) * 3 * cap), "hipMalloc upos");
  hip_check(hipMalloc(&ctx->d_force, sizeof(double) * 3 * cap), "hipMalloc force");
  ctx->cap = cap;
}

/** Copies host inputs to the staging buffers (unless they already are device
 *  pointers) and returns the device pointers to use. */
struct DevArgs {
  double const *pos, *q, *upos;
  double *force;
};
DevArgs stage_inputs(p3m_context *ctx, int64_t n, double const *pos, double const *q,
                     double const *upos, double *force, unsigned flags) {
  if (flags & P3M_DEVICE_POINTERS) return {pos, q, upos, force};
  ensure_staging(ctx, n);
  hipStream_t st = ctx->solver->stream();
  DevArgs d{ctx->d_pos, ctx->d_q, nullptr, nullptr};
  if (n > 0) {
    hip_check(hipMemcpyAsync(ctx->d_pos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st),
              "copy pos");
    hip_check(hipMemcpyAsync(ctx->d_q, q, sizeof(double) * n, hipMemcpyHostToDevice, st),
              "copy q");
    if (upos) {
      hip_check(hipMemcpyAsync(ctx->d_upos, upos, sizeof(double) * 3 * n,
                               hipMemcpyHostToDevice, st),
                "copy upos");
      d.upos = ctx->d_upos;
    }
    if (force) {
      hip_check(hipMemcpyAsync(ctx->d_force, force, sizeof(double) * 3 * n,
                               hipMemcpyHostToDevice, st),
                "copy force");
      d.force = ctx->d_force;
    }
  }
  return d;
}

std::vector<double> host_charges(p3m_context *ctx, int64_t n, double const *q, unsigned flags) {
  std::vector<double> h(static_cast<size_t>(n));
  if (n == 0) return h;
  if (flags & P3M_DEVICE_POINTERS) {
    hip_check(hipMemcpy(h.data(), q, sizeof(double) * n, hipMemcpyDeviceToHost), "copy q");
  } else {
    std::memcpy(h.data(), q, sizeof(double) * n);
  }
  (void)ctx;
  return h;
}

void require_params(p3m_context const *ctx) {
  if (!ctx->have_params) throw std::runtime_error("P3M: parameters not set");
}
void require_device(p3m_context const *ctx) {
  if (!ctx->solver) throw std::runtime_error("P3M: host-only context has no device");
}

MeshPrecision prec_of(p3m_params const &u) {
  return u.precision == P3M_PREC_FP32 ? MeshPrecision::FP32 : MeshPrecision::FP64;
}

/** init(): derive mesh constants, sanity checks, build device state
 *  (init_cpu_kernels / p3m_gpu_init, p3m.cpp:245-280). */
/** ELC parameters for the current box: box_h = L_z - gap
 *  (ElectrostaticLayerCorrection::recalc_box_h, elc.cpp:1020-1030), the
 *  space layer of dielectric contrasts (recalc_space_layer, :1032-1054;
 *  gap / 3 before P3M is tuned, as the elc_data constructor sets it) and
 *  far_cut, tuned from maxPWerror unless given (recalc_far_cut,
 *  elc.hpp:319-324). */
ElcParams elc_geometry(p3m_context *ctx) {
  ElcParams ep;
  for (int d = 0; d < 3; ++d) ep.box[d] = ctx->user.box_l[d];
  ep.prefactor = ctx->user.prefactor;
  ep.gap_size = ctx->elc.gap_size;
  ep.neutralize = ctx->elc.neutralize != 0;
  ep.delta_top = ctx->elc.delta_mid_top;
  ep.delta_bot = ctx->elc.delta_mid_bot;
  ep.const_pot = ctx->elc.const_pot != 0;
  ep.pot_diff = ctx->elc.pot_diff;
  ep.alpha = ctx->params.alpha;
  ep.r_cut = ctx->params.r_cut;
  double const box_h = ep.box[2] - ep.gap_size;
  if (box_h < 0.)
    throw std::runtime_error("ELC gap size (" + std::to_string(ep.gap_size) +
                             ") larger than box length in z-direction (" +
                             std::to_string(ep.box[2]) + ")");
  if (ep.dielectric())
    ep.space_layer = ctx->is_tuned ? elc_space_layer(ep.gap_size, box_h, ctx->params.r_cut)
                                   : ep.gap_size / 3.;
  ep.far_cut = (ctx->elc.far_cut == -1.)
                   ? elc_tune_far_cut(ep.box, ep.gap_size, ctx->elc.maxPWerror,
                                      ep.dielectric(), ep.space_layer)
                   : ctx->elc.far_cut;
  ctx->elc_far_cut = ep.far_cut;
  return ep;
}

void init_solver(p3m_context *ctx) {
  auto &p = ctx->params;
  auto const box = v3(ctx->user.box_l);
  p.cao3 = p.cao * p.cao * p.cao;
  p.recalc_a_ai_cao_cut(box);
  sanity_checks_boxl(p, box);
  p.alpha_L = p.alpha * box[0];
  p.r_cut_iL = p.r_cut * (1. / box[0]);
  require_device(ctx);
  ctx->solver->configure(p, box, ctx->user.skin, ctx->user.prefactor, prec_of(ctx->user));
  ctx->initialized = true;
  ctx->mesh_assigned = false; // new meshes
}
} // namespace

extern "C" {

int p3m_params_default(p3m_params *p) {
  if (!p) return P3M_ERR_INVALID;
  std::memset(p, 0, sizeof(*p));
  p->mesh[0] = p->mesh[1] = p->mesh[2] = -1;
  p->cao = -1;
  p->alpha = -1.;
  p->r_cut = -1.;
  p->mesh_off[0] = p->mesh_off[1] = p->mesh_off[2] = -1.;
  p->epsilon = 0.;
  p->accuracy = 1e-3;
  p->prefactor = 0.;
  p->box_l[0] = p->box_l[1] = p->box_l[2] = 1.;
  p->skin = 0.4;
  p->precision = P3M_PREC_FP64;
  p->tune = 1;
  p->timings = 10;
  p->verbose = 1;
  p->charge_neutrality_tolerance = 2e-12;
  return P3M_OK;
}

int p3m_create(int device, p3m_context **out) {
  if (!out) return P3M_ERR_INVALID;
  *out = nullptr;
  return guarded(static_cast<std::string *>(nullptr), [&] {
    auto ctx = std::make_unique<p3m_context>();
    ctx->device = device;
    if (device != -1) { // device -1: host-only context (validation, estimates)
      int ndev = 0;
      hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
      if (device < 0 || device >= ndev)
        throw std::invalid_argument("P3M: invalid HIP device " + std::to_string(device));
      ctx->solver = std::make_unique<Solver>(device);
    }
    *out = ctx.release();
  });
}

int p3m_destroy(p3m_context *ctx) {
  if (!ctx) return P3M_OK;
  if (ctx->solver) (void)hipSetDevice(ctx->device);
  free_staging(ctx);
  free_cache(ctx);
  delete ctx;
  return P3M_OK;
}

int p3m_set_params(p3m_context *ctx, p3m_params const *p) {
  if (!ctx || !p) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    // CoulombP3M script interface construction (CoulombP3M.hpp:102-128),
    // P3MParameters (common.hpp:84-150), CoulombP3MImpl (p3m.impl.hpp:81-92),
    // Actor::set_prefactor (actor.hpp:48-54)
    bool const tuning = p->tune != 0;
    P3MParams params(tuning, p->epsilon, p->r_cut, v3i(p->mesh), v3(p->mesh_off), p->cao,
                     p->alpha, p->accuracy);
    if (p->timings <= 0) throw std::domain_error("Parameter 'timings' must be > 0");
    if (p->prefactor <= 0.) throw std::domain_error("Parameter 'prefactor' must be > 0");
    for (int d = 0; d < 3; ++d)
      if (!(p->box_l[d] > 0.)) throw std::domain_error("box_l must be > 0");
    ctx->is_tuned = not params.tuning;
    params.tuning = false;
    ctx->params = params;
    ctx->user = *p;
    if (ctx->elc_on) { // adapt_epsilon_elc (p3m.impl.hpp:112-114)
      ctx->params.epsilon = P3M_EPSILON_METALLIC;
      ctx->user.epsilon = P3M_EPSILON_METALLIC;
    }
    ctx->have_params = true;
    ctx->initialized = false;
  });
}

int p3m_get_params(p3m_context const *ctx, p3m_params *out) {
  if (!ctx || !out) return P3M_ERR_INVALID;
  *out = ctx->user;
  auto const &p = ctx->params;
  for (int d = 0; d < 3; ++d) {
    out->mesh[d] = p.mesh[d];
    out->mesh_off[d] = p.mesh_off[d];
  }
  out->cao = p.cao;
  out->alpha = p.alpha;
  out->r_cut = p.r_cut;
  out->accuracy = p.accuracy;
  out->epsilon = p.epsilon;
  out->tune = ctx->is_tuned ? 0 : 1;
  return P3M_OK;
}

int p3m_is_tuned(p3m_context const *ctx, int *out) {
  if (!ctx || !out) return P3M_ERR_INVALID;
  *out = ctx->is_tuned ? 1 : 0;
  return P3M_OK;
}

namespace {
/** Mean time of the device pair loop (p3m_short_range) for trial
 *  parameters, in ms: one warm-up call, then `timings` calls between two
 *  events on the solver stream.  Forces go to scratch. */
double time_short_range(p3m_context *ctx, P3MParams const &trial, Vec3d const &box, int64_t n,
                        double const *d_pos, double const *d_q, double *scratch_force,
                        int timings) {
  if (!ctx->short_range) ctx->short_range = std::make_unique<ShortRange>(ctx->device);
  ShortRangeParams sp;
  for (int k = 0; k < 3; ++k) sp.box[k] = box[k];
  sp.prefactor = ctx->user.prefactor;
  sp.alpha = trial.alpha;
  sp.r_cut = trial.r_cut;
  sp.subdivision = ctx->short_range_subdivision;
  hipStream_t const st = ctx->solver->stream();
  ctx->short_range->compute(st, sp, n, d_pos, d_q, scratch_force, nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  float ms = 0.f;
  try {
    hip_check(hipEventRecord(e0, st), "hipEventRecord");
    for (int t = 0; t < timings; ++t)
      ctx->short_range->compute(st, sp, n, d_pos, d_q, scratch_force, nullptr);
    hip_check(hipEventRecord(e1, st), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return static_cast<double>(ms) / timings;
}
} // namespace

int p3m_activate(p3m_context *ctx, int64_t n, double const *pos, double const *q,
                 unsigned flags) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_params(ctx);
    if (ctx->solver) hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    auto const box = v3(ctx->user.box_l);
    // sanity checks before tuning (p3m.impl.hpp:124-139, p3m.hpp:67-73):
    // the k-space cutoff is not known yet, only the epsilon/cubic rule and
    // charge neutrality apply
    if (ctx->params.epsilon != P3M_EPSILON_METALLIC) {
      if (box[0] != box[1] or box[1] != box[2] or
          (ctx->params.mesh[0] != -1 and (ctx->params.mesh[0] != ctx->params.mesh[1] or
                                          ctx->params.mesh[1] != ctx->params.mesh[2])))
        throw std::runtime_error("CoulombP3M: non-metallic epsilon requires cubic box");
    }
    auto const charges = host_charges(ctx, n, q, flags);
    check_charge_neutrality(charges, ctx->user.charge_neutrality_tolerance);
    count_charges(ctx, charges);
    ctx->mesh_assigned = false; // tuning solves overwrite the mesh
    bool const elc_diel =
        ctx->elc_on && (ctx->elc.delta_mid_top != 0. || ctx->elc.delta_mid_bot != 0.);
    if (ctx->elc_on) {
      (void)elc_geometry(ctx); // gap and far_cut (elc.hpp:181-200)
      if (elc_diel) {
        // sanity_checks_dielectric_contrasts (elc.cpp:990-1008)
        double total = 0.;
        for (double const qq : charges) total += qq;
        if (std::fabs(total) >= std::sqrt(1e-14)) {
          if (ctx->elc.const_pot)
            throw std::runtime_error("ELC does not currently support non-neutral systems with a "
                                     "dielectric contrast.");
          throw std::runtime_error("ELC does not work for non-neutral systems and non-metallic "
                                   "dielectric contrast.");
        }
      }
    }
    // tune (p3m.cpp:757-791)
    auto &p = ctx->params;
    if (p.alpha_L == 0. and p.alpha != 0.) p.alpha_L = p.alpha * box[0];
    if (p.r_cut_iL == 0. and p.r_cut != 0.) p.r_cut_iL = p.r_cut * (1. / box[0]);
    if (!ctx->is_tuned) {
      TuningInput in;
      in.box_l = box;
      in.skin = ctx->user.skin;
      in.prefactor = ctx->user.prefactor;
      in.timings = ctx->user.timings;
      in.verbose = ctx->user.verbose != 0;
      in.name = "CoulombP3M";
      for (double const qq : charges)
        if (qq != 0.) {
          in.sum_qpart++;
          in.sum_q2 += qq * qq;
        }
      if (in.sum_qpart == 0)
        throw std::runtime_error("CoulombP3M: no charged particles in the system");
      require_device(ctx);
      DevArgs const d = stage_inputs(ctx, n, pos, q, nullptr, nullptr, flags);
      ensure_staging(ctx, n);
      double *scratch_force = ctx->d_force; // forces are computed but discarded
      TuningTimer timer = [&](Vec3i const &mesh, int cao, double r_cut_iL, double alpha_L,
                              int timings) -> double {
        // ELC with dielectric contrasts vetoes r_cut >= gap_size
        // (veto_r_cut, elc.hpp:243-248)
        if (elc_diel && r_cut_iL * box[0] >= ctx->elc.gap_size) return -1.;
        P3MParams trial = p;
        trial.mesh = mesh;
        trial.cao = cao;
        trial.alpha_L = alpha_L;
        trial.alpha = alpha_L * (1. / box[0]);
        trial.r_cut_iL = r_cut_iL;
        trial.r_cut = r_cut_iL * box[0];
        trial.recalc_a_ai_cao_cut(box);
        try {
          sanity_checks_boxl(trial, box);
          ctx->solver->configure(trial, box, ctx->user.skin, ctx->user.prefactor,
                                 prec_of(ctx->user));
        } catch (std::runtime_error const &) {
          return -1.;
        }
        SolveResult r;
        ctx->solver->solve(n, d.pos, d.q, nullptr, nullptr, P3M_ENERGY, &r); // warm-up
        double total = 0.;
        for (int t = 0; t < timings; ++t) {
          ctx->solver->solve(n, d.pos, d.q, nullptr, scratch_force, P3M_FORCES | P3M_ENERGY,
                             &r);
          total += ctx->solver->last_ms[0];
        }
        double rs_ms = model_real_space_ms(in.sum_qpart, trial.r_cut, box);
        if (ctx->tune_real_space == P3M_TUNE_RS_DEVICE &&
            trial.r_cut < 0.5 * std::min({box[0], box[1], box[2]}))
          rs_ms = time_short_range(ctx, trial, box, n, d.pos, d.q, scratch_force, timings);
        return total / timings + rs_ms;
      };
      tune_p3m(p, in, timer);
      ctx->is_tuned = true;
    }
    init_solver(ctx);
    if (ctx->elc_on) (void)elc_geometry(ctx); // space layer from the tuned r_cut
  });
}

int p3m_solve(p3m_context *ctx, int64_t n, double const *pos, double const *q,
              double const *pos_unfolded, double *force, unsigned flags, double *energy,
              double *virial) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_params(ctx);
    require_device(ctx);
    if (!ctx->is_tuned) throw std::runtime_error("P3M: solver used before activation");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->initialized) init_solver(ctx);
    if (n < 0) throw std::invalid_argument("P3M: negative particle count");
    if ((flags & P3M_FORCES) && !force && n > 0)
      throw std::invalid_argument("P3M: force array required");
    bool const async = (flags & P3M_ASYNC) != 0;
    if (async && !(flags & P3M_DEVICE_POINTERS))
      throw std::invalid_argument("P3M: P3M_ASYNC requires P3M_DEVICE_POINTERS");
    bool const dev_res = (flags & P3M_DEVICE_RESULTS) != 0;
    if (dev_res && !(flags & P3M_DEVICE_POINTERS))
      throw std::invalid_argument("P3M: P3M_DEVICE_RESULTS requires P3M_DEVICE_POINTERS");
    if (dev_res && ctx->elc_on)
      throw std::invalid_argument("P3M: P3M_DEVICE_RESULTS is not available with ELC");
    if (flags & ~(P3M_FORCES | P3M_ENERGY | P3M_PRESSURE | P3M_DEVICE_POINTERS | P3M_ASYNC |
                  P3M_DEVICE_RESULTS))
      throw std::invalid_argument("P3M: unknown solve flags");
    DevArgs const d = stage_inputs(ctx, n, pos, q, pos_unfolded, force, flags);
    ctx->mesh_assigned = false;
    SolveResult r;
    // one P3M pass over (nn, pp, qq): forces with the context's convention
    // (the GPU kernels' index rule makes a second pass for energy and
    // virial from the CPU kernels' mesh, p3m.cpp:867-892)
    auto p3m_pass = [&](int64_t nn, double const *pp, double const *qq, double const *up,
                        double *ff, unsigned fl, SolveResult *res) {
      unsigned const reductions = fl & (P3M_ENERGY | P3M_PRESSURE);
      if (ctx->force_conv == P3M_CONV_GPU && (fl & P3M_FORCES)) {
        ctx->solver->set_convention(1);
        ctx->solver->solve(nn, pp, qq, up, ff, fl & ~reductions, nullptr, !async);
        if (reductions && !async) {
          ctx->solver->set_convention(0);
          ctx->solver->solve(nn, pp, qq, up, nullptr, fl & ~P3M_FORCES, res, true);
        }
      } else {
        ctx->solver->set_convention(0);
        ctx->solver->solve(nn, pp, qq, up, (fl & P3M_FORCES) ? ff : nullptr, fl, res, !async);
      }
    };
    bool const diel =
        ctx->elc_on && (ctx->elc.delta_mid_top != 0. || ctx->elc.delta_mid_bot != 0.);
    ElcParams ep;
    if (ctx->elc_on) {
      ep = elc_geometry(ctx);
      if (!ctx->elc_solver) ctx->elc_solver = std::make_unique<ElcSolver>(ctx->device);
    }
    hipStream_t const st = ctx->solver->stream();
    if (dev_res) {
      // energy and virial stay on the device (an MD loop's NpT virial,
      // p3m.cpp:580-583, without a host round trip)
      unsigned const red = flags & (P3M_ENERGY | P3M_PRESSURE);
      if (ctx->force_conv == P3M_CONV_GPU && (flags & P3M_FORCES)) {
        // forces with the GPU rule, energy and virial from the CPU mesh
        ctx->solver->set_convention(1);
        ctx->solver->solve(n, d.pos, d.q, d.upos, d.force, flags & ~red, nullptr, false);
        if (red) {
          ctx->solver->set_convention(0);
          ctx->solver->solve(n, d.pos, d.q, d.upos, nullptr, flags & ~P3M_FORCES, nullptr,
                             false);
        }
      } else {
        ctx->solver->set_convention(0);
        ctx->solver->solve(n, d.pos, d.q, d.upos, (flags & P3M_FORCES) ? d.force : nullptr, flags,
                           nullptr, false);
      }
      ctx->solver->device_results((flags & P3M_ENERGY) ? energy : nullptr,
                                  (flags & P3M_PRESSURE) ? virial : nullptr);
      if (!async) hip_check(hipStreamSynchronize(ctx->solver->stream()), "P3M solve");
      return;
    }
    if (!diel) {
      p3m_pass(n, d.pos, d.q, d.upos, d.force, flags, &r);
    } else {
      // dielectric contrasts (ElectrostaticLayerCorrection::
      // add_long_range_forces / long_range_energy, elc.cpp:1189-1256): the
      // mesh holds the real and the image charges; the forces are gathered
      // at the real particles, plus each particle's pair term with its own
      // image; the energy is 0.5 E(real) + 0.5 E_self + 0.5 E(real + images)
      // - 0.5 E(images)
      ElcSolver &es = *ctx->elc_solver;
      double *pos3 = nullptr, *q_both = nullptr, *q_img = nullptr;
      es.images(st, ep, n, d.pos, d.q, &pos3, &q_both, &q_img);
      if (flags & P3M_FORCES) {
        double *f3 = es.image_forces(st, n);
        p3m_pass(3 * n, pos3, q_both, nullptr, f3, flags & ~(P3M_ENERGY | P3M_PRESSURE),
                 nullptr);
        es.add_real_rows(st, n, f3, d.force);
        es.layer_self(st, ep, n, d.pos, d.q, d.force, nullptr);
      }
      if ((flags & P3M_ENERGY) && !async) {
        unsigned const fe = (flags & ~(P3M_FORCES | P3M_PRESSURE | P3M_ASYNC)) | P3M_ENERGY;
        SolveResult r_real, r_both, r_img;
        p3m_pass(n, d.pos, d.q, d.upos, nullptr, fe, &r_real);
        double e_self = 0.;
        es.layer_self(st, ep, n, d.pos, d.q, nullptr, &e_self);
        p3m_pass(3 * n, pos3, q_both, nullptr, nullptr, fe, &r_both);
        p3m_pass(3 * n, pos3, q_img, nullptr, nullptr, fe, &r_img);
        r.energy = 0.5 * r_real.energy + 0.5 * e_self + 0.5 * r_both.energy - 0.5 * r_img.energy;
      }
    }
    double e_elc = 0.;
    if (ctx->elc_on) {
      // ElectrostaticLayerCorrection::add_long_range_forces / long_range_energy
      // (elc.cpp:1189-1256): the P3M solve above plus the far-field terms
      bool const want_e = (flags & P3M_ENERGY) && !async;
      ctx->elc_solver->compute(st, ep, n, d.pos, d.q,
                               (flags & P3M_FORCES) ? d.force : nullptr, want_e ? &e_elc : nullptr,
                               !async);
    }
    if (async) {
      if (energy) *energy = 0.;
      if (virial)
        for (int k = 0; k < 9; ++k) virial[k] = 0.;
      return;
    }
    if (ctx->elc_on) {
      r.energy += e_elc;
      // the reference has no ELC pressure: it warns and returns zero
      // (coulomb.cpp:125-131)
      for (int k = 0; k < 9; ++k) r.virial[k] = 0.;
    }
    if ((flags & P3M_FORCES) && !(flags & P3M_DEVICE_POINTERS) && n > 0) {
      hip_check(hipMemcpy(force, d.force, sizeof(double) * 3 * n, hipMemcpyDeviceToHost),
                "copy force");
    }
    if (energy) *energy = (flags & P3M_ENERGY) ? r.energy : 0.;
    if (virial)
      for (int k = 0; k < 9; ++k) virial[k] = (flags & P3M_PRESSURE) ? r.virial[k] : 0.;
  });
}

int p3m_solve_noalloc(p3m_context *ctx, int64_t n, double const *pos, double const *q,
                      double const *pos_unfolded, double *force, unsigned flags, double *energy,
                      double *virial) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_params(ctx);
    require_device(ctx);
    if (!ctx->is_tuned || !ctx->initialized)
      throw std::runtime_error("P3M: p3m_solve_noalloc needs a context set up by p3m_activate "
                               "(or by an earlier p3m_solve)");
    if (ctx->elc_on) throw std::invalid_argument("P3M: p3m_solve_noalloc does not support ELC");
    if (!(flags & P3M_DEVICE_POINTERS))
      throw std::invalid_argument("P3M: p3m_solve_noalloc takes device pointers only "
                                  "(P3M_DEVICE_POINTERS)");
    if (flags & ~(P3M_FORCES | P3M_ENERGY | P3M_PRESSURE | P3M_DEVICE_POINTERS))
      throw std::invalid_argument("P3M: p3m_solve_noalloc is synchronous: P3M_ASYNC and "
                                  "P3M_DEVICE_RESULTS are not available");
    if (n < 0) throw std::invalid_argument("P3M: negative particle count");
    if ((flags & P3M_FORCES) && !force && n > 0)
      throw std::invalid_argument("P3M: force array required");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ctx->mesh_assigned = false;
    SolveResult r;
    unsigned const red = flags & (P3M_ENERGY | P3M_PRESSURE);
    if (ctx->force_conv == P3M_CONV_GPU && (flags & P3M_FORCES)) {
      // forces with the GPU index rule, energy and virial from the CPU mesh
      ctx->solver->set_convention(1);
      ctx->solver->solve_fixed(n, pos, q, pos_unfolded, force, flags & ~red, nullptr);
      if (red) {
        ctx->solver->set_convention(0);
        ctx->solver->solve_fixed(n, pos, q, pos_unfolded, nullptr, flags & ~P3M_FORCES, &r);
      }
    } else {
      ctx->solver->set_convention(0);
      ctx->solver->solve_fixed(n, pos, q, pos_unfolded, (flags & P3M_FORCES) ? force : nullptr,
                               flags, &r);
    }
    if (energy) *energy = (flags & P3M_ENERGY) ? r.energy : 0.;
    if (virial)
      for (int k = 0; k < 9; ++k) virial[k] = (flags & P3M_PRESSURE) ? r.virial[k] : 0.;
  });
}

int p3m_noalloc_scratch_bytes(int64_t *bytes) {
  if (!bytes) return P3M_ERR_INVALID;
  *bytes = static_cast<int64_t>(BatchSolver::fixed_scratch_bytes());
  return P3M_OK;
}

namespace {
void require_solver_ready(p3m_context *ctx) {
  require_params(ctx);
  require_device(ctx);
  if (!ctx->is_tuned) throw std::runtime_error("P3M: solver used before activation");
  hip_check(hipSetDevice(ctx->device), "hipSetDevice");
  if (!ctx->initialized) init_solver(ctx);
}

void check_protocol_flags(unsigned flags, unsigned allowed) {
  if (flags & ~allowed) throw std::invalid_argument("P3M: unsupported flags for this call");
}
} // namespace

int p3m_prepare_fft_mesh(p3m_context *ctx, int reset_weights) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_solver_ready(ctx);
    // zero the charge mesh: a spread of no particles into a cleared mesh
    ctx->solver->set_convention(0);
    ctx->solver->solve(0, ctx->d_pos, ctx->d_q, nullptr, nullptr, kSolveMeshOnly, nullptr, true);
    if (reset_weights) ctx->cache_n = 0;
    ctx->mesh_assigned = true;
  });
}

int p3m_assign_charges(p3m_context *ctx, int64_t n, double const *pos, double const *q,
                       int skip_cache, unsigned flags) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_solver_ready(ctx);
    check_protocol_flags(flags, P3M_DEVICE_POINTERS);
    if (n < 0) throw std::invalid_argument("P3M: negative particle count");
    if (!ctx->mesh_assigned)
      throw std::runtime_error("P3M: assign_charge before prepare_fft_mesh");
    if (n == 0) return;
    DevArgs const d = stage_inputs(ctx, n, pos, q, nullptr, nullptr, flags);
    hipStream_t const st = ctx->solver->stream();
    ctx->solver->set_convention(0);
    ctx->solver->solve(n, d.pos, d.q, nullptr, nullptr, kSolveAccumulate | kSolveMeshOnly,
                       nullptr, true);
    if (!skip_cache) {
      int64_t const need = ctx->cache_n + n;
      if (need > ctx->cache_cap) {
        int64_t const cap = std::max<int64_t>(need + need / 2, 1024);
        double *np = nullptr, *nq = nullptr;
        hip_check(hipMalloc(&np, sizeof(double) * 3 * cap), "hipMalloc cache");
        hip_check(hipMalloc(&nq, sizeof(double) * cap), "hipMalloc cache");
        if (ctx->cache_n > 0) {
          hip_check(hipMemcpyAsync(np, ctx->d_cache_pos, sizeof(double) * 3 * ctx->cache_n,
                                   hipMemcpyDeviceToDevice, st),
                    "copy cache");
          hip_check(hipMemcpyAsync(nq, ctx->d_cache_q, sizeof(double) * ctx->cache_n,
                                   hipMemcpyDeviceToDevice, st),
                    "copy cache");
        }
        hip_check(hipStreamSynchronize(st), "sync");
        int64_t const keep = ctx->cache_n;
        free_cache(ctx);
        ctx->d_cache_pos = np;
        ctx->d_cache_q = nq;
        ctx->cache_cap = cap;
        ctx->cache_n = keep;
      }
      hip_check(hipMemcpyAsync(ctx->d_cache_pos + 3 * ctx->cache_n, d.pos, sizeof(double) * 3 * n,
                               hipMemcpyDeviceToDevice, st),
                "copy cache");
      hip_check(hipMemcpyAsync(ctx->d_cache_q + ctx->cache_n, d.q, sizeof(double) * n,
                               hipMemcpyDeviceToDevice, st),
                "copy cache");
      hip_check(hipStreamSynchronize(st), "sync");
      ctx->cache_n = need;
    }
  });
}

int p3m_charge_assign(p3m_context *ctx, int64_t n, double const *pos, double const *q,
                      unsigned flags) {
  int const st = p3m_prepare_fft_mesh(ctx, 1);
  if (st != P3M_OK) return st;
  return p3m_assign_charges(ctx, n, pos, q, 0, flags);
}

int p3m_count_charged_particles(p3m_context *ctx, int64_t n, double const *q, unsigned flags) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    check_protocol_flags(flags, P3M_DEVICE_POINTERS);
    if (n < 0) throw std::invalid_argument("P3M: negative particle count");
    if (n > 0 && !q) throw std::invalid_argument("P3M: charges required");
    if ((flags & P3M_DEVICE_POINTERS) && ctx->solver)
      hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    count_charges(ctx, host_charges(ctx, n, q, flags));
  });
}

int p3m_count_charged_particles_elc(p3m_context *ctx, int n, double sum_q2, double square_sum_q) {
  if (!ctx) return P3M_ERR_INVALID;
  ctx->sum_qpart = n;
  ctx->sum_q2 = sum_q2;
  ctx->square_sum_q = square_sum_q;
  return P3M_OK;
}

int p3m_adapt_epsilon_elc(p3m_context *ctx) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_params(ctx);
    ctx->params.epsilon = P3M_EPSILON_METALLIC;
    ctx->user.epsilon = P3M_EPSILON_METALLIC;
    ctx->initialized = false; // the dipole term of the descriptors changes
  });
}

int p3m_assigned_count(p3m_context const *ctx, int64_t *n) {
  if (!ctx || !n) return P3M_ERR_INVALID;
  *n = ctx->cache_n;
  return P3M_OK;
}

int p3m_solve_assigned(p3m_context *ctx, double const *pos_unfolded, double *force,
                       unsigned flags, double *energy) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_solver_ready(ctx);
    check_protocol_flags(flags, P3M_FORCES | P3M_ENERGY | P3M_DEVICE_POINTERS);
    if (!ctx->mesh_assigned)
      throw std::runtime_error("P3M: no charges assigned since the last solve");
    int64_t const n = ctx->cache_n;
    bool const want_f = (flags & P3M_FORCES) && n > 0;
    if (want_f && !force) throw std::invalid_argument("P3M: force array required");
    if (energy) *energy = 0.;
    // long_range_kernel skips the mesh work without charges (p3m.cpp:463)
    if (!(ctx->sum_q2 > 0.)) return;
    hipStream_t const st = ctx->solver->stream();
    bool const host = !(flags & P3M_DEVICE_POINTERS);
    double *d_force = force;
    double const *d_upos = pos_unfolded;
    if (host && n > 0) {
      ensure_staging(ctx, n);
      if (want_f) {
        hip_check(hipMemcpyAsync(ctx->d_force, force, sizeof(double) * 3 * n,
                                 hipMemcpyHostToDevice, st),
                  "copy force");
        d_force = ctx->d_force;
      }
      if (pos_unfolded) {
        hip_check(hipMemcpyAsync(ctx->d_upos, pos_unfolded, sizeof(double) * 3 * n,
                                 hipMemcpyHostToDevice, st),
                  "copy upos");
        d_upos = ctx->d_upos;
      }
    }
    SolveResult r;
    ctx->solver->set_convention(0);
    ctx->solver->solve(n, ctx->d_cache_pos, ctx->d_cache_q, d_upos, want_f ? d_force : nullptr,
                       (flags & (P3M_FORCES | P3M_ENERGY)) | kSolveNoSpread, &r, true);
    ctx->mesh_assigned = false;
    if (want_f && host)
      hip_check(hipMemcpy(force, d_force, sizeof(double) * 3 * n, hipMemcpyDeviceToHost),
                "copy force");
    if (energy && (flags & P3M_ENERGY)) {
      // the energy corrections of long_range_kernel (p3m.cpp:551-589) with
      // the charge sums of count_charged_particles(_elc)
      auto const box = v3(ctx->user.box_l);
      double const volume = box[0] * box[1] * box[2];
      double const alpha = ctx->params.alpha;
      double e = r.kspace;
      e -= ctx->sum_q2 * alpha * std::numbers::inv_sqrtpi;
      e -= ctx->square_sum_q * std::numbers::pi / (2. * volume * alpha * alpha);
      if (ctx->params.epsilon != P3M_EPSILON_METALLIC) {
        double const pref = 4. * std::numbers::pi / volume / (2. * ctx->params.epsilon + 1.);
        e += pref * (r.dipole[0] * r.dipole[0] + r.dipole[1] * r.dipole[1] +
                     r.dipole[2] * r.dipole[2]);
      }
      *energy = e * ctx->user.prefactor;
    }
  });
}

int p3m_debug_charge_mesh(p3m_context *ctx, int64_t n, double const *pos, double const *q,
                          double *mesh_out) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_params(ctx);
    require_device(ctx);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->initialized) init_solver(ctx);
    DevArgs const d = stage_inputs(ctx, n, pos, q, nullptr, nullptr, 0);
    std::vector<double> out;
    ctx->solver->set_convention(ctx->force_conv);
    ctx->solver->debug_charge_mesh(n, d.pos, d.q, out);
    ctx->mesh_assigned = false;
    std::memcpy(mesh_out, out.data(), sizeof(double) * out.size());
  });
}

int p3m_debug_indices(p3m_context *ctx, int64_t n, double const *pos, double const *q,
                      int32_t *nmp) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_params(ctx);
    require_device(ctx);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->initialized) init_solver(ctx);
    DevArgs const d = stage_inputs(ctx, n, pos, q, nullptr, nullptr, 0);
    std::vector<int32_t> out;
    ctx->solver->set_convention(ctx->force_conv);
    ctx->solver->debug_indices(n, d.pos, d.q, out);
    ctx->mesh_assigned = false;
    std::memcpy(nmp, out.data(), sizeof(int32_t) * out.size());
  });
}

int p3m_k_space_error(double prefactor, int const mesh[3], int cao, int n_c_part,
                      double sum_q2, double alpha_L, double const box_l[3], double *out) {
  if (!out) return P3M_ERR_INVALID;
  return guarded(static_cast<std::string *>(nullptr), [&] {
    *out = k_space_error(prefactor, v3i(mesh), cao, n_c_part, sum_q2, alpha_L, v3(box_l));
  });
}

int p3m_real_space_error(double prefactor, double r_cut_iL, int n_c_part, double sum_q2,
                         double alpha_L, double const box_l[3], double *out) {
  if (!out) return P3M_ERR_INVALID;
  return guarded(static_cast<std::string *>(nullptr), [&] {
    *out = real_space_error(prefactor, r_cut_iL, n_c_part, sum_q2, alpha_L, v3(box_l));
  });
}

int p3m_pair_forces_targets(p3m_context *ctx, int64_t n, double const *pos, double const *q,
                            int ntargets, int32_t const *targets, double *out) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_params(ctx);
    require_device(ctx);
    if (!ctx->is_tuned) throw std::runtime_error("P3M: solver used before activation");
    if (n < 0 || ntargets < 0) throw std::invalid_argument("P3M: negative count");
    auto const box = v3(ctx->user.box_l);
    for (int d = 0; d < 3; ++d)
      if (!(ctx->params.r_cut < 0.5 * box[d]))
        throw std::domain_error("P3M: r_cut must be below half the box for minimum image");
    pair_forces_targets(ctx->device, ctx->solver->stream(), n, pos, q, ntargets, targets, box,
                        ctx->user.prefactor, ctx->params.alpha, ctx->params.r_cut, out);
  });
}

int p3m_short_range(p3m_context *ctx, int64_t n, double const *pos, double const *q,
                    unsigned flags, double *force, double *energy, double *virial) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_params(ctx);
    require_device(ctx);
    if (!ctx->is_tuned) throw std::runtime_error("P3M: solver used before activation");
    if (n < 0) throw std::invalid_argument("P3M: negative particle count");
    if (!force && n > 0) throw std::invalid_argument("P3M: force array required");
    bool const async = (flags & P3M_ASYNC) != 0;
    bool const want_ev = (flags & (P3M_ENERGY | P3M_PRESSURE)) != 0;
    if (async && !(flags & P3M_DEVICE_POINTERS))
      throw std::invalid_argument("P3M: P3M_ASYNC requires P3M_DEVICE_POINTERS");
    if (async && want_ev)
      throw std::invalid_argument("P3M: energy and pressure need a synchronous call");
    auto const box = v3(ctx->user.box_l);
    for (int d = 0; d < 3; ++d)
      if (!(ctx->params.r_cut < 0.5 * box[d]))
        throw std::domain_error("P3M: r_cut must be below half the box for minimum image");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->initialized) init_solver(ctx);
    if (!ctx->short_range) ctx->short_range = std::make_unique<ShortRange>(ctx->device);
    DevArgs const d = stage_inputs(ctx, n, pos, q, nullptr, force, flags);
    ShortRangeParams sp;
    for (int k = 0; k < 3; ++k) sp.box[k] = box[k];
    sp.prefactor = ctx->user.prefactor;
    sp.alpha = ctx->params.alpha;
    sp.r_cut = ctx->params.r_cut;
    sp.subdivision = ctx->short_range_subdivision;
    if (ctx->elc_on && (ctx->elc.delta_mid_top != 0. || ctx->elc.delta_mid_bot != 0.)) {
      ElcParams const ep = elc_geometry(ctx);
      sp.elc_images = true;
      sp.elc_h = ep.box[2] - ep.gap_size;
      sp.elc_sl = ep.space_layer;
      sp.elc_dt = ep.delta_top;
      sp.elc_db = ep.delta_bot;
    }
    ShortRangeResult r;
    hipStream_t const st = ctx->solver->stream();
    ctx->short_range->compute(st, sp, n, d.pos, d.q, d.force, want_ev ? &r : nullptr);
    if (!async && !want_ev) hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (!(flags & P3M_DEVICE_POINTERS) && n > 0)
      hip_check(hipMemcpy(force, d.force, sizeof(double) * 3 * n, hipMemcpyDeviceToHost),
                "copy force");
    if (energy) *energy = (flags & P3M_ENERGY) ? r.energy : 0.;
    if (virial)
      for (int k = 0; k < 9; ++k) virial[k] = (flags & P3M_PRESSURE) ? r.virial[k] : 0.;
  });
}

int p3m_set_force_convention(p3m_context *ctx, int convention) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    if (convention != P3M_CONV_CPU && convention != P3M_CONV_GPU)
      throw std::invalid_argument("P3M: force convention must be P3M_CONV_CPU or P3M_CONV_GPU");
    ctx->force_conv = convention;
  });
}

int p3m_elc_params_default(p3m_elc_params *p) {
  if (!p) return P3M_ERR_INVALID;
  std::memset(p, 0, sizeof(*p));
  p->maxPWerror = -1.;
  p->gap_size = -1.;
  p->far_cut = -1.;
  p->neutralize = 1;
  return P3M_OK;
}

int p3m_set_elc(p3m_context *ctx, p3m_elc_params const *p) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    if (!p) {
      ctx->elc_on = false;
      return;
    }
    // elc_data constructor checks, in order (elc.cpp:1056-1102)
    double const round_error_prec = 1e-14;
    bool const far_calculated = (p->far_cut == -1.);
    bool const dielectric = (p->delta_mid_top != 0. || p->delta_mid_bot != 0.);
    bool const const_pot = p->const_pot != 0 && dielectric;
    double const delta_range = 1. + std::sqrt(round_error_prec);
    if (p->far_cut <= 0. && !far_calculated)
      throw std::domain_error("Parameter 'far_cut' must be > 0");
    if (p->maxPWerror <= 0.) throw std::domain_error("Parameter 'maxPWerror' must be > 0");
    if (p->gap_size <= 0.) throw std::domain_error("Parameter 'gap_size' must be > 0");
    if (p->pot_diff != 0. && !p->const_pot)
      throw std::invalid_argument(
          "Parameter 'const_pot' must be True when 'pot_diff' is non-zero");
    if (p->delta_mid_top < -delta_range || p->delta_mid_top > delta_range)
      throw std::domain_error("Parameter 'delta_mid_top' must be >= -1 and <= +1");
    if (p->delta_mid_bot < -delta_range || p->delta_mid_bot > delta_range)
      throw std::domain_error("Parameter 'delta_mid_bot' must be >= -1 and <= +1");
    double const dt = std::clamp(p->delta_mid_top, -1., 1.);
    double const db = std::clamp(p->delta_mid_bot, -1., 1.);
    if (dielectric && !const_pot && std::fabs(1. - dt * db) < round_error_prec)
      throw std::domain_error("ELC with two parallel metallic boundaries requires the "
                              "const_pot option");
    ctx->elc = *p;
    ctx->elc.delta_mid_top = dt;
    ctx->elc.delta_mid_bot = db;
    ctx->elc.const_pot = const_pot ? 1 : 0;
    ctx->elc.neutralize = (p->neutralize != 0 && !dielectric) ? 1 : 0;
    ctx->elc.pot_diff = p->const_pot ? p->pot_diff : 0.;
    ctx->elc_far_cut = far_calculated ? -1. : p->far_cut;
    ctx->elc_on = true;
    // adapt_epsilon_elc (p3m.impl.hpp:112-114): P3M under ELC is metallic
    if (ctx->params.epsilon != P3M_EPSILON_METALLIC) {
      ctx->params.epsilon = P3M_EPSILON_METALLIC;
      ctx->user.epsilon = P3M_EPSILON_METALLIC;
      ctx->initialized = false;
    }
  });
}

int p3m_get_elc(p3m_context const *ctx, p3m_elc_params *out, int *enabled) {
  if (!ctx || !out || !enabled) return P3M_ERR_INVALID;
  *enabled = ctx->elc_on ? 1 : 0;
  *out = ctx->elc;
  if (ctx->elc_on) out->far_cut = ctx->elc_far_cut;
  return P3M_OK;
}

int p3m_set_tuning_real_space(p3m_context *ctx, int mode) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    if (mode != P3M_TUNE_RS_MODEL && mode != P3M_TUNE_RS_DEVICE)
      throw std::invalid_argument(
          "P3M: tuning real space must be P3M_TUNE_RS_MODEL or P3M_TUNE_RS_DEVICE");
    ctx->tune_real_space = mode;
  });
}

int p3m_get_force_convention(p3m_context const *ctx, int *convention) {
  if (!ctx || !convention) return P3M_ERR_INVALID;
  *convention = ctx->force_conv;
  return P3M_OK;
}

int p3m_set_short_range_subdivision(p3m_context *ctx, int cells_per_cutoff) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    if (cells_per_cutoff < 0 || cells_per_cutoff > 2)
      throw std::invalid_argument("P3M: short-range subdivision must be 0, 1 or 2");
    ctx->short_range_subdivision = cells_per_cutoff;
  });
}

int p3m_short_range_cells(p3m_context const *ctx, int cells[3]) {
  if (!ctx || !cells) return P3M_ERR_INVALID;
  if (ctx->short_range)
    ctx->short_range->cells(cells);
  else
    cells[0] = cells[1] = cells[2] = 0;
  return P3M_OK;
}

int p3m_pair_force(p3m_context const *ctx, double q1q2, double const d[3], double dist,
                   double out[3]) {
  if (!ctx || !out) return P3M_ERR_INVALID;
  auto const f = pair_force(ctx->user.prefactor, ctx->params.alpha, ctx->params.r_cut, q1q2,
                            v3(d), dist);
  out[0] = f[0];
  out[1] = f[1];
  out[2] = f[2];
  return P3M_OK;
}

int p3m_pair_energy(p3m_context const *ctx, double q1q2, double dist, double *out) {
  if (!ctx || !out) return P3M_ERR_INVALID;
  *out = pair_energy(ctx->user.prefactor, ctx->params.alpha, ctx->params.r_cut, q1q2, dist);
  return P3M_OK;
}

int p3m_last_timings(p3m_context const *ctx, double *ms, int n) {
  if (!ctx || !ms || !ctx->solver) return P3M_ERR_INVALID;
  for (int i = 0; i < n && i < 8; ++i) ms[i] = ctx->solver->last_ms[i];
  return P3M_OK;
}

int p3m_profile_begin(p3m_context *ctx, int max_solves) {
  if (!ctx || !ctx->solver) return P3M_ERR_INVALID;
  return guarded(ctx, [&] { ctx->solver->profile_begin(max_solves); });
}

int p3m_profile_end(p3m_context *ctx, double *stage_ms, int *nsolves) {
  if (!ctx || !ctx->solver || !stage_ms) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    int const k = ctx->solver->profile_end(stage_ms);
    if (nsolves) *nsolves = k;
  });
}

int p3m_autotune(p3m_context *ctx, int64_t n, double const *pos, double const *q, int reps,
                 p3m_launch_shape *best, double *best_ms, double *default_ms) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_params(ctx);
    require_device(ctx);
    if (!ctx->is_tuned) throw std::runtime_error("P3M: solver used before activation");
    if (n < 0) throw std::invalid_argument("P3M: negative particle count");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->initialized) init_solver(ctx);
    ensure_staging(ctx, n);
    AutotuneResult const r = ctx->solver->autotune(n, pos, q, ctx->d_force, reps);
    ctx->mesh_assigned = false;
    if (best) *best = to_abi(r.best);
    if (best_ms) *best_ms = r.best_ms;
    if (default_ms) *default_ms = r.default_ms;
  });
}

int p3m_set_launch_shape(p3m_context *ctx, p3m_launch_shape const *shape) {
  if (!ctx || !shape) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_device(ctx);
    LaunchShape const s = from_abi(*shape);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ctx->solver->set_launch_shape(s);
    ctx->mesh_assigned = false;
  });
}

int p3m_set_deterministic(p3m_context *ctx, int on) {
  if (!ctx) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    require_device(ctx);
    if (on && ctx->user.precision != P3M_PREC_FP64)
      throw std::invalid_argument("P3M: deterministic solves need precision P3M_PREC_FP64");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ctx->solver->set_deterministic(on != 0);
  });
}

int p3m_get_launch_shape(p3m_context const *ctx, p3m_launch_shape *shape) {
  if (!ctx || !shape || !ctx->solver) return P3M_ERR_INVALID;
  *shape = to_abi(ctx->solver->launch_shape());
  return P3M_OK;
}

int p3m_synchronize(p3m_context *ctx) {
  if (!ctx || !ctx->solver) return P3M_ERR_INVALID;
  return guarded(ctx, [&] {
    hip_check(hipStreamSynchronize(ctx->solver->stream()), "hipStreamSynchronize");
  });
}

int p3m_get_stream(p3m_context const *ctx, void **stream) {
  if (!ctx || !stream || !ctx->solver) return P3M_ERR_INVALID;
  *stream = static_cast<void *>(ctx->solver->stream());
  return P3M_OK;
}

char const *p3m_last_error(p3m_context const *ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

char const *p3m_last_global_error(void) {
  std::lock_guard<std::mutex> lock(g_err_mutex);
  static thread_local std::string copy;
  copy = g_last_error;
  return copy.c_str();
}

} // extern "C"

/* ---------------------------------------------------------------- batch */

namespace {
/** Fixed (already tuned) parameters of one batch system, validated like
 *  p3m_set_params followed by the post-tuning sanity checks. */
P3MParams fixed_params(p3m_params const *p) {
  P3MParams params(false, p->epsilon, p->r_cut, v3i(p->mesh), v3(p->mesh_off), p->cao,
                   p->alpha, p->accuracy);
  if (p->prefactor <= 0.) throw std::domain_error("Parameter 'prefactor' must be > 0");
  for (int d = 0; d < 3; ++d) {
    if (!(p->box_l[d] > 0.)) throw std::domain_error("box_l must be > 0");
    if (params.mesh[d] < 1)
      throw std::domain_error("P3M batch: every system needs a fixed mesh");
  }
  auto const box = v3(p->box_l);
  params.cao3 = params.cao * params.cao * params.cao;
  params.recalc_a_ai_cao_cut(box);
  sanity_checks_boxl(params, box);
  params.alpha_L = params.alpha * box[0];
  params.r_cut_iL = params.r_cut * (1. / box[0]);
  return params;
}

std::vector<int> selection(p3m_batch const *b, int nsel, int const *ids) {
  if (nsel < 0) throw std::invalid_argument("P3M batch: negative selection size");
  if (nsel > 0 && !ids) throw std::invalid_argument("P3M batch: null system id list");
  std::vector<int> sel(ids, ids + nsel);
  int const ns = b->solver->num_systems();
  std::vector<char> seen(static_cast<size_t>(ns), 0);
  for (int s : sel) {
    if (s < 0 || s >= ns)
      throw std::invalid_argument("P3M batch: invalid system id " + std::to_string(s));
    if (seen[s]) throw std::invalid_argument("P3M batch: system selected twice");
    seen[s] = 1;
  }
  return sel;
}
} // namespace

extern "C" {

int p3m_batch_create(int device, int precision, p3m_batch **out) {
  if (!out) return P3M_ERR_INVALID;
  *out = nullptr;
  return guarded(static_cast<std::string *>(nullptr), [&] {
    if (precision != P3M_PREC_FP64 && precision != P3M_PREC_FP32)
      throw std::invalid_argument("P3M batch: unknown precision");
    int ndev = 0;
    hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (device < 0 || device >= ndev)
      throw std::invalid_argument("P3M: invalid HIP device " + std::to_string(device));
    auto b = std::make_unique<p3m_batch>();
    b->device = device;
    b->solver = std::make_unique<BatchSolver>(
        device, precision == P3M_PREC_FP32 ? MeshPrecision::FP32 : MeshPrecision::FP64);
    *out = b.release();
  });
}

int p3m_batch_destroy(p3m_batch *b) {
  if (!b) return P3M_OK;
  (void)hipSetDevice(b->device);
  delete b;
  return P3M_OK;
}

int p3m_batch_add_system(p3m_batch *b, p3m_params const *p, int64_t n, int *sys_id) {
  if (!b || !p) return P3M_ERR_INVALID;
  return guarded(&b->last_error, [&] {
    if (n < 0) throw std::invalid_argument("P3M: negative particle count");
    P3MParams const params = fixed_params(p);
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    int const id = b->solver->add_system(params, v3(p->box_l), p->skin, p->prefactor, n);
    if (sys_id) *sys_id = id;
  });
}

int p3m_batch_particle_offset(p3m_batch const *b, int sys_id, int64_t *offset) {
  if (!b || !offset || sys_id < 0 || sys_id >= b->solver->num_systems())
    return P3M_ERR_INVALID;
  *offset = b->solver->particle_offset(sys_id);
  return P3M_OK;
}

int p3m_batch_system_bytes(p3m_batch const *b, int sys_id, int64_t *bytes) {
  if (!b || !bytes || sys_id < 0 || sys_id >= b->solver->num_systems())
    return P3M_ERR_INVALID;
  *bytes = static_cast<int64_t>(b->solver->system_bytes(sys_id));
  return P3M_OK;
}

int p3m_batch_working_bytes(p3m_batch const *b, int nsel, int const *sys_ids, int64_t *bytes) {
  if (!b || !bytes) return P3M_ERR_INVALID;
  return guarded(static_cast<std::string *>(nullptr), [&] {
    auto const sel = selection(b, nsel, sys_ids);
    *bytes = static_cast<int64_t>(b->solver->working_bytes(sel));
  });
}

int p3m_batch_solve(p3m_batch *b, int nsel, int const *sys_ids, double const *pos,
                    double const *q, double *force, unsigned flags, double *energy,
                    double *virial) {
  if (!b) return P3M_ERR_INVALID;
  return guarded(&b->last_error, [&] {
    auto const sel = selection(b, nsel, sys_ids);
    bool const async = (flags & P3M_ASYNC) != 0;
    if ((flags & P3M_FORCES) && !force && b->solver->total_particles() > 0)
      throw std::invalid_argument("P3M: force array required");
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    ParticleArrays const pa{pos, q, nullptr, (flags & P3M_FORCES) ? force : nullptr};
    std::vector<SolveResult> res;
    b->solver->solve(sel, pa, flags, !async, async ? nullptr : &res);
    for (int i = 0; i < nsel; ++i) {
      bool const have = !async && static_cast<size_t>(i) < res.size();
      if (energy) energy[i] = (have && (flags & P3M_ENERGY)) ? res[i].energy : 0.;
      if (virial)
        for (int k = 0; k < 9; ++k)
          virial[9 * i + k] = (have && (flags & P3M_PRESSURE)) ? res[i].virial[k] : 0.;
    }
  });
}

int p3m_batch_graph_solve(p3m_batch *b, int nsel, int const *sys_ids, double const *pos,
                          double const *q, double *force) {
  if (!b) return P3M_ERR_INVALID;
  return guarded(&b->last_error, [&] {
    auto const sel = selection(b, nsel, sys_ids);
    if (!force && b->solver->total_particles() > 0)
      throw std::invalid_argument("P3M: force array required");
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    b->solver->graph_solve(sel, ParticleArrays{pos, q, nullptr, force}, P3M_FORCES);
  });
}

int p3m_batch_set_launch_shape(p3m_batch *b, p3m_launch_shape const *shape) {
  if (!b) return P3M_ERR_INVALID;
  return guarded(&b->last_error, [&] {
    if (!shape) throw std::invalid_argument("P3M batch: shape required");
    LaunchShape const s = from_abi(*shape);
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    b->solver->set_launch_shape(s);
  });
}

int p3m_batch_get_launch_shape(p3m_batch const *b, p3m_launch_shape *shape) {
  if (!b || !shape) return P3M_ERR_INVALID;
  *shape = to_abi(b->solver->launch_shape());
  return P3M_OK;
}

int p3m_batch_set_deterministic(p3m_batch *b, int on) {
  if (!b) return P3M_ERR_INVALID;
  return guarded(&b->last_error, [&] {
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    b->solver->set_deterministic(on != 0);
  });
}

int p3m_batch_autotune(p3m_batch *b, int nsel, int const *sys_ids, double const *pos,
                       double const *q, double *force, int reps, p3m_launch_shape *best,
                       double *best_ms, double *default_ms) {
  if (!b) return P3M_ERR_INVALID;
  return guarded(&b->last_error, [&] {
    auto const sel = selection(b, nsel, sys_ids);
    if (!force && b->solver->total_particles() > 0)
      throw std::invalid_argument("P3M: force array required");
    if (sel.empty()) throw std::invalid_argument("P3M batch: empty selection");
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    AutotuneResult const r =
        b->solver->autotune(sel, ParticleArrays{pos, q, nullptr, force}, reps);
    if (best) *best = to_abi(r.best);
    if (best_ms) *best_ms = r.best_ms;
    if (default_ms) *default_ms = r.default_ms;
  });
}

int p3m_batch_synchronize(p3m_batch *b) {
  if (!b) return P3M_ERR_INVALID;
  return guarded(&b->last_error, [&] {
    hip_check(hipStreamSynchronize(b->solver->stream()), "hipStreamSynchronize");
  });
}

int p3m_batch_get_stream(p3m_batch const *b, void **stream) {
  if (!b || !stream) return P3M_ERR_INVALID;
  *stream = static_cast<void *>(b->solver->stream());
  return P3M_OK;
}

int p3m_batch_profile_begin(p3m_batch *b, int max_solves) {
  if (!b) return P3M_ERR_INVALID;
  return guarded(&b->last_error, [&] { b->solver->profile_begin(max_solves); });
}

int p3m_batch_profile_end(p3m_batch *b, double *stage_ms, int *nsolves) {
  if (!b || !stage_ms) return P3M_ERR_INVALID;
  return guarded(&b->last_error, [&] {
    int const k = b->solver->profile_end(stage_ms);
    if (nsolves) *nsolves = k;
  });
}

char const *p3m_batch_last_error(p3m_batch const *b) {
  return b ? b->last_error.c_str() : "null batch";
}

int p3m_serve_create(p3m_batch *b, int lanes, int64_t budget_bytes, int max_batch,
                     double sigma, p3m_server **out) {
  if (!b || !out) return P3M_ERR_INVALID;
  *out = nullptr;
  return guarded(&b->last_error, [&] {
    if (budget_bytes <= 0) throw std::invalid_argument("P3M serve: budget must be > 0");
    if (!(sigma >= 0.)) throw std::invalid_argument("P3M serve: sigma must be >= 0");
    ServeConfig cfg;
    cfg.lanes = lanes;
    cfg.budget_bytes = static_cast<size_t>(budget_bytes);
    cfg.max_batch = max_batch;
    cfg.sigma = sigma;
    auto s = std::make_unique<p3m_server>();
    s->device = b->device;
    s->solver = b->solver;
    s->server = std::make_unique<Server>(*s->solver, cfg);
    *out = s.release();
  });
}

int p3m_serve_destroy(p3m_server *s) {
  delete s;
  return P3M_OK;
}

int p3m_serve_run(p3m_server *s, int n_req, double const *arrival_s, int const *sys_ids,
                  uint64_t const *seeds, double *pos, double const *q, double *force,
                  double *force_out, int64_t const *out_off, double *latency_s, double *wall_s,
                  int *n_batches, int64_t *peak_bytes, int *peak_lanes) {
  if (!s) return P3M_ERR_INVALID;
  return guarded(&s->last_error, [&] {
    if (n_req < 0 || (n_req > 0 && (!arrival_s || !sys_ids || !seeds)))
      throw std::invalid_argument("P3M serve: request arrays required");
    if (force_out && !out_off) throw std::invalid_argument("P3M serve: out_off required");
    if (s->solver->total_particles() > 0 && (!pos || !q || !force))
      throw std::invalid_argument("P3M serve: particle arrays required");
    hip_check(hipSetDevice(s->device), "hipSetDevice");
    std::vector<ServeRequest> reqs(static_cast<size_t>(n_req));
    for (int r = 0; r < n_req; ++r) reqs[r] = ServeRequest{arrival_s[r], sys_ids[r], seeds[r]};
    ServeStats const st = s->server->run(reqs, pos, q, force, force_out, out_off);
    if (latency_s)
      for (int r = 0; r < n_req; ++r) latency_s[r] = st.latency_s[r];
    if (wall_s) *wall_s = st.wall_s;
    if (n_batches) *n_batches = st.batches;
    if (peak_bytes) *peak_bytes = static_cast<int64_t>(st.peak_bytes);
    if (peak_lanes) *peak_lanes = st.peak_lanes;
  });
}

char const *p3m_serve_last_error(p3m_server const *s) {
  return s ? s->last_error.c_str() : "null server";
}

int p3m_tune_model(p3m_params *inout, int n_charged, double sum_q2, int mesh_min,
                   int mesh_max, int cao_min, int cao_max, double pair_ns,
                   int single_precision) {
  if (!inout) return P3M_ERR_INVALID;
  return guarded(static_cast<std::string *>(nullptr), [&] {
    P3MParams params(true, inout->epsilon, inout->r_cut, v3i(inout->mesh), v3(inout->mesh_off),
                     inout->cao, inout->alpha, inout->accuracy);
    if (inout->prefactor <= 0.) throw std::domain_error("Parameter 'prefactor' must be > 0");
    auto const box = v3(inout->box_l);
    for (int d = 0; d < 3; ++d)
      if (!(box[d] > 0.)) throw std::domain_error("box_l must be > 0");
    if (n_charged <= 0)
      throw std::runtime_error("CoulombP3M: no charged particles in the system");
    if (mesh_min < 1 || mesh_max < mesh_min || cao_min < 1 || cao_max > 7 || cao_max < cao_min)
      throw std::invalid_argument("P3M: invalid tuning ranges");
    TuningInput in;
    in.box_l = box;
    in.skin = inout->skin;
    in.prefactor = inout->prefactor;
    in.sum_qpart = n_charged;
    in.sum_q2 = sum_q2;
    in.timings = 1;
    in.verbose = inout->verbose != 0;
    in.mesh_min = mesh_min;
    in.mesh_max = mesh_max;
    in.cao_min = cao_min;
    in.cao_max = cao_max;
    P3MParams const base = params;
    TuningTimer timer = [&](Vec3i const &mesh, int cao, double r_cut_iL, double alpha_L,
                            int) -> double {
      P3MParams trial = base;
      trial.mesh = mesh;
      trial.cao = cao;
      trial.alpha_L = alpha_L;
      trial.alpha = alpha_L * (1. / box[0]);
      trial.r_cut_iL = r_cut_iL;
      trial.r_cut = r_cut_iL * box[0];
      trial.recalc_a_ai_cao_cut(box);
      try {
        sanity_checks_boxl(trial, box);
      } catch (std::runtime_error const &) {
        return -1.;
      }
      return model_step_ms(mesh, cao, trial.r_cut, n_charged, box, pair_ns > 0. ? pair_ns : 1.,
                           single_precision != 0);
    };
    tune_p3m(params, in, timer);
    for (int d = 0; d < 3; ++d) {
      inout->mesh[d] = params.mesh[d];
      inout->mesh_off[d] = params.mesh_off[d];
    }
    inout->cao = params.cao;
    inout->alpha = params.alpha;
    inout->r_cut = params.r_cut;
    inout->accuracy = params.accuracy;
    inout->tune = 0;
  });
}


} // extern "C"

// ---------------------------------------------------------------------------
// dipolar P3M (DipolarP3MImpl, src/core/magnetostatics/dp3m.cpp)
// ---------------------------------------------------------------------------
namespace {
void dp_free_staging(p3m_dipolar *c) {
  for (double **p : {&c->d_pos, &c->d_dip, &c->d_f, &c->d_t}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  c->cap = 0;
}

struct DpArgs {
  double const *pos, *dip;
  double *f, *t;
};

DpArgs dp_stage(p3m_dipolar *c, int64_t n, double const *pos, double const *dip, double *f,
                double *t, unsigned flags) {
  if (n > 0 && (!pos || !dip)) throw std::invalid_argument("DipolarP3M: null particle arrays");
  // the device sorts index particles with 32-bit integers (ranks times the
  // number of counter copies included)
  if (n > INT32_MAX - 8191)
    throw std::invalid_argument("DipolarP3M: more than 2^31 - 8192 particles");
  if (flags & P3M_DEVICE_POINTERS) return {pos, dip, f, t};
  if (n > c->cap) {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    dp_free_staging(c);
    for (double **p : {&c->d_pos, &c->d_dip, &c->d_f, &c->d_t})
      hip_check(hipMalloc(p, sizeof(double) * 3 * static_cast<size_t>(n)), "hipMalloc staging");
    c->cap = n;
  }
  size_t const bytes = sizeof(double) * 3 * static_cast<size_t>(n);
  if (n > 0) {
    hip_check(hipMemcpy(c->d_pos, pos, bytes, hipMemcpyHostToDevice), "copy pos");
    hip_check(hipMemcpy(c->d_dip, dip, bytes, hipMemcpyHostToDevice), "copy dip");
    if (f) hip_check(hipMemcpy(c->d_f, f, bytes, hipMemcpyHostToDevice), "copy force");
    if (t) hip_check(hipMemcpy(c->d_t, t, bytes, hipMemcpyHostToDevice), "copy torque");
  }
  return {c->d_pos, c->d_dip, f ? c->d_f : nullptr, t ? c->d_t : nullptr};
}

void dp_unstage(p3m_dipolar *c, int64_t n, DpArgs const &d, double *f, double *t,
                unsigned flags) {
  if ((flags & P3M_DEVICE_POINTERS) || n <= 0) return;
  size_t const bytes = sizeof(double) * 3 * static_cast<size_t>(n);
  if (f) hip_check(hipMemcpy(f, d.f, bytes, hipMemcpyDeviceToHost), "copy force");
  if (t) hip_check(hipMemcpy(t, d.t, bytes, hipMemcpyDeviceToHost), "copy torque");
  (void)c;
}

DipolarParams dp_params(p3m_dipolar_params const &u) {
  DipolarParams p;
  p.mesh = u.mesh;
  p.cao = u.cao;
  p.alpha = u.alpha;
  p.r_cut = u.r_cut;
  p.accuracy = u.accuracy;
  p.prefactor = u.prefactor;
  p.epsilon = u.epsilon;
  for (int d = 0; d < 3; ++d) p.mesh_off[d] = u.mesh_off[d];
  p.box = u.box_l;
  p.skin = u.skin;
  return p;
}
} // namespace

extern "C" {

int p3m_dipolar_params_default(p3m_dipolar_params *p) {
  if (!p) return P3M_ERR_INVALID;
  std::memset(p, 0, sizeof(*p));
  p->mesh = -1;
  p->cao = -1;
  p->alpha = -1.;
  p->r_cut = -1.;
  p->accuracy = -1.;
  p->mesh_off[0] = p->mesh_off[1] = p->mesh_off[2] = 0.5;
  p->skin = 0.4;
  p->tune = 1;
  p->timings = 10;
  return P3M_OK;
}

int p3m_dipolar_create(int device, p3m_dipolar **out) {
  if (!out) return P3M_ERR_INVALID;
  *out = nullptr;
  return guarded(static_cast<std::string *>(nullptr), [&] {
    auto c = std::make_unique<p3m_dipolar>();
    c->device = device;
    if (device >= 0) {
      int count = 0;
      hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
      if (device >= count) throw std::invalid_argument("P3M: invalid device id");
      c->solver = std::make_unique<DipolarSolver>(device);
    }
    p3m_dipolar_params_default(&c->user);
    *out = c.release();
  });
}

int p3m_dipolar_destroy(p3m_dipolar *c) {
  if (!c) return P3M_OK;
  if (c->device >= 0) (void)hipSetDevice(c->device);
  dp_free_staging(c);
  delete c;
  return P3M_OK;
}

char const *p3m_dipolar_last_error(p3m_dipolar const *c) {
  return c ? c->last_error.c_str() : "";
}

int p3m_dipolar_set_params(p3m_dipolar *c, p3m_dipolar_params const *p) {
  if (!c || !p) return P3M_ERR_INVALID;
  return guarded(&c->last_error, [&] {
    // DipolarP3MImpl constructor and P3MParameters checks (dp3m.impl.hpp:79-92,
    // common.hpp:94-150)
    P3MParams check(p->tune != 0, p->epsilon, p->r_cut, Vec3i{p->mesh, p->mesh, p->mesh},
                    Vec3d{p->mesh_off[0], p->mesh_off[1], p->mesh_off[2]}, p->cao, p->alpha,
                    p->accuracy);
    (void)check;
    if (p->timings <= 0) throw std::domain_error("Parameter 'timings' must be > 0");
    if (p->prefactor <= 0.) throw std::domain_error("Parameter 'prefactor' must be > 0");
    if (!(p->box_l > 0.)) throw std::domain_error("box_l must be > 0");
    c->user = *p;
    c->have_params = true;
    c->is_tuned = (p->tune == 0);
  });
}

int p3m_dipolar_get_params(p3m_dipolar const *c, p3m_dipolar_params *out) {
  if (!c || !out) return P3M_ERR_INVALID;
  *out = c->user;
  out->tune = c->is_tuned ? 0 : 1;
  return P3M_OK;
}

int p3m_dipolar_activate(p3m_dipolar *c, int64_t n, double const *pos, double const *dip,
                         unsigned flags) {
  if (!c) return P3M_ERR_INVALID;
  return guarded(&c->last_error, [&] {
    if (!c->have_params) throw std::runtime_error("DipolarP3M: parameters not set");
    if (!c->solver) throw std::runtime_error("DipolarP3M: host-only context cannot solve");
    if (n < 0) throw std::invalid_argument("DipolarP3M: negative particle count");
    (void)pos; // the cost-model tuning needs the moments only
    // count_magnetic_particles (dp3m.cpp:83-97)
    std::vector<double> h(3 * static_cast<size_t>(n));
    if (n > 0) {
      if (flags & P3M_DEVICE_POINTERS)
        hip_check(hipMemcpy(h.data(), dip, sizeof(double) * h.size(), hipMemcpyDeviceToHost),
                  "copy dip");
      else
        std::memcpy(h.data(), dip, sizeof(double) * h.size());
    }
    int n_dip = 0;
    double mu2 = 0.;
    for (int64_t i = 0; i < n; ++i) {
      double const m2 = h[3 * i] * h[3 * i] + h[3 * i + 1] * h[3 * i + 1] +
                        h[3 * i + 2] * h[3 * i + 2];
      if (m2 != 0.) {
        ++n_dip;
        mu2 += m2;
      }
    }
    auto &u = c->user;
    double const L = u.box_l;
    if (!c->is_tuned) {
      if (n_dip == 0) throw std::runtime_error("DipolarP3M: no dipolar particles in the system");
      // tuning (dp3m.cpp:533-682): for each (mesh, cao) the smallest r_cut
      // meeting the accuracy (bisection to 1e-3 L, TuningAlgorithm.cpp:116-172)
      // with alpha from the real-space error; the (mesh, cao) pair is chosen
      // by a cost model in place of the reference's timings
      int m_lo = u.mesh, m_hi = u.mesh;
      if (u.mesh == -1) {
        m_lo = static_cast<int>(std::round(std::pow(2., std::floor(std::log(std::cbrt(
                                                               static_cast<double>(n_dip))) /
                                                           std::log(2.)))));
        m_lo = std::max(2, m_lo + (m_lo % 2));
        m_hi = 128;
      }
      int c_lo = (u.cao == -1) ? 1 : u.cao, c_hi = (u.cao == -1) ? 7 : u.cao;
      double const r_max_iL = (L / 2. - u.skin) / L;
      double best = -1., best_r = 0., best_a = 0., best_acc = 0.;
      int best_m = 0, best_c = 0;
      double const density = n_dip / (L * L * L);
      for (int M = m_lo; M <= m_hi; M += 2)
        for (int cao = c_lo; cao <= c_hi; ++cao) {
          double out[4];
          double r_iL;
          if (u.r_cut > 0.) {
            r_iL = u.r_cut / L;
            dp3m_calculate_accuracy(L, M, cao, r_iL, n_dip, mu2, u.accuracy, out);
            if (out[0] > u.accuracy) continue;
          } else {
            dp3m_calculate_accuracy(L, M, cao, r_max_iL, n_dip, mu2, u.accuracy, out);
            if (out[0] > u.accuracy) continue;
            double lo = 0., hi = r_max_iL;
            while (hi - lo >= 1e-3) {
              double const mid = 0.5 * (lo + hi);
              dp3m_calculate_accuracy(L, M, cao, mid, n_dip, mu2, u.accuracy, out);
              if (out[0] > u.accuracy)
                lo = mid;
              else
                hi = mid;
            }
            r_iL = hi;
            dp3m_calculate_accuracy(L, M, cao, r_iL, n_dip, mu2, u.accuracy, out);
          }
          double const rc = r_iL * L;
          double const cost = 2e-9 * n_dip * cao * cao * cao +
                              5e-9 * static_cast<double>(M) * M * M +
                              1e-9 * n_dip * density * 4.18879 * rc * rc * rc;
          if (best < 0. || cost < best) {
            best = cost;
            best_m = M;
            best_c = cao;
            best_r = rc;
            best_a = (u.alpha > 0.) ? u.alpha : out[3] / L;
            best_acc = out[0];
          }
        }
      if (best < 0.) throw std::runtime_error("DipolarP3M: failed to reach requested accuracy");
      u.mesh = best_m;
      u.cao = best_c;
      u.r_cut = best_r;
      u.alpha = best_a;
      u.accuracy = best_acc;
      c->is_tuned = true;
    }
    // sanity_checks_boxl (dp3m.cpp:822-846)
    double const cao_cut = 0.5 * (L / u.mesh) * u.cao;
    if (cao_cut >= 0.5 * L) {
      std::ostringstream msg;
      msg << "dipolar P3M_init: k-space cutoff " << cao_cut
          << " is larger than half of box dimension " << L;
      throw std::runtime_error(msg.str());
    }
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    c->solver->configure(dp_params(u));
  });
}

int p3m_dipolar_solve(p3m_dipolar *c, int64_t n, double const *pos, double const *dip,
                      double *force, double *torque, unsigned flags, double *energy) {
  if (!c) return P3M_ERR_INVALID;
  return guarded(&c->last_error, [&] {
    if (!c->solver || !c->solver->configured())
      throw std::runtime_error("DipolarP3M: solver used before activation");
    if (n < 0) throw std::invalid_argument("DipolarP3M: negative particle count");
    bool const async = (flags & P3M_ASYNC) != 0;
    if (async && !(flags & P3M_DEVICE_POINTERS))
      throw std::invalid_argument("P3M: P3M_ASYNC requires P3M_DEVICE_POINTERS");
    bool const wf = (flags & P3M_FORCES) != 0;
    DpArgs const d = dp_stage(c, n, pos, dip, wf ? force : nullptr, wf ? torque : nullptr, flags);
    double e = 0.;
    bool const we = (flags & P3M_ENERGY) && !async;
    c->solver->solve(n, d.pos, d.dip, d.f, d.t, we ? &e : nullptr, !async);
    dp_unstage(c, n, d, wf ? force : nullptr, wf ? torque : nullptr, flags);
    if (energy) *energy = we ? e : 0.;
  });
}

int p3m_dipolar_short_range(p3m_dipolar *c, int64_t n, double const *pos, double const *dip,
                            double *force, double *torque, unsigned flags, double *energy) {
  if (!c) return P3M_ERR_INVALID;
  return guarded(&c->last_error, [&] {
    if (!c->solver || !c->solver->configured())
      throw std::runtime_error("DipolarP3M: solver used before activation");
    if (n < 0) throw std::invalid_argument("DipolarP3M: negative particle count");
    bool const async = (flags & P3M_ASYNC) != 0;
    if (async && !(flags & P3M_DEVICE_POINTERS))
      throw std::invalid_argument("P3M: P3M_ASYNC requires P3M_DEVICE_POINTERS");
    if (!(c->user.r_cut < 0.5 * c->user.box_l))
      throw std::domain_error("P3M: r_cut must be below half the box for minimum image");
    DpArgs const d = dp_stage(c, n, pos, dip, force, torque, flags);
    double e = 0.;
    bool const we = (flags & P3M_ENERGY) && !async;
    c->solver->short_range(n, d.pos, d.dip, d.f, d.t, we ? &e : nullptr, !async);
    dp_unstage(c, n, d, force, torque, flags);
    if (energy) *energy = we ? e : 0.;
  });
}

int p3m_dipolar_set_short_range_subdivision(p3m_dipolar *c, int cells_per_cutoff) {
  if (!c) return P3M_ERR_INVALID;
  return guarded(&c->last_error, [&] {
    if (!c->solver) throw std::runtime_error("DipolarP3M: host-only context cannot solve");
    c->solver->set_short_range_subdivision(cells_per_cutoff);
  });
}

int p3m_dipolar_get_stream(p3m_dipolar const *c, void **stream) {
  if (!c || !stream || !c->solver) return P3M_ERR_INVALID;
  *stream = reinterpret_cast<void *>(c->solver->stream());
  return P3M_OK;
}

int p3m_dipolar_synchronize(p3m_dipolar *c) {
  if (!c || !c->solver) return P3M_ERR_INVALID;
  return guarded(&c->last_error, [&] {
    hip_check(hipStreamSynchronize(c->solver->stream()), "hipStreamSynchronize");
  });
}

} // extern "C"
