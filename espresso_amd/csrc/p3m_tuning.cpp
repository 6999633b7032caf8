// P3M tuning; see p3m_tuning.hpp.
#include "p3m_tuning.hpp"
#include "p3m_batch.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <numbers>
#include <stdexcept>

namespace espresso_amd {

namespace {
constexpr double P3M_TUNE_CAO_TOO_LARGE = 1.;
constexpr double P3M_TUNE_ACCURACY_TOO_LARGE = 3.;
constexpr double P3M_RCUT_PREC = 1e-3;
constexpr double time_granularity = 2.;
constexpr std::size_t max_n_consecutive_trials = 20;
constexpr double time_sentinel = std::numeric_limits<double>::max();

struct Trial {
  Vec3i mesh{};
  int cao = -1;
  double alpha_L = -1.;
  double r_cut_iL = -1.;
  double accuracy = -1.;
  double time = time_sentinel;
};

class Tuner {
public:
  Tuner(P3MParams &p, TuningInput const &in, TuningTimer const &timer)
      : m_p{p}, m_in{in}, m_timer{timer} {}

  void determine_mesh_limits() {
    // p3m.cpp:676-707
    auto const &L = m_in.box_l;
    double const volume = L[0] * L[1] * L[2];
    if (m_p.mesh[0] == -1 and m_p.mesh[1] == -1 and m_p.mesh[2] == -1) {
      double const normalized_box_dim = std::cbrt(volume);
      double const max_npart_per_dim = 512.;
      double const min_npart_per_dim =
          std::min(max_npart_per_dim, std::cbrt(static_cast<double>(m_in.sum_qpart)));
      m_density_min = min_npart_per_dim / normalized_box_dim;
      m_density_max = max_npart_per_dim / normalized_box_dim;
      // optional mesh range (densities of the smallest box dimension)
      double const lmin = std::min({L[0], L[1], L[2]});
      double const lmax = std::max({L[0], L[1], L[2]});
      m_density_min = std::max(m_density_min, (m_in.mesh_min - 0.5) / lmin);
      m_density_max = std::min(m_density_max, (m_in.mesh_max + 0.5) / lmax);
      m_tune_mesh = true;
    } else {
      double const mesh_density = static_cast<double>(m_p.mesh[0]) * (1. / L[0]);
      m_density_min = m_density_max = mesh_density;
      if (m_p.mesh[1] == -1 and m_p.mesh[2] == -1) {
        for (int i : {1, 2}) {
          m_p.mesh[i] = static_cast<int>(std::round(mesh_density * L[i]));
          m_p.mesh[i] += m_p.mesh[i] % 2;
        }
      }
    }
  }

  void determine_r_cut_limits() {
    // TuningAlgorithm.cpp:57-74 (one rank: local box = box)
    auto const &L = m_in.box_l;
    if (m_p.r_cut_iL == 0.) {
      double const min_box_l = std::min({L[0], L[1], L[2]});
      m_rmin = 0.;
      m_rmax = std::min(min_box_l, min_box_l / 2.) - m_in.skin;
      m_rmin *= 1. / L[0];
      m_rmax *= 1. / L[0];
    } else {
      m_rmin = m_rmax = m_p.r_cut_iL;
    }
  }

  void determine_cao_limits(int initial_cao) {
    // TuningAlgorithm.cpp:76-86
    if (m_p.cao == -1) {
      cao_min = std::max(1, m_in.cao_min);
      cao_max = std::min(7, m_in.cao_max);
      cao_best = std::clamp(initial_cao, cao_min, cao_max);
    } else {
      cao_min = cao_max = cao_best = m_p.cao;
    }
  }

  Accuracy accuracy(Vec3i const &mesh, int cao, double r_cut_iL) const {
    return calculate_accuracy(m_in.prefactor, mesh, cao, r_cut_iL, m_in.sum_qpart,
                              m_in.sum_q2, m_p.accuracy, m_in.box_l);
  }

  double get_mc_time(Vec3i const &mesh, int cao, double &r_cut_iL_out, double &alpha_L_out,
                     double &acc_out) {
    // TuningAlgorithm.cpp:116-200
    auto const &L = m_in.box_l;
    double r_lo = m_rmin, r_hi = m_rmax;
    double k_cut = std::numeric_limits<double>::max();
    for (int d = 0; d < 3; ++d)
      k_cut = std::min(k_cut, (static_cast<double>(cao) / 2.) * (L[d] / mesh[d]));
    double const min_box_l = std::min({L[0], L[1], L[2]});
    double const k_cut_max = min_box_l - m_in.skin;
    if (cao >= std::min({mesh[0], mesh[1], mesh[2]}) or k_cut >= k_cut_max) {
      log("cao too large for this mesh", mesh, cao, 0., 0., 0.);
      return -P3M_TUNE_CAO_TOO_LARGE;
    }
    auto acc = accuracy(mesh, cao, r_hi);
    alpha_L_out = acc.alpha_L;
    acc_out = acc.total;
    if (acc.total > m_p.accuracy) {
      log("accuracy not achieved", mesh, cao, r_hi, acc.alpha_L, acc.total);
      return -P3M_TUNE_ACCURACY_TOO_LARGE;
    }
    double r_cut_iL;
    for (;;) {
      r_cut_iL = 0.5 * (r_lo + r_hi);
      if (r_hi - r_lo < P3M_RCUT_PREC) break;
      auto const a = accuracy(mesh, cao, r_cut_iL);
      if (a.total > m_p.accuracy)
        r_lo = r_cut_iL;
      else
        r_hi = r_cut_iL;
    }
    r_cut_iL_out = r_cut_iL = r_hi;
    acc = accuracy(mesh, cao, r_cut_iL);
    alpha_L_out = acc.alpha_L;
    acc_out = acc.total;
    double const t = m_timer(mesh, cao, r_cut_iL, acc.alpha_L, m_in.timings);
    log("success", mesh, cao, r_cut_iL, acc.alpha_L, acc.total, t);
    ++m_n_trials;
    return t;
  }

  double get_m_time(Vec3i const &mesh, int &tuned_cao, double &tuned_r_cut_iL,
                    double &tuned_alpha_L, double &tuned_accuracy) {
    // TuningAlgorithm.cpp:218-326
    double best_time = -1., tmp_r = 0., tmp_a = 0., tmp_acc = 0.;
    int final_dir = 0;
    int cao = tuned_cao;
    double tmp_time;
    do {
      tmp_time = get_mc_time(mesh, cao, tmp_r, tmp_a, tmp_acc);
      if (tmp_time == -P3M_TUNE_CAO_TOO_LARGE) return tmp_time;
      if (tmp_time >= 0.) {
        best_time = tmp_time;
        tuned_r_cut_iL = tmp_r;
        tuned_alpha_L = tmp_a;
        tuned_accuracy = tmp_acc;
        tuned_cao = cao;
        break;
      }
      cao++;
      final_dir = 1;
    } while (cao <= cao_max);
    if (cao > cao_max) return -P3M_TUNE_CAO_TOO_LARGE;
    if (cao == cao_min)
      final_dir = 1;
    else if (cao == cao_max)
      final_dir = -1;
    if (final_dir == 0) {
      double dir_times[3];
      for (final_dir = -1; final_dir <= 1; final_dir += 2) {
        dir_times[final_dir + 1] = tmp_time =
            get_mc_time(mesh, cao + final_dir, tmp_r, tmp_a, tmp_acc);
        if (tmp_time < 0.) continue;
        if (tmp_time < best_time) {
          best_time = tmp_time;
          tuned_r_cut_iL = tmp_r;
          tuned_alpha_L = tmp_a;
          tuned_accuracy = tmp_acc;
          tuned_cao = cao + final_dir;
        }
      }
      if (dir_times[0] == best_time) {
        final_dir = -1;
      } else if (dir_times[2] == best_time) {
        final_dir = 1;
      } else {
        if ((dir_times[0] >= 0 && dir_times[0] < best_time + time_granularity) &&
            (dir_times[2] < 0 || dir_times[2] > dir_times[0]))
          final_dir = -1;
        else if ((dir_times[2] >= 0 && dir_times[2] < best_time + time_granularity) &&
                 (dir_times[0] < 0 || dir_times[0] > dir_times[2]))
          final_dir = 1;
        else
          return best_time;
      }
      cao += 2 * final_dir;
    } else {
      cao += final_dir;
    }
    for (; cao >= cao_min && cao <= cao_max; cao += final_dir) {
      tmp_time = get_mc_time(mesh, cao, tmp_r, tmp_a, tmp_acc);
      if (tmp_time < 0.) break;
      if (tmp_time < best_time) {
        best_time = tmp_time;
        tuned_r_cut_iL = tmp_r;
        tuned_alpha_L = tmp_a;
        tuned_accuracy = tmp_acc;
        tuned_cao = cao;
      } else if (tmp_time > best_time + time_granularity) {
        break;
      }
    }
    return best_time;
  }

  Trial get_time() {
    // p3m.cpp:709-750
    auto const &L = m_in.box_l;
    Trial tuned;
    double time_best = time_sentinel;
    double mesh_density = m_density_min;
    while (mesh_density <= m_density_max) {
      Trial trial;
      if (m_tune_mesh) {
        for (int i : {0, 1, 2}) {
          trial.mesh[i] = static_cast<int>(std::round(L[i] * mesh_density));
          trial.mesh[i] += trial.mesh[i] % 2;
        }
      } else {
        trial.mesh = m_p.mesh;
      }
      trial.cao = cao_best;
      double const trial_time =
          get_m_time(trial.mesh, trial.cao, trial.r_cut_iL, trial.alpha_L, trial.accuracy);
      if (trial_time >= 0.) {
        m_rmax = trial.r_cut_iL;
        if (trial_time < time_best) {
          m_n_trials = 0;
          tuned = trial;
          time_best = tuned.time = trial_time;
        } else if (trial_time > time_best + time_granularity or
                   m_n_trials > max_n_consecutive_trials) {
          break;
        }
      }
      mesh_density += 0.1;
    }
    return tuned;
  }

  void commit(Vec3i const &mesh, int cao, double r_cut_iL, double alpha_L) {
    // TuningAlgorithm.cpp:88-98
    auto const &L = m_in.box_l;
    m_p.r_cut = r_cut_iL * L[0];
    m_p.r_cut_iL = r_cut_iL;
    m_p.cao = cao;
    m_p.alpha_L = alpha_L;
    m_p.alpha = alpha_L * (1. / L[0]);
    m_p.mesh = mesh;
  }

  void run() {
    m_p.tuning = true;
    Trial const t = get_time();
    m_p.tuning = false;
    if (t.time == time_sentinel)
      throw std::runtime_error(m_in.name + ": failed to reach requested accuracy");
    m_p.accuracy = t.accuracy;
    commit(t.mesh, t.cao, t.r_cut_iL, t.alpha_L);
    if (m_in.verbose)
      std::printf("%s tuned: mesh (%d %d %d) cao %d r_cut_iL %.6e alpha_L %.6e accuracy %.3e "
                  "time %.3f ms\n",
                  m_in.name.c_str(), t.mesh[0], t.mesh[1], t.mesh[2], t.cao, t.r_cut_iL,
                  t.alpha_L, t.accuracy, t.time);
  }

  int cao_best = -1;

private:
  void log(char const *what, Vec3i const &mesh, int cao, double r, double a, double acc,
           double t = -1.) const {
    if (m_in.verbose)
      std::printf("%-28s mesh %4d cao %d r_cut_iL %.5e alpha_L %.5e err %.3e t %.3f\n", what,
                  mesh[0], cao, r, a, acc, t);
  }

  P3MParams &m_p;
  TuningInput const &m_in;
  TuningTimer const &m_timer;
  double m_density_min = -1., m_density_max = -1.;
  bool m_tune_mesh = false;
  double m_rmin = -1., m_rmax = -1.;
  int cao_min = -1, cao_max = -1;
  std::size_t m_n_trials = 0;
};
} // namespace

double fft_stage_cost(int n) {
  // relative cost per point of the device FFT along a line of length n: one
  // unit per radix-2..8 stage (constant-twiddle register DFTs); a larger
  // prime radix R as the first stage computes symmetric output pairs with
  // (R - 1) / 2 table cosines, about R / 8 units, and one output at a time
  // with R - 1 table-twiddle products elsewhere, about (R - 1) / 2.5 units
  FftPlan1D const p = make_fft_plan(n);
  double c = 0.;
  for (int s = 0; s < p.nstages; ++s) {
    int const r = p.radix[s];
    c += r <= 8 ? 1. : s == 0 ? r / 8. : (r - 1) / 2.5;
  }
  return c;
}

double model_step_ms(Vec3i const &mesh, int cao, double r_cut, int n_q, Vec3d const &box_l,
                     double pair_ns, bool single_precision) {
  // device k-space cost from the measured stage rates of the solver, per
  // particle stencil point (spread + gather: 1.5 ps FP64 mesh, 1.8 ps FP32
  // mesh, whose systems are small and sparse) and per mesh point (FFTs and
  // convolution: 74 ps FP64, 40 ps FP32) for three-stage lines (n = 128),
  // half of it scaling with the FFT stage cost of the mesh, plus a fixed
  // launch cost
  double const R = static_cast<double>(mesh[0]) * mesh[1] * mesh[2];
  double const stencil = static_cast<double>(n_q) * cao * cao * cao;
  double const stages =
      (fft_stage_cost(mesh[0]) + fft_stage_cost(mesh[1]) + fft_stage_cost(mesh[2])) / 3.;
  double const stencil_ps = single_precision ? 1.8 : 1.5;
  double const mesh_ps = (single_precision ? 40. : 74.) * (0.5 + 0.5 * stages / 3.);
  return 1e-9 * (stencil_ps * stencil + mesh_ps * R) + 0.01 +
         model_real_space_ms(n_q, r_cut, box_l, pair_ns);
}

double model_real_space_ms(int n_q, double r_cut, Vec3d const &box_l, double pair_ns) {
  // pairs within r_cut for a uniform system; only used to rank candidate
  // parameter sets
  double const volume = box_l[0] * box_l[1] * box_l[2];
  double const density = n_q / volume;
  double const pairs =
      0.5 * n_q * density * (4. / 3.) * std::numbers::pi * r_cut * r_cut * r_cut;
  return pairs * pair_ns * 1e-6;
}

void tune_p3m(P3MParams &params, TuningInput const &in, TuningTimer const &timer) {
  // CoulombP3MImpl::tune, p3m.cpp:757-791
  if (params.alpha_L == 0. and params.alpha != 0.) params.alpha_L = params.alpha * in.box_l[0];
  if (params.r_cut_iL == 0. and params.r_cut != 0.)
    params.r_cut_iL = params.r_cut * (1. / in.box_l[0]);
  if (in.sum_qpart == 0)
    throw std::runtime_error("CoulombP3M: no charged particles in the system");
  Tuner t(params, in, timer);
  try {
    t.determine_mesh_limits();
    t.determine_r_cut_limits();
    t.determine_cao_limits(7);
    t.run();
  } catch (...) {
    params.tuning = false;
    throw;
  }
}

} // namespace espresso_amd
