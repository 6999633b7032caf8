// P3M parameter tuning (restatement of ESPResSo's CoulombTuningAlgorithm,
// src/core/electrostatics/p3m.cpp:594-791, and TuningAlgorithm,
// src/core/p3m/TuningAlgorithm.cpp:57-328).
//
// The reference times full integration steps (tuning.cpp:73-95); here the
// timer is a callback, by default the measured device k-space solve plus a
// modelled real-space pair cost, because the real-space loop is not part of
// this solver.
#pragma once

#include "p3m_host.hpp"

#include <functional>
#include <string>

namespace espresso_amd {

struct TuningInput {
  Vec3d box_l{};
  double skin = 0.4;
  double prefactor = 1.;
  int sum_qpart = 0;
  double sum_q2 = 0.;
  int timings = 10;
  bool verbose = false;
  std::string name = "CoulombP3M";
  // optional restriction of the explored meshes and orders
  int mesh_min = 1, mesh_max = 1 << 20;
  int cao_min = 1, cao_max = 7;
};

/** Relative cost per point of the device FFT along a line of length n. */
double fft_stage_cost(int n);

/** Modelled cost (ms) of one k-space solve on the device (FP64 or FP32 mesh)
 *  plus the real-space pair loop at `pair_ns` nanoseconds per pair, for
 *  tuning without timing runs. */
double model_step_ms(Vec3i const &mesh, int cao, double r_cut, int n_q, Vec3d const &box_l,
                     double pair_ns = 1., bool single_precision = false);

/** Returns the cost of one step in ms for a candidate (mesh, cao, r_cut_iL,
 *  alpha_L), or a negative value when the candidate cannot be timed. */
using TuningTimer = std::function<double(Vec3i const &mesh, int cao, double r_cut_iL,
                                         double alpha_L, int timings)>;

/** Tune the free parameters of `params` (mesh, cao, r_cut, alpha) to reach
 *  params.accuracy; throws std::runtime_error("<name>: failed to reach
 *  requested accuracy") like TuningAlgorithm::tune (TuningAlgorithm.hpp:154-175). */
void tune_p3m(P3MParams &params, TuningInput const &in, TuningTimer const &timer);

/** Modelled real-space cost (ms) of the short-range loop for N_q charges at
 *  `pair_ns` nanoseconds per pair within r_cut. */
double model_real_space_ms(int n_q, double r_cut, Vec3d const &box_l, double pair_ns = 1.);

} // namespace espresso_amd
