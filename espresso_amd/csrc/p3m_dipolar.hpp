// Dipolar P3M (magnetostatics) on the device: the k-space part of
// DipolarP3MImpl<double, CPU> (src/core/magnetostatics/dp3m.cpp:246-517)
// for a cubic box and mesh, and the real-space dipolar pair terms of
// DipolarP3M::pair_force / pair_energy (dp3m.hpp:160-235).
//
// One solve: the three dipole components are spread with the charge
// assignment weights of the Coulomb solver (same bit-exact CA-mesh index),
// transformed by complex 3D FFTs (LDS line transforms along each axis),
// combined in k-space with the dipolar influence functions (energy and
// torques: G_energy, forces: G_force), and transformed back: three field
// meshes for the torques and the six symmetric second-derivative meshes for
// the forces (the reference transforms nine; the force matrix is
// symmetric).  Influence functions, the self-energy correction and the
// tuning run on the host in FP64.
#pragma once

#include "p3m_host.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

namespace espresso_amd {

struct DipolarParams {
  int mesh = 0; // cubic mesh
  int cao = 0;
  double alpha = 0.;
  double r_cut = 0.;
  double accuracy = 0.;
  double prefactor = 0.;
  double epsilon = 0.; // 0 = metallic
  double mesh_off[3] = {0.5, 0.5, 0.5};
  double box = 0.; // cubic box edge
  double skin = 0.;
};

// Error estimates and tuning of dipolar P3M (dp3m.cpp:533-820).
double dp3m_real_space_error(double box, double r_cut_iL, int n_c_part, double sum_mu2,
                             double alpha_L);
double dp3m_k_space_error(double box, int mesh, int cao, int n_c_part, double sum_mu2,
                          double alpha_L);
/** DipolarTuningAlgorithm::calculate_accuracy: {total, real, k-space, alpha_L}. */
void dp3m_calculate_accuracy(double box, int mesh, int cao, double r_cut_iL, int n_c_part,
                             double sum_mu2, double accuracy, double out[4]);

class DipolarSolver {
public:
  explicit DipolarSolver(int device);
  ~DipolarSolver();
  DipolarSolver(DipolarSolver const &) = delete;
  DipolarSolver &operator=(DipolarSolver const &) = delete;

  /** Geometry, FFT plan and influence functions for tuned parameters. */
  void configure(DipolarParams const &p);
  bool configured() const;
  DipolarParams const &params() const;

  /** k-space forces and torques (added to the device arrays force and
   *  torque, [n][3], either may be null) and, with energy != null, the
   *  k-space energy including self, correction and surface terms.  pos and
   *  dip are device arrays [n][3].  Synchronises when energy != null or
   *  sync. */
  void solve(int64_t n, double const *pos, double const *dip, double *force, double *torque,
             double *energy, bool sync);
  /** Real-space pair forces, torques and (energy != null) energy of all
   *  pairs within r_cut (minimum image): cell lists of edge r_cut/2 or
   *  r_cut, all pairs in boxes narrower than three cells. */
  void short_range(int64_t n, double const *pos, double const *dip, double *force,
                   double *torque, double *energy, bool sync);

  hipStream_t stream() const;
  /** Cells per r_cut of the real space: 0 = by occupancy, 1 or 2. */
  void set_short_range_subdivision(int s);

private:
  struct Impl;
  std::unique_ptr<Impl> m;
};

} // namespace espresso_amd
