// Host-side P3M parameter handling for the MI355X solver.
//
// Restates the parameter semantics of ESPResSo's P3MParameters
// (src/core/p3m/common.hpp:59-172), the single-rank charge-assignment mesh
// geometry P3MLocalMesh (common.hpp:175-214, common.cpp:34-106), the
// closed-form error estimates (src/core/electrostatics/p3m.cpp:145-243) and
// the sanity checks (p3m.cpp:793-853).  Exceptions carry the reference's
// messages; the C ABI layer maps them to status codes.
#pragma once

#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace espresso_amd {

/** Value of epsilon that selects metallic boundary conditions
 *  (common.hpp:46). */
inline constexpr double P3M_EPSILON_METALLIC = 0.0;
/** Default mesh offset (config.hpp:45-47). */
inline constexpr double P3M_MESHOFF = 0.5;
/** Round-off guard of the mesh geometry (config.hpp:64-67). */
inline constexpr double ROUND_ERROR_PREC = 1.0e-14;

using Vec3d = std::array<double, 3>;
using Vec3i = std::array<int, 3>;

/** P3M parameters and derived mesh constants (common.hpp:59-172). */
struct P3MParams {
  bool tuning = false;
  double alpha_L = 0.;
  double r_cut_iL = 0.;
  Vec3i mesh{};
  Vec3d mesh_off{};
  int cao = 0;
  double accuracy = 0.;
  double epsilon = 0.;
  Vec3d cao_cut{};
  Vec3d a{};
  Vec3d ai{};
  double alpha = 0.;
  double r_cut = 0.;
  int cao3 = -1;

  P3MParams() = default;
  /** Validating constructor; throws std::domain_error with the reference's
   *  messages (common.hpp:94-150). */
  P3MParams(bool tuning, double epsilon, double r_cut, Vec3i const &mesh,
            Vec3d const &mesh_off, int cao, double alpha, double accuracy);

  /** a, ai and cao_cut from mesh and box (common.hpp:158-163). */
  void recalc_a_ai_cao_cut(Vec3d const &box_l);
};

/** Charge-assignment (CA) mesh of one rank owning the whole box:
 *  inner mesh plus margins for the stencil and the Verlet skin
 *  (common.cpp:34-106).  Indices of the first stencil point computed from
 *  ld_pos / ai are bit-identical to the reference's. */
struct LocalMesh {
  Vec3i dim{};
  int size = 0;
  Vec3i ld_ind{};
  Vec3d ld_pos{};
  Vec3i inner{};
  Vec3i in_ld{};
  Vec3i in_ur{};
  std::array<int, 6> margin{};
  int q_2_off = 0;
  int q_21_off = 0;

  void calc_local_ca_mesh(P3MParams const &params, Vec3d const &box_l, double skin,
                          double space_layer = 0.);
  void recalc_ld_pos(P3MParams const &params);
};

/** Shift tables 0..M/2, -M/2..-1 (common.hpp:248-264). */
std::vector<int> calc_meshift(int n, bool zero_out_midpoint);

/** sin(pi x)/(pi x) with the Taylor branch of math.hpp:53-81. */
double sinc(double x);
/** Closed-form aliasing sum of math.hpp:92-126. */
double analytic_cotangent_sum(int cao, int n, double mesh_i);

/** Kolafa-Perram real-space error (p3m.cpp:145-160). */
double real_space_error(double pref, double r_cut_iL, int n_c_part, double sum_q2,
                        double alpha_L, Vec3d const &box_l);
/** Hockney-Eastwood k-space error (p3m.cpp:187-243), term by term. */
double k_space_error(double pref, Vec3i const &mesh, int cao, int n_c_part, double sum_q2,
                     double alpha_L, Vec3d const &box_l);
/** The same sum grouped by |n|^2 with per-(mesh, cao) weights cached across
 *  calls: O(|n|^2_max) per call instead of O(M^3).  Agrees with
 *  k_space_error to rounding, which reaches 1e-6 relative where the terms
 *  cancel (small alpha_L, high cao); used by the tuner. */
double k_space_error_fast(double pref, Vec3i const &mesh, int cao, int n_c_part,
                          double sum_q2, double alpha_L, Vec3d const &box_l);

/** Result of CoulombTuningAlgorithm::calculate_accuracy (p3m.cpp:636-674),
 *  with the k-space part from k_space_error_fast. */
struct Accuracy {
  double total, rs, ks, alpha_L;
};
Accuracy calculate_accuracy(double pref, Vec3i const &mesh, int cao, double r_cut_iL,
                            int n_c_part, double sum_q2, double accuracy, Vec3d const &box_l);

/** Real-space pair force/energy with the Abramowitz-Stegun erfc part
 *  (p3m.hpp:146-180, AS_erfc_part.hpp:29-40). */
Vec3d pair_force(double prefactor, double alpha, double r_cut, double q1q2, Vec3d const &d,
                 double dist);
double pair_energy(double prefactor, double alpha, double r_cut, double q1q2, double dist);

/** k-space cutoff checks of CoulombP3M::sanity_checks_boxl (p3m.cpp:793-818),
 *  one rank: the local box is the whole box. */
void sanity_checks_boxl(P3MParams const &params, Vec3d const &box_l);

/** Relative charge excess of coulomb.cpp:265-284 (Kahan sum over charges
 *  divided by the smallest nonzero |q|). */
double charge_excess_ratio(std::vector<double> const &charges);
/** Throws the reference's neutrality error (coulomb.cpp:286-325). */
void check_charge_neutrality(std::vector<double> const &charges, double relative_tolerance);

} // namespace espresso_amd
