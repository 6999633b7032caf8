// ELC (electrostatic layer correction) far-field terms on the device: the
// correction that turns a 3D-periodic P3M solve of a slab (charges in
// 0 <= z <= h = L_z - gap) into a 2D-periodic one,
// ElectrostaticLayerCorrection::add_force / calc_energy
// (src/core/electrostatics/elc.cpp:830-940) without dielectric contrasts:
// the Yeh-Berkowitz dipole term and P3M background removal (elc.cpp:201-342)
// plus the sums over the reciprocal vectors (p/L_x, q/L_y) within far_cut
// (setup_PoQ / setup_PQ and their force and energy kernels, :500-828).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace espresso_amd {

struct ElcParams {
  double box[3] = {0., 0., 0.};
  double prefactor = 0.;
  double gap_size = 0.;
  double far_cut = 0.; // > 0 (tuned or given)
  bool neutralize = true;
  // dielectric contrasts (elc_data, elc.hpp:62-118): image charges of the
  // layers z < space_layer (delta_bot q at -z) and z > h - space_layer
  // (delta_top q at 2h - z); neutralize is off with contrasts
  double delta_top = 0.;
  double delta_bot = 0.;
  bool const_pot = false;
  double pot_diff = 0.;
  double space_layer = 0.;
  // P3M pair terms for the image self interaction (pair_force, p3m.hpp:146-180)
  double alpha = 0.;
  double r_cut = 0.;
  bool dielectric() const { return delta_top != 0. || delta_bot != 0.; }
};

/** One reciprocal vector of the far-field sums: k = (kx, ky), omega = |k|,
 *  pref_di = c prefactor pi / (L_x L_y) with c = 4 for the p- and q-modes
 *  and 8 for the pq-modes, and the sums' scale pref = -pref_di /
 *  expm1(omega L_z). */
struct ElcMode {
  double kx, ky, omega, pref, pref_di;
  double inv_omega;
  int p, q; // kx = 2 pi p / L_x, ky = 2 pi q / L_y
};

/** Reciprocal vectors in the reference's loop order (elc.cpp:842-883). */
std::vector<ElcMode> elc_modes(ElcParams const &p);

/** ElectrostaticLayerCorrection::tune_far_cut (elc.cpp:942-973); with
 *  dielectric contrasts the image height is h + space_layer.  Throws
 *  std::runtime_error "ELC tuning failed: maxPWerror too small". */
double elc_tune_far_cut(double const box[3], double gap_size, double max_pw_error,
                        bool dielectric = false, double space_layer = 0.);

/** recalc_space_layer (elc.cpp:1032-1054): gap / 3, capped by gap - r_cut
 *  and h / 2; std::runtime_error when no room is left. */
double elc_space_layer(double gap_size, double box_h, double r_cut);

class ElcSolver {
public:
  explicit ElcSolver(int device);
  ~ElcSolver();
  ElcSolver(ElcSolver const &) = delete;
  ElcSolver &operator=(ElcSolver const &) = delete;

  /** Adds the ELC forces to force (may be null) for the n particles of the
   *  device arrays pos [n][3] and q [n] on stream st.  With energy != null
   *  the energy is reduced.  With energy or sync the call synchronises the
   *  stream and raises std::runtime_error when a charged particle lies
   *  outside 0 <= z <= h after folding (the gap check of elc.cpp:184-195)
   *  or a position is not finite. */
  void compute(hipStream_t st, ElcParams const &p, int64_t n, double const *pos,
               double const *q, double *force, double *energy, bool sync);
  int num_modes() const;

  /** Dielectric contrasts: the real particles (folded) followed by the
   *  bottom and the top image of every particle, [3n] arrays on the device
   *  (charge_assign with ChargeProtocol BOTH, elc.cpp:1110-1141).  q_both
   *  holds the real charges and the image charges (0 outside the layers),
   *  q_img the image charges only.  Pointers stay valid until the next
   *  call. */
  void images(hipStream_t st, ElcParams const &p, int64_t n, double const *pos, double const *q,
              double **pos3, double **q_both, double **q_img);
  /** Scratch force array [3n][3] for a solve on the image arrays, cleared. */
  double *image_forces(hipStream_t st, int64_t n);
  /** force[i] += f3[i] for the first n rows. */
  void add_real_rows(hipStream_t st, int64_t n, double const *f3, double *force);
  /** dielectric_layers_self_forces / _self_energy (elc.hpp:140-171): the
   *  pair term of each particle in a layer with its own image.  With
   *  energy != null the call synchronises. */
  void layer_self(hipStream_t st, ElcParams const &p, int64_t n, double const *pos,
                  double const *q, double *force, double *energy);

private:
  struct Impl;
  Impl *m;
};

} // namespace espresso_amd
