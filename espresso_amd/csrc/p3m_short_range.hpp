// Real-space (short-range) part of P3M Coulomb on the device: every pair of
// charged particles closer than r_cut (minimum image) contributes the pair
// force and energy of CoulombP3M::pair_force / pair_energy (p3m.hpp:146-180,
// A&S erfc of AS_erfc_part.hpp:29-40), as the reference's short-range loop
// applies them over its cell system (forces.cpp:178-208).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace espresso_amd {

struct ShortRangeParams {
  double box[3] = {0., 0., 0.};
  double prefactor = 0.;
  double alpha = 0.;
  double r_cut = 0.;
  int subdivision = 0; // cells per r_cut: 0 = automatic (2 when the box allows), 1 or 2
  // ELC dielectric contrasts: each pair within r_cut also gets the terms of
  // add_pair_force_corrections / pair_energy_correction (elc.hpp:250-296),
  // particle i against the image of itself in the layer it sits in
  bool elc_images = false;
  double elc_h = 0.;  // L_z - gap
  double elc_sl = 0.; // space layer
  double elc_dt = 0., elc_db = 0.;
};

struct ShortRangeResult {
  double energy = 0.;    // sum over pairs of the pair energy
  double virial[9] = {}; // sum over pairs of F_ij (x) d_ij, d_ij = r_i - r_j (row-major)
};

/** Cell-list pair solver.  Charged particles are counting-sorted into cells
 *  of edge >= r_cut / s, s = 2 (or 1) cells per r_cut, at most 64 per
 *  dimension; one workgroup per cell
 *  takes its particles against the (2 s + 1)^3 neighbour cells, staged through LDS
 *  with the periodic shift applied, and accumulates each particle's force in
 *  registers (FP64).  With fewer than 3 cells along a dimension it falls
 *  back to all pairs with the minimum-image displacement. */
class ShortRange {
public:
  explicit ShortRange(int device);
  ~ShortRange();
  ShortRange(ShortRange const &) = delete;
  ShortRange &operator=(ShortRange const &) = delete;

  /** Adds the pair forces to force (device arrays of n particles, stream
   *  st).  With out != nullptr the energy and virial are reduced and the
   *  call synchronises the stream. */
  void compute(hipStream_t st, ShortRangeParams const &p, int64_t n, double const *pos,
               double const *q, double *force, ShortRangeResult *out);

  /** Cells per dimension of the last call (1 1 1: all pairs). */
  void cells(int out[3]) const;

  /** Dipolar real space (DipolarP3M::pair_force / pair_energy,
   *  dp3m.hpp:160-235) over the same cell lists, for particles with a
   *  non-zero dipole: adds forces and torques ([n][3] device arrays, either
   *  may be null) and, with energy != null, returns the pair energy (the
   *  call then synchronises).  exact_erfc selects libm erfc for B_r (the
   *  reference's choice at accuracy <= 5e-6).  Returns false, doing
   *  nothing, when the box is narrower than three cells of edge r_cut. */
  bool compute_dipolar(hipStream_t st, ShortRangeParams const &p, bool exact_erfc, int64_t n,
                       double const *pos, double const *dip, double *force, double *torque,
                       double *energy);

private:
  struct Impl;
  Impl *m;
};

} // namespace espresso_amd
