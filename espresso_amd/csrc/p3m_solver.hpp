// Device-side P3M k-space solver for MI355X (gfx950).
//
// One Solver owns the device buffers of one system: particle bins, the real
// charge mesh, the half-spectrum work meshes, the influence function and the
// three real field meshes.  A solve is one stream-ordered sequence:
//
//   bin -> scan -> scatter      sort charged particles into mesh bricks by the
//                               bit-exact index of their first stencil point
//   spread                      B-spline assignment into LDS brick tiles,
//                               flushed with periodic wrap into the mesh
//   fft_z (r2c) -> fft_y        forward transform, two real lines per complex
//   fft_x + convolution         forward x, influence function, energy/virial
//                               reduction, ik-differentiation, inverse x
//   fft_y inverse -> fft_z c2r  three real field meshes
//   gather                      LDS brick tiles of the fields, FP64 force
//                               accumulation per particle
//
// It is the k-space part of CoulombP3MImpl::long_range_kernel
// (src/core/electrostatics/p3m.cpp:452-592) and long_range_pressure
// (p3m.cpp:393-449); the host class around it owns parameters and tuning.
#pragma once

#include "p3m_host.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace espresso_amd {

enum class MeshPrecision : int { FP64 = 0, FP32 = 1 };

/** Radix plan of one 1D transform length. */
struct FftPlan1D {
  static constexpr int max_stages = 16;
  int n = 0;
  int nstages = 0;
  int radix[max_stages] = {};
};
FftPlan1D make_fft_plan(int n);

struct SolveResult {
  double energy = 0.;     // k-space energy incl. self/net/dipole terms * prefactor
  double virial[9] = {};  // k-space virial tensor (p3m.cpp:393-449 return value)
  double n_q = 0.;        // number of charged particles
  double sum_q = 0.;
  double sum_q2 = 0.;
  double dipole[3] = {};
};

class Solver {
public:
  explicit Solver(int device);
  ~Solver();
  Solver(Solver const &) = delete;
  Solver &operator=(Solver const &) = delete;

  /** Build geometry, FFT plans, buffers and the influence function.
   *  `params` must be tuned/valid (mesh, cao, alpha set). */
  void configure(P3MParams const &params, Vec3d const &box_l, double skin, double prefactor,
                 MeshPrecision precision);
  bool configured() const { return m_configured; }

  /** One k-space solve on device pointers (pos [n][3], q [n], force [n][3]
   *  accumulated in place, may be null without P3M_FORCES). */
  void solve(int64_t n, double const *d_pos, double const *d_q, double const *d_upos,
             double *d_force, unsigned flags, SolveResult *res, bool sync = true);

  /** Record per-stage HIP events of the next `max_solves` solves into a ring
   *  (no host synchronisation between solves); profile_end() synchronises
   *  and returns the number of profiled solves and the summed stage times
   *  (total, bin, spread, fft_fwd, convolve, fft_bwd, gather) in ms. */
  void profile_begin(int max_solves);
  int profile_end(double *stage_ms);

  /** Charge mesh after spreading ([Mx][My][Mz], device-to-host copy). */
  void debug_charge_mesh(int64_t n, double const *d_pos, double const *d_q,
                         std::vector<double> &out);
  /** CA-local nearest mesh point of each particle's first stencil point
   *  ([n][3]; -1 for uncharged particles). */
  void debug_indices(int64_t n, double const *d_pos, double const *d_q,
                     std::vector<int32_t> &out);

  hipStream_t stream() const { return m_stream; }
  int device() const { return m_device; }
  /** Stage times of the last solve in ms: total, bin, spread, fft_fwd,
   *  convolve, fft_bwd, gather. */
  float last_ms[8] = {};
  /** Brick shape used by spreading and gathering (autotuner knob). */
  void set_brick(int bx, int by, int bz);
  Vec3i brick() const { return m_brick; }

  LocalMesh const &local_mesh() const { return m_lm; }

  struct Impl;

private:
  Impl *m;
  int m_device;
  hipStream_t m_stream = nullptr;
  bool m_configured = false;
  Vec3i m_brick{8, 8, 8};
  LocalMesh m_lm;
};

/** Throws std::runtime_error for HIP failures. */
void hip_check(hipError_t err, char const *what);

} // namespace espresso_amd
