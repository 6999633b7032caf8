// Single-system P3M k-space solver for MI355X (gfx950): a batch of one.
//
// A solve is one stream-ordered sequence (see p3m_batch.hpp):
//   bin -> scan -> scatter      sort charged particles into 8^3 mesh bricks by
//                               the bit-exact index of their first stencil point
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
// (p3m.cpp:393-449); the host context around it owns parameters and tuning.
#pragma once

#include "p3m_batch.hpp"

#include <memory>
#include <vector>

namespace espresso_amd {

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
  bool configured() const { return static_cast<bool>(m_batch); }

  /** One k-space solve on device pointers (pos [n][3], q [n], force [n][3]
   *  accumulated in place, may be null without P3M_FORCES).  Without sync
   *  the work is only enqueued and `res` is not filled. */
  void solve(int64_t n, double const *d_pos, double const *d_q, double const *d_upos,
             double *d_force, unsigned flags, SolveResult *res, bool sync = true);

  /** solve() on device pointers without any allocation: every working
   *  buffer comes from the library's fixed scratch array
   *  (BatchSolver::solve_fixed); synchronous. */
  void solve_fixed(int64_t n, double const *d_pos, double const *d_q, double const *d_upos,
                   double *d_force, unsigned flags, SolveResult *res);

  /** Energy and virial of the last solve written to device memory in
   *  stream order (BatchSolver::device_results). */
  void device_results(double *d_energy, double *d_virial);

  /** Charge mesh after spreading ([Mx][My][Mz], device-to-host copy). */
  void debug_charge_mesh(int64_t n, double const *d_pos, double const *d_q,
                         std::vector<double> &out);
  /** CA-local nearest mesh point of each particle's first stencil point
   *  ([n][3]; -1 for uncharged particles). */
  void debug_indices(int64_t n, double const *d_pos, double const *d_q,
                     std::vector<int32_t> &out);

  /** Per-stage HIP events of asynchronous solves (see BatchSolver). */
  void profile_begin(int max_solves);
  int profile_end(double *stage_ms);

  hipStream_t stream() const;
  int device() const { return m_device; }
  /** Stage times of the last synchronous solve in ms: total, bin, spread,
   *  fft_fwd, convolve, fft_bwd, gather. */
  float last_ms[8] = {};
  void set_launch_shape(LaunchShape const &s);
  /** Deterministic solves (BatchSolver::set_deterministic; the default of
   *  FP64 solvers); an explicit choice is kept across later configure()
   *  calls. */
  void set_deterministic(bool on);
  LaunchShape launch_shape() const {
    return m_batch ? m_batch->launch_shape() : (m_shape_set ? m_shape : LaunchShape{});
  }
  /** Index rule of the solves (BatchSolver::set_convention); kept across
   *  later configure() calls. */
  void set_convention(int conv);
  int convention() const { return m_conv; }
  /** Pick the fastest launch shape for force solves of these device arrays
   *  (BatchSolver::autotune); d_force is scratch.  The shape is kept across
   *  later configure() calls. */
  AutotuneResult autotune(int64_t n, double const *d_pos, double const *d_q, double *d_force,
                          int reps);
  LocalMesh const &local_mesh() const;

private:
  int m_device;
  hipStream_t m_idle_stream = nullptr; // stream handed out before configure
  std::unique_ptr<BatchSolver> m_batch;
  LaunchShape m_shape;
  bool m_shape_set = false;
  int m_conv = 0;
  int m_det = -1; // -1: the batch solver's default
};

} // namespace espresso_amd
