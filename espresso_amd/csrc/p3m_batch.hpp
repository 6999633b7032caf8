// Descriptor-driven P3M k-space solver for many systems per launch (MI355X).
//
// Every kernel of the pipeline (bin, scan, scatter, spread, r2c/c2c/c2r LDS
// FFT passes with the fused influence/ik convolution, gather, per-system
// reductions) runs over a selection of systems in one launch: a work table
// (prefix sums of blocks per selected system) maps each workgroup to its
// system, whose mesh, stencil order, FFT plan and buffers come from a
// per-system descriptor in device memory.  Spread and gather are compiled
// per charge-assignment order and launched once per order present.
//
// A single system is a batch of one: the plain FP64 solver and the batched
// FP32-mesh solver use the same kernels.
#pragma once

#include "p3m_host.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
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
  double energy = 0.;    // k-space energy incl. self/net/dipole terms * prefactor
  double kspace = 0.;    // the mesh part sum G |rho_hat|^2 / (2 V), before the corrections
  double virial[9] = {}; // k-space virial tensor (p3m.cpp:393-449 return value)
  double n_q = 0.;       // number of charged particles
  double sum_q = 0.;
  double sum_q2 = 0.;
  double dipole[3] = {};
};

/** Solve flags beyond the C ABI's P3M_* bits, for the charge-assignment
 *  protocol of CoulombP3M (prepare_fft_mesh / assign_charge, p3m.hpp:91-143):
 *  kSolveAccumulate keeps the charge mesh (no zeroing) and adds the particles
 *  to it; kSolveMeshOnly stops after the spread; kSolveNoSpread bins the
 *  particles for the gather only and transforms the mesh as assigned. */
constexpr unsigned kSolveAccumulate = 1u << 16;
constexpr unsigned kSolveMeshOnly = 1u << 17;
constexpr unsigned kSolveNoSpread = 1u << 18;
/** Set by the solver itself for deterministic solves (set_deterministic). */
constexpr unsigned kSolveDeterministic = 1u << 19;

/** Device particle arrays of all systems, concatenated in system order
 *  (positions AoS [n][3], charges [n], forces [n][3]; upos may be null). */
struct ParticleArrays {
  double const *pos = nullptr;
  double const *q = nullptr;
  double const *upos = nullptr;
  double *force = nullptr;
};

/** Launch shapes picked by the autotuner for a class of systems. */
/** Gather mappings: lanes over the stencil layer i (several particles per
 *  wave), or over the (j, k) plane in segments of up to 64, 16 or 8 lanes per
 *  particle. */
constexpr int kGatherLayer = 0, kGatherPlane = 1, kGatherPlane16 = 2, kGatherPlane8 = 3;
/** Spread mappings: lanes over the (j, k) plane in segments of up to 64, 16
 *  or 8 lanes per particle. */
constexpr int kSpreadSeg64 = 0, kSpreadSeg16 = 1, kSpreadSeg8 = 2;

struct LaunchShape {
  int gather = kGatherPlane16; // gather mapping
  int ncol = 8;             // kz columns per block in the y/x FFT passes
  int zlines = 16;          // pairs of real z lines per block in the z passes
  int fused_lds = 96 * 1024; // fused y/x pass for meshes whose two x-y
                             // planes fit in this many LDS bytes (0: never)
  int spread = kSpreadSeg64; // spread mapping
  int threads = 512;         // block size of the spread and gather kernels (256, 512)
  int chunk = 64;            // particles staged per brick pass (1..64, <= threads / 3)
  int column = 0;            // 0: 8^3 bricks; 8 or 16: column march over W x W columns
  int seg = 32;              // column march: x layers per block (1..1024)
  int col_lanes = 8;         // column march: lanes per particle (8 or 16)
  int group = 1;             // spread: bricks per block, 1 or 4 (2x2x1, sharing one tile;
                             // 8^3 bricks and the plane spread mapping)
  int gather_threads = 0;    // block size of the brick gather (0: threads; 256, 512)
  int gather_block() const { return gather_threads ? gather_threads : threads; }
};

struct AutotuneResult {
  LaunchShape best;
  double best_ms = 0.;     // mean force-solve time of the best shape
  double default_ms = 0.;  // same for the shape in use before tuning
  int candidates = 0;
};

class BatchSolver {
public:
  BatchSolver(int device, MeshPrecision precision);
  ~BatchSolver();
  BatchSolver(BatchSolver const &) = delete;
  BatchSolver &operator=(BatchSolver const &) = delete;

  /** Add a system with tuned parameters; allocates its resident meshes and
   *  computes its influence function.  Returns the system id. */
  int add_system(P3MParams const &params, Vec3d const &box_l, double skin, double prefactor,
                 int64_t n_particles);
  int num_systems() const;
  /** Change the particle count of system s (offsets of later systems move). */
  void set_particle_count(int s, int64_t n);
  /** Index rule and ik of system s: 0 = the CPU kernels' (default), 1 = the
   *  reference's GPU kernels' (FP32 index nmp = floor(x M / L - (cao - 1) / 2
   *  + 1/2), Nyquist wave number kept at +M/2).  Drains the lanes when it
   *  changes. */
  void set_convention(int s, int conv);
  int convention(int s) const;
  int64_t particle_offset(int s) const;
  int64_t particle_count(int s) const;
  int64_t total_particles() const;
  /** Bytes of resident device memory held for system s (meshes, G). */
  size_t system_bytes(int s) const;
  LocalMesh const &local_mesh(int s) const;

  /** Solve the selected systems.  Forces are accumulated in place into
   *  pa.force (F -= ...) with P3M_FORCES.  With sync, per-system results
   *  are returned in `out` (indexed like `sel`). */
  void solve(std::vector<int> const &sel, ParticleArrays const &pa, unsigned flags, bool sync,
             std::vector<SolveResult> *out);

  /** Lanes: independent streams with their own plans and particle scratch
   *  over the shared resident systems, so that solves of disjoint selections
   *  overlap on the device.  Lane 0 exists from the start and serves solve(),
   *  graph_solve(), profiling and the debug calls. */
  int add_lane();
  int num_lanes() const;
  hipStream_t lane_stream(int lane) const;
  /** Enqueue a solve of `sel` on `lane` (no system of `sel` may be in another
   *  lane's unfinished solve).  With `scratch`, the particle scratch of the
   *  solve is taken from that device block of scratch_size >=
   *  working_bytes(sel) bytes (it must stay untouched until the solve ends)
   *  instead of the lane's own memory. */
  /** Energy and virial of system s from the last solve on the lane,
   *  computed on the device into d_energy (1 double) and d_virial (9),
   *  either may be null, in stream order. */
  void device_results(int lane, int s, double *d_energy, double *d_virial);
  void solve_on(int lane, std::vector<int> const &sel, ParticleArrays const &pa, unsigned flags,
                bool sync, std::vector<SolveResult> *out, void *scratch = nullptr,
                size_t scratch_size = 0);
  /** Enqueue on `lane`, for each system of `sel`: positions += sigma N(0,1)
   *  from a counter-based generator keyed by seeds[i], folded into the box,
   *  and forces zeroed.  The lane's previous work must be complete. */
  void prepare_on(int lane, std::vector<int> const &sel, double *pos, double *force,
                  double sigma, std::vector<unsigned long long> const &seeds);
  /** Device working memory a solve of `sel` needs beyond the resident
   *  systems (particle scratch). */
  size_t working_bytes(std::vector<int> const &sel) const;

  /** Set-up of allocation-free solves (solve_fixed): looks up this device's
   *  copy of the library's fixed scratch array, uploads the descriptors and
   *  creates the events of the fixed lane.  Allocates; call once after the
   *  systems are added (Solver::configure does). */
  void enable_fixed();
  /** Synchronous solve of `sel` whose every device buffer beyond the
   *  resident systems and the caller's arrays (particle scratch, brick
   *  counters and offsets, partial and per-system sums, work tables, error
   *  flag) is carved from the fixed scratch array of fixed_scratch_bytes():
   *  no device or pinned-host memory is allocated or freed.  Throws when the
   *  solve needs more than that array or enable_fixed() was not called (or
   *  systems were added since).  The array is shared by every context on
   *  the device: fixed solves that overlap on one device race. */
  void solve_fixed(std::vector<int> const &sel, ParticleArrays const &pa, unsigned flags,
                   std::vector<SolveResult> *out);
  /** Bytes of the fixed scratch a solve of `sel` needs. */
  size_t fixed_bytes(std::vector<int> const &sel) const;
  static size_t fixed_scratch_bytes();

  /** Copy the charge mesh of system s after spreading ([Mx][My][Mz]). */
  void debug_charge_mesh(int s, ParticleArrays const &pa, std::vector<double> &out);
  /** CA-local first stencil point of every particle of system s. */
  void debug_indices(int s, ParticleArrays const &pa, std::vector<int32_t> &out);

  /** Per-stage event ring for asynchronous solves (see Solver). */
  void profile_begin(int max_solves);
  int profile_end(double *stage_ms);
  float last_ms[8] = {};

  /** Capture the solve of `sel` (forces, no readback) in a hipGraph and
   *  replay it; the graph is rebuilt when the selection changes. */
  void graph_solve(std::vector<int> const &sel, ParticleArrays const &pa, unsigned flags);

  void set_launch_shape(LaunchShape const &s);
  /** Deterministic solves (FP64 meshes, 8^3 bricks; on by default for FP64
   *  batches, which cannot then take column-march shapes until it is turned
   *  off): the charge assignment
   *  sums 64-bit fixed-point contributions with integer atomics, so every
   *  solve of the same input gives the same bits, whatever the order of
   *  particles or of the device's atomics.  The mesh values are rounded to
   *  2^-s, s = 61 - ceil(log2(sum |q|)) (about 4e-13 at 10^6 unit charges).
   *  Solves that assign charges in several calls (the charge-assignment
   *  protocol) keep the floating-point mesh. */
  void set_deterministic(bool on);
  bool deterministic() const { return m_det; }
  /** Time force solves of `sel` for every candidate launch shape (gather
   *  mapping x FFT columns x z lines), `reps` solves each, and keep the
   *  fastest.  Forces are accumulated into pa.force on every solve. */
  AutotuneResult autotune(std::vector<int> const &sel, ParticleArrays const &pa, int reps);
  LaunchShape launch_shape() const { return m_shape; }
  hipStream_t stream() const { return m_stream; }
  int device() const { return m_device; }
  MeshPrecision precision() const { return m_prec; }

  struct Impl;

private:
  Impl *m;
  int m_device;
  MeshPrecision m_prec;
  hipStream_t m_stream = nullptr;
  LaunchShape m_shape;
  bool m_det = false;
};

void hip_check(hipError_t err, char const *what);

/** Real-space pair forces (A&S erfc) on `targets` from all n particles
 *  within r_cut, minimum image; device pointers, out[3 * ntargets]. */
void pair_forces_targets(int device, hipStream_t stream, int64_t n, double const *pos,
                         double const *q, int ntargets, int const *targets, Vec3d const &box,
                         double prefactor, double alpha, double r_cut, double *out);

} // namespace espresso_amd
