/*
 * espresso_amd_p3m.h -- C ABI of the MI355X P3M k-space Coulomb solver.
 *
 * Drop-in boundary for the long-range part of ESPResSo's P3M Coulomb actor.
 * Each entry point names the reference interface it replaces (paths are
 * relative to the ESPResSo source tree):
 *   src/core/electrostatics/p3m.hpp            struct CoulombP3M
 *   src/core/electrostatics/p3m.impl.hpp       CoulombP3MImpl / new_p3m_handle
 *   src/core/electrostatics/p3m_gpu_cuda.cuh   p3m_gpu_init / _add_farfield_force
 *   src/core/electrostatics/p3m_gpu_error.hpp  p3m_k_space_error_gpu
 *   src/core/p3m/common.hpp                    P3MParameters
 *
 * Conventions
 *  - Every function returns an int status (P3M_OK = 0); nothing throws across
 *    the ABI.  The message of the last failure on a context is returned by
 *    p3m_last_error(); context-less calls use p3m_last_global_error().
 *  - Status codes map to the reference's exception types:
 *      P3M_ERR_DOMAIN    -> std::domain_error      (parameter validation)
 *      P3M_ERR_RUNTIME   -> std::runtime_error     (sanity checks, tuning)
 *      P3M_ERR_INVALID   -> std::invalid_argument
 *      P3M_ERR_DEVICE    -> HIP runtime failure    (cuda_fatal_error analogue)
 *  - Positions are AoS [n][3] doubles, charges [n] doubles, forces [n][3]
 *    doubles accumulated in place (force -= ...), as the reference does on
 *    particles.  With P3M_DEVICE_POINTERS set, the three arrays are device
 *    pointers and nothing crosses PCIe; otherwise they are host pointers.
 *  - A context owns its device buffers and one HIP stream.  Calls on one
 *    context are serialised by the caller (one context per host thread).
 */
#ifndef ESPRESSO_AMD_P3M_H
#define ESPRESSO_AMD_P3M_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define P3M_OK 0
#define P3M_ERR_DOMAIN 1
#define P3M_ERR_RUNTIME 2
#define P3M_ERR_INVALID 3
#define P3M_ERR_DEVICE 4
#define P3M_ERR_MEMORY 5

/* solve flags */
#define P3M_FORCES 1u
#define P3M_ENERGY 2u
#define P3M_PRESSURE 4u
#define P3M_DEVICE_POINTERS 8u
#define P3M_ASYNC 16u /* enqueue only: no host sync, no energy/virial output;
                         requires P3M_DEVICE_POINTERS */
#define P3M_DEVICE_RESULTS 32u /* energy / virial arguments are device arrays
                         (1 and 9 doubles) written in stream order, also
                         with P3M_ASYNC; requires P3M_DEVICE_POINTERS, not
                         with ELC */

/* mesh precision */
#define P3M_PREC_FP64 0 /* plain FP64 solver (drop-in for P3M, double) */
#define P3M_PREC_FP32 1 /* FP32 mesh, FP64 weights/reductions/forces    */

/* real-space cost in timed tuning (p3m_set_tuning_real_space) */
#define P3M_TUNE_RS_MODEL 0  /* host pair loop modelled at 1 ns per pair     */
#define P3M_TUNE_RS_DEVICE 1 /* the device pair loop p3m_short_range, timed */

/* force conventions (p3m_set_force_convention) */
#define P3M_CONV_CPU 0 /* CPU kernels: CA mesh with mesh_off, Nyquist ik 0  */
#define P3M_CONV_GPU 1 /* GPU kernels: FP32 index, no mesh_off, +M/2 ik     */

typedef struct p3m_context p3m_context;

/* P3M parameters.  Field meaning and the -1 "to be tuned" sentinels follow
 * P3MParameters (src/core/p3m/common.hpp:59-150) and the Python actor
 * defaults (src/python/espressomd/electrostatics.py:134-146). */
typedef struct {
  int mesh[3];        /* mesh points per direction, {-1,-1,-1} = tune      */
  int cao;            /* charge assignment order 1..7, -1 = tune           */
  double alpha;       /* Ewald splitting parameter, -1 = tune              */
  double r_cut;       /* real-space cutoff, -1 = tune                      */
  double mesh_off[3]; /* first mesh point offset in [0,1], -1 = 0.5        */
  double epsilon;     /* surrounding dielectric, 0 = metallic              */
  double accuracy;    /* target rms force error                            */
  double prefactor;   /* electrostatics prefactor (Bjerrum length * kT)    */
  double box_l[3];    /* box lengths                                       */
  double skin;        /* Verlet skin: fixes the charge-assignment mesh     */
  int precision;      /* P3M_PREC_FP64 or P3M_PREC_FP32                    */
  int tune;           /* nonzero: parameters still to be tuned             */
  int timings;        /* tuning repetitions per trial (> 0)                */
  int verbose;        /* tuning log to stdout                              */
  double charge_neutrality_tolerance; /* -1 disables the check (actor.hpp:38) */
} p3m_params;

/* Fill defaults of the P3M actor (electrostatics.py:134-146). */
int p3m_params_default(p3m_params *p);

/* Create/destroy a solver context on HIP device `device` (device -1 gives a
 * host-only context: parameter validation and error estimates, no solves).
 * Replaces new_p3m_handle<...> (p3m.impl.hpp:186-197) and the lazily created
 * P3MGpuParams owner (p3m_gpu_cuda.cu:116-139). */
int p3m_create(int device, p3m_context **out);
int p3m_destroy(p3m_context *ctx);

/* Validate and store parameters: the P3MParameters constructor
 * (common.hpp:84-150) and CoulombP3MImpl constructor (p3m.impl.hpp:81-92),
 * with identical messages.  Does not build meshes. */
int p3m_set_params(p3m_context *ctx, const p3m_params *p);

/* Current (possibly tuned) parameters; mirrors the read-only getters of the
 * script interface (src/script_interface/electrostatics/CoulombP3M.hpp:66-100). */
int p3m_get_params(const p3m_context *ctx, p3m_params *out);
int p3m_is_tuned(const p3m_context *ctx, int *out);

/* Activation: sanity checks, tuning if requested, then init.
 * Replaces CoulombP3MImpl::on_activation / tune / init (p3m.impl.hpp:124-139,
 * p3m.cpp:757-791, 245-280, p3m_gpu_cuda.cu:549-667).  Particles are needed
 * to count charges (count_charged_particles, p3m.cpp:97-115) and for the
 * neutrality check (coulomb.cpp:286-325); pointers follow `flags`. */
int p3m_activate(p3m_context *ctx, int64_t n, const double *pos, const double *q,
                 unsigned flags);

/* Long-range solve.  P3M_FORCES: force[i] -= q_i E(r_i) * prefactor, the
 * k-space part of add_long_range_forces (p3m.cpp:452-549); P3M_ENERGY:
 * long_range_energy (p3m.cpp:551-589) into *energy; P3M_PRESSURE: the
 * k-space virial tensor of long_range_pressure (p3m.cpp:393-449) into
 * virial[9] (row-major, not yet divided by the volume, as the reference
 * returns it).  `pos_unfolded` (may be NULL = pos) feeds the dipole term
 * when epsilon is not metallic (p3m.cpp:375-386). */
int p3m_solve(p3m_context *ctx, int64_t n, const double *pos, const double *q,
              const double *pos_unfolded, double *force, unsigned flags, double *energy,
              double *virial);

/* p3m_solve without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code that runs under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every working buffer of the solve (particle scratch, brick counters and
 * offsets, partial and per-system sums, work tables, error flag) is carved
 * from one fixed device array of p3m_noalloc_scratch_bytes() (48 MiB) that
 * the library holds on each device until the process ends; the meshes are
 * the context's own, allocated when p3m_activate set it up.  Requirements:
 * an activated context, P3M_DEVICE_POINTERS, no ELC; synchronous (no
 * P3M_ASYNC or P3M_DEVICE_RESULTS).  A solve needing more scratch than the
 * array (about 44 bytes per particle: some 1.1 million particles) fails
 * with P3M_ERR_INVALID.  The array is shared by every context on a device
 * and keeps the last call's contents: calls that overlap on one device
 * race. */
int p3m_solve_noalloc(p3m_context *ctx, int64_t n, const double *pos, const double *q,
                      const double *pos_unfolded, double *force, unsigned flags, double *energy,
                      double *virial);
/* Size of the fixed scratch array of p3m_solve_noalloc in bytes. */
int p3m_noalloc_scratch_bytes(int64_t *bytes);

/* ------------------------------------------------------------------------
 * Charge-assignment protocol of CoulombP3M, as the reference's
 * ElectrostaticLayerCorrection drives it (elc.cpp:1014, 1110-1141, 1143-1187,
 * 1189-1256): the caller fills the charge mesh itself (real charges, then
 * image charges), sets the charge sums, and runs the k-space kernel on that
 * mesh without a fresh assignment (long_range_kernel under ELC,
 * p3m.cpp:463-468).  All calls are synchronous; arrays are host pointers, or
 * device pointers with P3M_DEVICE_POINTERS.
 * ------------------------------------------------------------------------ */
/* prepare_fft_mesh(reset_weights) (p3m.impl.hpp:161-168): zero the charge
 * mesh; with reset_weights also empty the interpolation cache. */
int p3m_prepare_fft_mesh(p3m_context *ctx, int reset_weights);
/* assign_charge(q, real_pos, skip_cache) (p3m.hpp:135-141, p3m.cpp:333-345)
 * for n charges at once: each (q[i], pos[i]) with q[i] != 0 is added to the
 * mesh with the CA-mesh weights of p3m_calculate_interpolation_weights;
 * positions outside the box wrap periodically.  Without skip_cache every
 * (pos[i], q[i]) is appended to the interpolation cache, in order. */
int p3m_assign_charges(p3m_context *ctx, int64_t n, const double *pos, const double *q,
                       int skip_cache, unsigned flags);
/* charge_assign(particles) (p3m.cpp:321-331): p3m_prepare_fft_mesh(ctx, 1)
 * then p3m_assign_charges(..., skip_cache = 0). */
int p3m_charge_assign(p3m_context *ctx, int64_t n, const double *pos, const double *q,
                      unsigned flags);
/* count_charged_particles (p3m.cpp:97-115): sum_qpart, sum_q2 and
 * square_sum_q of the charges q[n]; p3m_activate sets them too. */
int p3m_count_charged_particles(p3m_context *ctx, int64_t n, const double *q, unsigned flags);
/* count_charged_particles_elc(n, sum_q2, square_sum_q) (p3m.impl.hpp:105-110). */
int p3m_count_charged_particles_elc(p3m_context *ctx, int n, double sum_q2,
                                    double square_sum_q);
/* adapt_epsilon_elc (p3m.impl.hpp:111-113): epsilon becomes metallic. */
int p3m_adapt_epsilon_elc(p3m_context *ctx);
/* Number of entries in the interpolation cache. */
int p3m_assigned_count(const p3m_context *ctx, int64_t *n);
/* long_range_kernel on the assigned mesh (p3m.cpp:452-592 with the charge
 * assignment skipped, as under ELC): forward FFT, influence function and ik.
 * P3M_FORCES: force has one [3] row per cache entry, in cache order, and
 * row k -= q_k E(r_k) * prefactor / V (AssignForces with the cached weights,
 * p3m.cpp:344-373), plus the dipole term when epsilon is not metallic (the
 * moment of the cache entries; pos_unfolded, one row per entry or NULL, gives
 * their unfolded positions).  P3M_ENERGY: *energy = the mesh energy with the
 * self, net-charge and dipole corrections of the current charge sums
 * (p3m.cpp:551-589).  Nothing is done when sum_q2 is zero.  The mesh is
 * consumed: assign again before the next call. */
int p3m_solve_assigned(p3m_context *ctx, const double *pos_unfolded, double *force,
                       unsigned flags, double *energy);

/* Stage outputs for parity checks (device-side meshes copied to host):
 * charge mesh after the periodic wrap, [Mx][My][Mz] real. */
int p3m_debug_charge_mesh(p3m_context *ctx, int64_t n, const double *pos, const double *q,
                          double *mesh_out);
/* Per-particle global mesh index of the first stencil point (bit-exact
 * counterpart of p3m_calculate_interpolation_weights, interpolation.hpp:138-178),
 * -1 for uncharged particles; `nmp` gets the CA-mesh-local point [n][3]. */
int p3m_debug_indices(p3m_context *ctx, int64_t n, const double *pos, const double *q,
                      int32_t *nmp);

/* k-space rms force error estimate (p3m.cpp:187-243; GPU twin
 * p3m_k_space_error_gpu, p3m_gpu_error.hpp:31-33). */
int p3m_k_space_error(double prefactor, const int mesh[3], int cao, int n_c_part,
                      double sum_q2, double alpha_L, const double box_l[3], double *out);
/* real-space rms force error (Kolafa-Perram, p3m.cpp:145-160). */
int p3m_real_space_error(double prefactor, double r_cut_iL, int n_c_part, double sum_q2,
                         double alpha_L, const double box_l[3], double *out);

/* Real-space (short-range) part on the device for all n particles: every
 * pair of charged particles closer than the context's r_cut (minimum image,
 * r_cut below half the box) adds the force of CoulombP3M::pair_force
 * (p3m.hpp:146-165, A&S erfc of AS_erfc_part.hpp:29-40) to force, as the
 * reference's short-range loop over its cell system applies it
 * (forces.cpp:178-208, force on p1 += F(d), d = p1 - p2).  Cell lists of
 * edge r_cut/2 or r_cut (all pairs for boxes under three cells wide).  flags:
 * P3M_DEVICE_POINTERS, P3M_ASYNC (device pointers, no energy/pressure);
 * P3M_ENERGY sets *energy to the sum of CoulombP3M::pair_energy
 * (p3m.hpp:167-180), P3M_PRESSURE sets virial[9] to sum_pairs F (x) d. */
int p3m_short_range(p3m_context *ctx, int64_t n, const double *pos, const double *q,
                    unsigned flags, double *force, double *energy, double *virial);
/* Cells per r_cut of the short-range cell lists: 0 = automatic (2 where the
 * box is at least 2.5 r_cut wide), 1 or 2 (falls back to fewer cells, then
 * all pairs, in narrow boxes). */
int p3m_set_short_range_subdivision(p3m_context *ctx, int cells_per_cutoff);

/* Force convention of p3m_solve.  P3M_CONV_GPU gives the forces of the
 * reference's P3MGPU actor (p3m_gpu_add_farfield_force,
 * p3m_gpu_cuda.cu:672-723, called from add_long_range_forces_gpu,
 * p3m.cpp:867-888): first stencil point nmp = floor(m + 1/2) with
 * m = fmaf(float(x), float(M) / float(L), -(float)((cao - 1) / 2)) on the
 * folded position (p3m_gpu_cuda.cu:334-345), weights at m - nmp, ik with
 * the Nyquist wave number kept at +M/2 (apply_diff_op, :265-295) and no
 * dipole force term; mesh_off is ignored.  Energy and virial keep the CPU
 * kernels' mesh, as the reference's GPU actor computes them on the CPU
 * (p3m.cpp:889-892): a solve with P3M_FORCES and P3M_ENERGY or
 * P3M_PRESSURE then runs twice.  p3m_debug_indices reports the periodic
 * index of nmp under P3M_CONV_GPU.  Default P3M_CONV_CPU. */
int p3m_set_force_convention(p3m_context *ctx, int convention);
int p3m_get_force_convention(const p3m_context *ctx, int *convention);

/* ELC: electrostatic layer correction for slab systems (charges in
 * 0 <= z <= L_z - gap_size, periodic in x and y).  Field meaning, defaults
 * and checks follow elc_data (src/core/electrostatics/elc.hpp:62-118,
 * elc.cpp:1056-1102) and the Python actor
 * (src/python/espressomd/electrostatics.py, class ELC). */
typedef struct {
  double maxPWerror;    /* maximal pairwise error (required, > 0)          */
  double gap_size;      /* empty layer below L_z (required, > 0)           */
  double far_cut;       /* cutoff of the far-field sums, -1 = tune         */
  int neutralize;       /* keep P3M's neutralising background (default 1)  */
  double delta_mid_top; /* dielectric contrast at the top boundary, [-1, 1] */
  double delta_mid_bot; /* dielectric contrast at the bottom boundary       */
  int const_pot;        /* constant potential (with dielectric contrasts)  */
  double pot_diff;      /* applied potential difference (needs const_pot) */
} p3m_elc_params;

/* Defaults of the Python actor: far_cut -1, neutralize 1, no contrasts,
 * no const_pot; maxPWerror and gap_size -1 (must be set). */
int p3m_elc_params_default(p3m_elc_params *p);

/* Put ELC on top of the context's P3M solver, or remove it (p == NULL).
 * Replaces the ElectrostaticLayerCorrection constructor with
 * adapt_epsilon_elc (elc.cpp:1104-1108, 1010-1018; p3m.impl.hpp:112-114):
 * the P3M epsilon becomes metallic.  From then on p3m_solve adds the ELC
 * far-field forces and energy of ElectrostaticLayerCorrection::
 * add_long_range_forces / long_range_energy (elc.cpp:1189-1256,
 * 830-940); P3M_PRESSURE returns a zero virial, as the reference has no
 * ELC pressure (coulomb.cpp:125-131).  A charged particle in the gap is
 * reported by the solve (check_gap, elc.cpp:184-195).  Parameter errors
 * carry the reference's messages.  Dielectric contrasts (delta_mid_top/bot)
 * put image charges of the boundary layers into the mesh and add their pair
 * and self terms (elc.cpp:1110-1256, elc.hpp:140-171, 250-296); const_pot
 * with pot_diff adds the applied field (elc.cpp:201-342).  The reference's
 * own ELC actor can also drive the solver through the charge-assignment
 * protocol above (p3m_prepare_fft_mesh ... p3m_solve_assigned). */
int p3m_set_elc(p3m_context *ctx, const p3m_elc_params *p);
/* Current ELC parameters (far_cut as tuned by the last activation or
 * solve) and whether ELC is on; read-only getters of the script interface
 * (src/script_interface/electrostatics/ElectrostaticLayerCorrection.hpp:60-85). */
int p3m_get_elc(const p3m_context *ctx, p3m_elc_params *out, int *enabled);

/* ------------------------------------------------------------------------
 * Dipolar P3M (magnetostatics), DipolarP3MImpl<double, CPU>
 * (src/core/magnetostatics/dp3m.cpp, dp3m.hpp, dp3m.impl.hpp) for a cubic
 * box and mesh.  Parameters and -1 sentinels follow the Python actor
 * espressomd.magnetostatics.DipolarP3M (magnetostatics.py, class
 * DipolarP3M): mesh, cao, alpha, r_cut tuned when -1; epsilon 0 = metallic.
 * Arrays are [n][3] (positions, dipole moments, forces, torques), host or
 * device pointers by P3M_DEVICE_POINTERS; forces and torques are added in
 * place.
 * ------------------------------------------------------------------------ */
typedef struct p3m_dipolar p3m_dipolar;

typedef struct {
  int mesh;           /* cubic mesh, -1 = tune                              */
  int cao;            /* 1..7, -1 = tune                                    */
  double alpha;       /* -1 = tune                                          */
  double r_cut;       /* -1 = tune                                          */
  double accuracy;    /* target rms force error (required)                  */
  double prefactor;   /* mu_0 / (4 pi)                                      */
  double epsilon;     /* surrounding permeability, 0 = metallic             */
  double mesh_off[3]; /* default 0.5                                        */
  double box_l;       /* cubic box edge                                     */
  double skin;        /* Verlet skin: fixes the charge-assignment mesh      */
  int tune;           /* nonzero: tune what is -1                           */
  int timings;        /* > 0                                                */
  int verbose;
} p3m_dipolar_params;

int p3m_dipolar_params_default(p3m_dipolar_params *p);
/* Context on HIP device `device` (-1: host only, parameter checks only);
 * replaces the DipolarP3MImpl handle (dp3m.impl.hpp:79-92). */
int p3m_dipolar_create(int device, p3m_dipolar **out);
int p3m_dipolar_destroy(p3m_dipolar *ctx);
const char *p3m_dipolar_last_error(const p3m_dipolar *ctx);
/* Validate and store parameters (dp3m.impl.hpp:79-92, common.hpp:94-150). */
int p3m_dipolar_set_params(p3m_dipolar *ctx, const p3m_dipolar_params *p);
int p3m_dipolar_get_params(const p3m_dipolar *ctx, p3m_dipolar_params *out);
/* count_magnetic_particles and tune (dp3m.cpp:83-97, 646-682; error
 * estimates :684-820): for each (mesh, cao) the smallest r_cut meeting the
 * accuracy, alpha from the real-space error; (mesh, cao) by a cost model.
 * Then init (sanity_checks_boxl, :822-846; influence functions).  The
 * tuning is by cost model, so only the dipole moments are read; pos is
 * accepted (and may be null) to keep the signature of p3m_activate. */
int p3m_dipolar_activate(p3m_dipolar *ctx, int64_t n, const double *pos, const double *dip,
                         unsigned flags);
/* k-space part: P3M_FORCES adds forces and torques
 * (long_range_kernel with force_flag, dp3m.cpp:313-412, plus the surface
 * term :451-517), P3M_ENERGY returns the k-space energy with self,
 * correction and surface terms (:270-311, 892-901). */
int p3m_dipolar_solve(p3m_dipolar *ctx, int64_t n, const double *pos, const double *dip,
                      double *force, double *torque, unsigned flags, double *energy);
/* Real-space pair terms of DipolarP3M::pair_force / pair_energy
 * (dp3m.hpp:160-235) over all pairs within r_cut (minimum image, r_cut
 * below half the box): forces and torques added, energy with P3M_ENERGY. */
int p3m_dipolar_short_range(p3m_dipolar *ctx, int64_t n, const double *pos, const double *dip,
                            double *force, double *torque, unsigned flags, double *energy);
/* Cells per r_cut of p3m_dipolar_short_range: 0 = by occupancy (r_cut/2
 * cells when they hold about 40 dipoles or more), 1 or 2. */
int p3m_dipolar_set_short_range_subdivision(p3m_dipolar *ctx, int cells_per_cutoff);
int p3m_dipolar_get_stream(const p3m_dipolar *ctx, void **stream);
int p3m_dipolar_synchronize(p3m_dipolar *ctx);

/* Real-space cost that timed tuning (p3m_activate with tune set) adds to
 * each timed k-space solve.  The reference times whole integration steps,
 * real-space loop included (TuningAlgorithm.cpp:116-211 via benchmark_integration_step, tuning.cpp:73);
 * P3M_TUNE_RS_DEVICE times p3m_short_range on the activation particles for
 * each trial (r_cut, alpha), for callers whose pair loop runs there;
 * P3M_TUNE_RS_MODEL (default) models a host loop at 1 ns per pair. */
int p3m_set_tuning_real_space(p3m_context *ctx, int mode);
/* Cells per dimension used by the last p3m_short_range call (0 0 0 before
 * the first; 1 1 1 for all pairs). */
int p3m_short_range_cells(const p3m_context *ctx, int cells[3]);

/* Real-space pair forces of CoulombP3M::pair_force (p3m.hpp:146-165, A&S
 * erfc) on the ntargets particles listed in targets from all n particles
 * within the context's r_cut (minimum image, r_cut < half the box): device
 * pointers pos [n][3], q [n], targets [ntargets], out [ntargets][3]
 * (overwritten).  A scan over all sources per target: completes the total
 * force of selected particles, e.g. for error measurements. */
int p3m_pair_forces_targets(p3m_context *ctx, int64_t n, const double *pos, const double *q,
                            int ntargets, const int32_t *targets, double *out);

/* Real-space pair terms CoulombP3M::pair_force / pair_energy (p3m.hpp:146-180). */
int p3m_pair_force(const p3m_context *ctx, double q1q2, const double d[3], double dist,
                   double out[3]);
int p3m_pair_energy(const p3m_context *ctx, double q1q2, double dist, double *out);

/* Timing of the last solve on the device (ms), from HIP events on the
 * context stream: [0] total, then per stage (bin, spread, fft_fwd,
 * convolve, fft_bwd, gather). */
int p3m_last_timings(const p3m_context *ctx, double *ms, int n);

/* Profiling of asynchronous solves: per-stage HIP events of the next
 * `max_solves` solves are recorded into a ring on the context stream with no
 * host synchronisation; p3m_profile_end synchronises and returns the number
 * of profiled solves and the summed stage times in ms (7 values, order as in
 * p3m_last_timings). */
int p3m_profile_begin(p3m_context *ctx, int max_solves);
int p3m_profile_end(p3m_context *ctx, double *stage_ms, int *nsolves);

/* Launch shape of the device kernels (no counterpart in the reference). */
typedef struct p3m_launch_shape {
  int gather;    /* gather mapping: 0 = lanes over stencil layers, 1/2/3 = lanes
                    over the (j, k) plane in segments of up to 64/16/8 lanes */
  int ncol;      /* kz columns per block of the split y/x FFT passes
                    (power of two in [1, 64]) */
  int zlines;    /* pairs of real lines per block of the z passes (power of two
                    in [1, 64]) */
  int fused_lds; /* LDS bytes up to which a system uses the fused y/x pass (its
                    two x-y planes in LDS; 0 = never, at most 131072) */
  int spread;    /* spread mapping: 0/1/2 = segments of up to 64/16/8 lanes per
                    particle over the (j, k) plane */
  int threads;   /* block size of the spread and gather kernels: 256 or 512 */
  int chunk;     /* particles staged per brick pass: 1..64, at most threads/3 */
  int column;    /* particle bins of spread and gather: 0 = 8^3 bricks of first
                    stencil points; 8 or 16 = column march over W x W columns in
                    (y, z), one bin per x layer, a ring of the cao mesh planes a
                    layer touches in LDS (256-thread blocks) */
  int seg;       /* column march: x layers per block (1..1024) */
  int col_lanes; /* column march: lanes per particle (8 or 16) */
  int group;     /* spread: bricks per block sharing one tile and its halo: 1
                    or 4 (2x2x1, for sparse bricks; needs column 0 and the
                    plane spread mapping, spread 0) */
  int gather_threads; /* block size of the brick gather: 0 = threads, or 256
                    or 512 (chunk at most gather_threads/3); dense FP64 bricks
                    gather fastest in 512-thread blocks while their spread
                    prefers 256 */
} p3m_launch_shape;

/* Launch-shape autotuner of the context's device kernels (no counterpart in
 * the reference; see p3m_batch_autotune for the candidates): times `reps`
 * force solves of the device arrays pos [n][3], q [n] per candidate and keeps
 * the fastest for later solves.  Outputs may be NULL. */
int p3m_autotune(p3m_context *ctx, int64_t n, const double *pos, const double *q, int reps,
                 p3m_launch_shape *best, double *best_ms, double *default_ms);

/* Set (validated as documented in p3m_launch_shape) or read the launch
 * shape of the context's later solves (no counterpart in the reference). */
int p3m_set_launch_shape(p3m_context *ctx, const p3m_launch_shape *shape);
int p3m_get_launch_shape(const p3m_context *ctx, p3m_launch_shape *shape);

/* Deterministic solves (no counterpart in the reference, whose CPU solver
 * is deterministic by construction; on by default for FP64 contexts): the
 * charge assignment rounds every contribution once to a 64-bit fixed-point
 * value (resolution 2^-s with s = 61 - ceil(log2(sum |q|)), about 4e-13 for
 * 10^6 unit charges) and sums with integer atomics into a fixed-point mesh,
 * so repeated solves of the same input give the same bits.  FP64 precision
 * and 8^3 bricks only: a column-march launch shape needs it off; the
 * charge-assignment protocol calls keep the floating-point mesh. */
int p3m_set_deterministic(p3m_context *ctx, int on);

/* Block until all work enqueued on the context stream has finished. */
int p3m_synchronize(p3m_context *ctx);

/* Device stream of the context (hipStream_t) for callers that enqueue
 * around the solver. */
int p3m_get_stream(const p3m_context *ctx, void **stream);

const char *p3m_last_error(const p3m_context *ctx);
const char *p3m_last_global_error(void);

/* ------------------------------------------------------------------------
 * Batched solver: many independent systems per launch.  No counterpart in
 * the reference, which solves one system per process; it is the same
 * k-space algorithm as p3m_solve applied to every selected system.  All
 * particle arrays are device pointers to the systems' particles concatenated
 * in the order the systems were added (offsets from
 * p3m_batch_particle_offset).
 * ------------------------------------------------------------------------ */
typedef struct p3m_batch p3m_batch;

int p3m_batch_create(int device, int precision, p3m_batch **out);
int p3m_batch_destroy(p3m_batch *b);
/* Add a system with fixed (tuned) parameters and n particles; validation and
 * sanity checks as in p3m_set_params / activation. */
int p3m_batch_add_system(p3m_batch *b, const p3m_params *p, int64_t n, int *sys_id);
int p3m_batch_particle_offset(const p3m_batch *b, int sys_id, int64_t *offset);
int p3m_batch_system_bytes(const p3m_batch *b, int sys_id, int64_t *bytes);
/* Device working memory (particle scratch) of a solve of the selection,
 * beyond the resident systems. */
int p3m_batch_working_bytes(const p3m_batch *b, int nsel, const int *sys_ids, int64_t *bytes);
/* Solve the selected systems: flags P3M_FORCES / P3M_ENERGY / P3M_PRESSURE,
 * P3M_ASYNC (enqueue only).  energy[nsel] and virial[9*nsel] may be NULL. */
int p3m_batch_solve(p3m_batch *b, int nsel, const int *sys_ids, const double *pos,
                    const double *q, double *force, unsigned flags, double *energy,
                    double *virial);
/* Force solve of the selection replayed from a captured hipGraph (captured on
 * first use, re-captured when the selection or arrays change); enqueue only. */
int p3m_batch_graph_solve(p3m_batch *b, int nsel, const int *sys_ids, const double *pos,
                          const double *q, double *force);
/* Set (validated as documented in p3m_launch_shape) or read the launch shape
 * of later solves. */
int p3m_batch_set_launch_shape(p3m_batch *b, const p3m_launch_shape *shape);
int p3m_batch_get_launch_shape(const p3m_batch *b, p3m_launch_shape *shape);
/* Deterministic solves of an FP64 batch (see p3m_set_deterministic; on by
 * default for FP64 batches). */
int p3m_batch_set_deterministic(p3m_batch *b, int on);
/* Autotuner: time `reps` force solves of the selection for candidate launch
 * shapes (coordinate descent over fused pass, particle-mesh block size and
 * chunk, gather and spread mappings, FFT columns and z lines) and keep the
 * fastest; returns the chosen shape (`best` may be NULL) and the mean solve
 * time before (default_ms) and after (best_ms).  `force` is scratch: forces
 * accumulate into it on every timed solve. */
int p3m_batch_autotune(p3m_batch *b, int nsel, const int *sys_ids, const double *pos,
                       const double *q, double *force, int reps, p3m_launch_shape *best,
                       double *best_ms, double *default_ms);
int p3m_batch_synchronize(p3m_batch *b);
int p3m_batch_get_stream(const p3m_batch *b, void **stream);
int p3m_batch_profile_begin(p3m_batch *b, int max_solves);
int p3m_batch_profile_end(p3m_batch *b, double *stage_ms, int *nsolves);
const char *p3m_batch_last_error(const p3m_batch *b);

/* ------------------------------------------------------------------------
 * Serving: continuous batching over the resident systems of a batch.  No
 * counterpart in the reference.  Requests (arrival time, system, seed) are
 * admitted into batches as lanes (streams) and working memory free up; the
 * particle scratch of every batch in flight comes from one device arena of
 * budget_bytes.  Per request the system's positions get a Gaussian
 * displacement of width sigma (counter-based draws keyed by the seed, folded
 * into the box), its forces are cleared and recomputed.  A server keeps the
 * batch's solver alive: the batch handle may be destroyed before it.
 * ------------------------------------------------------------------------ */
typedef struct p3m_server p3m_server;

int p3m_serve_create(p3m_batch *b, int lanes, int64_t budget_bytes, int max_batch,
                     double sigma, p3m_server **out);
int p3m_serve_destroy(p3m_server *s);
/* Serve n_req requests sorted by arrival_s (seconds after the call starts)
 * to completion.  pos, q, force: device arrays of all systems of the batch.
 * The serving streams first wait for the work enqueued on the batch's stream
 * (p3m_batch_get_stream) before the call.
 * force_out (device, may be NULL): each request's forces are copied to
 * force_out + 3 * out_off[r] when its batch completes.  Outputs: latency_s
 * [n_req] (arrival to completion), wall_s, n_batches, peak_bytes (working
 * memory in flight), peak_lanes (batches in flight at once). */
int p3m_serve_run(p3m_server *s, int n_req, const double *arrival_s, const int *sys_ids,
                  const uint64_t *seeds, double *pos, const double *q, double *force,
                  double *force_out, const int64_t *out_off, double *latency_s, double *wall_s,
                  int *n_batches, int64_t *peak_bytes, int *peak_lanes);
const char *p3m_serve_last_error(const p3m_server *s);

/* Parameter tuning without a device: the reference's tuning algorithm
 * (p3m.cpp:594-791, TuningAlgorithm.cpp:57-328) with a modelled cost in place
 * of timed integration steps (device k-space stage rates of an FP64 mesh, or
 * of an FP32 mesh if single_precision, plus a real-space loop at pair_ns
 * nanoseconds per pair; pair_ns <= 0 selects 1 ns), meshes restricted to
 * [mesh_min, mesh_max] and orders to [cao_min, cao_max].
 * Fills mesh, cao, r_cut, alpha, accuracy. */
int p3m_tune_model(p3m_params *inout, int n_charged, double sum_q2, int mesh_min,
                   int mesh_max, int cao_min, int cao_max, double pair_ns,
                   int single_precision);

#ifdef __cplusplus
}
#endif

#endif /* ESPRESSO_AMD_P3M_H */
