// Single-system P3M solver: a one-system batch (see p3m_solver.hpp).
#include "p3m_solver.hpp"

#include <stdexcept>

namespace espresso_amd {

Solver::Solver(int device) : m_device(device) {
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&m_idle_stream, hipStreamNonBlocking), "hipStreamCreate");
}

Solver::~Solver() {
  m_batch.reset();
  if (m_idle_stream) {
    (void)hipSetDevice(m_device);
    (void)hipStreamDestroy(m_idle_stream);
  }
}

hipStream_t Solver::stream() const { return m_batch ? m_batch->stream() : m_idle_stream; }

LocalMesh const &Solver::local_mesh() const {
  if (!m_batch) throw std::runtime_error("P3M solver used before configure()");
  return m_batch->local_mesh(0);
}

void Solver::set_deterministic(bool on) {
  if (m_batch) m_batch->set_deterministic(on);
  m_det = on ? 1 : 0;
}

void Solver::set_launch_shape(LaunchShape const &s) {
  m_shape = s;
  m_shape_set = true;
  if (m_batch) m_batch->set_launch_shape(s);
}

AutotuneResult Solver::autotune(int64_t n, double const *d_pos, double const *d_q,
                                double *d_force, int reps) {
  if (!m_batch) throw std::runtime_error("P3M solver used before configure()");
  m_batch->set_particle_count(0, n);
  AutotuneResult const r = m_batch->autotune({0}, ParticleArrays{d_pos, d_q, nullptr, d_force}, reps);
  m_shape = r.best;
  m_shape_set = true;
  return r;
}

void Solver::configure(P3MParams const &params, Vec3d const &box_l, double skin,
                       double prefactor, MeshPrecision precision) {
  if (m_batch) hip_check(hipStreamSynchronize(m_batch->stream()), "sync");
  hip_check(hipStreamSynchronize(m_idle_stream), "sync");
  m_batch.reset();
  auto b = std::make_unique<BatchSolver>(m_device, precision);
  if (m_det >= 0 && (m_det == 1 || precision == MeshPrecision::FP64))
    b->set_deterministic(m_det == 1); // before a column shape of a non-deterministic solver
  if (m_shape_set) b->set_launch_shape(m_shape);
  b->add_system(params, box_l, skin, prefactor, 0);
  b->set_convention(0, m_conv);
  b->enable_fixed(); // set-up of the allocation-free solves
  m_batch = std::move(b);
}

void Solver::device_results(double *d_energy, double *d_virial) {
  if (!m_batch) throw std::runtime_error("P3M solver used before configure()");
  m_batch->device_results(0, 0, d_energy, d_virial);
}

void Solver::set_convention(int conv) {
  if (conv != 0 && conv != 1) throw std::invalid_argument("P3M: convention must be 0 or 1");
  m_conv = conv;
  if (m_batch) m_batch->set_convention(0, conv);
}

void Solver::solve(int64_t n, double const *d_pos, double const *d_q, double const *d_upos,
                   double *d_force, unsigned flags, SolveResult *res, bool sync) {
  if (!m_batch) throw std::runtime_error("P3M solver used before configure()");
  m_batch->set_particle_count(0, n);
  ParticleArrays pa{d_pos, d_q, d_upos, d_force};
  std::vector<SolveResult> out;
  m_batch->solve({0}, pa, flags, sync, res ? &out : nullptr);
  if (!sync) return;
  for (int k = 0; k < 8; ++k) last_ms[k] = m_batch->last_ms[k];
  if (res) *res = out.at(0);
}

void Solver::solve_fixed(int64_t n, double const *d_pos, double const *d_q,
                         double const *d_upos, double *d_force, unsigned flags,
                         SolveResult *res) {
  if (!m_batch) throw std::runtime_error("P3M solver used before configure()");
  m_batch->set_particle_count(0, n);
  ParticleArrays pa{d_pos, d_q, d_upos, d_force};
  std::vector<SolveResult> out;
  m_batch->solve_fixed({0}, pa, flags, res ? &out : nullptr);
  for (int k = 0; k < 8; ++k) last_ms[k] = m_batch->last_ms[k];
  if (res) *res = out.at(0);
}

void Solver::debug_charge_mesh(int64_t n, double const *d_pos, double const *d_q,
                               std::vector<double> &out) {
  if (!m_batch) throw std::runtime_error("P3M solver used before configure()");
  m_batch->set_particle_count(0, n);
  m_batch->debug_charge_mesh(0, ParticleArrays{d_pos, d_q, nullptr, nullptr}, out);
}

void Solver::debug_indices(int64_t n, double const *d_pos, double const *d_q,
                           std::vector<int32_t> &out) {
  if (!m_batch) throw std::runtime_error("P3M solver used before configure()");
  m_batch->set_particle_count(0, n);
  m_batch->debug_indices(0, ParticleArrays{d_pos, d_q, nullptr, nullptr}, out);
}

void Solver::profile_begin(int max_solves) {
  if (!m_batch) throw std::runtime_error("P3M solver used before configure()");
  m_batch->profile_begin(max_solves);
}

int Solver::profile_end(double *stage_ms) {
  if (!m_batch) throw std::runtime_error("P3M solver used before configure()");
  return m_batch->profile_end(stage_ms);
}

} // namespace espresso_amd
