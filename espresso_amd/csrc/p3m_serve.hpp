// Continuous-batching request loop over resident P3M systems (MI355X).
//
// Systems live in a BatchSolver with their meshes, influence functions,
// positions and forces resident in device memory.  Requests (system id,
// arrival time, displacement seed) are admitted into batches as lanes and
// working memory free up: a free lane takes pending requests in arrival
// order, skipping systems already in flight, until the batch reaches
// max_batch systems or its particle working set would push the memory in
// flight past the budget.  Each batch runs on its own lane (stream), so
// batches overlap on the device; a new batch starts as soon as a lane
// retires, while others are still running.
//
// Per request the positions of its system get a Gaussian displacement
// (sigma, counter-based draws keyed by the request seed), forces are
// cleared, and one k-space force solve runs.  Latency is measured from the
// request's arrival time to the host observing its batch's completion event.
#pragma once

#include "p3m_batch.hpp"

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace espresso_amd {

struct ServeConfig {
  int lanes = 4;                   // concurrent batches (streams)
  size_t budget_bytes = 256 << 20; // device working memory of batches in flight
  int max_batch = 64;              // systems per batch
  double sigma = 0.05;             // displacement per request
};

struct ServeRequest {
  double arrival_s = 0.; // seconds after the start of the run
  int sys = 0;
  unsigned long long seed = 0;
};

struct ServeStats {
  std::vector<double> latency_s; // per request, arrival to completion
  double wall_s = 0.;            // first arrival to last completion
  int batches = 0;
  size_t peak_bytes = 0;         // largest working memory in flight
  int peak_lanes = 0;            // most batches in flight at once
};

/** First-fit allocator of offsets in a fixed block of `size` bytes. */
class ArenaAllocator {
public:
  explicit ArenaAllocator(size_t size);
  /** Offset of a free range of `bytes` (256-byte aligned), or SIZE_MAX. */
  size_t allocate(size_t bytes);
  void release(size_t offset, size_t bytes);
  size_t free_bytes() const;

private:
  std::vector<std::pair<size_t, size_t>> m_free; // (offset, size), sorted by offset
};

class Server {
public:
  /** Allocates the working-memory arena (cfg.budget_bytes of device memory)
   *  and the lanes; every batch's particle scratch is a region of the arena,
   *  so the working memory of batches in flight never exceeds the budget. */
  Server(BatchSolver &batch, ServeConfig const &cfg);
  ~Server();
  Server(Server const &) = delete;
  Server &operator=(Server const &) = delete;
  /** Serve the requests (sorted by arrival) to completion.  pos, q, force
   *  are the device arrays of all systems.  With `capture`, each request's
   *  forces are copied (device to device) to capture + 3 * capture_off[r]
   *  when its batch completes. */
  ServeStats run(std::vector<ServeRequest> const &reqs, double *pos, double const *q,
                 double *force, double *capture, int64_t const *capture_off);

private:
  BatchSolver &m_batch;
  ServeConfig m_cfg;
  std::vector<int> m_lanes; // lane ids in the batch solver
  void *m_arena = nullptr;
};

} // namespace espresso_amd
