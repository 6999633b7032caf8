// Continuous-batching request loop; see p3m_serve.hpp.
#include "p3m_serve.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <stdexcept>
#include <string>
#include <thread>

namespace espresso_amd {

namespace {
struct InFlight {
  bool busy = false;
  std::vector<int> reqs;
  std::vector<int> sel;
  size_t bytes = 0;  // working-memory region of the batch in the arena
  size_t offset = 0;
  hipEvent_t done = nullptr;
};
constexpr size_t kAlign = 256;
} // namespace

ArenaAllocator::ArenaAllocator(size_t size) {
  if (size > 0) m_free.emplace_back(0, size);
}

size_t ArenaAllocator::allocate(size_t bytes) {
  bytes = (bytes + kAlign - 1) / kAlign * kAlign;
  for (auto it = m_free.begin(); it != m_free.end(); ++it) {
    if (it->second < bytes) continue;
    size_t const off = it->first;
    it->first += bytes;
    it->second -= bytes;
    if (it->second == 0) m_free.erase(it);
    return off;
  }
  return SIZE_MAX;
}

void ArenaAllocator::release(size_t offset, size_t bytes) {
  bytes = (bytes + kAlign - 1) / kAlign * kAlign;
  auto it = std::lower_bound(m_free.begin(), m_free.end(), std::make_pair(offset, size_t{0}));
  it = m_free.insert(it, {offset, bytes});
  // merge with the following and the preceding free range
  if (it + 1 != m_free.end() && it->first + it->second == (it + 1)->first) {
    it->second += (it + 1)->second;
    m_free.erase(it + 1);
  }
  if (it != m_free.begin() && (it - 1)->first + (it - 1)->second == it->first) {
    (it - 1)->second += it->second;
    m_free.erase(it);
  }
}

size_t ArenaAllocator::free_bytes() const {
  size_t f = 0;
  for (auto const &r : m_free) f += r.second;
  return f;
}

Server::Server(BatchSolver &batch, ServeConfig const &cfg) : m_batch(batch), m_cfg(cfg) {
  if (cfg.lanes < 1) throw std::invalid_argument("P3M serve: lanes must be >= 1");
  if (cfg.max_batch < 1) throw std::invalid_argument("P3M serve: max_batch must be >= 1");
  if (cfg.budget_bytes < kAlign) throw std::invalid_argument("P3M serve: budget too small");
  hip_check(hipSetDevice(batch.device()), "hipSetDevice");
  m_cfg.budget_bytes = cfg.budget_bytes / kAlign * kAlign;
  hip_check(hipMalloc(&m_arena, m_cfg.budget_bytes), "hipMalloc serving arena");
  // lane 0 belongs to the synchronous API; serving uses lanes of its own
  while (m_batch.num_lanes() < cfg.lanes + 1) m_batch.add_lane();
  for (int l = 1; l <= cfg.lanes; ++l) m_lanes.push_back(l);
}

Server::~Server() {
  (void)hipSetDevice(m_batch.device());
  for (int l : m_lanes) (void)hipStreamSynchronize(m_batch.lane_stream(l));
  if (m_arena) (void)hipFree(m_arena);
}

ServeStats Server::run(std::vector<ServeRequest> const &reqs, double *pos, double const *q,
                       double *force, double *capture, int64_t const *capture_off) {
  int const nreq = static_cast<int>(reqs.size());
  int const nsys = m_batch.num_systems();
  ServeStats st;
  st.latency_s.assign(nreq, 0.);
  for (int r = 0; r < nreq; ++r) {
    if (reqs[r].sys < 0 || reqs[r].sys >= nsys)
      throw std::invalid_argument("P3M serve: invalid system id " + std::to_string(reqs[r].sys));
    if (r > 0 && reqs[r].arrival_s < reqs[r - 1].arrival_s)
      throw std::invalid_argument("P3M serve: requests must be sorted by arrival time");
    if (m_batch.working_bytes({reqs[r].sys}) > m_cfg.budget_bytes)
      throw std::invalid_argument("P3M serve: a system's working memory exceeds the budget");
  }
  int const nl = static_cast<int>(m_lanes.size());
  std::vector<InFlight> fl(nl);
  for (auto &f : fl) hip_check(hipEventCreateWithFlags(&f.done, hipEventDisableTiming), "event");
  {
    // the serving lanes start after the work enqueued on the batch's own
    // stream (the caller fences that stream with its producers of pos/q/force)
    hipEvent_t ready = nullptr;
    hip_check(hipEventCreateWithFlags(&ready, hipEventDisableTiming), "event");
    hipError_t e = hipEventRecord(ready, m_batch.stream());
    for (int l : m_lanes)
      if (e == hipSuccess) e = hipStreamWaitEvent(m_batch.lane_stream(l), ready, 0);
    (void)hipEventDestroy(ready);
    if (e != hipSuccess) {
      for (auto &f : fl) (void)hipEventDestroy(f.done);
      hip_check(e, "P3M serve: fence on the batch stream");
    }
  }
  std::vector<char> busy_sys(nsys, 0), in_batch(nsys, 0), taken(nreq, 0);
  std::deque<int> pending;
  ArenaAllocator arena(m_cfg.budget_bytes);
  size_t bytes_in_flight = 0;
  int next = 0, completed = 0, active = 0;
  ParticleArrays const pa{pos, q, nullptr, force};
  using clock = std::chrono::steady_clock;
  auto const t0 = clock::now();
  auto elapsed = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
  double last_done = 0.;
  try {
    while (completed < nreq) {
      bool progress = false;
      double const now = elapsed();
      while (next < nreq && reqs[next].arrival_s <= now) pending.push_back(next++);
      // retire finished batches
      for (int l = 0; l < nl; ++l) {
        InFlight &f = fl[l];
        if (!f.busy) continue;
        hipError_t const e = hipEventQuery(f.done);
        if (e == hipErrorNotReady) continue;
        hip_check(e, "P3M serve batch");
        double const t = elapsed();
        for (int r : f.reqs) st.latency_s[r] = t - reqs[r].arrival_s;
        for (int s : f.sel) busy_sys[s] = 0;
        arena.release(f.offset, f.bytes);
        bytes_in_flight -= f.bytes;
        completed += static_cast<int>(f.reqs.size());
        last_done = t;
        f.busy = false;
        f.reqs.clear();
        f.sel.clear();
        --active;
        progress = true;
      }
      // admit pending requests into free lanes
      for (int l = 0; l < nl && !pending.empty(); ++l) {
        InFlight &f = fl[l];
        if (f.busy) continue;
        // candidates in arrival order while the working set fits the free
        // part of the arena; then one contiguous region is carved for it
        size_t const free_total = arena.free_bytes();
        std::vector<unsigned long long> seeds;
        for (int const r : pending) {
          if (static_cast<int>(f.sel.size()) >= m_cfg.max_batch) break;
          int const s = reqs[r].sys;
          if (busy_sys[s] || in_batch[s]) continue; // one request per system at a time, in order
          f.sel.push_back(s);
          if (m_batch.working_bytes(f.sel) > free_total) {
            f.sel.pop_back();
            continue;
          }
          in_batch[s] = 1;
          f.reqs.push_back(r);
          seeds.push_back(reqs[r].seed);
        }
        for (int s : f.sel) in_batch[s] = 0;
        size_t offset = SIZE_MAX;
        size_t bytes = 0;
        while (!f.sel.empty()) {
          bytes = m_batch.working_bytes(f.sel);
          offset = arena.allocate(bytes);
          if (offset != SIZE_MAX) break;
          // fragmented: shrink the batch until a region fits
          f.sel.pop_back();
          f.reqs.pop_back();
          seeds.pop_back();
        }
        if (f.sel.empty()) continue;
        for (int r : f.reqs) taken[r] = 1;
        pending.erase(std::remove_if(pending.begin(), pending.end(),
                                     [&](int r) { return taken[r] != 0; }),
                      pending.end());
        int const lane = m_lanes[l];
        m_batch.prepare_on(lane, f.sel, pos, force, m_cfg.sigma, seeds);
        m_batch.solve_on(lane, f.sel, pa, 1u /* forces */, false, nullptr,
                         static_cast<char *>(m_arena) + offset, bytes);
        hipStream_t const stream = m_batch.lane_stream(lane);
        if (capture) {
          for (int r : f.reqs) {
            int const s = reqs[r].sys;
            hip_check(hipMemcpyAsync(capture + 3 * capture_off[r],
                                     force + 3 * m_batch.particle_offset(s),
                                     sizeof(double) * 3 * m_batch.particle_count(s),
                                     hipMemcpyDeviceToDevice, stream),
                      "capture forces");
          }
        }
        hip_check(hipEventRecord(f.done, stream), "hipEventRecord");
        for (int s : f.sel) busy_sys[s] = 1;
        f.bytes = bytes;
        f.offset = offset;
        f.busy = true;
        bytes_in_flight += bytes;
        st.peak_bytes = std::max(st.peak_bytes, bytes_in_flight);
        ++active;
        st.peak_lanes = std::max(st.peak_lanes, active);
        ++st.batches;
        progress = true;
      }
      if (!progress) std::this_thread::yield();
    }
  } catch (...) {
    for (auto &f : fl) {
      if (f.busy) (void)hipEventSynchronize(f.done);
      (void)hipEventDestroy(f.done);
    }
    throw;
  }
  for (auto &f : fl) (void)hipEventDestroy(f.done);
  st.wall_s = last_done - (nreq > 0 ? reqs[0].arrival_s : 0.);
  return st;
}

} // namespace espresso_amd
