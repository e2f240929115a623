// serve.hpp -- continuous-batching query server over the engine.
//
// Queries arrive from any thread (submit) and are answered by one worker
// thread on the device.  Light queries (about 10 ms of device time or less)
// run as soon as the device reaches a point where other work may start: at
// once when it is idle, otherwise at the next pass point of the running heavy
// batch (between its kernel passes and between its per-query term launches,
// engine PassHookScope), on a work set of their own.  Heavy queries run in
// batches that share one launch sequence per kind (their D leaves in one
// persistent launch, the sieve of common segments shared), cut into passes of
// bounded size so that light work never waits long.  Work is admitted by its
// estimated HBM footprint (tables + scratch) against the context's budget; a
// query that can never fit is answered at once with PCG_E_NOMEM.  Results are
// collected with poll.
#pragma once

#include "engine.hpp"

#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace pcg {

struct ServeResult {
  uint64_t id = 0;
  int status = 0;           // PCG_OK or a negative PCG_E_* code
  i128 value = 0;
  double seconds = 0;       // submit -> result available
  std::string error;
};

/// One query: kind D / S2_hard with parameters in `q`, or pi(x) / nth_prime
/// with the argument in `arg`.
struct ServeQuery {
  enum Kind { HSL = 0, PI = 1, NTH_PRIME = 2 } kind = HSL;
  HslQuery q;
  i128 arg = 0;
};

struct ServeStats {
  uint64_t submitted = 0, completed = 0, batches = 0, max_batch = 0;
  double busy_seconds = 0;  // worker time inside batches
  uint64_t light_batches = 0, heavy_batches = 0;
  uint64_t interleaved = 0; // light queries answered at pass points of a heavy batch
  uint64_t rejected_nomem = 0;
};

class Server {
public:
  /// hbm_budget: device bytes the context's work may take (tables and
  /// scratch; 0 = 90% of the device memory); max_batch: most queries taken
  /// into one launch sequence (0 = 1024).
  Server(uint64_t hbm_budget, size_t max_batch);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  /// Queues a validated query (status == PCG_OK) or records an immediate
  /// error result for it.  Returns the query id.
  uint64_t submit(const ServeQuery& q, int status, const std::string& error);
  /// Moves up to `max` finished results into `out`, waiting up to
  /// timeout_ms (< 0: forever) for the first one.
  size_t poll(std::vector<ServeResult>& out, size_t max, int timeout_ms);
  ServeStats stats();

private:
  struct Pending {
    uint64_t id;
    ServeQuery q;
    double t_submit;
    double est_ms;          // rough device time (orders and sizes batches)
    uint64_t bytes;         // estimated footprint: tables + scratch
    bool deferred;          // could not run at a pass point (needs larger tables)
  };
  void run();
  /// Runs `batch` as one launch sequence per kind on the selected work set
  /// and publishes every result; queries that fail because the tables are
  /// frozen are handed back in `retry`.
  void run_batch(std::vector<Pending>& batch, bool in_hook, std::vector<Pending>* retry);
  /// Pass hook of a heavy batch: answers the light queries that are waiting.
  void interleave();
  void publish(const std::vector<Pending>& batch, std::vector<ServeResult>& res,
               const std::vector<bool>& keep);

  uint64_t budget_;
  uint64_t pass_slots_;
  size_t max_batch_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Pending> light_, heavy_;
  std::deque<ServeResult> done_;
  uint64_t next_id_ = 1;
  bool stop_ = false;
  ServeStats stats_;
  Work* work_light_ = nullptr;
  Work* work_heavy_ = nullptr;
  std::thread worker_;
};

/// PCG_E_* status of an exception thrown by the engine.
int status_of(const std::exception& e);

} // namespace pcg
