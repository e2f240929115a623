// serve.hpp -- continuous-batching query server over the engine.
//
// Queries arrive from any thread (submit); one worker thread takes everything
// that is pending when the device becomes free and runs it as one batch, so
// arrivals during a launch join the next one: D / S2_hard queries in one
// persistent launch per table set (hsl_compute_batch), pi(x) and nth_prime
// queries with their D parts in one launch (pi_gourdon_batch,
// nth_prime_batch).  Results are collected with poll.
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
};

class Server {
public:
  /// max_slots: slot budget of one kernel pass (0 = engine default);
  /// max_batch: most queries taken into one launch.
  Server(uint64_t max_slots, size_t max_batch);
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
  };
  void run();

  uint64_t max_slots_;
  size_t max_batch_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Pending> pending_;
  std::deque<ServeResult> done_;
  uint64_t next_id_ = 1;
  bool stop_ = false;
  ServeStats stats_;
  std::thread worker_;
};

/// PCG_E_* status of an exception thrown by the engine.
int status_of(const std::exception& e);

} // namespace pcg
