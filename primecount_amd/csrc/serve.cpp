// serve.cpp -- continuous-batching query server over the engine.

#include "serve.hpp"

#include "pix.hpp"
#include "tables.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <new>

namespace pcg {

namespace {

double now_s()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/// Light queries (about 10 ms or less on the device) run as soon as the
/// device reaches a pass point, heavy ones in batches.
bool light(const ServeQuery& q)
{
  if (q.kind == ServeQuery::PI) return q.arg <= (i128) 10000000000000000LL;          // 1e16
  if (q.kind == ServeQuery::NTH_PRIME) return q.arg <= (i128) 200000000000000LL;     // 2e14
  return q.q.x <= (i128) 10000000000000000LL;
}

/// Rough device time of a query in ms (about x^(2/3), from pi(1e19) in
/// 0.18 s and D(1e19) in 0.12 s on the MI355X); it only sizes the batches.
double estimate_ms(const ServeQuery& q)
{
  double x, at19;
  if (q.kind == ServeQuery::NTH_PRIME) {
    const double n = (double) q.arg;
    x = n * std::log(std::max(n, 3.0));
    at19 = 180;
  } else if (q.kind == ServeQuery::PI) {
    x = (double) q.arg;
    at19 = 180;
  } else {
    x = (double) q.q.x;
    at19 = 120;
  }
  return at19 * std::pow(std::max(x, 1.0) / 1e19, 2.0 / 3.0);
}

/// Scratch bytes a query adds to a batch beyond the shared tables.
constexpr uint64_t kQueryScratch = 64ull << 20;

/// Estimated device bytes: the tables the query needs (shared by a batch,
/// prefix valid) plus its scratch.
uint64_t footprint(const ServeQuery& q)
{
  switch (q.kind) {
    case ServeQuery::PI: return pi_footprint(q.arg);
    case ServeQuery::NTH_PRIME: return nth_prime_footprint((uint64_t) q.arg);
    default: break;
  }
  const int64_t z = std::max(q.q.y, q.q.z);
  return table_bytes_D(q.q.y, z) + kQueryScratch;
}

/// Slots of one sieve pass of a heavy batch: 2^31 (8 GiB of slot words), the
/// engine's default.  Smaller passes would give light queries more pass
/// points but cost the heavy work dearly: each pass starts on its heaviest
/// (lowest) segments and drains the persistent grid unevenly (D(1e19) on the
/// MI355X: 120.6 ms in one pass, 134.1 / 153.4 / 172.7 ms in passes of
/// 2^30 / 2^29 / 2^28 slots).
constexpr uint64_t kPassSlots = 1ull << 31;

/// Heavy work taken per batch: queries in arrival order until their
/// estimates pass this (at least one).  With more than kBacklog heavy queries
/// waiting the batch takes half of the estimated backlog (up to 4 s), for
/// throughput: the queries of a launch share the sieve of their common
/// segments.
constexpr double kHeavyRoundMs = 250.0;
constexpr size_t kBacklog = 8;

uint64_t device_memory()
{
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void) hipGetLastError();
    return 256ull << 30;
  }
  return (uint64_t) total_b;
}

} // namespace

int status_of(const std::exception& e)
{
  if (dynamic_cast<const std::bad_alloc*>(&e)) return PCG_E_NOMEM;
  if (const pcg_error* pe = dynamic_cast<const pcg_error*>(&e)) return pe->status();
  return PCG_E_INTERNAL;
}

Server::Server(uint64_t hbm_budget, size_t max_batch)
    : budget_(hbm_budget ? hbm_budget : device_memory() / 10 * 9),
      pass_slots_(std::max<uint64_t>(64, std::min<uint64_t>(kPassSlots, budget_ / 16))),
      max_batch_(max_batch ? max_batch : 1024)
{
  work_light_ = engine_work_create();
  work_heavy_ = engine_work_create();
  worker_ = std::thread([this] { run(); });
}

Server::~Server()
{
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  worker_.join();
  engine_work_destroy(work_light_);
  engine_work_destroy(work_heavy_);
}

uint64_t Server::submit(const ServeQuery& q, int status, const std::string& error)
{
  const double t = now_s();
  Pending p{0, q, t, 0.0, 0, false};
  std::string err = error;
  if (status == PCG_OK) {
    try {
      p.est_ms = estimate_ms(q);
      p.bytes = footprint(q);
      if (p.bytes + pass_slots_ * 4 > budget_) {
        status = PCG_E_NOMEM;
        err = "the query needs about " + std::to_string((p.bytes + pass_slots_ * 4) >> 20) +
              " MiB of device memory; the serving context's budget is " +
              std::to_string(budget_ >> 20) + " MiB";
      }
    } catch (const std::exception& e) {
      status = status_of(e);
      err = e.what();
    }
  }
  uint64_t id;
  {
    std::lock_guard<std::mutex> lock(mu_);
    id = next_id_++;
    p.id = id;
    stats_.submitted++;
    if (status != PCG_OK) {
      ServeResult r;
      r.id = id;
      r.status = status;
      r.error = err;
      done_.push_back(r);
      stats_.completed++;
      if (status == PCG_E_NOMEM) stats_.rejected_nomem++;
    } else if (light(q)) {
      light_.push_back(p);
    } else {
      heavy_.push_back(p);
    }
  }
  if (status != PCG_OK) cv_done_.notify_all();
  else cv_work_.notify_one();
  return id;
}

size_t Server::poll(std::vector<ServeResult>& out, size_t max, int timeout_ms)
{
  std::unique_lock<std::mutex> lock(mu_);
  auto ready = [&] { return !done_.empty(); };
  if (timeout_ms < 0) cv_done_.wait(lock, ready);
  else cv_done_.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready);
  size_t n = 0;
  while (n < max && !done_.empty()) {
    out.push_back(std::move(done_.front()));
    done_.pop_front();
    n++;
  }
  return n;
}

ServeStats Server::stats()
{
  std::lock_guard<std::mutex> lock(mu_);
  return stats_;
}

void Server::publish(const std::vector<Pending>& batch, std::vector<ServeResult>& res,
                     const std::vector<bool>& keep)
{
  const double t1 = now_s();
  size_t n = 0;
  {
    std::lock_guard<std::mutex> lock(mu_);
    for (size_t i = 0; i < batch.size(); i++) {
      if (!keep[i]) continue;
      res[i].id = batch[i].id;
      res[i].seconds = t1 - batch[i].t_submit;
      done_.push_back(std::move(res[i]));
      n++;
    }
    stats_.completed += n;
  }
  if (n) cv_done_.notify_all();
}

/// One launch sequence per kind of query (D / S2_hard in one persistent
/// launch per table set, pi(x) with the D leaves of all of them in one
/// launch, nth_prime through its pi(x0) batch).  When a launch fails as a
/// whole its queries are rerun one by one, so that every query gets its own
/// status; at a pass point, queries that need larger tables than the running
/// batch holds are handed back in `retry`.
void Server::run_batch(std::vector<Pending>& batch, bool in_hook, std::vector<Pending>* retry)
{
  const double t0 = now_s();
  std::vector<ServeResult> res(batch.size());
  std::vector<bool> keep(batch.size(), true);
  const uint64_t slots = pass_slots_;
  auto run_kind = [&](ServeQuery::Kind kind) {
    std::vector<size_t> idx;
    for (size_t i = 0; i < batch.size(); i++)
      if (batch[i].q.kind == kind) idx.push_back(i);
    if (idx.empty()) return;
    auto compute = [&](const std::vector<size_t>& part) {
      std::vector<i128> v;
      if (kind == ServeQuery::HSL) {
        std::vector<HslQuery> qs;
        for (size_t i : part) qs.push_back(batch[i].q.q);
        v = hsl_compute_batch(qs, nullptr, slots);
      } else if (kind == ServeQuery::PI) {
        std::vector<i128> xs;
        for (size_t i : part) xs.push_back(batch[i].q.arg);
        v = pi_gourdon_batch(xs, nullptr, slots);
      } else {
        std::vector<uint64_t> ns;
        for (size_t i : part) ns.push_back((uint64_t) batch[i].q.arg);
        for (uint64_t p : nth_prime_batch(ns, nullptr, slots)) v.push_back((i128) p);
      }
      for (size_t j = 0; j < part.size(); j++) res[part[j]].value = v[j];
    };
    try {
      compute(idx);
    } catch (const std::exception&) {
      for (size_t i : idx) {
        try {
          compute({i});
        } catch (const tables_frozen_error& e) {
          if (in_hook) {
            keep[i] = false;                   // runs after the heavy batch
          } else {
            res[i].status = status_of(e);
            res[i].error = e.what();
          }
        } catch (const std::exception& e) {
          res[i].status = status_of(e);
          res[i].error = e.what();
        }
      }
    }
  };
  run_kind(ServeQuery::HSL);
  run_kind(ServeQuery::PI);
  run_kind(ServeQuery::NTH_PRIME);
  publish(batch, res, keep);
  if (retry)
    for (size_t i = 0; i < batch.size(); i++)
      if (!keep[i]) retry->push_back(batch[i]);
  const double t1 = now_s();
  std::lock_guard<std::mutex> lock(mu_);
  stats_.batches++;
  stats_.max_batch = std::max<uint64_t>(stats_.max_batch, batch.size());
  if (!in_hook) stats_.busy_seconds += t1 - t0;
}

/// Pass hook of a heavy batch (engine lock held, tables frozen, the heavy
/// batch's work queued ahead on the stream): the waiting light queries run
/// now on the light work set; those that need larger tables wait for the
/// end of the heavy batch.  Nothing may escape into the heavy batch.
void Server::interleave()
{
  std::vector<Pending> batch;
  {
    std::lock_guard<std::mutex> lock(mu_);
    std::deque<Pending> keep;
    while (!light_.empty() && batch.size() < max_batch_) {
      Pending p = light_.front();
      light_.pop_front();
      if (p.deferred) keep.push_back(p);
      else batch.push_back(p);
    }
    for (auto it = keep.rbegin(); it != keep.rend(); ++it) light_.push_front(*it);
  }
  if (batch.empty()) return;
  Work* prev = engine_work_select(work_light_);
  std::vector<Pending> retry;
  try {
    run_batch(batch, true, &retry);
  } catch (...) {
    // run_batch records every query's own error; anything else is a bug
  }
  engine_work_select(prev);
  std::lock_guard<std::mutex> lock(mu_);
  for (auto it = retry.rbegin(); it != retry.rend(); ++it) {
    Pending p = *it;
    p.deferred = true;
    light_.push_front(p);
  }
  stats_.interleaved += batch.size() - retry.size();
  stats_.light_batches++;
}

/// Worker: light queries first, then a heavy batch whose pass points answer
/// the light queries that arrive meanwhile.  Drains the queues before
/// stopping.
void Server::run()
{
  for (;;) {
    std::vector<Pending> batch;
    bool heavy = false;
    {
      std::unique_lock<std::mutex> lock(mu_);
      cv_work_.wait(lock, [&] { return stop_ || !light_.empty() || !heavy_.empty(); });
      if (light_.empty() && heavy_.empty()) return;            // stop requested and drained
      for (Pending& p : light_) p.deferred = false;
      if (!light_.empty()) {
        while (!light_.empty() && batch.size() < max_batch_) {
          batch.push_back(light_.front());
          light_.pop_front();
        }
        stats_.light_batches++;
      } else {
        heavy = true;
        double backlog_ms = 0;
        for (const Pending& p : heavy_) backlog_ms += p.est_ms;
        const double round_ms = heavy_.size() > kBacklog ? std::min(4000.0, backlog_ms / 2)
                                                         : kHeavyRoundMs;
        // admission by footprint: the largest tables of the batch plus the
        // scratch of every member and one pass of slot words
        uint64_t tables = 0, scratch = pass_slots_ * 4;
        double ms = 0;
        while (!heavy_.empty() && batch.size() < max_batch_) {
          const Pending& p = heavy_.front();
          const uint64_t t = p.bytes > kQueryScratch ? p.bytes - kQueryScratch : 0;
          const uint64_t need = std::max(tables, t) + scratch + kQueryScratch;
          if (!batch.empty() && (ms >= round_ms || need > budget_)) break;
          tables = std::max(tables, t);
          scratch += kQueryScratch;
          ms += p.est_ms;
          batch.push_back(p);
          heavy_.pop_front();
        }
        stats_.heavy_batches++;
      }
    }
    if (!heavy) {
      Work* prev = engine_work_select(work_light_);
      run_batch(batch, false, nullptr);
      engine_work_select(prev);
      continue;
    }
    Work* prev = engine_work_select(work_heavy_);
    {
      PassHookScope hook([this] { interleave(); });
      run_batch(batch, false, nullptr);
    }
    engine_work_select(prev);
  }
}

} // namespace pcg
