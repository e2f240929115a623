// serve.cpp -- continuous-batching query server over the engine.

#include "serve.hpp"

#include "pix.hpp"

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

} // namespace

int status_of(const std::exception& e)
{
  if (dynamic_cast<const std::bad_alloc*>(&e)) return PCG_E_NOMEM;
  if (const pcg_error* pe = dynamic_cast<const pcg_error*>(&e)) return pe->status();
  return PCG_E_INTERNAL;
}

Server::Server(uint64_t max_slots, size_t max_batch)
    : max_slots_(max_slots), max_batch_(max_batch ? max_batch : 1024)
{
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
}

uint64_t Server::submit(const ServeQuery& q, int status, const std::string& error)
{
  const double t = now_s();
  uint64_t id;
  {
    std::lock_guard<std::mutex> lock(mu_);
    id = next_id_++;
    stats_.submitted++;
    if (status != PCG_OK) {
      ServeResult r;
      r.id = id;
      r.status = status;
      r.error = error;
      done_.push_back(r);
      stats_.completed++;
    } else {
      pending_.push_back(Pending{id, q, t});
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

namespace {

/// Light queries (about 10 ms or less on the device) are answered and
/// published before the heavy ones of the same round, so that they do not
/// wait for a launch of large x.
bool light(const ServeQuery& q)
{
  if (q.kind == ServeQuery::PI) return q.arg <= (i128) 10000000000000000LL;          // 1e16
  if (q.kind == ServeQuery::NTH_PRIME) return q.arg <= (i128) 200000000000000LL;     // 2e14
  return q.q.x <= (i128) 10000000000000000LL;
}

/// Rough device time of a query in ms (about x^(2/3), from pi(1e19) in
/// 0.23 s and D(1e19) in 0.17 s on the MI355X); it only orders the work.
double estimate_ms(const ServeQuery& q)
{
  double x, at19;
  if (q.kind == ServeQuery::NTH_PRIME) {
    const double n = (double) q.arg;
    x = n * std::log(std::max(n, 3.0));
    at19 = 230;
  } else if (q.kind == ServeQuery::PI) {
    x = (double) q.arg;
    at19 = 230;
  } else {
    x = (double) q.q.x;
    at19 = 170;
  }
  return at19 * std::pow(std::max(x, 1.0) / 1e19, 2.0 / 3.0);
}

/// Heavy work taken per round: queries in arrival order until their
/// estimates pass this (at least one), so that light queries arriving
/// meanwhile wait for at most about this long.  With a backlog of more than
/// kBacklog heavy queries the round takes ten times as much, for throughput
/// (the queries of a launch share the sieve of their common segments).
constexpr double kHeavyRoundMs = 300.0;
constexpr size_t kBacklog = 16;

} // namespace

/// Worker: per round, what is pending (up to max_batch) in two launch sets,
/// each published as soon as it is done: all light queries first, then heavy
/// ones in arrival order up to the round's budget of estimated work (the
/// others go back to the front of the queue); a set runs one launch
/// sequence per kind.  When a launch fails as a whole,
/// its queries are rerun one by one so that every query gets its own status.
/// Drains the queue before stopping.
void Server::run()
{
  for (;;) {
    std::vector<Pending> batch;
    {
      std::unique_lock<std::mutex> lock(mu_);
      cv_work_.wait(lock, [&] { return stop_ || !pending_.empty(); });
      if (pending_.empty()) return;            // stop requested and drained
      while (!pending_.empty() && batch.size() < max_batch_) {
        batch.push_back(pending_.front());
        pending_.pop_front();
      }
    }
    std::vector<size_t> sets[2];
    std::vector<Pending> later;
    size_t nheavy = 0;
    for (const Pending& p : batch) nheavy += !light(p.q);
    const double budget = nheavy > kBacklog ? 10 * kHeavyRoundMs : kHeavyRoundMs;
    double heavy_ms = 0;
    for (size_t i = 0; i < batch.size(); i++) {
      if (light(batch[i].q)) {
        sets[0].push_back(i);
      } else if (sets[1].empty() || heavy_ms < budget) {
        heavy_ms += estimate_ms(batch[i].q);
        sets[1].push_back(i);
      } else {
        later.push_back(batch[i]);
      }
    }
    if (!later.empty()) {
      // back to the front of the queue, in arrival order, for the next round
      std::lock_guard<std::mutex> lock(mu_);
      for (auto it = later.rbegin(); it != later.rend(); ++it) pending_.push_front(*it);
    }
    for (const std::vector<size_t>& set : sets) {
      if (set.empty()) continue;
      const double t0 = now_s();
      std::vector<ServeResult> res(batch.size());
      // one launch sequence per kind of query
      auto run_kind = [&](ServeQuery::Kind kind) {
        std::vector<size_t> idx;
        for (size_t i : set)
          if (batch[i].q.kind == kind) idx.push_back(i);
        if (idx.empty()) return;
        auto compute = [&](const std::vector<size_t>& part) {
          std::vector<i128> v;
          if (kind == ServeQuery::HSL) {
            std::vector<HslQuery> qs;
            for (size_t i : part) qs.push_back(batch[i].q.q);
            v = hsl_compute_batch(qs, nullptr, max_slots_);
          } else if (kind == ServeQuery::PI) {
            std::vector<i128> xs;
            for (size_t i : part) xs.push_back(batch[i].q.arg);
            v = pi_gourdon_batch(xs, nullptr, max_slots_);
          } else {
            std::vector<uint64_t> ns;
            for (size_t i : part) ns.push_back((uint64_t) batch[i].q.arg);
            for (uint64_t p : nth_prime_batch(ns, nullptr, max_slots_)) v.push_back((i128) p);
          }
          for (size_t j = 0; j < part.size(); j++) res[part[j]].value = v[j];
        };
        try {
          compute(idx);
        } catch (const std::exception&) {
          for (size_t i : idx) {
            try {
              compute({i});
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
      const double t1 = now_s();
      {
        std::lock_guard<std::mutex> lock(mu_);
        for (size_t i : set) {
          res[i].id = batch[i].id;
          res[i].seconds = t1 - batch[i].t_submit;
          done_.push_back(std::move(res[i]));
        }
        stats_.completed += set.size();
        stats_.batches++;
        stats_.max_batch = std::max<uint64_t>(stats_.max_batch, set.size());
        stats_.busy_seconds += t1 - t0;
      }
      cv_done_.notify_all();
    }
  }
}

} // namespace pcg
