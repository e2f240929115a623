// serve.cpp -- continuous-batching query server over the engine.

#include "serve.hpp"

#include <chrono>
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
  if (dynamic_cast<const pcg_error*>(&e)) {
    const std::string m = e.what();
    if (m.find("HIP error") != std::string::npos || m.find("device") != std::string::npos)
      return PCG_E_DEVICE;
    return PCG_E_INVALID;
  }
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

uint64_t Server::submit(const HslQuery& q, int status, const std::string& error)
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

/// Worker: one launch per round over everything pending (up to max_batch);
/// when a launch fails as a whole, its queries are rerun one by one so that
/// every query gets its own status.  Drains the queue before stopping.
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
    const double t0 = now_s();
    std::vector<HslQuery> qs;
    qs.reserve(batch.size());
    for (const auto& p : batch) qs.push_back(p.q);
    std::vector<ServeResult> res(batch.size());
    try {
      std::vector<i128> v = hsl_compute_batch(qs, nullptr, max_slots_);
      for (size_t i = 0; i < batch.size(); i++) res[i].value = v[i];
    } catch (const std::exception&) {
      for (size_t i = 0; i < batch.size(); i++) {
        try {
          res[i].value = hsl_compute_batch({qs[i]}, nullptr, max_slots_)[0];
        } catch (const std::exception& e) {
          res[i].status = status_of(e);
          res[i].error = e.what();
        }
      }
    }
    const double t1 = now_s();
    {
      std::lock_guard<std::mutex> lock(mu_);
      for (size_t i = 0; i < batch.size(); i++) {
        res[i].id = batch[i].id;
        res[i].seconds = t1 - batch[i].t_submit;
        done_.push_back(std::move(res[i]));
      }
      stats_.completed += batch.size();
      stats_.batches++;
      stats_.max_batch = std::max<uint64_t>(stats_.max_batch, batch.size());
      stats_.busy_seconds += t1 - t0;
    }
    cv_done_.notify_all();
  }
}

} // namespace pcg
