// engine.hpp -- host interface of the MI355X hard-special-leaf engine.
#pragma once

#include "pc_math.hpp"
#include "../../include/primecount_mi355x.h"

#include <hip/hip_runtime_api.h>

#include <functional>
#include <mutex>
#include <vector>

namespace pcg {

enum class LeafSum { D, S2_hard };

/// One hard-special-leaf query with explicit parameters.
struct HslQuery {
  LeafSum kind = LeafSum::D;
  i128 x = 0;
  int64_t y = 0, z = 0;
  int64_t k = 0;          // k (D) or c (S2_hard)
};

/// Engine settings (process-wide).
struct EngineConfig {
  uint32_t segment_words = 0;   // 0 = automatic
  int host_threads = 0;         // 0 = all hardware threads
  uint64_t max_slots_per_pass = 0;   // 0 = default (2^31 slots, 16 GiB)
  bool share_segments = true;   // queries of a launch share the sieve of a segment
  bool leaf_batching = true;    // count the leaves of several primes on one prefix build
  bool phase_timers = false;    // single-query launches run the kernel build with phase timers
  uint32_t slot_cbits = 0;      // 0 = automatic; else bits of C in a slot word (at least
                                // what the segment size needs): fewer bits for M, more escapes
  int tune_mode = 1;            // TuneMode (tuner.hpp): 0 rule, 1 cached winners, 2 tune new classes
  // schedule overrides (pcg_set_schedule); any of them set turns tuning off
  uint32_t group_size = 0;      // 0 = schedule's; 2 or 4 queries per shared segment
  uint32_t grid_pct = 0;        // 0 = schedule's; else percent of the occupancy grid
  int kernel_waves = -1;        // -1 = schedule's; 0 or 6 (single-query kernel build)
  uint64_t pix_table_bytes = 0; // largest prime-count table pi(x) keeps whole (0 = 65% of
                                // the device memory); above it the table is built in windows
};

EngineConfig& engine_config();

/// Computes one query on the device; fills `stats`.  Throws pcg_error.
i128 hsl_compute(const HslQuery& q, pcg_stats* stats);

/// Computes one D query with all its device memory in the scratch array of
/// the device (no device or pinned host allocation).  Throws pcg_error
/// (kErrInvalid) when the query does not fit the array.
i128 hsl_compute_noalloc(const HslQuery& q, pcg_stats* stats);

/// Computes many D/S2_hard queries with default or explicit parameters in one
/// persistent launch per table set.  max_slots overrides the pass slot budget
/// (0 = engine setting).  Throws pcg_error.
std::vector<i128> hsl_compute_batch(const std::vector<HslQuery>& qs, pcg_stats* stats,
                                    uint64_t max_slots = 0);

/// hsl_compute_batch for a caller that holds engine_lock(); work already
/// queued on the engine stream runs first.
std::vector<i128> hsl_compute_batch_locked(const std::vector<HslQuery>& qs, pcg_stats* stats,
                                           uint64_t max_slots = 0);

/// Builds the factor table of (kind, y, limit) on the device and on the host
/// (tables.cpp, checked entry by entry against the definitions) and returns
/// the number of entries that differ.
uint64_t factor_table_mismatches(LeafSum kind, int64_t y, int64_t limit);

/// Throws pcg_error if the engine cannot take the query (domain checks only;
/// no device work).
void hsl_validate(const HslQuery& q);

/// Builds (or keeps) resident tables that cover queries of this kind with
/// parameters up to (y, z).
void hsl_reserve(LeafSum kind, int64_t y, int64_t z);

bool device_available();
void release_device();

// ---- shared with the other pi(x) terms (pix.cpp) --------------------------

struct PrimeTables;
struct PixWord;

/// Device-resident data of the engine.  Valid while the caller holds the
/// lock returned by engine_lock().
struct EngineView {
  hipStream_t stream = nullptr;
  int cus = 0;
  const uint32_t* primes = nullptr;     // primes[0] = 0, primes[1] = 2, ...
  uint64_t nprimes = 0;                 // entries of primes[]
  const PrimeTables* host = nullptr;    // host copy of the prime tables
  const void* factor = nullptr;         // FactorTableD entries (D tables)
  bool factor32 = false;
  uint64_t factor_n = 0;                // entries of factor[]
  const PixWord* pix_tab = nullptr;     // prime-count table (see pix_kernels.hpp)
  uint64_t pix_n = 0;                   // it covers [0, pix_n]
};

/// The engine lock (recursive: engine calls made from a pass hook on the
/// thread that holds it take it again).
std::unique_lock<std::recursive_mutex> engine_lock();

/// What a work set's phi table buffer (extra slot 3) holds (pix.cpp).
struct PhiUpload {
  const void* buf = nullptr;
  uint32_t k = 0;
  uint64_t gen = 0;
};

/// A set of work buffers (slot words, work items, partial sums, the scratch
/// of the other pi(x) terms).  Engine calls on a thread use the set it
/// selected, the default set otherwise; all sets share the device's stream
/// and resident tables, so calls with different sets run one after another
/// on the device, and a call made from a pass hook does not disturb the
/// buffers of the batch it interrupts.
struct Work;
Work* engine_work_create();
void engine_work_destroy(Work* w);
/// Selects the calling thread's work set (nullptr: the default); returns the
/// previous one.
Work* engine_work_select(Work* w);

/// While alive, the calling thread's batches call `f` at their pass points:
/// between the kernel passes of a D / S2_hard launch and between the
/// per-query term launches of a pi(x) batch, with the engine lock held, the
/// work queued so far ahead of it on the stream, and the resident tables
/// frozen (an engine call inside f that needs larger tables fails with
/// kErrInternal instead of rebuilding them).  f runs on the same thread; calls
/// it makes see no hook.
class PassHookScope {
public:
  explicit PassHookScope(std::function<void()> f);
  ~PassHookScope();
  PassHookScope(const PassHookScope&) = delete;
  PassHookScope& operator=(const PassHookScope&) = delete;

private:
  std::function<void()> f_;
  const std::function<void()>* prev_;
};

/// Thrown by an engine call inside a pass hook that would need larger
/// resident tables than the running batch holds.
class tables_frozen_error : public pcg_error {
public:
  tables_frozen_error()
      : pcg_error("internal: resident tables are in use by the running batch", kErrInternal) {}
};

/// A pass point: runs the calling thread's pass hook, if any.
void engine_pass_point();

/// The selected work set's phi table record and its 6 term events (pix.cpp).
PhiUpload& engine_phi_upload();
hipEvent_t* engine_term_events();

/// Makes the D tables for (y, z) (y >= z shares one prefix-valid table), or
/// the S2_hard tables (kind), and a prime-count table covering [0, pix_n]
/// resident, with host and device prime tables covering at least [0, pmin];
/// caller holds the lock.
EngineView engine_view(int64_t y, int64_t z, uint64_t pix_n, LeafSum kind = LeafSum::D,
                       uint64_t pmin = 0);

/// Bytes of the largest prime-count table kept whole (pix_table_bytes, or
/// 65% of the device memory).
uint64_t engine_pix_budget();

/// One window of the prime-count table: the words [w_lo, w_lo + nwords)
/// (numbers up to n_max), counts absolute given `base`, the number of primes
/// >= 7 below word w_lo; returns the window (valid until the next call) and
/// sets *next_base to the number of primes >= 7 below its end.  Caller holds
/// the lock; the window replaces the previous one.
const PixWord* engine_pix_window(uint64_t w_lo, uint64_t nwords, uint64_t n_max, uint64_t base,
                                 uint64_t* next_base);

/// A device buffer of at least `bytes`, kept between calls (slot 0 .. 7);
/// caller holds the lock.
void* engine_buffer(int slot, size_t bytes);

/// Changes whenever release_device() frees the device buffers.
uint64_t engine_generation();

} // namespace pcg
