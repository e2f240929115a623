// engine.hpp -- host interface of the MI355X hard-special-leaf engine.
#pragma once

#include "pc_math.hpp"
#include "../../include/primecount_mi355x.h"

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
};

EngineConfig& engine_config();

/// Computes one query on the device; fills `stats`.  Throws pcg_error.
i128 hsl_compute(const HslQuery& q, pcg_stats* stats);

/// Computes many D/S2_hard queries with default or explicit parameters in one
/// persistent launch per table set.  max_slots overrides the pass slot budget
/// (0 = engine setting).  Throws pcg_error.
std::vector<i128> hsl_compute_batch(const std::vector<HslQuery>& qs, pcg_stats* stats,
                                    uint64_t max_slots = 0);

/// Throws pcg_error if the engine cannot take the query (domain checks only;
/// no device work).
void hsl_validate(const HslQuery& q);

/// Builds (or keeps) resident tables that cover queries of this kind with
/// parameters up to (y, z).
void hsl_reserve(LeafSum kind, int64_t y, int64_t z);

bool device_available();
void release_device();

} // namespace pcg
