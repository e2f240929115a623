// tuner.hpp -- schedule autotuner of the sieve launches.
//
// A launch's schedule (segment size, queries sieved together per segment,
// persistent grid) changes its time but never its results.  The engine's
// default schedule comes from a rule measured once on the MI355X; the tuner
// times a few candidates around it on the launch itself the first time a
// launch class is seen (when tuning is on), checks that every candidate gives
// the same sums, and keeps the fastest per device and class in memory and in
// a cache file, so that later launches of the class (in this process or in
// later ones) use it.  No counterpart in the reference, whose
// LoadBalancerS2 (src/LoadBalancerS2.cpp:65-264) sizes segments from rules.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

namespace pcg {

/// How one sieve launch runs.
struct Schedule {
  uint32_t W = 0;            // 64-bit sieve words per segment
  uint32_t gmax = 0;         // queries sieved together per shared segment (0 = rule)
  uint32_t grid_pct = 100;   // persistent grid, percent of the waves occupancy allows
  uint32_t waves = 0;        // 6: single-query launches use the kernel built for 6 waves
                             // per SIMD (at most 80 VGPRs), 0: the unconstrained build
};

/// Launch class: what the best schedule depends on.
struct TuneKey {
  std::string device;        // architecture, CU count and board name, no spaces
  int kind = 0;              // 0 = D, 1 = S2_hard
  int ft32 = 0;              // factor table with 32-bit entries
  int lclass = 0;            // floor(2 log2(sum of the launch's sieve limits))
  int qclass = 0;            // floor(log2(queries with work)), at most 8
  std::string str() const;
};

struct TuneEntry {
  Schedule s;
  double ms_default = 0;     // sieve + combine time of the default schedule
  double ms_best = 0;        // and of the winner, when it was tuned
};

enum TuneMode : int {
  kTuneOff = 0,              // always the default rule
  kTuneCached = 1,           // a cached winner when there is one, else the rule
  kTuneOn = 2,               // as kTuneCached, and tune classes not yet cached
};

/// The cached winner of a class (loads the cache file on first use).
bool tune_lookup(const TuneKey& k, TuneEntry* e);
/// Records a winner in memory and appends it to the cache file (a file that
/// cannot be written leaves the entry in memory only).
void tune_store(const TuneKey& k, const TuneEntry& e);
/// Cache file: `path`, or with nullptr / "" the default ($PCG_TUNE_CACHE,
/// else $XDG_CACHE_HOME/primecount_mi355x/schedules.txt, else
/// $HOME/.cache/primecount_mi355x/schedules.txt).  Drops the entries in
/// memory; the new file is read on the next lookup.
void tune_set_path(const char* path);
std::string tune_path();
/// Entries in memory for any device (after loading the file).
size_t tune_entries();

} // namespace pcg
