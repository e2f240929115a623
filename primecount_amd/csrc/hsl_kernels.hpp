// hsl_kernels.hpp -- data layout shared by the host planner and the HIP
// kernels of the hard-special-leaf engine.
//
// Terms (primecount's vocabulary):
//   segment   240*W consecutive numbers [low, high) of the sieve interval,
//             held in LDS as W 64-bit words, wheel-30 layout (bit 8j+i of a
//             word <-> 30j + {1,7,11,13,17,19,23,29}[i]), as in
//             src/sieve/Sieve.cpp:62-66
//   b         index of the sieving prime p_b (1-indexed: p_1 = 2)
//   leaf      x / (p_b * m) with a square-free m (type 1, factor-table loop)
//             or a prime m (type 2, prime loop); it contributes
//             sign * phi(x / (p_b * m), b - 1)
//   slot      one (segment, b) pair: C = unsieved numbers of the segment
//             before p_b is crossed off, M = sum of the leaf signs, packed
//             into one 32-bit word (SlotFormat)
#pragma once

#include <stdint.h>
#include <hip/hip_runtime_api.h>

namespace pcg {

/// Per sieving prime constants of one query (host computed, read uniformly).
/// For segment [low, high) the leaf range of prime p_b is
///   m_hi = min(A, xp / max(low, 1)),  m_lo = max(min(xp / high, C), B)
/// and the segment stops at the first b whose range is provably empty
/// (type 1: p >= m_hi; type 2: no prime in (p, m_hi]), exactly as
/// src/gourdon/D.cpp:125-132, 281-290 and
/// src/deleglise-rivat/S2_hard.cpp:100-107, 135-143 do.
struct alignas(16) BInfo {
  uint64_t xp_lo;   // low 64 bits of xp = x / p_b
  uint64_t A;       // static cap of m_hi
  uint64_t B;       // static floor of m_lo
  uint64_t C;       // cap of xp / high
  double xp_d;      // xp rounded to double
  uint32_t p;       // p_b
  uint32_t T;       // type 1: leaf iff factor[m] > T
  uint32_t type;    // 1 = factor-table leaves (sign -mu(m)), 2 = prime leaves (sign +1)
  uint32_t pad0;
  uint64_t pad1;
};

/// One query of a launch.
///
/// reach[reach_off + b - b_first] = number of segments that reach b (segment
/// s reaches b iff max(1, s * 240 W) <= x / p_b^2); it does not increase with
/// b, so segment s holds the slots b - b_first < nbs(s) = #{b : reach > s},
/// and the slots of segments [0, s) number
///   F(s) = sum_b min(s, reach_b) = s * nbs(s) + sfx[nbs(s)]
/// with the suffix sums sfx[sfx_off + k] = sum_{b - b_first >= k} reach_b.
struct QueryDesc {
  uint64_t limit;        // sieve interval [0, limit)
  uint32_t b_first;      // first b (k+1 or c+1)
  uint32_t b_last;       // last b that can hold leaves
  uint32_t binfo_off;    // BInfo of b is binfo[binfo_off + b - b_first]
  uint32_t pre_last;     // primes p_4..p_pre_last are pre-sieved (pre_last = b_first-1)
  uint32_t reach_off;    // per b: segments reaching b (b_last - b_first + 1 values)
  uint32_t sfx_off;      // suffix sums of reach (b_last - b_first + 2 values)
};

/// Segments [s0, s1) of query `query` within one pass: items
/// [item_base, item_base + s1 - s0); the slots of segment s start at
/// t_base + F(s) (t_base = pass offset of s0 minus F(s0), modulo 2^64).
struct alignas(16) PassSpan {
  uint32_t query;
  uint32_t s0;
  uint32_t s1;
  uint32_t item_base;
  uint64_t t_base;
  uint64_t pad;
};

/// One work-queue item: segment `seg` of query `query`.  Its slots
/// (seg, b) live at t_off + (b - b_first) for b - b_first < nbs; nbs bounds
/// the b the segment can reach (p_b <= sqrt(x / low)).
struct alignas(16) SegItem {
  uint32_t query;
  uint32_t seg;
  uint32_t nbs;
  uint32_t pad;
  uint64_t t_off;
  uint64_t pad1;
};

/// Queries that share one segment in a launch (kGroupMax of them at most;
/// a launch uses groups of 2 or of 4): the work item is item[0 .. n); all of
/// them cover the full segment and have the same first b, so the sieve of the
/// segment is the same for all.
constexpr int kGroupMax = 4;
struct alignas(16) SegGroup {
  uint32_t n;
  uint32_t item[kGroupMax];
  uint32_t pad[3];
};

/// One 32-bit word per slot: C in the low `cbits` bits (C <= 64 W, so
/// cbits = 16 for W < 1024, up to 19 for W = 4096) and M as a signed
/// (32 - cbits)-bit field above it.  |M| >= 2^(31 - cbits) does not fit: the
/// field then holds its most negative value (the escape marker, bit 31 alone)
/// and M goes to the escape table.
struct SlotFormat {
  uint32_t cbits;
};

inline SlotFormat slot_format(uint32_t W)
{
  uint32_t cb = 1;
  while ((1ull << cb) <= 64ull * W) cb++;     // C <= 64 W < 2^cb
  return SlotFormat{cb};
}

/// Escape table of one pass: open addressing over `cap` entries (a power of
/// two), key = slot offset + 1 (0 = empty), value = M.  The host sizes cap
/// to at least twice an upper bound of the escaping slots (a slot's |M| is at
/// most the number of m in its leaf range), so inserts always find room;
/// count > cap / 2 would set the overflow flag instead of looping.
struct EscTable {
  unsigned long long* keys;
  int32_t* vals;
  unsigned int* ctl;            // [0] entries inserted, [1] overflow flag
  uint32_t cap;
};

struct HslArgs {
  const QueryDesc* queries;
  const SegItem* items;
  const SegGroup* groups;       // null: one work item per SegItem
  const BInfo* binfo;
  const uint32_t* primes;       // primes[0] = 0, primes[1] = 2, ...
  const void* factor;           // uint16_t or uint32_t factor table
  const uint64_t* pi_bits;      // prime bits per 240 numbers (primes >= 7)
  const uint32_t* pi_cnt;       // primes >= 7 below 240 * i
  uint32_t* T_S;                // per slot: C | M << cbits (SlotFormat)
  EscTable esc;                 // M of the slots whose |M| does not fit
  uint32_t cbits;
  unsigned long long* S_acc;    // per query: 128-bit sum of sign * in-segment count (lo, hi)
  unsigned long long* counters; // leaves, hits, factor entries, slots
  unsigned int* queue;          // next work item
  uint32_t nitems;              // work items: SegItems, or SegGroups when groups is set
  uint32_t group_max;           // 2 or 4: the largest group (when groups is set)
  uint32_t W;                   // words per segment (multiple of 128, 128..4096)
  uint32_t waves;               // 6: the single-query kernel built for 6 waves per SIMD, else 0
  uint32_t bcap;                // batched leaf counting: most cross-off hits recorded per batch
                                // (<= kBatchRecords), 0 = off; single-query kernel, W <= 1024
  uint32_t timers;              // 1: the single-query kernel with phase timers
};

/// Phases of the sieve kernel's phase timers (the timed build): segment set-up
/// and end, window set-up and end, one-at-a-time cross-offs, grouped
/// cross-offs, a batch's recording cross-offs, pair-prefix builds, the
/// pre-scan of primes with factor entries, factor-entry (or prime) steps,
/// counting steps, and the entry and counting steps of leaf batches.
enum : uint32_t {
  kHslPhaseSegment, kHslPhaseWindow, kHslPhaseDense, kHslPhaseGroup, kHslPhaseBatchFlush,
  kHslPhasePrefix, kHslPhaseScan, kHslPhaseEntry, kHslPhaseDrain, kHslPhaseBatchEntry,
  kHslPhaseBatchDrain, kHslPhases
};
/// counters[kHslPhaseCounter0 + phase]: the phase's cycles summed over waves
constexpr uint32_t kHslPhaseCounter0 = 16;
/// 64-bit kernel counters: 14 work counts, 2 spare, then the phase cycles
constexpr uint32_t kHslCounters = kHslPhaseCounter0 + kHslPhases;

/// Batched leaf counting (single-query kernel): a batch holds at most
/// kBatchEntries factor entries (or prime leaves) and records at most
/// kBatchRecords first-time removals for the corrections.
constexpr uint32_t kBatchEntries = 128;
constexpr uint32_t kBatchRecords = 128;

constexpr int kBlock = 64;     // one wave per segment

/// Column of the carry combine: one (query, b) within one pass.
/// Its items are items[item_first .. item_first + nitems) (consecutive
/// segments of the query, in order); `col` indexes the pass-independent
/// carry and output arrays.
struct CombineCol {
  uint32_t b_rel;         // b - b_first
  uint32_t item_first;
  uint32_t nitems;
  uint32_t col;
};

/// LDS bytes the sieve kernel needs for W words per segment.
inline uint32_t hsl_lds_bytes(uint32_t W)
{
  return 8 * W            // segment bitmap
       + 2 * W            // word-pair prefix counts (32-bit, W / 2 pairs)
       + 240;             // sieve bits of the offsets <= r in a word, r < 240
}

/// Bytes of the device scratch array of the no-allocation entry point: it
/// holds every table and work buffer of one D(x) call (D(1e15) with the
/// default parameters needs about 30 MB).
constexpr size_t kHslScratchBytes = (size_t) 64 << 20;
constexpr int kHslMaxDevices = 64;

// Host-side launch wrappers (defined in hsl_kernels.hip).
hipError_t hsl_upload_constants();
/// Device address of the scratch array on `device` (the current device),
/// looked up with hipGetSymbolAddress on the first call for that device.
hipError_t hsl_scratch_address(int device, void** ptr);
hipError_t hsl_launch_sieve(const HslArgs& a, bool factor32, uint32_t grid, hipStream_t st);
/// Sieve workgroups (one wave each) a CU holds at once for W words per
/// segment, for the single-query kernel (waves = 6: its 6-waves-per-SIMD
/// build).
hipError_t hsl_sieve_occupancy(bool factor32, uint32_t W, uint32_t waves, int* blocks_per_cu);
hipError_t hsl_launch_fill_zero(void* p, size_t bytes, hipStream_t st);

/// Factor table entries for the indices 0 .. ft_to_index(limit) (numbers
/// coprime to 2310 up to limit < 2^32): the FactorTableD encoding for D
/// (d_table, leaves need mpf <= y) or the lpf encoding for S2_hard; primes
/// [0, nprimes) must cover sqrt(limit).
hipError_t hsl_launch_factor_table(void* out, bool factor32, bool d_table, uint64_t limit,
                                   uint64_t y, const uint32_t* primes, uint32_t nprimes,
                                   hipStream_t st);
hipError_t hsl_launch_build_items(const QueryDesc* queries, const PassSpan* spans, uint32_t nspans,
                                  const uint32_t* reach, const uint64_t* sfx, SegItem* items,
                                  uint32_t nitems, hipStream_t st);
hipError_t hsl_launch_combine(const CombineCol* cols, uint32_t ncols, const SegItem* items,
                              const uint32_t* T_S, const EscTable& esc, uint32_t cbits,
                              uint32_t slice_len, uint32_t nslices, long long* sC, long long* sM,
                              void* sMP, long long* carry, void* out, hipStream_t st);

} // namespace pcg
