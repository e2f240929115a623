// hsl_kernels.hip -- CDNA4 (gfx950) kernels of the hard-special-leaf engine.
//
// A persistent grid of single-wave workgroups takes segments from a device
// work queue.  A wave holds its segment in LDS as a wheel-30 bitmap (240
// numbers per 64-bit word, the layout of src/sieve/Sieve.cpp:62-66) and walks
// the sieving primes b = b_first, b_first + 1, ... in windows of up to 64 primes
// (lane l holds the per-prime values of b + l):
//   * leaves of p_b: factor-table entries (or primes) of the leaf range are
//     filtered, the exact quotient n = x / (p_b * m) is formed, and the
//     unsieved numbers <= n are counted from word-pair prefix counts (rebuilt
//     for every prime that has leaves in the segment) and popcounts of at most
//     2 words; M(s, b) is the sum of the leaf signs;
//   * cross-off of p_b: p_b and its multiples are cleared with LDS atomics;
//     every number removed for the first time is charged to p_b
//     (src/sieve/Sieve.cpp:497-845 counts the same first-time removals),
//     which gives C(s, b), the count of unsieved numbers before p_b is
//     crossed off.
// Small primes go one at a time: leaves, then cross-off.  Once every cofactor
// q of p_b in the segment is below p_b^2 (from then on for all larger primes
// of the segment), a multiple p_b * q that is still unsieved has q = 1 or q
// prime, so its least prime factor is min(p_b, q) and the first-time charge
// no longer depends on the order of the cross-offs.  The cross-offs of such
// primes run together across lanes, and only a prime that has leaves forces
// the pending ones to be applied before its count.
//
// The reference carries phi(low, b-1) sequentially through its segments
// (src/gourdon/D.cpp:96, 271).  Here each segment writes its C(s, b) and
// M(s, b) to HBM, and hsl_combine_* add the carry afterwards:
//   sum = sum_leaves sign * (phi(low_s - 1, b - 1) + count)
//       = sum_s S(s) + sum_b sum_s M(s, b) * sum_{s' < s} C(s', b).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "hsl_kernels.hpp"
#include "quotient_fixup.hpp"
#include "wheel_dev.hpp"

namespace pcg {

__constant__ int16_t c_cop_index[2310];   // largest index of a coprime <= r
__constant__ uint16_t c_cop[480];         // coprimes to 2310 in [0, 2310)
// c_cop continued for 64 entries past one period: entry i is the i-th number
// coprime to 2310 (i < 544), so a wave's 64 consecutive factor-table indices
// starting at residue r < 480 of a period read entries r .. r + 63
__constant__ uint16_t c_cop_ext[544];

/// Device memory of the no-allocation entry point (pcg_D_i64_noalloc): its
/// tables and work buffers, kHslScratchBytes per device for the life of the
/// process.  Host code never passes `scratch` itself (in host code the name is
/// a stand-in, not a device address); it uses the address that
/// hipGetSymbolAddress returns for the device (hsl_scratch_address).
static __device__ __align__(256) unsigned char scratch[kHslScratchBytes];

namespace {


/// LDS-only workgroup barrier: waits for this wave's LDS operations, not for
/// its global loads and stores, which stay in flight across the barrier.
__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

/// The lanes of a wave exchange values through LDS.  Without a fence the
/// compiler may reuse what a lane itself stored to an address (another lane's
/// store in between is invisible to it in the per-thread model); this fence
/// orders the wave's LDS accesses for the compiler.  The hardware executes a
/// wave's LDS operations in order, so no wait is needed at wave scope.
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/// floor(xp / m) - low for a leaf whose quotient lies in [low, low + 2^20)
/// (the leaf ranges guarantee low <= xp / m < high) and m < 2^32: r / m is
/// estimated in double with the hardware reciprocal, and the integer
/// residual corrects the estimate (quotient_fixup: a loop per direction,
/// exact whatever the reciprocal's error; with a relative error below 2^-21
/// the estimate is within one unit and each loop runs at most once).
__device__ __forceinline__ uint32_t leaf_stop(uint64_t xlo, uint64_t low, uint32_t m)
{
  return leaf_stop_from_rcp(xlo, low, m, __builtin_amdgcn_rcp((double) m));
}

/// min(floor(xp / d), cap) for cap < 2^48, with inv_d = 1 / d rounded to
/// double.  xd * inv_d carries a relative error below 3 * 2^-53, so below the
/// cap (quotient < 2^48 + 2) it is within 0.1 of xp / d and the residual
/// fixes the last unit; above it the comparison with cap + 2 is safe.
__device__ __forceinline__ uint64_t div_cap(double xd, uint64_t xlo, uint64_t d, double inv_d,
                                            uint64_t cap)
{
  double qd = xd * inv_d;
  if (qd >= (double) cap + 2.0)
    return cap;
  uint64_t q = (uint64_t) qd;
  int64_t r = (int64_t) (xlo - q * d);
  if (r < 0)
    q--;
  else if ((uint64_t) r >= d)
    q++;
  return q < cap ? q : cap;
}

/// to_index of src/BaseFactorTable.hpp:36-42: index of the largest number
/// coprime to 2310 that is <= n (n >= 1), for n < 2^32 -- every leaf-range
/// bound is (m <= z < 2^32 for D, m <= y < 2^32 for S2_hard) -- with 32-bit
/// arithmetic.
__device__ __forceinline__ uint32_t ft_to_index32(uint32_t n)
{
  const uint32_t q = n / 2310;
  return 480 * q + (uint32_t) c_cop_index[n - q * 2310];
}

/// pi(n) for n < 2^32 within the prime table (src/PiTable.hpp:46-57), with
/// 32-bit index arithmetic.
__device__ __forceinline__ uint64_t pi_lookup32(const uint64_t* __restrict__ bits,
                                                const uint32_t* __restrict__ cnt, uint32_t n)
{
  if (n < 7)
    return n < 2 ? 0 : (n < 3 ? 1 : (n < 5 ? 2 : 3));
  const uint32_t blk = n / 240;
  const uint32_t r = n - blk * 240;
  return 3 + (uint64_t) cnt[blk] + (uint64_t) __popcll(bits[blk] & le_mask(r));
}

/// 64-bit v_readlane (the builtin returns int; the low word must not be
/// sign-extended when the halves are joined).
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, uint32_t lane)
{
  uint32_t lo = (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) v, (int) lane);
  uint32_t hi = (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) (v >> 32), (int) lane);
  return ((uint64_t) hi << 32) | lo;
}

/// Calls f(std::integral_constant<int, i>{}) for i = 0 .. N - 1, in order.
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
  if constexpr (N > 0) {
    static_for<N - 1>(f);
    f(std::integral_constant<int, N - 1>{});
  }
}

/// A wave-uniform double moved to scalar registers (the compiler keeps the
/// result of a vector FP64 operation in vector registers and folds the
/// readfirstlane builtin on a value it knows to be uniform, so the move is
/// written as the instruction itself; no memory access).
__device__ __forceinline__ double uniform_f64(double v)
{
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  uint32_t lo, hi;
  asm volatile("v_readfirstlane_b32 %0, %1" : "=s"(lo) : "v"((uint32_t) b));
  asm volatile("v_readfirstlane_b32 %0, %1" : "=s"(hi) : "v"((uint32_t) (b >> 32)));
  return __builtin_bit_cast(double, ((uint64_t) hi << 32) | lo);
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t lane)
{
  return (uint32_t) __builtin_amdgcn_readlane((int) v, (int) lane);
}

__device__ __forceinline__ int wave_sum_i32(int v)
{
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ long long wave_sum_i64(long long v)
{
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

/// Inclusive prefix sum over the 64 lanes of a wave with DPP row shifts and
/// row broadcasts (no LDS round trips).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
  // row_shr:1,2,4,8 inside each row of 16 lanes (lanes without a source add 0)
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x111, 0xf, 0xf, false);
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x112, 0xf, 0xf, false);
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x114, 0xf, 0xf, false);
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x118, 0xf, 0xf, false);
  // row_bcast:15 adds lane 15 of a row to the next row, row_bcast:31 adds
  // lane 31 to rows 2 and 3
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x142, 0xa, 0xf, false);
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x143, 0xc, 0xf, false);
  return v;
}

/// build_pair_prefix for PER = W / 128 pairs per lane known at compile time:
/// the lane's PER 128-bit loads are issued together, then counted.
template <int PER, typename PT>
__device__ __forceinline__ void build_pair_prefix_n(const uint32_t* bm, PT* pre2, uint32_t lane)
{
  const uint4* w4 = reinterpret_cast<const uint4*>(bm) + lane * PER;
  uint4 q[PER];
#pragma unroll
  for (int i = 0; i < PER; i++)
    q[i] = w4[i];
  uint32_t c[PER];
  uint32_t tot = 0;
#pragma unroll
  for (int i = 0; i < PER; i++) {
    c[i] = __popc(q[i].x) + __popc(q[i].y) + __popc(q[i].z) + __popc(q[i].w);
    tot += c[i];
  }
  uint32_t run = wave_incl_scan(tot) - tot;
  PT* out = pre2 + lane * PER;
#pragma unroll
  for (int i = 0; i < PER; i++) {
    out[i] = (PT) run;
    run += c[i];
  }
}

/// Prefix counts of the segment at word-pair granularity: pre2[g] = unsieved
/// numbers in words [0, 2 g).  Lane l takes W / 128 consecutive pairs (one
/// 128-bit LDS load each), then a wave scan gives every lane its base.  The
/// segment sizes up to 640 words take a copy specialised for their pair
/// count (wave-uniform switch).
template <typename PT>
__device__ __forceinline__ void build_pair_prefix(const uint32_t* bm, PT* pre2, uint32_t W,
                                                  uint32_t lane)
{
  switch (W >> 7) {
    case 1: build_pair_prefix_n<1>(bm, pre2, lane); return;
    case 2: build_pair_prefix_n<2>(bm, pre2, lane); return;
    case 3: build_pair_prefix_n<3>(bm, pre2, lane); return;
    case 4: build_pair_prefix_n<4>(bm, pre2, lane); return;
    case 5: build_pair_prefix_n<5>(bm, pre2, lane); return;
    default: break;
  }
  const uint32_t per = W >> 7;
  const uint4* w4 = reinterpret_cast<const uint4*>(bm) + lane * per;
  PT* out = pre2 + lane * per;
  if (per <= 8) {
    uint32_t c[8];
    uint32_t tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; i++) {
      uint32_t v = 0;
      if (i < per) {
        const uint4 q = w4[i];
        v = __popc(q.x) + __popc(q.y) + __popc(q.z) + __popc(q.w);
      }
      c[i] = v;
      tot += v;
    }
    uint32_t run = wave_incl_scan(tot) - tot;
#pragma unroll
    for (uint32_t i = 0; i < 8; i++) {
      if (i < per)
        out[i] = (PT) run;
      run += c[i];
    }
  } else {
    uint32_t tot = 0;
    for (uint32_t i = 0; i < per; i++) {
      const uint4 q = w4[i];
      tot += __popc(q.x) + __popc(q.y) + __popc(q.z) + __popc(q.w);
    }
    uint32_t run = wave_incl_scan(tot) - tot;
    for (uint32_t i = 0; i < per; i++) {
      const uint4 q = w4[i];
      out[i] = (PT) run;
      run += __popc(q.x) + __popc(q.y) + __popc(q.z) + __popc(q.w);
    }
  }
}

/// Unsieved numbers in [low, low + stop] of the segment (count_stop.hpp:
/// the count is inclusive and relative to low): the prefix of the word pair,
/// the whole first word of the pair when stop lies in its second word, and
/// the low lut[stop % 240] bits of the word of stop (lut[r] = sieve bits of
/// the offsets <= r in a word).
template <typename PT>
__device__ __forceinline__ uint32_t seg_count(const uint32_t* bm, const PT* pre2,
                                              const uint8_t* lut, uint32_t stop)
{
  const uint32_t w = stop / 240;
  const uint32_t r = stop - w * 240;
  const uint32_t g = w >> 1;
  const uint4 q = reinterpret_cast<const uint4*>(bm)[g];
  const uint32_t n = lut[r];
  const bool odd = (w & 1) != 0;
  const uint32_t c = (uint32_t) pre2[g] + (odd ? __popc(q.x) + __popc(q.y) : 0u);
  const uint64_t word = odd ? (((uint64_t) q.w << 32) | q.z) : (((uint64_t) q.y << 32) | q.x);
  return c + (n ? (uint32_t) __popcll(word << (64 - n)) : 0u);
}

/// seg_count that also gives the bit-index bound of the count: the counted
/// numbers are the unsieved ones among the bits 64 w + i, i < lut[r], of the
/// segment, so a bit index below `bound` lies at or below stop.
template <typename PT>
__device__ __forceinline__ uint32_t seg_count_b(const uint32_t* bm, const PT* pre2,
                                                const uint8_t* lut, uint32_t stop, uint32_t& bound)
{
  const uint32_t w = stop / 240;
  const uint32_t r = stop - w * 240;
  const uint32_t g = w >> 1;
  const uint4 q = reinterpret_cast<const uint4*>(bm)[g];
  const uint32_t n = lut[r];
  const bool odd = (w & 1) != 0;
  const uint32_t c = (uint32_t) pre2[g] + (odd ? __popc(q.x) + __popc(q.y) : 0u);
  const uint64_t word = odd ? (((uint64_t) q.w << 32) | q.z) : (((uint64_t) q.y << 32) | q.x);
  bound = (w << 6) + n;
  return c + (n ? (uint32_t) __popcll(word << (64 - n)) : 0u);
}

/// Inclusive prefix maximum over the 64 lanes of a wave (same DPP pattern as
/// wave_incl_scan).
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v)
{
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x111, 0xf, 0xf, false));
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x112, 0xf, 0xf, false));
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x114, 0xf, 0xf, false));
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x118, 0xf, 0xf, false));
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x142, 0xa, 0xf, false));
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x143, 0xc, 0xf, false));
  return v;
}

/// Cross off p and its multiples p*q, q in [q_lo, q_hi] coprime to 30, in the
/// segment starting at low (Sieve::add/cross_off remove the multiples
/// > low, src/sieve/Sieve.cpp:192-216).  Lanes split the 8 residue classes
/// of q mod 30 and stride inside a class: consecutive multiples of one class
/// are p bytes apart.
__device__ __forceinline__ void cross_off_q(uint32_t p, uint64_t q_lo, uint64_t q_hi, uint64_t low,
                                            uint32_t* bm)
{
  if (p < 7 || q_hi < q_lo)
    return;
  const uint32_t cls = threadIdx.x & 7;
  const uint32_t j0 = threadIdx.x >> 3;
  uint64_t q = q_lo - q_lo % 30 + ((kWheelBytes >> (8 * cls)) & 0xff);
  if (q < q_lo)
    q += 30;
  if (q > q_hi)
    return;
  uint32_t n = (uint32_t) ((q_hi - q) / 30) + 1;
  uint32_t off = (uint32_t) ((uint64_t) p * q - low);
  uint32_t byte0 = off / 30;
  uint32_t bib = wheel_bit(off - byte0 * 30);
  for (uint32_t t = j0; t < n; t += kBlock / 8) {
    uint32_t B = byte0 + t * p;
    atomicAnd(&bm[B >> 2], ~(1u << (((B & 3) << 3) | bib)));
  }
}

/// Pre-sieve form: q range computed here.
__device__ __forceinline__ void cross_off_presieve(uint32_t p, uint64_t low, uint32_t seg_len,
                                                   uint32_t* bm)
{
  uint64_t q_lo = div_exact((double) low, low, p) + 1;
  uint64_t last = low + seg_len - 1;
  uint64_t q_hi = div_exact((double) last, last, p);
  cross_off_q(p, q_lo, q_hi, low, bm);
}

} // namespace

/// Inclusive-scan total of a wave (every lane contributes v).
__device__ __forceinline__ uint32_t wave_total_u32(uint32_t v)
{
  return readlane_u32(wave_incl_scan(v), 63);
}

__device__ __forceinline__ uint32_t esc_hash(unsigned long long key, uint32_t cap)
{
  return (uint32_t) ((key * 0x9E3779B97F4A7C15ull) >> 32) & (cap - 1);
}

/// Records M of slot `off` in the escape table (called by the lanes whose M
/// does not fit the slot word).  At most cap / 2 entries are inserted, so
/// the probe always reaches an empty entry; a further insert only raises
/// the overflow flag, which the host reports as an error.
__device__ __forceinline__ void esc_insert(EscTable t, uint64_t off, int32_t M)
{
  const unsigned int n = atomicAdd(&t.ctl[0], 1u);
  if (n >= t.cap / 2) {
    atomicOr(&t.ctl[1], 1u);
    return;
  }
  const unsigned long long key = (unsigned long long) off + 1;
  uint32_t h = esc_hash(key, t.cap);
  for (;;) {
    if (atomicCAS(&t.keys[h], 0ull, key) == 0ull) {
      t.vals[h] = M;
      return;
    }
    h = (h + 1) & (t.cap - 1);
  }
}

/// M of an escaped slot; 0 if it is missing (only after an overflow, which
/// the host turns into an error).  The probe ends at the first empty entry.
__device__ __forceinline__ int32_t esc_lookup(const EscTable& t, uint64_t off)
{
  const unsigned long long key = (unsigned long long) off + 1;
  uint32_t h = esc_hash(key, t.cap);
  for (uint32_t i = 0; i < t.cap; i++) {
    const unsigned long long k = t.keys[h];
    if (k == key) return t.vals[h];
    if (k == 0ull) return 0;
    h = (h + 1) & (t.cap - 1);
  }
  return 0;
}

/// Slot word of (C, M): C | M << cbits, or C | the escape marker (bit 31)
/// with M recorded in the escape table.
__device__ __forceinline__ uint32_t pack_slot(uint32_t C, int32_t M, uint32_t cbits, EscTable esc,
                                              uint64_t off)
{
  const int32_t lim = 1 << (31 - cbits);          // |M| < lim fits the signed field
  if (M > -lim && M < lim)
    return C | ((uint32_t) M << cbits);
  esc_insert(esc, off, M);
  return C | 0x80000000u;
}

/// One wave owns one segment at a time: the epochs need no workgroup
/// barriers (a wave's LDS operations execute in order), and the per-prime
/// bookkeeping runs once per segment and prime.
///
/// With NQ > 1 a work item is a group of up to NQ queries that share the
/// segment (the same interval [low, low + 240 W), the same first b and
/// pre-sieve): the sieve -- bitmap, cross-offs, first-time removals, pair
/// prefixes, the counts C(s, b) -- depends only on the segment and the
/// primes, so it runs once and every member answers its own leaves on it and
/// writes its own slots.  Each member stops at its own first empty leaf
/// range (its "goto next_segment"); the group stops when all have.
///
/// Read-only tables and written arrays come in as separate __restrict__
/// pointers so the compiler knows the tables are never written by the kernel
/// and can read uniform entries (BInfo, work items) through the scalar cache.
template <typename FT, int NQ, bool TIMERS = false>
__device__ __forceinline__ void sieve_segments(
    const QueryDesc* __restrict__ queries, const SegItem* __restrict__ items,
    const SegGroup* __restrict__ groups, const BInfo* __restrict__ binfo_all,
    const uint32_t* __restrict__ primes, const FT* __restrict__ factor,
    const uint64_t* __restrict__ pi_bits, const uint32_t* __restrict__ pi_cnt,
    uint32_t* __restrict__ T_S, EscTable esc, unsigned long long* __restrict__ S_acc,
    unsigned long long* __restrict__ counters, unsigned int* __restrict__ queue,
    uint32_t nitems, uint32_t W, uint32_t cbits, uint32_t bcap)
{
  extern __shared__ __align__(16) uint32_t lds[];
  constexpr bool BATCH = true;                                // batched leaf counting
  uint32_t* bm = lds;                                         // 2W u32
  uint32_t* pre2 = bm + 2 * W;                                // W/2 u32: word-pair prefixes
  uint8_t* lut = reinterpret_cast<uint8_t*>(pre2 + W / 2);    // 240 u8
  // LDS per wave bounds the waves per CU, and the kernel runs faster with
  // more of them, so what fits elsewhere lives elsewhere: the 480 numbers
  // coprime to 2310 are read from constant memory (c_cop), the leaf signs
  // of the window's primes stay in registers (lane j holds prime j's) and the
  // grouped cross-offs fetch their prime's data from its lane (ds_bpermute).
  __shared__ uint32_t w_rem[64];                              // per window prime: removals
  __shared__ uint32_t lq[128];                                // leaf queue: factor index | sign << 31
  // batched leaf counting (single-query kernel): the window lane (prime) of
  // each queued leaf, and the fresh removals recorded for the corrections
  __shared__ uint8_t lqb[BATCH ? 128 : 1];
  __shared__ uint16_t blist[BATCH ? kBatchRecords : 1];

  const uint32_t lane = threadIdx.x;
  // batched leaf counting (16-bit bit indices, W <= 1024) records at most
  // kBatchRecords removals
  if (!BATCH || W > 1024)
    bcap = 0;
  bcap = min(bcap, kBatchRecords);
  for (uint32_t r = lane; r < 240; r += kBlock)
    lut[r] = (uint8_t) (8 * (r / 30) + le_count(r % 30));

  const uint64_t SEG = 240ull * W;
  // Wave-uniform work counts, 32-bit so that they take one SGPR each; they
  // are added to the 64-bit totals whenever one passes 2^30 (checked after
  // each work item, which adds far less than 2^31) and at the end.
  uint32_t n_slots = 0, u_hits = 0, u_leaves = 0, u_entries = 0;
  // wave-level step counts: windows, primes with leaves, leaf-loop steps,
  // one-at-a-time cross-off primes and steps, grouped cross-off steps,
  // pair-prefix builds
  uint32_t u_win = 0, u_lprimes = 0, u_lsteps = 0, u_dprimes = 0, u_dsteps = 0, u_gsteps = 0,
           u_pbuilds = 0, u_qsteps = 0;
  // leaf batches and the first-time removals recorded for their corrections
  uint32_t u_batch = 0, u_brec = 0;
  auto flush_counters = [&]() {
    if (lane == 0) {
      const uint32_t v[14] = {u_leaves, u_hits, u_entries, n_slots, u_win, u_lprimes,
                              u_lsteps, u_dprimes, u_dsteps, u_gsteps, u_pbuilds, u_qsteps,
                              u_batch, u_brec};
#pragma unroll
      for (int i = 0; i < 14; i++)
        if (v[i])
          atomicAdd(&counters[i], (unsigned long long) v[i]);
    }
    n_slots = u_hits = u_leaves = u_entries = 0;
    u_win = u_lprimes = u_lsteps = u_dprimes = u_dsteps = u_gsteps = u_pbuilds = u_qsteps = 0;
    u_batch = u_brec = 0;
  };

  // Phase timers (the TIMERS build only): the wave's clock cycles per phase
  // (kHslPhase*), summed over the waves into counters[16 + phase].  A phase
  // switch reads s_memtime and adds the cycles since the last switch to the
  // phase that ends.
  uint32_t ph_acc[kHslPhases];
  uint64_t ph_t = 0;
  uint32_t ph_cur = kHslPhaseSegment;
  if constexpr (TIMERS) {
    for (uint32_t i = 0; i < kHslPhases; i++)
      ph_acc[i] = 0;
    ph_t = __builtin_amdgcn_s_memtime();
  }
  auto phase = [&](uint32_t next) {
    if constexpr (TIMERS) {
      const uint64_t t = __builtin_amdgcn_s_memtime();
      const uint32_t d = (uint32_t) (t - ph_t);
      // a switch on a uniform value keeps ph_acc in scalar registers
#pragma unroll
      for (uint32_t i = 0; i < kHslPhases; i++)
        if (ph_cur == i)
          ph_acc[i] += d;
      ph_t = t;
      ph_cur = next;
    }
  };

  for (;;) {
    phase(kHslPhaseSegment);
    uint32_t qi = 0;
    if (lane == 0)
      qi = atomicAdd(queue, 1u);
    qi = __builtin_amdgcn_readfirstlane(qi);
    if (qi >= nitems)
      break;
    // members of the work item (all wave-uniform)
    uint32_t nmem = 1;
    SegItem its[NQ];
    if constexpr (NQ == 1) {
      its[0] = items[qi];
    } else {
      const SegGroup gr = groups[qi];
      nmem = gr.n;
#pragma unroll
      for (int k = 0; k < NQ; k++)
        its[k] = items[gr.item[k < (int) nmem ? k : 0]];
    }
    const QueryDesc qd = queries[its[0].query];     // b_first, pre_last and limit are shared
    const BInfo* binfo[NQ];
    uint32_t* TS[NQ];
    uint32_t b_stop_m[NQ], stop_b[NQ];
    bool live[NQ];
    long long S[NQ];

    uint32_t b_stop = 0;
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      const QueryDesc q = queries[its[k].query];
      binfo[k] = binfo_all + q.binfo_off;
      TS[k] = T_S + its[k].t_off;
      b_stop_m[k] = qd.b_first + its[k].nbs - 1;
      live[k] = k < (int) nmem;
      stop_b[k] = live[k] ? b_stop_m[k] + 1 : 0u;
      S[k] = 0;
      if (live[k])
        b_stop = max(b_stop, b_stop_m[k]);
    }
    const uint64_t low = (uint64_t) its[0].seg * SEG;
    const uint64_t high = min(low + SEG, qd.limit);
    const uint32_t seg_len = (uint32_t) (high - low);
    const uint64_t low1 = low > 0 ? low : 1;
    // reciprocals of the segment bounds for the leaf-range quotients, in
    // scalar registers
    const double inv_low1 = uniform_f64(1.0 / (double) low1);
    const double inv_high = uniform_f64(1.0 / (double) high);
    const double low_d = (double) low, last_d = (double) (high - 1);   // exact: < 2^53

    // numbers coprime to 30 in [low, high)
    for (uint32_t w = lane; w < W; w += kBlock) {
      uint32_t base = w * 240;
      uint64_t v = base + 240 <= seg_len ? ~0ull : (base >= seg_len ? 0ull : le_mask(seg_len - base - 1));
      bm[2 * w] = (uint32_t) v;
      bm[2 * w + 1] = (uint32_t) (v >> 32);
    }
    // pre-sieve p_4 .. p_{b_first - 1} (2, 3, 5 are outside the wheel)
    for (uint32_t j = 4; j <= qd.pre_last; j++)
      cross_off_presieve(primes[j], low, seg_len, bm);
    wave_lds_sync();
    // the segment total
    uint32_t my_total = 0;
    {
      const uint4* w4 = reinterpret_cast<const uint4*>(bm);
      for (uint32_t i = lane; i < W / 2; i += kBlock) {
        const uint4 q = w4[i];
        my_total += __popc(q.x) + __popc(q.y) + __popc(q.z) + __popc(q.w);
      }
    }
    uint32_t total = wave_total_u32(my_total);   // unsieved numbers at the current b
    wave_lds_sync();

    uint32_t b = qd.b_first;
    bool done = false;
    // the word-pair prefixes hold the current sieve state unless a
    // first-time removal happened since they were built (wave-uniform)
    bool dirty = true;
    while (!done) {
      phase(kHslPhaseWindow);
      // Window of up to 64 primes b .. b + 63, lane l <-> b + l, kept in
      // registers; an epoch reads its prime's values with v_readlane.  A
      // member's window ends before its first b whose leaf range is provably
      // empty (that b and all later ones stop the member in this segment,
      // src/gourdon/D.cpp:131-132, 289-290); the shared window ends when
      // every member has stopped.
      const uint32_t bb = b + lane;
      uint32_t r_lo[NQ], r_hi[NQ], r_T[NQ];
      uint64_t r_xlo[NQ];
      bool brk[NQ];
      uint32_t r_nq = 0, r_h = 0, r_p = 0;
      // cofactor range [q_lo, q_hi] clipped to 32 bits (empty: lo > hi)
      uint32_t cq_lo = 0xFFFFFFFFu, cq_hi = 0;
      bool par = false;
      uint4 prm = make_uint4(0, 0, 0, 0);
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        r_lo[k] = r_hi[k] = r_T[k] = 0;
        r_xlo[k] = 0;
        brk[k] = true;
      }
      if (bb <= b_stop) {
        const uint32_t p = primes[bb];
#pragma unroll
        for (int k = 0; k < NQ; k++) {
          if (bb < stop_b[k]) {
            const BInfo bi = binfo[k][bb - qd.b_first];
            const uint64_t m_hi = div_cap(bi.xp_d, bi.xp_lo, low1, inv_low1, bi.A);
            uint64_t m_lo = div_cap(bi.xp_d, bi.xp_lo, high, inv_high, bi.C);
            m_lo = m_lo > bi.B ? m_lo : bi.B;
            // m_lo, m_hi <= max(A, B, C) <= z (D) or y: below 2^32
            if (bi.type == 1) {
              brk[k] = bi.p >= m_hi;
              if (!brk[k]) {
                r_hi[k] = ft_to_index32((uint32_t) m_hi);
                r_lo[k] = m_lo < m_hi ? ft_to_index32((uint32_t) m_lo) : r_hi[k];
              }
            } else {
              r_hi[k] = (uint32_t) pi_lookup32(pi_bits, pi_cnt, (uint32_t) m_hi);
              brk[k] = r_hi[k] <= bb;
              if (!brk[k])
                r_lo[k] = m_lo < m_hi ? (uint32_t) pi_lookup32(pi_bits, pi_cnt, (uint32_t) m_lo)
                                      : r_hi[k];
            }
            r_T[k] = bi.T | (bi.type == 1 ? 0x80000000u : 0u);
            r_xlo[k] = bi.xp_lo;
          }
        }
        r_p = p;
        // cofactors q of p with p*q in (low, high): q_lo .. q_lo + nq - 1
        // 1 / p from the hardware reciprocal and two Newton steps (relative
        // error near 2^-52); the quotients are below 2^48, so the products
        // are within 0.2 of low / p and (high - 1) / p
        double inv_p = __builtin_amdgcn_rcp((double) p);
        inv_p = fma(inv_p, fma(-(double) p, inv_p, 1.0), inv_p);
        inv_p = fma(inv_p, fma(-(double) p, inv_p, 1.0), inv_p);
        const uint64_t q_lo = div_exact_r(low_d, low, p, inv_p) + 1;
        const uint64_t q_hi = div_exact_r(last_d, high - 1, p, inv_p);
        r_nq = q_hi >= q_lo ? (uint32_t) (q_hi - q_lo + 1) : 0u;
        if (r_nq > 0) {
          cq_lo = q_lo < 0xFFFFFFFFull ? (uint32_t) q_lo : 0xFFFFFFFFu;
          cq_hi = q_hi < 0xFFFFFFFFull ? (uint32_t) q_hi : 0xFFFFFFFFu;
        }
        const uint32_t qm = (uint32_t) (q_lo % 30);
        const uint32_t off0 = (uint32_t) (p * q_lo - low);   // < 2^20 when nq > 0
        // Every cofactor is below p^2: a number of the segment that is still
        // unsieved when p is crossed off and is a multiple of p is p * q with
        // q = 1 or q prime, and its least prime factor is min(p, q).
        par = q_hi < (uint64_t) p * p;
        if (r_nq > 0 && p >= 7) {
          // hits: cofactors coprime to 30; c0 = wheel residues below q_lo % 30
          // (2, 3 and 5 are not in the wheel and have no hits)
          const uint32_t c0 = qm ? le_count(qm - 1) : 0u;
          const uint32_t v = qm + r_nq - 1;
          const uint32_t vb = v / 30;
          r_h = 8 * vb + le_count(v - 30 * vb) - c0;
          // the multiple p (q_lo + delta) lies at offset off0 + p delta of
          // the segment: byte (off0 + p delta) / 30, its wheel bit from the
          // remainder (off0 < 2^20, c0 < 8, qm < 30)
          prm.y = off0 | (c0 << 20) | (qm << 23);
        }
        prm.z = q_lo < p ? (uint32_t) (p - q_lo) : 0u;
      }
      // per member: its window end and its primes with factor entries
      uint32_t nwin_m[NQ];
      unsigned long long lmask_m[NQ];
      uint32_t nwin = 0;
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        const unsigned long long bal = __ballot(brk[k]);
        nwin_m[k] = bal ? (uint32_t) __ffsll((long long) bal) - 1 : 64u;
        nwin = max(nwin, nwin_m[k]);
        const unsigned long long in_k = nwin_m[k] >= 64 ? ~0ull : ((1ull << nwin_m[k]) - 1);
        lmask_m[k] = __ballot(r_hi[k] > r_lo[k]) & in_k;
      }
      if (nwin < 64)
        done = true;
      const unsigned long long in_win = nwin >= 64 ? ~0ull : ((1ull << nwin) - 1);
      // par (monotone in p) holds for the primes [seq_end, nwin) of the window
      const unsigned long long pbal = __ballot(par) & in_win;
      // primes with cofactors below p (q_lo < p)
      const unsigned long long zmask = __ballot(prm.z != 0) & in_win;
      const uint32_t seq_end = pbal ? (uint32_t) __ffsll((long long) pbal) - 1 : nwin;
      // primes with factor entries of any member; the cross-offs between two
      // of them are applied together just before the count of the second
      unsigned long long lmask = 0;
#pragma unroll
      for (int k = 0; k < NQ; k++)
        lmask |= lmask_m[k];
      const uint32_t h_win = lane < nwin ? r_h : 0u;
      // every prime of the window is crossed off by the window's end: its
      // hits are its cofactors coprime to 30
      const uint32_t h_scan = wave_incl_scan(h_win);
      u_hits += readlane_u32(h_scan, 63);
      // primes [0, d_win) include every one with more than 32 hits; the
      // others take 2^lg lanes each in a grouped cross-off, 2^lg >= hits
      const unsigned long long big = __ballot(h_win > 32);
      const uint32_t d_win = big ? 64u - (uint32_t) __clzll((long long) big) : 0u;
      const uint32_t hm = readlane_u32(wave_incl_max(lane >= d_win ? h_win : 0u), 63);
      const uint32_t lg = hm <= 1 ? 0u : 32u - (uint32_t) __clz(hm - 1);      // hm <= 32
      const uint32_t G = 1u << lg;
      const uint32_t grp = lane >> lg;
      const uint32_t ngrp = kBlock >> lg;
      prm.w = r_h;
      w_rem[lane] = 0;
      int w_m[NQ];                     // lane j: the leaf-sign sum M of prime j, per member
#pragma unroll
      for (int k = 0; k < NQ; k++)
        w_m[k] = 0;
      wave_lds_sync();

      // Cross off primes [r0, r1) of the window, in two parts:
      //  * primes up to the last one with more than 32 hits go one at a time,
      //    lanes split the 8 residue classes of q mod 30 and stride 8
      //    multiples apart inside a class (p bytes per multiple);
      //  * the rest together, 2^lg lanes per prime (2^lg >= hits), lane k of a
      //    prime's group taking its k-th cofactor coprime to 30.  If the part
      //    holds a prime below seq_end, the atomics of the primes are issued
      //    one prime group after the other, which keeps the first-time
      //    removals of the sequential order; otherwise they are issued at once
      //    and a number removed for the first time is charged to its least
      //    prime factor: p, or the cofactor q when 1 < q < p (q is then a
      //    prime of this part).
      const uint32_t kk = lane & (G - 1);
      uint32_t rec_n = 0;              // fresh removals recorded in blist (wave-uniform)
      // Grouped cross-off steps of the primes [s0, r1).  ORD: the groups of a
      // step issue their atomics one after the other; ATTR: a removal whose
      // cofactor q is a prime below p is charged to q; REC: the fresh
      // removals of the primes from rec_from on are appended to blist.
      auto steps = [&](auto ord_c, auto attr_c, auto rec_c, uint32_t s0, uint32_t r1,
                     uint32_t rec_from) -> unsigned long long {
        constexpr bool ORD = decltype(ord_c)::value, ATTR = decltype(attr_c)::value;
        constexpr bool REC = decltype(rec_c)::value;
        unsigned long long fresh_any = 0;
        for (uint32_t jb = s0; jb < r1; jb += ngrp) {
          u_gsteps++;
          const uint32_t jo = jb + grp;
          // the cross-off data of prime jo, read from lane jo (all lanes
          // active; ds_bpermute takes the lane modulo 64, and a lane past r1
          // has no hit, so what it reads does not matter)
          const int src = (int) (jo << 2);
          uint4 pp;
          pp.x = (uint32_t) __builtin_amdgcn_ds_bpermute(src, (int) r_p);
          pp.y = (uint32_t) __builtin_amdgcn_ds_bpermute(src, (int) prm.y);
          pp.z = ATTR ? (uint32_t) __builtin_amdgcn_ds_bpermute(src, (int) prm.z) : 0u;
          pp.w = (uint32_t) __builtin_amdgcn_ds_bpermute(src, (int) prm.w);
          const bool has = jo < r1 && kk < pp.w;
          // cofactor q_lo + delta: the k-th coprime to 30 from q_lo (delta <
          // 180 for a lane with a hit, kk < hits <= 32); its multiple lies at
          // offset off < 2^20 of the segment, byte off / 30 = mulhi(off,
          // ceil(2^32 / 30)) (exact below 2^28); other lanes do not touch the
          // segment
          const uint32_t idx = ((pp.y >> 20) & 7u) + kk;
          const uint32_t delta = 30 * (idx >> 3) + wheel_residue(idx & 7) - (pp.y >> 23);
          const uint32_t off = (pp.y & 0xFFFFFu) + __umul24(pp.x, delta);
          const uint32_t B = __umulhi(off, 143165577u);
          const uint32_t bit = ((B & 3) << 3) | wheel_bit(off - 30 * B);
          const uint32_t mask = 1u << bit;
          uint32_t old = 0;
          if constexpr (ORD) {
            // the groups of this step one after the other
            const uint32_t ng = min(ngrp, r1 - jb);
            for (uint32_t u = 0; u < ng; u++) {
              if (has && grp == u)
                old = atomicAnd(&bm[B >> 2], ~mask);
              wave_lds_sync();
            }
          } else {
            if (has)
              old = atomicAnd(&bm[B >> 2], ~mask);
          }
          // old is 0 on a lane without a hit
          const bool fresh = (old & mask) != 0;
          fresh_any |= __ballot(fresh);
          if constexpr (REC) {
            // bit index 8 B + wheel bit (below 2^16 for W <= 1024); the
            // lanes of a step hold its primes in order, so the records come
            // out sorted by prime
            const bool rec = fresh && jo >= rec_from;
            const unsigned long long rb = __ballot(rec);
            if (rec) {
              const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                  (uint32_t) (rb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t) rb, 0u));
              blist[rec_n + rank] = (uint16_t) (((B & ~3u) << 3) | bit);
            }
            rec_n += (uint32_t) __popcll(rb);
          }
          if (fresh) {
            uint32_t attr = jo;
            if constexpr (ATTR) {
              if (delta < pp.z) {                     // q = q_lo + delta < p
                const uint32_t q = pp.x - pp.z + delta;
                if (q > 1)
                  attr = (uint32_t) pi_lookup32(pi_bits, pi_cnt, q) - b;
              }
            }
            atomicAdd(&w_rem[attr], 1u);
          }
        }
        return fresh_any;
      };

      auto flush = [&](uint32_t r0, uint32_t r1) {
        if (r0 >= r1)
          return;
        const uint32_t ph_back = ph_cur;
        phase(kHslPhaseDense);
        const uint32_t d_end = min(max(d_win, r0), r1);
        for (uint32_t j = r0; j < d_end; j++) {
          const uint32_t p = readlane_u32(r_p, j);
          const uint32_t nq = readlane_u32(r_nq, j);
          u_dprimes++;
          uint32_t removed = 0;                         // wave-uniform
          if (nq > 0 && p >= 7) {
            const uint32_t py = readlane_u32(prm.y, j);
            const uint32_t qm = py >> 23;
            const uint32_t cls = lane & 7;
            const uint32_t r = wheel_residue(cls);
            const uint32_t delta = r >= qm ? r - qm : r + 30 - qm;   // < 30
            // byte and bit of the class's first multiple p (q_lo + delta), at
            // offset off < 2^20 + 29 p < 2^28 (these primes have more than 32
            // hits in the segment, so p < 2^16)
            const uint32_t off = (py & 0xFFFFFu) + __umul24(p, delta);
            const uint32_t B = __umulhi(off, 143165577u);
            const uint32_t j0 = lane >> 3;
            // lane j0 takes multiples j0, j0 + 8, ... of its class: p bytes
            // apart per multiple, so 8 p bytes (2 p words) per step, and the
            // bit within the word does not change
            const uint32_t B0 = B + j0 * p;
            const uint32_t mask = 1u << (((B0 & 3) << 3) | wheel_bit(off - 30 * B));
            const uint32_t nmax = (nq + 29) / 30;       // >= every class count
            const uint32_t nsteps = (nmax + 7) / 8;
            u_dsteps += nsteps;
            const uint32_t wstep = 2 * p;
            uint32_t wi = B0 >> 2;
            // Two multiples per lane per step: both atomics are in flight
            // together (one prime's removals may be counted in any order).
            // A lane stops at the end of the bitmap instead of at its class
            // count: a multiple p q with q > q_hi lies at or beyond `high`,
            // so it is outside the bitmap or on a bit that is already clear
            // (the numbers of a short last segment from seg_len on).
            const uint32_t nw = 2 * W;
            for (uint32_t u = 0; u < nsteps; u += 2) {
              uint32_t o0 = 0, o1 = 0;
              if (wi < nw)
                o0 = atomicAnd(&bm[wi], ~mask);
              if (wi + wstep < nw)
                o1 = atomicAnd(&bm[wi + wstep], ~mask);
              removed += (uint32_t) __popcll(__ballot((o0 & mask) != 0)) +
                         (uint32_t) __popcll(__ballot((o1 & mask) != 0));
              wi += 2 * wstep;
            }
          }
          dirty |= removed != 0;
          if (lane == 0)
            w_rem[j] = removed;
        }
        if (d_end >= r1) {
          phase(ph_back);
          return;
        }
        phase(kHslPhaseGroup);
        const bool ordered = d_end < seq_end;
        unsigned long long fresh_any;
        if (ordered)
          fresh_any = steps(std::true_type{}, std::false_type{}, std::false_type{}, d_end, r1, 64u);
        else
          fresh_any = steps(std::false_type{}, std::true_type{}, std::false_type{}, d_end, r1, 64u);
        dirty |= fresh_any != 0;
        phase(ph_back);
      };

      auto next_leaf = [&](uint32_t c) -> uint32_t {
        const unsigned long long m = c >= 64 ? 0ull : (lmask & (~0ull << c));
        return m ? (uint32_t) __ffsll((long long) m) - 1 : 64u;
      };

      // Batched leaf counting (single-query kernel, W <= 1024).  From
      // bat_from = max(seq_end, d_win) on, the window's primes are crossed
      // off in any order; if in addition no such prime's cofactor range
      // [q_lo, q_hi] meets the window's primes, no number of the segment is a
      // multiple of two of them that is still unsieved, so each first-time
      // removal belongs to the one prime that hits it whatever the order.
      // Then the leaves of consecutive primes j .. jb with leaves are counted
      // against one prefix build, made after crossing off every prime before
      // jb: a leaf of prime jj < jb adds the first-time removals of the
      // primes jj .. jb - 1 that lie at or below its stop (they are unsieved
      // in its own state, sieved in the prefix state), which the cross-off
      // steps record in blist, sorted by prime.  A batch holds at most
      // kBatchEntries factor entries (leaves) and its recorded primes at most
      // bcap hits (records), which bounds blist and the correction work.
      bool bat = false;
      uint32_t bat_from = 64;
      // per lane: factor entries of the member's leaf primes <= the lane's
      // prime (s_ent), and of all members at and up to it (e_tot, s_tot)
      uint32_t s_ent[NQ];
      uint32_t e_tot = 0, s_tot = 0;
#pragma unroll
      for (int k = 0; k < NQ; k++)
        s_ent[k] = 0;
      if constexpr (BATCH) {
        bat_from = max(seq_end, d_win);
        const unsigned long long cand = bat_from < 64 ? lmask & (~0ull << bat_from) : 0ull;
        if (bcap != 0 && __popcll(cand) >= 2) {
          // no prime from bat_from on has a window prime among its cofactors
          const uint32_t pf = readlane_u32(r_p, 0), pl = readlane_u32(r_p, nwin - 1);
          bat = __ballot(lane >= bat_from && lane < nwin && cq_lo <= pl && cq_hi >= pf) == 0;
          if (bat) {
#pragma unroll
            for (int k = 0; k < NQ; k++) {
              const uint32_t ek = ((lmask_m[k] >> lane) & 1) ? r_hi[k] - r_lo[k] : 0u;
              if (NQ > 1)
                s_ent[k] = wave_incl_scan(ek);
              e_tot += ek;
            }
            s_tot = wave_incl_scan(e_tot);
            if (NQ == 1)
              s_ent[0] = s_tot;
          }
        }
      }
      // the last prime with leaves that can join a batch starting at j: the
      // batch's primes with leaves have at most kBatchEntries factor entries
      // (or prime leaves) over all members, its recorded primes [j, jb) at
      // most bcap hits
      auto batch_end = [&](uint32_t j) -> uint32_t {
        const uint32_t ej = NQ == 1 ? readlane_u32(r_hi[0], j) - readlane_u32(r_lo[0], j)
                                    : readlane_u32(e_tot, j);
        const uint32_t be = readlane_u32(s_tot, j) - ej;
        const uint32_t bh = readlane_u32(h_scan, j) - readlane_u32(h_win, j);
        const bool ok = ((lmask >> lane) & 1) && lane >= j && s_tot - be <= kBatchEntries &&
                        h_scan - h_win - bh <= bcap;
        const unsigned long long bal = __ballot(ok);
        return bal ? 63u - (uint32_t) __clzll((long long) bal) : j;
      };
      // Counts the leaves of the primes with leaves in [j, jb].
      auto run_batch = [&](uint32_t j, uint32_t jb, uint32_t r0) {
        if (r0 < bat_from) {
          flush(r0, bat_from);
          r0 = bat_from;
        }
        rec_n = 0;
        phase(kHslPhaseBatchFlush);
        if (r0 < jb)
          dirty |= steps(std::false_type{}, std::false_type{}, std::true_type{}, r0, jb, j) != 0;
        wave_lds_sync();
        // first record of each prime of [j, jb]; the records of primes >= jj
        // are blist[foff(jj) .. R)
        const uint32_t fr = lane >= j && lane < jb ? w_rem[lane] : 0u;
        const uint32_t foff = wave_incl_scan(fr) - fr;
        const uint32_t R = rec_n;
        u_batch++;
        u_brec += R;
        if (dirty) {
          phase(kHslPhasePrefix);
          build_pair_prefix(bm, pre2, W, lane);
          wave_lds_sync();
          dirty = false;
          u_pbuilds++;
        }
        phase(kHslPhaseBatchEntry);
        // member by member (their leaves are counted against the same
        // prefix build and records): each member's leaves go through the
        // queue in prime order and are drained before the next member's
        const unsigned long long in_bat = (~0ull << j) & (jb >= 63 ? ~0ull : ((2ull << jb) - 1));
        static_for<NQ>([&](auto kc) {
          constexpr int K = decltype(kc)::value;
          const unsigned long long lm = lmask_m[K] & in_bat;
          if (!lm)
            return;
          auto next_k = [&](uint32_t c) -> uint32_t {
            const unsigned long long m = c >= 64 ? 0ull : (lm & (~0ull << c));
            return m ? (uint32_t) __ffsll((long long) m) - 1 : 64u;
          };
          const uint32_t xlo_lo = (uint32_t) r_xlo[K], xlo_hi = (uint32_t) (r_xlo[K] >> 32);
          uint32_t qh = 0, qt = 0;                   // queue head and tail (mod 128)
          auto drain_b = [&](uint32_t n) {
            phase(kHslPhaseBatchDrain);
            wave_lds_sync();
            uint32_t e = 0, jl = 0;
            if (lane < n) {
              e = lq[(qt + lane) & 127];
              jl = lqb[(qt + lane) & 127];
            }
            // the leaf's prime: x / p and its first record, from its lane
            const int src = (int) (jl << 2);
            const uint32_t xl = (uint32_t) __builtin_amdgcn_ds_bpermute(src, (int) xlo_lo);
            const uint32_t xh = (uint32_t) __builtin_amdgcn_ds_bpermute(src, (int) xlo_hi);
            const uint32_t st = (uint32_t) __builtin_amdgcn_ds_bpermute(src, (int) foff);
            uint32_t bound = 0;
            int cnt = 0;
            if (lane < n) {
              const uint32_t m = (e << 1) | 1u;
              cnt = (int) seg_count_b(bm, pre2, lut, leaf_stop(((uint64_t) xh << 32) | xl, low, m),
                                      bound);
            }
            // lane 0 holds the earliest prime of the step (the member's queue
            // is in prime order), so records before its first one count for
            // no lane
            const uint32_t e0 = (uint32_t) __builtin_amdgcn_readfirstlane((int) st);
            for (uint32_t cb = e0; cb < R; cb += kBlock) {
              const uint32_t ix = cb + lane;
              const uint32_t rv = ix < R ? (uint32_t) blist[ix] : 0u;
              // wave-uniform trip count, kept in a scalar register
              const uint32_t ne = (uint32_t) __builtin_amdgcn_readfirstlane(
                  (int) min((uint32_t) kBlock, R - cb));
              for (uint32_t i = 0; i < ne; i++) {
                const uint32_t v = readlane_u32(rv, i);
                cnt += (cb + i >= st && v < bound) ? 1 : 0;
              }
            }
            if (lane < n)
              S[K] += (long long) ((e >> 31) ? cnt : -cnt);
            wave_lds_sync();
            qt += n;
            u_qsteps++;
            phase(kHslPhaseBatchEntry);
          };
          // The member's leaves of the batch's primes with entries in [j, jb]:
          // their factor entries (or primes) are gathered 64 at a time across
          // the primes.  The ranges follow each other in prime order (s_ent is
          // the inclusive scan of the window's entry counts), so entry e of the
          // batch belongs to the last prime pe whose first entry pst is <= e,
          // at factor-table (or prime) index lo(pe) + 1 + e - pst.  One load
          // per lane and step replaces a step per prime; the leaves go to the
          // queue with their prime's lane.
          const uint32_t j0 = next_k(j);
          const uint32_t E1 = readlane_u32(s_ent[K], jb);
          uint32_t jc = j0;                                 // first prime of the step
          uint32_t jc_st = readlane_u32(s_ent[K], j0) - (readlane_u32(r_hi[K], j0) - readlane_u32(r_lo[K], j0));
          {
            // M of a prime whose leaves are primes: all of them, positive
            const bool in_b = ((lm >> lane) & 1) && !(r_T[K] >> 31);
            w_m[K] = in_b ? (int) (r_hi[K] - r_lo[K]) : w_m[K];
            u_lprimes += (uint32_t) __popcll(lm);
          }
          for (uint32_t e0 = jc_st; e0 < E1; e0 += kBlock) {
            const uint32_t e = e0 + lane;
            uint32_t pe = jc, pst = jc_st;
            uint32_t nst = readlane_u32(s_ent[K], jc);     // first entry of the next prime
            for (uint32_t jj = next_k(jc + 1); jj <= jb && nst < e0 + kBlock; jj = next_k(jj + 1)) {
              if (e >= nst) {
                pe = jj;
                pst = nst;
              }
              nst = readlane_u32(s_ent[K], jj);
            }
            const bool valid = e < E1;
            const int src = (int) (pe << 2);
            const uint32_t lo = (uint32_t) __builtin_amdgcn_ds_bpermute(src, (int) r_lo[K]);
            const uint32_t Tt = (uint32_t) __builtin_amdgcn_ds_bpermute(src, (int) r_T[K]);
            const uint32_t idx = lo + 1 + (e - pst);
            const bool t1 = (Tt >> 31) != 0;
            uint32_t f = 1, m = 0;                         // a prime leaf: positive
            if (valid) {
              if (t1) {
                f = (uint32_t) factor[idx];
                const uint32_t qd = idx / 480;
                m = 2310 * qd + c_cop[idx - 480 * qd];     // to_number(idx)
              } else {
                m = primes[idx];
              }
            }
            const bool leaf = valid && (!t1 || f > (Tt & 0x7fffffffu));
            const unsigned long long lb = __ballot(leaf);
            if (leaf) {
              const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                  (uint32_t) (lb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t) lb, 0u));
              lq[(qh + rank) & 127] = (m >> 1) | (f << 31);
              lqb[(qh + rank) & 127] = (uint8_t) pe;
            }
            const uint32_t nl = (uint32_t) __popcll(lb);
            u_entries += (uint32_t) __popcll(__ballot(valid && t1));
            u_leaves += nl;
            u_lsteps++;
            // M of the step's primes with factor entries: 2 (positive leaves) - leaves
            const uint32_t pe_last = readlane_u32(pe, 63);
            for (uint32_t jj = jc; jj <= pe_last; jj = next_k(jj + 1)) {
              const bool mine = leaf && t1 && pe == jj;
              const int a = (int) __popcll(__ballot(mine));
              const int pp = (int) __popcll(__ballot(mine && (f & 1)));
              w_m[K] += lane == jj ? 2 * pp - a : 0;
            }
            qh += nl;
            if (qh - qt >= 64)
              drain_b(64);
            jc = pe_last;
            jc_st = readlane_u32(pst, 63);
          }
          if (qh != qt)
            drain_b(qh - qt);
        });
      };


      uint32_t run = 0;                 // first prime whose cross-off is pending
      for (uint32_t j = next_leaf(0); j < nwin; j = next_leaf(j + 1)) {
        if constexpr (BATCH) {
          if (bat && j >= bat_from) {
            const uint32_t jb = batch_end(j);
            if (jb > j) {
              run_batch(j, jb, run);
              run = jb;
              j = jb;
              continue;
            }
          }
        }
        // members with leaves at j: factor entries that fit one step and
        // hold no leaf (every entry <= T, a quarter of the primes with
        // entries at 1e19) are passed over, and the cross-offs stay pending
        phase(kHslPhaseScan);
        uint32_t active = 0;
        // the first step of each member's leaf range: its factor entries (or
        // primes), loaded here so that the loads are in flight during the
        // cross-offs and the prefix build
        uint32_t first[NQ];
#pragma unroll
        for (int k = 0; k < NQ; k++) {
          first[k] = 0;
          if (!((lmask_m[k] >> j) & 1))
            continue;
          const uint32_t lo = readlane_u32(r_lo[k], j);
          const uint32_t hi = readlane_u32(r_hi[k], j);
          const uint32_t Tt = readlane_u32(r_T[k], j);
          const uint32_t idx = lo + 1 + lane;
          bool has = true;
          if (Tt >> 31) {
            first[k] = idx <= hi ? (uint32_t) factor[idx] : 0u;
            if (hi - lo <= (uint32_t) kBlock)
              has = __ballot(first[k] > (Tt & 0x7fffffffu)) != 0;
          } else {
            first[k] = idx <= hi ? primes[idx] : 0u;
          }
          if (has)
            active |= 1u << k;
        }
        if (!active)
          continue;
        flush(run, j);
        run = j;
        // word-pair prefixes of the current sieve state for the count queries
        if (dirty) {
          phase(kHslPhasePrefix);
          build_pair_prefix(bm, pre2, W, lane);
          wave_lds_sync();
          dirty = false;
          u_pbuilds++;
        }
        phase(kHslPhaseEntry);
#pragma unroll
        for (int k = 0; k < NQ; k++) {
          if (!((active >> k) & 1))
            continue;
          u_lprimes++;
          const uint32_t lo = readlane_u32(r_lo[k], j);
          const uint32_t hi = readlane_u32(r_hi[k], j);
          const uint32_t Tt = readlane_u32(r_T[k], j);
          const uint32_t T = Tt & 0x7fffffffu;
          const uint64_t xlo = readlane_u64(r_xlo[k], j);
          int M = 0;                                      // wave-uniform
          if (Tt >> 31) {
            // factor-table entries: leaves -mu(m) phi(xp / m, b - 1) for the m
            // whose entry exceeds T (src/gourdon/D.cpp:148-188)
            u_entries += hi - lo;
            u_lsteps += (hi - lo + kBlock - 1) / kBlock;
            // The leaves (about a third of the entries) are queued in LDS as
            // (m >> 1) | sign << 31 (m is odd and below 2^32) and counted 64
            // at a time, so that the quotient and the count run on full waves
            // (189 M counting steps instead of 427 M at x = 1e19).  qh, qt:
            // head and tail of the queue (mod 128), wave-uniform.
            uint32_t qh = 0, qt = 0;
            auto drain = [&](uint32_t n) {
              phase(kHslPhaseDrain);
              wave_lds_sync();
              if (lane < n) {
                const uint32_t e = lq[(qt + lane) & 127];
                const uint32_t m = (e << 1) | 1u;
                const int cnt = (int) seg_count(bm, pre2, lut, leaf_stop(xlo, low, m));
                S[k] += (long long) ((e >> 31) ? cnt : -cnt);
              }
              wave_lds_sync();
              qt += n;
              u_qsteps++;
              phase(kHslPhaseEntry);
            };
            // The factor-table index of the wave's step, idx = base + lane,
            // is walked as a scalar period start mb = 2310 (base / 480) and
            // residue rb = base % 480, so m = mb + c_cop_ext[rb + lane].
            // Entries and coprimes of the next step are loaded before this
            // step is filtered, queued and counted.
            uint32_t f = first[k];
            uint32_t np_tot = 0;                          // wave-uniform
            uint32_t rb = (lo + 1) % 480, mb = 2310 * ((lo + 1) / 480);
            uint32_t cop = c_cop_ext[rb + lane];
#pragma nounroll
            for (uint32_t base = lo + 1; base <= hi; base += kBlock) {
              const uint32_t nidx = base + lane + kBlock;
              // read in bounds (index clamped to hi), then zeroed beyond hi:
              // no branch around the load
              const uint32_t f_next = nidx <= hi ? (uint32_t) factor[nidx] : 0u;
              uint32_t rbn = rb + kBlock, mbn = mb;
              if (rbn >= 480) {
                rbn -= 480;
                mbn += 2310;
              }
              const uint32_t cop_next = c_cop_ext[rbn + lane];
              const bool leaf = f > T;
              const unsigned long long lb = __ballot(leaf);
              const uint32_t nl = (uint32_t) __popcll(lb);
              np_tot += (uint32_t) __popcll(__ballot(leaf && (f & 1)));
              if (leaf) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t) (lb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t) lb, 0u));
                const uint32_t m = mb + cop;
                lq[(qh + rank) & 127] = (m >> 1) | (f << 31);
              }
              qh += nl;
              if (qh - qt >= 64)
                drain(64);
              f = f_next;
              cop = cop_next;
              rb = rbn;
              mb = mbn;
            }
            if (qh != qt)
              drain(qh - qt);
            // qh counts the prime's leaves, np_tot the positive ones
            u_leaves += qh;
            M = (int) (2 * np_tot) - (int) qh;
          } else {
            // primes q: leaves +phi(xp / q, b - 1) (src/gourdon/D.cpp:279-302)
            u_leaves += hi - lo;
            u_lsteps += (hi - lo + kBlock - 1) / kBlock;
            M = (int) (hi - lo);
            uint32_t q = first[k];
#pragma nounroll
            for (uint32_t base = lo + 1; base <= hi; base += kBlock) {
              const uint32_t idx = base + lane;
              const uint32_t q_next = idx + kBlock <= hi ? primes[idx + kBlock] : 0u;
              if (idx <= hi)
                S[k] += seg_count(bm, pre2, lut, leaf_stop(xlo, low, q));
              q = q_next;
            }
          }
          w_m[k] = lane == j ? M : w_m[k];
        }
      }
      flush(run, nwin);
      phase(kHslPhaseWindow);
      wave_lds_sync();
      {
        // C(s, b + l) = count at the window start minus the removals of the
        // window's earlier primes; every member writes the primes before its
        // own window end
        const uint32_t rem = lane < nwin ? w_rem[lane] : 0u;
        const uint32_t incl = wave_incl_scan(rem);
        const uint32_t C = total - (incl - rem);
#pragma unroll
        for (int k = 0; k < NQ; k++) {
          if (live[k] && b + lane < stop_b[k] && lane < nwin_m[k]) {
            const uint32_t rel = b + lane - qd.b_first;
            TS[k][rel] = pack_slot(C, w_m[k], cbits, esc, its[k].t_off + rel);
          }
        }
        total -= readlane_u32(incl, 63);
      }
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        if (live[k] && nwin_m[k] < 64)
          stop_b[k] = min(stop_b[k], b + nwin_m[k]);
      }
      n_slots += nwin;
      u_win++;
      b += nwin;
    }
    phase(kHslPhaseSegment);
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      if (!live[k])
        continue;
      // slots the member never reached hold zero
      for (uint32_t bq = stop_b[k] + lane; bq <= b_stop_m[k]; bq += kBlock)
        TS[k][bq - qd.b_first] = 0;
      const long long tot = wave_sum_i64(S[k]);
      if (lane == 0 && tot != 0) {
        // 128-bit accumulation from two 64-bit atomics: the add whose low word
        // wraps carries into the high word (exact under any interleaving)
        unsigned long long* acc = S_acc + 2 * its[k].query;
        const unsigned long long v = (unsigned long long) tot;
        const unsigned long long old = atomicAdd(&acc[0], v);
        const unsigned long long inc = (old + v < old ? 1ull : 0ull) - (tot < 0 ? 1ull : 0ull);
        if (inc != 0)
          atomicAdd(&acc[1], inc);
      }
    }
    if ((n_slots | u_hits | u_leaves | u_entries | u_win | u_lprimes | u_lsteps | u_dprimes |
         u_dsteps | u_gsteps | u_pbuilds | u_qsteps | u_batch | u_brec) >= (1u << 30))
      flush_counters();
  }

  // the u_* counters are updated in wave-uniform control flow only
  flush_counters();
  if constexpr (TIMERS) {
    phase(kHslPhaseSegment);
    if (lane == 0)
      for (uint32_t i = 0; i < kHslPhases; i++)
        atomicAdd(&counters[kHslPhaseCounter0 + i], (unsigned long long) ph_acc[i]);
  }
}

#define PCG_SIEVE_PARAMS                                                                       \
    const QueryDesc* __restrict__ queries, const SegItem* __restrict__ items,                \
    const SegGroup* __restrict__ groups, const BInfo* __restrict__ binfo_all,                \
    const uint32_t* __restrict__ primes, const FT* __restrict__ factor,                      \
    const uint64_t* __restrict__ pi_bits, const uint32_t* __restrict__ pi_cnt,               \
    uint32_t* __restrict__ T_S, EscTable esc, unsigned long long* __restrict__ S_acc,        \
    unsigned long long* __restrict__ counters, unsigned int* __restrict__ queue,             \
    uint32_t nitems, uint32_t W, uint32_t cbits, uint32_t bcap
#define PCG_SIEVE_ARGS                                                                         \
    queries, items, groups, binfo_all, primes, factor, pi_bits, pi_cnt, T_S, esc, S_acc,     \
    counters, queue, nitems, W, cbits, bcap

template <typename FT, int NQ>
__global__ __launch_bounds__(kBlock) void hsl_sieve_kernel(PCG_SIEVE_PARAMS)
{
  sieve_segments<FT, NQ>(PCG_SIEVE_ARGS);
}

/// The single-query kernel compiled for at least 6 waves per SIMD (at most
/// 80 VGPRs), so segments small enough for LDS to allow more than 20 waves
/// per CU get them.  A schedule choice (HslArgs::waves); results are the
/// same.
template <typename FT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6)))
void hsl_sieve_kernel_w6(PCG_SIEVE_PARAMS)
{
  sieve_segments<FT, 1>(PCG_SIEVE_ARGS);
}

/// The kernel for groups of 2 queries compiled for at least 5 waves per SIMD
/// (at most 96 VGPRs; batched leaf counting takes it to 98 otherwise).
template <typename FT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
void hsl_sieve_kernel_g2(PCG_SIEVE_PARAMS)
{
  sieve_segments<FT, 2>(PCG_SIEVE_ARGS);
}

/// The single-query kernel with phase timers (HslArgs::timers): the same
/// sums, plus each wave's clock cycles per phase in counters[16 ..].
template <typename FT>
__global__ __launch_bounds__(kBlock) void hsl_sieve_kernel_timed(PCG_SIEVE_PARAMS)
{
  sieve_segments<FT, 1, true>(PCG_SIEVE_ARGS);
}

#undef PCG_SIEVE_PARAMS
#undef PCG_SIEVE_ARGS

template <typename FT, int NQ>
static hipError_t launch_sieve_t(const HslArgs& a, uint32_t grid, hipStream_t st)
{
#define PCG_SIEVE_LAUNCH(K)                                                                     \
  hipLaunchKernelGGL(K, dim3(grid), dim3(kBlock), hsl_lds_bytes(a.W), st, a.queries, a.items,  \
                     a.groups, a.binfo, a.primes, reinterpret_cast<const FT*>(a.factor),        \
                     a.pi_bits, a.pi_cnt, a.T_S, a.esc, a.S_acc, a.counters, a.queue, a.nitems, \
                     a.W, a.cbits, a.bcap)
  if (NQ == 1 && a.timers)
    PCG_SIEVE_LAUNCH((hsl_sieve_kernel_timed<FT>));
  else if (NQ == 1 && a.waves == 6)
    PCG_SIEVE_LAUNCH((hsl_sieve_kernel_w6<FT>));
  else if (NQ == 2)
    PCG_SIEVE_LAUNCH((hsl_sieve_kernel_g2<FT>));
  else
    PCG_SIEVE_LAUNCH((hsl_sieve_kernel<FT, NQ>));
#undef PCG_SIEVE_LAUNCH
  return hipGetLastError();
}

/// Zero-fill (used instead of hipMemsetAsync so the sequence can be captured).
__global__ void hsl_fill_zero(uint4* p, size_t n16)
{
  size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for (; i < n16; i += stride)
    p[i] = make_uint4(0, 0, 0, 0);
}

/// Number of b with reach > s: reach does not increase, so this is the first
/// index whose reach is <= s.
__device__ __forceinline__ uint32_t reach_count(const uint32_t* reach, uint32_t nb, uint32_t s)
{
  uint32_t lo = 0, hi = nb;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (reach[mid] > s)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/// Work items of one pass from its spans: item i is segment s0 + (i -
/// item_base) of its span's query, with nbs and slot offset from the reach
/// table (see QueryDesc).
__global__ void hsl_build_items(const QueryDesc* __restrict__ queries,
                                const PassSpan* __restrict__ spans, uint32_t nspans,
                                const uint32_t* __restrict__ reach_all,
                                const uint64_t* __restrict__ sfx_all, SegItem* __restrict__ items,
                                uint32_t nitems)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nitems)
    return;
  uint32_t lo = 0, hi = nspans - 1;               // last span with item_base <= i
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (spans[mid].item_base <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  const PassSpan sp = spans[lo];
  const QueryDesc q = queries[sp.query];
  const uint32_t nb = q.b_last - q.b_first + 1;
  const uint32_t seg = sp.s0 + (i - sp.item_base);
  const uint32_t nbs = reach_count(reach_all + q.reach_off, nb, seg);
  SegItem it;
  it.query = sp.query;
  it.seg = seg;
  it.nbs = nbs;
  it.pad = 0;
  it.t_off = sp.t_base + (uint64_t) seg * nbs + sfx_all[q.sfx_off + nbs];
  it.pad1 = 0;
  items[i] = it;
}

/// Column = one (query, b); slice = a range of the column's segments.  For
/// each column and slice: the slice's count total, sign total and the sum of
/// M(s, b) * (counts of the earlier segments inside the slice).
__global__ void hsl_combine_slices(const CombineCol* __restrict__ cols, uint32_t ncols,
                                   const SegItem* __restrict__ items,
                                   const uint32_t* __restrict__ T_S, EscTable esc, uint32_t cbits,
                                   uint32_t slice_len, uint32_t nslices,
                                   long long* sC, long long* sM, __int128* sMP)
{
  const uint32_t cmask = (1u << cbits) - 1;
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t sl = blockIdx.y;
  if (c >= ncols)
    return;
  const CombineCol cc = cols[c];
  uint32_t i0 = sl * slice_len;
  uint32_t i1 = min(cc.nitems, i0 + slice_len);
  long long P = 0, Ms = 0;
  __int128 MP = 0;
  uint32_t i = i0;
  // 8 items per step: the 16 loads of a step are independent and in flight
  // together
  for (; i + 8 <= i1; i += 8) {
    uint32_t v[8];
    uint64_t off[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      off[u] = items[cc.item_first + i + u].t_off + cc.b_rel;
      v[u] = T_S[off[u]];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      int tm = (int) v[u] >> cbits;
      if ((v[u] & ~cmask) == 0x80000000u)         // escape marker
        tm = esc_lookup(esc, off[u]);
      MP += (__int128) tm * P;
      P += v[u] & cmask;
      Ms += tm;
    }
  }
  for (; i < i1; i++) {
    const uint64_t off = items[cc.item_first + i].t_off + cc.b_rel;
    const uint32_t v = T_S[off];
    int tm = (int) v >> cbits;
    if ((v & ~cmask) == 0x80000000u)
      tm = esc_lookup(esc, off);
    MP += (__int128) tm * P;
    P += v & cmask;
    Ms += tm;
  }
  size_t o = (size_t) sl * ncols + c;
  sC[o] = P;
  sM[o] = Ms;
  sMP[o] = MP;
}

/// Adds the column's carry (counts of all earlier segments, including those
/// of earlier passes) and accumulates into out[col].
__global__ void hsl_combine_final(const CombineCol* __restrict__ cols, uint32_t ncols,
                                  uint32_t nslices, const long long* sC, const long long* sM,
                                  const __int128* sMP, long long* carry, __int128* out)
{
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols)
    return;
  const uint32_t g = cols[c].col;
  long long cr = carry[g];
  __int128 acc = 0;
  for (uint32_t sl = 0; sl < nslices; sl++) {
    size_t o = (size_t) sl * ncols + c;
    acc += sMP[o] + (__int128) cr * sM[o];
    cr += sC[o];
  }
  carry[g] = cr;
  out[g] += acc;
}

/// Factor table built on the device (FactorTableD for D, the lpf FactorTable
/// for S2_hard; src/gourdon/FactorTableD.hpp:69-260, src/FactorTable.hpp:
/// 66-219): one workgroup factors the numbers coprime to 2310 in
/// kFtBlocks consecutive blocks of 2310 numbers.  For every prime p >= 13 with
/// p^2 below the end of the range (in order, one barrier apart, so that the
/// updates of one prime never meet), the threads walk the multiples p q with q
/// >= 2 coprime to 2310 and record the smallest and largest prime index, the
/// number of prime factors, the cofactor and whether p^2 divides.  A number
/// with no recorded factor is 1 or a prime; a cofactor above 1 is the one
/// prime factor above the square root.  Entries: D (kD): 1 -> T_MAX - 1, a
/// prime <= y -> T_MAX, not square-free or a prime factor > y -> 0, else
/// 2 pi(lpf) + [mu = -1]; lpf: 1 -> T_MAX - 1, a prime -> T_MAX, not
/// square-free -> 0, else lpf - [mu = +1].
constexpr uint32_t kFtBlocks = 8;

template <typename T, bool kD>
__global__ __launch_bounds__(256) void hsl_factor_table_kernel(T* __restrict__ out, uint64_t limit,
                                                               uint64_t y,
                                                               const uint32_t* __restrict__ primes,
                                                               uint32_t nprimes)
{
  constexpr uint32_t NE = kFtBlocks * 480;
  __shared__ uint32_t rem[NE];
  __shared__ uint16_t lpfi[NE], mpfi[NE];
  __shared__ uint8_t nf[NE], sqf[NE];
  __shared__ uint16_t cop[480];
  __shared__ int16_t copi[2310];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < 480; i += 256)
    cop[i] = c_cop[i];
  for (uint32_t i = t; i < 2310; i += 256)
    copi[i] = c_cop_index[i];
  const uint64_t q0 = (uint64_t) blockIdx.x * kFtBlocks;     // first 2310-block
  const uint64_t n_lo = q0 * 2310;
  const uint64_t n_hi = min((q0 + kFtBlocks) * 2310, limit + 1);   // [n_lo, n_hi)
  const uint64_t i0 = q0 * 480;
  __syncthreads();
  for (uint32_t e = t; e < NE; e += 256) {
    rem[e] = (uint32_t) (2310 * (q0 + e / 480) + cop[e % 480]);
    lpfi[e] = 0;
    mpfi[e] = 0;
    nf[e] = 0;
    sqf[e] = 1;
  }
  __syncthreads();
  for (uint32_t j = 6; j < nprimes; j++) {
    const uint64_t p = primes[j];
    if (p * p >= n_hi)
      break;
    // cofactors q >= max(2, ceil(n_lo / p)) coprime to 2310 with p q < n_hi;
    // q = 1 is left out so that a prime keeps no recorded factor
    const uint64_t qmin = max<uint64_t>(2, (n_lo + p - 1) / p);
    const uint64_t qmax = (n_hi - 1) / p;
    if (qmax >= qmin) {
      const uint64_t qi_lo = 480 * ((qmin - 1) / 2310) + copi[(qmin - 1) % 2310] + 1;
      const uint64_t qi_hi = 480 * (qmax / 2310) + copi[qmax % 2310];
      for (uint64_t qi = qi_lo + t; qi <= qi_hi; qi += 256) {
        const uint64_t q = 2310 * (qi / 480) + cop[qi % 480];
        const uint64_t m = p * q;
        const uint32_t e = (uint32_t) (480 * (m / 2310) + copi[m % 2310] - i0);
        if (lpfi[e] == 0)
          lpfi[e] = (uint16_t) j;
        mpfi[e] = (uint16_t) j;
        nf[e]++;
        const uint32_t r = rem[e] / (uint32_t) p;
        rem[e] = r;
        if (r % (uint32_t) p == 0)
          sqf[e] = 0;
      }
    }
    __syncthreads();
  }
  const T T_MAX = (T) ~(T) 0;
  for (uint32_t e = t; e < NE; e += 256) {
    const uint64_t n = 2310 * (q0 + e / 480) + cop[e % 480];
    if (n > limit)
      continue;
    const uint32_t r = rem[e];
    const uint32_t nfac = nf[e] + (r > 1 ? 1u : 0u);
    T v;
    if (n == 1)
      v = (T) (T_MAX - 1);
    else if (lpfi[e] == 0)
      v = kD ? (n <= y ? T_MAX : (T) 0) : T_MAX;          // n is prime
    else if (!sqf[e])
      v = 0;
    else if (kD) {
      const uint64_t mpf = r > 1 ? r : primes[mpfi[e]];
      v = mpf > y ? (T) 0 : (T) (2u * lpfi[e] + (nfac & 1));
    } else {
      const uint32_t lpf = primes[lpfi[e]];
      v = (T) ((nfac & 1) ? lpf : lpf - 1);
    }
    out[i0 + e] = v;
  }
}

// ------------------------------------------------------------------------
// host-side launch wrappers

hipError_t hsl_launch_factor_table(void* out, bool factor32, bool d_table, uint64_t limit,
                                   uint64_t y, const uint32_t* primes, uint32_t nprimes,
                                   hipStream_t st)
{
  const uint64_t nblk = (limit / 2310 + kFtBlocks) / kFtBlocks;
  const dim3 g((uint32_t) nblk), b(256);
  if (factor32) {
    if (d_table)
      hipLaunchKernelGGL((hsl_factor_table_kernel<uint32_t, true>), g, b, 0, st, (uint32_t*) out,
                         limit, y, primes, nprimes);
    else
      hipLaunchKernelGGL((hsl_factor_table_kernel<uint32_t, false>), g, b, 0, st, (uint32_t*) out,
                         limit, y, primes, nprimes);
  } else {
    if (d_table)
      hipLaunchKernelGGL((hsl_factor_table_kernel<uint16_t, true>), g, b, 0, st, (uint16_t*) out,
                         limit, y, primes, nprimes);
    else
      hipLaunchKernelGGL((hsl_factor_table_kernel<uint16_t, false>), g, b, 0, st, (uint16_t*) out,
                         limit, y, primes, nprimes);
  }
  return hipGetLastError();
}

hipError_t hsl_scratch_address(int device, void** ptr)
{
  // one address per device: each GPU holds its own copy of the array
  static void* addr[kHslMaxDevices] = {};
  if (device < 0 || device >= kHslMaxDevices) return hipErrorInvalidDevice;
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

hipError_t hsl_upload_constants()
{
  int16_t ci[2310];
  uint16_t cp[480];
  int idx = -1, k = 0;
  for (int n = 0; n < 2310; n++) {
    if (n % 2 && n % 3 && n % 5 && n % 7 && n % 11) {
      idx++;
      cp[k++] = (uint16_t) n;
    }
    ci[n] = (int16_t) idx;
  }
  uint16_t cx[544];
  for (int i = 0; i < 544; i++) cx[i] = (uint16_t) (cp[i % 480] + (i >= 480 ? 2310 : 0));
  hipError_t e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_cop_index), ci, sizeof(ci))) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_cop_ext), cx, sizeof(cx))) != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_cop), cp, sizeof(cp));
}

hipError_t hsl_launch_sieve(const HslArgs& a, bool factor32, uint32_t grid, hipStream_t st)
{
  if (a.groups && a.group_max > 2)
    return factor32 ? launch_sieve_t<uint32_t, 4>(a, grid, st)
                    : launch_sieve_t<uint16_t, 4>(a, grid, st);
  if (a.groups)
    return factor32 ? launch_sieve_t<uint32_t, 2>(a, grid, st)
                    : launch_sieve_t<uint16_t, 2>(a, grid, st);
  return factor32 ? launch_sieve_t<uint32_t, 1>(a, grid, st) : launch_sieve_t<uint16_t, 1>(a, grid, st);
}

hipError_t hsl_sieve_occupancy(bool factor32, uint32_t W, uint32_t waves, int* blocks_per_cu)
{
  size_t lds = hsl_lds_bytes(W);
  if (waves == 6)
    return factor32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          blocks_per_cu, hsl_sieve_kernel_w6<uint32_t>, kBlock, lds)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          blocks_per_cu, hsl_sieve_kernel_w6<uint16_t>, kBlock, lds);
  if (factor32)
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, hsl_sieve_kernel<uint32_t, 1>,
                                                        kBlock, lds);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, hsl_sieve_kernel<uint16_t, 1>,
                                                      kBlock, lds);
}

hipError_t hsl_launch_fill_zero(void* p, size_t bytes, hipStream_t st)
{
  size_t n16 = (bytes + 15) / 16;
  if (n16 == 0) return hipSuccess;
  uint32_t grid = (uint32_t) std::min<size_t>((n16 + 255) / 256, 4096);
  hipLaunchKernelGGL(hsl_fill_zero, dim3(grid), dim3(256), 0, st, (uint4*) p, n16);
  return hipGetLastError();
}

hipError_t hsl_launch_build_items(const QueryDesc* queries, const PassSpan* spans, uint32_t nspans,
                                  const uint32_t* reach, const uint64_t* sfx, SegItem* items,
                                  uint32_t nitems, hipStream_t st)
{
  if (nitems == 0 || nspans == 0) return hipSuccess;
  hipLaunchKernelGGL(hsl_build_items, dim3((nitems + 255) / 256), dim3(256), 0, st, queries,
                     spans, nspans, reach, sfx, items, nitems);
  return hipGetLastError();
}

hipError_t hsl_launch_combine(const CombineCol* cols, uint32_t ncols, const SegItem* items,
                              const uint32_t* T_S, const EscTable& esc, uint32_t cbits,
                              uint32_t slice_len, uint32_t nslices, long long* sC, long long* sM,
                              void* sMP, long long* carry, void* out, hipStream_t st)
{
  if (ncols == 0) return hipSuccess;
  dim3 g1((ncols + 255) / 256, nslices);
  hipLaunchKernelGGL(hsl_combine_slices, g1, dim3(256), 0, st, cols, ncols, items, T_S, esc,
                     cbits, slice_len, nslices, sC, sM, (__int128*) sMP);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(hsl_combine_final, dim3((ncols + 255) / 256), dim3(256), 0, st, cols, ncols,
                     nslices, sC, sM, (const __int128*) sMP, carry, (__int128*) out);
  return hipGetLastError();
}

} // namespace pcg
