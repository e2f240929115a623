// pix_kernels.hpp -- device side of the other terms of Gourdon's prime
// counting function pi(x) = A - B + C + D + Phi0 + Sigma
// (src/gourdon/pi_gourdon.cpp:82-87); D is the hard-special-leaf engine.
//
// All of them read pi(n) from a prime-count table resident in HBM: for every
// 64-bit word w (numbers [240 w, 240 w + 240)) one 16-byte PixWord with the
// wheel-30 bits of the primes >= 7 and the count of primes >= 7 below 240 w,
// so that pi(n) = 3 + t[n / 240].cnt + popcount(t[n / 240].bits & mask(n % 240)),
// the lookup of src/PiTable.hpp:46-57 with 64-bit counts; bits and count sit
// side by side so that a lookup reads one 16-byte word.
#pragma once

#include <stdint.h>
#include <hip/hip_runtime_api.h>

namespace pcg {

struct alignas(16) PixWord {
  uint64_t bits;     // primes >= 7 in [240 w, 240 w + 240), wheel-30 bits
  uint64_t cnt;      // primes >= 7 below 240 w
};

/// Leaves of one sieving prime for the flat leaf-sum kernel: terms
/// w * pi(xp / primes[i]) for i = i_lo .. i_lo + (next off - off) - 1, with
/// w = 1 if primes[i] <= q1 and 2 otherwise.  Every quotient xp / primes[i]
/// is below 2^49 (it is covered by the prime-count table).
struct alignas(16) PixLeafDesc {
  uint64_t xp;
  double xd;         // xp rounded to double
  uint64_t off;      // first flat term index of this descriptor
  uint32_t i_lo;
  uint32_t q1;       // primes are below 2^32: q1 = min(bound, 2^32 - 1)
};

/// Leaves of one sieving prime for the tiled leaf-sum kernel: terms
/// w * pi(xp / primes[i]) for i_lo <= i <= i_hi, w = 1 if primes[i] <= q1 and
/// 2 otherwise; every quotient is at least 7 and below 2^49, and every
/// prime below 2^30 (kPixTileMaxPrime).
struct alignas(16) PixTileDesc {
  uint64_t xp;
  double xd;         // xp rounded to double
  uint32_t i_lo, i_hi;
  uint32_t q1;
  uint32_t pad;
};

/// kPixTileGroup consecutive descriptors d0 .. d0 + kPixTileGroup - 1 (nd of
/// them real, the rest empty: i_lo = 1, i_hi = 0) and the
/// tiles (64 consecutive prime indices: tile t holds i = 64 t .. 64 t + 63)
/// their ranges cover, t_lo .. t_hi, cut into items of kPixTileChunk tiles;
/// the group's items are item0, item0 + 1, ...
struct alignas(16) PixTileGroup {
  uint64_t item0;
  uint32_t d0, nd;
  uint32_t t_lo, t_hi;
};

static_assert(sizeof(PixTileDesc) == 32 && sizeof(PixTileGroup) == 32, "32-byte records");

constexpr uint32_t kPixTileGroup = 8;
constexpr uint32_t kPixTileMaxPrime = 1u << 30;
constexpr uint64_t kPixTileMaxQuotient = 1ull << 43;
constexpr uint32_t kPixTileChunk = 16;

constexpr uint32_t kPixSegWords = 4096;       // words per sieve segment (983,040 numbers)

/// The table words [wbase, wbase + nwords) (numbers up to N) into tab[0 ..
/// nwords): bits and per-segment totals; then the scan and the counts, whose
/// base is the number of primes >= 7 below word wbase.
hipError_t pix_launch_sieve(PixWord* tab, uint64_t* seg_tot, const uint32_t* primes,
                            uint32_t j_small_end, uint32_t j_end, uint64_t nwords, uint64_t N,
                            uint64_t wbase, hipStream_t st);
hipError_t pix_launch_scan(uint64_t* seg_tot, uint32_t nseg, hipStream_t st);
hipError_t pix_launch_counts(PixWord* tab, const uint64_t* seg_base, uint64_t nwords,
                             uint64_t base, hipStream_t st);
/// The leaf sums and B read pi from a window of the table: tab holds the
/// words [w_lo, w_hi) with absolute counts (the whole table: 0 and
/// UINT64_MAX); terms whose quotient lies outside add nothing.
hipError_t pix_launch_leafsum(const PixLeafDesc* desc, uint32_t ndesc, uint64_t total,
                              const uint32_t* primes, const PixWord* tab, uint64_t w_lo,
                              uint64_t w_hi, unsigned long long* acc, uint32_t cus, hipStream_t st);
/// max_quotient bounds xp / primes[i] over all terms (the table limit).
hipError_t pix_launch_leaftiles(const PixTileDesc* desc, const PixTileGroup* grp, uint32_t ngrp,
                                uint64_t nitems, const uint32_t* primes, uint32_t nprimes,
                                const PixWord* tab, uint64_t w_lo, uint64_t w_hi,
                                uint64_t max_quotient, unsigned long long* acc, uint32_t cus,
                                hipStream_t st);
/// B's primes come from the bits of ptab, the table from word 0 up to sqrtx.
hipError_t pix_launch_B(uint64_t x_lo, uint64_t x_hi, uint64_t y, uint64_t sqrtx,
                        const PixWord* ptab, const PixWord* tab, uint64_t w_lo, uint64_t w_hi,
                        unsigned long long* acc, hipStream_t st);
hipError_t pix_launch_sigma(uint64_t x_lo, uint64_t x_hi, uint64_t y, uint64_t sqrt_xy,
                            const uint32_t* primes, uint32_t b_lo, uint32_t b_hi,
                            const PixWord* tab, unsigned long long* acc, hipStream_t st);
/// lpf_table: the factor table has the lpf encoding (S1 of Deleglise-Rivat)
/// instead of FactorTableD's (Phi0 of Gourdon).
hipError_t pix_launch_phi0(uint64_t x_lo, uint64_t x_hi, const void* factor, bool factor32,
                           bool lpf_table, uint64_t nidx, uint32_t k, uint32_t pk,
                           const uint32_t* phi_tbl, uint32_t pp, uint64_t phi_pp,
                           unsigned long long* acc, hipStream_t st);
hipError_t pix_launch_extract(const PixWord* tab, uint64_t p_max, uint32_t* out, hipStream_t st);
hipError_t pix_launch_interval(uint64_t* out, uint64_t lo, uint64_t hi, const uint32_t* sp,
                               uint32_t n_small, uint64_t n_all, hipStream_t st);
/// The primes 7 <= p <= p_max as 64-bit values.
hipError_t pix_launch_extract64(const PixWord* tab, uint64_t p_max, uint64_t* out, hipStream_t st);
/// Sieve of [L + o_lo, L + o_hi] for L = L_hi 2^64 + L_lo above 2^63, a
/// multiple of 240 (L_hi < 2^20, o_hi < 2^62): word w of out holds the
/// numbers [L + 240 w, L + 240 w + 240); sp are the sieving primes from 7
/// up to sqrt(L + o_hi), the first n_small sieved in LDS.
hipError_t pix_launch_interval128(uint64_t* out, uint64_t L_hi, uint64_t L_lo, uint64_t o_lo,
                                  uint64_t o_hi, const uint64_t* sp, uint32_t n_small,
                                  uint64_t n_all, hipStream_t st);
hipError_t pix_launch_block_counts(const uint64_t* bits, uint64_t nwords, uint64_t* counts,
                                   hipStream_t st);
hipError_t pix_launch_select(const PixWord* tab, uint64_t nwords, uint64_t n, uint64_t* out,
                             hipStream_t st);
hipError_t pix_launch_pi_points(const uint64_t* n, uint64_t* out, uint32_t count,
                                const PixWord* tab, hipStream_t st);

} // namespace pcg
