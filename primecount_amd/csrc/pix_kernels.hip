// pix_kernels.hip -- CDNA4 (gfx950) kernels for the terms of Gourdon's pi(x)
// other than D: the prime-count table, the easy leaves of A and C, B, Phi0
// and the Sigma sums.  See pix_kernels.hpp for the table layout.

#include <hip/hip_runtime.h>

#include "pix_kernels.hpp"
#include "wheel_dev.hpp"

#include <algorithm>
#include <type_traits>

namespace pcg {

namespace {

constexpr uint32_t kThreads = 256;
typedef unsigned __int128 u128;
typedef __int128 i128;

/// pi(n) from the resident table.
__device__ __forceinline__ uint64_t pix_pi(const PixWord* __restrict__ tab, uint64_t n)
{
  if (n < 7)
    return n < 2 ? 0 : (n < 3 ? 1 : (n < 5 ? 2 : 3));
  const uint64_t w = n / 240;
  const uint32_t r = (uint32_t) (n - w * 240);
  const PixWord t = tab[w];
  return 3 + t.cnt + (uint64_t) __popcll(t.bits & le_mask(r));
}

/// pi(n) from a window of the table that holds the words [w_lo, w_hi) (its
/// counts are absolute); 0 for an n outside the window, whose lookup
/// another window answers.
__device__ __forceinline__ uint64_t pix_pi_win(const PixWord* __restrict__ tab, uint64_t n,
                                               uint64_t w_lo, uint64_t w_hi)
{
  const uint64_t w = n / 240;
  if (w - w_lo >= w_hi - w_lo)
    return 0;
  if (n < 7)
    return n < 2 ? 0 : (n < 3 ? 1 : (n < 5 ? 2 : 3));
  const uint32_t r = (uint32_t) (n - w * 240);
  const PixWord t = tab[w - w_lo];
  return 3 + t.cnt + (uint64_t) __popcll(t.bits & le_mask(r));
}

/// Adds a signed 128-bit value to a (lo, hi) pair with two 64-bit atomics; the
/// add whose low word wraps carries into the high word.
__device__ __forceinline__ void atomic_add_i128(unsigned long long* acc, i128 v)
{
  const unsigned long long lo = (unsigned long long) v;
  const unsigned long long hi = (unsigned long long) (v >> 64);
  const unsigned long long old = atomicAdd(&acc[0], lo);
  const unsigned long long h = hi + (old + lo < old ? 1ull : 0ull);
  if (h != 0)
    atomicAdd(&acc[1], h);
}

/// Block-wide sum of a signed 128-bit value, added to acc by thread 0.
__device__ __forceinline__ void block_add_i128(unsigned long long* acc, i128 v)
{
  __shared__ i128 red[kThreads];
  red[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != 0)
    atomic_add_i128(acc, red[0]);
}

/// floor(v / d) and v mod d for v < 2^128, 0 < d < 2^32 (32-bit limbs).
__device__ __forceinline__ u128 divmod128_32(u128 v, uint32_t d, uint32_t* rem)
{
  u128 q = 0;
  uint64_t r = 0;
  for (int s = 96; s >= 0; s -= 32) {
    const uint64_t cur = (r << 32) | (uint32_t) (v >> s);
    const uint64_t ql = cur / d;
    r = cur - ql * d;
    q |= (u128) ql << s;
  }
  *rem = (uint32_t) r;
  return q;
}

/// floor(sqrt(n)) for n < 2^64.
__device__ __forceinline__ uint64_t isqrt64(uint64_t n)
{
  uint64_t r = (uint64_t) sqrt((double) n);
  while ((u128) r * r > n)
    r--;
  while ((u128) (r + 1) * (r + 1) <= n)
    r++;
  return r;
}

// ---------------------------------------------------------------------------
// prime-count table

/// One workgroup sieves one segment of kPixSegWords words in LDS: the numbers
/// coprime to 30 in [low, high) (high <= N + 1), without 1, then the multiples
/// p * q (q >= p, q coprime to 30) of the primes p_4 = 7 .. p_{j_end - 1}.
/// The table (or window of it) starts at word wbase: its word w stands for
/// the numbers [240 (wbase + w), 240 (wbase + w) + 240).
/// Small primes are crossed off by all threads together (8 residue classes
/// of q, 32 threads per class), the others by one thread each.  Writes the
/// bits and the segment's prime count.
__global__ __launch_bounds__(kThreads) void pix_sieve_kernel(
    PixWord* __restrict__ tab, uint64_t* __restrict__ seg_tot,
    const uint32_t* __restrict__ primes, uint32_t j_small_end, uint32_t j_end,
    uint64_t nwords, uint64_t N, uint64_t wbase)
{
  __shared__ uint32_t bm[2 * kPixSegWords];
  __shared__ unsigned long long red[kThreads];
  const uint32_t t = threadIdx.x;
  const uint64_t w0 = (uint64_t) blockIdx.x * kPixSegWords;
  const uint32_t nw = (uint32_t) min<uint64_t>(kPixSegWords, nwords - w0);
  const uint64_t low = (wbase + w0) * 240;
  const uint64_t high = min(low + 240ull * nw, N + 1);
  for (uint32_t w = t; w < nw; w += kThreads) {
    const uint64_t base = low + 240ull * w;
    uint64_t v = base + 240 <= high ? ~0ull : (base >= high ? 0ull : le_mask((uint32_t) (high - base - 1)));
    if (base == 0)
      v &= ~1ull;                                    // 1 is not prime
    bm[2 * w] = (uint32_t) v;
    bm[2 * w + 1] = (uint32_t) (v >> 32);
  }
  __syncthreads();
  if (high > low) {
    const uint32_t cls = t & 7;
    const uint32_t t0 = t >> 3;
    for (uint32_t j = 4; j < j_small_end; j++) {
      const uint64_t p = primes[j];
      const uint64_t q_lo = max(p, div_exact((double) (low + p - 1), low + p - 1, p));
      const uint64_t q_hi = div_exact((double) (high - 1), high - 1, p);
      if (q_lo > q_hi)
        continue;
      uint64_t q = q_lo - q_lo % 30 + ((kWheelBytes >> (8 * cls)) & 0xff);
      if (q < q_lo)
        q += 30;
      if (q > q_hi)
        continue;
      const uint32_t n = (uint32_t) ((q_hi - q) / 30) + 1;
      const uint32_t off = (uint32_t) (p * q - low);
      const uint32_t byte0 = off / 30;
      const uint32_t bib = wheel_bit(off - byte0 * 30);
      for (uint32_t k = t0; k < n; k += kThreads / 8) {
        const uint32_t B = byte0 + k * (uint32_t) p;
        atomicAnd(&bm[B >> 2], ~(1u << (((B & 3) << 3) | bib)));
      }
    }
    for (uint32_t j = j_small_end + t; j < j_end; j += kThreads) {
      const uint64_t p = primes[j];
      const uint64_t q_lo = max(p, div_exact((double) (low + p - 1), low + p - 1, p));
      const uint64_t q_hi = div_exact((double) (high - 1), high - 1, p);
      if (q_lo > q_hi)
        continue;
      // walk the cofactors coprime to 30 in [q_lo, q_hi] by wheel index
      const uint32_t qm = (uint32_t) (q_lo % 30);
      const uint64_t qbase = q_lo - qm;
      for (uint32_t idx = qm ? le_count(qm - 1) : 0u;; idx++) {
        const uint64_t q = qbase + 30ull * (idx >> 3) + ((kWheelBytes >> (8 * (idx & 7))) & 0xff);
        if (q > q_hi)
          break;
        const uint32_t off = (uint32_t) (p * q - low);
        const uint32_t B = off / 30;
        atomicAnd(&bm[B >> 2], ~(1u << (((B & 3) << 3) | wheel_bit(off - B * 30))));
      }
    }
  }
  __syncthreads();
  unsigned long long tot = 0;
  for (uint32_t w = t; w < nw; w += kThreads) {
    const uint64_t v = ((uint64_t) bm[2 * w + 1] << 32) | bm[2 * w];
    tab[w0 + w].bits = v;
    tot += (unsigned long long) __popcll(v);
  }
  red[t] = tot;
  __syncthreads();
  for (uint32_t s = kThreads / 2; s > 0; s >>= 1) {
    if (t < s)
      red[t] += red[t + s];
    __syncthreads();
  }
  if (t == 0)
    seg_tot[blockIdx.x] = red[0];
}

/// Exclusive prefix sums of the segment counts, in place (one workgroup).
__global__ __launch_bounds__(1024) void pix_scan_kernel(uint64_t* __restrict__ seg, uint32_t nseg)
{
  __shared__ unsigned long long part[1024];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (nseg + 1023) / 1024;
  const uint32_t b0 = min(nseg, t * per), b1 = min(nseg, b0 + per);
  unsigned long long s = 0;
  for (uint32_t i = b0; i < b1; i++)
    s += seg[i];
  part[t] = s;
  __syncthreads();
  for (uint32_t d = 1; d < 1024; d <<= 1) {
    const unsigned long long v = t >= d ? part[t - d] : 0ull;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  unsigned long long run = part[t] - s;
  for (uint32_t i = b0; i < b1; i++) {
    const unsigned long long v = seg[i];
    seg[i] = run;
    run += v;
  }
}

/// tab[w].cnt = primes >= 7 below the word: base (those below the table's
/// first word) plus the segment base plus the prefix of the word popcounts
/// inside the segment (16 words per thread, one block scan).
__global__ __launch_bounds__(kThreads) void pix_count_kernel(PixWord* __restrict__ tab,
                                                             const uint64_t* __restrict__ seg_base,
                                                             uint64_t nwords, uint64_t base)
{
  __shared__ unsigned long long part[kThreads];
  const uint32_t t = threadIdx.x;
  const uint64_t w0 = (uint64_t) blockIdx.x * kPixSegWords;
  const uint32_t nw = (uint32_t) min<uint64_t>(kPixSegWords, nwords - w0);
  constexpr uint32_t per = kPixSegWords / kThreads;     // 16
  const uint32_t a = t * per;
  unsigned long long s = 0;
  for (uint32_t i = 0; i < per; i++)
    if (a + i < nw)
      s += (unsigned long long) __popcll(tab[w0 + a + i].bits);
  part[t] = s;
  __syncthreads();
  for (uint32_t d = 1; d < kThreads; d <<= 1) {
    const unsigned long long v = t >= d ? part[t - d] : 0ull;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  unsigned long long run = base + seg_base[blockIdx.x] + part[t] - s;
  for (uint32_t i = 0; i < per; i++) {
    if (a + i < nw) {
      const uint64_t b = tab[w0 + a + i].bits;
      tab[w0 + a + i].cnt = run;
      run += (unsigned long long) __popcll(b);
    }
  }
}

// ---------------------------------------------------------------------------
// leaf sums

/// floor(a / d) for 1 <= d < 2^32 and a quotient below 2^49; ad is a rounded
/// to double.  The reciprocal estimate gets two Newton steps (relative error
/// then near 2^-52, so the truncated product is within one of the quotient);
/// the residual a - q d, exact modulo 2^64 because |a - q d| < 2 d, fixes it.
/// qd receives the quotient as a double.
__device__ __forceinline__ uint64_t div49(uint64_t a, double ad, uint32_t d, double* qd)
{
  const double dd = (double) d;
  double r = __builtin_amdgcn_rcp(dd);
  r = __builtin_fma(r, __builtin_fma(-dd, r, 1.0), r);
  r = __builtin_fma(r, __builtin_fma(-dd, r, 1.0), r);
  const double qf = __builtin_trunc(ad * r);
  // the integer in the low 52 bits of qf + 2^52
  const uint64_t q = __builtin_bit_cast(uint64_t, qf + 4503599627370496.0) & ((1ull << 52) - 1);
  const int64_t rem = (int64_t) (a - q * d);
  const int32_t fix = rem < 0 ? -1 : ((uint64_t) rem >= d ? 1 : 0);
  *qd = qf + (double) fix;
  return q + (int64_t) fix;
}

/// Sum over the flat terms of all descriptors of w * pi(xp / primes[i]).
/// Each wave takes one contiguous range of terms: a lane finds the
/// descriptor of its first term by binary search and then steps forward by
/// 64 terms.  A lane works on kGroup terms at a time so that their prime
/// loads, and then their table loads, are in flight together; the masks of
/// the table words sit in LDS.  With WIN, tab holds the table words
/// [w_lo, w_hi): a term whose quotient lies outside adds nothing (another
/// window counts it); without, tab is the whole table.
template <bool WIN>
__global__ __launch_bounds__(kThreads) void pix_leafsum_kernel(
    const PixLeafDesc* __restrict__ desc, uint32_t ndesc, uint64_t total, uint64_t per_wave,
    const uint32_t* __restrict__ primes, const PixWord* __restrict__ tab, uint64_t w_lo,
    uint64_t w_hi, unsigned long long* __restrict__ acc)
{
  constexpr uint32_t kGroup = 8;
  __shared__ uint64_t masks[240];
  for (uint32_t r = threadIdx.x; r < 240; r += kThreads)
    masks[r] = le_mask(r);
  __syncthreads();

  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t) blockIdx.x * kThreads + threadIdx.x) >> 6;
  const uint64_t t0 = wave * per_wave;
  const uint64_t t1 = t0 + per_wave < total ? t0 + per_wave : total;
  u128 sum = 0;
  if (t0 < total) {
    // last descriptor with off <= t (clamped to the last term)
    const uint64_t tf = t0 + lane < t1 ? t0 + lane : t1 - 1;
    uint32_t lo = 0, hi = ndesc - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (desc[mid].off <= tf)
        lo = mid;
      else
        hi = mid - 1;
    }
    uint32_t d = lo;
    PixLeafDesc ds = desc[d];
    uint64_t next = d + 1 < ndesc ? desc[d + 1].off : total;
    for (uint64_t tb = t0 + lane; tb < t1 + lane; tb += 64 * kGroup) {
      uint64_t xp[kGroup];
      double xd[kGroup];
      uint32_t idx[kGroup], q1[kGroup], q[kGroup];
#pragma unroll
      for (uint32_t s = 0; s < kGroup; s++) {
        const uint64_t t = tb + 64 * s;
        const bool ok = t < t1;
        if (ok) {
          while (t >= next) {
            d++;
            ds = desc[d];
            next = d + 1 < ndesc ? desc[d + 1].off : total;
          }
        }
        xp[s] = ok ? ds.xp : 0;
        xd[s] = ok ? ds.xd : 0.0;
        q1[s] = ds.q1;
        idx[s] = ok ? ds.i_lo + (uint32_t) (t - ds.off) : ds.i_lo;
      }
#pragma unroll
      for (uint32_t s = 0; s < kGroup; s++)
        q[s] = primes[idx[s]];
      uint64_t w[kGroup];
      PixWord tw[kGroup];
      uint32_t rr[kGroup];
      bool in[kGroup];
#pragma unroll
      for (uint32_t s = 0; s < kGroup; s++) {
        const uint32_t qs = xp[s] ? q[s] : 1u;
        double vd;
        const uint64_t v = div49(xp[s], xd[s], qs, &vd);
        // w = floor(v / 240): (v + 1/2) / 240 is at least 1/480 from an
        // integer, far more than the rounding error for v < 2^49
        const double wf = __builtin_trunc(__builtin_fma(vd, 1.0 / 240.0, 0.5 / 240.0));
        const uint64_t wv = __builtin_bit_cast(uint64_t, wf + 4503599627370496.0) & ((1ull << 52) - 1);
        in[s] = !WIN || wv - w_lo < w_hi - w_lo;
        w[s] = !WIN ? wv : (in[s] ? wv - w_lo : 0);
        rr[s] = (uint32_t) v - 240u * (uint32_t) wv;
        xp[s] = v;
      }
#pragma unroll
      for (uint32_t s = 0; s < kGroup; s++) {
        tw[s] = tab[w[s]];
      }
      uint64_t part = 0;
#pragma unroll
      for (uint32_t s = 0; s < kGroup; s++) {
        const uint64_t v = xp[s];
        // pi(v) for v < 7 from a nibble table: 0, 0, 1, 2, 2, 3, 3
        uint64_t pv = v < 7 ? (uint64_t) ((0x3322100u >> (4 * (uint32_t) v)) & 15)
                            : 3 + tw[s].cnt + (uint64_t) __popcll(tw[s].bits & masks[rr[s]]);
        pv = in[s] ? pv : 0;
        part += q[s] <= q1[s] ? pv : 2 * pv;
      }
      sum += part;
    }
  }
  block_add_i128(acc, (i128) sum);
}

/// Sum of w * (pi(xp / q) - 3) over tiles x descriptors (the host adds the
/// 3 w).  Lane l of a wave holds the prime q = primes[64 t + l] of tile t and
/// its reciprocal (two Newton steps), and runs through the descriptors of
/// the tile's group, which are wave-uniform (scalar loads): the eight
/// quotients, then the eight table loads, then the sums.  Each wave takes
/// one contiguous range of items (group, run of kPixTileChunk tiles),
/// finding its first group by binary search.  Lanes outside a descriptor's
/// range, or (WIN) whose quotient lies outside the window [w_lo, w_hi) of
/// table words that tab holds, read its first word and add nothing.
template <bool WIN>
__global__ __launch_bounds__(kThreads) void pix_leaftile_kernel(
    const PixTileDesc* __restrict__ desc, const PixTileGroup* __restrict__ grp, uint32_t ngrp,
    uint64_t nitems, uint64_t per_wave, const uint32_t* __restrict__ primes, uint32_t nprimes,
    const PixWord* __restrict__ tab, uint64_t w_lo, uint64_t w_hi,
    unsigned long long* __restrict__ acc)
{
  constexpr uint32_t G = kPixTileGroup;
  __shared__ uint64_t masks[240];
  for (uint32_t r = threadIdx.x; r < 240; r += kThreads)
    masks[r] = le_mask(r);
  __syncthreads();

  const uint32_t lane = threadIdx.x & 63;
  // wave-uniform (scalar) from here on: group, item and descriptor data
  const uint64_t wave = (uint64_t) blockIdx.x * (kThreads / 64) +
                        (uint32_t) __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t j0 = wave * per_wave;
  const uint64_t j1 = j0 + per_wave < nitems ? j0 + per_wave : nitems;
  u128 sum = 0;
  if (j0 < nitems) {
    uint32_t lo = 0, hi = ngrp - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (grp[mid].item0 <= j0)
        lo = mid;
      else
        hi = mid - 1;
    }
    uint32_t g = lo;
    for (uint64_t j = j0; j < j1; j++) {
      while (g + 1 < ngrp && grp[g + 1].item0 <= j)
        g++;
      const PixTileGroup gr = grp[g];
      const uint32_t ta = gr.t_lo + (uint32_t) (j - gr.item0) * kPixTileChunk;
      const uint32_t tb = gr.t_hi + 1 < ta + kPixTileChunk ? gr.t_hi + 1 : ta + kPixTileChunk;
      // the tiles whose 64 primes lie in every descriptor's range (most of
      // them) take a body without the per-lane range masks
      uint32_t in_lo = 0, in_hi = 0xFFFFFFFFu;
#pragma unroll
      for (uint32_t k = 0; k < G; k++) {
        const PixTileDesc D = desc[gr.d0 + k];
        in_lo = max(in_lo, D.i_lo);
        in_hi = min(in_hi, D.i_hi);
      }
      in_hi = min(in_hi, nprimes - 1);
      // one tile: the sum of its leaves' w * (pi - 3) over the group
      auto tile = [&](auto full_c, uint32_t t) -> uint64_t {
        constexpr bool FULL = decltype(full_c)::value;
        const uint32_t idx = 64 * t + lane;
        const uint32_t q0 = primes[FULL || idx < nprimes ? idx : nprimes - 1];
        const uint32_t q = q0 > 1 ? q0 : 1u;
        const double qd = (double) q;
        double r = __builtin_amdgcn_rcp(qd);
        r = __builtin_fma(r, __builtin_fma(-qd, r, 1.0), r);
        r = __builtin_fma(r, __builtin_fma(-qd, r, 1.0), r);
        double vd[G];
        uint32_t v_lo[G], sh[G];
        bool ok[G], bad[G];
        bool bad_any = false;
#pragma unroll
        for (uint32_t k = 0; k < G; k++) {
          // groups are padded to G descriptors with empty ranges
          const PixTileDesc D = desc[gr.d0 + k];
          ok[k] = FULL || ((idx >= D.i_lo) & (idx <= D.i_hi));
          sh[k] = q > D.q1 ? 1u : 0u;
          // quotient xp / q.  xd * r is within v 2^-50 of xp / q, and
          // t = xd * r - 1/2 takes one more rounding (below v 2^-53), so
          // vd = rint(t) is floor(xp / q) when |t - vd| < 1/2 - v 2^-46
          // (the fraction of xd * r lies in [v 2^-46, 1 - v 2^-46]; t - vd
          // is exact); otherwise (rare) vd is within one of the quotient
          // and the residual fixes it below.  The low word of vd + 2^52 is
          // the low word of the integer vd.
          const double t = __builtin_fma(D.xd, r, -0.5);
          vd[k] = __builtin_rint(t);
          const double h = t - vd[k];
          const double tol = 0.5 - vd[k] * 0x1p-46;
          v_lo[k] = (uint32_t) __builtin_bit_cast(uint64_t, vd[k] + 4503599627370496.0);
          bad[k] = ok[k] & (__builtin_fabs(h) >= tol);
          bad_any |= bad[k];
        }
        if (__builtin_amdgcn_ballot_w64(bad_any) != 0) {
          // the residual xp - s q lies in (-q, 2 q) for the lanes in range;
          // its low 32 bits give it exactly for q < 2^30
#pragma unroll
          for (uint32_t k = 0; k < G; k++) {
            if (bad[k]) {
              const PixTileDesc D = desc[gr.d0 + k];
              const int32_t rem = (int32_t) ((uint32_t) D.xp - v_lo[k] * q);
              const int32_t fix = rem < 0 ? -1 : ((uint32_t) rem >= q ? 1 : 0);
              vd[k] += (double) fix;
              v_lo[k] += (uint32_t) fix;
            }
          }
        }
        uint64_t w[G];
        uint32_t rr[G];
#pragma unroll
        for (uint32_t k = 0; k < G; k++) {
          const double wf = __builtin_trunc(__builtin_fma(vd[k], 1.0 / 240.0, 0.5 / 240.0));
          const uint64_t wv =
              __builtin_bit_cast(uint64_t, wf + 4503599627370496.0) & ((1ull << 52) - 1);
          if constexpr (WIN)
            ok[k] = ok[k] & (wv - w_lo < w_hi - w_lo);
          w[k] = ok[k] ? (WIN ? wv - w_lo : wv) : 0;
          rr[k] = ok[k] ? v_lo[k] - 240u * (uint32_t) wv : 0u;
        }
        PixWord tw[G];
#pragma unroll
        for (uint32_t k = 0; k < G; k++)
          tw[k] = tab[w[k]];
        uint64_t part = 0;
#pragma unroll
        for (uint32_t k = 0; k < G; k++) {
          const uint64_t pv = tw[k].cnt + (uint64_t) __popcll(tw[k].bits & masks[rr[k]]);
          // without windows a lane out of range read word 0 with the empty mask: 0
          part += (!WIN || ok[k]) ? pv << sh[k] : 0;
        }
        return part;
      };
      for (uint32_t t = ta; t < tb; t++) {
        uint64_t part;
        if (64 * t >= in_lo && 64 * t + 63 <= in_hi)
          part = tile(std::true_type{}, t);
        else
          part = tile(std::false_type{}, t);
        sum += part;
      }
    }
  }
  block_add_i128(acc, (i128) sum);
}

/// B: sum of pi(x / p) over the primes y < p <= isqrt(x), read from the
/// bits of ptab, the table from word 0 (one thread per word); pi(x / p) from
/// tab, the window [w_lo, w_hi) of table words (quotients outside add
/// nothing).
__global__ __launch_bounds__(kThreads) void pix_B_kernel(uint64_t x_lo, uint64_t x_hi, uint64_t y,
                                                         uint64_t sqrtx,
                                                         const PixWord* __restrict__ ptab,
                                                         const PixWord* __restrict__ tab,
                                                         uint64_t win_lo, uint64_t win_hi,
                                                         unsigned long long* __restrict__ acc)
{
  const u128 x = ((u128) x_hi << 64) | x_lo;
  const double xd = (double) x;
  const uint64_t w_lo = (y + 1) / 240, w_hi = sqrtx / 240;
  u128 sum = 0;
  for (uint64_t w = w_lo + (uint64_t) blockIdx.x * kThreads + threadIdx.x; w <= w_hi;
       w += (uint64_t) gridDim.x * kThreads) {
    uint64_t v = ptab[w].bits;
    while (v) {
      const uint32_t bit = (uint32_t) __ffsll((long long) v) - 1;
      v &= v - 1;
      const uint64_t p = 240 * w + 30 * (bit >> 3) + ((kWheelBytes >> (8 * (bit & 7))) & 0xff);
      if (p > y && p <= sqrtx)
        sum += pix_pi_win(tab, div_exact(xd, x_lo, p), win_lo, win_hi);
    }
  }
  block_add_i128(acc, (i128) sum);
}

/// Sigma4..6 (src/gourdon/Sigma.cpp:58-90) over the primes b_lo <= b < b_hi:
/// acc[0..1] += pi(x / (p y)) (p <= sqrt(x / y)), acc[2..3] += pi(x / p^2)
/// (p > sqrt(x / y)), acc[4..5] += pi(isqrt(x / p))^2.
__global__ __launch_bounds__(kThreads) void pix_sigma_kernel(
    uint64_t x_lo, uint64_t x_hi, uint64_t y, uint64_t sqrt_xy, const uint32_t* __restrict__ primes,
    uint32_t b_lo, uint32_t b_hi, const PixWord* __restrict__ tab,
    unsigned long long* __restrict__ acc)
{
  const u128 x = ((u128) x_hi << 64) | x_lo;
  i128 s4 = 0, s5 = 0, s6 = 0;
  for (uint32_t b = b_lo + blockIdx.x * kThreads + threadIdx.x; b < b_hi; b += gridDim.x * kThreads) {
    const uint64_t p = primes[b];
    if (p <= sqrt_xy)
      s4 += pix_pi(tab, (uint64_t) (x / ((u128) p * y)));
    else
      s5 += pix_pi(tab, (uint64_t) (x / ((u128) p * p)));
    const uint64_t t = pix_pi(tab, isqrt64((uint64_t) (x / p)));
    s6 += (i128) ((u128) t * t);
  }
  block_add_i128(acc, s4);
  block_add_i128(acc + 2, s5);
  block_add_i128(acc + 4, s6);
}

/// Phi0 (src/gourdon/Phi0.cpp:43-92): sum of mu(n) phi(x / n, k) over the
/// square-free n > 1 whose prime factors lie in (p_k, y] and that are primes
/// or <= z, read from the FactorTableD entries of n (coprime to 2310):
/// T_MAX = prime <= y (mu = -1), 0 = not square-free or a factor > y,
/// otherwise 2 pi(lpf) + [mu = -1].  phi(v, k) = (v / pp) phi(pp) +
/// phi_tbl[v mod pp] with pp the product of the first k primes.
/// With kLpf the entries are those of the lpf FactorTable (T_MAX = prime,
/// 0 = not square-free, lpf for mu = -1, lpf - 1 for mu = +1) and the sum is
/// the Deleglise-Rivat S1 over the square-free n <= limit with lpf > p_k
/// (src/S1.cpp:29-70).
template <typename FT, bool kLpf>
__global__ __launch_bounds__(kThreads) void pix_phi0_kernel(
    uint64_t x_lo, uint64_t x_hi, const FT* __restrict__ factor, uint64_t nidx, uint32_t k,
    uint32_t pk, const uint32_t* __restrict__ phi_tbl, uint32_t pp, uint64_t phi_pp,
    unsigned long long* __restrict__ acc)
{
  __shared__ uint16_t cop[480];
  if (threadIdx.x == 0) {
    uint32_t c = 0;
    for (uint32_t n = 0; n < 2310; n++)
      if (n % 2 && n % 3 && n % 5 && n % 7 && n % 11)
        cop[c++] = (uint16_t) n;
  }
  __syncthreads();
  const u128 x = ((u128) x_hi << 64) | x_lo;
  const FT tmax = (FT) ~(FT) 0;
  i128 sum = 0;
  for (uint64_t i = 1 + (uint64_t) blockIdx.x * kThreads + threadIdx.x; i < nidx;
       i += (uint64_t) gridDim.x * kThreads) {
    const FT f = factor[i];
    if (f == 0 || f == (FT) (tmax - 1))
      continue;
    const uint64_t n = 2310ull * (i / 480) + cop[i % 480];
    int mu;
    if (f == tmax) {
      if (n <= pk)
        continue;
      mu = -1;
    } else if (kLpf) {
      const uint32_t lpf = (f & 1) ? (uint32_t) f : (uint32_t) f + 1;
      if (lpf <= pk)
        continue;
      mu = (f & 1) ? -1 : 1;
    } else {
      if ((uint64_t) f <= 2ull * k + 1)
        continue;
      mu = (f & 1) ? -1 : 1;
    }
    uint32_t r;
    const u128 v = divmod128_32(x, (uint32_t) n, &r);
    const u128 q = divmod128_32(v, pp, &r);
    const i128 ph = (i128) (q * phi_pp) + phi_tbl[r];
    sum += mu > 0 ? ph : -ph;
  }
  block_add_i128(acc, sum);
}

__global__ void pix_pi_points_kernel(const uint64_t* __restrict__ n, uint64_t* __restrict__ out,
                                     uint32_t count, const PixWord* __restrict__ tab)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count)
    out[i] = pix_pi(tab, n[i]);
}

// ---------------------------------------------------------------------------
// nth prime: sieving an interval and selecting a prime

/// The primes 7 <= p <= p_max in order, read from the table's bits: the
/// prime of bit j of word w goes to out[tab[w].cnt + rank of j in the word].
__global__ __launch_bounds__(kThreads) void pix_extract_kernel(const PixWord* __restrict__ tab,
                                                               uint64_t p_max, uint32_t* __restrict__ out)
{
  const uint64_t w_hi = p_max / 240;
  for (uint64_t w = (uint64_t) blockIdx.x * kThreads + threadIdx.x; w <= w_hi;
       w += (uint64_t) gridDim.x * kThreads) {
    uint64_t v = tab[w].bits;
    uint64_t pos = tab[w].cnt;
    while (v) {
      const uint32_t bit = (uint32_t) __ffsll((long long) v) - 1;
      v &= v - 1;
      const uint64_t p = 240 * w + 30 * (bit >> 3) + ((kWheelBytes >> (8 * (bit & 7))) & 0xff);
      if (p <= p_max)
        out[pos] = (uint32_t) p;
      pos++;
    }
  }
}

/// Sieve of the interval [lo, hi] (words lo / 240 .. hi / 240, one
/// workgroup per kPixSegWords words in LDS): the numbers coprime to 30 in
/// the interval minus the multiples p * q (q >= p) of the sieving primes
/// sp[0 .. n_small) (all threads per prime); the larger sieving primes are
/// applied afterwards by pix_interval_large_kernel.
__global__ __launch_bounds__(kThreads) void pix_interval_kernel(
    uint64_t* __restrict__ out, uint64_t lo, uint64_t hi, uint64_t nwords,
    const uint32_t* __restrict__ sp, uint32_t n_small)
{
  __shared__ uint32_t bm[2 * kPixSegWords];
  const uint32_t t = threadIdx.x;
  const uint64_t w0 = (uint64_t) blockIdx.x * kPixSegWords;
  const uint32_t nw = (uint32_t) min<uint64_t>(kPixSegWords, nwords - w0);
  const uint64_t low = (lo / 240 + w0) * 240;            // first number of the block
  for (uint32_t w = t; w < nw; w += kThreads) {
    const uint64_t base = low + 240ull * w;
    uint64_t v = ~0ull;
    if (base + 239 > hi)
      v = base > hi ? 0ull : le_mask((uint32_t) (hi - base));
    if (base < lo)
      v &= lo - base >= 240 ? 0ull : ~(lo - base == 0 ? 0ull : le_mask((uint32_t) (lo - base - 1)));
    if (base == 0)
      v &= ~1ull;                                          // 1 is not prime
    bm[2 * w] = (uint32_t) v;
    bm[2 * w + 1] = (uint32_t) (v >> 32);
  }
  __syncthreads();
  const uint64_t high = low + 240ull * nw;               // exclusive
  const uint32_t cls = t & 7;
  const uint32_t t0 = t >> 3;
  for (uint32_t j = 0; j < n_small; j++) {
    const uint64_t p = sp[j];
    // exact 64-bit divisions: the interval may lie anywhere below 2^63
    const uint64_t q_lo = max(p, (low + p - 1) / p);
    const uint64_t q_hi = (high - 1) / p;
    if (q_lo > q_hi)
      continue;
    uint64_t q = q_lo - q_lo % 30 + ((kWheelBytes >> (8 * cls)) & 0xff);
    if (q < q_lo)
      q += 30;
    if (q > q_hi)
      continue;
    const uint32_t n = (uint32_t) ((q_hi - q) / 30) + 1;
    const uint32_t off = (uint32_t) (p * q - low);
    const uint32_t byte0 = off / 30;
    const uint32_t bib = wheel_bit(off - byte0 * 30);
    for (uint32_t k = t0; k < n; k += kThreads / 8) {
      const uint32_t B = byte0 + k * (uint32_t) p;
      atomicAnd(&bm[B >> 2], ~(1u << (((B & 3) << 3) | bib)));
    }
  }
  __syncthreads();
  for (uint32_t w = t; w < nw; w += kThreads)
    out[w0 + w] = ((uint64_t) bm[2 * w + 1] << 32) | bm[2 * w];
}

/// The sieving primes sp[j_lo .. j_hi) on the interval, one thread per prime,
/// with 64-bit atomics on the interval's words in HBM.
__global__ __launch_bounds__(kThreads) void pix_interval_large_kernel(
    uint64_t* __restrict__ out, uint64_t lo, uint64_t hi, const uint32_t* __restrict__ sp,
    uint64_t j_lo, uint64_t j_hi)
{
  const uint64_t wbase = lo / 240;
  for (uint64_t j = j_lo + (uint64_t) blockIdx.x * kThreads + threadIdx.x; j < j_hi;
       j += (uint64_t) gridDim.x * kThreads) {
    const uint64_t p = sp[j];
    const uint64_t q_lo = max(p, (lo + p - 1) / p);
    const uint64_t q_hi = hi / p;
    if (q_lo > q_hi)
      continue;
    const uint32_t qm = (uint32_t) (q_lo % 30);
    const uint64_t qbase = q_lo - qm;
    for (uint32_t idx = qm ? le_count(qm - 1) : 0u;; idx++) {
      const uint64_t q = qbase + 30ull * (idx >> 3) + ((kWheelBytes >> (8 * (idx & 7))) & 0xff);
      if (q > q_hi)
        break;
      const uint64_t m = p * q;
      const uint64_t w = m / 240;
      const uint32_t r = (uint32_t) (m - w * 240);
      const uint32_t b = 8 * (r / 30) + wheel_bit(r % 30);
      atomicAnd((unsigned long long*) &out[w - wbase], ~(1ull << b));
    }
  }
}

/// The primes 7 <= p <= p_max as 64-bit values (sieving primes above 2^32).
__global__ __launch_bounds__(kThreads) void pix_extract64_kernel(const PixWord* __restrict__ tab,
                                                                 uint64_t p_max,
                                                                 uint64_t* __restrict__ out)
{
  const uint64_t w_hi = p_max / 240;
  for (uint64_t w = (uint64_t) blockIdx.x * kThreads + threadIdx.x; w <= w_hi;
       w += (uint64_t) gridDim.x * kThreads) {
    uint64_t v = tab[w].bits;
    uint64_t pos = tab[w].cnt;
    while (v) {
      const uint32_t bit = (uint32_t) __ffsll((long long) v) - 1;
      v &= v - 1;
      const uint64_t p = 240 * w + 30 * (bit >> 3) + ((kWheelBytes >> (8 * (bit & 7))) & 0xff);
      if (p <= p_max)
        out[pos] = p;
      pos++;
    }
  }
}

/// (L_hi 2^64 + L_lo) mod p for L_hi < 2^20 and p < 2^42: L_hi (2^64 mod p)
/// stays below 2^62.
__device__ __forceinline__ uint64_t mod128_small(uint64_t L_hi, uint64_t L_lo, uint64_t p)
{
  const uint64_t c = (~0ull % p + 1) % p;            // 2^64 mod p
  return (L_hi * c + L_lo % p) % p;
}

/// q^-1 mod 30 for q coprime to 30 (q mod 30 in {1, 7, 11, 13, 17, 19, 23, 29}).
__device__ __forceinline__ uint32_t inv_mod30(uint32_t q)
{
  switch (q) {
    case 7: return 13;
    case 11: return 11;
    case 13: return 7;
    case 17: return 23;
    case 19: return 19;
    case 23: return 17;
    case 29: return 29;
    default: return 1;
  }
}

/// Sieve of the interval [L + o_lo, L + o_hi] for a base L above 2^63 (a
/// multiple of 240, L = L_hi 2^64 + L_lo): word w of `out` stands for the
/// numbers [L + 240 w, L + 240 w + 240).  One workgroup per kPixSegWords
/// words in LDS crosses off every multiple of the sieving primes
/// sp[0 .. n_small): the interval lies far above their squares, so every
/// multiple in it is composite and p itself is not in it.  Thread class c
/// (8 per prime) takes the multiples p q with p q = 30 k + r_c (the wheel
/// residue of bit c): q = r_c p^-1 (mod 30), one multiple every 30 p.
__global__ __launch_bounds__(kThreads) void pix_interval128_kernel(
    uint64_t* __restrict__ out, uint64_t L_hi, uint64_t L_lo, uint64_t o_lo, uint64_t o_hi,
    uint64_t nwords, const uint64_t* __restrict__ sp, uint32_t n_small)
{
  __shared__ uint32_t bm[2 * kPixSegWords];
  const uint32_t t = threadIdx.x;
  const uint64_t w0 = (uint64_t) blockIdx.x * kPixSegWords;
  const uint32_t nw = (uint32_t) min<uint64_t>(kPixSegWords, nwords - w0);
  const uint64_t s = 240 * w0;                             // block offset from L
  for (uint32_t w = t; w < nw; w += kThreads) {
    const uint64_t base = s + 240ull * w;
    uint64_t v = ~0ull;
    if (base + 239 > o_hi)
      v = base > o_hi ? 0ull : le_mask((uint32_t) (o_hi - base));
    if (base < o_lo)
      v &= o_lo - base >= 240 ? 0ull : ~(o_lo - base == 0 ? 0ull : le_mask((uint32_t) (o_lo - base - 1)));
    bm[2 * w] = (uint32_t) v;
    bm[2 * w + 1] = (uint32_t) (v >> 32);
  }
  __syncthreads();
  const uint64_t seg_len = 240ull * nw;
  const uint32_t cls = t & 7;
  const uint32_t t0 = t >> 3;
  const uint32_t wr = (uint32_t) ((kWheelBytes >> (8 * cls)) & 0xff);
  for (uint32_t j = 0; j < n_small; j++) {
    const uint64_t p = sp[j];
    // first multiple at or after L + s: offset d from the block start
    const uint64_t r = (mod128_small(L_hi, L_lo, p) + s % p) % p;
    const uint64_t d = r ? p - r : 0;
    // its cofactor mod 30 (L + s is a multiple of 30), then the class's first
    const uint32_t pinv = inv_mod30((uint32_t) (p % 30));
    const uint32_t q0m = (uint32_t) (d % 30) * pinv % 30;
    const uint32_t qc = wr * pinv % 30;
    const uint64_t off = d + (uint64_t) ((qc + 30 - q0m) % 30) * p;
    if (off >= seg_len)
      continue;
    const uint32_t n = (uint32_t) ((seg_len - 1 - off) / (30 * p)) + 1;
    const uint32_t byte0 = (uint32_t) (off / 30);
    for (uint32_t k = t0; k < n; k += kThreads / 8) {
      const uint32_t B = byte0 + k * (uint32_t) p;
      atomicAnd(&bm[B >> 2], ~(1u << (((B & 3) << 3) | cls)));
    }
  }
  __syncthreads();
  for (uint32_t w = t; w < nw; w += kThreads)
    out[w0 + w] = ((uint64_t) bm[2 * w + 1] << 32) | bm[2 * w];
}

/// The sieving primes sp[j_lo .. j_hi) on the interval of
/// pix_interval128_kernel, one thread per prime, with 64-bit atomics on the
/// interval's words in HBM: the multiples p q with q coprime to 30.
__global__ __launch_bounds__(kThreads) void pix_interval128_large_kernel(
    uint64_t* __restrict__ out, uint64_t L_hi, uint64_t L_lo, uint64_t o_lo, uint64_t o_hi,
    const uint64_t* __restrict__ sp, uint64_t j_lo, uint64_t j_hi)
{
  for (uint64_t j = j_lo + (uint64_t) blockIdx.x * kThreads + threadIdx.x; j < j_hi;
       j += (uint64_t) gridDim.x * kThreads) {
    const uint64_t p = sp[j];
    const uint64_t r = (mod128_small(L_hi, L_lo, p) + o_lo % p) % p;
    const uint64_t off = o_lo + (r ? p - r : 0);           // first multiple at or after L + o_lo
    if (off > o_hi)
      continue;
    const uint32_t qm = (uint32_t) (off % 30) * inv_mod30((uint32_t) (p % 30)) % 30;
    for (uint32_t idx = qm ? le_count(qm - 1) : 0u;; idx++) {
      const uint64_t dq = 30ull * (idx >> 3) + ((kWheelBytes >> (8 * (idx & 7))) & 0xff) - qm;
      const uint64_t o = off + dq * p;
      if (o > o_hi)
        break;
      const uint64_t w = o / 240;
      const uint32_t r240 = (uint32_t) (o - w * 240);
      const uint32_t b = 8 * (r240 / 30) + wheel_bit(r240 % 30);
      atomicAnd((unsigned long long*) &out[w], ~(1ull << b));
    }
  }
}

/// Primes per block of kPixSegWords words of the interval bitmap.
__global__ __launch_bounds__(kThreads) void pix_block_count_kernel(const uint64_t* __restrict__ bits,
                                                                   uint64_t nwords,
                                                                   uint64_t* __restrict__ counts)
{
  __shared__ unsigned long long red[kThreads];
  const uint64_t w0 = (uint64_t) blockIdx.x * kPixSegWords;
  unsigned long long s = 0;
  for (uint64_t w = w0 + threadIdx.x; w < min<uint64_t>(w0 + kPixSegWords, nwords); w += kThreads)
    s += (unsigned long long) __popcll(bits[w]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t d = kThreads / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d)
      red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    counts[blockIdx.x] = red[0];
}

/// The n-th prime (n >= 4) from the table: the last word w with
/// 3 + cnt[w] < n holds it (binary search), then its bit of rank
/// n - 3 - cnt[w].  One thread.
__global__ void pix_select_kernel(const PixWord* __restrict__ tab, uint64_t nwords, uint64_t n,
                                  uint64_t* __restrict__ out)
{
  uint64_t lo = 0, hi = nwords - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) / 2;
    if (3 + tab[mid].cnt < n)
      lo = mid;
    else
      hi = mid - 1;
  }
  uint64_t v = tab[lo].bits;
  uint64_t k = n - 3 - tab[lo].cnt;             // 1-based rank inside the word
  uint64_t res = 0;
  while (v) {
    const uint32_t bit = (uint32_t) __ffsll((long long) v) - 1;
    v &= v - 1;
    if (--k == 0) {
      res = 240 * lo + 30 * (bit >> 3) + ((kWheelBytes >> (8 * (bit & 7))) & 0xff);
      break;
    }
  }
  *out = res;
}

uint32_t grid_for(uint64_t work, uint32_t cap)
{
  const uint64_t g = (work + kThreads - 1) / kThreads;
  return (uint32_t) (g < 1 ? 1 : (g > cap ? cap : g));
}

} // namespace

hipError_t pix_launch_sieve(PixWord* tab, uint64_t* seg_tot, const uint32_t* primes,
                            uint32_t j_small_end, uint32_t j_end, uint64_t nwords, uint64_t N,
                            uint64_t wbase, hipStream_t st)
{
  const uint64_t nseg = (nwords + kPixSegWords - 1) / kPixSegWords;
  hipLaunchKernelGGL(pix_sieve_kernel, dim3((uint32_t) nseg), dim3(kThreads), 0, st, tab, seg_tot,
                     primes, j_small_end, j_end, nwords, N, wbase);
  return hipGetLastError();
}

hipError_t pix_launch_scan(uint64_t* seg_tot, uint32_t nseg, hipStream_t st)
{
  hipLaunchKernelGGL(pix_scan_kernel, dim3(1), dim3(1024), 0, st, seg_tot, nseg);
  return hipGetLastError();
}

hipError_t pix_launch_counts(PixWord* tab, const uint64_t* seg_base, uint64_t nwords,
                             uint64_t base, hipStream_t st)
{
  const uint64_t nseg = (nwords + kPixSegWords - 1) / kPixSegWords;
  hipLaunchKernelGGL(pix_count_kernel, dim3((uint32_t) nseg), dim3(kThreads), 0, st, tab, seg_base,
                     nwords, base);
  return hipGetLastError();
}

hipError_t pix_launch_leafsum(const PixLeafDesc* desc, uint32_t ndesc, uint64_t total,
                              const uint32_t* primes, const PixWord* tab, uint64_t w_lo,
                              uint64_t w_hi, unsigned long long* acc, uint32_t cus, hipStream_t st)
{
  if (total == 0 || ndesc == 0) return hipSuccess;
  // contiguous ranges of at least 4,096 terms, up to 64 waves per CU
  const uint64_t max_waves = (uint64_t) cus * 64;
  const uint64_t waves = std::min<uint64_t>((total + 4095) / 4096, max_waves);
  const uint64_t per_wave = ((total + waves - 1) / waves + 63) / 64 * 64;
  const uint32_t grid = (uint32_t) ((waves + kThreads / 64 - 1) / (kThreads / 64));
  const bool win = w_lo != 0 || w_hi != ~0ull;
  hipLaunchKernelGGL(win ? pix_leafsum_kernel<true> : pix_leafsum_kernel<false>, dim3(grid),
                     dim3(kThreads), 0, st, desc, ndesc, total,
                     per_wave, primes, tab, w_lo, w_hi, acc);
  return hipGetLastError();
}

hipError_t pix_launch_leaftiles(const PixTileDesc* desc, const PixTileGroup* grp, uint32_t ngrp,
                                uint64_t nitems, const uint32_t* primes, uint32_t nprimes,
                                const PixWord* tab, uint64_t w_lo, uint64_t w_hi,
                                uint64_t max_quotient, unsigned long long* acc, uint32_t cus,
                                hipStream_t st)
{
  if (nitems == 0 || ngrp == 0) return hipSuccess;
  // the quotient guard v 2^-46 stays at most 1/8 for quotients below 2^43 (at
  // most a quarter of the lanes take the residual fix-up)
  if (max_quotient >= kPixTileMaxQuotient) return hipErrorInvalidValue;
  // contiguous item ranges, up to 64 waves per CU
  const uint64_t max_waves = (uint64_t) cus * 64;
  const uint64_t per_wave = (nitems + max_waves - 1) / max_waves;
  const uint64_t waves = (nitems + per_wave - 1) / per_wave;
  const uint32_t grid = (uint32_t) ((waves + kThreads / 64 - 1) / (kThreads / 64));
  const bool win = w_lo != 0 || w_hi != ~0ull;
  hipLaunchKernelGGL(win ? pix_leaftile_kernel<true> : pix_leaftile_kernel<false>, dim3(grid),
                     dim3(kThreads), 0, st, desc, grp, ngrp,
                     nitems, per_wave, primes, nprimes, tab, w_lo, w_hi, acc);
  return hipGetLastError();
}

hipError_t pix_launch_B(uint64_t x_lo, uint64_t x_hi, uint64_t y, uint64_t sqrtx,
                        const PixWord* ptab, const PixWord* tab, uint64_t w_lo, uint64_t w_hi,
                        unsigned long long* acc, hipStream_t st)
{
  if (sqrtx <= y) return hipSuccess;
  const uint64_t words = sqrtx / 240 - (y + 1) / 240 + 1;
  hipLaunchKernelGGL(pix_B_kernel, dim3(grid_for(words, 65536)), dim3(kThreads), 0, st, x_lo, x_hi,
                     y, sqrtx, ptab, tab, w_lo, w_hi, acc);
  return hipGetLastError();
}

hipError_t pix_launch_sigma(uint64_t x_lo, uint64_t x_hi, uint64_t y, uint64_t sqrt_xy,
                            const uint32_t* primes, uint32_t b_lo, uint32_t b_hi,
                            const PixWord* tab, unsigned long long* acc, hipStream_t st)
{
  if (b_hi <= b_lo) return hipSuccess;
  hipLaunchKernelGGL(pix_sigma_kernel, dim3(grid_for(b_hi - b_lo, 4096)), dim3(kThreads), 0, st,
                     x_lo, x_hi, y, sqrt_xy, primes, b_lo, b_hi, tab, acc);
  return hipGetLastError();
}

hipError_t pix_launch_phi0(uint64_t x_lo, uint64_t x_hi, const void* factor, bool factor32,
                           bool lpf_table, uint64_t nidx, uint32_t k, uint32_t pk,
                           const uint32_t* phi_tbl, uint32_t pp, uint64_t phi_pp,
                           unsigned long long* acc, hipStream_t st)
{
  if (nidx <= 1) return hipSuccess;
  const uint32_t grid = grid_for(nidx, 16384);
  if (factor32 && lpf_table)
    hipLaunchKernelGGL((pix_phi0_kernel<uint32_t, true>), dim3(grid), dim3(kThreads), 0, st, x_lo,
                       x_hi, (const uint32_t*) factor, nidx, k, pk, phi_tbl, pp, phi_pp, acc);
  else if (factor32)
    hipLaunchKernelGGL((pix_phi0_kernel<uint32_t, false>), dim3(grid), dim3(kThreads), 0, st, x_lo,
                       x_hi, (const uint32_t*) factor, nidx, k, pk, phi_tbl, pp, phi_pp, acc);
  else if (lpf_table)
    hipLaunchKernelGGL((pix_phi0_kernel<uint16_t, true>), dim3(grid), dim3(kThreads), 0, st, x_lo,
                       x_hi, (const uint16_t*) factor, nidx, k, pk, phi_tbl, pp, phi_pp, acc);
  else
    hipLaunchKernelGGL((pix_phi0_kernel<uint16_t, false>), dim3(grid), dim3(kThreads), 0, st, x_lo,
                       x_hi, (const uint16_t*) factor, nidx, k, pk, phi_tbl, pp, phi_pp, acc);
  return hipGetLastError();
}

hipError_t pix_launch_extract(const PixWord* tab, uint64_t p_max, uint32_t* out, hipStream_t st)
{
  hipLaunchKernelGGL(pix_extract_kernel, dim3(grid_for(p_max / 240 + 1, 65536)), dim3(kThreads), 0,
                     st, tab, p_max, out);
  return hipGetLastError();
}

hipError_t pix_launch_interval(uint64_t* out, uint64_t lo, uint64_t hi, const uint32_t* sp,
                               uint32_t n_small, uint64_t n_all, hipStream_t st)
{
  const uint64_t nwords = hi / 240 - lo / 240 + 1;
  const uint64_t nblk = (nwords + kPixSegWords - 1) / kPixSegWords;
  hipLaunchKernelGGL(pix_interval_kernel, dim3((uint32_t) nblk), dim3(kThreads), 0, st, out, lo, hi,
                     nwords, sp, n_small);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n_all <= n_small) return e;
  hipLaunchKernelGGL(pix_interval_large_kernel, dim3(grid_for(n_all - n_small, 65536)),
                     dim3(kThreads), 0, st, out, lo, hi, sp, (uint64_t) n_small, n_all);
  return hipGetLastError();
}

hipError_t pix_launch_extract64(const PixWord* tab, uint64_t p_max, uint64_t* out, hipStream_t st)
{
  hipLaunchKernelGGL(pix_extract64_kernel, dim3(grid_for(p_max / 240 + 1, 65536)), dim3(kThreads), 0,
                     st, tab, p_max, out);
  return hipGetLastError();
}

hipError_t pix_launch_interval128(uint64_t* out, uint64_t L_hi, uint64_t L_lo, uint64_t o_lo,
                                  uint64_t o_hi, const uint64_t* sp, uint32_t n_small,
                                  uint64_t n_all, hipStream_t st)
{
  // (L itself is a multiple of 240; its low word need not be when L >= 2^64)
  if (L_hi >= (1ull << 20) || o_lo > o_hi || o_hi >= (1ull << 62))
    return hipErrorInvalidValue;
  const uint64_t nwords = o_hi / 240 + 1;
  const uint64_t nseg = (nwords + kPixSegWords - 1) / kPixSegWords;
  hipLaunchKernelGGL(pix_interval128_kernel, dim3((uint32_t) nseg), dim3(kThreads), 0, st, out,
                     L_hi, L_lo, o_lo, o_hi, nwords, sp, n_small);
  if (n_all > n_small)
    hipLaunchKernelGGL(pix_interval128_large_kernel, dim3(grid_for(n_all - n_small, 65536)),
                       dim3(kThreads), 0, st, out, L_hi, L_lo, o_lo, o_hi, sp, (uint64_t) n_small,
                       n_all);
  return hipGetLastError();
}

hipError_t pix_launch_block_counts(const uint64_t* bits, uint64_t nwords, uint64_t* counts,
                                   hipStream_t st)
{
  const uint64_t nblk = (nwords + kPixSegWords - 1) / kPixSegWords;
  hipLaunchKernelGGL(pix_block_count_kernel, dim3((uint32_t) nblk), dim3(kThreads), 0, st, bits,
                     nwords, counts);
  return hipGetLastError();
}

hipError_t pix_launch_select(const PixWord* tab, uint64_t nwords, uint64_t n, uint64_t* out,
                             hipStream_t st)
{
  hipLaunchKernelGGL(pix_select_kernel, dim3(1), dim3(1), 0, st, tab, nwords, n, out);
  return hipGetLastError();
}

hipError_t pix_launch_pi_points(const uint64_t* n, uint64_t* out, uint32_t count,
                                const PixWord* tab, hipStream_t st)
{
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(pix_pi_points_kernel, dim3((count + 255) / 256), dim3(256), 0, st, n, out,
                     count, tab);
  return hipGetLastError();
}

} // namespace pcg
