// hsl_kernels.hip -- CDNA4 (gfx950) kernels of the hard-special-leaf engine.
//
// One workgroup of 256 threads owns one segment of the sieve interval at a
// time, held in LDS as a wheel-30 bitmap (240 numbers per 64-bit word, the
// layout of src/sieve/Sieve.cpp:62-66).  For b = b_first, b_first+1, ... the
// workgroup
//   1. answers the leaves of p_b in the segment: each thread takes factor-table
//      entries (or primes) strided by 256, filters them, computes the exact
//      quotient x / (p_b * m) and counts the unsieved numbers <= it with a
//      three-level count (super-group prefix, group counters, popcount);
//   2. records C(s, b), the unsieved count of the segment before p_b is
//      crossed off, and M(s, b), the sum of the leaf signs;
//   3. crosses off p_b and its multiples with LDS atomics, decrementing the
//      group counters for every number removed for the first time
//      (src/sieve/Sieve.cpp:497-845 counts the same first-time removals).
// The reference carries phi(low, b-1) sequentially through its segments
// (src/gourdon/D.cpp:96, 271).  Here the carry is split: a workgroup keeps
// the carry within its chunk, and phi at the chunk start is an exclusive
// scan over chunks of the per-chunk counts, done by hsl_combine_* afterwards:
//   sum = sum_leaves sign * (phi_chunk(b) + phi_local(b) + count)
//       = sum_chunks [ S(chunk) + sum_b M(chunk, b) * phi_chunk(b) ].

#include <hip/hip_runtime.h>

#include <algorithm>

#include "hsl_kernels.hpp"

namespace pcg {

__constant__ int16_t c_cop_index[2310];   // largest index of a coprime <= r
__constant__ uint16_t c_cop[480];         // coprimes to 2310 in [0, 2310)
__constant__ uint8_t c_wheel[8];          // 1, 7, 11, ..., 29
__constant__ uint8_t c_wheel_bit[30];     // residue -> bit in byte (0xff: none)

namespace {

constexpr int kWaves = kBlock / 64;

/// floor(xp / d) for a quotient < 2^50 and d < 2^53.  xd is xp rounded to
/// double; the IEEE division estimate is within one of the true quotient
/// (relative error < 2^-51), and the residual xp - q*d, evaluated modulo 2^64
/// from the low word of xp, is exact because |residual| < 2 d < 2^63.
__device__ __forceinline__ uint64_t div_exact(double xd, uint64_t xlo, uint64_t d)
{
  double qd = xd / (double) d;
  uint64_t q = (uint64_t) qd;
  int64_t r = (int64_t) (xlo - q * d);
  if (r < 0)
    q--;
  else if ((uint64_t) r >= d)
    q++;
  return q;
}

/// min(floor(xp / d), cap) for cap < 2^48.
__device__ __forceinline__ uint64_t div_cap(double xd, uint64_t xlo, uint64_t d, uint64_t cap)
{
  double qd = xd / (double) d;
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
/// coprime to 2310 that is <= n (n >= 1).
__device__ __forceinline__ uint64_t ft_to_index(uint64_t n)
{
  uint64_t q = n / 2310;
  return 480 * q + (int64_t) c_cop_index[n - q * 2310];
}

/// pi(n) for n within the prime table.
__device__ __forceinline__ uint64_t pi_lookup(const uint64_t* __restrict__ bits,
                                              const uint32_t* __restrict__ cnt,
                                              const uint64_t* lemask, uint64_t n)
{
  if (n < 7)
    return n < 2 ? 0 : (n < 3 ? 1 : (n < 5 ? 2 : 3));
  uint64_t blk = n / 240;
  uint32_t r = (uint32_t) (n - blk * 240);
  return 3 + (uint64_t) cnt[blk] + (uint64_t) __popcll(bits[blk] & lemask[r]);
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

/// Unsieved numbers in [low, low + stop] of the segment.
__device__ __forceinline__ uint32_t seg_count(const uint32_t* bm, const uint32_t* cnt1,
                                              const uint32_t* pre2, const uint64_t* lemask,
                                              uint32_t stop)
{
  uint32_t w = stop / 240;
  uint32_t r = stop - w * 240;
  uint32_t g = w >> 3;
  uint32_t sg = w >> 6;
  uint32_t c = pre2[sg];
  uint4 v = reinterpret_cast<const uint4*>(cnt1)[sg];
  uint32_t ng = g & 7;
  uint32_t h[8] = {v.x & 0xffff, v.x >> 16, v.y & 0xffff, v.y >> 16,
                   v.z & 0xffff, v.z >> 16, v.w & 0xffff, v.w >> 16};
#pragma unroll
  for (uint32_t i = 0; i < 7; i++)
    c += i < ng ? h[i] : 0;
  const uint2* words = reinterpret_cast<const uint2*>(bm);
  for (uint32_t u = g << 3; u < w; u++) {
    uint2 x = words[u];
    c += __popc(x.x) + __popc(x.y);
  }
  uint2 x = words[w];
  uint64_t m = lemask[r];
  c += __popc(x.x & (uint32_t) m) + __popc(x.y & (uint32_t) (m >> 32));
  return c;
}

/// Cross off p and its multiples in the segment (numbers p*q > low with q
/// coprime to 30, as Sieve::add/cross_off, src/sieve/Sieve.cpp:192-216).
/// Threads split the 8 residue classes of q mod 30 and stride inside a class:
/// consecutive multiples of one class are p bytes apart.  With COUNT the
/// first-time removals decrement the group counters and are returned.
template <bool COUNT>
__device__ __forceinline__ uint32_t cross_off(uint32_t p, uint64_t low, uint32_t seg_len,
                                              uint32_t* bm, uint32_t* cnt1,
                                              unsigned long long& hits)
{
  if (p < 7)
    return 0;
  const uint32_t cls = threadIdx.x & 7;
  const uint32_t j0 = threadIdx.x >> 3;
  uint64_t q_lo = div_exact((double) low, low, p) + 1;
  uint64_t q_hi = div_exact((double) (low + seg_len - 1), low + seg_len - 1, p);
  if (q_hi < q_lo)
    return 0;
  uint64_t q = q_lo - q_lo % 30 + c_wheel[cls];
  if (q < q_lo)
    q += 30;
  if (q > q_hi)
    return 0;
  uint32_t n = (uint32_t) ((q_hi - q) / 30) + 1;
  uint64_t off = (uint64_t) p * q - low;
  uint32_t byte0 = (uint32_t) (off / 30);
  uint32_t bib = c_wheel_bit[off - (uint64_t) byte0 * 30];
  uint32_t removed = 0;
  for (uint32_t t = j0; t < n; t += kBlock / 8) {
    uint32_t B = byte0 + t * p;
    uint32_t mask = 1u << (((B & 3) << 3) | bib);
    if (COUNT) {
      uint32_t old = atomicAnd(&bm[B >> 2], ~mask);
      if (old & mask) {
        removed++;
        uint32_t g = B >> 6;
        atomicSub(&cnt1[g >> 1], 1u << ((g & 1) << 4));
      }
    } else {
      atomicAnd(&bm[B >> 2], ~mask);
    }
  }
  if (COUNT && j0 < n)
    hits += (n - j0 + kBlock / 8 - 1) / (kBlock / 8);
  return removed;
}

} // namespace

template <typename FT>
__global__ __launch_bounds__(kBlock) void hsl_sieve_kernel(HslArgs a)
{
  extern __shared__ __align__(16) uint32_t lds[];
  const uint32_t W = a.W;
  uint32_t* bm = lds;                                        // 2W u32
  uint32_t* cnt1 = bm + 2 * W;                               // W/16 u32
  uint32_t* pre2 = cnt1 + W / 16;                            // 64 u32
  uint64_t* lemask = reinterpret_cast<uint64_t*>(pre2 + 64); // 240 u64
  uint16_t* cop = reinterpret_cast<uint16_t*>(lemask + 240); // 480 u16

  __shared__ int sh_acc[4];             // [parity] leaf sign sum, [2+parity] removals
  __shared__ uint32_t sh_chunk;
  __shared__ long long sh_red[kWaves];

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t wave = tid >> 6;

  for (uint32_t r = tid; r < 240; r += kBlock) {
    uint64_t m = 0;
    for (int j = 0; j < 8; j++)
      for (int i = 0; i < 8; i++)
        if ((uint32_t) (30 * j + c_wheel[i]) <= r)
          m |= 1ull << (8 * j + i);
    lemask[r] = m;
  }
  for (uint32_t i = tid; i < 480; i += kBlock)
    cop[i] = c_cop[i];
  if (tid < 4)
    sh_acc[tid] = 0;
  __syncthreads();

  const uint64_t SEG = 240ull * W;
  const FT* __restrict__ factor = reinterpret_cast<const FT*>(a.factor);
  unsigned long long n_leaves = 0, n_hits = 0, n_entries = 0, n_pairs = 0;

  for (;;) {
    if (tid == 0)
      sh_chunk = atomicAdd(a.queue, 1u);
    __syncthreads();
    const uint32_t qi = sh_chunk;
    __syncthreads();
    if (qi >= a.nchunks)
      break;
    const uint32_t ci = a.order[qi];

    const ChunkDesc ch = a.chunks[ci];
    const QueryDesc qd = a.queries[ch.query];
    const BInfo* __restrict__ binfo = a.binfo + qd.binfo_off;
    uint32_t* TC = a.T_C + ch.tc_off;
    int* TM = a.T_M + ch.tc_off;
    const uint32_t b_stop = min(qd.b_last, qd.b_first + ch.nbc - 1);
    long long S = 0;         // this thread: sum of sign * in-segment count
    __int128 Sacc = 0;       // thread 0: sum of M(s, b) * phi_local(b)

    for (uint32_t s = ch.seg_begin; s < ch.seg_end; s++) {
      const uint64_t low = (uint64_t) s * SEG;
      const uint64_t high = min(low + SEG, qd.limit);
      const uint32_t seg_len = (uint32_t) (high - low);
      const uint64_t low1 = low > 0 ? low : 1;

      // numbers coprime to 30 in [low, high)
      for (uint32_t w = tid; w < W; w += kBlock) {
        uint32_t base = w * 240;
        uint64_t v;
        if (base + 240 <= seg_len)
          v = ~0ull;
        else if (base >= seg_len)
          v = 0;
        else
          v = lemask[seg_len - base - 1];
        bm[2 * w] = (uint32_t) v;
        bm[2 * w + 1] = (uint32_t) (v >> 32);
      }
      __syncthreads();

      // pre-sieve p_4 .. p_{b_first - 1} (2, 3, 5 are outside the wheel)
      for (uint32_t j = 4; j <= qd.pre_last; j++)
        cross_off<false>(a.primes[j], low, seg_len, bm, cnt1, n_hits);
      __syncthreads();

      // group counters and the segment total
      uint32_t my_total = 0;
      {
        const uint2* words = reinterpret_cast<const uint2*>(bm);
        for (uint32_t gp = tid; gp < W / 16; gp += kBlock) {
          uint32_t c0 = 0, c1 = 0;
#pragma unroll
          for (int u = 0; u < 8; u++) {
            uint2 x0 = words[16 * gp + u];
            uint2 x1 = words[16 * gp + 8 + u];
            c0 += __popc(x0.x) + __popc(x0.y);
            c1 += __popc(x1.x) + __popc(x1.y);
          }
          cnt1[gp] = c0 | (c1 << 16);
          my_total += c0 + c1;
        }
      }
      long long tsum = wave_sum_i64((long long) my_total);
      if (lane == 0)
        sh_red[wave] = tsum;
      __syncthreads();
      uint32_t total = 0;      // meaningful in thread 0 only
      if (tid == 0)
        for (int i = 0; i < kWaves; i++)
          total += (uint32_t) sh_red[i];

      for (uint32_t b = qd.b_first; b <= b_stop; b++) {
        const BInfo bi = binfo[b - qd.b_first];
        const uint64_t m_hi = div_cap(bi.xp_d, bi.xp_lo, low1, bi.A);
        uint64_t lo_i, hi_i;
        if (bi.type == 1) {
          if (bi.p >= m_hi)
            break;
          uint64_t m_lo = div_cap(bi.xp_d, bi.xp_lo, high, bi.C);
          m_lo = m_lo > bi.B ? m_lo : bi.B;
          hi_i = ft_to_index(m_hi);
          lo_i = m_lo < m_hi ? ft_to_index(m_lo) : hi_i;
        } else {
          hi_i = pi_lookup(a.pi_bits, a.pi_cnt, lemask, m_hi);
          if (hi_i <= b)
            break;
          uint64_t m_lo = div_cap(bi.xp_d, bi.xp_lo, high, bi.C);
          m_lo = m_lo > bi.B ? m_lo : bi.B;
          lo_i = m_lo < m_hi ? pi_lookup(a.pi_bits, a.pi_cnt, lemask, m_lo) : hi_i;
        }
        const uint32_t bidx = b - qd.b_first;
        const int par = b & 1;
        uint32_t phiL = 0;
        int Mold = 0;
        if (tid == 0) {
          phiL = TC[bidx];
          Mold = TM[bidx];
        }

        if (hi_i > lo_i) {
          // super-group prefix over the group counters
          if (wave == 0) {
            uint32_t v = 0;
            if (lane < W / 64) {
              uint4 q = reinterpret_cast<const uint4*>(cnt1)[lane];
              v = (q.x & 0xffff) + (q.x >> 16) + (q.y & 0xffff) + (q.y >> 16) +
                  (q.z & 0xffff) + (q.z >> 16) + (q.w & 0xffff) + (q.w >> 16);
            }
            uint32_t incl = v;
            for (int off = 1; off < 64; off <<= 1) {
              uint32_t t = __shfl_up(incl, off, 64);
              if ((int) lane >= off)
                incl += t;
            }
            if (lane < W / 64)
              pre2[lane] = incl - v;
          }
          __syncthreads();

          int M = 0;
          if (bi.type == 1) {
            n_entries += (hi_i - lo_i + kBlock - 1 - tid) / kBlock;
            for (uint64_t idx = lo_i + 1 + tid; idx <= hi_i; idx += kBlock) {
              uint32_t f = factor[idx];
              if (f > bi.T) {
                uint64_t qq = idx / 480;
                uint64_t m = 2310 * qq + cop[idx - qq * 480];
                uint64_t n = div_exact(bi.xp_d, bi.xp_lo, m);
                uint32_t cnt = seg_count(bm, cnt1, pre2, lemask, (uint32_t) (n - low));
                int sg = (f & 1) ? 1 : -1;
                S += sg * (long long) cnt;
                M += sg;
                n_leaves++;
              }
            }
          } else {
            for (uint64_t j = lo_i + 1 + tid; j <= hi_i; j += kBlock) {
              uint64_t q = a.primes[j];
              uint64_t n = div_exact(bi.xp_d, bi.xp_lo, q);
              S += seg_count(bm, cnt1, pre2, lemask, (uint32_t) (n - low));
              M += 1;
              n_leaves++;
            }
          }
          M = wave_sum_i32(M);
          if (lane == 0 && M != 0)
            atomicAdd(&sh_acc[par], M);
          __syncthreads();
        }

        uint32_t removed = cross_off<true>(bi.p, low, seg_len, bm, cnt1, n_hits);
        int rem = wave_sum_i32((int) removed);
        if (lane == 0 && rem != 0)
          atomicAdd(&sh_acc[2 + par], rem);
        __syncthreads();
        if (tid == 0) {
          int M = sh_acc[par];
          int R = sh_acc[2 + par];
          sh_acc[par] = 0;
          sh_acc[2 + par] = 0;
          Sacc += (__int128) M * (__int128) phiL;
          TC[bidx] = phiL + total;
          TM[bidx] = Mold + M;
          total -= (uint32_t) R;
          n_pairs++;
        }
      }
      __syncthreads();
    }

    long long ws = wave_sum_i64(S);
    if (lane == 0)
      sh_red[wave] = ws;
    __syncthreads();
    if (tid == 0) {
      __int128 tot = Sacc;
      for (int i = 0; i < kWaves; i++)
        tot += sh_red[i];
      a.S_out[2 * (size_t) ci] = (unsigned long long) tot;
      a.S_out[2 * (size_t) ci + 1] = (unsigned long long) (tot >> 64);
    }
    __syncthreads();
  }

  long long l = wave_sum_i64((long long) n_leaves);
  long long h = wave_sum_i64((long long) n_hits);
  long long e = wave_sum_i64((long long) n_entries);
  if (lane == 0) {
    atomicAdd(&a.counters[0], (unsigned long long) l);
    atomicAdd(&a.counters[1], (unsigned long long) h);
    atomicAdd(&a.counters[2], (unsigned long long) e);
  }
  if (tid == 0)
    atomicAdd(&a.counters[3], n_pairs);
}

template __global__ void hsl_sieve_kernel<uint16_t>(HslArgs);
template __global__ void hsl_sieve_kernel<uint32_t>(HslArgs);

/// Zero-fill (used instead of hipMemsetAsync so the sequence can be captured).
__global__ void hsl_fill_zero(uint4* p, size_t n16)
{
  size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for (; i < n16; i += stride)
    p[i] = make_uint4(0, 0, 0, 0);
}

/// Column = one (query, b).  Slice = a range of that query's chunks.
/// For each column and slice: the slice's count total, sign total and the
/// sum of M(chunk, b) * (counts of the earlier chunks inside the slice).
__global__ void hsl_combine_slices(const CombineCol* __restrict__ cols, uint32_t ncols,
                                   const ChunkDesc* __restrict__ chunks,
                                   const uint32_t* __restrict__ T_C, const int* __restrict__ T_M,
                                   uint32_t slice_len, uint32_t nslices,
                                   long long* sC, long long* sM, __int128* sMP)
{
  uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t sl = blockIdx.y;
  if (col >= ncols)
    return;
  const CombineCol cc = cols[col];
  uint32_t c0 = sl * slice_len;
  uint32_t c1 = min(cc.nch, c0 + slice_len);
  long long P = 0, Ms = 0;
  __int128 MP = 0;
  for (uint32_t c = c0; c < c1; c++) {
    const ChunkDesc ch = chunks[cc.chunk_first + c];
    uint64_t off = ch.tc_off + cc.b_rel;
    long long tc = T_C[off];
    long long tm = T_M[off];
    MP += (__int128) tm * P;
    P += tc;
    Ms += tm;
  }
  size_t o = (size_t) sl * ncols + col;
  sC[o] = P;
  sM[o] = Ms;
  sMP[o] = MP;
}

__global__ void hsl_combine_final(uint32_t ncols, uint32_t nslices,
                                  const long long* sC, const long long* sM, const __int128* sMP,
                                  __int128* out)
{
  uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncols)
    return;
  long long carry = 0;
  __int128 acc = 0;
  for (uint32_t sl = 0; sl < nslices; sl++) {
    size_t o = (size_t) sl * ncols + col;
    acc += sMP[o] + (__int128) carry * sM[o];
    carry += sC[o];
  }
  out[col] = acc;
}

// ------------------------------------------------------------------------
// host-side launch wrappers

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
  const uint8_t wheel[8] = {1, 7, 11, 13, 17, 19, 23, 29};
  uint8_t wb[30];
  for (int r = 0; r < 30; r++) wb[r] = 0xff;
  for (int i = 0; i < 8; i++) wb[wheel[i]] = (uint8_t) i;
  hipError_t e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_cop_index), ci, sizeof(ci))) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_cop), cp, sizeof(cp))) != hipSuccess) return e;
  if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_wheel), wheel, sizeof(wheel))) != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_wheel_bit), wb, sizeof(wb));
}

hipError_t hsl_launch_sieve(const HslArgs& a, bool factor32, uint32_t grid, hipStream_t st)
{
  size_t lds = hsl_lds_bytes(a.W);
  if (factor32)
    hipLaunchKernelGGL(hsl_sieve_kernel<uint32_t>, dim3(grid), dim3(kBlock), lds, st, a);
  else
    hipLaunchKernelGGL(hsl_sieve_kernel<uint16_t>, dim3(grid), dim3(kBlock), lds, st, a);
  return hipGetLastError();
}

hipError_t hsl_sieve_occupancy(bool factor32, uint32_t W, int* blocks_per_cu)
{
  size_t lds = hsl_lds_bytes(W);
  if (factor32)
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, hsl_sieve_kernel<uint32_t>, kBlock, lds);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, hsl_sieve_kernel<uint16_t>, kBlock, lds);
}

hipError_t hsl_launch_fill_zero(void* p, size_t bytes, hipStream_t st)
{
  size_t n16 = (bytes + 15) / 16;
  if (n16 == 0) return hipSuccess;
  uint32_t grid = (uint32_t) std::min<size_t>((n16 + 255) / 256, 4096);
  hipLaunchKernelGGL(hsl_fill_zero, dim3(grid), dim3(256), 0, st, (uint4*) p, n16);
  return hipGetLastError();
}

hipError_t hsl_launch_combine(const CombineCol* cols, uint32_t ncols, const ChunkDesc* chunks,
                              const uint32_t* T_C, const int32_t* T_M, uint32_t slice_len,
                              uint32_t nslices, long long* sC, long long* sM, void* sMP,
                              void* out, hipStream_t st)
{
  if (ncols == 0) return hipSuccess;
  dim3 g1((ncols + 255) / 256, nslices);
  hipLaunchKernelGGL(hsl_combine_slices, g1, dim3(256), 0, st, cols, ncols, chunks, T_C, T_M,
                     slice_len, nslices, sC, sM, (__int128*) sMP);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(hsl_combine_final, dim3((ncols + 255) / 256), dim3(256), 0, st, ncols,
                     nslices, sC, sM, (const __int128*) sMP, (__int128*) out);
  return hipGetLastError();
}

} // namespace pcg
