// tables.cpp -- host table builders (see tables.hpp).  All builders are
// block-parallel over std::thread; each block is independent, so the result
// does not depend on the thread count.

#include "tables.hpp"

#include <algorithm>
#include <cstring>
#include <atomic>
#include <array>
#include <sched.h>
#include <thread>

namespace pcg {

const uint8_t kWheel[8] = {1, 7, 11, 13, 17, 19, 23, 29};

int default_threads()
{
  // the CPUs this process may run on (not the whole host), at most 16: the
  // host's share per GPU
  cpu_set_t set;
  unsigned n = 0;
  if (sched_getaffinity(0, sizeof(set), &set) == 0)
    n = (unsigned) CPU_COUNT(&set);
  if (n == 0)
    n = std::thread::hardware_concurrency();
  return n ? (int) std::min(n, 16u) : 1;
}

namespace {

template <typename F>
void parallel_for(int64_t n, int threads, F&& f)
{
  threads = (int) std::max<int64_t>(1, std::min<int64_t>(threads, n));
  if (threads == 1) {
    for (int64_t i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> pool;
  std::atomic<int64_t> next{0};
  for (int t = 0; t < threads; t++)
    pool.emplace_back([&]() {
      for (int64_t i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto& th : pool) th.join();
}

int wheel_bit_of_residue(int r)
{
  for (int i = 0; i < 8; i++)
    if (kWheel[i] == r) return i;
  return -1;
}

} // namespace

const std::vector<uint16_t>& coprime2310()
{
  static const std::vector<uint16_t> t = [] {
    std::vector<uint16_t> v;
    for (int n = 0; n < 2310; n++)
      if (n % 2 && n % 3 && n % 5 && n % 7 && n % 11) v.push_back((uint16_t) n);
    return v;
  }();
  return t;
}

const std::vector<int16_t>& coprime2310_index()
{
  static const std::vector<int16_t> t = [] {
    std::vector<int16_t> v(2310);
    int idx = -1;
    for (int n = 0; n < 2310; n++) {
      if (n % 2 && n % 3 && n % 5 && n % 7 && n % 11) idx++;
      v[n] = (int16_t) idx;
    }
    return v;
  }();
  return t;
}

namespace {

/// le_masks()[r]: the bits of a wheel-30 word for the offsets <= r.
const std::array<uint64_t, 240>& le_masks()
{
  static const std::array<uint64_t, 240> m = [] {
    std::array<uint64_t, 240> a{};
    for (int r = 0; r < 240; r++)
      for (int j = 0; j < 8; j++)
        for (int i = 0; i < 8; i++)
          if (30 * j + kWheel[i] <= r) a[r] |= 1ull << (8 * j + i);
    return a;
  }();
  return m;
}

} // namespace

uint64_t PrimeTables::pi(uint64_t n) const
{
  if (n < 2) return 0;
  if (n < 7) return n < 3 ? 1 : (n < 5 ? 2 : 3);
  const uint64_t blk = n / 240, r = n % 240;
  return 3 + pi_cnt[blk] + (uint64_t) __builtin_popcountll(pi_bits[blk] & le_masks()[r]);
}

PrimeTables build_prime_tables(uint64_t max, int threads)
{
  if (max >= (1ull << 32) - 1)
    throw pcg_error("prime tables are limited to numbers < 2^32");
  PrimeTables P;
  P.max = std::max<uint64_t>(max, 7);
  uint64_t nwords = P.max / 240 + 1;
  P.pi_bits.assign(nwords, 0);
  P.pi_cnt.assign(nwords + 1, 0);

  // sieving primes <= sqrt(max), by a tiny simple sieve
  uint64_t sq = isqrt(P.max) + 1;
  std::vector<char> small(sq + 1, 1);
  std::vector<uint32_t> sp;
  for (uint64_t i = 2; i <= sq; i++)
    if (small[i]) {
      if (i >= 7) sp.push_back((uint32_t) i);
      for (uint64_t j = i * i; j <= sq; j += i) small[j] = 0;
    }

  // wheel-30 sieve over blocks of words; bit (8j + i) of word w <-> 240w + 30j + kWheel[i]
  const uint64_t block_words = 1 << 14;
  int64_t nblocks = (int64_t) ((nwords + block_words - 1) / block_words);
  std::vector<uint64_t> block_count(nblocks, 0);
  parallel_for(nblocks, threads, [&](int64_t bi) {
    uint64_t w0 = bi * block_words, w1 = std::min<uint64_t>(nwords, w0 + block_words);
    uint8_t* bytes = (uint8_t*) &P.pi_bits[w0];
    uint64_t nbytes = (w1 - w0) * 8;
    std::memset(bytes, 0xff, nbytes);
    uint64_t lo = w0 * 240, hi = w1 * 240;       // numbers [lo, hi)
    for (uint32_t p : sp) {
      if ((uint64_t) p * p >= hi) break;
      // multiples p*q >= max(p*p, lo) with q coprime to 30
      uint64_t qmin = std::max<uint64_t>(p, (lo + p - 1) / p);
      for (int i = 0; i < 8; i++) {
        uint64_t q = qmin - qmin % 30 + kWheel[i];
        if (q < qmin) q += 30;
        for (uint64_t m = p * q; m < hi; m += 30ull * p) {
          uint64_t off = m - lo;
          bytes[off / 30] &= (uint8_t) ~(1u << wheel_bit_of_residue((int) (off % 30)));
        }
      }
    }
    if (w0 == 0) bytes[0] &= 0xfe;               // 1 is not prime
    // clear numbers > max in the last word
    if (w1 == nwords) {
      for (uint64_t off = P.max + 1 - lo; off < nbytes * 30; off++)
        if (off % 2 && off % 3 && off % 5)
          bytes[off / 30] &= (uint8_t) ~(1u << wheel_bit_of_residue((int) (off % 30)));
    }
    uint64_t c = 0;
    for (uint64_t w = w0; w < w1; w++) c += __builtin_popcountll(P.pi_bits[w]);
    block_count[bi] = c;
  });

  // counts and the prime list
  std::vector<uint64_t> block_start(nblocks + 1, 0);
  for (int64_t b = 0; b < nblocks; b++) block_start[b + 1] = block_start[b] + block_count[b];
  uint64_t nprimes = 3 + block_start[nblocks];
  P.primes.assign(nprimes + 1, 0);
  P.primes[1] = 2; P.primes[2] = 3; P.primes[3] = 5;
  parallel_for(nblocks, threads, [&](int64_t bi) {
    uint64_t w0 = bi * block_words, w1 = std::min<uint64_t>(nwords, w0 + block_words);
    uint64_t c = block_start[bi];
    for (uint64_t w = w0; w < w1; w++) {
      P.pi_cnt[w] = (uint32_t) c;
      uint64_t bits = P.pi_bits[w];
      while (bits) {
        int b = __builtin_ctzll(bits);
        bits &= bits - 1;
        P.primes[4 + c] = (uint32_t) (240 * w + 30 * (b >> 3) + kWheel[b & 7]);
        c++;
      }
    }
  });
  P.pi_cnt[nwords] = (uint32_t) block_start[nblocks];
  return P;
}

namespace {

/// Shared factorisation sweep over the numbers coprime to 2310 in [1, limit].
/// For each index it reports (n, pi(lpf), lpf, mpf, number of prime factors,
/// squarefree) to `emit`, in parallel over blocks of 2310 * 256 numbers.
template <typename Emit>
void factor_sweep(int64_t limit, const PrimeTables& P, int threads, Emit&& emit)
{
  const int64_t QB = 256;                          // 2310-blocks per work item
  const int64_t nq = limit / 2310 + 1;
  const int64_t items = (nq + QB - 1) / QB;
  parallel_for(items, threads, [&](int64_t it) {
    int64_t q0 = it * QB, q1 = std::min(nq, q0 + QB);
    int64_t n_lo = q0 * 2310, n_hi = std::min<int64_t>(q1 * 2310, limit + 1);   // [n_lo, n_hi)
    int64_t i0 = q0 * 480;
    int64_t cnt_idx = (q1 - q0) * 480;
    std::vector<uint64_t> rem(cnt_idx);
    std::vector<uint32_t> lpfi(cnt_idx, 0), mpfs(cnt_idx, 0);
    std::vector<uint8_t> nf(cnt_idx, 0), sqf(cnt_idx, 1);
    for (int64_t e = 0; e < cnt_idx; e++) rem[e] = ft_to_number(i0 + e);
    // primes from 13 (pi = 6) while p * p < n_hi
    for (uint64_t j = 6; j < P.primes.size(); j++) {
      uint64_t p = P.primes[j];
      if ((int64_t) (p * p) >= n_hi) break;
      // first cofactor q >= max(2, ceil(n_lo / p)) that is coprime to 2310;
      // q = 1 is skipped so that primes keep lpf index 0
      uint64_t qmin = std::max<uint64_t>(2, ((uint64_t) n_lo + p - 1) / p);
      int64_t qi = qmin == 0 ? 0 : ft_to_index(qmin - 1) + 1;
      for (;; qi++) {
        uint64_t m = p * ft_to_number(qi);
        if ((int64_t) m >= n_hi) break;
        int64_t e = ft_to_index(m) - i0;
        if (lpfi[e] == 0) lpfi[e] = (uint32_t) j;
        mpfs[e] = (uint32_t) p;
        nf[e]++;
        rem[e] /= p;
        if (rem[e] % p == 0) sqf[e] = 0;
      }
    }
    for (int64_t e = 0; e < cnt_idx; e++) {
      int64_t n = (int64_t) ft_to_number(i0 + e);
      if (n > limit) break;
      uint64_t r = rem[e];
      uint64_t mpf = r > 1 ? r : mpfs[e];
      int nfac = nf[e] + (r > 1 ? 1 : 0);
      emit(i0 + e, n, lpfi[e], lpfi[e] ? P.primes[lpfi[e]] : 0, mpf, nfac, sqf[e] != 0);
    }
  });
}

} // namespace

template <typename T>
std::vector<T> build_factor_table_D(int64_t y, int64_t z, const PrimeTables& P, int threads)
{
  const T T_MAX = (T) ~(T) 0;
  if (sizeof(T) == 2 && z > kFactorTableD16Max)
    throw pcg_error("z must be <= FactorTableD::max()");
  z = std::max<int64_t>(1, z);
  if ((uint64_t) isqrt((u128) z) > P.max)
    throw pcg_error("factor table needs primes up to sqrt(z)");
  std::vector<T> f((size_t) ft_to_index(z) + 1);
  factor_sweep(z, P, threads, [&](int64_t idx, int64_t n, uint32_t lpf_pi, uint64_t,
                                  uint64_t mpf, int nfac, bool sqfree) {
    T v;
    if (n == 1) v = (T) (T_MAX - 1);
    else if (lpf_pi == 0) v = n <= y ? T_MAX : (T) 0;          // n is prime
    else if (!sqfree || (int64_t) mpf > y) v = 0;
    else v = (T) (2 * (uint64_t) lpf_pi + (nfac & 1));
    f[idx] = v;
  });
  return f;
}

template <typename T>
std::vector<T> build_factor_table_lpf(int64_t y, const PrimeTables& P, int threads)
{
  const T T_MAX = (T) ~(T) 0;
  if (sizeof(T) == 2 && y > kFactorTable16Max)
    throw pcg_error("y must be <= FactorTable::max()");
  y = std::max<int64_t>(1, y);
  if ((uint64_t) isqrt((u128) y) > P.max)
    throw pcg_error("factor table needs primes up to sqrt(y)");
  std::vector<T> f((size_t) ft_to_index(y) + 1);
  factor_sweep(y, P, threads, [&](int64_t idx, int64_t n, uint32_t lpf_pi, uint64_t lpf,
                                  uint64_t, int nfac, bool sqfree) {
    T v;
    if (n == 1) v = (T) (T_MAX - 1);
    else if (lpf_pi == 0) v = T_MAX;                             // n is prime
    else if (!sqfree) v = 0;
    else v = (T) ((nfac & 1) ? lpf : lpf - 1);
    f[idx] = v;
  });
  return f;
}

template std::vector<uint16_t> build_factor_table_D<uint16_t>(int64_t, int64_t, const PrimeTables&, int);
template std::vector<uint32_t> build_factor_table_D<uint32_t>(int64_t, int64_t, const PrimeTables&, int);
template std::vector<uint16_t> build_factor_table_lpf<uint16_t>(int64_t, const PrimeTables&, int);
template std::vector<uint32_t> build_factor_table_lpf<uint32_t>(int64_t, const PrimeTables&, int);

} // namespace pcg
