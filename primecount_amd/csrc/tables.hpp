// tables.hpp -- host builders for the lookup tables the hard-special-leaf
// kernels stream from HBM:
//   - primes[] (1-indexed, primes[0] = 0), as the reference's
//     PiTable::get_primes (src/PiTable.cpp:215)
//   - a prime-counting table, 240 numbers per (count, bits) pair, answering
//     pi(n) in O(1) like PiTable::operator[] (src/PiTable.hpp:46-57); the bit
//     layout is the wheel-30 layout of the device sieve
//   - FactorTableD<T>: mu/lpf/mpf table of src/gourdon/FactorTableD.hpp:20-39
//   - FactorTable<T>:  mu/lpf table of src/FactorTable.hpp:138-150
// The factor tables keep the reference's encodings and its index map over the
// numbers coprime to 2*3*5*7*11 (src/BaseFactorTable.hpp:36-49), because the
// leaf filter "factor[m] > threshold" is defined on those encodings.
#pragma once

#include "pc_math.hpp"

#include <vector>

namespace pcg {

/// Wheel-30 residues in bit order: bit i of a sieve byte is 30*j + kWheel[i].
extern const uint8_t kWheel[8];

/// Residues coprime to 2310 in [0, 2310), ascending (480 values, first is 1).
const std::vector<uint16_t>& coprime2310();
/// coprime2310_index()[r] = index of the largest coprime residue <= r (-1 for 0).
const std::vector<int16_t>& coprime2310_index();

/// to_index / to_number of src/BaseFactorTable.hpp:36-49.
inline int64_t ft_to_index(uint64_t n)
{
  return 480 * (int64_t) (n / 2310) + coprime2310_index()[n % 2310];
}
inline uint64_t ft_to_number(uint64_t idx)
{
  return 2310 * (idx / 480) + coprime2310()[idx % 480];
}

struct PrimeTables {
  uint64_t max = 0;                 // tables cover [0, max]
  std::vector<uint32_t> primes;     // primes[0] = 0, primes[1] = 2, ...
  std::vector<uint64_t> pi_bits;    // per 240 numbers: primes >= 7, wheel-30 bits
  std::vector<uint32_t> pi_cnt;     // primes >= 7 below 240*i
  uint64_t pi(uint64_t n) const;    // number of primes <= n, n <= max
};

/// Sieve primes up to max (max < 2^32) with `threads` host threads.
PrimeTables build_prime_tables(uint64_t max, int threads);

/// FactorTableD<T>(y, z): entries for to_number(i), i in [0, to_index(z)].
/// 1 -> T_MAX-1; prime <= y -> T_MAX; mu = 0 or a prime factor > y -> 0;
/// otherwise 2*pi(lpf) + [mu = -1].
template <typename T>
std::vector<T> build_factor_table_D(int64_t y, int64_t z, const PrimeTables& P, int threads);

/// FactorTable<T>(y): 1 -> T_MAX-1; prime -> T_MAX; mu = 0 -> 0;
/// otherwise lpf - 1 (mu = +1) or lpf (mu = -1).
template <typename T>
std::vector<T> build_factor_table_lpf(int64_t y, const PrimeTables& P, int threads);

/// Largest z a 16-bit FactorTableD can hold (src/gourdon/FactorTableD.hpp:248-256).
constexpr int64_t kFactorTableD16Max = 386083LL * 386083 - 1;
/// Largest y a 16-bit FactorTable can hold (src/FactorTable.hpp:196-203).
constexpr int64_t kFactorTable16Max = 65534LL * 65534 - 1;

int default_threads();

} // namespace pcg
