// wheel_dev.hpp -- device helpers shared by the HIP kernels: the wheel-30
// sieve word layout (bit 8j + i of a 64-bit word <-> 30 j + {1, 7, 11, 13,
// 17, 19, 23, 29}[i], src/sieve/Sieve.cpp:62-66) and exact division from a
// double estimate.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcg {

/// Nibble table: number of wheel residues {1,7,11,...,29} that are <= r.
constexpr uint64_t le_nibbles(int first)
{
  uint64_t v = 0;
  for (int r = first; r < first + 16 && r < 30; r++) {
    const int w[8] = {1, 7, 11, 13, 17, 19, 23, 29};
    int k = 0;
    for (int i = 0; i < 8; i++)
      k += w[i] <= r;
    v |= (uint64_t) k << (4 * (r - first));
  }
  return v;
}
constexpr uint64_t kLeLo = le_nibbles(0);
constexpr uint64_t kLeHi = le_nibbles(16);

/// Number of wheel residues {1, 7, 11, ..., 29} that are <= r (r < 30).
__device__ __forceinline__ uint32_t le_count(uint32_t r)
{
  return (uint32_t) ((r < 16 ? (kLeLo >> (4 * r)) : (kLeHi >> (4 * (r - 16)))) & 15);
}

/// Bits of a sieve word that stand for the offsets <= r (r < 240).
__device__ __forceinline__ uint64_t le_mask(uint32_t r)
{
  uint32_t j = r / 30;
  uint32_t rem = r - 30 * j;
  uint32_t k = (uint32_t) ((rem < 16 ? (kLeLo >> (4 * rem)) : (kLeHi >> (4 * (rem - 16)))) & 15);
  uint64_t below = (1ull << (8 * j)) - 1;
  return below | ((((uint64_t) 1 << k) - 1) << (8 * j));
}

/// Residues coprime to 30 in bit order, packed as bytes: 1, 7, 11, ..., 29.
constexpr uint64_t kWheelBytes = 0x1D1713110D0B0701ull;

/// Bit r set for the residues r in [0, 30) coprime to 30.
constexpr uint32_t kCop30 = (1u << 1) | (1u << 7) | (1u << 11) | (1u << 13) |
                            (1u << 17) | (1u << 19) | (1u << 23) | (1u << 29);

/// Bit (0..7) of a wheel residue r in a sieve byte.
__device__ __forceinline__ uint32_t wheel_bit(uint32_t r)
{
  uint32_t k = (uint32_t) ((r < 16 ? (kLeLo >> (4 * r)) : (kLeHi >> (4 * (r - 16)))) & 15);
  return k - 1;
}

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

/// floor(n / d) for n < 2^53 (nd = n, exact) and a quotient below 2^48,
/// with inv_d within a relative 2^-51 of 1 / d: the product is within 0.25
/// of n / d and the residual, exact because |n - q d| < 2 d, fixes it.
__device__ __forceinline__ uint64_t div_exact_r(double nd, uint64_t n, uint64_t d, double inv_d)
{
  uint64_t q = (uint64_t) (nd * inv_d);
  int64_t r = (int64_t) (n - q * d);
  if (r < 0)
    q--;
  else if ((uint64_t) r >= d)
    q++;
  return q;
}

} // namespace pcg
