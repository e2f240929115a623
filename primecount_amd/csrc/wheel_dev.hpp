// wheel_dev.hpp -- device helpers shared by the HIP kernels: the wheel-30
// sieve word layout (bit 8j + i of a 64-bit word <-> 30 j + {1, 7, 11, 13,
// 17, 19, 23, 29}[i], src/sieve/Sieve.cpp:62-66) and exact division from a
// double estimate.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcg {

/// Bit r set for the residues r in [0, 30) coprime to 30.
constexpr uint32_t kCop30 = (1u << 1) | (1u << 7) | (1u << 11) | (1u << 13) |
                            (1u << 17) | (1u << 19) | (1u << 23) | (1u << 29);

/// Number of wheel residues {1, 7, 11, ..., 29} that are <= r (r < 30).
__device__ __forceinline__ uint32_t le_count(uint32_t r)
{
  return (uint32_t) __popc(kCop30 & ((2u << r) - 1u));
}

/// Bits of a sieve word that stand for the offsets <= r (r < 240): the low
/// 8 (r / 30) + le_count(r % 30) bits.
__device__ __forceinline__ uint64_t le_mask(uint32_t r)
{
  const uint32_t j = r / 30;
  const uint32_t nb = 8 * j + le_count(r - 30 * j);
  return nb >= 64 ? ~0ull : (1ull << nb) - 1;
}

/// Residues coprime to 30 in bit order, packed as bytes: 1, 7, 11, ..., 29.
constexpr uint64_t kWheelBytes = 0x1D1713110D0B0701ull;

/// Bit (0..7) of a wheel residue r in a sieve byte: the residues below r.
__device__ __forceinline__ uint32_t wheel_bit(uint32_t r)
{
  return (uint32_t) __popc(kCop30 & ((1u << r) - 1u));
}

/// The i-th wheel residue (i < 8): byte i of kWheelBytes.
__device__ __forceinline__ uint32_t wheel_residue(uint32_t i)
{
  return (uint32_t) ((kWheelBytes >> (8 * i)) & 0xff);
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
