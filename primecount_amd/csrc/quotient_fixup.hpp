// quotient_fixup.hpp -- exact integer quotient from an estimate, shared by
// the HIP kernels and the CPU checks of tests/native.
//
// The functions are constexpr, which HIP's compiler treats as callable from
// host and device code alike, so the CPU check runs the same code as the
// kernel.
#pragma once

#include <stdint.h>

namespace pcg {

/// floor(r / m) from an estimate s of it, for r < 2^63 and m >= 1.  The
/// residual r - s * m (exact modulo 2^64, and as a signed value while it lies
/// within 2^63) moves s one unit at a time until it lies in [0, m).  Each
/// loop is taken only by lanes whose estimate is off, so exactness does not
/// depend on how accurate the estimate is, only the number of steps does.
constexpr uint32_t quotient_fixup(uint64_t r, uint32_t m, uint32_t s)
{
  int64_t t = (int64_t) (r - (uint64_t) s * m);
  while (t < 0) {
    s--;
    t += m;
  }
  while ((uint64_t) t >= m) {
    s++;
    t -= m;
  }
  return s;
}

/// r as a double for r < 2^52, exactly: the bits of 2^52 + r, minus 2^52.
constexpr double u52_to_double(uint64_t r)
{
  return __builtin_bit_cast(double, r | 0x4330000000000000ull) - 4503599627370496.0;
}

/// floor(xp / m) - low for a leaf whose quotient lies in [low, low + 2^20)
/// and m < 2^32, from rcp ~ 1 / m: r = xp - low * m is exact modulo 2^64
/// from the low word of xp because 0 <= r < m * 2^20 < 2^52, and r * rcp is
/// truncated to the estimate that quotient_fixup corrects.
constexpr uint32_t leaf_stop_from_rcp(uint64_t xlo, uint64_t low, uint32_t m, double rcp)
{
  const uint64_t r = xlo - low * m;
  const double e = u52_to_double(r) * rcp;     // >= 0 for a positive reciprocal
  // an estimate beyond the quotient's range (only from a badly wrong
  // reciprocal) is clamped into it, which keeps s * m below 2^52
  const uint32_t s = e >= 1048576.0 ? 1048575u : (uint32_t) e;
  return quotient_fixup(r, m, s);
}

} // namespace pcg
