// pc_math.hpp -- integer helpers for the host side of the hard-special-leaf
// engine: 128-bit types, exact floor roots, ceil division.
#pragma once

#include <stdint.h>
#include <cmath>
#include <stdexcept>
#include <string>

namespace pcg {

typedef __int128 i128;
typedef unsigned __int128 u128;

/// Status codes carried by pcg_error; the values are the PCG_E_* codes of the
/// C ABI (include/primecount_mi355x.h).
enum : int {
  kErrInvalid = -1,    // argument outside the supported domain
  kErrDevice = -2,     // HIP runtime / device error
  kErrNoMem = -3,      // device or host allocation failed
  kErrInternal = -4,   // broken internal invariant
};

/// Error type thrown by the C++ layer; the C ABI returns its status().
/// Mirrors primecount_error (include/primecount.hpp:26-32), with the kind of
/// failure kept as a code instead of being guessed from the message.
class pcg_error : public std::runtime_error {
public:
  explicit pcg_error(const std::string& msg, int status = kErrInvalid)
      : std::runtime_error(msg), status_(status) {}
  int status() const { return status_; }

private:
  int status_;
};

/// floor(sqrt(n)) for n < 2^127, exact.
inline uint64_t isqrt(u128 n)
{
  uint64_t r = (uint64_t) std::sqrt((long double) n);
  while (r > 0 && (u128) r * r > n) r--;
  while ((u128) (r + 1) * (r + 1) <= n) r++;
  return r;
}

/// floor(n^(1/N)) for N in {3, 4, 6}, exact (checked by integer powers).
template <int N>
inline uint64_t iroot(u128 n)
{
  if (n == 0) return 0;
  long double est = std::pow((long double) n, 1.0L / N);
  uint64_t r = (uint64_t) est;
  auto pw = [](uint64_t v) {
    u128 p = 1;
    for (int i = 0; i < N; i++) {
      if (p > (~(u128) 0) / (v ? v : 1)) return ~(u128) 0;   // saturate
      p *= v;
    }
    return p;
  };
  while (r > 0 && pw(r) > n) r--;
  while (pw(r + 1) <= n) r++;
  return r;
}

inline i128 ceil_div(i128 a, i128 b) { return (a + b - 1) / b; }

template <typename T>
inline T in_between(T lo, T x, T hi)
{
  if (x < lo || hi < lo) return lo;
  if (x > hi) return hi;
  return x;
}

std::string to_string128(i128 v);

} // namespace pcg
