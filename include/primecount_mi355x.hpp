// primecount_mi355x.hpp -- C++ overloads with the reference's signatures,
// header-only, over the C ABI of primecount_mi355x.h.
//
//   int64_t  D(int64_t x, int64_t y, int64_t z, int64_t k, int threads, bool print)
//   int128_t D(int128_t x, ...)            include/gourdon.hpp:28,38
//   int64_t  S2_hard(int64_t x, int64_t y, int64_t z, int64_t c, int threads, bool print)
//   int128_t S2_hard(int128_t x, ...)      include/S.hpp:25,32
//
// Errors throw primecount_mi355x::primecount_error, as the reference throws
// primecount_error (include/primecount.hpp:26-32).  `threads` and `print`
// are accepted for signature compatibility; the device decides its own
// parallelism.
#ifndef PRIMECOUNT_MI355X_HPP
#define PRIMECOUNT_MI355X_HPP

#include "primecount_mi355x.h"

#include <stdexcept>
#include <stdint.h>
#include <string>

namespace primecount_mi355x {

typedef __int128 int128_t;

class primecount_error : public std::runtime_error {
public:
  explicit primecount_error(const std::string& msg) : std::runtime_error(msg) {}
};

namespace detail {

inline void check(int rc)
{
  if (rc != PCG_OK)
    throw primecount_error(std::string("primecount_mi355x: ") + pcg_last_error());
}

inline pcg_i128 to_abi(int128_t v)
{
  pcg_i128 r;
  r.lo = (uint64_t) v;
  r.hi = (int64_t) (v >> 64);
  return r;
}

inline int128_t from_abi(pcg_i128 v)
{
  return (int128_t) (((unsigned __int128) (uint64_t) v.hi << 64) | v.lo);
}

} // namespace detail

inline int64_t D(int64_t x, int64_t y, int64_t z, int64_t k, int threads = 0, bool is_print = false)
{
  (void) threads; (void) is_print;
  int64_t r = 0;
  detail::check(pcg_D_i64(x, y, z, k, &r));
  return r;
}

inline int128_t D(int128_t x, int64_t y, int64_t z, int64_t k, int threads = 0, bool is_print = false)
{
  (void) threads; (void) is_print;
  pcg_i128 r;
  detail::check(pcg_D_i128(detail::to_abi(x), y, z, k, &r));
  return detail::from_abi(r);
}

inline int64_t S2_hard(int64_t x, int64_t y, int64_t z, int64_t c, int threads = 0, bool is_print = false)
{
  (void) threads; (void) is_print;
  int64_t r = 0;
  detail::check(pcg_S2_hard_i64(x, y, z, c, &r));
  return r;
}

inline int128_t S2_hard(int128_t x, int64_t y, int64_t z, int64_t c, int threads = 0, bool is_print = false)
{
  (void) threads; (void) is_print;
  pcg_i128 r;
  detail::check(pcg_S2_hard_i128(detail::to_abi(x), y, z, c, &r));
  return detail::from_abi(r);
}

} // namespace primecount_mi355x

#endif
