// primecount_mi355x.hpp -- C++ overloads with the reference's signatures,
// header-only, over the C ABI of primecount_mi355x.h.
//
//   int64_t  D(int64_t x, int64_t y, int64_t z, int64_t k, int threads, bool print)
//   int128_t D(int128_t x, ...)            include/gourdon.hpp:28,38
//   int64_t  S2_hard(int64_t x, int64_t y, int64_t z, int64_t c, int threads, bool print)
//   int128_t S2_hard(int128_t x, ...)      include/S.hpp:25,32
//   pi(x), nth_prime(n)                    include/primecount.hpp:57,74,84
//   AC, B, Phi0, Sigma                     include/gourdon.hpp:24-27,34-37
//   S1, S2_trivial, S2_easy                include/S.hpp:22-24,29-31
//   P2, pi_deleglise_rivat                 include/primecount-internal.hpp:33,43,59,61
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

/// A 128-bit result of a 64-bit overload (fits by the domain of x).
inline int64_t narrow(pcg_i128 v)
{
  const int128_t r = from_abi(v);
  if (r < INT64_MIN || r > INT64_MAX)
    throw primecount_error("primecount_mi355x: result does not fit in int64_t");
  return (int64_t) r;
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

/* pi(x) by Gourdon's algorithm and the n-th prime, every term on the device */

inline int64_t pi(int64_t x)
{
  pcg_i128 r;
  detail::check(pcg_pi(detail::to_abi(x), &r));
  return detail::narrow(r);
}

inline int128_t pi(int128_t x)
{
  pcg_i128 r;
  detail::check(pcg_pi(detail::to_abi(x), &r));
  return detail::from_abi(r);
}

inline int64_t nth_prime(int64_t n)
{
  int64_t r = 0;
  detail::check(pcg_nth_prime(n, &r));
  return r;
}

inline int128_t nth_prime(int128_t n)
{
  pcg_i128 r;
  detail::check(pcg_nth_prime_128(detail::to_abi(n), &r));
  return detail::from_abi(r);
}

/* The other Gourdon terms (include/gourdon.hpp:24-27, 34-37) */

#define PRIMECOUNT_MI355X_TERM(T, NAME, PARAMS, ARGS, OUT)                         \
  inline T NAME PARAMS                                                            \
  {                                                                               \
    (void) threads; (void) is_print;                                              \
    pcg_i128 r;                                                                   \
    detail::check(pcg_##NAME ARGS);                                               \
    return OUT(r);                                                                \
  }

PRIMECOUNT_MI355X_TERM(int64_t, AC, (int64_t x, int64_t y, int64_t z, int64_t k, int threads = 0,
                       bool is_print = false), (detail::to_abi(x), y, z, k, &r), detail::narrow)
PRIMECOUNT_MI355X_TERM(int128_t, AC, (int128_t x, int64_t y, int64_t z, int64_t k,
                       int threads = 0, bool is_print = false),
                       (detail::to_abi(x), y, z, k, &r), detail::from_abi)
PRIMECOUNT_MI355X_TERM(int64_t, B, (int64_t x, int64_t y, int threads = 0, bool is_print = false),
                       (detail::to_abi(x), y, &r), detail::narrow)
PRIMECOUNT_MI355X_TERM(int128_t, B, (int128_t x, int64_t y, int threads = 0,
                       bool is_print = false), (detail::to_abi(x), y, &r), detail::from_abi)
PRIMECOUNT_MI355X_TERM(int64_t, Phi0, (int64_t x, int64_t y, int64_t z, int64_t k,
                       int threads = 0, bool is_print = false),
                       (detail::to_abi(x), y, z, k, &r), detail::narrow)
PRIMECOUNT_MI355X_TERM(int128_t, Phi0, (int128_t x, int64_t y, int64_t z, int64_t k,
                       int threads = 0, bool is_print = false),
                       (detail::to_abi(x), y, z, k, &r), detail::from_abi)
PRIMECOUNT_MI355X_TERM(int64_t, Sigma, (int64_t x, int64_t y, int threads = 0,
                       bool is_print = false), (detail::to_abi(x), y, &r), detail::narrow)
PRIMECOUNT_MI355X_TERM(int128_t, Sigma, (int128_t x, int64_t y, int threads = 0,
                       bool is_print = false), (detail::to_abi(x), y, &r), detail::from_abi)

/* The Deleglise-Rivat terms (include/S.hpp:22-24, 29-31) */

PRIMECOUNT_MI355X_TERM(int64_t, S1, (int64_t x, int64_t y, int64_t c, int threads = 0,
                       bool is_print = false), (detail::to_abi(x), y, c, &r), detail::narrow)
PRIMECOUNT_MI355X_TERM(int128_t, S1, (int128_t x, int64_t y, int64_t c, int threads = 0,
                       bool is_print = false), (detail::to_abi(x), y, c, &r), detail::from_abi)
PRIMECOUNT_MI355X_TERM(int64_t, S2_trivial, (int64_t x, int64_t y, int64_t z, int64_t c,
                       int threads = 0, bool is_print = false),
                       (detail::to_abi(x), y, z, c, &r), detail::narrow)
PRIMECOUNT_MI355X_TERM(int128_t, S2_trivial, (int128_t x, int64_t y, int64_t z, int64_t c,
                       int threads = 0, bool is_print = false),
                       (detail::to_abi(x), y, z, c, &r), detail::from_abi)
PRIMECOUNT_MI355X_TERM(int64_t, S2_easy, (int64_t x, int64_t y, int64_t z, int64_t c,
                       int threads = 0, bool is_print = false),
                       (detail::to_abi(x), y, z, c, &r), detail::narrow)
PRIMECOUNT_MI355X_TERM(int128_t, S2_easy, (int128_t x, int64_t y, int64_t z, int64_t c,
                       int threads = 0, bool is_print = false),
                       (detail::to_abi(x), y, z, c, &r), detail::from_abi)

#undef PRIMECOUNT_MI355X_TERM

/// P2(x, y, a) of include/primecount-internal.hpp:43, 61; the reference
/// requires a == pi(y), which the device recomputes.
inline int64_t P2(int64_t x, int64_t y, int64_t a, int threads = 0, bool is_print = false)
{
  (void) a; (void) threads; (void) is_print;
  pcg_i128 r;
  detail::check(pcg_P2(detail::to_abi(x), y, &r));
  return detail::narrow(r);
}

inline int128_t P2(int128_t x, int64_t y, int64_t a, int threads = 0, bool is_print = false)
{
  (void) a; (void) threads; (void) is_print;
  pcg_i128 r;
  detail::check(pcg_P2(detail::to_abi(x), y, &r));
  return detail::from_abi(r);
}

/// pi_deleglise_rivat(x, threads), include/primecount-internal.hpp:33, 59.
inline int64_t pi_deleglise_rivat(int64_t x, int threads = 0)
{
  (void) threads;
  pcg_i128 r;
  detail::check(pcg_pi_deleglise_rivat(detail::to_abi(x), &r));
  return detail::narrow(r);
}

inline int128_t pi_deleglise_rivat(int128_t x, int threads = 0)
{
  (void) threads;
  pcg_i128 r;
  detail::check(pcg_pi_deleglise_rivat(detail::to_abi(x), &r));
  return detail::from_abi(r);
}

} // namespace primecount_mi355x

#endif
