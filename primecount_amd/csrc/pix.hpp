// pix.hpp -- Gourdon's prime counting function pi(x) on the device:
//   pi(x) = A - B + C + D + Phi0 + Sigma   (src/gourdon/pi_gourdon.cpp:82-87)
// with D from the hard-special-leaf engine and the other terms from the
// kernels of pix_kernels.hip over a prime-count table resident in HBM.
#pragma once

#include "engine.hpp"

namespace pcg {

enum class GourdonTerm { AC, B, Phi0, Sigma };

struct PixStats {
  double seconds_total = 0, seconds_table = 0, seconds_sigma = 0, seconds_phi0 = 0,
         seconds_ac = 0, seconds_b = 0, seconds_d = 0;
  uint64_t ac_leaves = 0;     // A and C terms summed
  uint64_t table_limit = 0;   // prime-count table covers [0, table_limit]
  uint64_t table_windows = 0; // windows the table was built in (1 = whole)
};

/// One term with explicit parameters (src/gourdon/AC.cpp, B.cpp, Phi0.cpp,
/// Sigma.cpp signatures of include/gourdon.hpp:24-27,34-37).  Throws pcg_error.
i128 gourdon_term(GourdonTerm t, i128 x, int64_t y, int64_t z, int64_t k, PixStats* st);

/// pi(x) with the reference's default parameters.  Throws pcg_error.
i128 pi_gourdon(i128 x, PixStats* st);

/// The Deleglise-Rivat terms other than S2_hard, on the device with the
/// S2_hard tables (src/S1.cpp, src/deleglise-rivat/S2_trivial.cpp,
/// S2_easy.cpp, src/P2.cpp).
enum class DRTerm { S1, S2_trivial, S2_easy, P2 };
i128 dr_term(DRTerm term, i128 x, int64_t y, int64_t z, int64_t c);

/// pi(x) by the Deleglise-Rivat algorithm with the reference's parameters
/// (src/deleglise-rivat/pi_deleglise_rivat.cpp:40-53, 66-92).
i128 pi_deleglise_rivat(i128 x);

/// pi(x) for many x: the terms other than D query by query over shared
/// resident tables, D for all queries in one persistent launch.
/// max_slots bounds the slot arrays of a D pass (0 = engine setting).
std::vector<i128> pi_gourdon_batch(const std::vector<i128>& xs, PixStats* st,
                                   uint64_t max_slots = 0);

/// The n-th prime, n <= 216289611853439384 (primes below 2^63): pi(x0) at
/// x0 = R^-1(n), then a sieve of the interval between x0 and the answer on
/// the device (src/nth_prime.cpp:95-145, nth_prime_sieve.cpp).  Throws
/// pcg_error.
uint64_t nth_prime(uint64_t n, PixStats* st);

/// The n-th prime for any n >= 1 whose answer lies in pi(x)'s domain (x up
/// to about 1e23): the 64-bit path up to n = 216289611853439384, beyond it
/// pi at x0 = R^-1(n) and a sieve of the interval from x0 anchored at a
/// 128-bit base (src/nth_prime.cpp:141-175).  Throws pcg_error.
i128 nth_prime_128(i128 n, PixStats* st);

/// nth_prime for many n (pi at the starting points as one pi_gourdon_batch).
std::vector<uint64_t> nth_prime_batch(const std::vector<uint64_t>& ns, PixStats* st,
                                      uint64_t max_slots = 0);

/// Throws pcg_error (kErrInvalid) unless pi(x) is inside the device path's
/// domain (x >= 0; the parameter checks and the prime-count table limit).
void pi_validate(i128 x);

/// Estimated device bytes of the tables and per-query scratch that pi(x),
/// nth_prime(n) and a D query with tables (y, z) need: the sizes a serving
/// context admits work by.
uint64_t pi_footprint(i128 x);
uint64_t nth_prime_footprint(uint64_t n);
uint64_t table_bytes_D(int64_t y, int64_t z);

/// Li(x) = li(x) - li(2), the offset logarithmic integral (long double).
long double offset_log_integral(long double x);

/// Whether pi(x) = pix lies within Schoenfeld's bound |pi(x) - Li(x)| <
/// sqrt(x) log(x) / (8 pi) (always true for x < 2657), the check the
/// reference runs after every pi(x) (verify_pix, src/util.cpp:397-420).
bool pix_plausible(i128 x, i128 pix);

/// Throws pcg_error (kErrInternal) naming the function if pix fails the check.
void verify_pix(const char* name, i128 x, i128 pix);

} // namespace pcg
