// verify.cpp -- plausibility check of a pi(x) result, as the reference runs
// after every pi(x) (verify_pix, src/util.cpp:397-420, called at
// src/gourdon/pi_gourdon.cpp:89,160 and
// src/deleglise-rivat/pi_deleglise_rivat.cpp:92,138).
//
// Schoenfeld's bound (Math. Comp. 30, 1976): under the Riemann hypothesis
// |pi(x) - Li(x)| < sqrt(x) log(x) / (8 pi) for x >= 2657.  A result outside
// it is wrong; the reference aborts, here the call fails with a status.

#include "pix.hpp"

#include <cmath>
#include <string>

namespace pcg {

/// The logarithmic integral li(x) for x > 1 by Ramanujan's series
///   li(x) = gamma + ln ln x
///           + sqrt(x) sum_{n>=1} (-1)^(n-1) (ln x)^n / (n! 2^(n-1))
///                     sum_{k=0}^{floor((n-1)/2)} 1 / (2k + 1),
/// in long double.  Its terms peak near n = ln(x) / 2 and alternate; at
/// x = 1e22 the largest is about 50 times li(x), so the sum keeps about
/// 60 of the 64 bits: far more than the check needs.
long double log_integral(long double x)
{
  const long double gamma = 0.577215664901532860606512090082402431L;
  const long double lx = logl(x);
  long double t = lx, inner = 0, sum = 0;     // t = (-1)^(n-1) (ln x)^n / (n! 2^(n-1))
  for (int n = 1; n < 1000; n++) {
    if (n > 1) t *= -lx / (2.0L * n);
    if (n & 1) inner += 1.0L / n;
    const long double add = t * inner;
    sum += add;
    if (n > lx && fabsl(add) < 1e-22L * fabsl(sum)) break;
  }
  return gamma + logl(lx) + sqrtl(x) * sum;
}

/// The offset logarithmic integral Li(x) = li(x) - li(2).
long double offset_log_integral(long double x)
{
  const long double li2 = 1.045163780117492784844588889194613136L;
  return log_integral(x) - li2;
}

bool pix_plausible(i128 x, i128 pix)
{
  if (x < 2657) return true;
  const long double xl = (long double) x;
  const long double Li = offset_log_integral(xl);
  const long double bound = sqrtl(xl) * logl(xl) / (8.0L * 3.14159265358979323846L);
  return fabsl((long double) pix - Li) < bound;
}

void verify_pix(const char* name, i128 x, i128 pix)
{
  if (pix_plausible(x, pix)) return;
  throw pcg_error(std::string("result check failed: ") + name + "(" + to_string128(x) + ") = " +
                      to_string128(pix) + " is outside |pi(x) - Li(x)| < sqrt(x) log(x) / (8 pi)",
                  kErrInternal);
}

} // namespace pcg
