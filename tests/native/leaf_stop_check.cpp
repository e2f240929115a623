// CPU check of the sieve kernel's leaf quotient (leaf_stop in
// primecount_amd/csrc/hsl_kernels.hip, which calls leaf_stop_from_rcp of
// primecount_amd/csrc/quotient_fixup.hpp -- the same code is compiled here):
// for m < 2^32 and a quotient floor(xp / m) in [low, low + 2^20), the
// estimate trunc((double) r * rcp(m)), r = xp - low * m (mod 2^64), followed
// by the integer fix-up loops is exact.  The hardware v_rcp_f64 was measured
// on the MI355X up to 2.5e8 ulps (a relative error near 2^-24) off 1 / m;
// the reciprocal is modelled as the correctly rounded 1 / m times (1 + e),
// with |e| from 2^-52 up to 2^-18 (64 times the measured error), toward zero
// and away, and, to show that exactness does not rest on any error bound,
// also 2^-10 and 2^-4, where the loops take many steps.
#include <cstdint>
#include <cstdio>
#include <random>

#include "../../primecount_amd/csrc/quotient_fixup.hpp"

typedef unsigned __int128 u128;

int main()
{
  std::mt19937_64 g(20261203);
  const double errs[] = {0.0,     0x1p-52, -0x1p-52, 0x1p-40, -0x1p-40, 0x1p-24, -0x1p-24,
                         0x1p-22, -0x1p-22, 0x1p-20, -0x1p-20, 0x1p-18, -0x1p-18,
                         0x1p-10, -0x1p-10, 0x1p-4,  -0x1p-4};
  long bad = 0, n = 0;
  for (int it = 0; it < 300000; it++) {
    const int mb = 1 + (int) (g() % 32);
    const uint32_t m = (uint32_t) ((g() >> (64 - mb)) | 1);
    // quotients < 240 * 4096 (the largest segment), often near its ends
    uint64_t stop = g() % 983040;
    if (it % 4 == 1) stop = g() % 64;
    if (it % 4 == 2) stop = 983039 - g() % 64;
    const uint64_t low = g() % (1ull << 50);
    const uint64_t rem = it % 3 == 0 ? 0 : (it % 3 == 1 ? m - 1 : g() % m);
    const u128 xp = (u128) (low + stop) * m + rem;         // floor(xp / m) = low + stop
    for (double e : errs) {
      n++;
      const double rcp = (1.0 / (double) m) * (1.0 + e);
      if (pcg::leaf_stop_from_rcp((uint64_t) xp, low, m, rcp) != stop) {
        if (bad < 5)
          std::printf("mismatch m=%u stop=%llu e=%g\n", m, (unsigned long long) stop, e);
        bad++;
      }
    }
  }
  std::printf("%ld cases, %ld mismatches\n", n, bad);
  std::printf(bad ? "FAIL\n" : "OK\n");
  return bad != 0;
}
