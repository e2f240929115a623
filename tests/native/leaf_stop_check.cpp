// CPU check of the error analysis behind the sieve kernel's leaf quotient
// (leaf_stop in primecount_amd/csrc/hsl_kernels.hip): for m < 2^32 and a
// quotient floor(xp / m) in [low, low + 2^20), the estimate
//   s = trunc((double) r * rcp(m)),  r = xp - low * m  (mod 2^64)
// followed by one integer fix-up against r - s * m is exact, even though the
// reciprocal is an approximation: the hardware v_rcp_f64 was measured on the
// MI355X to be up to 2.5e8 ulps (a relative error near 2^-24) off 1/m.  The
// reciprocal is modelled as the correctly rounded 1/m times (1 + e), with
// |e| up to 2^-22 (four times the measured error), toward zero and away.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

typedef unsigned __int128 u128;

static uint32_t leaf_stop(uint64_t xlo, uint64_t low, uint32_t m, double rcp)
{
  const uint64_t r = xlo - low * m;
  uint32_t s = (uint32_t) ((double) r * rcp);
  const int64_t t = (int64_t) (r - (uint64_t) s * m);
  return t < 0 ? s - 1 : ((uint64_t) t >= m ? s + 1 : s);
}

static double perturb(double v, double e)
{
  return v * (1.0 + e);
}

int main()
{
  std::mt19937_64 g(20261203);
  const double errs[] = {0.0, 0x1p-52, -0x1p-52, 0x1p-40, -0x1p-40, 0x1p-24, -0x1p-24,
                         0x1p-22, -0x1p-22};
  long bad = 0, n = 0;
  for (int it = 0; it < 400000; it++) {
    const int mb = 1 + (int) (g() % 32);
    const uint32_t m = (uint32_t) ((g() >> (64 - mb)) | 1);
    const uint64_t stop = g() % 983040;                    // < 240 * 4096
    const uint64_t low = g() % (1ull << 50);
    const u128 xp = (u128) (low + stop) * m + g() % m;     // floor(xp / m) = low + stop
    for (double e : errs) {
      n++;
      if (leaf_stop((uint64_t) xp, low, m, perturb(1.0 / (double) m, e)) != stop) {
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
