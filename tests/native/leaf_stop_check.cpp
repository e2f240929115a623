// CPU check of the error analysis behind the sieve kernel's leaf quotient
// (leaf_stop in primecount_amd/csrc/hsl_kernels.hip): for m < 2^32 and a
// quotient floor(xp / m) in [low, low + 2^20), the estimate
//   s = trunc((double) r * rcp(m)),  r = xp - low * m  (mod 2^64)
// followed by one integer fix-up against r - s * m is exact, even when the
// reciprocal is several ulps off (the hardware v_rcp_f64 is an approximation).
// The reciprocal is modelled as the correctly rounded 1/m moved k ulps toward
// zero and away from zero, k up to 4096 (relative error up to ~2^-40).
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

static double perturb(double v, int ulps)
{
  for (int i = 0; i < (ulps < 0 ? -ulps : ulps); i++)
    v = std::nextafter(v, ulps < 0 ? 0.0 : 1e300);
  return v;
}

int main()
{
  std::mt19937_64 g(20261203);
  const int ulps[] = {0, 1, -1, 64, -64, 4096, -4096};
  long bad = 0, n = 0;
  for (int it = 0; it < 400000; it++) {
    const int mb = 1 + (int) (g() % 32);
    const uint32_t m = (uint32_t) ((g() >> (64 - mb)) | 1);
    const uint64_t stop = g() % 983040;                    // < 240 * 4096
    const uint64_t low = g() % (1ull << 50);
    const u128 xp = (u128) (low + stop) * m + g() % m;     // floor(xp / m) = low + stop
    for (int u : ulps) {
      n++;
      if (leaf_stop((uint64_t) xp, low, m, perturb(1.0 / (double) m, u)) != stop) {
        if (bad < 5)
          std::printf("mismatch m=%u stop=%llu ulps=%d\n", m, (unsigned long long) stop, u);
        bad++;
      }
    }
  }
  std::printf("%ld cases, %ld mismatches\n", n, bad);
  std::printf(bad ? "FAIL\n" : "OK\n");
  return bad != 0;
}
