// CPU check of the error analysis behind the A/C leaf-sum quotient (div49 in
// primecount_amd/csrc/pix_kernels.hip) and of the table-word index computed
// from it: for 1 <= d < 2^32 and a quotient q = floor(a / d) < 2^49,
//   r = rcp(d) refined by two Newton steps, s = trunc(double(a) * r),
//   one fix-up against a - s * d (exact modulo 2^64),
//   w = trunc((q + 1/2) / 240) in FP64,  q - 240 w,
// must give q, q / 240 and q % 240 exactly.  The hardware reciprocal was
// measured on the MI355X to be up to 2.5e8 ulps (relative error near 2^-24)
// off; it is modelled here as 1/d times (1 + e) with |e| up to 2^-20.  The
// tiled kernel's variant (fraction test, 32-bit residual) is checked too.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

typedef unsigned __int128 u128;

static uint64_t bits_of(double x)
{
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return u;
}

static uint64_t div49(uint64_t a, double ad, uint32_t d, double rcp, double* qd)
{
  const double dd = (double) d;
  double r = rcp;
  r = std::fma(r, std::fma(-dd, r, 1.0), r);
  r = std::fma(r, std::fma(-dd, r, 1.0), r);
  const double qf = std::trunc(ad * r);
  const uint64_t q = bits_of(qf + 4503599627370496.0) & ((1ull << 52) - 1);
  const int64_t rem = (int64_t) (a - q * d);
  const int32_t fix = rem < 0 ? -1 : ((uint64_t) rem >= d ? 1 : 0);
  *qd = qf + (double) fix;
  return q + (int64_t) fix;
}

// The tiled kernel's variant (pix_leaftile_kernel): q < 2^30, quotient below
// 2^41; the estimate's integer part is taken as it is when the fraction of
// xd * r lies in [v 2^-46, 1 - v 2^-46], and fixed by the 32-bit residual
// otherwise.  Returns the quotient; *fixed tells whether the fix ran.
static uint64_t tile_quotient(uint64_t a, uint32_t q, double rcp, bool* fixed)
{
  const double qd = (double) q;
  double r = rcp;
  r = std::fma(r, std::fma(-qd, r, 1.0), r);
  r = std::fma(r, std::fma(-qd, r, 1.0), r);
  const double xd = (double) a;
  const double pf = xd * r;
  double vd = std::trunc(pf);
  // the compiler may fuse the fraction into fma(xd, r, -vd): test both
  const double h1 = (pf - vd) - 0.5, h2 = std::fma(xd, r, -vd) - 0.5;
  const double tol = 0.5 - vd * 0x1p-46;
  uint32_t v_lo = (uint32_t) bits_of(vd + 4503599627370496.0);
  *fixed = std::fabs(h1) > tol || std::fabs(h2) > tol;
  if (*fixed) {
    const int32_t rem = (int32_t) ((uint32_t) a - v_lo * q);
    const int32_t fix = rem < 0 ? -1 : ((uint32_t) rem >= q ? 1 : 0);
    vd += (double) fix;
    v_lo += (uint32_t) fix;
  }
  const uint64_t v = bits_of(vd + 4503599627370496.0) & ((1ull << 52) - 1);
  if ((uint32_t) v != v_lo) return ~0ull;
  return v;
}

static long check_tile(std::mt19937_64& g, long* n, long* nfixed)
{
  const double errs[] = {0.0, 0x1p-52, -0x1p-52, 0x1p-24, -0x1p-24, 0x1p-20, -0x1p-20};
  long bad = 0;
  for (int it = 0; it < 600000; it++) {
    const int db = 1 + (int) (g() % 30);
    uint32_t q = (uint32_t) (g() >> (64 - db));
    if (q == 0) q = 1;
    const int qb = 1 + (int) (g() % 41);
    const uint64_t v = (g() >> (64 - qb)) | 7;
    uint64_t rem = g() % q;
    if (it % 4 == 0) rem = 0;
    if (it % 4 == 1) rem = q - 1;
    const u128 a = (u128) v * q + rem;
    if (a >> 64) continue;
    for (double e : errs) {
      (*n)++;
      bool fixed;
      const uint64_t got = tile_quotient((uint64_t) a, q, (1.0 / (double) q) * (1.0 + e), &fixed);
      *nfixed += fixed;
      if (got != v) {
        if (bad < 5)
          std::printf("tile mismatch a=%llu q=%u e=%g\n", (unsigned long long) a, q, e);
        bad++;
      }
    }
  }
  return bad;
}

int main()
{
  std::mt19937_64 g(49);
  const double errs[] = {0.0, 0x1p-52, -0x1p-52, 0x1p-24, -0x1p-24, 0x1p-20, -0x1p-20};
  long bad = 0, n = 0;
  for (int it = 0; it < 600000; it++) {
    const int db = 1 + (int) (g() % 32);
    uint32_t d = (uint32_t) (g() >> (64 - db));
    if (d == 0) d = 1;
    // quotient below 2^49 and a below 2^64
    const u128 qmax = std::min<u128>(((u128) 1 << 49) - 1, (((u128) 1 << 64) - d) / d);
    const int qb = 1 + (int) (g() % 49);
    u128 q = (g() >> (64 - qb)) % (qmax + 1);
    uint64_t rem = g() % d;
    if (it % 4 == 0) rem = 0;
    if (it % 4 == 1) rem = d - 1;
    const uint64_t a = (uint64_t) (q * d + rem);
    for (double e : errs) {
      n++;
      double qd;
      const uint64_t got = div49(a, (double) a, d, (1.0 / (double) d) * (1.0 + e), &qd);
      const double wf = std::trunc(std::fma(qd, 1.0 / 240.0, 0.5 / 240.0));
      const uint64_t w = bits_of(wf + 4503599627370496.0) & ((1ull << 52) - 1);
      const uint32_t rr = (uint32_t) got - 240u * (uint32_t) w;
      if (got != (uint64_t) q || qd != (double) (uint64_t) q || w != (uint64_t) q / 240 ||
          rr != (uint64_t) q % 240) {
        if (bad < 5)
          std::printf("mismatch a=%llu d=%u e=%g\n", (unsigned long long) a, d, e);
        bad++;
      }
    }
  }
  std::printf("%ld cases, %ld mismatches\n", n, bad);
  long nt = 0, nfixed = 0;
  bad += check_tile(g, &nt, &nfixed);
  std::printf("tiled: %ld cases, %ld took the residual fix\n", nt, nfixed);
  std::printf(bad ? "FAIL\n" : "OK\n");
  return bad != 0;
}
