// Host table builders vs the arithmetic definitions (CPU only).
// Mirrors the checks of the reference's test/gourdon/FactorTableD.cpp:66-145,
// test/PiTable.cpp and test/generate_lpf.cpp, against a plain sieve.
#include "../../primecount_amd/csrc/tables.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace pcg;

static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { if (fails++ < 20) { std::printf(__VA_ARGS__); std::printf("\n"); } } } while (0)

struct Arith {
  std::vector<int> lpf, mpf, mu, pi;
  explicit Arith(int n) : lpf(n + 1, 0), mpf(n + 1, 0), mu(n + 1, 1), pi(n + 1, 0) {
    lpf[1] = mpf[1] = 1;
    for (int p = 2; p <= n; p++) {
      if (lpf[p]) continue;
      for (int m = p; m <= n; m += p) {
        if (!lpf[m]) lpf[m] = p;
        mpf[m] = p;
        mu[m] = -mu[m];
      }
      for (long long m = (long long) p * p; m <= n; m += (long long) p * p) mu[m] = 0;
    }
    for (int i = 2; i <= n; i++) pi[i] = pi[i - 1] + (lpf[i] == i);
  }
};

int main()
{
  const int N = 3000000;
  Arith A(N);
  for (int threads : {1, 5}) {
    PrimeTables P = build_prime_tables(N, threads);
    for (int n = 0; n <= N; n++)
      CHECK(P.pi(n) == (uint64_t) A.pi[n], "pi(%d) = %llu != %d", n, (unsigned long long) P.pi(n), A.pi[n]);
    CHECK(P.primes[0] == 0 && P.primes[1] == 2, "primes[0..1]");
    for (size_t i = 1; i < P.primes.size(); i++)
      CHECK(A.lpf[P.primes[i]] == (int) P.primes[i], "primes[%zu] = %u not prime", i, P.primes[i]);
    CHECK(P.primes.size() == (size_t) A.pi[N] + 1, "prime count");

    // to_index / to_number (src/BaseFactorTable.hpp:36-49)
    for (uint64_t n = 1; n < 50000; n++) {
      int64_t i = ft_to_index(n);
      uint64_t m = ft_to_number((uint64_t) i);
      CHECK(m <= n && (m == 1 || (m % 2 && m % 3 && m % 5 && m % 7 && m % 11)), "to_index(%llu)", (unsigned long long) n);
      CHECK(ft_to_number((uint64_t) i + 1) > n, "to_index(%llu) not the largest", (unsigned long long) n);
    }

    // FactorTableD<uint16_t>(y, z): y == z and y < z
    for (auto yz : std::vector<std::pair<int, int>>{{2619, 2619}, {1001, 31031}, {6029, 99130}, {207, 1001}}) {
      int y = yz.first, z = yz.second;
      auto f = build_factor_table_D<uint16_t>(y, z, P, threads);
      CHECK(f.size() == (size_t) ft_to_index(z) + 1, "FactorTableD size");
      for (size_t i = 0; i < f.size(); i++) {
        int n = (int) ft_to_number(i);
        int v;
        if (n == 1) v = 65534;
        else if (A.lpf[n] == n) v = n <= y ? 65535 : 0;
        else if (A.mu[n] == 0 || A.mpf[n] > y) v = 0;
        else v = 2 * A.pi[A.lpf[n]] + (A.mu[n] == -1);
        CHECK(f[i] == v, "FactorTableD(%d,%d)[%d] = %d != %d", y, z, n, f[i], v);
      }
      // filter equivalence (test/gourdon/FactorTableD.cpp:118-145):
      // factor[m] > 2 pi(p) + 1  <=>  mu(m) != 0, lpf(m) > p, mpf(m) <= y (m > 1 composite)
      for (int p : {13, 17, 101, 173}) {
        int b = A.pi[p];
        for (size_t i = 1; i < f.size(); i++) {
          int n = (int) ft_to_number(i);
          if (A.lpf[n] == n) continue;
          bool want = A.mu[n] != 0 && A.lpf[n] > p && A.mpf[n] <= y;
          CHECK((f[i] > 2 * b + 1) == want, "filter p=%d n=%d", p, n);
        }
      }
    }
    // FactorTable<uint16_t>(y) (src/FactorTable.hpp:138-150)
    for (int y : {1000, 99999, 2000000}) {
      auto f = build_factor_table_lpf<uint16_t>(y, P, threads);
      for (size_t i = 0; i < f.size(); i++) {
        int n = (int) ft_to_number(i);
        int v;
        if (n == 1) v = 65534;
        else if (A.lpf[n] == n) v = 65535;
        else if (A.mu[n] == 0) v = 0;
        else v = A.mu[n] == 1 ? A.lpf[n] - 1 : A.lpf[n];
        CHECK(f[i] == v, "FactorTable(%d)[%d] = %d != %d", y, n, f[i], v);
      }
    }
  }
  std::printf(fails ? "FAILED (%d)\n" : "OK\n", fails);
  return fails ? 1 : 0;
}
