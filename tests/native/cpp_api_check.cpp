// cpp_api_check.cpp -- a C++ program that links libprimecount_mi355x.so and
// calls the engine through include/primecount_mi355x.hpp (the overloads that
// mirror primecount's internal D() / S2_hard() and its pi() / nth_prime()),
// on the device.  Expected values: D and S2_hard rows of the reference's own
// tests (test/gourdon/D.cpp:36-89, test/deleglise-rivat/S2_hard_alpha.cpp),
// pi(10^n) from the reference's README, nth_prime values printed by the
// reference build (tests/golden/nth_prime.json).
//
// Built by __graft_entry__.build() into tests/native/_bin; run by
// tests/test_gpu_cpp_api.py.  Prints one line per check and "cpp api ok".
#include <cstdio>
#include <cstdlib>
#include <string>

#include "primecount_mi355x.hpp"

namespace pm = primecount_mi355x;

static int failures = 0;

static std::string str128(pm::int128_t v)
{
  if (v == 0) return "0";
  const bool neg = v < 0;
  unsigned __int128 u = neg ? (unsigned __int128) (-(v + 1)) + 1 : (unsigned __int128) v;
  std::string s;
  while (u) {
    s.insert(s.begin(), (char) ('0' + (int) (u % 10)));
    u /= 10;
  }
  return neg ? "-" + s : s;
}

static void expect(const char* what, pm::int128_t got, pm::int128_t want)
{
  const bool ok = got == want;
  std::printf("%s %s = %s%s\n", ok ? "ok  " : "FAIL", what, str128(got).c_str(),
              ok ? "" : (" (want " + str128(want) + ")").c_str());
  std::fflush(stdout);
  failures += !ok;
}

int main()
{
  const pm::int128_t E19 = (pm::int128_t) 10000000000000000000ull;
  // D(x, y, z, k), 64-bit and 128-bit overloads (D.cpp:36-89 rows)
  expect("D(1e13, 241615, 241615, 8)", pm::D((int64_t) 10000000000000ll, 241615, 241615, 8),
         304739017794ll);
  expect("D(int128 1e13, 107720, 209946, 8)",
         pm::D((pm::int128_t) 10000000000000ll, 107720, 209946, 8), 270354670695ll);
  // S2_hard(x, y, z, c) of the reference's own Deleglise-Rivat parameters
  expect("S2_hard(1e12, 10000, 1e8, 8)", pm::S2_hard((int64_t) 1000000000000ll, 10000, 100000000, 8),
         47185925583ll);
  // pi(x) and the Deleglise-Rivat pi(x): README values
  expect("pi(1e12)", pm::pi((int64_t) 1000000000000ll), 37607912018ll);
  expect("pi(int128 1e19)", pm::pi(E19), (pm::int128_t) 234057667276344607ll);
  expect("pi_deleglise_rivat(1e15)", pm::pi_deleglise_rivat((int64_t) 1000000000000000ll),
         29844570422669ll);
  // nth_prime (values printed by the reference build)
  expect("nth_prime(1e9)", pm::nth_prime((int64_t) 1000000000ll), 22801763489ll);
  // the first prime above 2^63 through the 128-bit overload (pi(p) = n and
  // pi(p - 1) = n - 1 pin it; n = pi(2^63 - 1) + 1)
  expect("nth_prime(int128 216289611853439385)",
         pm::nth_prime((pm::int128_t) 216289611853439385ll),
         ((pm::int128_t) 1 << 63) + 29);
  // an argument outside the domain throws primecount_error, and the engine
  // keeps working afterwards
  bool threw = false;
  try {
    pm::D((int64_t) 1000000000000ll, 100, 50, 8);        // z < y
  } catch (const pm::primecount_error& e) {
    threw = true;
    std::printf("ok   D(1e12, 100, 50, 8) throws: %s\n", e.what());
  }
  if (!threw) {
    std::printf("FAIL D(1e12, 100, 50, 8) did not throw\n");
    failures++;
  }
  expect("D(1e6, 207, 207, 8) after the error", pm::D((int64_t) 1000000, 207, 207, 8), 2465);
  if (failures) {
    std::printf("cpp api: %d failures\n", failures);
    return 1;
  }
  std::printf("cpp api ok\n");
  return 0;
}
