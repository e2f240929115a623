/*
 * dr_oracle.c -- CPU restatement of the Deleglise-Rivat terms other than
 * S2_hard, pi(x) = S1 + S2_trivial + S2_easy + S2_hard + pi(y) - 1 - P2
 * (src/deleglise-rivat/pi_deleglise_rivat.cpp:40-53, 86-92).
 *
 * TEST INFRASTRUCTURE ONLY, like hsl_oracle.c and gourdon_oracle.c (same
 * library): the checker for the MI355X engine's Deleglise-Rivat terms.  The
 * product never links or loads it.
 *
 * Each term is the plain sum over its definition, with pi(n) from a sieve
 * table covering [0, max(x / y, sqrt x, y)] (x <= 1e13 in a test):
 *   S1          src/S1.cpp:29-70: sum over the square-free n <= y whose least
 *               prime factor exceeds p_c of mu(n) phi(x / n, c) (n = 1
 *               included), phi by Legendre's recursion;
 *   S2_trivial  src/deleglise-rivat/S2_trivial.cpp:39-88: the leaves
 *               p_b p_l with phi(x / (p_b p_l), b - 1) = 1, counted per prime
 *               p_b, max(p_c, sqrt z) < p_b < y, as pi(y) - pi(max(x / p_b^2,
 *               p_b)); no arithmetic series;
 *   S2_easy     src/deleglise-rivat/S2_easy.cpp:53-160: for max(c, pi(sqrt y))
 *               < b <= pi(x^(1/3)) and pi(in_between(p_b, z / p_b, y)) < l <=
 *               pi(min(x / p_b^2, y)), pi(x / (p_b p_l)) - b + 2, leaf by leaf
 *               (no clustering);
 *   P2          src/P2.cpp:95-127 restated by its definition: the n <= x with
 *               exactly two prime factors, both above y, i.e. the sum over
 *               y < p <= sqrt x of pi(x / p) - pi(p) + 1 (no closed form).
 * z <= x / y as in the algorithm (then min_sparse <= min_clustered in S2_easy
 * and its clustered and sparse ranges join into one).
 * Parity status: pinned by the values the reference build prints for these
 * terms (tests/golden/dr_terms.json) and by phi(x, pi(y)) = S1 + S2 and
 * pi(x) = phi(x, pi(y)) + pi(y) - 1 - P2 against sieved pi(x).
 */
#include <stdint.h>
#include <stdlib.h>
#include <math.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

static int64_t d_isqrt(u128 n)
{
  int64_t r = (int64_t) sqrtl((long double) n);
  while (r > 0 && (u128) r * (u128) r > n) r--;
  while ((u128) (r + 1) * (u128) (r + 1) <= n) r++;
  return r;
}

static int64_t d_iroot3(u128 n)
{
  int64_t r = (int64_t) cbrtl((long double) n);
  while (r > 0 && (u128) r * r * r > n) r--;
  while ((u128) (r + 1) * (r + 1) * (r + 1) <= n) r++;
  return r;
}

static int64_t d_max(int64_t a, int64_t b) { return a > b ? a : b; }
static int64_t d_min(int64_t a, int64_t b) { return a < b ? a : b; }

typedef struct {
  int64_t n, np;
  int64_t* primes;   /* primes[1..np], primes[0] = 0 */
  int64_t* pi;       /* pi[m], m <= n */
} dtab_t;

static void dtab_init(dtab_t* t, int64_t n)
{
  if (n < 32) n = 32;
  t->n = n;
  char* comp = (char*) calloc((size_t) n + 1, 1);
  t->pi = (int64_t*) calloc((size_t) n + 1, sizeof(int64_t));
  int64_t c = 0;
  for (int64_t i = 2; i <= n; i++) {
    if (!comp[i]) {
      c++;
      if (i <= n / i)
        for (int64_t j = i * i; j <= n; j += i) comp[j] = 1;
    }
    t->pi[i] = c;
  }
  t->np = c;
  t->primes = (int64_t*) calloc((size_t) c + 2, sizeof(int64_t));
  int64_t k = 1;
  for (int64_t i = 2; i <= n; i++)
    if (!comp[i]) t->primes[k++] = i;
  free(comp);
}

static void dtab_free(dtab_t* t) { free(t->primes); free(t->pi); }

static i128 d_phi(i128 x, int64_t a, const int64_t* primes)
{
  if (a == 0 || x == 0) return x;
  return d_phi(x, a - 1, primes) - d_phi(x / primes[a], a - 1, primes);
}

/* square-free n = sf * (primes with index > b), n <= y, sign mu */
static i128 s1_rec(i128 x, int64_t y, int64_t b, int64_t c, int64_t sf, int mu, const dtab_t* P)
{
  i128 s = 0;
  for (int64_t j = b + 1; j <= P->np; j++) {
    const int64_t p = P->primes[j];
    if (p > y / sf) break;
    const int64_t n = sf * p;
    s += -mu * d_phi(x / n, c, P->primes);
    s += s1_rec(x, y, j, c, n, -mu, P);
  }
  return s;
}

static i128 S1_impl(i128 x, int64_t y, int64_t c, const dtab_t* P)
{
  return d_phi(x, c, P->primes) + s1_rec(x, y, c, c, 1, 1, P);
}

static i128 S2_trivial_impl(i128 x, int64_t y, int64_t z, int64_t c, const dtab_t* P)
{
  if (y < 2) return 0;
  const int64_t start = d_max(c > 0 ? P->primes[c] : 0, d_isqrt((u128) z)) + 1;
  i128 s = 0;
  for (int64_t b = 1; b <= P->np; b++) {
    const int64_t p = P->primes[b];
    if (p >= y) break;
    if (p < start) continue;
    const i128 xpp = x / ((i128) p * p);
    const int64_t m = xpp > p ? (int64_t) xpp : p;
    s += P->pi[y] - P->pi[m];
  }
  return s;
}

static int64_t d_between(int64_t lo, int64_t v, int64_t hi)
{
  return v < lo ? lo : (v > hi ? hi : v);
}

static i128 S2_easy_impl(i128 x, int64_t y, int64_t z, int64_t c, const dtab_t* P)
{
  const int64_t pi_sqrty = P->pi[d_isqrt((u128) y)];
  const int64_t pi_x13 = P->pi[d_iroot3((u128) x)];
  i128 s = 0;
  for (int64_t b = d_max(c, pi_sqrty) + 1; b <= pi_x13; b++) {
    const int64_t p = P->primes[b];
    const i128 xp = x / p;
    const int64_t l_hi = P->pi[d_min((int64_t) (xp / p), y)];
    const int64_t l_lo = P->pi[d_between(p, z / p, y)];
    for (int64_t l = l_lo + 1; l <= l_hi; l++)
      s += P->pi[(int64_t) (xp / P->primes[l])] - b + 2;
  }
  return s;
}

static i128 P2_impl(i128 x, int64_t y, const dtab_t* P)
{
  const int64_t sqrtx = d_isqrt((u128) x);
  i128 s = 0;
  for (int64_t b = P->pi[y] + 1; b <= P->np && P->primes[b] <= sqrtx; b++) {
    const int64_t p = P->primes[b];
    s += P->pi[(int64_t) (x / p)] - P->pi[p] + 1;
  }
  return s;
}

/* which: 0 = S1, 1 = S2_trivial, 2 = S2_easy, 3 = P2 */
void oracle_dr_term(int which, uint64_t x_lo, int64_t x_hi, int64_t y, int64_t z, int64_t c,
                    uint64_t* out_lo, int64_t* out_hi)
{
  const i128 x = (i128) (((u128) (uint64_t) x_hi << 64) | x_lo);
  int64_t n = (int64_t) (x / y);
  n = d_max(n, d_isqrt((u128) x));
  n = d_max(n, y);
  dtab_t P;
  dtab_init(&P, n + 1);
  i128 r = 0;
  switch (which) {
    case 0: r = S1_impl(x, y, c, &P); break;
    case 1: r = S2_trivial_impl(x, y, z, c, &P); break;
    case 2: r = S2_easy_impl(x, y, z, c, &P); break;
    default: r = P2_impl(x, y, &P); break;
  }
  dtab_free(&P);
  *out_lo = (uint64_t) r;
  *out_hi = (int64_t) (r >> 64);
}
