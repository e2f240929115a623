/*
 * hsl_oracle.c -- CPU restatement of primecount's hard-special-leaf sums.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the MI355X engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only to compare against or to time beside the product.  The
 * product (primecount_amd/) never links, loads or calls it.
 *
 * It restates, in plain C with simple data structures, the loop structure of
 *   - Gourdon's D(x, y, z, k)          src/gourdon/D.cpp:69-308 (D_thread_default)
 *   - Deleglise-Rivat S2_hard(x,y,z,c)  src/deleglise-rivat/S2_hard.cpp:51-161
 * run as ONE thread that owns the whole sieve interval [0, limit) (so the
 * phi[] carry starts at zero, phi_vector.cpp is not needed).  Where the
 * reference uses compressed tables it uses their mathematical definition:
 *   - FactorTableD  (src/gourdon/FactorTableD.hpp:20-39): the filter
 *       factor[m] > 2*pi(p_b)+1  <=>  m coprime to 2*3*5*7*11, mu(m) != 0,
 *       lpf(m) > p_b and mpf(m) <= y   (m = 1 and primes <= y always pass)
 *   - FactorTable   (src/FactorTable.hpp:138-150): mu_lpf[m] > p
 *       <=>  m coprime to 2*3*5*7*11, mu(m) != 0, lpf(m) > p
 *   - the m index range (to_index(min_m), to_index(max_m)] of
 *     src/BaseFactorTable.hpp:36-49 is the set of m coprime to 2310 in
 *     (min_m, max_m].
 *   - Sieve (src/sieve/Sieve.cpp:52-69, 497-845; count_stop.hpp:41-89): the
 *     segment holds the numbers coprime to 30 in [low, high); pre-sieving and
 *     cross-off remove a prime together with its multiples; count(stop)
 *     counts unsieved numbers in [low, low + stop].  Here a Fenwick tree over
 *     the segment answers count(stop).
 *
 * Parity status: pinned by the reference's known-answer rows
 * (test/gourdon/D.cpp:36-89, test/deleglise-rivat/S2_hard_alpha.cpp:36-73),
 * by fixtures made with the reference build in oracle/_ref, and by an
 * independent closed-form brute force in the tests.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

static int64_t isqrt_u128(u128 n)
{
  int64_t r = (int64_t) sqrtl((long double) n);
  while (r > 0 && (u128) r * (u128) r > n) r--;
  while ((u128) (r + 1) * (u128) (r + 1) <= n) r++;
  return r;
}

/* ---------- arithmetic tables (simple sieves, no compression) ---------- */

typedef struct {
  int64_t n;          /* tables cover [0, n] */
  int32_t* lpf;       /* least prime factor, lpf[1] = 1 */
  int32_t* mpf;       /* largest prime factor */
  int8_t* mu;         /* Moebius function */
} arith_t;

static void arith_init(arith_t* t, int64_t n)
{
  t->n = n;
  t->lpf = (int32_t*) calloc((size_t) n + 1, sizeof(int32_t));
  t->mpf = (int32_t*) calloc((size_t) n + 1, sizeof(int32_t));
  t->mu = (int8_t*) malloc((size_t) n + 1);
  for (int64_t i = 0; i <= n; i++) t->mu[i] = 1;
  if (n >= 1) { t->lpf[1] = 1; t->mpf[1] = 1; }
  for (int64_t p = 2; p <= n; p++) {
    if (t->lpf[p] != 0) continue;               /* composite */
    for (int64_t m = p; m <= n; m += p) {
      if (t->lpf[m] == 0) t->lpf[m] = (int32_t) p;
      t->mpf[m] = (int32_t) p;                   /* ascending p: last wins */
      t->mu[m] = (int8_t) -t->mu[m];
    }
    if (p <= n / p)
      for (int64_t m = p * p; m <= n; m += p * p) t->mu[m] = 0;
  }
}

static void arith_free(arith_t* t)
{
  free(t->lpf); free(t->mpf); free(t->mu);
}

static int coprime2310(int64_t m)
{
  return (m % 2) && (m % 3) && (m % 5) && (m % 7) && (m % 11);
}

/* primes[1..np], primes[0] = 0; pi[n] for n <= pmax */
typedef struct {
  int64_t pmax;
  int64_t np;
  int64_t* primes;
  int32_t* pi;
} primes_t;

static void primes_init(primes_t* P, int64_t pmax)
{
  if (pmax < 2) pmax = 2;
  P->pmax = pmax;
  char* comp = (char*) calloc((size_t) pmax + 1, 1);
  P->pi = (int32_t*) calloc((size_t) pmax + 1, sizeof(int32_t));
  int64_t cnt = 0;
  for (int64_t i = 2; i <= pmax; i++) {
    if (!comp[i]) {
      cnt++;
      if (i <= pmax / i)
        for (int64_t j = i * i; j <= pmax; j += i) comp[j] = 1;
    }
    P->pi[i] = (int32_t) cnt;
  }
  P->np = cnt;
  P->primes = (int64_t*) calloc((size_t) cnt + 2, sizeof(int64_t));
  int64_t k = 1;
  for (int64_t i = 2; i <= pmax; i++)
    if (!comp[i]) P->primes[k++] = i;
  free(comp);
}

static void primes_free(primes_t* P)
{
  free(P->primes); free(P->pi);
}

/* ---------- segment sieve with a Fenwick tree for count(stop) ---------- */

typedef struct {
  int64_t low, high, size;
  uint8_t* bit;      /* bit[i]: number low+i is unsieved */
  int32_t* fen;      /* Fenwick tree over bit[] */
  int64_t total;     /* unsieved numbers in [low, high) */
} seg_t;

static void fen_add(seg_t* s, int64_t i, int32_t v)
{
  for (i++; i <= s->size; i += i & -i) s->fen[i] += v;
}

/* unsieved count in [low, low + stop] */
static int64_t seg_count(const seg_t* s, int64_t stop)
{
  int64_t r = 0;
  if (stop >= s->size) stop = s->size - 1;
  for (int64_t i = stop + 1; i > 0; i -= i & -i) r += s->fen[i];
  return r;
}

static void seg_reset(seg_t* s, int64_t low, int64_t high)
{
  s->low = low; s->high = high; s->size = high - low;
  memset(s->fen, 0, sizeof(int32_t) * (size_t) (s->size + 1));
  s->total = 0;
  for (int64_t i = 0; i < s->size; i++) {
    int64_t n = low + i;
    s->bit[i] = (uint8_t) ((n % 2) && (n % 3) && (n % 5));
  }
}

/* finish initialisation: build the Fenwick tree from bit[] in O(size) */
static void seg_build(seg_t* s)
{
  s->total = 0;
  for (int64_t i = 1; i <= s->size; i++) {
    s->fen[i] += s->bit[i - 1];
    s->total += s->bit[i - 1];
    int64_t j = i + (i & -i);
    if (j <= s->size) s->fen[j] += s->fen[i];
  }
}

/* remove prime p and its multiples > low (Sieve::add: first multiple > low) */
static void seg_cross_off(seg_t* s, int64_t p, int counting)
{
  int64_t q = s->low / p + 1;
  for (int64_t m = p * q; m < s->high; m += p) {
    int64_t i = m - s->low;
    if (s->bit[i]) {
      s->bit[i] = 0;
      if (counting) { fen_add(s, i, -1); s->total--; }
    }
  }
}

/* ---------------------------- D(x, y, z, k) ---------------------------- */

static i128 D_impl(i128 x, int64_t y, int64_t z, int64_t k, int64_t seg_size)
{
  if (x < 1) return 0;
  int64_t xz = (int64_t) (x / z);
  /* x_star: src/util.cpp:362-386 */
  int64_t yy = y > 1 ? y : 1;
  i128 x_div_yy = (x + (i128) yy * yy - 1) / ((i128) yy * yy);
  int64_t r4 = isqrt_u128((u128) isqrt_u128((u128) x));  /* floor(x^(1/4)) */
  int64_t x_star = (int64_t) (x_div_yy > r4 ? x_div_yy : r4);
  int64_t sqrt_xy = isqrt_u128((u128) (x / yy));
  if (x_star > yy) x_star = yy;
  if (x_star > sqrt_xy) x_star = sqrt_xy;
  if (x_star < 1) x_star = 1;

  int64_t pmax = y;
  if (pmax < x_star) pmax = x_star;
  int64_t sqz = isqrt_u128((u128) z);
  if (pmax < sqz) pmax = sqz;
  primes_t P; primes_init(&P, pmax + 1);
  arith_t A; arith_init(&A, z > 1 ? z : 1);

  int64_t pi_sqrtz = P.pi[sqz];
  int64_t limit = xz;
  int64_t t1 = isqrt_u128((u128) x), t2 = isqrt_u128((u128) limit);
  int64_t mb = t1 < t2 ? t1 : t2;
  if (mb > x_star) mb = x_star;
  int64_t max_b = P.pi[mb];
  int64_t min_b = k + 1;                  /* low = 0: pi[min(xz/limit, x*)] <= k */
  i128 sum = 0;
  if (min_b > max_b || limit <= 0) { primes_free(&P); arith_free(&A); return 0; }

  int64_t* phi = (int64_t*) calloc((size_t) max_b + 2, sizeof(int64_t));
  seg_t S;
  S.bit = (uint8_t*) malloc((size_t) seg_size);
  S.fen = (int32_t*) malloc(sizeof(int32_t) * ((size_t) seg_size + 1));

  for (int64_t low = 0; low < limit; low += seg_size) {
    int64_t high = low + seg_size < limit ? low + seg_size : limit;
    int64_t low1 = low > 1 ? low : 1;
    seg_reset(&S, low, high);
    for (int64_t i = 4; i <= min_b - 1; i++) seg_cross_off(&S, P.primes[i], 0);
    seg_build(&S);
    int64_t b = min_b;
    int64_t last = pi_sqrtz < max_b ? pi_sqrtz : max_b;
    for (; b <= last; b++) {
      int64_t prime = P.primes[b];
      u128 xp = (u128) x / (u128) prime;
      int64_t xp_low = xp / (u128) low1 < (u128) z ? (int64_t) (xp / (u128) low1) : z;
      int64_t xp_high = xp / (u128) high < (u128) z ? (int64_t) (xp / (u128) high) : z;
      int64_t min_m = xp_high > z / prime ? xp_high : z / prime;
      u128 xpp = xp / ((u128) prime * (u128) prime);
      int64_t max_m = xpp < (u128) xp_low ? (int64_t) xpp : xp_low;
      if (prime >= max_m) goto next_segment;
      for (int64_t m = max_m; m > min_m; m--) {
        if (!coprime2310(m)) continue;
        int pass = (m == 1) || (A.mu[m] != 0 && A.lpf[m] > prime && A.mpf[m] <= y);
        if (!pass) continue;
        int64_t xpm = (int64_t) (xp / (u128) m);
        int64_t cnt = seg_count(&S, xpm - low);
        sum -= (i128) A.mu[m] * (phi[b] + cnt);
      }
      phi[b] += S.total;
      seg_cross_off(&S, prime, 1);
    }
    for (; b <= max_b; b++) {
      int64_t prime = P.primes[b];
      u128 xp = (u128) x / (u128) prime;
      int64_t xp_low = xp / (u128) low1 < (u128) y ? (int64_t) (xp / (u128) low1) : y;
      int64_t xp_high = xp / (u128) high < (u128) y ? (int64_t) (xp / (u128) high) : y;
      int64_t min_m = xp_high > prime ? xp_high : prime;
      u128 xpp = xp / ((u128) prime * (u128) prime);
      int64_t max_m = xpp < (u128) xp_low ? (int64_t) xpp : xp_low;
      int64_t l = P.pi[max_m];
      if (prime >= P.primes[l]) goto next_segment;
      for (; P.primes[l] > min_m; l--) {
        int64_t xpq = (int64_t) (xp / (u128) P.primes[l]);
        sum += phi[b] + seg_count(&S, xpq - low);
      }
      phi[b] += S.total;
      seg_cross_off(&S, prime, 1);
    }
  next_segment:;
  }

  free(phi); free(S.bit); free(S.fen);
  primes_free(&P); arith_free(&A);
  return sum;
}

/* -------------------------- S2_hard(x, y, z, c) ------------------------ */

static i128 S2_hard_impl(i128 x, int64_t y, int64_t z, int64_t c, int64_t seg_size)
{
  if (x < 1 || z <= 0) return 0;
  int64_t sqy = isqrt_u128((u128) y);
  int64_t sqrt_y = sqy;
  int64_t max_prime = z / (sqy > 0 ? sqy : 1);
  if (max_prime > y) max_prime = y;
  int64_t pmax = max_prime;
  if (pmax < sqy) pmax = sqy;
  primes_t P; primes_init(&P, pmax + 1);
  arith_t A; arith_init(&A, y > 1 ? y : 1);

  int64_t limit = z;
  int64_t pi_sqrty = P.pi[sqrt_y];
  int64_t max_b;
  if (limit <= y) max_b = pi_sqrty;
  else {
    int64_t t1 = isqrt_u128((u128) x), t2 = isqrt_u128((u128) z);
    int64_t mb = t1 < t2 ? t1 : t2;
    if (mb > y) mb = y;
    if (mb > P.pmax) mb = P.pmax;
    max_b = P.pi[mb];
  }
  int64_t min_b = P.pi[(z / limit) < P.primes[max_b] ? (z / limit) : P.primes[max_b]];
  min_b = (c > min_b ? c : min_b) + 1;
  i128 sum = 0;
  if (min_b > max_b) { primes_free(&P); arith_free(&A); return 0; }

  int64_t* phi = (int64_t*) calloc((size_t) max_b + 2, sizeof(int64_t));
  seg_t S;
  S.bit = (uint8_t*) malloc((size_t) seg_size);
  S.fen = (int32_t*) malloc(sizeof(int32_t) * ((size_t) seg_size + 1));

  for (int64_t low = 0; low < limit; low += seg_size) {
    int64_t high = low + seg_size < limit ? low + seg_size : limit;
    int64_t low1 = low > 1 ? low : 1;
    seg_reset(&S, low, high);
    for (int64_t i = 4; i <= min_b - 1; i++) seg_cross_off(&S, P.primes[i], 0);
    seg_build(&S);
    int64_t b = min_b;
    int64_t last = pi_sqrty < max_b ? pi_sqrty : max_b;
    for (; b <= last; b++) {
      int64_t prime = P.primes[b];
      u128 xp = (u128) x / (u128) prime;
      int64_t xp_high = xp / (u128) high < (u128) y ? (int64_t) (xp / (u128) high) : y;
      int64_t min_m = xp_high > y / prime ? xp_high : y / prime;
      int64_t max_m = xp / (u128) low1 < (u128) y ? (int64_t) (xp / (u128) low1) : y;
      if (prime >= max_m) goto next_segment;
      for (int64_t m = max_m; m > min_m; m--) {
        if (!coprime2310(m)) continue;
        int pass = (m == 1) || (A.mu[m] != 0 && A.lpf[m] > prime);
        if (!pass) continue;
        int64_t xpm = (int64_t) (xp / (u128) m);
        int64_t cnt = seg_count(&S, xpm - low);
        sum -= (i128) A.mu[m] * (phi[b] + cnt);
      }
      phi[b] += S.total;
      seg_cross_off(&S, prime, 1);
    }
    for (; b <= max_b; b++) {
      int64_t prime = P.primes[b];
      u128 xp = (u128) x / (u128) prime;
      int64_t xp_low = xp / (u128) low1 < (u128) y ? (int64_t) (xp / (u128) low1) : y;
      int64_t xp_high = xp / (u128) high < (u128) y ? (int64_t) (xp / (u128) high) : y;
      int64_t zp = z / prime;
      int64_t l = P.pi[xp_low < zp ? xp_low : zp];
      int64_t min_hard = xp_high > prime ? xp_high : prime;
      if (prime >= P.primes[l]) goto next_segment;
      for (; P.primes[l] > min_hard; l--) {
        int64_t xpq = (int64_t) (xp / (u128) P.primes[l]);
        sum += phi[b] + seg_count(&S, xpq - low);
      }
      phi[b] += S.total;
      seg_cross_off(&S, prime, 1);
    }
  next_segment:;
  }

  free(phi); free(S.bit); free(S.fen);
  primes_free(&P); arith_free(&A);
  return sum;
}

/* ------------------------------ C ABI ---------------------------------- */
/* 128-bit values cross the ABI as (lo, hi) pairs of 64-bit words.        */

static i128 mk(uint64_t lo, int64_t hi) { return (i128) (((u128) (uint64_t) hi << 64) | lo); }
static void put(i128 v, uint64_t* lo, int64_t* hi) { *lo = (uint64_t) v; *hi = (int64_t) (v >> 64); }

void oracle_D(uint64_t x_lo, int64_t x_hi, int64_t y, int64_t z, int64_t k,
              int64_t seg_size, uint64_t* out_lo, int64_t* out_hi)
{
  put(D_impl(mk(x_lo, x_hi), y, z, k, seg_size > 0 ? seg_size : (1 << 16)), out_lo, out_hi);
}

void oracle_S2_hard(uint64_t x_lo, int64_t x_hi, int64_t y, int64_t z, int64_t c,
                    int64_t seg_size, uint64_t* out_lo, int64_t* out_hi)
{
  put(S2_hard_impl(mk(x_lo, x_hi), y, z, c, seg_size > 0 ? seg_size : (1 << 16)), out_lo, out_hi);
}

/* phi(n, a): numbers <= n not divisible by any of the first a primes,
   by direct sieving (tiny n only; used to pin the sieve semantics). */
int64_t oracle_phi_direct(int64_t n, int64_t a)
{
  if (n < 1) return 0;
  primes_t P; primes_init(&P, n + 1 > 64 ? n + 1 : 64);
  int64_t cnt = 0;
  for (int64_t m = 1; m <= n; m++) {
    int ok = 1;
    for (int64_t i = 1; i <= a && i <= P.np && ok; i++)
      if (m % P.primes[i] == 0) ok = 0;
    cnt += ok;
  }
  primes_free(&P);
  return cnt;
}
