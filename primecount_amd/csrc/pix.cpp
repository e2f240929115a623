// pix.cpp -- host orchestration of Gourdon's pi(x) on the device, of the
// Deleglise-Rivat terms S1, S2_trivial, S2_easy and P2 (with S2_hard, the
// Deleglise-Rivat pi(x)) and of nth_prime.
//
// The easy-leaf sums of A and C (src/gourdon/AC.cpp) are cut into one
// descriptor per sieving prime (or per pair of primes for the m = q r leaves
// of C1) whose leaves are consecutive primes; the flat leaf-sum kernel then
// reads pi(x / (p m)) from the resident prime-count table.  The reference
// walks the same leaves segment by segment with a small pi table
// (SegmentedPiTable, x^(1/4) memory) and clusters equal pi values; here the
// whole table up to max(x / y, sqrt(x)) lives in HBM (7.5 GB at x = 1e19),
// which turns every leaf into two loads.  A table above the engine's budget
// (65% of the device memory by default: 1e22 needs 354 GB) is built in
// windows instead: the part up to sqrt(x) that Sigma and B's primes read
// stays whole, and the leaf sums of A, C and B run once per window, each
// leaf adding its pi value in the window that holds its quotient.

#include "pix.hpp"

#include "hsl_kernels.hpp"
#include "params.hpp"
#include "pix_kernels.hpp"
#include "tables.hpp"

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <future>
#include <map>
#include <thread>

namespace pcg {

#define PIX_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw pcg_error(std::string("HIP error '") + hipGetErrorString(e_) +         \
                      "' in " #expr,                                               \
                      e_ == hipErrorOutOfMemory ? kErrNoMem : kErrDevice);         \
  } while (0)

namespace {

double now_s()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

uint64_t umin(uint64_t a, uint64_t b) { return a < b ? a : b; }
uint64_t umax(uint64_t a, uint64_t b) { return a > b ? a : b; }

/// Largest n the terms of x read pi(n) for.
uint64_t table_need(i128 x, int64_t y, int64_t x_star)
{
  u128 n = (u128) x / (u128) y;
  n = std::max<u128>(n, isqrt((u128) x));
  n = std::max<u128>(n, (u128) x / ((u128) x_star * (u128) y));
  n = std::max<u128>(n, isqrt((u128) x / (u128) x_star));
  if (n >= ((u128) 1 << 49))
    throw pcg_error("pi(x): x / y must be < 2^49 (prime-count table limit)");
  return (uint64_t) n + 1;
}

/// Largest n that Sigma, pi(isqrt(x)) and B's primes read pi(n) or the
/// table's bits for: the part of the table kept whole when the rest is built
/// in windows.
uint64_t table_low_need(i128 x, int64_t y, int64_t x_star)
{
  u128 n = isqrt((u128) x);
  n = std::max<u128>(n, (u128) x / ((u128) x_star * (u128) y));
  n = std::max<u128>(n, isqrt((u128) x / (u128) x_star));
  n = std::max<u128>(n, (u128) y);
  return (uint64_t) n + 1;
}

/// Table words per window when a table of `need` numbers is built in
/// windows (0: it is kept whole): the fewest windows of at most the budget,
/// of equal size up to a sieve segment.
uint64_t window_words(uint64_t need)
{
  const uint64_t words = need / 240 + 1;
  const uint64_t budget = std::max<uint64_t>(engine_pix_budget() / sizeof(PixWord), kPixSegWords);
  if (words <= budget) return 0;
  // windows of equal size, whole sieve segments
  const uint64_t nwin = (words + budget - 1) / budget;
  const uint64_t per = (words + nwin - 1) / nwin;
  return (per + kPixSegWords - 1) / kPixSegWords * kPixSegWords;
}

/// Builds the prime-count table up to `need` in windows of win words, lowest
/// first, and calls f(tab, w_lo, w_hi) for each (tab holds the words
/// [w_lo, w_hi) with absolute counts, the primes below a window carried from
/// the one before); adds the build time to *table_s; returns the windows.
template <typename F>
uint64_t for_windows(uint64_t need, uint64_t win, double* table_s, F&& f)
{
  const uint64_t words = need / 240 + 1;
  uint64_t base = 0, n = 0;
  for (uint64_t w_lo = 0; w_lo < words; w_lo += win) {
    const uint64_t nw = std::min<uint64_t>(win, words - w_lo);
    const double t = now_s();
    const PixWord* tab = engine_pix_window(w_lo, nw, need, base, &base);
    *table_s += now_s() - t;
    f(tab, w_lo, w_lo + nw);
    n++;
  }
  return n;
}

i128 acc_value(const uint64_t* a)
{
  return (i128) (((u128) a[1] << 64) | a[0]);
}

/// Accumulators of one run: 8 signed 128-bit sums in device memory.
struct Acc {
  unsigned long long* d = nullptr;
  uint64_t h[16] = {};
};

void acc_zero(Acc& A, hipStream_t s)
{
  PIX_HIP(hsl_launch_fill_zero(A.d, 128, s));
}

void acc_read(Acc& A, hipStream_t s)
{
  PIX_HIP(hipMemcpyAsync(A.h, A.d, 128, hipMemcpyDeviceToHost, s));
  PIX_HIP(hipStreamSynchronize(s));
}

/// pi(n) for a few n, on the device.
std::vector<uint64_t> pi_points(const EngineView& v, const std::vector<uint64_t>& ns)
{
  for (uint64_t n : ns)
    if (n > v.pix_n) throw pcg_error("internal: prime-count table too small", kErrInternal);
  uint64_t* dn = (uint64_t*) engine_buffer(1, ns.size() * 8);
  uint64_t* dout = (uint64_t*) engine_buffer(2, ns.size() * 8);
  PIX_HIP(hipMemcpyAsync(dn, ns.data(), ns.size() * 8, hipMemcpyHostToDevice, v.stream));
  PIX_HIP(pix_launch_pi_points(dn, dout, (uint32_t) ns.size(), v.pix_tab, v.stream));
  std::vector<uint64_t> out(ns.size());
  PIX_HIP(hipMemcpyAsync(out.data(), dout, ns.size() * 8, hipMemcpyDeviceToHost, v.stream));
  PIX_HIP(hipStreamSynchronize(v.stream));
  return out;
}

/// phi(r, k) for r < pp (pp = product of the first k primes) and phi(pp, k):
/// the tiny phi of src/PhiTiny.hpp, tabulated once per k and kept on the
/// device.
struct PhiTiny {
  uint32_t k = 0, pp = 1;
  uint64_t phi_pp = 1;
  std::vector<uint32_t> tbl;
};

const PhiTiny& phi_tiny_table(uint32_t k)
{
  static std::map<uint32_t, PhiTiny> cache;
  auto it = cache.find(k);
  if (it != cache.end()) return it->second;
  static const uint32_t small[8] = {2, 3, 5, 7, 11, 13, 17, 19};
  PhiTiny t;
  t.k = k;
  for (uint32_t i = 0; i < k; i++) {
    t.pp *= small[i];
    t.phi_pp *= small[i] - 1;
  }
  t.tbl.resize(t.pp);
  uint32_t c = 0;
  for (uint32_t r = 0; r < t.pp; r++) {
    bool ok = r > 0;
    for (uint32_t i = 0; i < k && ok; i++)
      ok = r % small[i] != 0;
    c += ok;
    t.tbl[r] = c;                      // phi(r, k): 1 <= m <= r coprime to pp
  }
  return cache.emplace(k, std::move(t)).first->second;
}

i128 phi_tiny_host(i128 x, const PhiTiny& t)
{
  const u128 u = (u128) x;
  return (i128) ((u / t.pp) * t.phi_pp + t.tbl[(uint32_t) (u % t.pp)]);
}

/// Sigma4 .. Sigma6 over the primes x_star < p <= x^(1/3), into acc[0..5].
void launch_sigma(i128 x, int64_t y, int64_t x_star, const EngineView& v, unsigned long long* acc)
{
  const uint64_t x13 = iroot<3>((u128) x);
  const uint64_t sqrt_xy = isqrt((u128) x / (u128) y);
  const uint32_t b_lo = (uint32_t) v.host->pi((uint64_t) x_star) + 1;
  const uint32_t b_hi = (uint32_t) v.host->pi(x13) + 1;
  if (b_hi > v.nprimes) throw pcg_error("internal: prime table too small for Sigma", kErrInternal);
  PIX_HIP(pix_launch_sigma((uint64_t) x, (uint64_t) ((u128) x >> 64), (uint64_t) y, sqrt_xy, v.primes,
                           b_lo, b_hi, v.pix_tab, acc, v.stream));
}

/// Sigma from a = pi(y), b = pi(x^(1/3)), c = pi(sqrt(x / y)), d = pi(x_star),
/// pi(isqrt(x)) and the Sigma4 .. Sigma6 sums h[0..5].
i128 sigma_value(i128 a, i128 b, i128 c, i128 d, i128 pi_sqrtx, const uint64_t* h)
{
  // Sigma0 .. Sigma3 (src/gourdon/Sigma.cpp:30-52)
  const i128 s0 = a - 1 + (pi_sqrtx * (pi_sqrtx - 1)) / 2 - (a * (a - 1)) / 2;
  const i128 s1 = (a - b) * (a - b - 1) / 2;
  const i128 s2 = a * (b - c - (c * (c - 3)) / 2 + (d * (d - 3)) / 2);
  const i128 s3 = (b * (b - 1) * (2 * b - 1)) / 6 - b - (d * (d - 1) * (2 * d - 1)) / 6 + d;
  return s0 + s1 + s2 + s3 + acc_value(h) * a + acc_value(h + 2) - acc_value(h + 4);
}

i128 run_sigma(i128 x, int64_t y, int64_t x_star, const EngineView& v)
{
  const uint64_t x13 = iroot<3>((u128) x);
  const uint64_t sqrt_xy = isqrt((u128) x / (u128) y);
  const std::vector<uint64_t> p = pi_points(v, {(uint64_t) y, x13, sqrt_xy, (uint64_t) x_star,
                                                isqrt((u128) x)});
  Acc A;
  A.d = (unsigned long long*) engine_buffer(0, 128);
  acc_zero(A, v.stream);
  launch_sigma(x, y, x_star, v, A.d);
  acc_read(A, v.stream);
  return sigma_value(p[0], p[1], p[2], p[3], p[4], A.h);
}

/// The tiny phi table of k on the device (uploaded once per buffer).
const uint32_t* phi_tiny_device(const PhiTiny& t, const EngineView& v)
{
  uint32_t* dtbl = (uint32_t*) engine_buffer(3, (size_t) t.pp * 4);
  // the device copy stays until the buffer moves (per work set)
  PhiUpload& up = engine_phi_upload();
  if (dtbl != up.buf || up.k != t.k || up.gen != engine_generation()) {
    PIX_HIP(hipMemcpyAsync(dtbl, t.tbl.data(), (size_t) t.pp * 4, hipMemcpyHostToDevice, v.stream));
    PIX_HIP(hipStreamSynchronize(v.stream));
    up.buf = dtbl;
    up.k = t.k;
    up.gen = engine_generation();
  }
  return dtbl;
}

/// The sum of mu(n) phi(x / n, k) over the table entries n <= z, into acc[0..1].
void launch_phi0(i128 x, int64_t z, int64_t k, const EngineView& v, const PhiTiny& t,
                 const uint32_t* dtbl, unsigned long long* acc)
{
  static const uint32_t small[8] = {2, 3, 5, 7, 11, 13, 17, 19};
  const uint64_t nidx = (uint64_t) ft_to_index((uint64_t) z) + 1;
  if (nidx > v.factor_n) throw pcg_error("internal: factor table too small for Phi0", kErrInternal);
  PIX_HIP(pix_launch_phi0((uint64_t) x, (uint64_t) ((u128) x >> 64), v.factor, v.factor32, false,
                          nidx, (uint32_t) k, small[k - 1], dtbl, t.pp, t.phi_pp, acc, v.stream));
}

const PhiTiny& phi0_table(int64_t k)
{
  if (k < 5 || k > 8)
    throw pcg_error("Phi0(x, y, z, k): the device path needs 5 <= k <= 8");
  return phi_tiny_table((uint32_t) k);
}

i128 run_phi0(i128 x, int64_t y, int64_t z, int64_t k, const EngineView& v)
{
  (void) y;
  const PhiTiny& t = phi0_table(k);
  const uint32_t* dtbl = phi_tiny_device(t, v);
  Acc A;
  A.d = (unsigned long long*) engine_buffer(0, 128);
  acc_zero(A, v.stream);
  launch_phi0(x, z, k, v, t, dtbl, A.d);
  acc_read(A, v.stream);
  return phi_tiny_host(x, t) + acc_value(A.h);
}

/// B: the sum of pi(x / p) over y < p <= isqrt(x), into acc[0..1]; the
/// primes from the resident table, pi(x / p) from tab, which holds the
/// table words [w_lo, w_hi) (the resident table: 0, UINT64_MAX).
void launch_B(i128 x, int64_t y, const EngineView& v, unsigned long long* acc,
              const PixWord* tab = nullptr, uint64_t w_lo = 0, uint64_t w_hi = ~0ull)
{
  if (isqrt((u128) x) > v.pix_n) throw pcg_error("internal: prime-count table too small for B", kErrInternal);
  PIX_HIP(pix_launch_B((uint64_t) x, (uint64_t) ((u128) x >> 64), (uint64_t) y, isqrt((u128) x),
                       v.pix_tab, tab ? tab : v.pix_tab, w_lo, w_hi, acc, v.stream));
}

i128 run_B(i128 x, int64_t y, const EngineView& v)
{
  Acc A;
  A.d = (unsigned long long*) engine_buffer(0, 128);
  acc_zero(A, v.stream);
  launch_B(x, y, v, A.d);
  acc_read(A, v.stream);
  return acc_value(A.h);
}

/// Descriptor lists of the A and C leaves (AC.cpp, as one segment [0, isqrt(x))).
struct LeafLists {
  std::vector<PixLeafDesc> plus, minus;
  uint64_t n_plus = 0, n_minus = 0;
  i128 correction = 0;       // the (2 - b) parts of C, summed on the host
};

void add_desc(std::vector<PixLeafDesc>& list, uint64_t& total, uint64_t xp, uint64_t q1,
              uint64_t i_lo, uint64_t i_hi)
{
  if (i_hi < i_lo) return;
  PixLeafDesc d;
  d.xp = xp;
  d.xd = (double) xp;
  d.q1 = (uint32_t) std::min<uint64_t>(q1, 0xffffffffull);
  d.off = total;
  d.i_lo = (uint32_t) i_lo;
  list.push_back(d);
  total += i_hi - i_lo + 1;
}

/// The descriptors of the sieving primes b = t, t + T, t + 2T, ... of each
/// part of A and C (term order does not matter to the sums).
void ac_lists_part(i128 x, int64_t y, int64_t z, int64_t k, int64_t x_star, const PrimeTables& P,
                   uint64_t t, uint64_t T, LeafLists& L)
{
  const u128 ux = (u128) x;
  const uint64_t sqrtx = isqrt(ux);
  const uint64_t x13 = iroot<3>(ux);
  const u128 xhigh = ux / sqrtx;                 // one segment [0, isqrt(x))
  const uint64_t uy = (uint64_t) y, uz = (uint64_t) z;
  auto pi = [&](uint64_t n) -> uint64_t {
    if (n > P.max) throw pcg_error("internal: prime table too small for A/C", kErrInternal);
    return P.pi(n);
  };
  auto xp_of = [&](uint64_t p) -> uint64_t {
    const u128 xp = ux / p;
    if (xp >> 64) throw pcg_error("A/C: x / p does not fit in 64 bits");
    return (uint64_t) xp;
  };

  // A: pi[x_star] < b <= pi[x^(1/3)] (AC.cpp:57-113, 456-465)
  const uint64_t a0 = pi((uint64_t) x_star) + 1;
  for (uint64_t b = a0 + t; b <= pi(x13); b += T) {
    const uint64_t p = P.primes[b];
    const uint64_t xp = xp_of(p);
    const uint64_t sqrt_xp = isqrt(xp);
    const uint64_t min2 = umin((uint64_t) (xhigh / p), sqrt_xp);
    add_desc(L.plus, L.n_plus, xp, xp / uy, pi(umax(p, min2)) + 1, pi(sqrt_xp));
  }
  // C1: max(k, pi[(x/z)^(1/3)]) < b <= pi[sqrt(z)] (AC.cpp:115-214, 402-418)
  const uint64_t sqrtz = isqrt((u128) uz);
  const uint64_t xz = (uint64_t) (ux / uz);
  const uint64_t c0 = umax((uint64_t) k, pi(iroot<3>((u128) xz))) + 1;
  for (uint64_t b = c0 + t; b <= pi(sqrtz); b += T) {
    const uint64_t p = P.primes[b];
    const uint64_t xp = xp_of(p);
    const uint64_t max_m = umin(xp, uz);
    const uint64_t min_m = umax(umax((uint64_t) (xhigh / p), xp / (p * p)), uz / p);
    if (min_m >= max_m) continue;
    const uint64_t max_prime = umin(uy, max_m);
    if (min_m < max_prime) {
      const uint64_t lo = pi(min_m) + 1, hi = pi(max_prime);
      add_desc(L.plus, L.n_plus, xp, ~0ull, lo, hi);
      if (hi >= lo) L.correction += (i128) (hi - lo + 1) * (2 - (i128) b);
    }
    const uint64_t max_q = umin(uy, isqrt((u128) max_m));
    const uint64_t min_q = umax(p, min_m / uy);
    if (min_q < max_q) {
      for (uint64_t i = pi(min_q) + 1; i <= pi(max_q); i++) {
        const uint64_t q = P.primes[i];
        const uint64_t min_r = umax(q, min_m / q);
        const uint64_t max_r = umin(uy, max_m / q);
        if (min_r >= max_r) continue;
        const uint64_t lo = pi(min_r) + 1, hi = pi(max_r);
        add_desc(L.minus, L.n_minus, xp / q, ~0ull, lo, hi);
        if (hi >= lo) L.correction -= (i128) (hi - lo + 1) * (2 - (i128) b);
      }
    }
  }
  // C2: max(k, pi[(x/y)^(1/3)], pi[sqrt(z)]) < b <= pi[x_star] (AC.cpp:216-346, 420-454)
  const uint64_t xy = (uint64_t) (ux / uy);
  const uint64_t min_c2 = umax(umax((uint64_t) k, pi(iroot<3>((u128) xy))), pi(sqrtz));
  for (uint64_t b = min_c2 + 1 + t; b <= pi((uint64_t) x_star); b += T) {
    const uint64_t p = P.primes[b];
    const uint64_t xp = xp_of(p);
    const uint64_t max_m = umin(xp / p, uy);
    const uint64_t min_m = umax(umax((uint64_t) (xhigh / p), xp / (p * p)), p);
    if (min_m >= max_m) continue;
    const uint64_t lo = pi(min_m) + 1, hi = pi(max_m);
    add_desc(L.plus, L.n_plus, xp, ~0ull, lo, hi);
    if (hi >= lo) L.correction += (i128) (hi - lo + 1) * (2 - (i128) b);
  }
}

/// Runs part(t, T, lists) for t = 0 .. T - 1 on T host threads (up to 16;
/// one when nb, the number of sieving primes, is small) and concatenates the
/// parts' lists: the order of the terms does not matter to the sums.
template <typename Part>
LeafLists threaded_lists(uint64_t nb, Part&& part)
{
  const uint64_t T = nb < 4096 ? 1 : (uint64_t) std::min(16, default_threads());
  std::vector<LeafLists> parts(T);
  if (T == 1) {
    part(0, 1, parts[0]);
    return std::move(parts[0]);
  }
  std::vector<std::exception_ptr> err(T);
  std::vector<std::thread> th;
  for (uint64_t t = 0; t < T; t++)
    th.emplace_back([&, t] {
      try {
        part(t, T, parts[t]);
      } catch (...) {
        err[t] = std::current_exception();
      }
    });
  for (auto& h : th) h.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
  LeafLists L;
  size_t np = 0, nm = 0;
  for (const auto& q : parts) {
    np += q.plus.size();
    nm += q.minus.size();
  }
  L.plus.reserve(np);
  L.minus.reserve(nm);
  for (const auto& q : parts) {
    for (PixLeafDesc d : q.plus) {
      d.off += L.n_plus;
      L.plus.push_back(d);
    }
    for (PixLeafDesc d : q.minus) {
      d.off += L.n_minus;
      L.minus.push_back(d);
    }
    L.n_plus += q.n_plus;
    L.n_minus += q.n_minus;
    L.correction += q.correction;
  }
  return L;
}

/// Descriptor lists of A and C, built by up to 16 host threads for large x.
LeafLists ac_lists(i128 x, int64_t y, int64_t z, int64_t k, int64_t x_star, const PrimeTables& P)
{
  return threaded_lists(P.pi(iroot<3>((u128) x)), [&](uint64_t t, uint64_t T, LeafLists& out) {
    ac_lists_part(x, y, z, k, x_star, P, t, T, out);
  });
}

/// The A/C leaves as the device takes them: the plus list as tile
/// descriptors in groups of kPixTileGroup (tiled kernel), the minus list (the
/// short q·r runs of C1) as flat descriptors.
struct AcWork {
  std::vector<PixLeafDesc> plus;      // flat plus list when a prime is >= 2^30
  uint64_t n_plus = 0;
  std::vector<PixTileDesc> tiles;
  std::vector<PixTileGroup> groups;
  uint64_t nitems = 0;
  std::vector<PixLeafDesc> minus;
  uint64_t n_minus = 0;
  i128 offset = 0;          // 3 * (sum of the plus weights) + the (2 - b) parts of C
  uint64_t leaves = 0;

  size_t bytes() const
  {
    return (plus.size() + tiles.size() + groups.size() + minus.size() + 4) * 32;
  }
};

/// max_quotient bounds every leaf quotient (the prime-count table limit).
AcWork ac_work(LeafLists&& L, const PrimeTables& P, uint64_t max_quotient)
{
  AcWork W;
  const size_t n = L.plus.size();
  W.offset = L.correction;
  W.leaves = L.n_plus + L.n_minus;
  W.minus = std::move(L.minus);
  W.n_minus = L.n_minus;
  uint64_t top = 0;
  for (size_t i = 0; i < n; i++) {
    const uint64_t cnt = (i + 1 < n ? L.plus[i + 1].off : L.n_plus) - L.plus[i].off;
    top = std::max<uint64_t>(top, P.primes[L.plus[i].i_lo + cnt - 1]);
  }
  if (top >= kPixTileMaxPrime || max_quotient >= kPixTileMaxQuotient) {
    // the tiled kernel's 32-bit residual needs primes below 2^30, its
    // quotient guard quotients below 2^43; the flat kernel is exact below 2^49
    W.plus = std::move(L.plus);
    W.n_plus = L.n_plus;
    return W;
  }
  W.tiles.resize(n);
  for (size_t i = 0; i < n; i++) {
    const PixLeafDesc& d = L.plus[i];
    const uint64_t cnt = (i + 1 < n ? L.plus[i + 1].off : L.n_plus) - d.off;
    PixTileDesc& t = W.tiles[i];
    t.xp = d.xp;
    t.xd = d.xd;
    t.i_lo = d.i_lo;
    t.i_hi = (uint32_t) (d.i_lo + cnt - 1);
    t.q1 = d.q1;
    t.pad = 0;
    // the tiled kernel leaves out the pi(v) < 3 corrections: every quotient
    // of the plus list is at least p >= 7
    if (d.xp / P.primes[t.i_hi] < 7) throw pcg_error("internal: A/C quotient below 7", kErrInternal);
    // weight-2 leaves: primes[i] > q1
    uint64_t two = 0;
    if (d.q1 < P.max) {
      const uint64_t pq = P.pi(d.q1);
      if (pq < t.i_hi) two = t.i_hi - std::max<uint64_t>((uint64_t) t.i_lo - 1, pq);
    }
    W.offset += 3 * (i128) (cnt + two);
  }
  for (size_t g0 = 0; g0 < n; g0 += kPixTileGroup) {
    PixTileGroup g;
    g.item0 = W.nitems;
    g.d0 = (uint32_t) g0;
    g.nd = (uint32_t) std::min<size_t>(kPixTileGroup, n - g0);
    g.t_lo = ~0u;
    g.t_hi = 0;
    for (size_t i = g0; i < g0 + g.nd; i++) {
      g.t_lo = std::min(g.t_lo, W.tiles[i].i_lo / 64);
      g.t_hi = std::max(g.t_hi, W.tiles[i].i_hi / 64);
    }
    W.nitems += (g.t_hi - g.t_lo + kPixTileChunk) / kPixTileChunk;
    W.groups.push_back(g);
  }
  // the last group padded with empty descriptors
  while (W.tiles.size() % kPixTileGroup) {
    PixTileDesc e{};
    e.i_lo = 1;
    e.i_hi = 0;
    e.q1 = ~0u;
    W.tiles.push_back(e);
  }
  return W;
}

/// Uploads the A/C work to dd (W.bytes() bytes).  W must stay alive until
/// the stream has passed the copies.
void upload_ac(const AcWork& W, const EngineView& v, char* dd)
{
  PixLeafDesc* dp = (PixLeafDesc*) dd;
  PixTileDesc* dt = (PixTileDesc*) (dp + W.plus.size() + 1);
  PixTileGroup* dg = (PixTileGroup*) (dt + W.tiles.size() + 1);
  PixLeafDesc* dm = (PixLeafDesc*) (dg + W.groups.size() + 1);
  if (!W.plus.empty())
    PIX_HIP(hipMemcpyAsync(dp, W.plus.data(), W.plus.size() * sizeof(PixLeafDesc),
                           hipMemcpyHostToDevice, v.stream));
  if (!W.tiles.empty())
    PIX_HIP(hipMemcpyAsync(dt, W.tiles.data(), W.tiles.size() * sizeof(PixTileDesc),
                           hipMemcpyHostToDevice, v.stream));
  if (!W.groups.empty())
    PIX_HIP(hipMemcpyAsync(dg, W.groups.data(), W.groups.size() * sizeof(PixTileGroup),
                           hipMemcpyHostToDevice, v.stream));
  if (!W.minus.empty())
    PIX_HIP(hipMemcpyAsync(dm, W.minus.data(), W.minus.size() * sizeof(PixLeafDesc),
                           hipMemcpyHostToDevice, v.stream));
}

/// The leaf sums of the A/C work uploaded to dd: plus into acc[0..1], minus
/// into acc[2..3], reading pi from tab, which holds the table words
/// [w_lo, w_hi) (the resident table: nullptr); max_q bounds every quotient.
void sum_ac(const AcWork& W, const EngineView& v, const char* dd, unsigned long long* acc,
            uint64_t max_q, const PixWord* tab = nullptr, uint64_t w_lo = 0,
            uint64_t w_hi = ~0ull)
{
  const PixLeafDesc* dp = (const PixLeafDesc*) dd;
  const PixTileDesc* dt = (const PixTileDesc*) (dp + W.plus.size() + 1);
  const PixTileGroup* dg = (const PixTileGroup*) (dt + W.tiles.size() + 1);
  const PixLeafDesc* dm = (const PixLeafDesc*) (dg + W.groups.size() + 1);
  if (!tab) tab = v.pix_tab;
  PIX_HIP(pix_launch_leafsum(dp, (uint32_t) W.plus.size(), W.n_plus, v.primes, tab, w_lo, w_hi,
                             acc, (uint32_t) v.cus, v.stream));
  PIX_HIP(pix_launch_leaftiles(dt, dg, (uint32_t) W.groups.size(), W.nitems, v.primes,
                               (uint32_t) v.nprimes, tab, w_lo, w_hi, max_q, acc,
                               (uint32_t) v.cus, v.stream));
  PIX_HIP(pix_launch_leafsum(dm, (uint32_t) W.minus.size(), W.n_minus, v.primes, tab, w_lo, w_hi,
                             acc + 2, (uint32_t) v.cus, v.stream));
}

void launch_ac(const AcWork& W, const EngineView& v, char* dd, unsigned long long* acc)
{
  upload_ac(W, v, dd);
  sum_ac(W, v, dd, acc, v.pix_n);
}

i128 ac_value(const AcWork& W, const uint64_t* h)
{
  return acc_value(h) - acc_value(h + 2) + W.offset;
}

i128 run_ac(i128 x, int64_t y, int64_t z, int64_t k, int64_t x_star, const EngineView& v,
            uint64_t* leaves)
{
  const AcWork W = ac_work(ac_lists(x, y, z, k, x_star, *v.host), *v.host, v.pix_n);
  *leaves = W.leaves;
  Acc A;
  A.d = (unsigned long long*) engine_buffer(0, 128);
  acc_zero(A, v.stream);
  char* dd = (char*) engine_buffer(4, W.bytes());
  launch_ac(W, v, dd, A.d);
  acc_read(A, v.stream);
  return ac_value(W, A.h);
}

void check_params(i128 x, int64_t y, int64_t z, int64_t k)
{
  if (x < 1 || y < 1 || z < y)
    throw pcg_error("Gourdon terms need x >= 1 and 1 <= y <= z");
  if (k < 0 || k > 8)
    throw pcg_error("k must be in [0, 8]");
  if (x > ((i128) y << 62))
    throw pcg_error("x is too large (x > y * 2^62)");
  if ((u128) z * (u128) z > (u128) x && z > 1)
    throw pcg_error("Gourdon terms need z < sqrt(x)");
}

/// zeta(s) for integers s >= 2: the first 64 terms plus the Euler-Maclaurin
/// tail (relative error far below what the R^-1 starting point needs).
long double zeta_int(int s)
{
  static long double cache[1024] = {};
  if (s < 1024 && cache[s] != 0) return cache[s];
  const int N = 64;
  long double z = 0;
  for (int n = 1; n <= N; n++) z += powl((long double) n, (long double) -s);
  z += powl((long double) N, (long double) (1 - s)) / (s - 1) - 0.5L * powl((long double) N, (long double) -s);
  if (s < 1024) cache[s] = z;
  return z;
}

/// Riemann's R(x) = 1 + sum_{k>=1} (ln x)^k / (k k! zeta(k + 1)) (Gram).
long double riemann_r(long double x)
{
  if (x < 2) return 0;
  const long double lx = logl(x);
  long double term = 1, sum = 1;
  for (int k = 1; k < 1000; k++) {
    term *= lx / k;
    const long double add = term / (k * zeta_int(k + 1));
    sum += add;
    if (k > lx && add < 1e-16L * sum) break;
  }
  return sum;
}

/// x with R(x) ~ n (Newton steps): the starting point of the nth prime
/// search.  The answer does not depend on it, only the sieve interval does
/// (the reference uses Li^-1, R^-1 or R(psi)^-1, src/nth_prime.cpp:78-88).
uint64_t riemann_r_inverse(uint64_t n)
{
  long double x = (long double) n * logl((long double) n);
  for (int it = 0; it < 100; it++) {
    const long double nx = x - (riemann_r(x) - (long double) n) * logl(x);
    if (nx < 2) { x = 2; break; }
    if (fabsl(nx - x) < 0.5L) { x = nx; break; }
    x = nx;
  }
  return (uint64_t) llroundl(x);
}

/// riemann_r_inverse for n above the 64-bit range of answers (x up to about
/// 1e23: the long double keeps x to about 1e4, far inside the sieve
/// interval around it).
u128 riemann_r_inverse_128(u128 n)
{
  const long double nd = (long double) n;
  long double x = nd * logl(nd);
  for (int it = 0; it < 100; it++) {
    const long double nx = x - (riemann_r(x) - nd) * logl(x);
    if (nx < 2) { x = 2; break; }
    const bool done = fabsl(nx - x) < 0.5L + x * 1e-18L;
    x = nx;
    if (done) break;
  }
  return (u128) roundl(x);
}

} // namespace

i128 gourdon_term(GourdonTerm term, i128 x, int64_t y, int64_t z, int64_t k, PixStats* st)
{
  check_params(x, y, z, k);
  const int64_t x_star = get_x_star_gourdon(x, y);
  const double t0 = now_s();
  auto lock = engine_lock();
  const EngineView v = engine_view(y, z, table_need(x, y, x_star));
  const double t1 = now_s();
  i128 r = 0;
  uint64_t leaves = 0;
  switch (term) {
    case GourdonTerm::AC: r = run_ac(x, y, z, k, x_star, v, &leaves); break;
    case GourdonTerm::B: r = run_B(x, y, v); break;
    case GourdonTerm::Phi0: r = run_phi0(x, y, z, k, v); break;
    case GourdonTerm::Sigma: r = run_sigma(x, y, x_star, v); break;
  }
  if (st) {
    *st = PixStats();
    st->seconds_table = t1 - t0;
    st->seconds_total = now_s() - t0;
    st->ac_leaves = leaves;
    st->table_limit = v.pix_n;
  }
  return r;
}

namespace {

/// Device words per query in pi_group: Sigma4-6 [0, 6), Phi0 [6, 8),
/// B [8, 10), A/C plus [10, 12), A/C minus [12, 14).
constexpr size_t kTermWords = 16;

/// Adds the elapsed time between two recorded events to *sec.
void add_elapsed(hipEvent_t a, hipEvent_t b, double* sec)
{
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) == hipSuccess) *sec += ms * 1e-3;
}

/// pi(x) for the queries xs[idx[j]] > 10^8 whose parameters gps[idx[j]] one
/// table set serves (view v).  Sigma, Phi0 and B of every query are launched
/// back to back on the engine stream, then the D leaves of all of them in
/// one persistent launch, then the A/C leaf sums, whose descriptor lists a
/// host thread builds while the device runs the terms before them; the
/// term sums are read back once.  With win > 0
/// the prime-count table up to `need` is built in windows of win words
/// (v's resident table covers the low part only) and B and the A/C sums run
/// once per window.  Caller holds the engine lock.
void pi_group(uint64_t max_slots, const std::vector<i128>& xs, const std::vector<size_t>& idx,
              const std::vector<GourdonParams>& gps, const EngineView& v,
              std::vector<i128>& result, PixStats& s, uint64_t need = 0, uint64_t win = 0)
{
  const size_t n = idx.size();
  if (n == 0) return;
  hipEvent_t* ev = engine_term_events();
  unsigned long long* acc = (unsigned long long*) engine_buffer(0, n * kTermWords * 8);
  uint64_t* dn = (uint64_t*) engine_buffer(1, n * 8);
  uint64_t* dpts = (uint64_t*) engine_buffer(2, n * 8);
  PIX_HIP(hsl_launch_fill_zero(acc, n * kTermWords * 8, v.stream));
  // pi(isqrt(x)) for Sigma0
  std::vector<uint64_t> pts(n);
  for (size_t j = 0; j < n; j++) pts[j] = isqrt((u128) xs[idx[j]]);
  for (uint64_t p : pts)
    if (p > v.pix_n) throw pcg_error("internal: prime-count table too small", kErrInternal);
  PIX_HIP(hipMemcpyAsync(dn, pts.data(), n * 8, hipMemcpyHostToDevice, v.stream));
  PIX_HIP(pix_launch_pi_points(dn, dpts, (uint32_t) n, v.pix_tab, v.stream));

  // A and C: descriptor lists on host threads while the device runs Sigma,
  // Phi0, B and D (the lists read only the host prime tables, which stay
  // as they are while the engine lock is held)
  const uint64_t max_q = win ? need : v.pix_n;
  std::future<std::vector<AcWork>> lists =
      std::async(std::launch::async, [&xs, &idx, &gps, &v, max_q, n] {
        std::vector<AcWork> W(n);
        for (size_t j = 0; j < n; j++) {
          const GourdonParams& gp = gps[idx[j]];
          W[j] = ac_work(ac_lists(xs[idx[j]], gp.y, gp.z, gp.k, gp.x_star, *v.host), *v.host,
                         max_q);
        }
        return W;
      });

  PIX_HIP(hipEventRecord(ev[0], v.stream));
  for (size_t j = 0; j < n; j++) {
    const GourdonParams& gp = gps[idx[j]];
    launch_sigma(xs[idx[j]], gp.y, gp.x_star, v, acc + kTermWords * j);
  }
  PIX_HIP(hipEventRecord(ev[1], v.stream));
  for (size_t j = 0; j < n; j++) {
    const GourdonParams& gp = gps[idx[j]];
    const PhiTiny& t = phi0_table(gp.k);
    launch_phi0(xs[idx[j]], gp.z, gp.k, v, t, phi_tiny_device(t, v), acc + kTermWords * j + 6);
  }
  PIX_HIP(hipEventRecord(ev[2], v.stream));
  if (!win)
    for (size_t j = 0; j < n; j++)
      launch_B(xs[idx[j]], gps[idx[j]].y, v, acc + kTermWords * j + 8);
  PIX_HIP(hipEventRecord(ev[3], v.stream));

  // D: the hard special leaves of all queries in one persistent launch
  std::vector<HslQuery> dq(n);
  for (size_t j = 0; j < n; j++) {
    const GourdonParams& gp = gps[idx[j]];
    dq[j].kind = LeafSum::D;
    dq[j].x = xs[idx[j]];
    dq[j].y = gp.y;
    dq[j].z = gp.z;
    dq[j].k = gp.k;
  }
  const std::vector<i128> d = hsl_compute_batch_locked(dq, nullptr, max_slots);
  PIX_HIP(hipEventRecord(ev[4], v.stream));

  std::vector<AcWork> L = lists.get();
  size_t bytes = 0;
  for (size_t j = 0; j < n; j++) {
    bytes += L[j].bytes();
    s.ac_leaves += L[j].leaves;
  }
  char* dd0 = (char*) engine_buffer(4, bytes + 64);
  char* dd = dd0;
  for (size_t j = 0; j < n; j++) {
    upload_ac(L[j], v, dd);
    if (!win) {
      sum_ac(L[j], v, dd, acc + kTermWords * j + 10, max_q);
      // a pass point after each query's A/C launch: other work queued now
      // runs after it
      if (j + 1 < n) engine_pass_point();
    }
    dd += L[j].bytes();
  }
  if (win)
    s.table_windows += for_windows(need, win, &s.seconds_table,
                                   [&](const PixWord* tab, uint64_t w_lo, uint64_t w_hi) {
      dd = dd0;
      for (size_t j = 0; j < n; j++) {
        launch_B(xs[idx[j]], gps[idx[j]].y, v, acc + kTermWords * j + 8, tab, w_lo, w_hi);
        sum_ac(L[j], v, dd, acc + kTermWords * j + 10, max_q, tab, w_lo, w_hi);
        dd += L[j].bytes();
      }
    });
  PIX_HIP(hipEventRecord(ev[5], v.stream));

  std::vector<uint64_t> h(n * kTermWords), pis(n);
  PIX_HIP(hipMemcpyAsync(h.data(), acc, n * kTermWords * 8, hipMemcpyDeviceToHost, v.stream));
  PIX_HIP(hipMemcpyAsync(pis.data(), dpts, n * 8, hipMemcpyDeviceToHost, v.stream));
  PIX_HIP(hipStreamSynchronize(v.stream));
  add_elapsed(ev[0], ev[1], &s.seconds_sigma);
  add_elapsed(ev[1], ev[2], &s.seconds_phi0);
  add_elapsed(ev[2], ev[3], &s.seconds_b);
  add_elapsed(ev[3], ev[4], &s.seconds_d);
  add_elapsed(ev[4], ev[5], &s.seconds_ac);
  const PrimeTables& P = *v.host;
  for (size_t j = 0; j < n; j++) {
    const i128 x = xs[idx[j]];
    const GourdonParams& gp = gps[idx[j]];
    const uint64_t* hj = h.data() + kTermWords * j;
    const uint64_t x13 = iroot<3>((u128) x);
    const uint64_t sqrt_xy = isqrt((u128) x / (u128) gp.y);
    const uint64_t top = umax(umax((uint64_t) gp.y, x13), umax(sqrt_xy, (uint64_t) gp.x_star));
    if (top > P.max) throw pcg_error("internal: prime table too small for Sigma", kErrInternal);
    const i128 sigma = sigma_value(P.pi((uint64_t) gp.y), P.pi(x13), P.pi(sqrt_xy),
                                   P.pi((uint64_t) gp.x_star), (i128) pis[j], hj);
    const i128 phi0 = phi_tiny_host(x, phi0_table(gp.k)) + acc_value(hj + 6);
    const i128 ac = ac_value(L[j], hj + 10);
    result[idx[j]] = sigma + phi0 + ac - acc_value(hj + 8) + d[j];
  }
}

} // namespace

std::vector<i128> pi_gourdon_batch(const std::vector<i128>& xs, PixStats* st, uint64_t max_slots)
{
  PixStats s;
  const double t0 = now_s();
  std::vector<i128> result(xs.size(), 0);
  // Small x: the prime-count table answers directly (it is built for at
  // least 2^20 numbers, with the primes up to the square root of that).
  // Larger x: one table group for the queries with y >= z (prefix-valid
  // tables of the largest), one group per query with y < z (its
  // FactorTableD depends on y).
  std::vector<GourdonParams> gps(xs.size());
  std::vector<size_t> small, shared;
  std::vector<std::vector<size_t>> alone;
  int64_t ymax = 0;
  uint64_t need = 0;
  for (size_t i = 0; i < xs.size(); i++) {
    const i128 x = xs[i];
    if (x < 2) continue;
    if (x <= (i128) 100000000) {
      small.push_back(i);
      ymax = std::max<int64_t>(ymax, (int64_t) isqrt((u128) std::max<i128>(x, (i128) 1 << 20)) + 1);
      need = std::max<uint64_t>(need, (uint64_t) x);
      continue;
    }
    gps[i] = gourdon_params(x);
    check_params(x, gps[i].y, gps[i].z, gps[i].k);
    if (gps[i].y >= gps[i].z) {
      shared.push_back(i);
      ymax = std::max<int64_t>(ymax, gps[i].z);
      need = std::max<uint64_t>(need, table_need(x, gps[i].y, gps[i].x_star));
    } else {
      alone.push_back({i});
    }
  }
  if (!small.empty() || !shared.empty()) {
    auto lock = engine_lock();
    // a table over the budget: the low part whole, the rest in windows
    const uint64_t win = window_words(need);
    uint64_t low = 0;
    if (win) {
      for (size_t i : shared) low = std::max<uint64_t>(low, table_low_need(xs[i], gps[i].y, gps[i].x_star));
      for (size_t i : small) low = std::max<uint64_t>(low, (uint64_t) xs[i]);
    }
    const EngineView v = engine_view(ymax, ymax, win ? low : need);
    s.seconds_table += now_s() - t0;
    s.table_limit = win ? need : v.pix_n;
    s.table_windows = win ? 0 : 1;
    if (!small.empty()) {
      std::vector<uint64_t> ns;
      for (size_t i : small) ns.push_back((uint64_t) xs[i]);
      const std::vector<uint64_t> p = pi_points(v, ns);
      for (size_t j = 0; j < small.size(); j++) result[small[j]] = (i128) p[j];
    }
    pi_group(max_slots, xs, shared, gps, v, result, s, need, win);
  }
  for (const auto& g : alone) {
    const GourdonParams& gp = gps[g[0]];
    auto lock = engine_lock();
    const double t = now_s();
    const uint64_t gneed = table_need(xs[g[0]], gp.y, gp.x_star);
    const uint64_t win = window_words(gneed);
    const EngineView v = engine_view(gp.y, gp.z, win ? table_low_need(xs[g[0]], gp.y, gp.x_star) : gneed);
    s.seconds_table += now_s() - t;
    s.table_limit = std::max<uint64_t>((uint64_t) s.table_limit, win ? gneed : (uint64_t) v.pix_n);
    if (!win) s.table_windows = std::max<uint64_t>(s.table_windows, 1);
    pi_group(max_slots, xs, g, gps, v, result, s, gneed, win);
  }
  // the reference checks every pi(x) against Schoenfeld's bound
  // (src/gourdon/pi_gourdon.cpp:89,160)
  for (size_t i = 0; i < xs.size(); i++) verify_pix("pi_gourdon", xs[i], result[i]);
  s.seconds_total = now_s() - t0;
  if (st) *st = s;
  return result;
}

i128 pi_gourdon(i128 x, PixStats* st)
{
  return pi_gourdon_batch({x}, st)[0];
}

void pi_validate(i128 x)
{
  if (x < 0) throw pcg_error("pi(x): x must be >= 0");
  if (x <= (i128) 100000000) return;
  const GourdonParams gp = gourdon_params(x);
  check_params(x, gp.y, gp.z, gp.k);
  (void) table_need(x, gp.y, gp.x_star);
}

/// Upper estimate of pi(n): n / (ln n - 1.1) for n >= 100 (Rosser-Schoenfeld
/// type bound; pi(n) < 1.26 n / ln n), small n counted exactly enough.
static uint64_t pi_upper(uint64_t n)
{
  if (n < 100) return 26;
  const double l = std::log((double) n);
  return (uint64_t) (1.26 * (double) n / l) + 1;
}

uint64_t table_bytes_D(int64_t y, int64_t z)
{
  const uint64_t pmax = (uint64_t) std::max<int64_t>(y, (int64_t) isqrt((u128) z)) + 1;
  const uint64_t ft = (uint64_t) ft_to_index((uint64_t) z) + 1;
  return pi_upper(pmax) * 4 + (pmax / 240 + 1) * 12 + ft * (z > kFactorTableD16Max ? 4 : 2);
}

uint64_t pi_footprint(i128 x)
{
  if (x <= (i128) 100000000) return (uint64_t) 100000000 / 240 * sizeof(PixWord) + (64u << 20);
  const GourdonParams gp = gourdon_params(x);
  const uint64_t need = table_need(x, gp.y, gp.x_star);
  // prime-count table (one 16-byte word per 240 numbers), the D tables, and
  // the per-query scratch: A/C descriptors and term sums (under 64 MiB up to
  // x = 1e19, about x^(1/3) growth beyond)
  const double scale = std::max(1.0, std::cbrt((double) x / 1e19));
  // a table over the budget: its low part plus one window
  const uint64_t win = window_words(need);
  const uint64_t table = win ? (table_low_need(x, gp.y, gp.x_star) / 240 + 1 + win) * sizeof(PixWord)
                             : (need / 240 + 1) * sizeof(PixWord);
  return table + table_bytes_D(gp.y, gp.z) + (uint64_t) (scale * (64u << 20));
}

uint64_t nth_prime_footprint(uint64_t n)
{
  if (n <= 5761455) return pi_footprint(100000000);
  // pi(x0) at x0 ~ n ln n plus the interval sieve (bitmap and counts)
  const double l = std::log((double) n);
  const double x0 = (double) n * (l + std::log(l));
  return pi_footprint((i128) x0) + (256u << 20);
}

// ---------------------------------------------------------------------------
// Deleglise-Rivat: pi(x) = S1 + S2_trivial + S2_easy + S2_hard + pi(y) - 1 - P2
// (src/deleglise-rivat/pi_deleglise_rivat.cpp:40-53, 86-92)

namespace {

/// S2_easy leaves as descriptors (src/deleglise-rivat/S2_easy.cpp:53-160):
/// for max(c, pi(sqrt y)) < b <= pi(x^(1/3)) the leaves pi(x / (p_b q)) - b + 2
/// over the primes q of the clustered range (A, B] = (min_clustered,
/// min_trivial] and then of the sparse range (min_sparse, i].  The sparse
/// leaves are summed one by one.  The clustered ones have q > sqrt(x / p_b), so
/// their quotients v = x / (p_b q) are below sqrt(x / p_b) and repeat; the
/// reference walks the runs of equal pi(v).  Here the same sum is counted from
/// the other side, over the pairs of primes (q, r) with r q <= x / p_b:
///   sum_{A < q <= B} pi(xp / q) = pi(xp / B) (pi(B) - pi(A))
///       + sum_{xp / B < r <= xp / (A + 1)} (pi(xp / r) - pi(A)),
/// one leaf pi(xp / r) per prime r below sqrt(x / p_b) instead of one per q
/// (at x = 1e19: 3.2e10 leaves instead of 2.0e11); the reflection that
/// src/gourdon/AC.cpp:271-320 applies to the clustered leaves of C2.  The -b + 2 parts and the
/// pi(A), pi(B) products go to the correction.
void s2_easy_part(i128 x, int64_t y, int64_t z, int64_t c, const PrimeTables& P, uint64_t t,
                  uint64_t T, LeafLists& L)
{
  const u128 ux = (u128) x;
  const uint64_t uy = (uint64_t) y, uz = (uint64_t) z;
  auto pi = [&](uint64_t n) -> uint64_t {
    if (n > P.max) throw pcg_error("internal: prime table too small for S2_easy", kErrInternal);
    return P.pi(n);
  };
  const uint64_t pi_sqrty = pi(isqrt((u128) uy));
  const uint64_t pi_x13 = pi(iroot<3>(ux));
  for (uint64_t b = umax((uint64_t) c, pi_sqrty) + 1 + t; b <= pi_x13; b += T) {
    const uint64_t prime = P.primes[b];
    const u128 xp128 = ux / prime;
    if (xp128 >> 64) throw pcg_error("S2_easy: x / p does not fit in 64 bits");
    const uint64_t xp = (uint64_t) xp128;
    const uint64_t min_trivial = umin(xp / prime, uy);
    const uint64_t min_clustered = in_between<uint64_t>(prime, isqrt((u128) xp), uy);
    const uint64_t min_sparse = in_between<uint64_t>(prime, uz / prime, uy);
    uint64_t i = pi(min_trivial);
    const uint64_t pi_min_clustered = pi(min_clustered);
    const uint64_t pi_min_sparse = pi(min_sparse);
    if (i > pi_min_clustered) {
      // q in (A, B]: A = min_clustered >= isqrt(xp), B = min_trivial
      const uint64_t nq = i - pi_min_clustered;
      const uint64_t r_lo = pi(xp / min_trivial);            // pi(xp / B)
      const uint64_t r_hi = pi(xp / (min_clustered + 1));    // pi(xp / (A + 1))
      L.correction += (i128) r_lo * nq + (i128) nq * (2 - (i128) b);
      if (r_hi > r_lo) {
        add_desc(L.plus, L.n_plus, xp, ~0ull, r_lo + 1, r_hi);
        L.correction -= (i128) (r_hi - r_lo) * (i128) pi_min_clustered;
      }
      i = pi_min_clustered;
    }
    if (i > pi_min_sparse) {
      add_desc(L.plus, L.n_plus, xp, ~0ull, pi_min_sparse + 1, i);
      L.correction += (i128) (i - pi_min_sparse) * (2 - (i128) b);
    }
  }
}

LeafLists s2_easy_lists(i128 x, int64_t y, int64_t z, int64_t c, const PrimeTables& P)
{
  return threaded_lists(P.pi(iroot<3>((u128) x)), [&](uint64_t t, uint64_t T, LeafLists& out) {
    s2_easy_part(x, y, z, c, P, t, T, out);
  });
}

/// S2_trivial (src/deleglise-rivat/S2_trivial.cpp:30-77), on the host: for
/// the primes max(p_c, sqrt z) < p < y with x / p^2 > p, pi(y) - pi(x / p^2);
/// from the first other prime below y on, the arithmetic series of pi(y) - b.
i128 s2_trivial_host(i128 x, int64_t y, int64_t z, int64_t c, const PrimeTables& P)
{
  if (y < 2) return 0;
  const uint64_t uy = (uint64_t) y;
  if (uy > P.max) throw pcg_error("internal: prime table too small for S2_trivial", kErrInternal);
  const i128 pi_y = (i128) P.pi(uy);
  const uint64_t sqrtz = isqrt((u128) z);
  const uint64_t prime_c = P.primes[(size_t) c];
  const uint64_t start = umax(prime_c, sqrtz) + 1;
  if (start >= uy) return 0;
  i128 sum = 0;
  uint64_t prime = uy;                 // the prime the loop stopped at (>= y: none)
  for (uint64_t b = P.pi(start - 1) + 1; b < P.primes.size(); b++) {
    const uint64_t p = P.primes[b];
    if (p >= uy) break;
    const uint64_t xpp = (uint64_t) ((u128) x / ((u128) p * p));
    if (xpp <= p) {
      prime = p;
      break;
    }
    sum += pi_y - (i128) P.pi(xpp);
  }
  if (prime < uy) {
    const i128 n = (i128) P.pi(uy - 1) - (i128) P.pi(prime) + 1;
    const i128 a1 = pi_y - (i128) P.pi(uy - 1);
    const i128 a2 = pi_y - (i128) P.pi(prime);
    sum += n * (a1 + a2) / 2;
  }
  return sum;
}

/// P2 (src/P2.cpp:95-127): the sum of pi(x / p) over y < p <= sqrt(x) (the
/// B kernel) plus (a - 2)(a + 1)/2 - (b - 2)(b + 1)/2, a = pi(y), b = pi(sqrt x).
i128 run_p2(i128 x, int64_t y, const EngineView& v)
{
  if (x < 4) return 0;
  const uint64_t sqrtx = isqrt((u128) x);
  if ((uint64_t) y >= sqrtx) return 0;
  const i128 a = (i128) v.host->pi((uint64_t) y);
  const i128 b = (i128) pi_points(v, {sqrtx})[0];
  return (a - 2) * (a + 1) / 2 - (b - 2) * (b + 1) / 2 + run_B(x, y, v);
}

/// S1 (src/S1.cpp:29-70): phi(x, c) plus the sum of mu(n) phi(x / n, c) over
/// the square-free n <= y with least prime factor above p_c, from the lpf
/// factor table of the S2_hard tables.
i128 run_s1(i128 x, int64_t y, int64_t c, const EngineView& v)
{
  static const uint32_t small[8] = {2, 3, 5, 7, 11, 13, 17, 19};
  if (c < 5 || c > 8)
    throw pcg_error("S1(x, y, c): the device path needs 5 <= c <= 8");
  const PhiTiny& t = phi_tiny_table((uint32_t) c);
  const uint32_t* dtbl = phi_tiny_device(t, v);
  const uint64_t nidx = (uint64_t) ft_to_index((uint64_t) y) + 1;
  if (nidx > v.factor_n) throw pcg_error("internal: factor table too small for S1", kErrInternal);
  Acc A;
  A.d = (unsigned long long*) engine_buffer(0, 128);
  acc_zero(A, v.stream);
  PIX_HIP(pix_launch_phi0((uint64_t) x, (uint64_t) ((u128) x >> 64), v.factor, v.factor32, true,
                          nidx, (uint32_t) c, small[c - 1], dtbl, t.pp, t.phi_pp, A.d, v.stream));
  acc_read(A, v.stream);
  return phi_tiny_host(x, t) + acc_value(A.h);
}

/// S2_easy from its descriptor lists W (built by ac_work over s2_easy_lists).
i128 run_s2_easy_lists(const AcWork& W, const EngineView& v)
{
  Acc A;
  A.d = (unsigned long long*) engine_buffer(0, 128);
  acc_zero(A, v.stream);
  char* dd = (char*) engine_buffer(4, W.bytes());
  launch_ac(W, v, dd, A.d);
  acc_read(A, v.stream);
  return ac_value(W, A.h);
}

i128 run_s2_easy(i128 x, int64_t y, int64_t z, int64_t c, const EngineView& v)
{
  return run_s2_easy_lists(ac_work(s2_easy_lists(x, y, z, c, *v.host), *v.host, v.pix_n), v);
}

/// S2_easy - P2 with the prime-count table up to `need` built in windows of
/// win words (v's resident table covers [0, max(sqrt(x), y)]).
i128 s2_easy_minus_p2_windowed(i128 x, int64_t y, int64_t z, int64_t c, const EngineView& v,
                               uint64_t need, uint64_t win)
{
  const AcWork W = ac_work(s2_easy_lists(x, y, z, c, *v.host), *v.host, need);
  Acc A;                      // S2_easy plus [0, 2), minus [2, 4), P2's sum [4, 6)
  A.d = (unsigned long long*) engine_buffer(0, 128);
  acc_zero(A, v.stream);
  char* dd = (char*) engine_buffer(4, W.bytes());
  upload_ac(W, v, dd);
  const uint64_t sqrtx = isqrt((u128) x);
  const bool p2 = x >= 4 && (uint64_t) y < sqrtx;
  double table_s = 0;
  for_windows(need, win, &table_s, [&](const PixWord* tab, uint64_t w_lo, uint64_t w_hi) {
    sum_ac(W, v, dd, A.d, need, tab, w_lo, w_hi);
    if (p2) launch_B(x, y, v, A.d + 4, tab, w_lo, w_hi);
  });
  acc_read(A, v.stream);
  i128 r = ac_value(W, A.h);
  if (p2) {
    const i128 a = (i128) v.host->pi((uint64_t) y);
    const i128 b = (i128) pi_points(v, {sqrtx})[0];
    r -= (a - 2) * (a + 1) / 2 - (b - 2) * (b + 1) / 2 + acc_value(A.h + 4);
  }
  return r;
}

void check_dr(i128 x, int64_t y, int64_t z, int64_t c)
{
  if (x < 1 || y < 1 || z < 1)
    throw pcg_error("Deleglise-Rivat terms need x, y, z >= 1");
  if (c < 0 || c > 8)
    throw pcg_error("c must be in [0, 8]");
  if (x > ((i128) y << 62))
    throw pcg_error("x is too large (x > y * 2^62)");
  if ((uint64_t) y >= (1ull << 32) || (uint64_t) z >= (1ull << 50))
    throw pcg_error("Deleglise-Rivat terms need y < 2^32 and z < 2^50");
}

/// Largest n the terms read pi(n) for: x / y (P2), sqrt(x), y.
uint64_t dr_table_need(i128 x, int64_t y)
{
  u128 n = (u128) x / (u128) y;
  n = std::max<u128>(n, isqrt((u128) x));
  n = std::max<u128>(n, (u128) y);
  if (n >= ((u128) 1 << 49))
    throw pcg_error("pi(x): x / y must be < 2^49 (prime-count table limit)");
  return (uint64_t) n + 1;
}

} // namespace

i128 dr_term(DRTerm term, i128 x, int64_t y, int64_t z, int64_t c)
{
  check_dr(x, y, z, c);
  auto lock = engine_lock();
  const EngineView v =
      engine_view(y, z, dr_table_need(x, y), LeafSum::S2_hard, (uint64_t) y + 1);
  if ((uint64_t) y > v.host->max) throw pcg_error("internal: prime table below y", kErrInternal);
  switch (term) {
    case DRTerm::S1: return run_s1(x, y, c, v);
    case DRTerm::S2_trivial: return s2_trivial_host(x, y, z, c, *v.host);
    case DRTerm::S2_easy: return run_s2_easy(x, y, z, c, v);
    case DRTerm::P2: return run_p2(x, y, v);
  }
  return 0;
}

i128 pi_deleglise_rivat(i128 x)
{
  if (x < 2) return 0;
  if (x <= (i128) 100000000) return pi_gourdon(x, nullptr);     // the table answers
  const DRParams p = deleglise_rivat_params(x);
  check_dr(x, p.y, p.z, p.c);
  auto lock = engine_lock();
  // a table over the budget: [0, max(sqrt(x), y)] whole, the rest in windows
  const uint64_t need = dr_table_need(x, p.y);
  const uint64_t win = window_words(need);
  const uint64_t low = std::max<uint64_t>(isqrt((u128) x), (uint64_t) p.y) + 1;
  const EngineView v = engine_view(p.y, p.z, win ? low : need, LeafSum::S2_hard,
                                   (uint64_t) p.y + 1);
  if ((uint64_t) p.y > v.host->max) throw pcg_error("internal: prime table below y", kErrInternal);
  i128 r = run_s1(x, p.y, p.c, v);
  r += s2_trivial_host(x, p.y, p.z, p.c, *v.host);
  // without windows the S2_easy descriptor lists are built on host threads
  // while the device runs S2_hard and P2 (they read only the host prime
  // tables, which stay as they are while the engine lock is held)
  std::future<AcWork> easy;
  if (!win)
    easy = std::async(std::launch::async, [&x, &p, &v] {
      return ac_work(s2_easy_lists(x, p.y, p.z, p.c, *v.host), *v.host, v.pix_n);
    });
  HslQuery q;
  q.kind = LeafSum::S2_hard;
  q.x = x;
  q.y = p.y;
  q.z = p.z;
  q.k = p.c;
  r += hsl_compute_batch_locked({q}, nullptr)[0];
  if (win) {
    r += s2_easy_minus_p2_windowed(x, p.y, p.z, p.c, v, need, win);
  } else {
    r -= run_p2(x, p.y, v);
    const AcWork W = easy.get();
    r += run_s2_easy_lists(W, v);
  }
  r += (i128) v.host->pi((uint64_t) p.y) - 1;
  // src/deleglise-rivat/pi_deleglise_rivat.cpp:92,138
  verify_pix("pi_deleglise_rivat", x, r);
  return r;
}

std::vector<uint64_t> nth_prime_batch(const std::vector<uint64_t>& ns, PixStats* st,
                                      uint64_t max_slots)
{
  // primes below 2^63, as the reference's 64-bit nth_prime (src/nth_prime.cpp:95-100)
  const uint64_t max_n = 216289611853439384ull;
  for (uint64_t n : ns) {
    if (n < 1) throw pcg_error("nth_prime(n): n must be >= 1");
    if (n > max_n) throw pcg_error("nth_prime(n): n must be <= 216289611853439384");
  }
  const double t0 = now_s();
  PixStats s;
  std::vector<uint64_t> result(ns.size(), 0);
  std::vector<size_t> table_i, big_i;
  for (size_t i = 0; i < ns.size(); i++) {
    const uint64_t n = ns[i];
    if (n <= 3) result[i] = n == 1 ? 2 : (n == 2 ? 3 : 5);
    else if (n <= 5761455) table_i.push_back(i);           // pi(10^8): the table answers
    else big_i.push_back(i);
  }
  if (!table_i.empty()) {
    auto lock = engine_lock();
    const EngineView v = engine_view(10001, 10001, 100000000);
    uint64_t* dout = (uint64_t*) engine_buffer(2, 8);
    for (size_t i : table_i) {
      PIX_HIP(pix_launch_select(v.pix_tab, v.pix_n / 240 + 1, ns[i], dout, v.stream));
      PIX_HIP(hipMemcpyAsync(&result[i], dout, 8, hipMemcpyDeviceToHost, v.stream));
      PIX_HIP(hipStreamSynchronize(v.stream));
    }
  }
  if (big_i.empty()) {
    s.seconds_total = now_s() - t0;
    if (st) *st = s;
    return result;
  }
  // answers from 2^62 on (n above about 1.1e17): the sieve interval anchored
  // at a 128-bit base, one query at a time
  {
    std::vector<size_t> keep;
    for (size_t i : big_i) {
      if (riemann_r_inverse(ns[i]) < ((uint64_t) 1 << 62)) {
        keep.push_back(i);
      } else {
        PixStats s1;
        result[i] = (uint64_t) nth_prime_128((i128) ns[i], &s1);
        s.table_limit = std::max(s.table_limit, s1.table_limit);
      }
    }
    big_i.swap(keep);
  }
  if (big_i.empty()) {
    s.seconds_total = now_s() - t0;
    if (st) *st = s;
    return result;
  }
  // pi at the starting points of all queries in one batch
  std::vector<i128> x0s;
  for (size_t i : big_i) x0s.push_back((i128) riemann_r_inverse(ns[i]));
  const std::vector<i128> cs = pi_gourdon_batch(x0s, &s, max_slots);
  int64_t ymax = 0;
  for (i128 x0 : x0s) ymax = std::max<int64_t>(ymax, gourdon_params(x0).z);
  auto lock = engine_lock();
  for (size_t j = 0; j < big_i.size(); j++) {
    const uint64_t n = ns[big_i[j]];
    const uint64_t x0 = (uint64_t) x0s[j];
    const uint64_t c = (uint64_t) cs[j];
    uint64_t res = 0;
    const bool down = c >= n;
    const uint64_t k = down ? c - n + 1 : n - c;   // rank from x0 downward / upward
    const double lx = std::log((double) x0);
    uint64_t delta = (uint64_t) ((double) (k + 16) * lx * 2.0) + 100000;
    for (;;) {
      uint64_t lo, hi;
      if (down) {
        hi = x0;
        lo = x0 > delta + 2 ? x0 - delta : 2;
      } else {
        lo = x0 + 1;
        hi = x0 + delta;
      }
      const uint64_t r = isqrt((u128) hi);
      const EngineView v = engine_view(ymax, ymax, r + 1);
      const std::vector<uint64_t> pr = pi_points(v, {r, 15359});
      const uint64_t n_all = pr[0] >= 3 ? pr[0] - 3 : 0;       // sieving primes >= 7
      const uint64_t n_small = std::min<uint64_t>(pr[1] - 3, n_all);
      uint32_t* sp = (uint32_t*) engine_buffer(5, (n_all + 1) * 4);
      if (n_all) PIX_HIP(pix_launch_extract(v.pix_tab, r, sp, v.stream));
      const uint64_t nwords = hi / 240 - lo / 240 + 1;
      const uint64_t nblk = (nwords + kPixSegWords - 1) / kPixSegWords;
      uint64_t* bits = (uint64_t*) engine_buffer(6, nwords * 8);
      uint64_t* cnts = (uint64_t*) engine_buffer(7, nblk * 8);
      PIX_HIP(pix_launch_interval(bits, lo, hi, sp, (uint32_t) n_small, n_all, v.stream));
      PIX_HIP(pix_launch_block_counts(bits, nwords, cnts, v.stream));
      std::vector<uint64_t> bc(nblk);
      PIX_HIP(hipMemcpyAsync(bc.data(), cnts, nblk * 8, hipMemcpyDeviceToHost, v.stream));
      PIX_HIP(hipStreamSynchronize(v.stream));
      uint64_t total = 0;
      for (uint64_t b : bc) total += b;
      if (total < k) {
        if (down && lo == 2) throw pcg_error("internal: nth_prime interval", kErrInternal);
        delta *= 2;
        continue;
      }
      // block holding the k-th prime from the interval's top (down) or bottom
      uint64_t need = k, blk = 0;
      if (down) {
        for (blk = nblk; blk-- > 0;) {
          if (bc[blk] >= need) break;
          need -= bc[blk];
        }
      } else {
        for (blk = 0; blk < nblk; blk++) {
          if (bc[blk] >= need) break;
          need -= bc[blk];
        }
      }
      const uint64_t w0 = blk * kPixSegWords;
      const uint64_t nw = std::min<uint64_t>(kPixSegWords, nwords - w0);
      std::vector<uint64_t> words(nw);
      PIX_HIP(hipMemcpyAsync(words.data(), bits + w0, nw * 8, hipMemcpyDeviceToHost, v.stream));
      PIX_HIP(hipStreamSynchronize(v.stream));
      static const uint8_t wheel[8] = {1, 7, 11, 13, 17, 19, 23, 29};
      const uint64_t base = (lo / 240 + w0) * 240;
      for (uint64_t i = 0; i < nw && !res; i++) {
        const uint64_t wi = down ? nw - 1 - i : i;
        const uint64_t wv = words[wi];
        const uint64_t c1 = (uint64_t) __builtin_popcountll(wv);
        if (c1 < need) { need -= c1; continue; }
        for (int jb = 0; jb < 64; jb++) {
          const int bit = down ? 63 - jb : jb;
          if (!((wv >> bit) & 1)) continue;
          if (--need == 0) {
            res = base + 240 * wi + 30 * (bit >> 3) + wheel[bit & 7];
            break;
          }
        }
      }
      if (!res) throw pcg_error("internal: nth_prime selection", kErrInternal);
      break;
    }
    result[big_i[j]] = res;
  }
  s.seconds_total = now_s() - t0;
  if (st) *st = s;
  return result;
}

uint64_t nth_prime(uint64_t n, PixStats* st)
{
  return nth_prime_batch({n}, st)[0];
}

i128 nth_prime_128(i128 n, PixStats* st)
{
  // src/nth_prime.cpp:141-175: the 64-bit path below 2^63, then pi at a
  // close estimate and a sieve of the interval between it and the answer
  const i128 max_n = (i128) 216289611853439384ull;
  if (n < 1) throw pcg_error("nth_prime(n): n must be >= 1");
  if (n <= max_n && riemann_r_inverse((uint64_t) n) < ((uint64_t) 1 << 62))
    return (i128) nth_prime((uint64_t) n, st);
  const double t0 = now_s();
  const u128 x0 = riemann_r_inverse_128((u128) n);
  pi_validate((i128) x0);
  PixStats s;
  const i128 c = pi_gourdon_batch({(i128) x0}, &s)[0];
  const int64_t ymax = gourdon_params((i128) x0).z;
  const bool down = c >= n;
  const u128 k128 = down ? (u128) (c - n + 1) : (u128) (n - c);   // rank from x0 down / up
  if (k128 >= ((u128) 1 << 50)) throw pcg_error("internal: nth_prime estimate", kErrInternal);
  const uint64_t k = (uint64_t) k128;
  const double lx = std::log((double) x0);
  uint64_t delta = (uint64_t) ((double) (k + 16) * lx * 2.0) + 100000;
  auto lock = engine_lock();
  u128 res = 0;
  for (;;) {
    if (delta >= ((uint64_t) 1 << 61)) throw pcg_error("internal: nth_prime interval", kErrInternal);
    const u128 lo = down ? x0 - delta : x0 + 1;
    const u128 hi = down ? x0 : x0 + delta;
    const u128 L = lo - lo % 240;                   // a multiple of 240 (and of 30)
    const uint64_t o_lo = (uint64_t) (lo - L), o_hi = (uint64_t) (hi - L);
    const uint64_t r = (uint64_t) isqrt(hi);
    const EngineView v = engine_view(ymax, ymax, r + 1);
    const std::vector<uint64_t> pr = pi_points(v, {r, 15359});
    const uint64_t n_all = pr[0] >= 3 ? pr[0] - 3 : 0;       // sieving primes >= 7
    const uint64_t n_small = std::min<uint64_t>(pr[1] - 3, n_all);
    uint64_t* sp = (uint64_t*) engine_buffer(5, (n_all + 1) * 8);
    if (n_all) PIX_HIP(pix_launch_extract64(v.pix_tab, r, sp, v.stream));
    const uint64_t nwords = o_hi / 240 + 1;
    const uint64_t nblk = (nwords + kPixSegWords - 1) / kPixSegWords;
    uint64_t* bits = (uint64_t*) engine_buffer(6, nwords * 8);
    uint64_t* cnts = (uint64_t*) engine_buffer(7, nblk * 8);
    PIX_HIP(pix_launch_interval128(bits, (uint64_t) (L >> 64), (uint64_t) L, o_lo, o_hi, sp,
                                   (uint32_t) n_small, n_all, v.stream));
    PIX_HIP(pix_launch_block_counts(bits, nwords, cnts, v.stream));
    std::vector<uint64_t> bc(nblk);
    PIX_HIP(hipMemcpyAsync(bc.data(), cnts, nblk * 8, hipMemcpyDeviceToHost, v.stream));
    PIX_HIP(hipStreamSynchronize(v.stream));
    uint64_t total = 0;
    for (uint64_t b : bc) total += b;
    if (total < k) {
      delta *= 2;
      continue;
    }
    // the block, then the word and bit, of the k-th prime from the top
    // (down) or the bottom of the interval
    uint64_t need = k, blk = 0;
    if (down) {
      for (blk = nblk; blk-- > 0;) {
        if (bc[blk] >= need) break;
        need -= bc[blk];
      }
    } else {
      for (blk = 0; blk < nblk; blk++) {
        if (bc[blk] >= need) break;
        need -= bc[blk];
      }
    }
    const uint64_t w0 = blk * kPixSegWords;
    const uint64_t nw = std::min<uint64_t>(kPixSegWords, nwords - w0);
    std::vector<uint64_t> words(nw);
    PIX_HIP(hipMemcpyAsync(words.data(), bits + w0, nw * 8, hipMemcpyDeviceToHost, v.stream));
    PIX_HIP(hipStreamSynchronize(v.stream));
    static const uint8_t wheel[8] = {1, 7, 11, 13, 17, 19, 23, 29};
    for (uint64_t i = 0; i < nw && !res; i++) {
      const uint64_t wi = down ? nw - 1 - i : i;
      const uint64_t wv = words[wi];
      const uint64_t c1 = (uint64_t) __builtin_popcountll(wv);
      if (c1 < need) { need -= c1; continue; }
      for (int jb = 0; jb < 64; jb++) {
        const int bit = down ? 63 - jb : jb;
        if (!((wv >> bit) & 1)) continue;
        if (--need == 0) {
          res = L + (u128) (240 * (w0 + wi) + 30 * (bit >> 3) + wheel[bit & 7]);
          break;
        }
      }
    }
    if (!res) throw pcg_error("internal: nth_prime selection", kErrInternal);
    break;
  }
  s.seconds_total = now_s() - t0;
  if (st) *st = s;
  return (i128) res;
}

} // namespace pcg
