// engine.cpp -- host orchestration of the hard-special-leaf engine:
// validation, table building and upload, the segment/chunk plan, the kernel
// launches and the final 128-bit reduction.
//
// Replaces the OpenMP driver of the reference (D_OpenMP, src/gourdon/D.cpp:
// 382-419; S2_hard_OpenMP, src/deleglise-rivat/S2_hard.cpp:181-215) and its
// LoadBalancerS2 (src/LoadBalancerS2.cpp:65-344): instead of threads taking
// growing [low, low + segments * segment_size) intervals, the whole interval
// is cut up front into chunks of equal estimated work that a persistent grid
// of workgroups takes from a device work queue.

#include "engine.hpp"
#include "hsl_kernels.hpp"
#include "pix_kernels.hpp"
#include "params.hpp"
#include "tables.hpp"
#include "tuner.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <functional>
#include <numeric>

namespace pcg {

#define PCG_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw pcg_error(std::string("HIP error '") + hipGetErrorString(e_) +         \
                      "' in " #expr,                                               \
                      e_ == hipErrorOutOfMemory ? kErrNoMem : kErrDevice);         \
  } while (0)

EngineConfig& engine_config()
{
  static EngineConfig cfg = [] {
    EngineConfig c;
    // PCG_AUTOTUNE=0|1|2 sets the initial schedule mode (pcg_set_autotune)
    if (const char* e = std::getenv("PCG_AUTOTUNE"); e && e[0] >= '0' && e[0] <= '2' && !e[1])
      c.tune_mode = e[0] - '0';
    return c;
  }();
  return cfg;
}

namespace {

double now()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/// Device and pinned host allocations made by the library in this process
/// (reported in pcg_stats; the no-allocation entry point leaves it unchanged).
std::atomic<uint64_t> g_allocations{0};

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  void* ensure(size_t bytes)
  {
    bytes = std::max<size_t>(bytes, 256);
    if (bytes > cap) {
      if (p) (void) hipFree(p);
      p = nullptr;
      cap = 0;
      hipError_t e = hipMalloc(&p, bytes);
      g_allocations++;
      if (e != hipSuccess) {
        (void) hipGetLastError();
        throw pcg_error("device allocation of " + std::to_string(bytes) + " bytes failed",
                        kErrNoMem);
      }
      cap = bytes;
    }
    return p;
  }
  void release()
  {
    if (p) (void) hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

/// Page-locked host staging memory (asynchronous uploads).
struct HostBuf {
  void* p = nullptr;
  size_t cap = 0;
  void* ensure(size_t bytes)
  {
    bytes = std::max<size_t>(bytes, 256);
    if (bytes > cap) {
      if (p) (void) hipHostFree(p);
      p = nullptr;
      cap = 0;
      hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
      g_allocations++;
      if (e != hipSuccess) {
        (void) hipGetLastError();
        throw pcg_error("pinned host allocation of " + std::to_string(bytes) + " bytes failed",
                        kErrNoMem);
      }
      cap = bytes;
    }
    return p;
  }
  void release()
  {
    if (p) (void) hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

} // namespace

/// Work buffers of one set of engine calls (engine_work_select): the
/// per-query constants, work items, slot words, escape tables, columns and
/// partial sums of the sieve passes, the scratch of the other pi(x) terms and
/// their events.  Sets share the device's stream and resident tables.
struct Work {
  std::vector<hipEvent_t> ev;         // timing events, 3 per pass
  hipEvent_t staged[2] = {};          // staging slot s may be overwritten
  hipEvent_t term_ev[6] = {};         // pi(x) term boundaries (pix.cpp)
  PhiUpload phi;                      // what extra[3] holds (pix.cpp)
  DevBuf queries, items, binfo;
  DevBuf tc, sacc, counters, queue, cols, sc, sm, smp, out, carry, reach, sfx, spans, groups;
  DevBuf esc_keys, esc_vals, esc_ctl;
  DevBuf extra[8];                    // scratch of the other pi(x) terms
  HostBuf stage[2];                   // per pass: spans, columns

  hipEvent_t event(size_t i)
  {
    while (ev.size() <= i) {
      hipEvent_t e;
      PCG_HIP(hipEventCreate(&e));
      ev.push_back(e);
    }
    return ev[i];
  }

  void init()
  {
    if (!staged[0])
      for (auto& e : staged) PCG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (!term_ev[0])
      for (auto& e : term_ev) PCG_HIP(hipEventCreate(&e));
  }

  void release()
  {
    for (DevBuf* b : {&queries, &items, &binfo, &tc, &sacc, &counters, &queue, &cols, &sc, &sm,
                      &smp, &out, &carry, &reach, &sfx, &spans, &groups, &esc_keys, &esc_vals,
                      &esc_ctl})
      b->release();
    for (HostBuf* b : {&stage[0], &stage[1]})
      b->release();
    for (DevBuf& b : extra)
      b.release();
    phi = PhiUpload();
  }

  ~Work()
  {
    release();
    for (hipEvent_t e : ev) (void) hipEventDestroy(e);
    for (hipEvent_t e : staged) if (e) (void) hipEventDestroy(e);
    for (hipEvent_t e : term_ev) if (e) (void) hipEventDestroy(e);
  }
};

namespace {

/// Device-wide state: the stream, the resident tables, the scratch array.
struct Device {
  bool ready = false;
  int dev = 0;
  int cus = 0;
  std::string ident;                  // architecture, CU count, board name (schedule cache key)
  hipStream_t stream = nullptr;
  DevBuf primes, factor, pi_bits, pi_cnt;
  DevBuf pix_tab, pix_seg;            // prime-count table over [0, pix_n]
  DevBuf pix_win;                     // one window of a prime-count table too large to keep whole
  char* scratch = nullptr;            // the device's scratch array (no-allocation entry point)
  uint64_t pix_n = 0;

  void init()
  {
    if (ready) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
      (void) hipGetLastError();
      throw pcg_error("no HIP device visible: the MI355X engine needs a gfx950 GPU", kErrDevice);
    }
    PCG_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    PCG_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      throw pcg_error(std::string("unsupported device ") + prop.gcnArchName + " (need gfx950)",
                      kErrDevice);
    cus = prop.multiProcessorCount;
    ident = std::string(prop.gcnArchName) + ":" + std::to_string(cus) + "cu:" + prop.name;
    for (char& c : ident)
      if (c == ' ' || c == '\t' || c == '\n') c = '_';
    PCG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    PCG_HIP(hsl_upload_constants());
    void* sp = nullptr;
    PCG_HIP(hsl_scratch_address(dev, &sp));
    scratch = (char*) sp;
    ready = true;
  }

  void release()
  {
    for (DevBuf* b : {&primes, &factor, &pi_bits, &pi_cnt, &pix_tab, &pix_seg, &pix_win})
      b->release();
    pix_n = 0;
  }
};

std::recursive_mutex g_mutex;
Device g_dev;

// Work sets: the default one, the sets engine_work_create made (all released
// by release_device), and the set the calling thread uses.
Work g_work0;
std::vector<Work*> g_works = {&g_work0};
thread_local Work* t_work = nullptr;

Work& cur_work()
{
  Work& w = t_work ? *t_work : g_work0;
  w.init();
  return w;
}

// Pass hook of the calling thread (PassHookScope) and the table freeze that
// holds while it runs.
thread_local const std::function<void()>* t_pass_hook = nullptr;
thread_local int t_tables_frozen = 0;

void throw_if_frozen()
{
  if (t_tables_frozen) throw tables_frozen_error();
}

/// Host tables shared by the queries of one launch.
struct TableSet {
  LeafSum kind = LeafSum::D;
  bool factor32 = false;
  int64_t ft_y = 0, ft_limit = 0;   // factor table parameters
  uint64_t factor_n = 0;            // factor table entries (built on the device)
  PrimeTables P;
};

/// Everything the device needs for one query.
struct QueryPlan {
  bool empty = true;
  QueryDesc qd{};
  std::vector<BInfo> binfo;
  std::vector<uint32_t> reach;     // per b: segments that reach b (non-increasing)
  std::vector<uint64_t> sfx;       // suffix sums of reach
  std::vector<uint32_t> esc_n;     // per b: the first esc_n segments may hold |M| too large
                                   // for the slot word (SlotFormat); no later one can
  uint32_t nseg = 0;               // segments that reach at least b_first

  /// b values segment s reaches: #{b : reach_b > s}
  uint32_t nbs(uint64_t s) const
  {
    return (uint32_t) (std::upper_bound(reach.begin(), reach.end(), s,
                                        [](uint64_t v, uint32_t r) { return v >= r; }) -
                       reach.begin());
  }
  /// slots of segments [0, s)
  uint64_t F(uint64_t s) const
  {
    const uint32_t k = nbs(s);
    return s * k + sfx[k];
  }
};

uint64_t imin(uint64_t a, uint64_t b) { return a < b ? a : b; }

/// Slot budget of one pass (a slot word takes 4 bytes).
constexpr uint64_t kMaxSlotsPerPass = (uint64_t) 1 << 31;

/// Prime-table range a query needs.
uint64_t prime_range(const HslQuery& q)
{
  uint64_t y = (uint64_t) std::max<int64_t>(q.y, 1), z = (uint64_t) std::max<int64_t>(q.z, 1);
  if (q.kind == LeafSum::D)
    return std::max<uint64_t>(y, isqrt(z)) + 1;
  uint64_t sqy = std::max<uint64_t>(1, isqrt(y));
  uint64_t max_prime = imin(y, z / sqy);
  return std::max({max_prime, imin(isqrt(z), y), sqy}) + 1;
}

void validate(const HslQuery& q)
{
  if (q.x < 1) return;
  if (q.y < 1 || q.z < 1)
    throw pcg_error("y and z must be >= 1");
  if (q.k < 0 || q.k > 1000)
    throw pcg_error("k/c out of range");
  if (q.kind == LeafSum::D) {
    if (q.z < q.y)
      throw pcg_error("D(x, y, z, k) needs z >= y");
    if (q.x / q.z > ((i128) 1 << 50))
      throw pcg_error("D(x, y, z, k): x / z must be < 2^50");
    if (q.z >= ((int64_t) 1 << 32))
      throw pcg_error("D(x, y, z, k): z must be < 2^32");
  } else {
    if ((i128) q.z > ((i128) 1 << 50))
      throw pcg_error("S2_hard(x, y, z, c): z must be < 2^50");
  }
  if (q.x > ((i128) std::max<int64_t>(q.y, 1) << 62))
    throw pcg_error("x is too large (x > y * 2^62)");
  if (q.y >= (int64_t) 0xFFFFFFF0LL)
    throw pcg_error("y must be < 2^32");
}

/// Per-b constants of one query (see BInfo).
QueryPlan plan_query(const HslQuery& q, const PrimeTables& P)
{
  QueryPlan qp;
  if (q.x < 1) return qp;
  const i128 x = q.x;
  const uint64_t y = (uint64_t) q.y, z = (uint64_t) q.z;
  uint64_t limit, b_first, b_last, split;
  if (q.kind == LeafSum::D) {
    limit = (uint64_t) (x / (i128) z);
    uint64_t x_star = (uint64_t) get_x_star_gourdon(x, (int64_t) y);
    uint64_t mb = imin(imin(x_star, isqrt((u128) x)), isqrt(limit));
    b_first = (uint64_t) q.k + 1;
    b_last = P.pi(imin(mb, P.max));
    split = P.pi(isqrt(z));                     // pi(sqrt(z))
  } else {
    limit = z;
    uint64_t sqy = isqrt(y);
    b_first = (uint64_t) q.k + 1;
    split = P.pi(sqy);                          // pi(sqrt(y))
    if (z <= y)
      b_last = split;
    else
      b_last = P.pi(imin(imin(imin(isqrt((u128) x), isqrt(z)), y), P.max));
  }
  if (limit == 0 || b_first > b_last) return qp;
  if (b_last >= P.primes.size())
    throw pcg_error("internal: prime table too small", kErrInternal);

  qp.empty = false;
  qp.qd.limit = limit;
  qp.qd.b_first = (uint32_t) b_first;
  qp.qd.b_last = (uint32_t) b_last;
  qp.qd.pre_last = (uint32_t) (b_first - 1);
  qp.binfo.resize(b_last - b_first + 1);
  for (uint64_t b = b_first; b <= b_last; b++) {
    BInfo& bi = qp.binfo[b - b_first];
    std::memset(&bi, 0, sizeof(bi));
    uint64_t p = P.primes[b];
    u128 xp = (u128) x / p;
    bi.xp_lo = (uint64_t) xp;
    bi.xp_d = (double) xp;
    bi.p = (uint32_t) p;
    u128 xpp = xp / ((u128) p * p);
    if (q.kind == LeafSum::D) {
      if (b <= split) {
        bi.type = 1;
        bi.T = (uint32_t) (2 * b + 1);
        bi.A = (uint64_t) std::min<u128>(xpp, z);
        bi.B = z / p;
        bi.C = z;
      } else {
        bi.type = 2;
        bi.A = (uint64_t) std::min<u128>(xpp, y);
        bi.B = p;
        bi.C = y;
      }
    } else {
      if (b <= split) {
        bi.type = 1;
        bi.T = (uint32_t) p;
        bi.A = y;
        bi.B = y / p;
        bi.C = y;
      } else {
        bi.type = 2;
        bi.A = imin(y, z / p);
        bi.B = p;
        bi.C = y;
      }
    }
  }
  return qp;
}

/// Slots per segment.  A segment starting at `low` can only process b with
/// p_b <= sqrt(x / low) (beyond that its leaf range is empty and the kernel
/// stops): segment s reaches b iff max(1, s * 240 W) <= x / p_b^2.  reach_b
/// counts those segments; it does not increase with b, and segments beyond
/// reach_{b_first} hold no leaves and are dropped.
void plan_reach(QueryPlan& qp, const HslQuery& q, const PrimeTables& P, uint32_t W, uint32_t cbits)
{
  const uint64_t SEG = 240ull * W;
  const uint64_t nseg_all = (qp.qd.limit + SEG - 1) / SEG;
  const uint32_t b_first = qp.qd.b_first, b_last = qp.qd.b_last;
  const uint32_t nb = b_last - b_first + 1;
  qp.reach.assign(nb, 0);
  for (uint32_t b = b_first; b <= b_last; b++) {
    const uint64_t p = P.primes[b];
    const u128 L = (u128) q.x / ((u128) p * p);
    if (L >= 1)
      qp.reach[b - b_first] = (uint32_t) std::min<u128>(L / SEG + 1, nseg_all);
  }
  qp.sfx.assign(nb + 1, 0);
  for (uint32_t k = nb; k-- > 0;)
    qp.sfx[k] = qp.sfx[k + 1] + qp.reach[k];
  qp.nseg = qp.reach[0];
  // Slot (s, b) holds at most floor(xp / low1) - floor(xp / high) leaves
  // (the integers m of its leaf range), so |M| <= that.  For s >= 1 this is
  // below xp / (SEG s (s + 1)) + 1, which reaches thr only while
  // s (s + 1) <= K = xp / (SEG (thr - 1)); segment 0 always counts.
  const double thr = (double) (1u << (31 - cbits));
  qp.esc_n.assign(nb, 0);
  for (uint32_t b = b_first; b <= b_last; b++) {
    const long double xp = (long double) ((u128) q.x / P.primes[b]);
    const long double K = xp / ((long double) SEG * (thr - 1));
    const long double sm = (sqrtl(1 + 4 * K) - 1) / 2;     // largest s with s (s + 1) <= K
    const uint64_t n = sm < 4e9L ? (uint64_t) sm + 2 : UINT32_MAX;   // + 1 segment of margin
    qp.esc_n[b - b_first] = (uint32_t) std::min<uint64_t>(n, qp.reach[b - b_first]);
  }
}

/// Upper bound of the slots of a span [s0, s1) of a query whose M may not
/// fit the slot word.
uint64_t escape_bound(const QueryPlan& qp, uint32_t s0, uint32_t s1)
{
  uint64_t n = 0;
  for (size_t br = 0; br < qp.esc_n.size(); br++) {
    const uint32_t hi = std::min(s1, qp.esc_n[br]);
    if (hi <= s0) break;                       // esc_n does not increase with b
    n += hi - s0;
  }
  return n;
}

/// Segment size (the schedule's default; the autotuner, tuner.hpp, times
/// candidates around it).  One wave owns a segment; smaller segments mean
/// more waves in flight and a finer-grained work queue, larger ones fewer
/// (segment, b) slots per number.  Measured on the MI355X for D with the
/// default parameters (one call, kernel + combine ms; the sieve limit is
/// x / z; 10 W + 1,008 bytes of LDS per wave: 28 waves per CU up to W = 384,
/// 26 at 512, 21 at 640, and at most 20 for the 81 VGPRs of the kernel):
///
///   x      limit    W=128   256    384    512    640    768
///   1e17   8.4e9    14.5   13.5   18.5   .      .      .
///   1e18   3.1e10   .      41.7   37.6   47.7   .      .
///   1e19   1.1e11   .      .     145.0  129.5  119.2  136.3
///   1e21   1.5e12   .      .      .     2275   2067   2205
///
/// (an older build of the kernel: 1e20, limit 4.8e11: 572 / 563 / 579 ms at
/// W = 512 / 640 / 768).  S2_hard (sieve limit z) turns to 640 later:
///
///   x      limit    W=128   256    384    512    640    768
///   1e17   9.4e9    18.2   21.1   27.3   .      .      .
///   1e19   1.3e11   .      .     184.3  162.6  180.1  210.1
///   1e21   1.8e12   .      .      .     3082   2766   2865
///
/// Small segments win while a few heavy first segments would otherwise
/// finish last; in a batch the heavy segments of many queries overlap and
/// larger segments win, at the same total: the 256 reference x values
/// (1e12 .. 1e18) in one launch took 773, 711, 794 ms at W = 512, 640, 768.
/// So the rule takes the sum of the sieve limits of the launch.
uint32_t auto_segment_words(uint64_t limit, LeafSum kind, bool single)
{
  uint32_t cfg = engine_config().segment_words;
  if (cfg) return cfg;
  const bool d = kind == LeafSum::D;
  if (single && engine_config().leaf_batching) {
    // one query per launch, batched leaf counting (measured on the MI355X at
    // 1e15 .. 1e21: D 128 words up to 1e17, 256 at 1e18, 512 at 1e19, 640
    // from 1e20; S2_hard 128 up to 1e17, 256, 384, 512, 640 at 1e18 .. 1e21)
    if (limit < (d ? 16000000000ull : 18000000000ull)) return 128;
    if (limit < (d ? 45000000000ull : 66000000000ull)) return 256;
    if (limit < (d ? 60000000000ull : 250000000000ull)) return 384;
    return limit < (d ? 220000000000ull : 900000000000ull) ? 512 : 640;
  }
  if (limit < (d ? 5000000000ull : 15000000000ull)) return 128;
  if (limit < (d ? 15000000000ull : 35000000000ull)) return 256;
  if (limit < 80000000000ull) return 384;
  return limit < (d ? 100000000000ull : 300000000000ull) ? 512 : 640;
}

/// Table requirements of a group of queries.
struct TableNeed {
  LeafSum kind = LeafSum::D;
  uint64_t pmax = 7;
  int64_t fy = 1, flim = 1;
  bool shared = true;     // D: every query has y >= z (no entry <= z depends on y)
};

TableNeed table_need(const std::vector<HslQuery>& qs)
{
  TableNeed n;
  n.kind = qs[0].kind;
  for (const auto& q : qs) {
    if (q.x < 1) continue;
    n.pmax = std::max(n.pmax, prime_range(q));
    if (q.kind == LeafSum::D) {
      n.flim = std::max(n.flim, q.z);
      n.fy = std::max(n.fy, q.y);
      n.shared = n.shared && q.y >= q.z;
    } else {
      n.flim = std::max(n.flim, q.y);
      n.fy = std::max(n.fy, q.y);
    }
  }
  if (n.kind == LeafSum::D && n.shared)
    n.fy = std::max(n.fy, n.flim);
  return n;
}

/// Can the resident tables answer these queries?  Prime tables are prefix
/// valid; an lpf FactorTable is prefix valid; a FactorTableD built with
/// fy >= flim is prefix valid for every query with y >= z.
bool covers(const TableSet& T, const TableNeed& n)
{
  if (T.kind != n.kind || T.P.max < n.pmax || T.ft_limit < n.flim)
    return false;
  if (n.kind == LeafSum::S2_hard)
    return true;
  if (n.shared)
    return T.ft_y >= T.ft_limit;
  return T.ft_y == n.fy && T.ft_limit == n.flim;
}

TableSet build_tables(const TableNeed& n)
{
  TableSet T;
  T.kind = n.kind;
  int threads = engine_config().host_threads > 0 ? engine_config().host_threads : default_threads();
  T.P = build_prime_tables(n.pmax, threads);
  T.ft_y = n.fy;
  T.ft_limit = std::max<int64_t>(1, n.flim);
  T.factor32 = T.kind == LeafSum::D ? n.flim > kFactorTableD16Max : n.flim > kFactorTable16Max;
  T.factor_n = (uint64_t) ft_to_index((uint64_t) T.ft_limit) + 1;
  if ((uint64_t) T.ft_limit >= (1ull << 32))
    throw pcg_error("factor table limit must be below 2^32");
  if ((uint64_t) isqrt((u128) T.ft_limit) > T.P.max)
    throw pcg_error("internal: factor table needs primes up to its square root", kErrInternal);
  return T;
}

void upload(DevBuf& b, const void* src, size_t bytes)
{
  b.ensure(bytes);
  if (bytes) PCG_HIP(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
}

/// Tables resident in HBM from an earlier call, rebuilt only when they do
/// not cover the new queries.  Caller holds g_mutex.
TableSet g_tables;
bool g_tables_valid = false;

const TableSet& resident_tables(const std::vector<HslQuery>& qs, uint64_t pmin = 0)
{
  TableNeed need = table_need(qs);
  need.pmax = std::max(need.pmax, pmin);
  if (g_tables_valid && covers(g_tables, need))
    return g_tables;
  throw_if_frozen();
  g_tables_valid = false;
  g_tables = build_tables(need);
  Device& D = g_dev;
  const TableSet& T = g_tables;
  upload(D.primes, T.P.primes.data(), T.P.primes.size() * 4);
  upload(D.pi_bits, T.P.pi_bits.data(), T.P.pi_bits.size() * 8);
  upload(D.pi_cnt, T.P.pi_cnt.data(), T.P.pi_cnt.size() * 4);
  // the factor table is built where it is used: on the device
  D.factor.ensure(T.factor_n * (T.factor32 ? 4 : 2));
  PCG_HIP(hsl_launch_factor_table(D.factor.p, T.factor32, T.kind == LeafSum::D,
                                  (uint64_t) T.ft_limit, (uint64_t) T.ft_y,
                                  (const uint32_t*) D.primes.p, (uint32_t) T.P.primes.size(),
                                  D.stream));
  PCG_HIP(hipStreamSynchronize(D.stream));
  g_tables_valid = true;
  return g_tables;
}

/// Run a set of queries that share one table set.  Caller holds g_mutex.
///
/// Work items are (query, segment) pairs; their (segment, b) slots are
/// written by the sieve kernel and folded by the combine kernels.  Items are
/// processed in passes of at most kMaxSlotsPerPass slots so that the slot
/// arrays stay bounded; the per-column carry survives between passes.
std::vector<i128> tune_group(const std::vector<HslQuery>& qs, pcg_stats* st, uint64_t slots_override,
                             const TuneKey& key, const Schedule& def, bool shareable);

/// Sieve-limit sums below this are not tuned: their launches take well under
/// a millisecond and the rule's schedule is as good as any.
constexpr uint64_t kTuneMinLimit = 1ull << 30;

std::vector<i128> run_group(const std::vector<HslQuery>& qs, pcg_stats* st, uint64_t slots_override = 0,
                            const Schedule* forced = nullptr)
{
  Device& D = g_dev;
  D.init();
  Work& Wk = cur_work();
  double t0 = now();
  for (const auto& q : qs) validate(q);

  std::vector<i128> result(qs.size(), 0);
  bool any = false;
  for (const auto& q : qs) any |= q.x >= 1;
  if (!any) return result;

  const TableSet& T = resident_tables(qs);
  double t_tab = now();

  std::vector<QueryPlan> plans(qs.size());
  uint64_t sum_limit = 1;
  uint32_t n_work = 0;
  for (size_t i = 0; i < qs.size(); i++) {
    plans[i] = plan_query(qs[i], T.P);
    if (!plans[i].empty) {
      sum_limit += plans[i].qd.limit;
      n_work++;
    }
  }
  // The schedule: forced (a tuning trial), set by the caller (segment words),
  // a cached winner of this launch class, or the rule; with tuning on, a
  // class not yet cached is tuned on this launch.
  const EngineConfig& cfg = engine_config();
  const bool overridden =
      cfg.segment_words || cfg.group_size || cfg.grid_pct || cfg.kernel_waves >= 0;
  Schedule sch;
  sch.W = auto_segment_words(sum_limit, T.kind, n_work == 1);
  // A single query with batched leaf counting runs the build for 6 waves
  // per SIMD when LDS allows more than 20 waves per CU (up to 512 words):
  // measured 1.5-7.5% faster from 1e17 to 1e20, slower at 640 words.
  if (n_work == 1 && cfg.leaf_batching && sch.W <= 512 && sum_limit >= 5000000000ull)
    sch.waves = 6;
  int sched_source = 0;
  if (forced) {
    sch = *forced;
  } else if (!overridden && cfg.tune_mode != kTuneOff && sum_limit >= kTuneMinLimit) {
    TuneKey key;
    key.device = D.ident;
    key.kind = T.kind == LeafSum::D ? 0 : 1;
    key.ft32 = T.factor32 ? 1 : 0;
    key.lclass = (int) std::floor(2.0 * std::log2((double) sum_limit));
    key.qclass = std::min(8, 31 - __builtin_clz(std::max<uint32_t>(n_work, 1)));
    TuneEntry e;
    if (tune_lookup(key, &e)) {
      sch = e.s;
      sched_source = 1;
    } else if (engine_config().tune_mode == kTuneOn) {
      return tune_group(qs, st, slots_override, key, sch,
                        n_work > 1 && engine_config().share_segments);
    }
  } else if (overridden) {
    if (cfg.group_size) sch.gmax = cfg.group_size;
    if (cfg.grid_pct) sch.grid_pct = cfg.grid_pct;
    if (cfg.kernel_waves >= 0) sch.waves = (uint32_t) cfg.kernel_waves;
  }
  int bpc = 0;
  uint32_t W = sch.W;
  const uint32_t waves = sch.waves == 6 ? 6u : 0u;
  if (W < 128 || W > 4096 || (W % 128) != 0)
    throw pcg_error("segment words must be a multiple of 128 in [128, 4096]");
  const uint32_t cbits = std::max(slot_format(W).cbits, engine_config().slot_cbits);
  if (cbits > 28) throw pcg_error("slot bits must be at most 28");
  PCG_HIP(hsl_sieve_occupancy(T.factor32, W, waves, &bpc));
  bpc = std::max(bpc, 1);
  const uint64_t slots_wg = std::max<uint64_t>(1, (uint64_t) D.cus * bpc * sch.grid_pct / 100);

  // Queries with work, largest first: items are taken from the work queue in
  // (query, segment) order, so the heaviest items (the first segments of the
  // largest queries) start first and the light tails end the launch.
  std::vector<size_t> live;
  for (size_t i = 0; i < qs.size(); i++) {
    if (plans[i].empty) continue;
    plan_reach(plans[i], qs[i], T.P, W, cbits);
    if (plans[i].nseg > 0) live.push_back(i);
  }
  if (live.empty()) return result;
  std::stable_sort(live.begin(), live.end(), [&](size_t a, size_t b) {
    return plans[a].sfx[0] > plans[b].sfx[0];
  });

  std::vector<QueryDesc> qdesc;
  std::vector<BInfo> binfo;
  std::vector<uint32_t> reach;
  std::vector<uint64_t> sfx;
  std::vector<uint32_t> qslot(qs.size(), UINT32_MAX);
  std::vector<size_t> slot_query;                // query slot -> index into qs
  std::vector<uint32_t> colbase;                 // per slot: first global column
  uint32_t ncols_total = 0;
  uint64_t nsegments = 0;
  for (size_t i : live) {
    QueryDesc qd = plans[i].qd;
    qd.binfo_off = (uint32_t) binfo.size();
    qd.reach_off = (uint32_t) reach.size();
    qd.sfx_off = (uint32_t) sfx.size();
    qslot[i] = (uint32_t) qdesc.size();
    slot_query.push_back(i);
    qdesc.push_back(qd);
    binfo.insert(binfo.end(), plans[i].binfo.begin(), plans[i].binfo.end());
    reach.insert(reach.end(), plans[i].reach.begin(), plans[i].reach.end());
    sfx.insert(sfx.end(), plans[i].sfx.begin(), plans[i].sfx.end());
    colbase.push_back(ncols_total);
    ncols_total += plans[i].nbs(0);
    nsegments += plans[i].nseg;
  }

  upload(Wk.queries, qdesc.data(), qdesc.size() * sizeof(QueryDesc));
  upload(Wk.binfo, binfo.data(), binfo.size() * sizeof(BInfo));
  upload(Wk.reach, reach.data(), reach.size() * 4);
  upload(Wk.sfx, sfx.data(), sfx.size() * 8);
  const size_t nq = qdesc.size();
  Wk.carry.ensure((size_t) ncols_total * 8);
  Wk.out.ensure((size_t) ncols_total * 16);
  Wk.sacc.ensure(nq * 16);
  Wk.counters.ensure(kHslCounters * 8);
  hipStream_t s = D.stream;
  PCG_HIP(hsl_launch_fill_zero(Wk.carry.p, (size_t) ncols_total * 8, s));
  PCG_HIP(hsl_launch_fill_zero(Wk.out.p, (size_t) ncols_total * 16, s));
  PCG_HIP(hsl_launch_fill_zero(Wk.sacc.p, nq * 16, s));
  PCG_HIP(hsl_launch_fill_zero(Wk.counters.p, kHslCounters * 8, s));

  uint64_t max_slots = slots_override ? slots_override : engine_config().max_slots_per_pass;
  if (max_slots == 0) max_slots = kMaxSlotsPerPass;

  // Passes: runs of (query, segment) items with at most max_slots slots (a
  // single segment may exceed it).  A pass is a list of spans, one per query
  // it touches; boundaries are found by binary search on the slot count F.
  // With shared segments a pass is a segment range [S0, S1) of every query,
  // so that the queries of a segment meet in one pass.
  struct Pass { std::vector<PassSpan> spans; uint32_t nitems = 0; uint64_t slots = 0; };
  std::vector<Pass> passes(1);
  const bool share = nq > 1 && engine_config().share_segments;
  if (share) {
    uint64_t smax = 0;
    for (uint32_t qsl = 0; qsl < nq; qsl++) smax = std::max<uint64_t>(smax, plans[slot_query[qsl]].nseg);
    auto slots_upto = [&](uint64_t S) {          // slots of segments [0, S) of all queries
      uint64_t t = 0;
      for (uint32_t qsl = 0; qsl < nq; qsl++) {
        const QueryPlan& qp = plans[slot_query[qsl]];
        t += qp.F(std::min<uint64_t>(S, qp.nseg));
      }
      return t;
    };
    uint64_t S0 = 0;
    passes.clear();
    while (S0 < smax) {
      const uint64_t f0 = slots_upto(S0);
      uint64_t lo = S0, hi = smax;               // largest S1 with slots(S1) - slots(S0) <= cap
      while (lo < hi) {
        const uint64_t mid = (lo + hi + 1) / 2;
        if (slots_upto(mid) - f0 <= max_slots) lo = mid;
        else hi = mid - 1;
      }
      const uint64_t S1 = std::max(lo, S0 + 1);
      Pass ps;
      for (uint32_t qsl = 0; qsl < nq; qsl++) {
        const QueryPlan& qp = plans[slot_query[qsl]];
        const uint64_t s0 = std::min<uint64_t>(S0, qp.nseg), s1 = std::min<uint64_t>(S1, qp.nseg);
        if (s1 <= s0) continue;
        PassSpan sp{};
        sp.query = qsl;
        sp.s0 = (uint32_t) s0;
        sp.s1 = (uint32_t) s1;
        sp.item_base = ps.nitems;
        sp.t_base = ps.slots - qp.F(s0);       // modulo 2^64
        ps.spans.push_back(sp);
        ps.nitems += (uint32_t) (s1 - s0);
        ps.slots += qp.F(s1) - qp.F(s0);
      }
      passes.push_back(ps);
      S0 = S1;
    }
  } else {
    for (uint32_t qsl = 0; qsl < nq; qsl++) {
      const QueryPlan& qp = plans[slot_query[qsl]];
      uint64_t s0 = 0;
      while (s0 < qp.nseg) {
        Pass& ps = passes.back();
        const uint64_t f0 = qp.F(s0);
        const uint64_t cap = max_slots > ps.slots ? max_slots - ps.slots : 0;
        // largest s1 in (s0, nseg] with F(s1) - F(s0) <= cap
        uint64_t lo = s0, hi = qp.nseg;
        while (lo < hi) {
          const uint64_t mid = (lo + hi + 1) / 2;
          if (qp.F(mid) - f0 <= cap) lo = mid;
          else hi = mid - 1;
        }
        uint64_t s1 = lo;
        if (s1 == s0) {
          if (ps.nitems > 0) {           // full: start a new pass
            passes.emplace_back();
            continue;
          }
          s1 = s0 + 1;                   // one segment larger than a pass
        }
        PassSpan sp{};
        sp.query = qsl;
        sp.s0 = (uint32_t) s0;
        sp.s1 = (uint32_t) s1;
        sp.item_base = ps.nitems;
        sp.t_base = ps.slots - f0;       // modulo 2^64
        ps.spans.push_back(sp);
        ps.nitems += (uint32_t) (s1 - s0);
        ps.slots += qp.F(s1) - f0;
        s0 = s1;
        if (s0 < qp.nseg) passes.emplace_back();
      }
    }
  }
  if (passes.back().nitems == 0) passes.pop_back();
  const size_t npass = passes.size();

  uint64_t max_pass_slots = 0, max_pass_items = 0, max_pass_spans = 0;
  for (const Pass& ps : passes) {
    max_pass_slots = std::max(max_pass_slots, ps.slots);
    max_pass_items = std::max<uint64_t>(max_pass_items, ps.nitems);
    max_pass_spans = std::max<uint64_t>(max_pass_spans, ps.spans.size());
  }
  const size_t spans_bytes = (size_t) max_pass_spans * sizeof(PassSpan);
  const size_t cols_bytes = (size_t) ncols_total * sizeof(CombineCol);
  Wk.spans.ensure(spans_bytes);
  Wk.items.ensure((size_t) max_pass_items * sizeof(SegItem));
  Wk.cols.ensure(cols_bytes);
  Wk.tc.ensure(max_pass_slots * 4);
  Wk.queue.ensure(npass * 4);
  PCG_HIP(hsl_launch_fill_zero(Wk.queue.p, npass * 4, s));
  // escape tables: per pass at least twice the bound of its escaping slots
  std::vector<uint32_t> esc_cap(npass);
  uint32_t max_cap = 0;
  uint64_t esc_bound_total = 0;
  for (size_t pi = 0; pi < npass; pi++) {
    uint64_t bound = 0;
    for (const PassSpan& sp : passes[pi].spans)
      bound += escape_bound(plans[slot_query[sp.query]], sp.s0, sp.s1);
    esc_bound_total += bound;
    uint64_t cap = 1024;
    while (cap < 2 * bound + 2) cap *= 2;
    if (cap > (1ull << 31)) throw pcg_error("internal: escape table too large", kErrInternal);
    esc_cap[pi] = (uint32_t) cap;
    max_cap = std::max(max_cap, (uint32_t) cap);
  }
  Wk.esc_keys.ensure((size_t) max_cap * 8);
  Wk.esc_vals.ensure((size_t) max_cap * 4);
  Wk.esc_ctl.ensure(npass * 8);
  PCG_HIP(hsl_launch_fill_zero(Wk.esc_ctl.p, npass * 8, s));
  for (auto& hb : Wk.stage) hb.ensure(spans_bytes + cols_bytes);

  // Shared segments: segment s of every query that covers it in full is
  // sieved once for groups of up to gmax queries with the same first b; a
  // query's last (possibly partial) segment goes alone.  Groups are ordered
  // by segment, the largest queries first within one.  Groups of 4 pay off
  // for large batches (256 reference x values: 947 ms against 1073 ms with
  // groups of 2); below that their heavier items and registers cost more
  // (64 values: 294 ms against 253 ms).
  std::vector<std::vector<SegGroup>> pass_groups(npass);
  size_t max_groups = 0;
  const uint32_t gmax = sch.gmax ? sch.gmax : (nq >= 128 ? 4u : 2u);
  if (share) {
    std::vector<uint32_t> order(nq);
    for (uint32_t i = 0; i < nq; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      return plans[slot_query[a]].nseg > plans[slot_query[b]].nseg;
    });
    std::vector<uint32_t> item_base(nq);
    for (size_t pi = 0; pi < npass; pi++) {
      std::vector<SegGroup>& groups = pass_groups[pi];
      uint32_t S0 = UINT32_MAX, S1 = 0;
      for (const PassSpan& sp : passes[pi].spans) {
        item_base[sp.query] = sp.item_base - sp.s0;      // item of segment sg: item_base + sg
        S0 = std::min(S0, sp.s0);
        S1 = std::max(S1, sp.s1);
      }
      size_t active = nq;                        // order[0 .. active) have nseg > sg
      for (uint64_t sg = S0; sg < S1; sg++) {
        while (active > 0 && plans[slot_query[order[active - 1]]].nseg <= sg) active--;
        SegGroup g{};
        uint32_t g_first = 0;
        for (size_t r = 0; r < active; r++) {
          const uint32_t q = order[r];
          const uint32_t idx = item_base[q] + (uint32_t) sg;
          if (sg + 1 == plans[slot_query[q]].nseg) {   // last segment: alone
            SegGroup one{};
            one.n = 1;
            one.item[0] = idx;
            groups.push_back(one);
            continue;
          }
          if (g.n == gmax || (g.n > 0 && qdesc[q].b_first != g_first)) {
            groups.push_back(g);
            g = SegGroup{};
          }
          if (g.n == 0) g_first = qdesc[q].b_first;
          g.item[g.n++] = idx;
        }
        if (g.n > 0) groups.push_back(g);
      }
      max_groups = std::max(max_groups, groups.size());
    }
    Wk.groups.ensure(max_groups * sizeof(SegGroup));
  }
  double t_plan = now() - t_tab, ms_kernel = 0, ms_combine = 0;

  for (size_t pi = 0; pi < npass; pi++) {
    double tp = now();
    const Pass& ps = passes[pi];
    const int slot = (int) (pi & 1);
    if (pi >= 2) PCG_HIP(hipEventSynchronize(Wk.staged[slot]));
    char* stage = (char*) Wk.stage[slot].p;
    PassSpan* spans = (PassSpan*) stage;
    CombineCol* cols = (CombineCol*) (stage + spans_bytes);
    std::memcpy(spans, ps.spans.data(), ps.spans.size() * sizeof(PassSpan));
    // combine columns: per span, b with at least one segment in [s0, s1)
    uint32_t ncols = 0, max_n = 1;
    uint64_t col_items = 0;
    for (const PassSpan& sp : ps.spans) {
      const QueryPlan& qp = plans[slot_query[sp.query]];
      const uint32_t nbs0 = qp.nbs(sp.s0);
      for (uint32_t br = 0; br < nbs0; br++) {
        const uint32_t n = std::min(qp.reach[br], sp.s1) - sp.s0;   // reach[br] > s0
        cols[ncols++] = CombineCol{br, sp.item_base, n, colbase[sp.query] + br};
        max_n = std::max(max_n, n);
        col_items += n;
      }
    }
    // 128 slices per column, 512 for passes of more than 2^28 slots: the
    // combine reads the slot words with one dependent step per 8 items a
    // thread, so large passes need more slices to keep enough reads in
    // flight (D(1e19): 3.28 -> 2.35 ms), while small ones lose more to the
    // final reduction over the slices (D(1e17): 0.66 against 0.78 ms)
    const uint32_t slices = col_items > (1ull << 28) ? 512 : 128;
    uint32_t slice_len = std::max<uint32_t>(1, (max_n + slices - 1) / slices);
    uint32_t nslices = (max_n + slice_len - 1) / slice_len;

    PCG_HIP(hipMemcpyAsync(Wk.spans.p, spans, ps.spans.size() * sizeof(PassSpan), hipMemcpyHostToDevice, s));
    PCG_HIP(hipMemcpyAsync(Wk.cols.p, cols, (size_t) ncols * sizeof(CombineCol), hipMemcpyHostToDevice, s));
    PCG_HIP(hipEventRecord(Wk.staged[slot], s));
    if ((size_t) nslices * ncols * 16 > Wk.smp.cap)
      PCG_HIP(hipStreamSynchronize(s));          // earlier passes may still read them
    Wk.sc.ensure((size_t) nslices * ncols * 8);
    Wk.sm.ensure((size_t) nslices * ncols * 8);
    Wk.smp.ensure((size_t) nslices * ncols * 16);
    PCG_HIP(hsl_launch_build_items((const QueryDesc*) Wk.queries.p, (const PassSpan*) Wk.spans.p,
                                   (uint32_t) ps.spans.size(), (const uint32_t*) Wk.reach.p,
                                   (const uint64_t*) Wk.sfx.p, (SegItem*) Wk.items.p, ps.nitems, s));
    t_plan += now() - tp;

    HslArgs a{};
    a.queries = (const QueryDesc*) Wk.queries.p;
    a.items = (const SegItem*) Wk.items.p;
    a.binfo = (const BInfo*) Wk.binfo.p;
    a.primes = (const uint32_t*) D.primes.p;
    a.factor = D.factor.p;
    a.pi_bits = (const uint64_t*) D.pi_bits.p;
    a.pi_cnt = (const uint32_t*) D.pi_cnt.p;
    a.T_S = (uint32_t*) Wk.tc.p;
    a.cbits = cbits;
    a.esc.keys = (unsigned long long*) Wk.esc_keys.p;
    a.esc.vals = (int32_t*) Wk.esc_vals.p;
    a.esc.ctl = (unsigned int*) Wk.esc_ctl.p + 2 * pi;
    a.esc.cap = esc_cap[pi];
    // the previous pass's combine has read its table (stream order)
    PCG_HIP(hsl_launch_fill_zero(Wk.esc_keys.p, (size_t) esc_cap[pi] * 8, s));
    a.S_acc = (unsigned long long*) Wk.sacc.p;
    a.counters = (unsigned long long*) Wk.counters.p;
    a.queue = (unsigned int*) Wk.queue.p + pi;
    a.nitems = ps.nitems;
    a.groups = nullptr;
    if (share) {
      // the previous pass's kernel may still read the groups buffer
      const std::vector<SegGroup>& groups = pass_groups[pi];
      if (pi > 0) PCG_HIP(hipStreamSynchronize(s));
      PCG_HIP(hipMemcpyAsync(Wk.groups.p, groups.data(), groups.size() * sizeof(SegGroup),
                             hipMemcpyHostToDevice, s));
      a.groups = (const SegGroup*) Wk.groups.p;
      a.nitems = (uint32_t) groups.size();
      a.group_max = gmax;
    }
    a.W = W;
    a.waves = share ? 0u : waves;
    a.bcap = engine_config().leaf_batching ? kBatchRecords : 0u;
    a.timers = engine_config().phase_timers ? 1u : 0u;
    uint32_t grid = (uint32_t) std::min<uint64_t>(slots_wg, a.nitems);

    PCG_HIP(hipEventRecord(Wk.event(3 * pi), s));
    PCG_HIP(hsl_launch_sieve(a, T.factor32, grid, s));
    PCG_HIP(hipEventRecord(Wk.event(3 * pi + 1), s));
    PCG_HIP(hsl_launch_combine((const CombineCol*) Wk.cols.p, ncols, (const SegItem*) Wk.items.p,
                               (const uint32_t*) Wk.tc.p, a.esc, cbits, slice_len, nslices,
                               (long long*) Wk.sc.p, (long long*) Wk.sm.p, Wk.smp.p,
                               (long long*) Wk.carry.p, Wk.out.p, s));
    PCG_HIP(hipEventRecord(Wk.event(3 * pi + 2), s));
    // between passes: the calling thread's pass hook (other work on the
    // same stream, after this pass) may run
    if (pi + 1 < npass) engine_pass_point();
  }

  std::vector<i128> col_out(ncols_total);
  std::vector<uint64_t> sacc(2 * nq);
  uint64_t counters[kHslCounters] = {};
  PCG_HIP(hipMemcpyAsync(col_out.data(), Wk.out.p, (size_t) ncols_total * 16, hipMemcpyDeviceToHost, s));
  PCG_HIP(hipMemcpyAsync(sacc.data(), Wk.sacc.p, nq * 16, hipMemcpyDeviceToHost, s));
  PCG_HIP(hipMemcpyAsync(counters, Wk.counters.p, sizeof(counters), hipMemcpyDeviceToHost, s));
  std::vector<uint32_t> esc_ctl(2 * npass);
  PCG_HIP(hipMemcpyAsync(esc_ctl.data(), Wk.esc_ctl.p, npass * 8, hipMemcpyDeviceToHost, s));
  PCG_HIP(hipStreamSynchronize(s));
  uint64_t n_escaped = 0;
  for (size_t pi = 0; pi < npass; pi++) {
    if (esc_ctl[2 * pi + 1] != 0)
      throw pcg_error("internal: escape table overflow (" + std::to_string(esc_ctl[2 * pi]) +
                          " escaping slots, capacity " + std::to_string(esc_cap[pi]) + ")",
                      kErrInternal);
    n_escaped += esc_ctl[2 * pi];
  }
  for (size_t pi = 0; pi < npass; pi++) {
    float mk = 0, mc = 0;
    PCG_HIP(hipEventElapsedTime(&mk, Wk.ev[3 * pi], Wk.ev[3 * pi + 1]));
    PCG_HIP(hipEventElapsedTime(&mc, Wk.ev[3 * pi + 1], Wk.ev[3 * pi + 2]));
    ms_kernel += mk;
    ms_combine += mc;
  }

  std::vector<i128> per_slot(nq, 0);
  for (size_t q = 0; q < nq; q++)
    per_slot[q] = (i128) (((u128) sacc[2 * q + 1] << 64) | sacc[2 * q]);
  for (size_t q = 0; q < nq; q++) {
    uint32_t ncol_q = (q + 1 < qdesc.size() ? colbase[q + 1] : ncols_total) - colbase[q];
    for (uint32_t c = 0; c < ncol_q; c++) per_slot[q] += col_out[colbase[q] + c];
  }
  for (size_t i = 0; i < qs.size(); i++)
    if (qslot[i] != UINT32_MAX) result[i] = per_slot[qslot[i]];

  if (st) {
    st->seconds_total = now() - t0;
    st->seconds_tables = t_tab - t0;
    st->seconds_plan = t_plan;
    st->seconds_kernel = ms_kernel / 1000.0;
    st->seconds_combine = ms_combine / 1000.0;
    st->segments = nsegments;
    st->chunks = nsegments;
    st->segment_words = W;
    st->sieve_limit = qdesc.size() == 1 ? qdesc[0].limit : 0;
    st->b_first = qdesc.size() == 1 ? qdesc[0].b_first : 0;
    st->b_last = qdesc.size() == 1 ? qdesc[0].b_last : 0;
    st->leaves = counters[0];
    st->hits = counters[1];
    st->factor_entries = counters[2];
    st->pairs = counters[3];
    st->windows = counters[4];
    st->leaf_primes = counters[5];
    st->leaf_steps = counters[6];
    st->dense_primes = counters[7];
    st->dense_steps = counters[8];
    st->group_steps = counters[9];
    st->prefix_builds = counters[10];
    st->queue_steps = counters[11];
    st->leaf_batches = counters[12];
    st->batch_records = counters[13];
    for (uint32_t i = 0; i < kHslPhases; i++)
      st->phase_cycles[i] = counters[kHslPhaseCounter0 + i];
    st->device_allocations = g_allocations.load();
    st->escaped_slots = n_escaped;
    st->escape_bound = esc_bound_total;
    st->schedule_source = (uint64_t) sched_source;
    st->group_size = share ? gmax : 1;
    st->grid_waves = slots_wg;
    st->kernel_waves = share ? 0 : waves;
  }
  return result;
}

/// Tunes one launch class on the launch itself: the default schedule first
/// (its sums are the answer), then segment sizes around the rule's, then the
/// group size of shared segments and a smaller grid, each change kept only
/// if it is faster.  Every trial must give the same sums as the first (a
/// schedule never changes results; a difference is an engine fault).  Short
/// launches are timed three times and take the fastest.  The winner is
/// cached for the device and class; `st` receives the winner's run.
std::vector<i128> tune_group(const std::vector<HslQuery>& qs, pcg_stats* st, uint64_t slots_override,
                             const TuneKey& key, const Schedule& def, bool shareable)
{
  std::vector<i128> ref;
  Schedule best = def;
  double best_ms = 0;
  bool have_best = false;
  pcg_stats best_st{};
  auto trial = [&](const Schedule& s) {
    double ms = 1e300;
    pcg_stats ls{};
    for (int rep = 0; rep < 3; rep++) {
      pcg_stats rs{};
      std::vector<i128> r = run_group(qs, &rs, slots_override, &s);
      if (ref.empty()) ref = r;
      else if (r != ref)
        throw pcg_error("internal: schedule W=" + std::to_string(s.W) + " g=" + std::to_string(s.gmax) +
                            " grid=" + std::to_string(s.grid_pct) + "% changed a result",
                        kErrInternal);
      const double t = (rs.seconds_kernel + rs.seconds_combine) * 1e3;
      if (t < ms) {
        ms = t;
        ls = rs;
      }
      if (t >= 20.0) break;                 // long launches: one run is enough
    }
    if (!have_best || ms < best_ms) {
      best = s;
      best_ms = ms;
      best_st = ls;
      have_best = true;
    }
    return ms;
  };
  const double def_ms = trial(def);
  static const uint32_t kWords[] = {128, 256, 384, 512, 640, 768, 1024};
  const int nw = (int) (sizeof kWords / sizeof kWords[0]);
  int i0 = 0;
  while (i0 + 1 < nw && kWords[i0] < def.W) i0++;
  for (int i = std::max(0, i0 - 2); i <= std::min(nw - 1, i0 + 2); i++) {
    if (kWords[i] == def.W) continue;
    Schedule s = def;
    s.W = kWords[i];
    trial(s);
  }
  if (shareable) {
    const uint32_t g0 = best_st.group_size;
    for (uint32_t g : {2u, 4u}) {
      if (g == g0) continue;
      Schedule s = best;
      s.gmax = g;
      trial(s);
    }
  }
  if (!shareable) {
    Schedule s = best;
    s.waves = best.waves == 6 ? 0u : 6u;
    trial(s);
  }
  {
    Schedule s = best;
    s.grid_pct = 75;
    trial(s);
  }
  TuneEntry e;
  e.s = best;
  e.ms_default = def_ms;
  e.ms_best = best_ms;
  tune_store(key, e);
  if (st) {
    *st = best_st;
    st->schedule_source = 2;
  }
  return ref;
}

/// Queries can share a D factor table when y >= z (then no entry <= z
/// depends on y, src/gourdon/FactorTableD.hpp:168-190) -- the default
/// parameters always have y == z.  Everything else runs on its own table.
std::vector<std::vector<size_t>> group_queries(const std::vector<HslQuery>& qs)
{
  std::vector<std::vector<size_t>> groups;
  std::vector<size_t> shared_d, shared_s2;
  for (size_t i = 0; i < qs.size(); i++) {
    if (qs[i].kind == LeafSum::S2_hard) shared_s2.push_back(i);
    else if (qs[i].y >= qs[i].z) shared_d.push_back(i);
    else groups.push_back({i});
  }
  if (!shared_d.empty()) groups.push_back(shared_d);
  if (!shared_s2.empty()) groups.push_back(shared_s2);
  return groups;
}

} // namespace

namespace {

/// Tables of the no-allocation entry point, at the start of the device's
/// scratch array; kept while they cover the next query.
struct ScratchTables {
  bool valid = false;
  TableSet T;
  size_t primes = 0, pi_bits = 0, pi_cnt = 0, factor = 0;   // byte offsets
  size_t end = 0;
};
ScratchTables g_scratch_tables;

/// Bump placement of buffers in the scratch array, 256-byte aligned.
struct ScratchLayout {
  size_t used;
  size_t take(size_t bytes)
  {
    const size_t off = (used + 255) & ~(size_t) 255;
    used = off + std::max<size_t>(bytes, 16);
    return off;
  }
};

std::string mib(size_t bytes)
{
  return std::to_string((bytes + (1 << 20) - 1) >> 20) + " MiB";
}

} // namespace

/// D(x, y, z, k) for one query with every table and work buffer in the
/// device's scratch array: one pass over all segments, no device or pinned
/// host allocation.  Calls are serialised by the engine lock (they share the
/// array, which keeps the last call's tables; everything a call needs zeroed
/// is cleared by fill kernels).  Queries whose buffers exceed the array are
/// rejected (kErrInvalid).
i128 hsl_compute_noalloc(const HslQuery& q, pcg_stats* stats)
{
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  Device& D = g_dev;
  D.init();
  const double t0 = now();
  if (q.kind != LeafSum::D) throw pcg_error("internal: no-allocation path is for D", kErrInternal);
  validate(q);
  if (q.x < 1) return 0;
  hipStream_t s = D.stream;
  char* const base = D.scratch;
  ScratchTables& ST = g_scratch_tables;
  const TableNeed need = table_need({q});
  if (!(ST.valid && covers(ST.T, need))) {
    ST.valid = false;
    TableSet T = build_tables(need);
    ScratchLayout L{0};
    const size_t o_primes = L.take(T.P.primes.size() * 4);
    const size_t o_bits = L.take(T.P.pi_bits.size() * 8);
    const size_t o_cnt = L.take(T.P.pi_cnt.size() * 4);
    const size_t o_fac = L.take(T.factor_n * (T.factor32 ? 4 : 2));
    if (L.used > kHslScratchBytes)
      throw pcg_error("the no-allocation entry point holds " + mib(kHslScratchBytes) +
                      " of device scratch; the tables of this query need " + mib(L.used));
    PCG_HIP(hipMemcpyAsync(base + o_primes, T.P.primes.data(), T.P.primes.size() * 4,
                           hipMemcpyHostToDevice, s));
    PCG_HIP(hipMemcpyAsync(base + o_bits, T.P.pi_bits.data(), T.P.pi_bits.size() * 8,
                           hipMemcpyHostToDevice, s));
    PCG_HIP(hipMemcpyAsync(base + o_cnt, T.P.pi_cnt.data(), T.P.pi_cnt.size() * 4,
                           hipMemcpyHostToDevice, s));
    PCG_HIP(hsl_launch_factor_table(base + o_fac, T.factor32, true, (uint64_t) T.ft_limit,
                                    (uint64_t) T.ft_y, (const uint32_t*) (base + o_primes),
                                    (uint32_t) T.P.primes.size(), s));
    PCG_HIP(hipStreamSynchronize(s));
    ST.T = std::move(T);
    ST.primes = o_primes;
    ST.pi_bits = o_bits;
    ST.pi_cnt = o_cnt;
    ST.factor = o_fac;
    ST.end = L.used;
    ST.valid = true;
  }
  const TableSet& T = ST.T;
  const double t_tab = now();

  QueryPlan qp = plan_query(q, T.P);
  if (qp.empty) return 0;
  const uint32_t W = auto_segment_words(qp.qd.limit + 1, T.kind, true);
  if (W < 128 || W > 4096 || (W % 128) != 0)
    throw pcg_error("segment words must be a multiple of 128 in [128, 4096]");
  const uint32_t cbits = std::max(slot_format(W).cbits, engine_config().slot_cbits);
  if (cbits > 28) throw pcg_error("slot bits must be at most 28");
  plan_reach(qp, q, T.P, W, cbits);
  if (qp.nseg == 0) return 0;

  // one pass: every segment of the query, one column per b
  const uint64_t slots = qp.F(qp.nseg);
  const uint32_t nitems = qp.nseg;
  const uint32_t ncols = qp.nbs(0);
  const uint32_t max_n = qp.reach[0];
  const uint32_t slice_len = std::max<uint32_t>(1, (max_n + 127) / 128);
  const uint32_t nslices = (max_n + slice_len - 1) / slice_len;
  const uint64_t bound = escape_bound(qp, 0, qp.nseg);
  uint64_t cap = 1024;
  while (cap < 2 * bound + 2) cap *= 2;
  const size_t nb = qp.binfo.size();
  ScratchLayout L{ST.end};
  const size_t o_qd = L.take(sizeof(QueryDesc)), o_binfo = L.take(nb * sizeof(BInfo));
  const size_t o_reach = L.take(nb * 4), o_sfx = L.take((nb + 1) * 8);
  const size_t o_span = L.take(sizeof(PassSpan)), o_items = L.take((size_t) nitems * sizeof(SegItem));
  const size_t o_cols = L.take((size_t) ncols * sizeof(CombineCol));
  const size_t o_slots = L.take(slots * 4);
  const size_t o_ekeys = L.take(cap * 8), o_evals = L.take(cap * 4), o_ectl = L.take(8);
  const size_t part = (size_t) nslices * ncols;
  const size_t o_sc = L.take(part * 8), o_sm = L.take(part * 8), o_smp = L.take(part * 16);
  const size_t o_carry = L.take((size_t) ncols * 8), o_out = L.take((size_t) ncols * 16);
  const size_t o_sacc = L.take(16), o_cnt = L.take(128), o_queue = L.take(4);
  if (L.used > kHslScratchBytes || slots > ((uint64_t) 1 << 32))
    throw pcg_error("the no-allocation entry point holds " + mib(kHslScratchBytes) +
                    " of device scratch; D(x, y, z, k) at x = " + to_string128(q.x) + " needs " +
                    mib(L.used) + " (" + std::to_string(slots) + " slots)");

  QueryDesc qd = qp.qd;
  qd.binfo_off = 0;
  qd.reach_off = 0;
  qd.sfx_off = 0;
  PassSpan sp{};
  sp.query = 0;
  sp.s0 = 0;
  sp.s1 = qp.nseg;
  sp.item_base = 0;
  sp.t_base = 0;
  std::vector<CombineCol> cols(ncols);
  for (uint32_t br = 0; br < ncols; br++) cols[br] = CombineCol{br, 0, qp.reach[br], br};
  PCG_HIP(hipMemcpyAsync(base + o_qd, &qd, sizeof(qd), hipMemcpyHostToDevice, s));
  PCG_HIP(hipMemcpyAsync(base + o_binfo, qp.binfo.data(), nb * sizeof(BInfo), hipMemcpyHostToDevice, s));
  PCG_HIP(hipMemcpyAsync(base + o_reach, qp.reach.data(), nb * 4, hipMemcpyHostToDevice, s));
  PCG_HIP(hipMemcpyAsync(base + o_sfx, qp.sfx.data(), (nb + 1) * 8, hipMemcpyHostToDevice, s));
  PCG_HIP(hipMemcpyAsync(base + o_span, &sp, sizeof(sp), hipMemcpyHostToDevice, s));
  PCG_HIP(hipMemcpyAsync(base + o_cols, cols.data(), (size_t) ncols * sizeof(CombineCol),
                         hipMemcpyHostToDevice, s));
  // the array keeps the previous call's contents: clear what starts at zero
  PCG_HIP(hsl_launch_fill_zero(base + o_ekeys, cap * 8, s));
  PCG_HIP(hsl_launch_fill_zero(base + o_ectl, 8, s));
  PCG_HIP(hsl_launch_fill_zero(base + o_carry, (size_t) ncols * 8, s));
  PCG_HIP(hsl_launch_fill_zero(base + o_out, (size_t) ncols * 16, s));
  PCG_HIP(hsl_launch_fill_zero(base + o_sacc, 16, s));
  PCG_HIP(hsl_launch_fill_zero(base + o_cnt, 128, s));
  PCG_HIP(hsl_launch_fill_zero(base + o_queue, 4, s));
  PCG_HIP(hsl_launch_build_items((const QueryDesc*) (base + o_qd), (const PassSpan*) (base + o_span),
                                 1, (const uint32_t*) (base + o_reach),
                                 (const uint64_t*) (base + o_sfx), (SegItem*) (base + o_items),
                                 nitems, s));
  int bpc = 0;
  PCG_HIP(hsl_sieve_occupancy(T.factor32, W, 0, &bpc));
  bpc = std::max(bpc, 1);
  HslArgs a{};
  a.queries = (const QueryDesc*) (base + o_qd);
  a.items = (const SegItem*) (base + o_items);
  a.groups = nullptr;
  a.binfo = (const BInfo*) (base + o_binfo);
  a.primes = (const uint32_t*) (base + ST.primes);
  a.factor = base + ST.factor;
  a.pi_bits = (const uint64_t*) (base + ST.pi_bits);
  a.pi_cnt = (const uint32_t*) (base + ST.pi_cnt);
  a.T_S = (uint32_t*) (base + o_slots);
  a.cbits = cbits;
  a.esc.keys = (unsigned long long*) (base + o_ekeys);
  a.esc.vals = (int32_t*) (base + o_evals);
  a.esc.ctl = (unsigned int*) (base + o_ectl);
  a.esc.cap = (uint32_t) cap;
  a.S_acc = (unsigned long long*) (base + o_sacc);
  a.counters = (unsigned long long*) (base + o_cnt);
  a.queue = (unsigned int*) (base + o_queue);
  a.nitems = nitems;
  a.W = W;
  a.bcap = engine_config().leaf_batching ? kBatchRecords : 0u;
  const uint32_t grid = (uint32_t) std::min<uint64_t>((uint64_t) D.cus * bpc, nitems);
  PCG_HIP(hsl_launch_sieve(a, T.factor32, grid, s));
  PCG_HIP(hsl_launch_combine((const CombineCol*) (base + o_cols), ncols, a.items, a.T_S, a.esc,
                             cbits, slice_len, nslices, (long long*) (base + o_sc),
                             (long long*) (base + o_sm), base + o_smp,
                             (long long*) (base + o_carry), base + o_out, s));
  std::vector<i128> col_out(ncols);
  uint64_t sacc[2] = {}, counters[14] = {};
  uint32_t ectl[2] = {};
  PCG_HIP(hipMemcpyAsync(col_out.data(), base + o_out, (size_t) ncols * 16, hipMemcpyDeviceToHost, s));
  PCG_HIP(hipMemcpyAsync(sacc, base + o_sacc, 16, hipMemcpyDeviceToHost, s));
  PCG_HIP(hipMemcpyAsync(counters, base + o_cnt, sizeof(counters), hipMemcpyDeviceToHost, s));
  PCG_HIP(hipMemcpyAsync(ectl, base + o_ectl, 8, hipMemcpyDeviceToHost, s));
  PCG_HIP(hipStreamSynchronize(s));
  if (ectl[1] != 0)
    throw pcg_error("internal: escape table overflow", kErrInternal);
  i128 r = (i128) (((u128) sacc[1] << 64) | sacc[0]);
  for (const i128& v : col_out) r += v;
  if (stats) {
    *stats = pcg_stats{};
    stats->seconds_total = now() - t0;
    stats->seconds_tables = t_tab - t0;
    stats->segments = nitems;
    stats->chunks = nitems;
    stats->segment_words = W;
    stats->sieve_limit = qp.qd.limit;
    stats->b_first = qp.qd.b_first;
    stats->b_last = qp.qd.b_last;
    stats->leaves = counters[0];
    stats->hits = counters[1];
    stats->factor_entries = counters[2];
    stats->pairs = counters[3];
    stats->device_allocations = g_allocations.load();
    stats->escaped_slots = ectl[0];
    stats->escape_bound = bound;
  }
  return r;
}

i128 hsl_compute(const HslQuery& q, pcg_stats* stats)
{
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  auto r = run_group({q}, stats);
  return r[0];
}

std::vector<i128> hsl_compute_batch(const std::vector<HslQuery>& qs, pcg_stats* stats,
                                    uint64_t max_slots)
{
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  return hsl_compute_batch_locked(qs, stats, max_slots);
}

std::vector<i128> hsl_compute_batch_locked(const std::vector<HslQuery>& qs, pcg_stats* stats,
                                           uint64_t max_slots)
{
  std::vector<i128> out(qs.size(), 0);
  double t0 = now();
  pcg_stats acc{};
  for (const auto& g : group_queries(qs)) {
    std::vector<HslQuery> sub;
    for (size_t i : g) sub.push_back(qs[i]);
    pcg_stats st{};
    auto r = run_group(sub, &st, max_slots);
    for (size_t j = 0; j < g.size(); j++) out[g[j]] = r[j];
    acc.seconds_tables += st.seconds_tables;
    acc.seconds_plan += st.seconds_plan;
    acc.seconds_kernel += st.seconds_kernel;
    acc.seconds_combine += st.seconds_combine;
    acc.segments += st.segments;
    acc.chunks += st.chunks;
    acc.segment_words = st.segment_words;
    acc.leaves += st.leaves;
    acc.hits += st.hits;
    acc.factor_entries += st.factor_entries;
    acc.pairs += st.pairs;
    acc.windows += st.windows;
    acc.leaf_primes += st.leaf_primes;
    acc.leaf_steps += st.leaf_steps;
    acc.dense_primes += st.dense_primes;
    acc.dense_steps += st.dense_steps;
    acc.group_steps += st.group_steps;
    acc.prefix_builds += st.prefix_builds;
    acc.queue_steps += st.queue_steps;
    acc.leaf_batches += st.leaf_batches;
    acc.batch_records += st.batch_records;
    for (uint32_t i = 0; i < kHslPhases; i++)
      acc.phase_cycles[i] += st.phase_cycles[i];
    acc.escaped_slots += st.escaped_slots;
    acc.escape_bound += st.escape_bound;
    acc.schedule_source = std::max(acc.schedule_source, st.schedule_source);
    acc.group_size = std::max(acc.group_size, st.group_size);
    acc.grid_waves = std::max(acc.grid_waves, st.grid_waves);
    acc.kernel_waves = std::max(acc.kernel_waves, st.kernel_waves);
  }
  acc.device_allocations = g_allocations.load();
  acc.seconds_total = now() - t0;
  if (stats) *stats = acc;
  return out;
}

uint64_t factor_table_mismatches(LeafSum kind, int64_t y, int64_t limit)
{
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  g_dev.init();
  if (y < 1 || limit < 1 || (uint64_t) limit >= (1ull << 32))
    throw pcg_error("factor table check needs y >= 1 and 1 <= limit < 2^32");
  const bool d_table = kind == LeafSum::D;
  const bool f32 = d_table ? limit > kFactorTableD16Max : limit > kFactorTable16Max;
  const int threads = default_threads();
  const PrimeTables P = build_prime_tables(std::max<uint64_t>(isqrt((u128) limit), 13) + 1, threads);
  std::vector<uint32_t> host;
  if (d_table) {
    if (f32) host = build_factor_table_D<uint32_t>(y, limit, P, threads);
    else for (uint16_t v : build_factor_table_D<uint16_t>(y, limit, P, threads)) host.push_back(v);
  } else {
    if (f32) host = build_factor_table_lpf<uint32_t>(limit, P, threads);
    else for (uint16_t v : build_factor_table_lpf<uint16_t>(limit, P, threads)) host.push_back(v);
  }
  DevBuf dp, dout;
  dp.ensure(P.primes.size() * 4);
  dout.ensure(host.size() * (f32 ? 4 : 2));
  uint64_t bad = 0;
  try {
    PCG_HIP(hipMemcpy(dp.p, P.primes.data(), P.primes.size() * 4, hipMemcpyHostToDevice));
    PCG_HIP(hsl_launch_factor_table(dout.p, f32, d_table, (uint64_t) limit, (uint64_t) y,
                                    (const uint32_t*) dp.p, (uint32_t) P.primes.size(), g_dev.stream));
    PCG_HIP(hipStreamSynchronize(g_dev.stream));
    if (f32) {
      std::vector<uint32_t> dev(host.size());
      PCG_HIP(hipMemcpy(dev.data(), dout.p, dev.size() * 4, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < host.size(); i++) bad += dev[i] != host[i];
    } else {
      std::vector<uint16_t> dev(host.size());
      PCG_HIP(hipMemcpy(dev.data(), dout.p, dev.size() * 2, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < host.size(); i++) bad += dev[i] != host[i];
    }
  } catch (...) {
    dp.release();
    dout.release();
    throw;
  }
  dp.release();
  dout.release();
  return bad;
}

void hsl_validate(const HslQuery& q)
{
  validate(q);
}

void hsl_reserve(LeafSum kind, int64_t y, int64_t z)
{
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  g_dev.init();
  HslQuery q;
  q.kind = kind;
  q.x = 1;
  q.y = y;
  q.z = z;
  validate(q);
  (void) resident_tables({q});
}

bool device_available()
{
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  try {
    g_dev.init();
    return true;
  } catch (...) {
    return false;
  }
}

std::unique_lock<std::recursive_mutex> engine_lock()
{
  return std::unique_lock<std::recursive_mutex>(g_mutex);
}

Work* engine_work_create()
{
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  Work* w = new Work();
  g_works.push_back(w);
  return w;
}

void engine_work_destroy(Work* w)
{
  if (!w) return;
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  g_works.erase(std::remove(g_works.begin(), g_works.end(), w), g_works.end());
  if (t_work == w) t_work = nullptr;
  delete w;
}

Work* engine_work_select(Work* w)
{
  Work* prev = t_work;
  t_work = w;
  return prev;
}

PassHookScope::PassHookScope(std::function<void()> f) : f_(std::move(f)), prev_(t_pass_hook)
{
  t_pass_hook = &f_;
}

PassHookScope::~PassHookScope()
{
  t_pass_hook = prev_;
}

void engine_pass_point()
{
  if (!t_pass_hook) return;
  const std::function<void()>* h = t_pass_hook;
  t_pass_hook = nullptr;              // calls made by the hook see no hook
  t_tables_frozen++;
  try {
    (*h)();
  } catch (...) {
    t_tables_frozen--;
    t_pass_hook = h;
    throw;
  }
  t_tables_frozen--;
  t_pass_hook = h;
}

PhiUpload& engine_phi_upload()
{
  return cur_work().phi;
}

hipEvent_t* engine_term_events()
{
  return cur_work().term_ev;
}

void* engine_buffer(int slot, size_t bytes)
{
  if (slot < 0 || slot >= 8) throw pcg_error("internal: engine buffer slot", kErrInternal);
  return cur_work().extra[slot].ensure(bytes);
}

/// Prime-count table over [0, n] built on the device: segmented sieve of
/// the numbers coprime to 30 with the primes 7 .. isqrt(n) (from the resident
/// prime table), then the per-word counts.  Prefix valid, so a larger table
/// serves every smaller query.
static void ensure_pix_table(uint64_t n)
{
  Device& D = g_dev;
  if (D.pix_n >= n) return;
  throw_if_frozen();
  n = std::max<uint64_t>(n, 1u << 20);
  const TableSet& T = g_tables;
  const uint64_t r = isqrt(n);
  if (r > T.P.max) throw pcg_error("internal: prime table too small for the prime-count table", kErrInternal);
  const uint64_t nwords = n / 240 + 1;
  const uint64_t nseg = (nwords + kPixSegWords - 1) / kPixSegWords;
  D.pix_n = 0;
  D.pix_tab.ensure(nwords * sizeof(PixWord));
  D.pix_seg.ensure(nseg * 8);
  const uint32_t j_end = (uint32_t) T.P.pi(r) + 1;               // primes[4 .. pi(r)]
  const uint32_t j_small = (uint32_t) std::min<uint64_t>(T.P.pi(std::min<uint64_t>(r, 240ull * kPixSegWords / 64)) + 1, j_end);
  hipStream_t s = D.stream;
  PCG_HIP(pix_launch_sieve((PixWord*) D.pix_tab.p, (uint64_t*) D.pix_seg.p,
                           (const uint32_t*) D.primes.p, std::max<uint32_t>(j_small, 4),
                           std::max<uint32_t>(j_end, 4), nwords, n, 0, s));
  PCG_HIP(pix_launch_scan((uint64_t*) D.pix_seg.p, (uint32_t) nseg, s));
  PCG_HIP(pix_launch_counts((PixWord*) D.pix_tab.p, (const uint64_t*) D.pix_seg.p, nwords, 0, s));
  PCG_HIP(hipStreamSynchronize(s));
  D.pix_n = nwords * 240 - 1 < n ? nwords * 240 - 1 : n;
}

uint64_t engine_pix_budget()
{
  const EngineConfig& cfg = engine_config();
  if (cfg.pix_table_bytes) return cfg.pix_table_bytes;
  g_dev.init();
  size_t free_b = 0, total_b = 0;
  PCG_HIP(hipMemGetInfo(&free_b, &total_b));
  return (uint64_t) (0.65 * (double) total_b);
}

const PixWord* engine_pix_window(uint64_t w_lo, uint64_t nwords, uint64_t n_max, uint64_t base,
                                 uint64_t* next_base)
{
  Device& D = g_dev;
  throw_if_frozen();
  const TableSet& T = g_tables;
  if (nwords == 0) throw pcg_error("internal: empty prime-count window", kErrInternal);
  // numbers [240 w_lo, min(240 (w_lo + nwords), n_max + 1)): sieving primes
  // up to the square root of the last one
  const uint64_t last = std::min<uint64_t>(240 * (w_lo + nwords) - 1, n_max);
  if (last < 240 * w_lo) throw pcg_error("internal: prime-count window past its limit", kErrInternal);
  const uint64_t r = isqrt(last);
  if (r > T.P.max) throw pcg_error("internal: prime table too small for the prime-count window", kErrInternal);
  const uint64_t nseg = (nwords + kPixSegWords - 1) / kPixSegWords;
  D.pix_win.ensure(nwords * sizeof(PixWord));
  D.pix_seg.ensure(nseg * 8);
  const uint32_t j_end = (uint32_t) T.P.pi(r) + 1;
  const uint32_t j_small = (uint32_t) std::min<uint64_t>(T.P.pi(std::min<uint64_t>(r, 240ull * kPixSegWords / 64)) + 1, j_end);
  hipStream_t s = D.stream;
  PixWord* tab = (PixWord*) D.pix_win.p;
  PCG_HIP(pix_launch_sieve(tab, (uint64_t*) D.pix_seg.p, (const uint32_t*) D.primes.p,
                           std::max<uint32_t>(j_small, 4), std::max<uint32_t>(j_end, 4), nwords,
                           n_max, w_lo, s));
  PCG_HIP(pix_launch_scan((uint64_t*) D.pix_seg.p, (uint32_t) nseg, s));
  PCG_HIP(pix_launch_counts(tab, (const uint64_t*) D.pix_seg.p, nwords, base, s));
  PixWord lastw;
  PCG_HIP(hipMemcpyAsync(&lastw, tab + nwords - 1, sizeof(PixWord), hipMemcpyDeviceToHost, s));
  PCG_HIP(hipStreamSynchronize(s));
  *next_base = lastw.cnt + (uint64_t) __builtin_popcountll(lastw.bits);
  return tab;
}

EngineView engine_view(int64_t y, int64_t z, uint64_t pix_n, LeafSum kind, uint64_t pmin)
{
  Device& D = g_dev;
  D.init();
  HslQuery q;
  q.kind = kind;
  q.x = 1;
  q.y = y;
  q.z = z;
  validate(q);
  const TableSet& T = resident_tables({q}, pmin);
  if (pix_n) ensure_pix_table(pix_n);
  EngineView v;
  v.stream = D.stream;
  v.cus = D.cus;
  v.primes = (const uint32_t*) D.primes.p;
  v.nprimes = T.P.primes.size();
  v.host = &T.P;
  v.factor = D.factor.p;
  v.factor32 = T.factor32;
  v.factor_n = T.factor_n;
  v.pix_tab = (const PixWord*) D.pix_tab.p;
  v.pix_n = D.pix_n;
  return v;
}

static uint64_t g_generation = 1;

uint64_t engine_generation()
{
  return g_generation;
}

void release_device()
{
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  g_generation++;
  g_dev.release();
  for (Work* w : g_works) w->release();
  g_tables = TableSet();
  g_tables_valid = false;
}

} // namespace pcg
