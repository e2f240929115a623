// capi.cpp -- the C ABI declared in include/primecount_mi355x.h.
//
// Every entry point converts C++ exceptions into a negative status and a
// thread-local message, the way primecount's C API converts primecount_error
// into -1 (src/api_c.cpp:19-56).

#include "../../include/primecount_mi355x.h"

#include "engine.hpp"
#include "params.hpp"
#include "pix.hpp"
#include "serve.hpp"
#include "tuner.hpp"

#include <exception>
#include <new>
#include <string>

using namespace pcg;

namespace pcg {
hipError_t lds_read_gbs(double* gbs);      // lds_probe.hip
}

namespace {

thread_local std::string t_error;
thread_local pcg_stats t_stats{};
thread_local pcg_pi_stats t_pi_stats{};

i128 from_abi(pcg_i128 v)
{
  return (i128) (((u128) (uint64_t) v.hi << 64) | (u128) v.lo);
}

pcg_i128 to_abi(i128 v)
{
  pcg_i128 r;
  r.lo = (uint64_t) v;
  r.hi = (int64_t) (v >> 64);
  return r;
}

template <typename F>
int guarded(F&& f)
{
  try {
    t_error.clear();
    f();
    return PCG_OK;
  } catch (const std::bad_alloc&) {
    t_error = "out of memory";
    return PCG_E_NOMEM;
  } catch (const pcg_error& e) {
    t_error = e.what();
    return e.status();
  } catch (const std::exception& e) {
    t_error = e.what();
    return PCG_E_INTERNAL;
  } catch (...) {
    t_error = "unknown error";
    return PCG_E_INTERNAL;
  }
}

i128 run(LeafSum kind, i128 x, int64_t y, int64_t z, int64_t k)
{
  HslQuery q;
  q.kind = kind;
  q.x = x;
  q.y = y;
  q.z = z;
  q.k = k;
  pcg_stats st{};
  i128 r = hsl_compute(q, &st);
  t_stats = st;
  return r;
}

HslQuery default_query(LeafSum kind, i128 x)
{
  HslQuery q;
  q.kind = kind;
  q.x = x;
  if (x < 1) return q;
  if (kind == LeafSum::D) {
    GourdonParams p = gourdon_params(x);
    q.y = p.y; q.z = p.z; q.k = p.k;
  } else {
    DRParams p = deleglise_rivat_params(x);
    q.y = p.y; q.z = p.z; q.k = p.c;
  }
  return q;
}

} // namespace

extern "C" {

static_assert(kErrInvalid == PCG_E_INVALID && kErrDevice == PCG_E_DEVICE &&
                  kErrNoMem == PCG_E_NOMEM && kErrInternal == PCG_E_INTERNAL,
              "pcg_error status codes are the C ABI's");
static_assert(kTuneOff == PCG_TUNE_OFF && kTuneCached == PCG_TUNE_CACHED && kTuneOn == PCG_TUNE_ON,
              "autotune modes are the C ABI's");

const char* pcg_last_error(void) { return t_error.c_str(); }

int pcg_lds_read_gbs(double* gbs)
{
  return guarded([&] {
    if (!gbs) throw pcg_error("gbs must not be NULL");
    const hipError_t e = lds_read_gbs(gbs);
    if (e != hipSuccess)
      throw pcg_error(std::string("HIP error '") + hipGetErrorString(e) + "' in the LDS probe",
                      e == hipErrorOutOfMemory ? kErrNoMem : kErrDevice);
  });
}

int pcg_check_pi(pcg_i128 x, pcg_i128 pix)
{
  return guarded([&] { verify_pix("pi", from_abi(x), from_abi(pix)); });
}

const char* pcg_version(void) { return "0.1.0 (gfx950)"; }

int pcg_device_available(void) { return device_available() ? 1 : 0; }

int pcg_D_i64(int64_t x, int64_t y, int64_t z, int64_t k, int64_t* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = (int64_t) run(LeafSum::D, x, y, z, k);
  });
}

int pcg_D_i64_noalloc(int64_t x, int64_t y, int64_t z, int64_t k, int64_t* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    HslQuery q;
    q.kind = LeafSum::D;
    q.x = x;
    q.y = y;
    q.z = z;
    q.k = k;
    pcg_stats st{};
    *out = (int64_t) hsl_compute_noalloc(q, &st);
    t_stats = st;
  });
}

int pcg_D_i128(pcg_i128 x, int64_t y, int64_t z, int64_t k, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(run(LeafSum::D, from_abi(x), y, z, k));
  });
}

int pcg_S2_hard_i64(int64_t x, int64_t y, int64_t z, int64_t c, int64_t* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = (int64_t) run(LeafSum::S2_hard, x, y, z, c);
  });
}

int pcg_S2_hard_i128(pcg_i128 x, int64_t y, int64_t z, int64_t c, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(run(LeafSum::S2_hard, from_abi(x), y, z, c));
  });
}

int pcg_D(pcg_i128 x, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    HslQuery q = default_query(LeafSum::D, from_abi(x));
    *out = to_abi(q.x < 1 ? 0 : run(LeafSum::D, q.x, q.y, q.z, q.k));
  });
}

int pcg_S2_hard(pcg_i128 x, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    HslQuery q = default_query(LeafSum::S2_hard, from_abi(x));
    *out = to_abi(q.x < 1 ? 0 : run(LeafSum::S2_hard, q.x, q.y, q.z, q.k));
  });
}

int pcg_gourdon_params(pcg_i128 x, int64_t* y, int64_t* z, int64_t* k, int64_t* x_star)
{
  return guarded([&] {
    GourdonParams p = gourdon_params(from_abi(x));
    if (y) *y = p.y;
    if (z) *z = p.z;
    if (k) *k = p.k;
    if (x_star) *x_star = p.x_star;
  });
}

int pcg_dr_params(pcg_i128 x, int64_t* y, int64_t* z, int64_t* c)
{
  return guarded([&] {
    DRParams p = deleglise_rivat_params(from_abi(x));
    if (y) *y = p.y;
    if (z) *z = p.z;
    if (c) *c = p.c;
  });
}

static void set_pi_stats(const PixStats& st)
{
  t_pi_stats.seconds_total = st.seconds_total;
  t_pi_stats.seconds_table = st.seconds_table;
  t_pi_stats.seconds_sigma = st.seconds_sigma;
  t_pi_stats.seconds_phi0 = st.seconds_phi0;
  t_pi_stats.seconds_ac = st.seconds_ac;
  t_pi_stats.seconds_b = st.seconds_b;
  t_pi_stats.seconds_d = st.seconds_d;
  t_pi_stats.ac_leaves = st.ac_leaves;
  t_pi_stats.table_limit = st.table_limit;
  t_pi_stats.table_windows = st.table_windows;
}

int pcg_nth_prime_128(pcg_i128 n, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    PixStats st;
    *out = to_abi(nth_prime_128(from_abi(n), &st));
    set_pi_stats(st);
  });
}

int pcg_pi(pcg_i128 x, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    PixStats st;
    *out = to_abi(pi_gourdon(from_abi(x), &st));
    set_pi_stats(st);
  });
}

int pcg_nth_prime(int64_t n, int64_t* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    if (n < 1) throw pcg_error("nth_prime(n): n must be >= 1");
    PixStats st;
    *out = (int64_t) nth_prime((uint64_t) n, &st);
    set_pi_stats(st);
  });
}

int pcg_last_pi_stats(pcg_pi_stats* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = t_pi_stats;
  });
}

int pcg_AC(pcg_i128 x, int64_t y, int64_t z, int64_t k, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(gourdon_term(GourdonTerm::AC, from_abi(x), y, z, k, nullptr));
  });
}

int pcg_B(pcg_i128 x, int64_t y, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(gourdon_term(GourdonTerm::B, from_abi(x), y, y, 0, nullptr));
  });
}

int pcg_Phi0(pcg_i128 x, int64_t y, int64_t z, int64_t k, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(gourdon_term(GourdonTerm::Phi0, from_abi(x), y, z, k, nullptr));
  });
}

int pcg_Sigma(pcg_i128 x, int64_t y, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(gourdon_term(GourdonTerm::Sigma, from_abi(x), y, y, 0, nullptr));
  });
}

int pcg_pi_deleglise_rivat(pcg_i128 x, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(pi_deleglise_rivat(from_abi(x)));
  });
}

int pcg_S1(pcg_i128 x, int64_t y, int64_t c, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(dr_term(DRTerm::S1, from_abi(x), y, y, c));
  });
}

int pcg_S2_trivial(pcg_i128 x, int64_t y, int64_t z, int64_t c, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(dr_term(DRTerm::S2_trivial, from_abi(x), y, z, c));
  });
}

int pcg_S2_easy(pcg_i128 x, int64_t y, int64_t z, int64_t c, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(dr_term(DRTerm::S2_easy, from_abi(x), y, z, c));
  });
}

int pcg_P2(pcg_i128 x, int64_t y, pcg_i128* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = to_abi(dr_term(DRTerm::P2, from_abi(x), y, y, 8));
  });
}

int pcg_ctx_create(int64_t y_max, int64_t z_max, uint64_t hbm_budget, uint32_t max_batch,
                   pcg_ctx** ctx)
{
  return guarded([&] {
    if (!ctx) throw pcg_error("ctx must not be NULL");
    *ctx = nullptr;
    if (y_max < 0 || z_max < 0) throw pcg_error("y_max and z_max must be >= 0");
    if (!device_available())
      throw pcg_error("no HIP device visible: the MI355X engine needs a gfx950 GPU", kErrDevice);
    // default parameters have y == z: one shared D table covers them all
    if (y_max > 0 || z_max > 0) {
      const int64_t m = std::max(y_max, z_max);
      hsl_reserve(LeafSum::D, m, m);
    }
    *ctx = reinterpret_cast<pcg_ctx*>(new Server(hbm_budget, max_batch));
  });
}

int pcg_submit(pcg_ctx* ctx, const pcg_query* q, size_t n, uint64_t* ids)
{
  return guarded([&] {
    if (!ctx) throw pcg_error("ctx must not be NULL");
    if (n && (!q || !ids)) throw pcg_error("q and ids must not be NULL");
    Server* sv = reinterpret_cast<Server*>(ctx);
    for (size_t i = 0; i < n; i++) {
      ServeQuery sq;
      int status = PCG_OK;
      std::string err;
      try {
        const i128 x = from_abi(q[i].x);
        const bool dflt = q[i].y == 0 && q[i].z == 0 && q[i].k == 0;
        if (q[i].kind == PCG_KIND_PI || q[i].kind == PCG_KIND_NTH_PRIME) {
          if (!dflt) throw pcg_error("pi and nth_prime queries take no parameters");
          sq.kind = q[i].kind == PCG_KIND_PI ? ServeQuery::PI : ServeQuery::NTH_PRIME;
          sq.arg = x;
          if (sq.kind == ServeQuery::NTH_PRIME && (x < 1 || x > (i128) 216289611853439384LL))
            throw pcg_error("nth_prime(n): n must be in [1, 216289611853439384]");
          // the domain checks of the device path, so that an out-of-domain
          // query is answered at once instead of failing inside a batch
          if (sq.kind == ServeQuery::PI)
            pi_validate(x);
        } else if (q[i].kind == PCG_KIND_D || q[i].kind == PCG_KIND_S2_HARD) {
          const LeafSum kind = q[i].kind == PCG_KIND_D ? LeafSum::D : LeafSum::S2_hard;
          if (dflt) {
            sq.q = default_query(kind, x);
          } else {
            sq.q.kind = kind;
            sq.q.x = x;
            sq.q.y = q[i].y;
            sq.q.z = q[i].z;
            sq.q.k = q[i].k;
          }
          hsl_validate(sq.q);
        } else {
          throw pcg_error("unknown query kind");
        }
      } catch (const std::exception& e) {
        status = status_of(e);
        err = e.what();
      }
      ids[i] = sv->submit(sq, status, err);
    }
  });
}

int pcg_poll(pcg_ctx* ctx, pcg_result* out, size_t max, int timeout_ms, size_t* n)
{
  return guarded([&] {
    if (!ctx || !n) throw pcg_error("ctx and n must not be NULL");
    if (max && !out) throw pcg_error("out must not be NULL");
    *n = 0;
    std::vector<ServeResult> res;
    reinterpret_cast<Server*>(ctx)->poll(res, max, timeout_ms);
    for (size_t i = 0; i < res.size(); i++) {
      out[i].id = res[i].id;
      out[i].status = res[i].status;
      out[i].reserved = 0;
      out[i].value = to_abi(res[i].value);
      out[i].seconds = res[i].seconds;
    }
    *n = res.size();
  });
}

int pcg_ctx_get_stats(pcg_ctx* ctx, pcg_ctx_stats* out)
{
  return guarded([&] {
    if (!ctx || !out) throw pcg_error("ctx and out must not be NULL");
    ServeStats st = reinterpret_cast<Server*>(ctx)->stats();
    out->submitted = st.submitted;
    out->completed = st.completed;
    out->batches = st.batches;
    out->max_batch = st.max_batch;
    out->light_batches = st.light_batches;
    out->heavy_batches = st.heavy_batches;
    out->interleaved = st.interleaved;
    out->rejected_nomem = st.rejected_nomem;
    out->busy_seconds = st.busy_seconds;
  });
}

int pcg_ctx_destroy(pcg_ctx* ctx)
{
  return guarded([&] { delete reinterpret_cast<Server*>(ctx); });
}

int pcg_set_alpha_y(double a) { return guarded([&] { set_alpha_y(a); }); }
int pcg_set_alpha_z(double a) { return guarded([&] { set_alpha_z(a); }); }
int pcg_set_alpha(double a) { return guarded([&] { set_alpha(a); }); }

int pcg_D_batch(const pcg_i128* xs, size_t n, pcg_i128* out)
{
  return guarded([&] {
    if (n && (!xs || !out)) throw pcg_error("xs and out must not be NULL");
    std::vector<HslQuery> qs;
    qs.reserve(n);
    for (size_t i = 0; i < n; i++) qs.push_back(default_query(LeafSum::D, from_abi(xs[i])));
    pcg_stats st{};
    auto r = hsl_compute_batch(qs, &st);
    t_stats = st;
    for (size_t i = 0; i < n; i++) out[i] = to_abi(r[i]);
  });
}

int pcg_set_segment_words(uint32_t words)
{
  return guarded([&] {
    if (words != 0 && (words < 128 || words > 4096 || (words % 128) != 0))
      throw pcg_error("segment words must be 0 or a multiple of 128 in [128, 4096]");
    engine_config().segment_words = words;
  });
}

int pcg_set_schedule(uint32_t words, uint32_t group, uint32_t grid_pct, int32_t waves)
{
  return guarded([&] {
    if (words != 0 && (words < 128 || words > 4096 || (words % 128) != 0))
      throw pcg_error("segment words must be 0 or a multiple of 128 in [128, 4096]");
    if (group != 0 && group != 2 && group != 4) throw pcg_error("group must be 0, 2 or 4");
    if (grid_pct > 100) throw pcg_error("grid percentage must be in [0, 100]");
    if (waves != -1 && waves != 0 && waves != 6) throw pcg_error("waves must be -1, 0 or 6");
    auto lock = engine_lock();
    EngineConfig& c = engine_config();
    c.segment_words = words;
    c.group_size = group;
    c.grid_pct = grid_pct;
    c.kernel_waves = waves;
  });
}

int pcg_set_autotune(int mode)
{
  return guarded([&] {
    if (mode < PCG_TUNE_OFF || mode > PCG_TUNE_ON) throw pcg_error("autotune mode must be 0, 1 or 2");
    auto lock = engine_lock();
    engine_config().tune_mode = mode;
  });
}

int pcg_set_schedule_cache(const char* path, uint64_t* entries)
{
  return guarded([&] {
    auto lock = engine_lock();
    tune_set_path(path);
    if (entries) *entries = tune_entries();
  });
}

int pcg_factor_table_check(int kind, int64_t y, int64_t limit, uint64_t* mismatches)
{
  return guarded([&] {
    if (!mismatches) throw pcg_error("mismatches must not be NULL");
    if (kind != PCG_KIND_D && kind != PCG_KIND_S2_HARD) throw pcg_error("unknown table kind");
    *mismatches = factor_table_mismatches(kind == PCG_KIND_D ? LeafSum::D : LeafSum::S2_hard, y,
                                          limit);
  });
}

int pcg_set_shared_segments(int on)
{
  return guarded([&] { engine_config().share_segments = on != 0; });
}

int pcg_set_leaf_batching(int on)
{
  return guarded([&] { engine_config().leaf_batching = on != 0; });
}

int pcg_set_phase_timers(int on)
{
  return guarded([&] { engine_config().phase_timers = on != 0; });
}

int pcg_set_pass_slots(uint64_t slots)
{
  return guarded([&] {
    if (slots != 0 && slots < 64)
      throw pcg_error("pass slots must be 0 or >= 64");
    engine_config().max_slots_per_pass = slots;
  });
}

int pcg_set_pi_table_bytes(uint64_t bytes)
{
  return guarded([&] {
    if (bytes != 0 && bytes < (1u << 20))
      throw pcg_error("prime-count table bytes must be 0 or >= 2^20");
    engine_config().pix_table_bytes = bytes;
  });
}

int pcg_set_slot_bits(uint32_t cbits)
{
  return guarded([&] {
    if (cbits != 0 && (cbits < 14 || cbits > 28))
      throw pcg_error("slot bits must be 0 or in [14, 28]");
    engine_config().slot_cbits = cbits;
  });
}

int pcg_set_host_threads(int threads)
{
  return guarded([&] {
    if (threads < 0) throw pcg_error("threads must be >= 0");
    engine_config().host_threads = threads;
  });
}

int pcg_last_stats(pcg_stats* out)
{
  return guarded([&] {
    if (!out) throw pcg_error("out must not be NULL");
    *out = t_stats;
  });
}

int pcg_release(void)
{
  return guarded([&] { release_device(); });
}

} // extern "C"
