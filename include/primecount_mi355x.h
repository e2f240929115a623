/*
 * primecount_mi355x.h -- C ABI of the MI355X hard-special-leaf engine.
 *
 * The engine computes Gourdon's D(x, y, z, k) and the Deleglise-Rivat
 * S2_hard(x, y, z, c) on one AMD Instinct MI355X (gfx950) with hand-written
 * HIP kernels.  Results are bit-exact with primecount's CPU implementation.
 *
 * Conventions
 *   - Every function returns PCG_OK (0) or a negative PCG_E_* status; on error
 *     pcg_last_error() returns a message (thread-local, valid until the next
 *     call on the same thread).  This mirrors the reference C API, which
 *     catches primecount_error and returns -1 (src/api_c.cpp:19-56).
 *   - 128-bit integers cross the ABI as pcg_i128 {lo, hi} (two's complement,
 *     hi holds bits 64..127), because C has no portable 128-bit type.
 *   - Callers pass scalars and caller-owned output pointers only; the library
 *     owns every device buffer (tables, work queues), as the reference's D()
 *     owns its FactorTableD/PiTable (src/gourdon/D.cpp:443-446).
 */
#ifndef PRIMECOUNT_MI355X_H
#define PRIMECOUNT_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pcg_i128 {
  uint64_t lo;
  int64_t hi;
} pcg_i128;

enum {
  PCG_OK = 0,
  PCG_E_INVALID = -1,      /* argument outside the supported domain       */
  PCG_E_DEVICE = -2,       /* HIP runtime / device error                  */
  PCG_E_NOMEM = -3,        /* device or host allocation failed            */
  PCG_E_INTERNAL = -4      /* any other failure                           */
};

/* Message of the last failed call on this thread ("" if none). */
const char* pcg_last_error(void);

/* Library version string, e.g. "0.1.0 (gfx950)". */
const char* pcg_version(void);

/* 1 if a gfx950 device is visible and the kernels loaded, else 0. */
int pcg_device_available(void);

/* ---- hot path: hard special leaves ---------------------------------- */

/* Gourdon's D(x, y, z, k) for x <= 2^63 - 1.
 * Replaces: int64_t D(int64_t x, int64_t y, int64_t z, int64_t k,
 *                     int threads, bool is_print)
 *           include/gourdon.hpp:28 (src/gourdon/D.cpp:425-452). */
int pcg_D_i64(int64_t x, int64_t y, int64_t z, int64_t k, int64_t* out);

/* Gourdon's D(x, y, z, k) for 128-bit x (x <= y * 2^62).
 * Replaces: int128_t D(int128_t x, int64_t y, int64_t z, int64_t k,
 *                      int threads, bool is_print)
 *           include/gourdon.hpp:38 (src/gourdon/D.cpp:456-505). */
int pcg_D_i128(pcg_i128 x, int64_t y, int64_t z, int64_t k, pcg_i128* out);

/* Deleglise-Rivat S2_hard(x, y, z, c) for x <= 2^63 - 1.
 * Replaces: int64_t S2_hard(int64_t x, int64_t y, int64_t z, int64_t c,
 *                           int threads, bool is_print)
 *           include/S.hpp:25 (src/deleglise-rivat/S2_hard.cpp:221-248). */
int pcg_S2_hard_i64(int64_t x, int64_t y, int64_t z, int64_t c, int64_t* out);

/* Deleglise-Rivat S2_hard(x, y, z, c) for 128-bit x.
 * Replaces: int128_t S2_hard(int128_t x, ...)  include/S.hpp:32
 *           (src/deleglise-rivat/S2_hard.cpp:252-293). */
int pcg_S2_hard_i128(pcg_i128 x, int64_t y, int64_t z, int64_t c, pcg_i128* out);

/* Gourdon's D(x, y, z, k) for x <= 2^63 - 1, as pcg_D_i64, without
 * allocating: no hipMalloc, hipHostMalloc, hipFree or other allocating
 * device routine.  Every table and work buffer of the call lives in a
 * static device array of 64 MiB per GPU (held until the process ends), one
 * kernel pass over all segments; a query that needs more (D(x) with the
 * default parameters fits up to about x = 1e15) fails with PCG_E_INVALID.
 * The array keeps the last call's tables (rebuilt when a query needs larger
 * ones); calls are serialised by the library's lock, so they never overlap
 * on one GPU within a process.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (HIP graph capture, tight
 * serving loops, harnesses that forbid allocation). */
int pcg_D_i64_noalloc(int64_t x, int64_t y, int64_t z, int64_t k, int64_t* out);

/* D(x) with primecount's default (y, z, k): the CLI's `primecount x --D`.
 * Replaces: maxint_t D(maxint_t x, int threads)  src/app/main.cpp:205-246. */
int pcg_D(pcg_i128 x, pcg_i128* out);

/* S2_hard(x) with primecount's default (y, z, c): `primecount x --S2-hard`.
 * Replaces: maxint_t S2_hard(maxint_t x, int threads)  src/app/main.cpp:435-461. */
int pcg_S2_hard(pcg_i128 x, pcg_i128* out);

/* ---- parameter selection (host, exact IEEE double as the reference) -- */

/* x -> (y, z, k, x_star) of Gourdon's algorithm.
 * Restates: src/util.cpp:276-349 (get_alpha_gourdon), 362-386
 * (get_x_star_gourdon), src/app/main.cpp:205-246, src/PhiTiny.hpp:59-63. */
int pcg_gourdon_params(pcg_i128 x, int64_t* y, int64_t* z, int64_t* k, int64_t* x_star);

/* x -> (y, z, c) of the Deleglise-Rivat algorithm.
 * Restates: src/util.cpp:222-263, src/app/main.cpp:435-461. */
int pcg_dr_params(pcg_i128 x, int64_t* y, int64_t* z, int64_t* c);

/* Tuning-factor overrides; values < 1 restore the default formula.
 * Replace: set_alpha_y / set_alpha_z / set_alpha  src/util.cpp:103-133. */
int pcg_set_alpha_y(double alpha_y);
int pcg_set_alpha_z(double alpha_z);
int pcg_set_alpha(double alpha);

/* ---- batched queries: one persistent launch over a device work queue -- */

/* D(x_i) with default parameters for n values of x.  All queries share one
 * persistent kernel launch; results land in out[i].  Replaces n calls of
 * maxint_t D(maxint_t x, int threads) (src/app/main.cpp:205-246). */
int pcg_D_batch(const pcg_i128* xs, size_t n, pcg_i128* out);

/* ---- engine configuration and statistics ----------------------------- */

/* Segment size in 64-bit sieve words (240 numbers each); a multiple of 128
 * in [128, 4096].  0 selects the schedule (the autotuner's cached winner or
 * the segment-size rule).  Segment size never changes results (replaces the
 * segment sizing of src/LoadBalancerS2.cpp:65-110). */
int pcg_set_segment_words(uint32_t words);

/* Schedule autotuner of the sieve launches (segment size, queries per shared
 * segment, persistent grid; none of them changes results).  Launches are
 * classed by device, kind (D / S2_hard), factor-table width, the sum of
 * their sieve limits (half-octaves) and their number of queries (octaves).
 *   PCG_TUNE_OFF    always the segment-size rule;
 *   PCG_TUNE_CACHED (default) a cached winner of the class when there is
 *                   one, else the rule;
 *   PCG_TUNE_ON     as CACHED, and a class not yet cached is tuned on the
 *                   launch itself: the rule's schedule first (its sums are
 *                   the answer), then a few candidates around it, every one
 *                   checked to give the same sums; the fastest is cached.
 * Launches whose sieve limits sum to less than 2^30 use the rule.  The
 * initial mode can also be set with the environment variable PCG_AUTOTUNE
 * (0, 1, 2).  No counterpart in the reference (src/LoadBalancerS2.cpp sizes
 * segments by rules). */
/* Pins the schedule of the following launches: segment words (as
 * pcg_set_segment_words), queries per shared segment (2 or 4), the persistent
 * grid as a percentage (1..100) of the waves occupancy allows, and the
 * single-query kernel build (0 unconstrained, 6 built for 6 waves per SIMD).
 * 0 (waves: -1) leaves a knob to the schedule; any knob set turns the
 * autotuner off.  Never changes results. */
int pcg_set_schedule(uint32_t words, uint32_t group, uint32_t grid_pct, int32_t waves);

#define PCG_TUNE_OFF 0
#define PCG_TUNE_CACHED 1
#define PCG_TUNE_ON 2
int pcg_set_autotune(int mode);

/* The autotuner's cache file: path, or NULL / "" for the default
 * ($PCG_TUNE_CACHE, else $XDG_CACHE_HOME/primecount_mi355x/schedules.txt,
 * else $HOME/.cache/primecount_mi355x/schedules.txt).  Winners are appended
 * as text lines keyed by device and class; a file that cannot be written
 * keeps them in memory only.  Forgets the winners in memory; the file is
 * read on next use.  *entries (if not NULL) receives the number of winners
 * the file holds for any device. */
int pcg_set_schedule_cache(const char* path, uint64_t* entries);

/* Whether the queries of one launch (pcg_D_batch, serving, pi(x) batches)
 * share the sieve of the segments they have in common: each segment is
 * sieved once for up to 4 queries that cover it in full, and each answers its
 * own leaves.  1 (default) on, 0 off.  Never changes results. */
int pcg_set_shared_segments(int on);

/* Self-check of the device factor-table builder: builds the FactorTableD
 * (kind PCG_KIND_D, entries up to limit with prime factors <= y) or the lpf
 * FactorTable (PCG_KIND_S2_HARD) on the device and on the host and stores the
 * number of differing entries in *mismatches.  No counterpart in the
 * reference (src/gourdon/FactorTableD.hpp, src/FactorTable.hpp build on the
 * host); limit < 2^32. */
int pcg_factor_table_check(int kind, int64_t y, int64_t limit, uint64_t* mismatches);

/* (segment, prime) slots per kernel pass (0 = default, 2^31).  C and M take
 * 4 bytes of HBM per slot; larger queries run in several passes whose carry is
 * kept on the device.  The pass size never changes results.  No counterpart in
 * the reference (its threads carry phi through their own segments). */
int pcg_set_pass_slots(uint64_t slots);

/* Largest prime-count table (bytes; 16 bytes per 240 numbers) that pi(x)
 * keeps whole in device memory (0 = default, 65% of the device memory).  A
 * pi(x) or Deleglise-Rivat pi(x) whose table would be larger builds it in
 * windows of this size, one after the other, and sums the leaves of A, C, B
 * (and S2_easy, P2) window by window: device memory stays bounded, the
 * results do not change.  Replaces the role of the reference's
 * SegmentedPiTable (src/gourdon/SegmentedPiTable.cpp), whose x^(1/4)-sized
 * segments bound its memory the same way. */
int pcg_set_pi_table_bytes(uint64_t bytes);

/* Bits of C in the 32-bit slot word (C | M << bits).  0 (default) takes the
 * fewest the segment size allows (C <= 64 words); more bits leave fewer for
 * M, so that more slots escape to the per-pass escape table.  14..28.  Never
 * changes results; it exists to exercise the escape path.  No counterpart
 * in the reference. */
int pcg_set_slot_bits(uint32_t cbits);

/* Batched leaf counting (default on): in the part of a window of 64 sieving
 * primes where cross-off order does not matter, the leaves of several primes
 * are counted against one word-pair prefix build, each corrected by the
 * recorded first-time removals of the primes between it and the build.
 * Never changes results; it exists to compare the two paths.  No
 * counterpart in the reference (its leaves are counted one prime at a time,
 * src/gourdon/D.cpp:122-273). */
int pcg_set_leaf_batching(int on);

/* Phase timers (default off): single-query launches run a build of the sieve
 * kernel that reads the clock at every phase change and reports the cycles
 * per phase in pcg_stats.phase_cycles.  Same results; the timers slow the
 * kernel down.  A profiling aid with no counterpart in the reference. */
int pcg_set_phase_timers(int on);

/* Host threads used by the table builders (0 = all hardware threads). */
int pcg_set_host_threads(int threads);

typedef struct pcg_stats {
  double seconds_total;      /* wall time of the last D/S2_hard call          */
  double seconds_tables;     /* host table build + upload (0 when resident)   */
  double seconds_plan;       /* host planning + per-call uploads              */
  double seconds_kernel;     /* sieve kernel (HIP events)                     */
  double seconds_combine;    /* carry scan + combine kernels (HIP events)     */
  uint64_t segments;         /* sieve segments processed                      */
  uint64_t chunks;           /* work-queue items                              */
  uint64_t segment_words;    /* 64-bit words per segment                      */
  uint64_t sieve_limit;      /* x/z (D) or z (S2_hard)                        */
  uint64_t b_first, b_last;  /* sieving-prime index range                     */
  uint64_t leaves;           /* hard special leaves summed                    */
  uint64_t hits;             /* cross-off hits (prime multiples visited)      */
  uint64_t factor_entries;   /* factor-table entries scanned                  */
  uint64_t pairs;            /* (segment, sieving prime) steps                */
  /* wave-level work of the sieve kernel */
  uint64_t windows;          /* windows of up to 64 sieving primes            */
  uint64_t leaf_primes;      /* (segment, prime) pairs with factor entries    */
  uint64_t leaf_steps;       /* leaf-loop steps (64 entries each)             */
  uint64_t dense_primes;     /* primes crossed off one at a time              */
  uint64_t dense_steps;      /* their cross-off steps                         */
  uint64_t group_steps;      /* grouped cross-off steps                       */
  uint64_t prefix_builds;    /* word-pair prefix rebuilds (a prime with leaves
                                after first-time removals)                    */
  uint64_t queue_steps;      /* steps counting queued factor-table leaves
                                (up to 64 leaves each)                        */
  uint64_t device_allocations; /* device and pinned host allocations the library
                                has made in this process so far (at the end of
                                the call)                                     */
  uint64_t escaped_slots;    /* slots whose M did not fit the 32-bit slot word
                                and went to the escape table                  */
  uint64_t escape_bound;     /* the host's upper bound of those slots         */
  uint64_t schedule_source;  /* 0 the segment-size rule (or pcg_set_segment_words),
                                1 a cached autotuner winner, 2 tuned by this call */
  uint64_t group_size;       /* queries sieved together per shared segment     */
  uint64_t grid_waves;       /* waves of the persistent sieve grid (at most)   */
  uint64_t kernel_waves;     /* 6: the sieve kernel built for 6 waves per SIMD,
                                0: the unconstrained build                    */
  uint64_t leaf_batches;     /* batches of primes whose leaves were counted on
                                one prefix build (pcg_set_leaf_batching)      */
  uint64_t batch_records;    /* first-time removals recorded for the batches'
                                corrections                                   */
  uint64_t phase_cycles[11]; /* with pcg_set_phase_timers(1): the sieve kernel's
                                clock cycles per phase, summed over its waves:
                                segment set-up and end, window set-up and end,
                                one-at-a-time cross-offs, grouped cross-offs,
                                batches' recording cross-offs, pair-prefix
                                builds, pre-scan of primes with factor entries,
                                factor-entry steps, counting steps, batches'
                                entry steps, batches' counting steps (else 0) */
} pcg_stats;

/* Statistics of the last D/S2_hard call on this thread. */
int pcg_last_stats(pcg_stats* out);

/* LDS read bandwidth of the current device in GB/s, measured by a probe
 * kernel (every CU reads its LDS with bank-conflict-free 128-bit loads): the
 * peak of the sieve kernel's LDS roofline.  No counterpart in the
 * reference. */
int pcg_lds_read_gbs(double* gbs);

/* Release every device buffer the library holds, including the prime and
 * factor tables that stay resident in HBM between calls. */
int pcg_release(void);

/* ---- the prime counting function ---------------------------------------
 *
 * pi(x) by Xavier Gourdon's algorithm, pi(x) = A - B + C + D + Phi0 + Sigma
 * (src/gourdon/pi_gourdon.cpp:82-87), with the reference's default
 * parameters; every term runs on the device (D on the hard-special-leaf
 * engine, the others over a prime-count table resident in HBM).  Replaces
 * pi(x) / primecount_pi (include/primecount.hpp:57,74, include/primecount.h:
 * 51, 69 -> src/gourdon/pi_gourdon.cpp:22-155).  x <= 10^8 is answered from
 * the prime-count table directly.  Domain: x / y < 2^49 (the prime-count
 * table covers [0, x / y] in HBM: 7.5 GB at x = 1e19, 98 GB at 1e21). */
int pcg_pi(pcg_i128 x, pcg_i128* out);

/* The other Gourdon terms with explicit parameters, as in
 * include/gourdon.hpp:24-27 (64-bit) and 34-37 (128-bit): A + C
 * (src/gourdon/AC.cpp), B (src/gourdon/B.cpp), Phi0 (src/gourdon/Phi0.cpp,
 * device path for 5 <= k <= 8) and Sigma (src/gourdon/Sigma.cpp). */
int pcg_AC(pcg_i128 x, int64_t y, int64_t z, int64_t k, pcg_i128* out);
int pcg_B(pcg_i128 x, int64_t y, pcg_i128* out);
int pcg_Phi0(pcg_i128 x, int64_t y, int64_t z, int64_t k, pcg_i128* out);
int pcg_Sigma(pcg_i128 x, int64_t y, pcg_i128* out);

/* pi(x) by the Deleglise-Rivat algorithm, pi(x) = S1 + S2_trivial + S2_easy
 * + S2_hard + pi(y) - 1 - P2, with the reference's parameters (y, z, c of
 * pcg_dr_params); every term on the device (S2_hard on the hard-special-leaf
 * engine, S1 from its lpf factor table, S2_easy and P2 over the prime-count
 * table; S2_trivial is a closed form over pi(n) lookups).  Replaces
 * pi_deleglise_rivat (include/primecount-internal.hpp:33, 59 ->
 * src/deleglise-rivat/pi_deleglise_rivat.cpp:40-140).  Domain as pcg_pi. */
int pcg_pi_deleglise_rivat(pcg_i128 x, pcg_i128* out);

/* The Deleglise-Rivat terms with explicit parameters: S1, S2_trivial,
 * S2_easy (include/S.hpp:22-24, 29-31; src/S1.cpp,
 * src/deleglise-rivat/S2_trivial.cpp, S2_easy.cpp; S1 needs 5 <= c <= 8) and
 * P2 (include/primecount-internal.hpp:43, 61; src/P2.cpp; pi(y) is computed
 * here).  Domain: x <= y * 2^62, y < 2^32, z < 2^50, x / y < 2^49. */
int pcg_S1(pcg_i128 x, int64_t y, int64_t c, pcg_i128* out);
int pcg_S2_trivial(pcg_i128 x, int64_t y, int64_t z, int64_t c, pcg_i128* out);
int pcg_S2_easy(pcg_i128 x, int64_t y, int64_t z, int64_t c, pcg_i128* out);
int pcg_P2(pcg_i128 x, int64_t y, pcg_i128* out);

/* The n-th prime, 1 <= n <= 216289611853439384 (the primes below 2^63):
 * pi(x0) on the device at x0 = R^-1(n) (Riemann's R), then a device sieve of
 * the interval between x0 and the answer.  Replaces nth_prime /
 * primecount_nth_prime (include/primecount.hpp:84, include/primecount.h:81
 * -> src/nth_prime.cpp:95-145). */
int pcg_nth_prime(int64_t n, int64_t* out);

/* The n-th prime for n beyond the 64-bit range of answers (any n >= 1; up to
 * n = 216289611853439384 it is pcg_nth_prime): pi at x0 = R^-1(n) on the
 * device, then a sieve of the interval between x0 and the answer anchored at
 * a 128-bit base, with 64-bit sieving primes.  Domain: the answer's x0
 * inside pcg_pi's domain (y < 2^32: up to about 1e23).
 * Replaces: pc_int128_t primecount_nth_prime_128(pc_int128_t n)
 *           include/primecount.h:99, src/nth_prime.cpp:141-175. */
int pcg_nth_prime_128(pcg_i128 n, pcg_i128* out);

/* Plausibility check of a pi(x) value: PCG_OK if |pix - Li(x)| <
 * sqrt(x) log(x) / (8 pi) (Schoenfeld's bound; always for x < 2657), else
 * PCG_E_INTERNAL with a message.  pcg_pi and pcg_pi_deleglise_rivat run it on
 * every result and fail with PCG_E_INTERNAL instead of returning a value
 * outside the bound.  Host arithmetic only (no device needed).
 * Restates: verify_pix  src/util.cpp:397-420 (called at
 * src/gourdon/pi_gourdon.cpp:89,160), which aborts instead. */
int pcg_check_pi(pcg_i128 x, pcg_i128 pix);

typedef struct pcg_pi_stats {
  double seconds_total;      /* wall time of the last pcg_pi call             */
  double seconds_table;      /* tables, including the prime-count table       */
  double seconds_sigma, seconds_phi0, seconds_ac, seconds_b, seconds_d;
                             /* per term: device time between stream events
                                (the terms run back to back; A/C includes the
                                descriptor upload, D its host planning)       */
  uint64_t ac_leaves;        /* A and C leaves summed                         */
  uint64_t table_limit;      /* the prime-count table covers [0, table_limit] */
  uint64_t table_windows;    /* windows the table was built in (1 = whole)    */
} pcg_pi_stats;

/* Statistics of the last pcg_pi call on this thread. */
int pcg_last_pi_stats(pcg_pi_stats* out);

/* ---- serving: continuous batching ---------------------------------------
 *
 * A context owns a worker thread.  pcg_submit queues queries from any thread
 * and returns at once.  Light queries (pi(x) with x <= 1e16, nth_prime(n) with
 * n <= 2e14, D / S2_hard with x <= 1e16: about 10 ms or less) run as soon as
 * the device reaches a point where other work may start: at once when it is
 * idle, otherwise at the next pass point of the running heavy batch (between
 * the term launches of its queries and between its sieve passes of at most
 * 2^31 slots), with work buffers of their own.  Heavy queries
 * run in batches that share one launch sequence per kind -- D / S2_hard in
 * one persistent launch per table set, as pcg_D_batch; pi(x) and nth_prime
 * with the D parts of all of them in one launch -- taken in arrival order up
 * to about 250 ms of estimated work (half the estimated backlog, up to 4 s,
 * when more than 8 heavy queries wait).  Work is admitted by its estimated
 * device memory (tables + scratch) against the context's budget; a query
 * that can never fit is answered at once with PCG_E_NOMEM.  pcg_poll hands
 * out finished results in completion order.  No counterpart in the
 * reference, whose entry points compute one x per call (D / S2_hard,
 * include/gourdon.hpp:28,38, include/S.hpp:25,32; pi, src/api.cpp:75-91). */

#define PCG_KIND_D 0
#define PCG_KIND_S2_HARD 1
#define PCG_KIND_PI 2          /* pi(x)                                       */
#define PCG_KIND_NTH_PRIME 3   /* the x-th prime (x < 2^63)                   */

typedef struct pcg_query {
  int32_t kind;        /* PCG_KIND_*                                           */
  int32_t reserved;    /* 0                                                    */
  pcg_i128 x;          /* x, or n for PCG_KIND_NTH_PRIME                      */
  int64_t y, z, k;     /* D / S2_hard: all 0 = the reference's default
                          parameters for x; must be 0 for pi and nth_prime    */
} pcg_query;

typedef struct pcg_result {
  uint64_t id;         /* as returned by pcg_submit                            */
  int32_t status;      /* PCG_OK or a negative PCG_E_* code                    */
  int32_t reserved;
  pcg_i128 value;      /* the answer when status == PCG_OK                     */
  double seconds;      /* submit to result available                           */
} pcg_result;

typedef struct pcg_ctx pcg_ctx;

/* Creates a serving context.  y_max, z_max > 0 build the D tables for default
 * queries up to that size now (else on first use); hbm_budget bounds the
 * device memory the context's work may take, tables and scratch, in bytes
 * (0 = 90% of the device's memory); max_batch bounds the queries of one
 * launch (0 = 1024). */
int pcg_ctx_create(int64_t y_max, int64_t z_max, uint64_t hbm_budget, uint32_t max_batch,
                   pcg_ctx** ctx);

/* Queues n queries; ids[i] receives the id of q[i].  A query outside the
 * engine's domain is answered at once with PCG_E_INVALID, one whose
 * estimated device memory exceeds the context's budget with PCG_E_NOMEM. */
int pcg_submit(pcg_ctx* ctx, const pcg_query* q, size_t n, uint64_t* ids);

/* Moves up to max finished results into out; waits up to timeout_ms (< 0:
 * without limit) for the first one.  *n receives the number written. */
int pcg_poll(pcg_ctx* ctx, pcg_result* out, size_t max, int timeout_ms, size_t* n);

typedef struct pcg_ctx_stats {
  uint64_t submitted, completed, batches, max_batch;
  double busy_seconds;  /* worker time inside launches                         */
  uint64_t light_batches, heavy_batches;
  uint64_t interleaved;     /* light queries answered at pass points of a heavy
                               batch                                          */
  uint64_t rejected_nomem;  /* queries refused for their device memory        */
} pcg_ctx_stats;

int pcg_ctx_get_stats(pcg_ctx* ctx, pcg_ctx_stats* out);

/* Finishes every queued query, stops the worker and frees the context
 * (results not yet polled are dropped). */
int pcg_ctx_destroy(pcg_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* PRIMECOUNT_MI355X_H */
