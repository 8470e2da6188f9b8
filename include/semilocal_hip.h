/*
 * semilocal_hip.h -- C ABI of the MI355X semi-local LCS engine.
 *
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.
 * Every entry point returns an int status (SL_OK = 0) and leaves a message for
 * sl_last_error() on failure; nothing throws across the ABI.  Calls on
 * different handles may run concurrently from different threads: each handle
 * owns its HIP stream and device buffers, and every one-shot entry point
 * builds, uses and frees a private handle.  Calls on one handle must not
 * overlap.
 *
 * Result convention (identical to the reference's AbstractPermutation /
 * common::Permutation filled by the combing solvers): for strings a (length
 * m, rows) and b (length n, columns), `row_to_col[s]` is the exit position of
 * strand s among the m+n strands: positions 0..n-1 are bottom-edge exits at
 * column j, positions n..n+m-1 are right-edge exits at row m-1-(pos-n).
 * `col_to_row` is its inverse.  LLCS(a, b) = m - #{s < m : row_to_col[s] >= n}.
 */
#ifndef SEMILOCAL_HIP_H
#define SEMILOCAL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SL_OK 0
#define SL_ERR_ARG 1
#define SL_ERR_HIP 2
#define SL_ERR_NOMEM 3
#define SL_ERR_TIMEOUT 4
#define SL_ERR_UNSUPPORTED 5
#define SL_PENDING 6          /* sl_server_result: the request is still queued or running */

/* Symbol element width accepted by the entry points that take `sym_bytes`:
 * 1 = uint8_t symbols, 2 = uint16_t, 4 = int32_t (remapped internally to a
 * dense alphabet; at most 65535 distinct symbols per call). */

/* Library version (major*10000 + minor*100 + patch). */
int sl_version(void);

/* Message of the last failure on the calling thread ("" if none). */
const char *sl_last_error(void);

/* Selects the HIP device used by later calls on this thread and creates the
 * library's streams.  Optional: every entry point initialises device 0 (or the
 * current device) lazily. */
int sl_init(int device);

/* Writes "name;compute_units;clock_khz;hbm_bytes;arch" into buf (name is the
 * device's marketing name, or its ISA name where that is empty; arch is the
 * ISA name with its features, e.g. "gfx950:sramecc+:xnack-"). */
int sl_device_info(char *buf, int len);

/* Diagnostics (no reference counterpart): one plain device-to-device copy of
 * `bytes` (a multiple of 16) by a streaming kernel, timed with HIP events.
 * Used to scale the HBM byte counters of the profiler (bytes read = bytes
 * written = `bytes`). */
int sl_probe_copy(int64_t bytes, double *ms);

/* ------------------------------------------------------------------------ */
/* Plain semi-local LCS of one pair (anti-diagonal sticky-braid combing).     */
/* Replaces: semi_local::sticky_braid_mpi<Input,WithIf,WithWait>(matrix, a,   */
/*   a_size, b, b_size, threads_num, is_reverse)  -- semi_local.h:212-214;    */
/*   semi_local::sticky_braid_sequential -- semi_local.h:135-137;             */
/*   semi_local::OpenMPIterativeCombing::compute -- SemiLocalLCS.h:238-244;   */
/*   semi_local::SimpleIterativeCombing::compute -- SemiLocalLCS.h:68.         */
/* Host buffers; col_to_row may be NULL.  m or n may be 0.                    */
/* ------------------------------------------------------------------------ */
int sl_semi_local_comb(const void *a, int m, const void *b, int n, int sym_bytes,
                       int32_t *row_to_col, int32_t *col_to_row);

/* Allocation-free variant of sl_semi_local_comb.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Inputs and outputs are DEVICE pointers owned by the caller: d_a (m symbols),
 * d_b (n symbols), d_row_to_col (m+n entries) and optionally d_col_to_row
 * (m+n entries, or NULL).  The work is enqueued on `stream` (a hipStream_t;
 * NULL = the null stream) and the call waits for it, then reports a
 * kernel-side timeout or defect like sl_semi_local_comb.  No hipMalloc, no
 * hipFree, no allocating library routine: the per-call state (launch control
 * block, pair descriptor, pass queue and inter-pass strand buffer) lives in a
 * fixed device array of about 8.4 MB that each GPU holds until the process
 * ends, rewritten at the start of every call.  Calls on one device share that
 * array and are serialised by a host lock.  Limits: max(m, n) <=
 * SL_NOALLOC_MAX_LEN (longer strings: SL_ERR_ARG), sym_bytes 1 or 2 (symbols
 * are compared as given, with no remapping).  Results equal
 * sl_semi_local_comb's on the same strings. */
#define SL_NOALLOC_MAX_LEN 1048576
int sl_semi_local_comb_noalloc(const void *d_a, int m, const void *d_b, int n, int sym_bytes,
                               int32_t *d_row_to_col, int32_t *d_col_to_row, void *stream);

/* LLCS read from a kernel: replaces lcs::LCSBySemiLocal::compute's count,
 * include/LCS/LCS.h:223-233.  Pure host arithmetic over row_to_col. */
int sl_llcs_from_kernel(const int32_t *row_to_col, int m, int n);

/* ------------------------------------------------------------------------ */
/* Ragged batch: thousands of independent pairs combed in ONE persistent      */
/* launch fed by a device queue of (pair, row-pass) tiles.  No reference      */
/* counterpart (the reference combs one pair per call); the per-pair results */
/* are exactly those of sl_semi_local_comb on each pair.                      */
/* Inputs are concatenated: pair i's a is a_cat[a_off[i] .. a_off[i]+a_len[i])*/
/* in elements of sym_bytes.  Outputs are concatenated in caller order: pair  */
/* i's row_to_col occupies (a_len[i]+b_len[i]) entries starting at the sum of */
/* the previous pairs' sizes.                                                 */
/* ------------------------------------------------------------------------ */
typedef struct sl_batch sl_batch;

#define SL_BATCH_WITH_COL_TO_ROW 1   /* also produce col_to_row */

/* Uploads the batch to HBM and prepares the queue.  `rows_per_lane` = 0
 * gives every pair its own tile height from a cost model (see
 * sl_batch_set_tiles); 4, 8 or 16 forces one for all pairs. */
int sl_batch_create(sl_batch **out, int npairs, const void *a_cat, const int64_t *a_off,
                    const int32_t *a_len, const void *b_cat, const int64_t *b_off,
                    const int32_t *b_len, int sym_bytes, int flags, int rows_per_lane);
/* Combs every pair in one persistent launch (asynchronous on the batch's stream). */
int sl_batch_run(sl_batch *bt);
/* Combs the same pairs with one launch per pair, in caller order (the plain
 * one-pair-per-launch form used as the comparison point). */
int sl_batch_run_per_pair(sl_batch *bt);
/* Waits for the batch's stream and reports a kernel-side timeout, if any. */
int sl_batch_sync(sl_batch *bt);
/* Copies results to host (after sync).  col_to_row_cat may be NULL. */
int sl_batch_download(sl_batch *bt, int32_t *row_to_col_cat, int32_t *col_to_row_cat);
/* Times `iters` runs after `warmup` runs with HIP events on the batch's
 * stream.  mode 0 = persistent batched launch, 1 = one launch per pair.
 * kernel_ms_out (may be NULL) receives the combing kernels' own time per run
 * (events around the kernel launches only). */
int sl_batch_time(sl_batch *bt, int mode, int warmup, int iters, double *ms_per_run,
                  double *kernel_ms_out);
/* Tile shape of the batch: rows per lane R (one queue tile = one pass of 64R
 * rows of one pair) and persistent waves per compute unit (default 8).
 * R = 4, 8 or 16 puts every pair at that height.  R = 0 (the default) gives
 * each pair its own, in the one launch: the height with the fewest
 * instructions for its padded rows (ceil(m / 64R) passes of n + 63 steps);
 * a pair whose chain of passes would outlast the critical fraction of the
 * batch's modelled time (sl_batch_set_critical_fraction) becomes a critical
 * chain instead -- the height with the shortest chain, a short queue lag and
 * raised wave priority; a batch too small to give every SIMD two waves at
 * R = 16 takes one height for all from a latency model.  sl_batch_pair_tiles
 * reports the choice.  sl_batch_get_tiles reports the policy (0 = per pair).
 * Results never depend on any of it (the semi-local kernel is unique). */
int sl_batch_set_tiles(sl_batch *bt, int rows_per_lane, int waves_per_cu);
int sl_batch_get_tiles(const sl_batch *bt, int *rows_per_lane, int *waves_per_cu);
/* Per-pair policy: critical chains are the pairs whose modelled chain of
 * passes exceeds `fraction` x the batch's modelled throughput time (default
 * 0.6; 0 = none).  Re-plans the queue. */
int sl_batch_set_critical_fraction(sl_batch *bt, double fraction);
int sl_batch_get_critical_fraction(const sl_batch *bt, double *fraction);
/* Tile height (4, 8 or 16; 0 for a pair with an empty string) and priority
 * flag (1 = critical chain) of every pair, in caller order; either array may
 * be NULL. */
int sl_batch_pair_tiles(const sl_batch *bt, int32_t *rows_per_lane, int32_t *priority);
/* Queue order of the batched launch: 0 = pair by pair, largest pairs first;
 * s = 1 .. 8 = a list-scheduling simulation that hands out a pair's passes
 * about when each can start behind the previous one, assuming a pass starts
 * 128 s steps after its predecessor (256 for a critical chain; default 2).
 * Results never depend on it. */
int sl_batch_set_schedule(sl_batch *bt, int schedule);
int sl_batch_get_schedule(const sl_batch *bt, int *schedule);
/* Autotuner: times every tile shape -- rows per lane {4, 8, 16, and per pair
 * with critical fraction 0, 0.4, 0.6, 0.8} x persistent waves per CU
 * {4, 8, 12, 16} x queue order {0, 2, 3, 4, 6} -- on this batch (`iters` runs
 * each, HIP events), re-times the four fastest warm in 7 interleaved rounds
 * (per round and shape one warm-up run, then max(3, iters) timed runs) and
 * keeps the one with the fastest median round; reports the shape through the
 * out-pointers (0 rows per lane = per pair; the order through
 * sl_batch_get_schedule, the fraction through
 * sl_batch_get_critical_fraction). */
int sl_batch_autotune(sl_batch *bt, int iters, int *rows_per_lane, int *waves_per_cu);
/* Sum of m*n over the batch (cells), number of queue tiles, number of pairs. */
int sl_batch_stats(const sl_batch *bt, int64_t *cells, int64_t *tiles, int64_t *pairs);
/* Pairs (with m, n > 0) per symbol test of the combing cell at the batch's
 * current tile height: symbol codes (4 VALU ops per cell), one-hot tags above
 * the strand-id bits (3 ops; needs bits(m+n-1) + sigma <= 32), or a per-lane
 * LDS table of row-match masks (3 ops and one LDS read per step; needs
 * sigma <= 32 and bits(m+n-1) <= 32 - R).  A launch with fewer waves than
 * CUs (each launch of sl_batch_run_per_pair, a small batch) gives table pairs
 * the code cell instead: one LDS read fewer per step, lower latency when few
 * waves share a CU.  Results never depend on it. */
int sl_batch_cell_modes(const sl_batch *bt, int64_t *code_pairs, int64_t *onehot_pairs, int64_t *table_pairs);
/* Host-only planning, no device needed: the queue a batch of pairs with
 * these string lengths gets at tile policy `rows_per_lane` (0 = per pair, as
 * in sl_batch_set_tiles, with the default critical fraction), `waves_per_cu`
 * waves on `cus` compute units and queue order `schedule` (as in
 * sl_batch_set_schedule).  Writes the (pair, pass) tickets in launch order
 * (pair_out / pass_out may be NULL to ask for *nitems only; otherwise they
 * hold *nitems entries: one per pass of 64 R rows of every pair with both
 * lengths > 0, R its tile height). */
int sl_plan_queue(int npairs, const int32_t *a_len, const int32_t *b_len, int rows_per_lane, int waves_per_cu,
                  int cus, int schedule, int32_t *pair_out, int32_t *pass_out, int64_t *nitems);
void sl_batch_destroy(sl_batch *bt);

/* One-shot convenience: create + run + sync + download + destroy. */
int sl_comb_batch(int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                  const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes,
                  int32_t *row_to_col_cat, int32_t *col_to_row_cat);

/* ------------------------------------------------------------------------ */
/* Bit-parallel LLCS of two binary strings (1-bit strands, formula 2).        */
/* Replaces: prefix_lcs_via_semi_local::binary::llcs_2symbol_smart_combing<W>  */
/*   (W *a_reverse, int a_size, W *b, int b_size, int a_total_symbols,        */
/*    int threads_num, bool formula_one) -> int                                */
/*   -- prefix_lcs/bitwise/transposition_network_binary_alphabet.h:598-599.    */
/* Same packing: a reversed, MSB-first uint64 words (encode_reverse,          */
/* encoders_and_decoders.h:100-139); b LSB-first (encode, :50-87).  Like the  */
/* reference, lengths are expected to be multiples of 64.  threads_num and    */
/* formula_one are accepted and ignored (the score is the same).              */
/* ------------------------------------------------------------------------ */
int sl_llcs_bitparallel(const uint64_t *a_rev, int a_words, const uint64_t *b, int b_words,
                        int a_total_symbols, int threads_num, int formula_one, int *llcs);

/* Resident form (inputs uploaded once; run/time repeatedly). */
typedef struct sl_bitpar sl_bitpar;
int sl_bitpar_create(sl_bitpar **out, const uint64_t *a_rev, int a_words, const uint64_t *b,
                     int b_words, int a_total_symbols);
int sl_bitpar_run(sl_bitpar *plan);                     /* asynchronous */
int sl_bitpar_result(sl_bitpar *plan, int *llcs);      /* waits, then reads the score */
int sl_bitpar_time(sl_bitpar *plan, int warmup, int iters, double *ms_per_run);
void sl_bitpar_destroy(sl_bitpar *plan);
/* The kernel a plan runs (same score either way).  No counterpart in the     */
/* reference, whose one routine is the combing:                              */
/*   SL_BITPAR_CARRY   -- the Hyyro kernel (sl_hyyro_*) on the two full       */
/*                        strings, padding included: one step per symbol of   */
/*                        the shorter, the longer bit-packed.  The default    */
/*                        where the longer string has at most 2^20 symbols;   */
/*   SL_BITPAR_COMBING -- sliding-diagonal combing by formula 2 (any size).   */
/* result() reads the form of the last run.                                  */
#define SL_BITPAR_CARRY 0
#define SL_BITPAR_COMBING 1
int sl_bitpar_set_form(sl_bitpar *plan, int form);   /* SL_ERR_ARG if the form is not available */
int sl_bitpar_get_form(sl_bitpar *plan, int *form);

/* ------------------------------------------------------------------------ */
/* Unit-Monge (sticky braid) multiplication R = P (.) Q of n x n permutations  */
/* (row -> col arrays), level-by-level steady ant on the GPU.                 */
/* Replaces: steady_ant::steady_ant_parallel_wrapper(p, q, product, map,      */
/*   nested_lvls) -- unit_monge_mult/steady_ant.h:811-812 (and the sequential */
/*   steady_ant_optimized_seq, :378-453);                                     */
/*   common::BraidMultiplicationStrategy::multiply(p, q, product) --          */
/*   common/include/Permutations/MongeMultuplicationStrategy.h:18-21.         */
/* Inputs must be full permutations of [0, n) (no NOPOINT entries); anything  */
/* else is rejected with SL_ERR_ARG.  The precalc map of the reference is not */
/* needed (base cases are solved on the device).                              */
/* ------------------------------------------------------------------------ */
int sl_braid_multiply(const int32_t *p, const int32_t *q, int n, int32_t *r);
/* A batch of independent products of sizes `sizes[i]`, concatenated densely:
 * product i reads p_cat/q_cat and writes r_cat at the sum of the previous
 * sizes.  All products of the batch run in the same level launches. */
int sl_braid_multiply_batch(int nprob, const int32_t *p_cat, const int32_t *q_cat,
                            const int32_t *sizes, int32_t *r_cat);
/* Times `iters` device runs of the batch (inputs resident in HBM) after
 * `warmup` runs, with HIP events.  elements_levels (may be NULL) receives the
 * sum over recursion levels of the elements processed (the work unit). */
int sl_braid_multiply_time(int nprob, const int32_t *p_cat, const int32_t *q_cat,
                           const int32_t *sizes, int warmup, int iters, double *ms_per_run,
                           int64_t *elements_levels);

/* Staggered multiplication: glues two adjacent block braids sharing k strands.
 * Replaces: steady_ant::staggered_sticky_multiplication<RowGlue, UseParallel>
 *   (p, q, k, map, product, nested_parall_regions) -- steady_ant.h:609-612;
 *   StaggeredStickyBraidMultiplication::staggeredStickyMultiplication<RowGlue>
 *   -- MongeMultuplicationStrategy.h:679-758.
 * p is pr x pr, q is qr x qr (row -> col); the product is (pr + qr - k)
 * square.  row_glue != 0: vertical neighbours, p combs the earlier rows of a
 * (upper block), q the later ones, k = the shared columns; row_glue == 0:
 * horizontal neighbours, p combs the earlier columns of b, k = the shared
 * rows.  k == pr == qr is rejected with SL_ERR_ARG: the reference prints a
 * message and leaves the product unset there (steady_ant.h:616-619). */
int sl_staggered_multiply(const int32_t *p, int pr, const int32_t *q, int qr, int k, int row_glue,
                          int32_t *product);

/* ------------------------------------------------------------------------ */
/* Hybrid block decomposition: small_m x small_n blocks combed in one batched */
/* launch, then merged level by level with staggered multiplications (all     */
/* merges of a level in one batched multiply).  The permutation equals that   */
/* of sl_semi_local_comb on (a, b).                                            */
/* Replaces: semi_local::hybrid_iterative<Input,WithIf>(perm, a, a_size, b,    */
/*   b_size, map, small_m, small_n, threads_num) -- semi_local.h:577-580;      */
/*   semi_local::OpenMPHybridIterative::compute -- SemiLocalLCS.h:594-794.     */
/* ------------------------------------------------------------------------ */
int sl_hybrid_iterative(const void *a, int m, const void *b, int n, int sym_bytes, int small_m,
                        int small_n, int32_t *row_to_col);

/* Resident form: strings uploaded and every level planned once. */
typedef struct sl_hybrid sl_hybrid;
int sl_hybrid_create(sl_hybrid **out, const void *a, int m, const void *b, int n, int sym_bytes,
                     int small_m, int small_n);
int sl_hybrid_run(sl_hybrid *h);                        /* combing + merges, waits */
int sl_hybrid_download(sl_hybrid *h, int32_t *row_to_col);
/* Times `iters` runs after `warmup` runs (HIP events): total, combing part
 * and merge part per run (comb_ms / merge_ms may be NULL). */
int sl_hybrid_time(sl_hybrid *h, int warmup, int iters, double *ms_per_run, double *comb_ms,
                   double *merge_ms);
/* Cells combed per run (sum of block areas) and number of merge levels. */
int sl_hybrid_stats(const sl_hybrid *h, int64_t *cells, int *merge_levels);
/* enable != 0: the first run captures the whole pipeline (combing launch and
 * every merge level) into a HIP graph on the handle's stream, and every run
 * replays it with one graph launch.  Results are unchanged.  In graph mode
 * sl_hybrid_time reports the total only (comb_ms and merge_ms are -1). */
int sl_hybrid_set_graph(sl_hybrid *h, int enable);
/* Runs sl_batch_autotune on the hybrid's combing launch (its blocks as one
 * ragged batch) and keeps the fastest shape; a captured graph is dropped and
 * captured again on the next graph run.  Results never depend on it. */
int sl_hybrid_autotune(sl_hybrid *h, int iters, int *rows_per_lane, int *waves_per_cu);
void sl_hybrid_destroy(sl_hybrid *h);
/* Block split per length bucket.  Replaces the fixed block rule of          */
/*   hybrid_iterative_wrapper (semi_local.h:740-790) with a measured one:    */
/* for the pair (a, b), builds the hybrid for every candidate grid           */
/* cand_m[i] x cand_n[i], times `iters` runs of each after a warm-up (HIP    */
/* events, the default tile policy) and returns the fastest grid; ms         */
/* (optional, ncand entries) receives each grid's time.  The choice is kept  */
/* for the pair's length bucket, floor(log2(max(m, n))), for the process.    */
/* Results never depend on the grid.                                          */
int sl_hybrid_autotune_blocks(const void *a, int m, const void *b, int n, int sym_bytes, int ncand,
                              const int32_t *cand_m, const int32_t *cand_n, int iters, int32_t *best_m,
                              int32_t *best_n, double *ms);
/* The grid kept for the length bucket of (m, n); 0 x 0 when that bucket was */
/* never tuned.                                                               */
int sl_hybrid_bucket_blocks(int m, int n, int32_t *small_m, int32_t *small_n);

/* ------------------------------------------------------------------------ */
/* Request loop (serving): continuous batching of comparisons on one GPU.     */
/* No reference counterpart (the reference answers one comparison per call   */
/* and queries on a host Permutation, approximate_matching/utils.h:268-359).  */
/* Comparisons that arrive while others run are admitted as soon as device    */
/* memory (one pool carved at create time, at most 90% of free HBM) and a     */
/* launch slot are free; each admitted micro-batch is combed by one           */
/* persistent launch on its own stream, so micro-batches overlap on the       */
/* device.  Braids stay resident until released (or evicted, least recently  */
/* used first, when a new comparison needs the memory); follow-up queries are */
/* answered from them without recomputation.  Results equal the one-shot      */
/* entry points' on the same input.                                           */
/* ------------------------------------------------------------------------ */
typedef struct sl_server sl_server;

#define SL_QUERY_H 0                 /* raw H(i, j), i in [-m, n], j in [0, m+n]     */
#define SL_QUERY_STRING_SUBSTRING 1  /* LCS(a, b[i:j)),  0 <= i <= j <= n            */
#define SL_QUERY_SUBSTRING_STRING 2  /* LCS(a[i:j), b),  0 <= i <= j <= m            */

/* Resident-braid index: a wavelet matrix over a kernel (row -> col, size
 * m+n) that answers every query kind above in O(log(m+n)) rank steps (one
 * lane per query) instead of an O(m+n) scan.  No reference counterpart as an
 * index; the counts are those of approximate_matching/utils.h:268-359.
 * Memory: about 0.19 * (m+n) * log2(m+n) bytes. */
typedef struct sl_braid_index sl_braid_index;
int sl_braid_index_create(sl_braid_index **out, const int32_t *row_to_col, int m, int n);
/* nq queries: kind[k] in SL_QUERY_*, arguments (qi[k], qj[k]) as above;
 * out[k] = the answer (LCS length, or H(i, j)). */
int sl_braid_index_query(sl_braid_index *ix, int nq, const int32_t *kind, const int32_t *qi, const int32_t *qj,
                         int64_t *out);
int sl_braid_index_stats(const sl_braid_index *ix, int *levels, int64_t *bytes);
void sl_braid_index_destroy(sl_braid_index *ix);

typedef struct sl_server_stats_t {
    int64_t compares_done, queries_done, cells_done;
    int64_t launches;          /* micro-batches launched (comparisons and queries) */
    int64_t deferrals;         /* admissions postponed for lack of device memory */
    int64_t evictions;         /* resident braids evicted to admit new work */
    int64_t pool_bytes, pool_in_use, pool_peak;
    int64_t resident_braids;
    double busy_seconds;       /* first submission to last completion */
    int64_t indexed_braids;    /* index builds completed (sl_server_set_index) */
    int64_t index_queries;     /* queries answered from an index */
    int64_t peak_inflight;     /* most micro-batches on the device at one time (launched, not yet complete) */
    int64_t peak_inflight_compares;   /* the same, counting comparison micro-batches only */
} sl_server_stats_t;

/* pool_bytes 0: 90% of the free HBM; max_inflight 0: 4 concurrent
 * micro-batches; max_batch 0: 1024 requests per micro-batch; max_batch_cells
 * 0: 4e10 cells per micro-batch. */
int sl_server_create(sl_server **out, int64_t pool_bytes, int max_inflight, int max_batch,
                     int64_t max_batch_cells);
/* Kept braids with m + n >= min_size (default: off) get a wavelet-matrix
 * index (as sl_braid_index_*), built on the device on a stream of its own
 * once their comparison's result is out; queries on a braid whose index is
 * built take O(log(m+n)) rank steps instead of an O(m+n) scan (earlier ones
 * scan).  min_size <= 0 turns indexing off.  Applies to comparisons launched
 * after the call.  The index lives in the pool with its braid.  A build is
 * ~3 kernels per bit of m+n: it pays for braids that get many queries, and
 * costs latency when most braids get few. */
int sl_server_set_index(sl_server *s, int64_t min_size);
/* Queues the comparison of a (m symbols, rows) with b (n symbols, columns);
 * the input is copied.  *id names the request and its braid.  keep == 0
 * frees the braid as soon as the score is known.  Result: the LLCS. */
int sl_server_submit_compare(sl_server *s, const void *a, int m, const void *b, int n, int sym_bytes,
                             int keep, int64_t *id);
/* Queues a query (SL_QUERY_*) on the braid of comparison `braid`; it runs
 * once that braid is resident.  Result: the LCS (or H) value. */
int sl_server_submit_query(sl_server *s, int64_t braid, int kind, int i, int j, int64_t *id);
/* SL_OK with *result (and the submit-to-completion latency) when the request
 * is done, its failure status if it failed, SL_PENDING while it is queued or
 * running; wait != 0 blocks until it is done. */
int sl_server_result(sl_server *s, int64_t id, int wait, int64_t *result, double *latency_ms);
/* Copies a resident braid (row -> col, m + n entries) to the host. */
int sl_server_download_braid(sl_server *s, int64_t braid, int32_t *row_to_col);
/* Gives a braid up (freed once no queued query reads it). */
int sl_server_release(sl_server *s, int64_t braid);
int sl_server_stats(sl_server *s, sl_server_stats_t *stats);
void sl_server_destroy(sl_server *s);

/* ------------------------------------------------------------------------ */
/* Score-only LLCS by Hyyro's bit-vector recurrence, batched on the GPU.      */
/* Replaces: lllcs_hyyro::hyyro_magic / hyyro_magic_mpi_with_precalc          */
/*   (prefix_lcs/bitwise/transposition_network_binary_alphabet.h:743-916) and */
/*   the namespaced Hyyro<Word, Input>::compute (include/BitwiseLCS/           */
/*   BitParalellLCS.h:337-433); any alphabet with at most 256 distinct symbols */
/*   per pair, strings up to 2^20 symbols (the longer one is bit-packed).     */
/* Inputs are laid out as for sl_batch_create; llcs[i] receives pair i's LLCS.*/
/* ------------------------------------------------------------------------ */
int sl_llcs_hyyro(const void *a, int m, const void *b, int n, int sym_bytes, int *llcs);
int sl_llcs_hyyro_batch(int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                        const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes,
                        int32_t *llcs);

/* Resident form: symbols recoded and match-mask tables built once. */
typedef struct sl_hyyro sl_hyyro;
int sl_hyyro_create(sl_hyyro **out, int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                    const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes);
int sl_hyyro_run(sl_hyyro *h);                      /* asynchronous */
int sl_hyyro_result(sl_hyyro *h, int32_t *llcs);    /* waits, then copies npairs scores */
int sl_hyyro_time(sl_hyyro *h, int warmup, int iters, double *ms_per_run);
/* Work: sum over pairs of rows x 64-bit words (after dropping rows whose
 * symbol is absent from the packed string), and the mask-table words. */
int sl_hyyro_stats(const sl_hyyro *h, int64_t *word_rows, int64_t *table_words);
/* Length buckets by packed width (64-bit words of the longer string): up to */
/* 64, 128, 256, 2048 and 16384 words.  Each bucket runs one kernel form,    */
/* chosen by its 32-bit words per lane (the reference's analogue is the word */
/* type and threads of hyyro_magic_mpi_with_precalc, one per call): 2 (64    */
/* bits, the default), 1 or 4 in buckets 0, 1, 3, 4; 2 or 4 in bucket 2.     */
/* Buckets 0-2 run one workgroup per pair (1, 2, 4 waves at 2 words per      */
/* lane), buckets 3-4 segments of 4 waves.  Arrays of SL_HYYRO_BUCKETS       */
/* entries; any pointer may be null.  max_words[b] is in 64-bit words,       */
/* words_per_lane[b] in 32-bit words, waves[b] per pair (buckets 0-2) or per */
/* segment (3-4).                                                            */
#define SL_HYYRO_BUCKETS 5
int sl_hyyro_buckets(const sl_hyyro *h, int32_t *max_words, int32_t *pairs, int32_t *words_per_lane,
                     int32_t *waves);
/* Forces the 32-bit words per lane per bucket (0 keeps a bucket's choice); */
/* a value the bucket does not offer is SL_ERR_ARG.                          */
int sl_hyyro_set_words_per_lane(sl_hyyro *h, const int32_t *words_per_lane);
/* Autotuner: times every offered form on each bucket's own pairs (rounds    */
/* rounds of two timed launches after a warm-up) and keeps the fastest;      */
/* words_per_lane (optional) receives the choices.  Results never depend on  */
/* the choice.                                                               */
int sl_hyyro_autotune(sl_hyyro *h, int rounds, int32_t *words_per_lane);
void sl_hyyro_destroy(sl_hyyro *h);

/* ------------------------------------------------------------------------ */
/* Approximate matching over resident braids.                                 */
/* The pattern p is combed against overlapping chunks of the text (one ragged */
/* batch launch; every substring of length <= max_window lies inside one      */
/* chunk) and the chunk braids stay in HBM.  Scores are those of the          */
/* reference's duplicate searches (match 0, mismatch -2, gap -1):             */
/*   score(s, e) = 2 LCS(p, text[s, e)) - m - (e - s)   (-indel distance).    */
/* ------------------------------------------------------------------------ */
typedef struct sl_amatch sl_amatch;

/* chunk = windows per chunk (0: a default from the text length and the CU
 * count).  m, n >= 1; sym_bytes 1, 2 or 4. */
int sl_amatch_create(sl_amatch **out, const void *p, int m, const void *text, int n, int sym_bytes,
                     int max_window, int chunk);
/* Chunks, windows per chunk, combed cells (m x chunk text lengths). */
int sl_amatch_stats(const sl_amatch *am, int *chunks, int *chunk_windows, int64_t *cells);
/* For every window [i, i+w), i = 0 .. n-w (n-w+1 entries per array): the
 * window's own score, and its best substring [best_start, best_end) of length
 * >= min_len: the highest score, then the longest, then the rightmost -- the
 * window maximum of InteractiveDuplicateSearchSemiDense / SemiSparse
 * (approximate_matching/ApproximateMatching.h:158-207, 373-413).  Needs
 * 1 <= w <= min(n, max_window), 0 <= min_len <= w, w - min_len < 4096. */
int sl_amatch_windows(sl_amatch *am, int w, int min_len, int32_t *window_score, int32_t *best_start,
                      int32_t *best_end, int32_t *best_score);
/* HIP-event times per run: re-combing every chunk, and the window kernel. */
int sl_amatch_time(sl_amatch *am, int w, int min_len, int warmup, int iters, double *comb_ms,
                   double *windows_ms);
/* Replaces: InteractiveDuplicateSearchSemiDense::find (variant 0,            */
/*   ApproximateMatching.h:339-432) and InteractiveDuplicateSearchSemiSparse::*/
/*   find (variant 1, :252-336) -- (p, p_size, text, text_size, k, result).   */
/* Needs 0 < k <= 1 and int(m / k) <= max_window.  *count receives the number */
/* of intervals; sl_duplicate_results copies them (start, end, score).        */
int sl_duplicate_search(sl_amatch *am, double k, int variant, int *count);
int sl_duplicate_results(const sl_amatch *am, int32_t *start, int32_t *end, double *score);
/* Wall time of the last sl_duplicate_search: window kernel + copy to host,
 * then the sequential phases on the host. */
int sl_amatch_last_times(const sl_amatch *am, double *device_ms, double *host_ms);
void sl_amatch_destroy(sl_amatch *am);

#ifdef __cplusplus
}
#endif

#endif /* SEMILOCAL_HIP_H */
