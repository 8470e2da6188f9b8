/*
 * oracle.c -- CPU restatement of the reference's semi-local LCS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * engine in semilocal_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only to check or to
 * time the CPU side.  The product path never calls into it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference's source/cpu/ directory).  Conventions (all int32):
 *   a[0..m) are the rows (vertical string), b[0..n) the columns.
 *   perm[s] = exit position of strand s  (row -> col of the m+n square
 *   permutation matrix): cols 0..n-1 are bottom exits, n..n+m-1 right exits.
 *
 * Pinning (what checks each function independently of this file's own logic):
 *   orc_comb_rowmajor / orc_comb_antidiag / orc_hybrid_iterative
 *       -> golden vectors produced by the compiled reference (tests/golden),
 *          brute-force kernel from the DP matrix (orc_kernel_bruteforce),
 *          LLCS read-out vs. DP, and row-major == anti-diagonal == hybrid.
 *   orc_kernel_bruteforce -> LLCS/H(i,j) vs. independent DP on windows.
 *   orc_steady_ant / orc_staggered -> O(n^3) tropical product
 *          (orc_mult_naive), identity neutrality, associativity, and
 *          golden products from the compiled reference.
 *   orc_bitpar_llcs -> LLCS by DP and by the integer permutation.
 *   orc_semi_local_H -> DP on string-substring / substring-string windows.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NOPOINT (-1)

/* ------------------------------------------------------------------------- */
/* A1: row-major combing.  semi_local.h:135-196 (sticky_braid_sequential),    */
/* SemiLocalLCS.h:68-125 (SimpleIterativeCombing::compute).                   */
/* ------------------------------------------------------------------------- */
void orc_comb_rowmajor(const int32_t *a, int m, const int32_t *b, int n, int32_t *perm)
{
    int32_t *top = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *left = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    for (int i = 0; i < m; ++i) left[i] = i;          /* semi_local.h:146-148 */
    for (int j = 0; j < n; ++j) top[j] = m + j;       /* semi_local.h:149-151 */
    for (int i = 0; i < m; ++i) {
        int slot = m - 1 - i;                         /* strand slot of row i, :154 */
        int32_t l = left[slot];
        int32_t sym = a[i];
        for (int j = 0; j < n; ++j) {
            int32_t t = top[j];
            if (sym == b[j] || l > t) {               /* combing rule, :160 */
                top[j] = l;
                l = t;
            }
        }
        left[slot] = l;
    }
    for (int l = 0; l < m; ++l) perm[left[l]] = n + l;   /* :190 */
    for (int j = 0; j < n; ++j) perm[top[j]] = j;        /* :192 */
    free(top);
    free(left);
}

/* ------------------------------------------------------------------------- */
/* A2: anti-diagonal combing in three phases with optional OpenMP.            */
/* semi_local.h:212-276 (sticky_braid_mpi), :35-67 (anti_diagonal_computation)*/
/* :70-82 (initialization), :84-111 (construct_permutation incl. reverse),    */
/* :113-119 (fill_a_reverse).                                                  */
/* ------------------------------------------------------------------------- */
static void antidiag_step(int32_t *map, const int32_t *ar, const int32_t *b, int len,
                          int le, int te, int oa, int ob)
{
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int k = 0; k < len; ++k) {
        int32_t l = map[le + k], t = map[te + k];
        int r = (ar[oa + k] == b[ob + k]) || (l > t);
        if (r) { map[le + k] = t; map[te + k] = l; }
    }
}

static void comb_antidiag_impl(const int32_t *a, int m, const int32_t *b, int n,
                               int32_t *perm, int threads, int reverse)
{
    if (m > n) {   /* semi_local.h:218-221: swap and mark reversed */
        comb_antidiag_impl(b, n, a, m, perm, threads, !reverse);
        return;
    }
    int size = m + n;
    int32_t *map = (int32_t *)malloc(sizeof(int32_t) * (size_t)(size > 0 ? size : 1));
    int32_t *ar = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    int same = (m + n - 1) - 2 * (m - 1);     /* :230-231 */
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
#endif
    {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int k = 0; k < size; ++k) map[k] = k;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int i = 0; i < m; ++i) ar[i] = a[m - 1 - i];
        /* phase 1: growing diagonals, :244-251 */
        int le = m - 1;
        for (int d = 0; d < m - 1; ++d, --le) antidiag_step(map, ar, b, d + 1, le, m, le, 0);
        /* phase 2: full-length diagonals, :254-258 */
        for (int j = 0; j < same; ++j) antidiag_step(map, ar, b, m, 0, m + j, 0, j);
        /* phase 3: shrinking diagonals, :261-269 */
        int sj = same;
        for (int d = m - 2; d >= 0; --d, ++sj) antidiag_step(map, ar, b, d + 1, 0, sj + m, 0, sj);
        /* (the omp for above carries an implicit barrier per diagonal, as :64-66) */
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int s = 0; s < size; ++s) {
            /* construct_permutation, :86-111 */
            int strand = map[s];
            int col = (s >= m) ? (s - m) : (n + s);
            if (!reverse) perm[strand] = col;
            else perm[size - 1 - strand] = size - 1 - col;
        }
    }
    free(map);
    free(ar);
}

void orc_comb_antidiag(const int32_t *a, int m, const int32_t *b, int n, int32_t *perm, int threads)
{
    comb_antidiag_impl(a, m, b, n, perm, threads, 0);
}

/* Batch of independent pairs: OpenMP parallel-for over pairs, each combed by
 * the single-threaded anti-diagonal routine (the reference has no batch API;
 * this is the CPU baseline shape for ragged batches).                         */
void orc_comb_batch(const int32_t *a_cat, const int64_t *a_off, const int32_t *a_len,
                    const int32_t *b_cat, const int64_t *b_off, const int32_t *b_len,
                    const int64_t *p_off, int32_t *perm_cat, int npairs, int threads)
{
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
#endif
    for (int p = 0; p < npairs; ++p)
        comb_antidiag_impl(a_cat + a_off[p], a_len[p], b_cat + b_off[p], b_len[p],
                           perm_cat + p_off[p], 1, 0);
}

/* ------------------------------------------------------------------------- */
/* LLCS read-out from the kernel.  include/LCS/LCS.h:223-233.                 */
/* ------------------------------------------------------------------------- */
int orc_llcs_from_perm(const int32_t *perm, int m, int n)
{
    int acc = 0;
    for (int s = 0; s < m; ++s) if (perm[s] >= n) acc++;
    return m - acc;
}

/* DP LLCS, two rows.  naive_prefix_lcs.h:49-88 (prefix_lcs_sequential).       */
int orc_llcs_dp(const int32_t *a, int m, const int32_t *b, int n)
{
    int32_t *prev = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    int32_t *cur = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    for (int i = 1; i <= m; ++i) {
        cur[0] = 0;
        for (int j = 1; j <= n; ++j) {
            int v = prev[j] > cur[j - 1] ? prev[j] : cur[j - 1];
            int d = prev[j - 1] + (a[i - 1] == b[j - 1]);
            cur[j] = v > d ? v : d;
        }
        int32_t *t = prev; prev = cur; cur = t;
    }
    int r = prev[n];
    free(prev);
    free(cur);
    return r;
}

/* ------------------------------------------------------------------------- */
/* Brute-force kernel from the explicit semi-local matrix: LCS of a against    */
/* every window of b padded with m wildcards on both sides, then the cross-    */
/* difference permutation.  SemiLocalLCS.h:13-63 (DummyPrefixLCSStrategy) and   */
/* common/include/Permutations/Matrices.h:116-124 (getCrossDiffPermutation).   */
/* O((m+n)^2 * m * n): tiny inputs only.                                       */
/* ------------------------------------------------------------------------- */
void orc_kernel_bruteforce(const int32_t *a, int m, const int32_t *b, int n, int32_t *perm)
{
    const int32_t WILD = INT32_MIN;
    int ext = m + n + m;
    int32_t *bx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ext > 0 ? ext : 1));
    for (int i = 0; i < m; ++i) { bx[i] = WILD; bx[m + n + i] = WILD; }
    for (int j = 0; j < n; ++j) bx[m + j] = b[j];
    int N = m + n + 1;
    int32_t *H = (int32_t *)malloc(sizeof(int32_t) * (size_t)N * (size_t)N);
    int32_t *row = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ext + 1));
    int32_t *nrow = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ext + 1));
    for (int i = 0; i < N; ++i) {
        /* one DP of a vs bx[i .. j+m) for all j at once (prefix DP from i) */
        for (int j = 0; j <= ext; ++j) row[j] = 0;
        for (int r = 1; r <= m; ++r) {
            nrow[i] = 0;
            for (int c = i + 1; c <= ext; ++c) {
                int v = row[c] > nrow[c - 1] ? row[c] : nrow[c - 1];
                int match = (a[r - 1] == bx[c - 1]) || (bx[c - 1] == WILD);
                int d = row[c - 1] + match;
                nrow[c] = v > d ? v : d;
            }
            int32_t *t = row; row = nrow; nrow = t;
        }
        for (int j = 0; j < N; ++j) {
            int v;
            if (j <= i - m) v = -(j - i + m);      /* SemiLocalLCS.h:49-50 */
            else v = -row[j + m];                  /* LCS(a, bx[i, j+m)), :52 */
            H[(size_t)i * N + j] = v;
        }
    }
    for (int s = 0; s < m + n; ++s) perm[s] = NOPOINT;
    for (int i = 0; i < N - 1; ++i)
        for (int j = 0; j < N - 1; ++j) {
            int cd = (H[(size_t)i * N + j + 1] + H[(size_t)(i + 1) * N + j]) -
                     (H[(size_t)i * N + j] + H[(size_t)(i + 1) * N + j + 1]);
            if (cd == 1) perm[i] = j;
        }
    free(bx); free(H); free(row); free(nrow);
}

/* ------------------------------------------------------------------------- */
/* Semi-local matrix entry from the kernel.                                    */
/* approximate_matching/utils.h:281-293 (dominanceSumRandom) and :331-334      */
/* (canonicalDecompositionWithKnown with v = 1):                               */
/*   H(i,j) = j - i - #{ s : s >= i + m and perm[s] < j },  i in [-m, n].      */
/* LCS(a, b[l:r)) = H(l, r); LCS(a[s:e), b) = H(-s, n+m-e) - s - (m-e).        */
/* ------------------------------------------------------------------------- */
int orc_semi_local_H(const int32_t *perm, int m, int n, int i, int j)
{
    int cnt = 0;
    for (int s = i + m; s < m + n; ++s) if (s >= 0 && perm[s] < j) cnt++;
    return j - i - cnt;
}

/* ------------------------------------------------------------------------- */
/* Unit-Monge (sticky braid) multiplication.                                   */
/* Permutations are given row->col (int32, n entries, full permutations).      */
/* ------------------------------------------------------------------------- */

/* O(n^3) tropical product of the dominance matrices + cross difference.
 * unit_monge_mult/steady_ant.h:28-64 (naive::mult_dist), matrices.h:342-359.  */
void orc_mult_naive(const int32_t *p, const int32_t *q, int n, int32_t *r)
{
    int N = n + 1;
    int32_t *dp = (int32_t *)calloc((size_t)N * N, sizeof(int32_t));
    int32_t *dq = (int32_t *)calloc((size_t)N * N, sizeof(int32_t));
    int32_t *dc = (int32_t *)malloc(sizeof(int32_t) * (size_t)N * N);
    /* dominance sum: D(i,j) = #{ (row,col) : row >= i, col < j }  (top_right_summator) */
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            int c1 = 0, c2 = 0;
            for (int row = i; row < n; ++row) { if (p[row] < j) c1++; if (q[row] < j) c2++; }
            dp[(size_t)i * N + j] = c1;
            dq[(size_t)i * N + j] = c2;
        }
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < N; ++k) {
            int best = INT32_MAX;
            for (int j = 0; j < N; ++j) {
                int v = dp[(size_t)i * N + j] + dq[(size_t)j * N + k];
                if (v < best) best = v;
            }
            dc[(size_t)i * N + k] = best;
        }
    for (int i = 0; i < n; ++i) r[i] = NOPOINT;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            int cd = (dc[(size_t)i * N + j + 1] + dc[(size_t)(i + 1) * N + j]) -
                     (dc[(size_t)i * N + j] + dc[(size_t)(i + 1) * N + j + 1]);
            if (cd == 1) r[i] = j;
        }
    free(dp); free(dq); free(dc);
}

/* A sub-permutation view: row->col and col->row arrays of size sz. */
typedef struct { int32_t *rc; int32_t *cr; int sz; } sperm;

static void sp_clear(sperm *x)
{
    for (int i = 0; i < x->sz; ++i) { x->rc[i] = NOPOINT; x->cr[i] = NOPOINT; }
}
static inline void sp_set(sperm *x, int r, int c) { x->rc[r] = c; x->cr[c] = r; }

/* get_vertical_slice, steady_ant.h:100-112 */
static void vertical_slice(const sperm *p, int c0, int c1, int32_t *rowmap, sperm *out)
{
    int ptr = 0;
    for (int row = 0; row < p->sz; ++row) {
        int col = p->rc[row];
        if (col >= c0 && col < c1) { rowmap[ptr] = row; sp_set(out, ptr, col - c0); ptr++; }
    }
}
/* get_horizontal_slice, steady_ant.h:123-136 */
static void horizontal_slice(const sperm *q, int r0, int r1, int32_t *colmap, sperm *out)
{
    int ptr = 0;
    for (int col = 0; col < q->sz; ++col) {
        int row = q->cr[col];
        if (row >= r0 && row < r1) { colmap[ptr] = col; sp_set(out, row - r0, ptr); ptr++; }
    }
}
/* inverse_mapping, steady_ant.h:146-160 */
static void inverse_map(const sperm *sh, const int32_t *rowmap, const int32_t *colmap, sperm *flat)
{
    sp_clear(flat);
    for (int c = 0; c < sh->sz; ++c) sp_set(flat, rowmap[sh->cr[c]], colmap[c]);
}

/* dominance_sum_queries.h:18-82: incremental moves.
 *   "top_left"   sum(row,col) = #{ (r,c) : r >= row, c >= col }
 *   "bottom_right" sum(row,col) = #{ (r,c) : r <  row, c <  col }          */
static inline int tl_left(int row, int col, int s, const sperm *x)
{ if (col == 0) return s; int rc = x->cr[col - 1]; if (rc >= row && rc != NOPOINT) s++; return s; }
static inline int tl_down(int row, int col, int s, const sperm *x)
{ if (row >= x->sz) return 0; int cc = x->rc[row]; if (cc >= col && cc != NOPOINT) s--; return s; }
static inline int tl_up(int row, int col, int s, const sperm *x)
{ if (row == 0) return s; int cc = x->rc[row - 1]; if (cc >= col && cc != NOPOINT) s++; return s; }
static inline int tl_right(int row, int col, int s, const sperm *x)
{ if (col >= x->sz) return 0; int rc = x->cr[col]; if (rc >= row && rc != NOPOINT) s--; return s; }
static inline int br_left(int row, int col, int s, const sperm *x)
{ if (col == 0) return s; int rc = x->cr[col - 1]; if (rc != NOPOINT && rc < row) s--; return s; }
static inline int br_down(int row, int col, int s, const sperm *x)
{ if (row >= x->sz) return 0; int cc = x->rc[row]; if (cc < col && cc != NOPOINT) s++; return s; }
static inline int br_up(int row, int col, int s, const sperm *x)
{ if (row == 0) return s; int cc = x->rc[row - 1]; if (cc < col && cc != NOPOINT) s--; return s; }
static inline int br_right(int row, int col, int s, const sperm *x)
{ if (col >= x->sz) return 0; int rc = x->cr[col]; if (rc < row && rc != NOPOINT) s++; return s; }

/* ant_passage + merge, steady_ant.h:186-265 and :434-450.  Result in r_lo. */
static void ant_passage_merge(sperm *rlo, sperm *rhi, int n)
{
    int32_t *gr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * n + 2));
    int32_t *gc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * n + 2));
    int ng = 0;
    int cur_row = n, cur_col = -1;
    int shi = 0, slo = 0;
    int went_right = 0;
    for (;;) {
        if (cur_row == 0) break;
        if (cur_col == n) break;
        int dr = cur_row - 1, dc = cur_col + 1;
        if (went_right) {
            shi = br_right(dr, dc - 1, shi, rhi);
            slo = tl_right(dr, dc - 1, slo, rlo);
        } else {
            shi = br_up(dr + 1, dc, shi, rhi);
            slo = tl_up(dr + 1, dc, slo, rlo);
        }
        if (shi - slo < 0) { went_right = 1; cur_col++; }
        else if (shi - slo == 0) { went_right = 0; cur_row--; }
        else { abort(); /* "Impissble" branch of :234-236 never taken for permutations */ }
        if (dc > 0) {
            int above_left = br_left(dr, dc, shi, rhi) - tl_left(dr, dc, slo, rlo);
            int below_right = br_down(dr, dc, shi, rhi) - tl_down(dr, dc, slo, rlo);
            if (above_left < 0 && below_right > 0) { gr[ng] = dr; gc[ng] = dc - 1; ng++; }
        }
    }
    for (int i = 0; i < n; ++i) {
        int c = rhi->rc[i];
        if (c != NOPOINT) sp_set(rlo, i, c);
    }
    for (int g = 0; g < ng; ++g) sp_set(rlo, gr[g], gc[g]);
    free(gr);
    free(gc);
}

/* steady_ant_optimized_seq, steady_ant.h:378-453; result overwrites p.
 * Base case n <= 1 stands in for the precalc table (precalc, :860-894), which
 * holds the same unique products.                                            */
static void steady_ant_rec(sperm *p, sperm *q)
{
    int n = p->sz;
    if (n <= 1) {
        if (n == 1) { int r = p->cr[0], c = q->rc[0]; sp_clear(p); if (r != NOPOINT && c != NOPOINT) sp_set(p, r, c); }
        return;
    }
    int h = n / 2, k = n - h;
    int32_t *mem = (int32_t *)malloc(sizeof(int32_t) * (size_t)(8 * n + 4 * n));
    sperm plo = { mem, mem + h, h }, qlo = { mem + 2 * h, mem + 3 * h, h };
    sperm phi = { mem + 4 * h, mem + 4 * h + k, k }, qhi = { mem + 4 * h + 2 * k, mem + 4 * h + 3 * k, k };
    int32_t *plo_rm = mem + 8 * n, *qlo_cm = plo_rm + h, *phi_rm = qlo_cm + h, *qhi_cm = phi_rm + k;
    sp_clear(&plo); sp_clear(&qlo); sp_clear(&phi); sp_clear(&qhi);
    vertical_slice(p, 0, h, plo_rm, &plo);
    horizontal_slice(q, 0, h, qlo_cm, &qlo);
    vertical_slice(p, h, n, phi_rm, &phi);
    horizontal_slice(q, h, n, qhi_cm, &qhi);
    steady_ant_rec(&plo, &qlo);
    steady_ant_rec(&phi, &qhi);
    inverse_map(&plo, plo_rm, qlo_cm, p);   /* r_lo reuses p, :419 */
    inverse_map(&phi, phi_rm, qhi_cm, q);   /* r_hi reuses q, :420 */
    ant_passage_merge(p, q, n);
    free(mem);
}

/* Product r = p (.) q of two n x n permutations given row->col. */
void orc_steady_ant(const int32_t *p, const int32_t *q, int n, int32_t *r)
{
    int32_t *mem = (int32_t *)malloc(sizeof(int32_t) * (size_t)(4 * (n > 0 ? n : 1)));
    sperm P = { mem, mem + n, n }, Q = { mem + 2 * n, mem + 3 * n, n };
    sp_clear(&P); sp_clear(&Q);
    for (int i = 0; i < n; ++i) { if (p[i] != NOPOINT) sp_set(&P, i, p[i]); if (q[i] != NOPOINT) sp_set(&Q, i, q[i]); }
    steady_ant_rec(&P, &Q);
    for (int i = 0; i < n; ++i) r[i] = P.rc[i];
    free(mem);
}

/* staggered_sticky_multiplication<RowGlue,false>, steady_ant.h:609-737.
 * p is (pr x pr), q is (qr x qr), k shared strands; product is
 * (pr+qr-k) square.  row_glue != 0 glues vertical neighbours (p = the block *
 * of the earlier rows of a, as semi_local.h:711-715 calls it).               */
void orc_staggered(const int32_t *p, int pr, const int32_t *q, int qr, int k, int row_glue, int32_t *prod)
{
    int N = pr + qr - k;
    for (int i = 0; i < N; ++i) prod[i] = NOPOINT;
    int32_t *pcr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(pr > 0 ? pr : 1));
    int32_t *qcr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(qr > 0 ? qr : 1));
    for (int i = 0; i < pr; ++i) pcr[p[i]] = i;
    for (int i = 0; i < qr; ++i) qcr[q[i]] = i;
    if (k == 0) {    /* :621-646 */
        for (int i = 0; i < pr; ++i) { if (row_glue) prod[i + qr] = p[i] + qr; else prod[i] = p[i]; }
        for (int i = 0; i < qr; ++i) { if (row_glue) prod[i] = q[i]; else prod[i + pr] = q[i] + pr; }
        free(pcr); free(qcr);
        return;
    }
    int32_t *mem = (int32_t *)malloc(sizeof(int32_t) * (size_t)(6 * k));
    sperm pred = { mem, mem + k, k }, qred = { mem + 2 * k, mem + 3 * k, k };
    int32_t *map_row = mem + 4 * k, *map_col = mem + 5 * k;
    sp_clear(&pred); sp_clear(&qred);
    sperm P = { (int32_t *)p, pcr, pr }, Q = { (int32_t *)q, qcr, qr };
    if (row_glue) {  /* :672-675 */
        vertical_slice(&P, 0, k, map_row, &pred);
        horizontal_slice(&Q, qr - k, qr, map_col, &qred);
    } else {         /* :676-680 */
        vertical_slice(&P, pr - k, pr, map_row, &pred);
        horizontal_slice(&Q, 0, k, map_col, &qred);
    }
    steady_ant_rec(&pred, &qred);
    for (int i = 0; i < k; ++i) {    /* :695-704 */
        int cur_col = map_col[pred.rc[i]], cur_row = map_row[i];
        if (row_glue) prod[qr - k + cur_row] = cur_col;
        else prod[cur_row] = cur_col + pr - k;
    }
    if (row_glue) {  /* :707-729 */
        for (int j = k; j < pr; ++j) prod[pcr[j] + qr - k] = j + qr - k;
        for (int i = 0; i < qr - k; ++i) prod[i] = q[i];
    } else {
        for (int j = 0; j < pr - k; ++j) prod[pcr[j]] = j;
        for (int i = k; i < qr; ++i) prod[i - k + pr] = pr - k + q[i];
    }
    free(mem); free(pcr); free(qcr);
}

/* ------------------------------------------------------------------------- */
/* A9: hybrid_iterative, semi_local.h:577-738.  Blocks are combed with the    */
/* anti-diagonal routine, then merged pairwise level by level.                */
/* ------------------------------------------------------------------------- */
typedef struct { int32_t *rc; int size, bm, bn; } blockperm;

void orc_hybrid_iterative(const int32_t *a, int m, const int32_t *b, int n,
                          int small_m, int small_n, int32_t *perm, int threads)
{
    int cols_per_block = n / small_n;           /* :586-587 */
    int rows_per_block = m / small_m;
    int ntasks = small_m * small_n;
    blockperm *cur = (blockperm *)calloc((size_t)ntasks, sizeof(blockperm));
    blockperm *nxt = (blockperm *)calloc((size_t)ntasks, sizeof(blockperm));
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
#endif
    for (int i = 0; i < ntasks; ++i) {           /* :596-627 */
        int sc = (i % small_n) * cols_per_block;
        int ec = sc + cols_per_block < n ? sc + cols_per_block : n;
        int sr = (i / small_n) * rows_per_block;
        int er = sr + rows_per_block < m ? sr + rows_per_block : m;
        if (i % small_n == small_n - 1) ec = n;
        if (i / small_n == small_m - 1) er = m;
        int bm = er - sr, bn = ec - sc;
        cur[i].rc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(bm + bn));
        cur[i].size = bm + bn; cur[i].bm = bm; cur[i].bn = bn;
        comb_antidiag_impl(a + sr, bm, b + sc, bn, cur[i].rc, 1, 0);
    }
    int steps = 0;
    { int t = 1; while (t < small_n) { t *= 2; steps++; } t = 1; while (t < small_m) { t *= 2; steps++; } }
    int block_rows = rows_per_block, block_cols = cols_per_block;
    int sm = small_m, sn = small_n;
    for (int st = 0; st < steps; ++st) {          /* :641-725 */
        int in_row = sn > sm;
        if (sn > 1 && sm > 1) in_row = block_rows >= 2 * block_cols;
        if (in_row) block_cols *= 2; else block_rows *= 2;
        int ncols = in_row ? (sn + 1) / 2 : sn;
        int nrows = !in_row ? (sm + 1) / 2 : sm;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
#endif
        for (int i = 0; i < ncols * nrows; ++i) {
            int r = i / ncols, c = i % ncols;
            blockperm *P, *Q;
            if (in_row) {
                P = &cur[r * sn + 2 * c];
                if (2 * c + 1 >= sn) { nxt[i] = *P; P->rc = NULL; continue; }
                Q = &cur[r * sn + 2 * c + 1];
                int N = P->bm + Q->bn + P->bn;
                int32_t *pr = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
                orc_staggered(P->rc, P->size, Q->rc, Q->size, P->bm, 0, pr);
                nxt[i].rc = pr; nxt[i].size = N; nxt[i].bm = P->bm; nxt[i].bn = P->bn + Q->bn;
            } else {
                P = &cur[2 * r * sn + c];
                if (2 * r + 1 >= sm) { nxt[i] = *P; P->rc = NULL; continue; }
                Q = &cur[(2 * r + 1) * sn + c];
                int N = P->bm + Q->bm + P->bn;
                int32_t *pr = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
                orc_staggered(P->rc, P->size, Q->rc, Q->size, P->bn, 1, pr);
                nxt[i].rc = pr; nxt[i].size = N; nxt[i].bm = P->bm + Q->bm; nxt[i].bn = P->bn;
            }
            free(P->rc); P->rc = NULL;
            free(Q->rc); Q->rc = NULL;
        }
        sn = ncols; sm = nrows;
        blockperm *t = cur; cur = nxt; nxt = t;
    }
    for (int j = 0; j < m + n; ++j) perm[j] = cur[0].rc[j];   /* :731-732 */
    free(cur[0].rc);
    free(cur);
    free(nxt);
}

/* ------------------------------------------------------------------------- */
/* A5: bit-parallel LLCS for binary strings, formula 2.                        */
/* prefix_lcs/bitwise/transposition_network_binary_alphabet.h:252-323          */
/* (process_antidiag_formula2) and :598-663 (llcs_2symbol_smart_combing).      */
/* a_rev: a reversed, packed MSB-first per word; b packed LSB-first            */
/* (encoders_and_decoders.h:50-139).  Lengths must be multiples of 64.         */
/* ------------------------------------------------------------------------- */
static inline void bit_cell(uint64_t *lp, uint64_t *tp, uint64_t sa, uint64_t sb)
{
    uint64_t l = *lp, t = *tp, mask = 1, lc, tc, cond;
    for (int sh = 63; sh > 0; --sh) {                 /* upper half, :271-286 */
        lc = l >> sh; tc = t << sh;
        cond = ~((sa >> sh) ^ sb);
        t = (lc | ~mask) & (t | (cond & mask));
        l = tc ^ (t << sh) ^ l;
        mask = (mask << 1) | 1u;
    }
    cond = ~(sa ^ sb);                                /* centre, :288-295 */
    lc = l; tc = t;
    t = (lc | ~mask) & (t | (cond & mask));
    l = tc ^ t ^ l;
    mask = ~(uint64_t)0;
    for (int sh = 1; sh < 64; ++sh) {                 /* lower half, :300-308 */
        mask <<= 1;
        lc = l << sh; tc = t >> sh;
        cond = ~((sa << sh) ^ sb);
        t = (lc | ~mask) & (t | (cond & mask));
        l = tc ^ (t >> sh) ^ l;
    }
    *lp = l; *tp = t;
}

static void bit_diag(int lo, int hi, int le, int te, uint64_t *ls, uint64_t *ts,
                     const uint64_t *ar, const uint64_t *b)
{
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int j = lo; j < hi; ++j) bit_cell(&ls[le + j], &ts[te + j], ar[le + j], b[te + j]);
}

int orc_bitpar_llcs(const uint64_t *a_rev, int m, const uint64_t *b, int n, int a_total, int threads)
{
    uint64_t *ls = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)m);
    uint64_t *ts = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    long dis = 0;
    int same = (m + n - 1) - 2 * (m - 1);
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1) reduction(+:dis)
#endif
    {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int k = 0; k < m; ++k) ls[k] = ~(uint64_t)0;   /* :619-620 */
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int k = 0; k < n; ++k) ts[k] = 0;              /* :621-622 */
        for (int d = 0; d < m - 1; ++d) bit_diag(0, d + 1, m - 1 - d, 0, ls, ts, a_rev, b);
        for (int k = 0; k < same; ++k) bit_diag(0, m, 0, k, ls, ts, a_rev, b);
        int sj = same;
        for (int d = m - 1; d >= 1; --d, ++sj) bit_diag(0, d, 0, sj, ls, ts, a_rev, b);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int i = 0; i < m; ++i) dis += __builtin_popcountll(ls[i]);   /* :648-659 */
    }
    free(ls);
    free(ts);
    return a_total - (int)dis;
}

/* Bit packing of a 0/1 symbol string.  encode_reverse (MSB-first, reversed)
 * encoders_and_decoders.h:100-139 and encode (LSB-first) :50-87, for the
 * identity alphabet map 0->0, 1->1 (LLCS is invariant under relabeling).   */
void orc_pack_bits_reverse(const uint8_t *a, int len, uint64_t *out)
{
    int words = (len + 63) / 64;
    for (int w = 0; w < words; ++w) {
        uint64_t word = 0;
        for (int s = 0; s < 64 && w * 64 + s < len; ++s)
            word |= (uint64_t)(a[w * 64 + s] & 1) << (63 - s);
        out[words - 1 - w] = word;
    }
}
void orc_pack_bits(const uint8_t *b, int len, uint64_t *out)
{
    int words = (len + 63) / 64;
    for (int w = 0; w < words; ++w) {
        uint64_t word = 0;
        for (int s = 0; s < 64 && w * 64 + s < len; ++s)
            word |= (uint64_t)(b[w * 64 + s] & 1) << s;
        out[w] = word;
    }
}

int orc_openmp_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
