/*
 * CPU oracle for the database scan — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the scalar Gotoh reference kernel of hyalite 0.3.0 and of the
 * database-level reductions built on it. Only the test-suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load this library, and only as the checker: the
 * product path (hyalite_amd/csrc) never links, loads or calls it.
 *
 * Reference files restated (paths relative to the hyalite source tree):
 *   scoring validation ........ src/scoring.rs:65-100
 *   width proof ............... src/width.rs:36-71 (ScoreWidth), 107-169 (magnitude_bound,
 *                               required_width)
 *   mode flags ................ src/kernel.rs:43-98 (Flags::for_mode)
 *   gap penalty ............... src/kernel.rs:103-109
 *   DP fill ................... src/kernel.rs:168-217 (fill_dp)
 *   answer reduction .......... src/kernel.rs:113-129 (Best::consider), 223-260 (reduce_answer)
 *   scan / scan_all / scan_scores (scalar backend) ... src/database.rs:115-144, 405-414,
 *                               453-462, 511-513
 *   traceback ................. src/align.rs:169-420 (full-matrix DP, canonical walk),
 *                               230-290 (budget arithmetic)
 *   brute-force enumerators ... tests/common/mod.rs:65-171 (independent structure: exhaustive
 *                               path / substring enumeration, used to pin the DP)
 *
 * Every function below is pinned by the tests in tests/test_oracle.py (hand vectors from the
 * reference's own tests, exhaustive brute force over short inputs, width-proof boundary KATs).
 *
 * Arithmetic: cells are computed in int64 and stored as int32. For every input the width proof
 * accepts, the reference's wrapping i32 arithmetic never wraps, so the results are identical.
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#include "oracle.h"

/* src/kernel.rs:39 — sentinel for unreachable E/F cells. */
#define HYO_NEG ((int64_t)(INT32_MIN / 4))

/* src/kernel.rs:43-98 */
typedef struct {
    int top_row_free, left_col_free, answer_last_row, answer_last_col, local;
} hyo_flags;

static hyo_flags flags_for_mode(uint32_t mode) {
    hyo_flags f = {0, 0, 0, 0, 0};
    switch (mode) {
    case HYO_MODE_NW: break;
    case HYO_MODE_HW: f.top_row_free = 1; f.answer_last_row = 1; break;
    case HYO_MODE_OV:
        f.top_row_free = 1; f.left_col_free = 1; f.answer_last_row = 1; f.answer_last_col = 1;
        break;
    case HYO_MODE_SW: f.top_row_free = 1; f.left_col_free = 1; f.local = 1; break;
    case HYO_MODE_SHW: f.left_col_free = 1; f.answer_last_col = 1; break;
    default: break;
    }
    return f;
}

/* src/kernel.rs:103-109 — Opal convention go + (len-1)*ge, 0 for an empty gap. */
int32_t hyo_gap_penalty(int32_t go, int32_t ge, uint64_t len) {
    if (len == 0) return 0;
    return (int32_t)((int64_t)go + ((int64_t)len - 1) * (int64_t)ge);
}

/* src/scoring.rs:65-100 — error order: EmptyAlphabet, MatrixShape, NegativeGapPenalty,
 * GapOpenLessThanExtend. */
int hyo_scoring_check(uint32_t al, const int32_t* matrix, uint64_t matrix_len, int32_t go,
                      int32_t ge, int32_t* min_e, int32_t* max_e) {
    if (al == 0) return HYO_E_EMPTY_ALPHABET;
    if ((uint64_t)al * (uint64_t)al != matrix_len) return HYO_E_MATRIX_SHAPE;
    if (go < 0 || ge < 0) return HYO_E_NEGATIVE_GAP;
    if (go < ge) return HYO_E_GAP_OPEN_LT_EXT;
    int32_t lo = matrix[0], hi = matrix[0];
    for (uint64_t k = 1; k < matrix_len; ++k) {
        if (matrix[k] < lo) lo = matrix[k];
        if (matrix[k] > hi) hi = matrix[k];
    }
    if (min_e) *min_e = lo;
    if (max_e) *max_e = hi;
    return HYO_OK;
}

/* src/width.rs:107-137 — computed in 128-bit so no product can wrap. */
static __int128 magnitude_bound(uint32_t mode, int32_t min_e, int32_t max_e, int32_t go,
                                int32_t ge, uint64_t mql, uint64_t mtl) {
    __int128 m = (__int128)mql, n = (__int128)mtl;
    __int128 min_ij = m < n ? m : n;
    __int128 max_pos = max_e > 0 ? (__int128)max_e : 0;
    __int128 max_neg = min_e < 0 ? -(__int128)min_e : 0;
    __int128 positive = min_ij * max_pos;
    __int128 negative;
    if (mode == HYO_MODE_SW) {
        negative = go;
    } else if (mode == HYO_MODE_OV) {
        negative = min_ij * max_neg + go;
    } else {
        __int128 run = m + n - 1;
        if (run < 0) run = 0;
        negative = (m + n) * max_neg + go + run * ge;
    }
    return positive > negative ? positive : negative;
}

/* src/width.rs:145-169 with ScoreWidth::max_abs (width.rs:48-55) and narrowest_for (69-71). */
int hyo_required_width(uint32_t mode, int32_t min_e, int32_t max_e, int32_t go, int32_t ge,
                       uint64_t mql, uint64_t mtl, int64_t* bound_out, int32_t* width_bits) {
    __int128 b = magnitude_bound(mode, min_e, max_e, go, ge, mql, mtl);
    int64_t b64 = b > (__int128)INT64_MAX ? INT64_MAX : (int64_t)b;
    if (bound_out) *bound_out = b64;
    if (b64 <= 127) { *width_bits = 8; return HYO_OK; }
    if (b64 <= 32767) { *width_bits = 16; return HYO_OK; }
    if (b64 <= 536870911) { *width_bits = 32; return HYO_OK; }
    *width_bits = 0;
    return HYO_E_SCORE_RANGE;
}

/* src/kernel.rs:113-129 — maximise score, then minimise (grid_col, grid_row). */
typedef struct { int64_t score; uint64_t row, col; } hyo_best;

static void consider(hyo_best* b, int64_t s, uint64_t row, uint64_t col) {
    if (s > b->score || (s == b->score && (col < b->col || (col == b->col && row < b->row)))) {
        b->score = s; b->row = row; b->col = col;
    }
}

/* src/kernel.rs:168-217 (fill_dp) followed by src/kernel.rs:223-260 (reduce_answer).
 * h must hold (m+1)*(n+1) int32, f must hold n+1 int32. */
static void align_core_buf(const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n,
                           uint32_t al, const int32_t* matrix, int32_t go, int32_t ge,
                           uint32_t mode, int32_t* h, int32_t* f, int32_t* score,
                           int64_t* qend, int64_t* tend) {
    hyo_flags fl = flags_for_mode(mode);
    const uint64_t cols = n + 1;
#define H(i, j) h[(uint64_t)(i) * cols + (uint64_t)(j)]
    H(0, 0) = 0;
    for (uint64_t j = 1; j <= n; ++j) H(0, j) = fl.top_row_free ? 0 : -hyo_gap_penalty(go, ge, j);
    for (uint64_t i = 1; i <= m; ++i) H(i, 0) = fl.left_col_free ? 0 : -hyo_gap_penalty(go, ge, i);
    for (uint64_t j = 0; j <= n; ++j) f[j] = (int32_t)HYO_NEG;
    for (uint64_t i = 1; i <= m; ++i) {
        int64_t e = HYO_NEG;
        const int32_t* qrow = matrix + (uint64_t)q[i - 1] * al;
        for (uint64_t j = 1; j <= n; ++j) {
            int64_t e1 = (int64_t)H(i, j - 1) - go, e2 = e - ge;
            e = e1 > e2 ? e1 : e2;
            int64_t f1 = (int64_t)H(i - 1, j) - go, f2 = (int64_t)f[j] - ge;
            int64_t fj = f1 > f2 ? f1 : f2;
            f[j] = (int32_t)fj;
            int64_t diag = (int64_t)H(i - 1, j - 1) + qrow[t[j - 1]];
            int64_t cell = diag;
            if (e > cell) cell = e;
            if (fj > cell) cell = fj;
            if (fl.local && cell < 0) cell = 0;
            H(i, j) = (int32_t)cell;
        }
    }
    hyo_best b = {HYO_NEG, 0, 0};
    if (fl.local) {
        for (uint64_t i = 0; i <= m; ++i)
            for (uint64_t j = 0; j <= n; ++j) consider(&b, H(i, j), i, j);
    } else {
        consider(&b, H(m, n), m, n);
        if (fl.answer_last_row)
            for (uint64_t j = 0; j <= n; ++j) consider(&b, H(m, j), m, j);
        if (fl.answer_last_col)
            for (uint64_t i = 0; i <= m; ++i) consider(&b, H(i, n), i, n);
    }
#undef H
    *score = (int32_t)b.score;
    *qend = b.row == 0 ? -1 : (int64_t)b.row - 1;
    *tend = b.col == 0 ? -1 : (int64_t)b.col - 1;
}

/* src/kernel.rs:267-277 (align_core). */
int hyo_align_core(const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n, uint32_t al,
                   const int32_t* matrix, int32_t go, int32_t ge, uint32_t mode, int32_t* score,
                   int64_t* qend, int64_t* tend) {
    int32_t* h = (int32_t*)malloc(sizeof(int32_t) * (m + 1) * (n + 1));
    int32_t* f = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
    if (!h || !f) { free(h); free(f); return HYO_E_NOMEM; }
    align_core_buf(q, m, t, n, al, matrix, go, ge, mode, h, f, score, qend, tend);
    free(h);
    free(f);
    return HYO_OK;
}

/* ---------------------------------------------------------------------------------------- */
/* Brute force: tests/common/mod.rs:65-171. Exhaustive path enumeration for global alignment, */
/* then maxima over substring windows for the other modes. Exponential; tiny inputs only.     */
/* ---------------------------------------------------------------------------------------- */

#define BRUTE_NEG_INF INT64_MIN

typedef struct {
    const uint8_t* q; uint64_t m; const uint8_t* t; uint64_t n;
    const int32_t* mat; uint32_t al; int32_t go, ge;
} brute_prob;

enum { MV_START = 0, MV_RIGHT = 1, MV_DOWN = 2 };

static int64_t brute_rec(const brute_prob* p, uint64_t i, uint64_t j, int last) {
    if (i == p->m && j == p->n) return 0;
    int64_t best = BRUTE_NEG_INF;
    if (i < p->m && j < p->n) {
        int64_t r = brute_rec(p, i + 1, j + 1, MV_START);
        int64_t v = (int64_t)p->mat[(uint64_t)p->q[i] * p->al + p->t[j]] + r;
        if (v > best) best = v;
    }
    if (j < p->n) {
        int64_t cost = last == MV_RIGHT ? p->ge : p->go;
        int64_t v = -cost + brute_rec(p, i, j + 1, MV_RIGHT);
        if (v > best) best = v;
    }
    if (i < p->m) {
        int64_t cost = last == MV_DOWN ? p->ge : p->go;
        int64_t v = -cost + brute_rec(p, i + 1, j, MV_DOWN);
        if (v > best) best = v;
    }
    return best;
}

/* Global score of q[a:b] against t[c:d]. */
static int64_t brute_nw_window(const uint8_t* q, uint64_t a, uint64_t b, const uint8_t* t,
                               uint64_t c, uint64_t d, uint32_t al, const int32_t* mat,
                               int32_t go, int32_t ge) {
    brute_prob p = {q + a, b - a, t + c, d - c, mat, al, go, ge};
    return brute_rec(&p, 0, 0, MV_START);
}

int64_t hyo_brute_score(uint32_t mode, const uint8_t* q, uint64_t m, const uint8_t* t,
                        uint64_t n, uint32_t al, const int32_t* mat, int32_t go, int32_t ge) {
    int64_t best;
    switch (mode) {
    case HYO_MODE_NW:
        return brute_nw_window(q, 0, m, t, 0, n, al, mat, go, ge);
    case HYO_MODE_SW:
    case HYO_MODE_OV:
        best = 0;
        for (uint64_t a = 0; a <= m; ++a)
            for (uint64_t b = a; b <= m; ++b)
                for (uint64_t c = 0; c <= n; ++c)
                    for (uint64_t d = c; d <= n; ++d) {
                        if (mode == HYO_MODE_OV && !((a == 0 || c == 0) && (b == m || d == n)))
                            continue;
                        int64_t v = brute_nw_window(q, a, b, t, c, d, al, mat, go, ge);
                        if (v > best) best = v;
                    }
        return best;
    case HYO_MODE_HW:
        best = BRUTE_NEG_INF;
        for (uint64_t c = 0; c <= n; ++c)
            for (uint64_t d = c; d <= n; ++d) {
                int64_t v = brute_nw_window(q, 0, m, t, c, d, al, mat, go, ge);
                if (v > best) best = v;
            }
        return best;
    case HYO_MODE_SHW:
        best = BRUTE_NEG_INF;
        for (uint64_t a = 0; a <= m; ++a)
            for (uint64_t b = a; b <= m; ++b) {
                int64_t v = brute_nw_window(q, a, b, t, 0, n, al, mat, go, ge);
                if (v > best) best = v;
            }
        return best;
    default:
        return BRUTE_NEG_INF;
    }
}

/* Brute-force value of the DP cell (i, j): the best global score of q[a:i] against t[c:j]
 * over the start windows each mode's free borders allow (SW additionally floors at 0). This
 * extends the reference's whole-pair enumerators (tests/common/mod.rs:98-171) to every end cell
 * so that end positions can be pinned by enumeration too. */
static int64_t brute_cell(uint32_t mode, const uint8_t* q, uint64_t i, const uint8_t* t,
                          uint64_t j, uint32_t al, const int32_t* mat, int32_t go, int32_t ge) {
    hyo_flags fl = flags_for_mode(mode);
    int64_t best = fl.local ? 0 : BRUTE_NEG_INF;
    for (uint64_t a = 0; a <= i; ++a)
        for (uint64_t c = 0; c <= j; ++c) {
            int ok;
            if (fl.local) ok = 1;
            else if (fl.top_row_free && fl.left_col_free) ok = (a == 0 || c == 0);
            else if (fl.top_row_free) ok = (a == 0);
            else if (fl.left_col_free) ok = (c == 0);
            else ok = (a == 0 && c == 0);
            if (!ok) continue;
            int64_t v = brute_nw_window(q, a, i, t, c, j, al, mat, go, ge);
            if (v > best) best = v;
        }
    return best;
}

/* Score and ends by enumeration: brute-force cell values reduced over each mode's answer
 * region (src/kernel.rs:13-19) with the (score desc, col asc, row asc) tie-break. */
void hyo_brute_end(uint32_t mode, const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n,
                   uint32_t al, const int32_t* mat, int32_t go, int32_t ge, int32_t* score,
                   int64_t* qend, int64_t* tend) {
    hyo_flags fl = flags_for_mode(mode);
    hyo_best b = {BRUTE_NEG_INF, 0, 0};
    for (uint64_t i = 0; i <= m; ++i)
        for (uint64_t j = 0; j <= n; ++j) {
            int in;
            if (fl.local) in = 1;
            else in = (i == m && j == n) || (fl.answer_last_row && i == m) ||
                      (fl.answer_last_col && j == n);
            if (!in) continue;
            consider(&b, brute_cell(mode, q, i, t, j, al, mat, go, ge), i, j);
        }
    *score = (int32_t)b.score;
    *qend = b.row == 0 ? -1 : (int64_t)b.row - 1;
    *tend = b.col == 0 ? -1 : (int64_t)b.col - 1;
}

/* ---------------------------------------------------------------------------------------- */
/* Database level (scalar backend of src/database.rs).                                        */
/* The database is flat residues plus n+1 offsets.                                            */
/* ---------------------------------------------------------------------------------------- */

typedef struct {
    const uint8_t* flat; const uint64_t* off; uint64_t lo, hi;
    const uint8_t* q; uint64_t m; uint32_t al; const int32_t* mat; int32_t go, ge;
    uint32_t mode; int32_t* scores; int64_t* qends; int64_t* tends; int rc;
} scan_job;

static void* scan_worker(void* arg) {
    scan_job* jb = (scan_job*)arg;
    uint64_t maxn = 0;
    for (uint64_t k = jb->lo; k < jb->hi; ++k) {
        uint64_t len = jb->off[k + 1] - jb->off[k];
        if (len > maxn) maxn = len;
    }
    int32_t* h = (int32_t*)malloc(sizeof(int32_t) * (jb->m + 1) * (maxn + 1));
    int32_t* f = (int32_t*)malloc(sizeof(int32_t) * (maxn + 1));
    if (!h || !f) { free(h); free(f); jb->rc = HYO_E_NOMEM; return NULL; }
    for (uint64_t k = jb->lo; k < jb->hi; ++k) {
        int32_t s; int64_t qe, te;
        align_core_buf(jb->q, jb->m, jb->flat + jb->off[k], jb->off[k + 1] - jb->off[k], jb->al,
                       jb->mat, jb->go, jb->ge, jb->mode, h, f, &s, &qe, &te);
        jb->scores[k] = s;
        if (jb->qends) jb->qends[k] = qe;
        if (jb->tends) jb->tends[k] = te;
    }
    free(h);
    free(f);
    jb->rc = HYO_OK;
    return NULL;
}

/* Per-target (score, query_end, target_end) in database order: the scalar arm of
 * Database::scan_all (src/database.rs:405-414) and, with qends/tends NULL, of
 * Database::scan_scores (src/database.rs:511-513). Split over nthreads contiguous shards. */
int hyo_scan_all(const uint8_t* flat, const uint64_t* off, uint64_t n, const uint8_t* q,
                 uint64_t m, uint32_t al, const int32_t* mat, int32_t go, int32_t ge,
                 uint32_t mode, int32_t* scores, int64_t* qends, int64_t* tends, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > n) nthreads = (int)(n ? n : 1);
    scan_job* jobs = (scan_job*)calloc((size_t)nthreads, sizeof(scan_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    if (!jobs || !th) { free(jobs); free(th); return HYO_E_NOMEM; }
    /* Balance shards by residue count. */
    uint64_t total = off[n] - off[0], k = 0;
    for (int w = 0; w < nthreads; ++w) {
        uint64_t target = off[0] + (total * (uint64_t)(w + 1)) / (uint64_t)nthreads;
        uint64_t lo = k;
        if (w == nthreads - 1) k = n;
        else while (k < n && off[k + 1] <= target) ++k;
        scan_job j = {flat, off, lo, k, q, m, al, mat, go, ge, mode, scores, qends, tends, 0};
        jobs[w] = j;
    }
    for (int w = 0; w < nthreads; ++w) pthread_create(&th[w], NULL, scan_worker, &jobs[w]);
    int rc = HYO_OK;
    for (int w = 0; w < nthreads; ++w) {
        pthread_join(th[w], NULL);
        if (jobs[w].rc != HYO_OK) rc = jobs[w].rc;
    }
    free(jobs);
    free(th);
    return rc;
}

/* Database::scan on the scalar backend (src/database.rs:115-144): per-target align_core, then
 * the scalar argmax keeping the smallest index on a tie (strictly-greater replaces); ends are
 * reported only when the search type tracks them. */
int hyo_scan(const uint8_t* flat, const uint64_t* off, uint64_t n, const uint8_t* q, uint64_t m,
             uint32_t al, const int32_t* mat, int32_t go, int32_t ge, uint32_t mode,
             uint32_t tracks_end, int32_t* score, uint64_t* db_index, int64_t* qend,
             int64_t* tend, int nthreads) {
    if (n == 0) return HYO_E_EMPTY_DATABASE;
    int32_t* s = (int32_t*)malloc(sizeof(int32_t) * n);
    int64_t* qe = (int64_t*)malloc(sizeof(int64_t) * n);
    int64_t* te = (int64_t*)malloc(sizeof(int64_t) * n);
    if (!s || !qe || !te) { free(s); free(qe); free(te); return HYO_E_NOMEM; }
    int rc = hyo_scan_all(flat, off, n, q, m, al, mat, go, ge, mode, s, qe, te, nthreads);
    if (rc == HYO_OK) {
        int64_t best = INT64_MIN; uint64_t bi = 0;
        for (uint64_t k = 0; k < n; ++k)
            if ((int64_t)s[k] > best) { best = s[k]; bi = k; }
        *score = (int32_t)best;
        *db_index = bi;
        *qend = tracks_end ? qe[bi] : -1;
        *tend = tracks_end ? te[bi] : -1;
    }
    free(s); free(qe); free(te);
    return rc;
}

/* ---------------------------------------------------------------------------------------- */
/* Traceback: src/align.rs:169-420 (traceback, full-matrix DP, canonical backward walk) and   */
/* src/align.rs:230-290 (budget arithmetic). Full H/E/F matrices; the checkpoint path of the  */
/* reference returns byte-identical alignments, so one path suffices for the checker.         */
/* ---------------------------------------------------------------------------------------- */

static uint64_t sat_mul(uint64_t a, uint64_t b) {
    if (a != 0 && b > UINT64_MAX / a) return UINT64_MAX;
    return a * b;
}

static uint64_t full_matrix_bytes(uint64_t m, uint64_t n) {
    return sat_mul(sat_mul(sat_mul(m + 1, n + 1), 3), 4);
}

static uint64_t isqrt_u64(uint64_t v) {
    uint64_t r = 0;
    for (uint64_t bit = (uint64_t)1 << 62; bit; bit >>= 2) {
        if (v >= r + bit) { v -= r + bit; r = (r >> 1) + bit; } else { r >>= 1; }
    }
    return r;
}

static uint64_t traceback_min_bytes(uint64_t m, uint64_t n) {
    uint64_t full = full_matrix_bytes(m, n);
    if (m == 0 || n == 0) return full;
    uint64_t k = isqrt_u64(m);
    if (k < 1) k = 1;
    uint64_t ckpt = sat_mul(sat_mul(2 * (m / k + 1) + 3 * (k + 1), n + 1), 4);
    return full < ckpt ? full : ckpt;
}

uint64_t hyo_traceback_min_bytes_for_database(uint64_t max_m, uint64_t max_n) {
    uint64_t v = traceback_min_bytes(max_m, max_n);
    uint64_t a = full_matrix_bytes(max_m, 0), b = full_matrix_bytes(0, max_n);
    if (a > v) v = a;
    if (b > v) v = b;
    return v;
}

int hyo_traceback(const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n, uint32_t al,
                  const int32_t* matrix, int32_t go, int32_t ge, uint32_t mode, int32_t* score,
                  uint64_t* qstart, uint64_t* qend, uint64_t* tstart, uint64_t* tend,
                  uint8_t* ops, uint64_t* nops) {
    hyo_flags fl = flags_for_mode(mode);
    const uint64_t cols = n + 1, cells = (m + 1) * cols;
    int64_t* h = (int64_t*)malloc(sizeof(int64_t) * cells);
    int64_t* e = (int64_t*)malloc(sizeof(int64_t) * cells);
    int64_t* f = (int64_t*)malloc(sizeof(int64_t) * cells);
    if (!h || !e || !f) { free(h); free(e); free(f); return HYO_E_NOMEM; }
#define IX(i, j) ((uint64_t)(i) * cols + (uint64_t)(j))
    for (uint64_t k = 0; k < cells; ++k) { h[k] = 0; e[k] = HYO_NEG; f[k] = HYO_NEG; }
    /* Borders. E on row 0 and F on column 0 carry the penalised border gap so the walk can trace
     * and close it; both stay at the sentinel where the border is free. */
    for (uint64_t j = 1; j <= n; ++j) {
        h[IX(0, j)] = fl.top_row_free ? 0 : -(int64_t)hyo_gap_penalty(go, ge, j);
        e[IX(0, j)] = fl.top_row_free ? HYO_NEG : h[IX(0, j)];
    }
    for (uint64_t i = 1; i <= m; ++i) {
        h[IX(i, 0)] = fl.left_col_free ? 0 : -(int64_t)hyo_gap_penalty(go, ge, i);
        f[IX(i, 0)] = fl.left_col_free ? HYO_NEG : h[IX(i, 0)];
        e[IX(i, 0)] = HYO_NEG;
        const int32_t* qrow = matrix + (uint64_t)q[i - 1] * al;
        for (uint64_t j = 1; j <= n; ++j) {
            int64_t e1 = h[IX(i, j - 1)] - go, e2 = e[IX(i, j - 1)] - ge;
            e[IX(i, j)] = e1 > e2 ? e1 : e2;
            int64_t f1 = h[IX(i - 1, j)] - go, f2 = f[IX(i - 1, j)] - ge;
            f[IX(i, j)] = f1 > f2 ? f1 : f2;
            int64_t cell = h[IX(i - 1, j - 1)] + qrow[t[j - 1]];
            if (e[IX(i, j)] > cell) cell = e[IX(i, j)];
            if (f[IX(i, j)] > cell) cell = f[IX(i, j)];
            if (fl.local && cell < 0) cell = 0;
            h[IX(i, j)] = cell;
        }
    }
    /* End cell: the answer region with the (score, col, row) tie-break, row by row. */
    hyo_best b = {HYO_NEG, 0, 0};
    for (uint64_t i = 0; i <= m; ++i) {
        if (fl.local) {
            for (uint64_t j = 0; j <= n; ++j) consider(&b, h[IX(i, j)], i, j);
        } else {
            if (fl.answer_last_col) consider(&b, h[IX(i, n)], i, n);
            if (i == m) {
                consider(&b, h[IX(m, n)], m, n);
                if (fl.answer_last_row)
                    for (uint64_t j = 0; j <= n; ++j) consider(&b, h[IX(m, j)], m, j);
            }
        }
    }
    /* Backward walk: diagonal before a gap, a target-consuming gap (Del) before a query-consuming
     * one (Ins), closing a gap before extending it; local alignments stop at the first 0 cell. */
    enum { ST_H, ST_E, ST_F } st = ST_H;
    uint64_t i = b.row, j = b.col, k = 0;
    for (;;) {
        if (st == ST_H) {
            int64_t hij = h[IX(i, j)];
            if (fl.local && hij == 0) break;
            if (i == 0 && j == 0) break;
            if (i == 0) { if (fl.top_row_free) break; st = ST_E; continue; }
            if (j == 0) { if (fl.left_col_free) break; st = ST_F; continue; }
            int64_t sub = matrix[(uint64_t)q[i - 1] * al + t[j - 1]];
            if (hij == h[IX(i - 1, j - 1)] + sub) {
                ops[k++] = q[i - 1] == t[j - 1] ? HYO_OP_MATCH : HYO_OP_MISMATCH;
                --i; --j;
            } else if (hij == e[IX(i, j)]) {
                st = ST_E;
            } else {
                st = ST_F;
            }
        } else if (st == ST_E) {
            ops[k++] = HYO_OP_DEL;
            int64_t ev = e[IX(i, j)];
            --j;
            if (ev == h[IX(i, j)] - go) st = ST_H;
        } else {
            ops[k++] = HYO_OP_INS;
            int64_t fv = f[IX(i, j)];
            --i;
            if (fv == h[IX(i, j)] - go) st = ST_H;
        }
    }
#undef IX
    for (uint64_t a = 0, z = k; a + 1 < z; ++a, --z) {  /* forward order */
        uint8_t tmp = ops[a]; ops[a] = ops[z - 1]; ops[z - 1] = tmp;
    }
    *score = (int32_t)b.score;
    *qstart = i; *qend = b.row; *tstart = j; *tend = b.col; *nops = k;
    free(h); free(e); free(f);
    return HYO_OK;
}

/* ---------------------------------------------------------------------------------------- */
/* Pairwise extras: src/kernel.rs:474-575 (align_pair_position_max, scalar arm) and          */
/* src/kernel.rs:741-792 (score2).                                                            */
/* ---------------------------------------------------------------------------------------- */

/* Per-target-position maxima of the SW DP: out[t] = max over rows of H[i][t+1] (floor 0), the
 * best score = the array maximum, the target end = its first position (-1 when the best is 0). */
int hyo_position_max(const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n, uint32_t al,
                     const int32_t* matrix, int32_t go, int32_t ge, int32_t* out, int32_t* score,
                     int64_t* tend) {
    int64_t* h = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    int64_t* f = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    if (!h || !f) { free(h); free(f); return HYO_E_NOMEM; }
    for (uint64_t j = 0; j <= n; ++j) { h[j] = 0; f[j] = HYO_NEG; }
    for (uint64_t j = 0; j < n; ++j) out[j] = 0;
    for (uint64_t i = 1; i <= m; ++i) {
        int64_t diag = h[0], left = 0, ev = HYO_NEG;  /* H[i][0] = 0, E[i][0] = sentinel */
        const int32_t* row = matrix + (uint64_t)q[i - 1] * al;
        for (uint64_t j = 1; j <= n; ++j) {
            const int64_t e1 = left - go, e2 = ev - ge;
            ev = e1 > e2 ? e1 : e2;
            const int64_t f1 = h[j] - go, f2 = f[j] - ge;
            f[j] = f1 > f2 ? f1 : f2;
            int64_t c = diag + row[t[j - 1]];
            if (ev > c) c = ev;
            if (f[j] > c) c = f[j];
            if (c < 0) c = 0;
            diag = h[j];
            h[j] = c;
            left = c;
            if (c > out[j - 1]) out[j - 1] = (int32_t)c;
        }
    }
    int32_t best = 0;
    int64_t pos = -1;
    for (uint64_t j = 0; j < n; ++j)
        if (out[j] > best) { best = out[j]; pos = (int64_t)j; }
    *score = best;
    *tend = best > 0 ? pos : -1;
    free(h); free(f);
    return HYO_OK;
}

/* bwa's score2 recipe as the reference streams it: above-threshold columns contiguous with the
 * open peak's maximum position extend it, anything else closes it; a closed peak outside
 * [te - w, te + w], w = ceil(best / matrix_max), replaces the runner-up only when strictly
 * higher. Returns 1 with (score, pos) set, or 0 for None. */
int hyo_score2(const int32_t* colmax, uint64_t n, int32_t best_score, uint64_t best_tend,
               int32_t matrix_max, int32_t min_score, int32_t* s2, uint64_t* pos) {
    if (matrix_max <= 0) return 0;
    const int64_t bsc = best_score > 0 ? best_score : 0;
    int64_t w = (bsc + matrix_max - 1) / matrix_max;
    if (w < 0) w = 0;
    const int64_t lo = (int64_t)best_tend - w, hi = (int64_t)best_tend + w;
    int have2 = 0, open = 0;
    int32_t b2s = 0, cs = 0;
    uint64_t b2p = 0, cp = 0;
    for (uint64_t i = 0; i <= n; ++i) {
        const int close_now = i == n || (colmax[i] >= min_score && !(open && cp + 1 == i));
        if (i < n && colmax[i] < min_score) continue;
        if (open && close_now) {
            const int64_t p = (int64_t)cp;
            if ((p < lo || p > hi) && (!have2 || cs > b2s)) { have2 = 1; b2s = cs; b2p = cp; }
            open = 0;
        }
        if (i == n) break;
        if (open) {
            if (cs < colmax[i]) { cs = colmax[i]; cp = i; }
        } else {
            open = 1; cs = colmax[i]; cp = i;
        }
    }
    if (have2) { *s2 = b2s; *pos = b2p; }
    return have2;
}

/* Brute-force column maxima: out[t] = max over query ends i of the best local alignment ending
 * exactly at cell (i, t+1), by enumerating start windows (tiny inputs only). */
void hyo_brute_colmax(const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n, uint32_t al,
                      const int32_t* mat, int32_t go, int32_t ge, int32_t* out) {
    for (uint64_t j = 1; j <= n; ++j) {
        int64_t best = 0;
        for (uint64_t i = 0; i <= m; ++i) {
            const int64_t v = brute_cell(HYO_MODE_SW, q, i, t, j, al, mat, go, ge);
            if (v > best) best = v;
        }
        out[j - 1] = (int32_t)best;
    }
}
