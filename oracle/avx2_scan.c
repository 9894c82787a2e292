/*
 * AVX2 CPU baseline for the database scan — TEST / BENCHMARK INFRASTRUCTURE ONLY.
 *
 * A C restatement of the shape of hyalite's avx2 i16 inter-sequence backend, used as the measured
 * CPU baseline (the reference is Rust and cannot be built in this image):
 *   - 16 x i16 lanes per 256-bit register, saturating add/sub (src/inter.rs:2220-2415, Avx2I16);
 *   - targets batched 16 at a time in database order (PackedDb::build, src/inter.rs:431-508);
 *     optionally length-sorted ("best-effort" variant);
 *   - the Precomputed substitution table table[(qsym * w + j) * 16 + lane] and per-column
 *     mask_le / mask_eq vectors, built once per database before any scan (src/inter.rs:466-474);
 *   - the Score sweep of scan_batch (src/inter.rs:683-831): two H rows ping-ponged, one F row,
 *     scalar E per lane vector, mode answer via masks.
 * Only databases whose width proof gives at most 16 bits are accepted (as the reference's i16
 * path). Each call returns per-target scores in database order; tests/test_oracle.py checks them
 * against the scalar oracle.
 */
#include <immintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

#define LANES 16

typedef struct {
    uint32_t idx[LANES];  /* database indices of the batch's targets */
    uint32_t len[LANES];
    int real;             /* live lanes */
    uint32_t w;           /* batch width (longest target) */
    int16_t* table;       /* al * w * LANES */
    int16_t* mask_le;     /* (w + 1) * LANES: -1 where col <= len */
    int16_t* mask_eq;     /* (w + 1) * LANES: -1 where col == len */
} batch_t;

static int16_t sat16(int64_t v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : (int16_t)v); }

static int32_t gap_pen(int32_t go, int32_t ge, int64_t len) {
    return len == 0 ? 0 : (int32_t)((int64_t)go + (len - 1) * (int64_t)ge);
}

typedef struct {
    const batch_t* b; size_t lo, hi;
    const uint8_t* q; uint64_t m; int32_t go, ge; uint32_t mode;
    int32_t* scores; double seconds;
} job_t;

static void scan_batch16(const batch_t* bt, const uint8_t* q, uint64_t m, int32_t go, int32_t ge,
                         uint32_t mode, int16_t* h, int16_t* f, int32_t* out) {
    const int top_free = mode == HYO_MODE_HW || mode == HYO_MODE_OV || mode == HYO_MODE_SW;
    const int left_free = mode == HYO_MODE_OV || mode == HYO_MODE_SW || mode == HYO_MODE_SHW;
    const int last_row = mode == HYO_MODE_HW || mode == HYO_MODE_OV;
    const int last_col = mode == HYO_MODE_OV || mode == HYO_MODE_SHW;
    const int local = mode == HYO_MODE_SW;
    const uint32_t w = bt->w;
    const size_t cols = (size_t)(w + 1) * LANES;
    int16_t* prev = h;
    int16_t* cur = h + cols;
    const __m256i go_v = _mm256_set1_epi16(sat16(go)), ge_v = _mm256_set1_epi16(sat16(ge));
    const __m256i zero = _mm256_setzero_si256(), neg = _mm256_set1_epi16(-32768);
    for (uint32_t j = 0; j <= w; ++j) {
        const __m256i border = top_free ? zero : _mm256_set1_epi16(sat16(-(int64_t)gap_pen(go, ge, j)));
        _mm256_storeu_si256((__m256i*)(prev + (size_t)j * LANES), border);
        _mm256_storeu_si256((__m256i*)(f + (size_t)j * LANES), neg);
    }
    __m256i sw_ans = zero, lastcol = zero;
    if (last_col) {
        lastcol = neg;
        for (uint32_t j = 0; j <= w; ++j) {
            const __m256i mk = _mm256_loadu_si256((const __m256i*)(bt->mask_eq + (size_t)j * LANES));
            const __m256i v = _mm256_loadu_si256((const __m256i*)(prev + (size_t)j * LANES));
            lastcol = _mm256_blendv_epi8(lastcol, _mm256_max_epi16(lastcol, v), mk);
        }
    }
    for (uint64_t i = 1; i <= m; ++i) {
        const __m256i border = left_free ? zero : _mm256_set1_epi16(sat16(-(int64_t)gap_pen(go, ge, i)));
        _mm256_storeu_si256((__m256i*)cur, border);
        const int16_t* trow = bt->table + (size_t)q[i - 1] * w * LANES;
        __m256i e = neg;
        for (uint32_t j = 1; j <= w; ++j) {
            const size_t col = (size_t)j * LANES, pc = (size_t)(j - 1) * LANES;
            e = _mm256_max_epi16(_mm256_subs_epi16(_mm256_loadu_si256((const __m256i*)(cur + pc)), go_v),
                                 _mm256_subs_epi16(e, ge_v));
            const __m256i fj = _mm256_max_epi16(
                _mm256_subs_epi16(_mm256_loadu_si256((const __m256i*)(prev + col)), go_v),
                _mm256_subs_epi16(_mm256_loadu_si256((const __m256i*)(f + col)), ge_v));
            _mm256_storeu_si256((__m256i*)(f + col), fj);
            const __m256i sub = _mm256_loadu_si256((const __m256i*)(trow + pc));
            const __m256i diag = _mm256_adds_epi16(_mm256_loadu_si256((const __m256i*)(prev + pc)), sub);
            __m256i cell = _mm256_max_epi16(diag, _mm256_max_epi16(e, fj));
            if (local) {
                cell = _mm256_max_epi16(cell, zero);
                const __m256i mk = _mm256_loadu_si256((const __m256i*)(bt->mask_le + col));
                sw_ans = _mm256_blendv_epi8(sw_ans, _mm256_max_epi16(sw_ans, cell), mk);
            }
            if (last_col) {
                const __m256i mk = _mm256_loadu_si256((const __m256i*)(bt->mask_eq + col));
                lastcol = _mm256_blendv_epi8(lastcol, _mm256_max_epi16(lastcol, cell), mk);
            }
            _mm256_storeu_si256((__m256i*)(cur + col), cell);
        }
        int16_t* t = prev; prev = cur; cur = t;
    }
    __m256i ans;
    if (local) {
        ans = sw_ans;
    } else if (last_row) {
        __m256i r = _mm256_loadu_si256((const __m256i*)prev);
        for (uint32_t j = 1; j <= w; ++j) {
            const __m256i mk = _mm256_loadu_si256((const __m256i*)(bt->mask_le + (size_t)j * LANES));
            r = _mm256_blendv_epi8(r, _mm256_max_epi16(r, _mm256_loadu_si256((const __m256i*)(prev + (size_t)j * LANES))), mk);
        }
        ans = last_col ? _mm256_max_epi16(r, lastcol) : r;
    } else if (last_col) {
        ans = lastcol;
    } else {
        __m256i r = _mm256_loadu_si256((const __m256i*)prev);
        for (uint32_t j = 1; j <= w; ++j) {
            const __m256i mk = _mm256_loadu_si256((const __m256i*)(bt->mask_eq + (size_t)j * LANES));
            r = _mm256_blendv_epi8(r, _mm256_loadu_si256((const __m256i*)(prev + (size_t)j * LANES)), mk);
        }
        ans = r;
    }
    int16_t a[LANES];
    _mm256_storeu_si256((__m256i*)a, ans);
    for (int k = 0; k < bt->real; ++k) out[bt->idx[k]] = a[k];
}

static void* worker(void* arg) {
    job_t* jb = (job_t*)arg;
    uint32_t maxw = 0;
    for (size_t b = jb->lo; b < jb->hi; ++b) if (jb->b[b].w > maxw) maxw = jb->b[b].w;
    int16_t* h = (int16_t*)aligned_alloc(64, sizeof(int16_t) * 2 * (maxw + 1) * LANES + 64);
    int16_t* f = (int16_t*)aligned_alloc(64, sizeof(int16_t) * (maxw + 1) * LANES + 64);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (size_t b = jb->lo; b < jb->hi; ++b)
        scan_batch16(&jb->b[b], jb->q, jb->m, jb->go, jb->ge, jb->mode, h, f, jb->scores);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    jb->seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    free(h);
    free(f);
    return NULL;
}

/* Per-target scores of the AVX2 baseline. sort != 0 batches length-sorted targets instead of
 * database order. *scan_seconds (may be NULL) receives the wall time of the scan itself (table
 * construction, which the reference does once at database build, is excluded). */
int hyo_avx2_scan_scores_timed(const uint8_t* flat, const uint64_t* off, uint64_t n,
                               const uint8_t* q, uint64_t m, uint32_t al, const int32_t* mat,
                               int32_t go, int32_t ge, uint32_t mode, int32_t* scores,
                               int nthreads, int sort, double* scan_seconds) {
    int32_t min_e = mat[0], max_e = mat[0];
    for (uint64_t k = 1; k < (uint64_t)al * al; ++k) {
        if (mat[k] < min_e) min_e = mat[k];
        if (mat[k] > max_e) max_e = mat[k];
    }
    uint64_t mtl = 0;
    for (uint64_t k = 0; k < n; ++k) if (off[k + 1] - off[k] > mtl) mtl = off[k + 1] - off[k];
    int64_t bound; int32_t bits;
    if (hyo_required_width(mode, min_e, max_e, go, ge, m, mtl, &bound, &bits) != HYO_OK || bits > 16)
        return HYO_E_SCORE_RANGE;
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint64_t k = 0; k < n; ++k) order[k] = (uint32_t)k;
    if (sort) { /* stable insertion-free: counting sort by length, descending */
        uint64_t* cnt = (uint64_t*)calloc(mtl + 2, sizeof(uint64_t));
        for (uint64_t k = 0; k < n; ++k) cnt[mtl - (off[k + 1] - off[k])]++;
        uint64_t acc = 0;
        for (uint64_t v = 0; v <= mtl; ++v) { uint64_t c = cnt[v]; cnt[v] = acc; acc += c; }
        for (uint64_t k = 0; k < n; ++k) order[cnt[mtl - (off[k + 1] - off[k])]++] = (uint32_t)k;
        free(cnt);
    }
    const size_t nb = (size_t)((n + LANES - 1) / LANES);
    batch_t* bs = (batch_t*)calloc(nb ? nb : 1, sizeof(batch_t));
    for (size_t b = 0; b < nb; ++b) {
        batch_t* bt = &bs[b];
        bt->real = (int)((n - b * LANES) < LANES ? (n - b * LANES) : LANES);
        bt->w = 0;
        for (int k = 0; k < bt->real; ++k) {
            const uint32_t gi = order[b * LANES + k];
            bt->idx[k] = gi;
            bt->len[k] = (uint32_t)(off[gi + 1] - off[gi]);
            if (bt->len[k] > bt->w) bt->w = bt->len[k];
        }
        const uint32_t w = bt->w;
        bt->table = (int16_t*)malloc(sizeof(int16_t) * ((size_t)al * w * LANES + 1));
        bt->mask_le = (int16_t*)malloc(sizeof(int16_t) * (w + 1) * LANES);
        bt->mask_eq = (int16_t*)malloc(sizeof(int16_t) * (w + 1) * LANES);
        for (uint32_t j = 0; j <= w; ++j)
            for (int k = 0; k < LANES; ++k) {
                const uint32_t L = k < bt->real ? bt->len[k] : 0;
                bt->mask_le[(size_t)j * LANES + k] = (k < bt->real && j <= L) ? -1 : 0;
                bt->mask_eq[(size_t)j * LANES + k] = (k < bt->real && j == L) ? -1 : 0;
            }
        for (uint32_t s = 0; s < al; ++s)
            for (uint32_t j = 0; j < w; ++j)
                for (int k = 0; k < LANES; ++k) {
                    int16_t v = 0;
                    if (k < bt->real && j < bt->len[k]) {
                        const uint8_t t = flat[off[bt->idx[k]] + j];
                        v = sat16(mat[(size_t)s * al + t]);
                    }
                    bt->table[((size_t)s * w + j) * LANES + k] = v;
                }
    }
    if (nthreads < 1) nthreads = 1;
    if ((size_t)nthreads > nb) nthreads = (int)(nb ? nb : 1);
    job_t* jobs = (job_t*)calloc((size_t)nthreads, sizeof(job_t));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    uint64_t total = 0;
    for (size_t b = 0; b < nb; ++b) total += (uint64_t)bs[b].w + 1;
    size_t b = 0;
    for (int t = 0; t < nthreads; ++t) {
        const uint64_t target = total * (uint64_t)(t + 1) / (uint64_t)nthreads;
        const size_t lo = b;
        uint64_t acc = 0;
        for (size_t x = 0; x < lo; ++x) acc += (uint64_t)bs[x].w + 1;
        if (t == nthreads - 1) b = nb;
        else while (b < nb && acc + bs[b].w + 1 <= target) { acc += bs[b].w + 1; ++b; }
        job_t j = {bs, lo, b, q, m, go, ge, mode, scores, 0.0};
        jobs[t] = j;
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, worker, &jobs[t]);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (scan_seconds)
        *scan_seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    for (size_t x = 0; x < nb; ++x) { free(bs[x].table); free(bs[x].mask_le); free(bs[x].mask_eq); }
    free(bs); free(jobs); free(th); free(order);
    return HYO_OK;
}

int hyo_avx2_scan_scores(const uint8_t* flat, const uint64_t* off, uint64_t n, const uint8_t* q,
                         uint64_t m, uint32_t al, const int32_t* mat, int32_t go, int32_t ge,
                         uint32_t mode, int32_t* scores, int nthreads, int sort) {
    return hyo_avx2_scan_scores_timed(flat, off, n, q, m, al, mat, go, ge, mode, scores, nthreads,
                                      sort, NULL);
}
