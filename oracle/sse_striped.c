/*
 * SSE4.1 striped (Farrar) CPU baseline for single-pair alignment — TEST / BENCHMARK
 * INFRASTRUCTURE ONLY.
 *
 * A C restatement of the shape of hyalite's striped single-pair path (src/striped.rs:1-429), the
 * code align_pair / align_pairs run on x86-64 for SearchType::Score when the width proof gives
 * i16: the query striped over 8 x i16 lanes (segLen = ceil(m / 8), lane l holding query rows
 * l * segLen + k), a per-symbol striped profile built once per query, a horizontally vectorised
 * column sweep with saturating arithmetic and Farrar's lazy-F correction for the vertical gaps
 * that cross stripe boundaries. Local mode (SW) only, with every quantity floored at 0 (the SW
 * clamp), which is the form the benchmark's pair batches need; the other modes stay on the scalar
 * restatement. Pairs are aligned one after another on one thread, as align_pairs does
 * (src/kernel.rs:429-452). Checked against the scalar oracle before it is timed.
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

#define SL 8  /* i16 lanes per 128-bit register */

static int16_t sat16i(int32_t v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : (int16_t)v); }

typedef struct {
    __m128i* prof;   /* al * seg vectors */
    __m128i* hs;     /* H of the current column */
    __m128i* hl;     /* H of the previous column */
    __m128i* e;      /* E of the next column */
    size_t cap_prof, cap_seg;
} sse_bufs;

static int grow(sse_bufs* b, size_t al, size_t seg) {
    if (al * seg > b->cap_prof) {
        free(b->prof);
        b->prof = (__m128i*)aligned_alloc(16, sizeof(__m128i) * al * seg);
        if (!b->prof) return -1;
        b->cap_prof = al * seg;
    }
    if (seg > b->cap_seg) {
        free(b->hs); free(b->hl); free(b->e);
        b->hs = (__m128i*)aligned_alloc(16, sizeof(__m128i) * seg);
        b->hl = (__m128i*)aligned_alloc(16, sizeof(__m128i) * seg);
        b->e = (__m128i*)aligned_alloc(16, sizeof(__m128i) * seg);
        if (!b->hs || !b->hl || !b->e) return -1;
        b->cap_seg = seg;
    }
    return 0;
}

/* Best local score of q against t (both non-empty), i16 saturating model. */
static int32_t striped_sw(sse_bufs* b, const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n,
                          uint32_t al, const int32_t* mat, int32_t go, int32_t ge) {
    const size_t seg = (size_t)((m + SL - 1) / SL);
    if (grow(b, al, seg) != 0) return INT32_MIN;
    /* Profile: prof[s * seg + k] lane l = sub(q[l * seg + k], s); rows past m get the lowest
     * value, so their cells only carry decayed gaps of real cells. */
    for (uint32_t s = 0; s < al; ++s) {
        for (size_t k = 0; k < seg; ++k) {
            int16_t v[SL];
            for (int l = 0; l < SL; ++l) {
                const uint64_t i = (uint64_t)l * seg + k;
                v[l] = i < m ? sat16i(mat[(uint64_t)q[i] * al + s]) : (int16_t)-32768;
            }
            b->prof[s * seg + k] = _mm_loadu_si128((const __m128i*)v);
        }
    }
    const __m128i zero = _mm_setzero_si128();
    const __m128i vgo = _mm_set1_epi16((int16_t)(go > 32767 ? 32767 : go));
    const __m128i vge = _mm_set1_epi16((int16_t)(ge > 32767 ? 32767 : ge));
    for (size_t k = 0; k < seg; ++k) b->hs[k] = b->hl[k] = b->e[k] = zero;
    __m128i vmax = zero;
    for (uint64_t j = 0; j < n; ++j) {
        const __m128i* pv = b->prof + (size_t)t[j] * seg;
        __m128i vf = zero;
        /* H of the row above lane 0's first row is the previous column's last stripe, shifted
         * one lane up, with the border 0 entering lane 0. */
        __m128i vh = _mm_slli_si128(b->hs[seg - 1], 2);
        __m128i* tmp = b->hl; b->hl = b->hs; b->hs = tmp;
        for (size_t k = 0; k < seg; ++k) {
            vh = _mm_adds_epi16(vh, pv[k]);
            vh = _mm_max_epi16(vh, b->e[k]);
            vh = _mm_max_epi16(vh, vf);
            vh = _mm_max_epi16(vh, zero);
            vmax = _mm_max_epi16(vmax, vh);
            b->hs[k] = vh;
            const __m128i hg = _mm_subs_epu16(vh, vgo);
            b->e[k] = _mm_max_epi16(_mm_subs_epu16(b->e[k], vge), hg);
            vf = _mm_max_epi16(_mm_subs_epu16(vf, vge), hg);
            vh = b->hl[k];
        }
        /* Lazy F: carry F across stripe boundaries while it can still raise an H. */
        vf = _mm_slli_si128(vf, 2);
        size_t k = 0;
        for (;;) {
            const __m128i lim = _mm_subs_epu16(b->hs[k], vgo);
            if (_mm_movemask_epi8(_mm_cmpgt_epi16(vf, lim)) == 0) break;
            const __m128i h = _mm_max_epi16(b->hs[k], vf);
            b->hs[k] = h;
            vmax = _mm_max_epi16(vmax, h);
            b->e[k] = _mm_max_epi16(b->e[k], _mm_subs_epu16(h, vgo));
            vf = _mm_subs_epu16(vf, vge);
            if (++k == seg) {
                k = 0;
                vf = _mm_slli_si128(vf, 2);
            }
        }
    }
    int16_t lanes[SL];
    _mm_storeu_si128((__m128i*)lanes, vmax);
    int32_t best = 0;
    for (int l = 0; l < SL; ++l) best = lanes[l] > best ? lanes[l] : best;
    return best;
}

/* Scores of n pairs (query k = qflat[qoff[k] .. qoff[k+1]), target likewise) and the seconds
 * the loop took. Empty sides score 0 (the SW empty alignment). Returns 0, or -1 on allocation
 * failure. */
int hyo_sse_striped_sw_pairs(const uint8_t* qflat, const uint64_t* qoff, const uint8_t* tflat,
                             const uint64_t* toff, uint64_t npairs, uint32_t al,
                             const int32_t* mat, int32_t go, int32_t ge, int32_t* scores,
                             double* seconds) {
    sse_bufs b;
    memset(&b, 0, sizeof(b));
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int rc = 0;
    for (uint64_t k = 0; k < npairs; ++k) {
        const uint64_t m = qoff[k + 1] - qoff[k], n = toff[k + 1] - toff[k];
        if (m == 0 || n == 0) { scores[k] = 0; continue; }
        const int32_t s = striped_sw(&b, qflat + qoff[k], m, tflat + toff[k], n, al, mat, go, ge);
        if (s == INT32_MIN) { rc = -1; break; }
        scores[k] = s;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    free(b.prof); free(b.hs); free(b.hl); free(b.e);
    return rc;
}
