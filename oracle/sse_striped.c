/*
 * SSE4.1 striped (Farrar) CPU baseline for single-pair alignment — TEST / BENCHMARK
 * INFRASTRUCTURE ONLY.
 *
 * A C restatement of the shape of hyalite's striped single-pair path (src/striped.rs:1-429), the
 * code align_pair / align_pairs run on x86-64 for SearchType::Score at the width the proof gives
 * (16 x i8 or 8 x i16 lanes per 128-bit register): the query striped over the lanes
 * (segLen = ceil(m / lanes), lane l holding query rows l * segLen + k), a per-symbol striped
 * profile built once per query, a horizontally vectorised
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

static int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

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

/*
 * One striped local-mode kernel per element width. LANES elements per register; ELEM the element
 * type; the intrinsics are the width's saturating add, unsigned saturating subtract (every SW
 * quantity is >= 0, so it floors gaps at 0), signed max and compare, and the element shift.
 * Best local score of q against t (both non-empty); INT32_MIN on allocation failure.
 */
#define STRIPED_SW(NAME, LANES, ELEM, EMIN, EMAX, SET1, ADDS, SUBSU, MAX, CMPGT, SHIFT1)          \
static int32_t NAME(sse_bufs* b, const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n,      \
                    uint32_t al, const int32_t* mat, int32_t go, int32_t ge) {                    \
    const size_t seg = (size_t)((m + LANES - 1) / LANES);                                         \
    if (grow(b, al, seg) != 0) return INT32_MIN;                                                  \
    /* Profile: prof[s * seg + k] lane l = sub(q[l * seg + k], s); rows past m get the lowest   \
     * value, so their cells only carry decayed gaps of real cells. */                           \
    for (uint32_t s = 0; s < al; ++s) {                                                           \
        for (size_t k = 0; k < seg; ++k) {                                                        \
            ELEM v[LANES];                                                                        \
            for (int l = 0; l < LANES; ++l) {                                                     \
                const uint64_t i = (uint64_t)l * seg + k;                                         \
                v[l] = i < m ? (ELEM)clampi(mat[(uint64_t)q[i] * al + s], EMIN, EMAX) : (ELEM)EMIN;\
            }                                                                                     \
            b->prof[s * seg + k] = _mm_loadu_si128((const __m128i*)v);                            \
        }                                                                                         \
    }                                                                                             \
    const __m128i zero = _mm_setzero_si128();                                                     \
    const __m128i vgo = SET1((ELEM)clampi(go, 0, EMAX));                                          \
    const __m128i vge = SET1((ELEM)clampi(ge, 0, EMAX));                                          \
    for (size_t k = 0; k < seg; ++k) b->hs[k] = b->hl[k] = b->e[k] = zero;                        \
    __m128i vmax = zero;                                                                          \
    for (uint64_t j = 0; j < n; ++j) {                                                            \
        const __m128i* pv = b->prof + (size_t)t[j] * seg;                                         \
        __m128i vf = zero;                                                                        \
        /* H of the row above lane 0's first row is the previous column's last stripe, shifted  \
         * one lane up, with the border 0 entering lane 0. */                                    \
        __m128i vh = SHIFT1(b->hs[seg - 1]);                                                      \
        __m128i* tmp = b->hl; b->hl = b->hs; b->hs = tmp;                                         \
        for (size_t k = 0; k < seg; ++k) {                                                        \
            vh = ADDS(vh, pv[k]);                                                                 \
            vh = MAX(vh, b->e[k]);                                                                \
            vh = MAX(vh, vf);                                                                     \
            vh = MAX(vh, zero);                                                               \
            vmax = MAX(vmax, vh);                                                                 \
            b->hs[k] = vh;                                                                        \
            const __m128i hg = SUBSU(vh, vgo);                                                    \
            b->e[k] = MAX(SUBSU(b->e[k], vge), hg);                                               \
            vf = MAX(SUBSU(vf, vge), hg);                                                         \
            vh = b->hl[k];                                                                        \
        }                                                                                         \
        /* Lazy F: carry F across stripe boundaries while it can still raise an H. */            \
        vf = SHIFT1(vf);                                                                          \
        size_t k = 0;                                                                             \
        for (;;) {                                                                                \
            const __m128i lim = SUBSU(b->hs[k], vgo);                                             \
            if (_mm_movemask_epi8(CMPGT(vf, lim)) == 0) break;                                    \
            const __m128i h = MAX(b->hs[k], vf);                                                  \
            b->hs[k] = h;                                                                         \
            vmax = MAX(vmax, h);                                                                  \
            b->e[k] = MAX(b->e[k], SUBSU(h, vgo));                                                \
            vf = SUBSU(vf, vge);                                                                  \
            if (++k == seg) {                                                                     \
                k = 0;                                                                            \
                vf = SHIFT1(vf);                                                                  \
            }                                                                                     \
        }                                                                                         \
    }                                                                                             \
    ELEM lanes[LANES];                                                                            \
    _mm_storeu_si128((__m128i*)lanes, vmax);                                                      \
    int32_t best = 0;                                                                             \
    for (int l = 0; l < LANES; ++l) best = lanes[l] > best ? lanes[l] : best;                     \
    return best;                                                                                  \
}

static __m128i shl_i16(__m128i v) { return _mm_slli_si128(v, 2); }
static __m128i shl_i8(__m128i v) { return _mm_slli_si128(v, 1); }

STRIPED_SW(striped_sw16, 8, int16_t, -32768, 32767, _mm_set1_epi16, _mm_adds_epi16,
           _mm_subs_epu16, _mm_max_epi16, _mm_cmpgt_epi16, shl_i16)
STRIPED_SW(striped_sw8, 16, int8_t, -128, 127, _mm_set1_epi8, _mm_adds_epi8, _mm_subs_epu8,
           _mm_max_epi8, _mm_cmpgt_epi8, shl_i8)

/* The width the reference's proof gives a local-mode pair (src/width.rs): the larger of
 * min(m, n) * max(0, largest entry) and gap_open; 8 bits up to 127, 16 up to 32767, else 32. */
static int sw_width(uint64_t m, uint64_t n, int32_t max_e, int32_t go) {
    const int64_t pos = (int64_t)(m < n ? m : n) * (int64_t)(max_e > 0 ? max_e : 0);
    const int64_t b = pos > go ? pos : go;
    return b <= 127 ? 8 : (b <= 32767 ? 16 : 32);
}

/* Scores of n pairs (query k = qflat[qoff[k] .. qoff[k+1]), target likewise) and the seconds
 * the loop took, each pair at the width its proof gives (16 x i8 or 8 x i16). Empty sides score
 * 0 (the SW empty alignment). Returns 0, -1 on allocation failure, -2 when a pair needs 32 bits
 * (the reference runs those elsewhere; this baseline does not). */
int hyo_sse_striped_sw_pairs(const uint8_t* qflat, const uint64_t* qoff, const uint8_t* tflat,
                             const uint64_t* toff, uint64_t npairs, uint32_t al,
                             const int32_t* mat, int32_t go, int32_t ge, int32_t* scores,
                             double* seconds) {
    sse_bufs b;
    memset(&b, 0, sizeof(b));
    int32_t max_e = INT32_MIN;
    for (uint64_t k = 0; k < (uint64_t)al * al; ++k) max_e = mat[k] > max_e ? mat[k] : max_e;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int rc = 0;
    for (uint64_t k = 0; k < npairs; ++k) {
        const uint64_t m = qoff[k + 1] - qoff[k], n = toff[k + 1] - toff[k];
        if (m == 0 || n == 0) { scores[k] = 0; continue; }
        const int wbits = sw_width(m, n, max_e, go);
        if (wbits == 32) { rc = -2; break; }
        const int32_t s = wbits == 8
            ? striped_sw8(&b, qflat + qoff[k], m, tflat + toff[k], n, al, mat, go, ge)
            : striped_sw16(&b, qflat + qoff[k], m, tflat + toff[k], n, al, mat, go, ge);
        if (s == INT32_MIN) { rc = -1; break; }
        scores[k] = s;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    free(b.prof); free(b.hs); free(b.hl); free(b.e);
    return rc;
}
