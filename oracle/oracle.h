/* CPU oracle for the database scan — TEST INFRASTRUCTURE ONLY (see oracle.c). */
#ifndef HYALITE_ORACLE_H
#define HYALITE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mode codes (src/mode.rs:11-27); the same numbering as the product C ABI. */
enum { HYO_MODE_SW = 0, HYO_MODE_NW = 1, HYO_MODE_HW = 2, HYO_MODE_OV = 3, HYO_MODE_SHW = 4 };

/* Status codes; the same numbering as the product C ABI's construction errors. */
enum {
    HYO_OK = 0,
    HYO_E_EMPTY_ALPHABET = 1,
    HYO_E_MATRIX_SHAPE = 2,
    HYO_E_NEGATIVE_GAP = 3,
    HYO_E_GAP_OPEN_LT_EXT = 4,
    HYO_E_SCORE_RANGE = 5,
    HYO_E_EMPTY_DATABASE = 8,
    HYO_E_NOMEM = 100
};

/* Alignment operations (src/align.rs:47-57), in declaration order. */
enum { HYO_OP_MATCH = 0, HYO_OP_MISMATCH = 1, HYO_OP_INS = 2, HYO_OP_DEL = 3 };

int32_t hyo_gap_penalty(int32_t go, int32_t ge, uint64_t len);
int hyo_scoring_check(uint32_t al, const int32_t* matrix, uint64_t matrix_len, int32_t go,
                      int32_t ge, int32_t* min_e, int32_t* max_e);
int hyo_required_width(uint32_t mode, int32_t min_e, int32_t max_e, int32_t go, int32_t ge,
                       uint64_t mql, uint64_t mtl, int64_t* bound_out, int32_t* width_bits);
int hyo_align_core(const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n, uint32_t al,
                   const int32_t* matrix, int32_t go, int32_t ge, uint32_t mode, int32_t* score,
                   int64_t* qend, int64_t* tend);
int64_t hyo_brute_score(uint32_t mode, const uint8_t* q, uint64_t m, const uint8_t* t,
                        uint64_t n, uint32_t al, const int32_t* mat, int32_t go, int32_t ge);
void hyo_brute_end(uint32_t mode, const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n,
                   uint32_t al, const int32_t* mat, int32_t go, int32_t ge, int32_t* score,
                   int64_t* qend, int64_t* tend);
int hyo_scan_all(const uint8_t* flat, const uint64_t* off, uint64_t n, const uint8_t* q,
                 uint64_t m, uint32_t al, const int32_t* mat, int32_t go, int32_t ge,
                 uint32_t mode, int32_t* scores, int64_t* qends, int64_t* tends, int nthreads);
int hyo_scan(const uint8_t* flat, const uint64_t* off, uint64_t n, const uint8_t* q, uint64_t m,
             uint32_t al, const int32_t* mat, int32_t go, int32_t ge, uint32_t mode,
             uint32_t tracks_end, int32_t* score, uint64_t* db_index, int64_t* qend,
             int64_t* tend, int nthreads);

/* src/align.rs:169-420: the canonical alignment (ops in forward order, at most m + n of them;
 * spans half-open). */
int hyo_traceback(const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n, uint32_t al,
                  const int32_t* matrix, int32_t go, int32_t ge, uint32_t mode, int32_t* score,
                  uint64_t* qstart, uint64_t* qend, uint64_t* tstart, uint64_t* tend,
                  uint8_t* ops, uint64_t* nops);
/* src/align.rs:230-275 (traceback_min_bytes_for_database). */
uint64_t hyo_traceback_min_bytes_for_database(uint64_t max_m, uint64_t max_n);

/* src/kernel.rs:474-575: SW per-target-position maxima (out has n entries), best score, first
 * target position of the best (-1 when the best is 0). */
int hyo_position_max(const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n, uint32_t al,
                     const int32_t* matrix, int32_t go, int32_t ge, int32_t* out, int32_t* score,
                     int64_t* tend);
/* src/kernel.rs:741-792: 1 with (score2, pos) when a secondary peak qualifies, else 0. */
int hyo_score2(const int32_t* colmax, uint64_t n, int32_t best_score, uint64_t best_tend,
               int32_t matrix_max, int32_t min_score, int32_t* s2, uint64_t* pos);
/* SSE4.1 striped (Farrar) local-mode scores of n pairs, one thread (the CPU baseline of the
 * single-pair path); 0 or -1 on allocation failure. */
int hyo_sse_striped_sw_pairs(const uint8_t* qflat, const uint64_t* qoff, const uint8_t* tflat,
                             const uint64_t* toff, uint64_t npairs, uint32_t al,
                             const int32_t* mat, int32_t go, int32_t ge, int32_t* scores,
                             double* seconds);
/* Column maxima by enumeration of local alignments ending in each column (tiny inputs). */
void hyo_brute_colmax(const uint8_t* q, uint64_t m, const uint8_t* t, uint64_t n, uint32_t al,
                      const int32_t* mat, int32_t go, int32_t ge, int32_t* out);

#ifdef __cplusplus
}
#endif
#endif
