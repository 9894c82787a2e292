/*
 * hyalite_hip — C ABI of the MI355X (gfx950) database-scan backend.
 *
 * This is the drop-in boundary a `Backend::Hip` arm of hyalite would bind through a thin FFI
 * crate. It replaces, for the Hip backend only, the packed-database build and the
 * inter-sequence scan of hyalite 0.3.0:
 *
 *   hy_db_build      <- DatabaseBuilder::build            src/database.rs:737-877
 *                       (+ build_width_groups             src/database.rs:884-948,
 *                          PackedDb::build                 src/inter.rs:431-508)
 *   hy_scratch_new   <- Scratch::new                      src/database.rs:965-999
 *   hy_scan          <- Database::scan / scan_dispatch    src/database.rs:83-208, src/inter.rs:1074-1124
 *   hy_scan_all      <- Database::scan_all / fill_ends    src/database.rs:213-301, 347-464,
 *                                                          src/inter.rs:1226-1262
 *   hy_scan_scores   <- Database::scan_scores / fill_scores src/database.rs:305-329, 481-515,
 *                                                          src/inter.rs:1161-1205
 *   hy_scan_aligned  <- Database::scan_aligned            src/database.rs:528-540,
 *                                                          src/align.rs:192-226, 487-545
 *   hy_scan_all_aligned <- Database::scan_all_aligned     src/database.rs:551-580
 *   hy_pair_scratch_new <- PairScratch::new               src/kernel.rs:285-311
 *   hy_align_pair    <- align_pair / align_pair_with      src/kernel.rs:331-420
 *   hy_align_pairs   <- align_pairs                       src/kernel.rs:429-452
 *   hy_align_pair_position_max <- align_pair_position_max(_with) src/kernel.rs:474-575
 *   hy_score2        <- score2                            src/kernel.rs:741-792
 *   hy_scan_batch    <- a loop of Database::scan over many queries (batched form; no single
 *                       reference counterpart — each hit equals Database::scan of that query)
 *   hy_backend_available <- Backend::is_available         src/backend.rs:77-84
 *   hy_strerror      <- impl Display for Error             src/error.rs:98-157
 *
 * Conventions
 *   - Sequences and queries are pre-encoded alphabet indices (uint8_t), as in the reference.
 *   - Ends are 0-based inclusive positions; -1 stands for `None` (hit.rs:16-25).
 *   - Every function returns 0 on success or one of the HY_E_* codes below; no exception or
 *     longjmp crosses the ABI. Construction errors use the numbering of the reference's `Error`
 *     variants (error.rs:16-96); scan-time codes (>= 100) are contract violations that the
 *     reference turns into panics (database.rs:74-81), so a Rust wrapper should panic on them.
 *   - A hy_db is immutable after build and may be shared by many threads; each thread uses its
 *     own hy_scratch (one HIP stream each). hy_db_retain/hy_db_free reference-count it, which is
 *     what a Rust `Clone` of `Database` maps to.
 */
#ifndef HYALITE_HIP_H
#define HYALITE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hy_db hy_db;
typedef struct hy_scratch hy_scratch;
typedef struct hy_pair_scratch hy_pair_scratch;

/* Mode (src/mode.rs:11-27). */
enum { HY_MODE_SW = 0, HY_MODE_NW = 1, HY_MODE_HW = 2, HY_MODE_OV = 3, HY_MODE_SHW = 4 };
/* SearchType (src/search.rs:13-27). ALIGNMENT carries its traceback budget in hy_db_build's
 * traceback_max_bytes and tracks ends like SCORE_END (search.rs:33-35). */
enum { HY_SEARCH_SCORE = 0, HY_SEARCH_SCORE_END = 1, HY_SEARCH_ALIGNMENT = 2 };
/* AlignOp (src/align.rs:47-57): the values of an alignment's op bytes. Ins consumes a query
 * symbol, Del a target symbol (the target is the reference, as in SAM). */
enum { HY_OP_MATCH = 0, HY_OP_MISMATCH = 1, HY_OP_INS = 2, HY_OP_DEL = 3 };
/* Backend choice: AUTO or FORCE(HIP). The CPU backends stay in the reference crate. */
enum { HY_BACKEND_AUTO = 0, HY_BACKEND_HIP = 9 };

/* Status codes. 1..11 mirror error.rs:16-96 in declaration order. */
enum {
    HY_OK = 0,
    HY_E_EMPTY_ALPHABET = 1,
    HY_E_MATRIX_SHAPE = 2,
    HY_E_NEGATIVE_GAP_PENALTY = 3,
    HY_E_GAP_OPEN_LESS_THAN_EXTEND = 4,
    HY_E_SCORE_RANGE_TOO_WIDE = 5,   /* err_detail = the width-proof bound */
    HY_E_SYMBOL_OUT_OF_RANGE = 6,    /* err_detail = the offending symbol */
    HY_E_INCOMPLETE_BUILDER = 7,     /* err_detail = 0 sequences, 1 scoring, 2 mode, 3 max_query_len */
    HY_E_EMPTY_DATABASE = 8,
    HY_E_BACKEND_UNAVAILABLE = 9,
    HY_E_INVALID_BACKEND_NAME = 10,
    HY_E_TRACEBACK_BUDGET = 11,      /* err_detail = bytes needed (as uint64_t bits) */
    /* Scan-time contract violations and runtime failures. */
    HY_E_QUERY_TOO_LONG = 100,       /* err: query longer than max_query_len */
    HY_E_QUERY_SYMBOL = 101,         /* query symbol >= alphabet_len */
    HY_E_NULL_ARGUMENT = 102,
    HY_E_INVALID_ARGUMENT = 103,
    HY_E_DEVICE = 104,               /* a HIP runtime call failed */
    HY_E_OUT_OF_MEMORY = 105,
    HY_E_INTERNAL = 106,             /* a device-side self-check failed */
    HY_E_TIMEOUT = 107,              /* hy_server_wait: not completed within the timeout */
    HY_E_NOT_ALIGNMENT = 108,        /* an aligned scan on a database built without ALIGNMENT */
    HY_E_BAD_IMAGE = 109,            /* hy_db_load: not a database image, or a damaged one */
    HY_E_STATIC_LIMIT = 110          /* hy_scan_scores_static: the problem needs more than its
                                        static device array holds */
};

/* Field flags for hy_db_build's `have` mask: which builder fields were set. A field left out
 * yields HY_E_INCOMPLETE_BUILDER in the reference's check order. */
enum {
    HY_HAVE_SEQUENCES = 1u << 0,
    HY_HAVE_SCORING = 1u << 1,
    HY_HAVE_MODE = 1u << 2,
    HY_HAVE_MAX_QUERY_LEN = 1u << 3,
    HY_HAVE_ALL = 0xFu
};

/* Whether a gfx950 HIP device is usable (Backend::is_available for Hip). */
int hy_backend_available(void);

/* Parse a backend name as HYALITE_BACKEND would (src/backend.rs:143-155 plus "hip").
 * Writes HY_BACKEND_AUTO or HY_BACKEND_HIP to *out; the reference CPU names
 * (scalar, sse4.1, sse41, avx2, neon) parse but are not served here: *out = -1. */
int hy_backend_parse(const char* name, int32_t* out);

/* Build a device-resident database. Validation order, error codes and the per-sequence width
 * proof follow DatabaseBuilder::build (database.rs:737-877). Sequences are copied; the caller's
 * buffers may be freed afterwards. `backend` is HY_BACKEND_AUTO or HY_BACKEND_HIP; AUTO also
 * consults HYALITE_BACKEND. traceback_max_bytes is SearchType::Alignment's max_bytes (ignored for
 * the other search types): the build fails with HY_E_TRACEBACK_BUDGET when the reference's
 * traceback could not serve the whole [0, max_query_len] x [0, max_target_len] box within it
 * (database.rs:773-787). err_detail (may be NULL) receives the payload documented above. */
int hy_db_build(const uint8_t* const* seqs, const size_t* lens, size_t n, uint32_t have,
                const int32_t* matrix, size_t matrix_len, uint32_t alphabet_len, int32_t gap_open,
                int32_t gap_ext, uint32_t mode, uint32_t search_type, uint64_t traceback_max_bytes,
                size_t max_query_len, uint32_t backend, hy_db** out, int64_t* err_detail);

/* Reference-count a database (a Rust `Clone`); every retain needs one hy_db_free. */
int hy_db_retain(hy_db* db);
void hy_db_free(hy_db* db);

/* Accessors (Database::sequence_count, score_width, max_query_len, max_target_len, mode,
 * search_type; database.rs:581-656). widest_width_bits is 8, 16 or 32. */
int hy_db_info(const hy_db* db, uint32_t* widest_width_bits, uint64_t* n, uint64_t* max_query_len,
               uint64_t* max_target_len, uint32_t* mode, uint32_t* search_type);

/* On-disk image (no reference counterpart: the reference keeps sequences in memory only). The
 * image holds the scheme, the builder fields and the residues in database order with a
 * checksum; hy_db_load re-runs hy_db_build on them (the interleaved layouts are gathered on the
 * device), so a loaded database equals the saved one and gives the same hits. hy_db_save needs a
 * writable path (HY_E_INVALID_ARGUMENT otherwise); hy_db_load returns HY_E_BAD_IMAGE for a
 * file that is not an image or fails its size or checksum checks, then the build's own codes. */
int hy_db_save(const hy_db* db, const char* path);
int hy_db_load(const char* path, uint32_t backend, hy_db** out, int64_t* err_detail);

/* The scheme a database was built with: matrix (alphabet_len^2 entries as given; may be NULL),
 * gaps and the traceback budget (0 unless built for SearchType::Alignment). */
int hy_db_scoring(const hy_db* db, int32_t* matrix, uint32_t* alphabet_len, int32_t* gap_open,
                  int32_t* gap_ext, uint64_t* traceback_max_bytes);

/* Per-width partition sizes (number of targets proven i8, i16, i32). */
int hy_db_partitions(const hy_db* db, uint64_t* n_i8, uint64_t* n_i16, uint64_t* n_i32);

/* Per-thread working memory: a HIP stream plus device scratch sized for `db`. */
int hy_scratch_new(const hy_db* db, hy_scratch** out);
void hy_scratch_free(hy_scratch* s);

/* Single best hit (Database::scan). Highest score, smallest db_index on ties; ends follow the
 * (target_end, query_end) tie-break and are -1 for SearchType::Score. */
int hy_scan(const hy_db* db, hy_scratch* s, const uint8_t* query, size_t qlen, int32_t* score,
            uint64_t* db_index, int64_t* query_end, int64_t* target_end);

/* One hit per target in database order (Database::scan_all). Each output array has n entries;
 * query_ends/target_ends may be NULL for SearchType::Score (they would be all -1). */
int hy_scan_all(const hy_db* db, hy_scratch* s, const uint8_t* query, size_t qlen, int32_t* scores,
                int64_t* query_ends, int64_t* target_ends);

/* Per-target scores in database order (Database::scan_scores; SearchType-independent). */
int hy_scan_scores(const hy_db* db, hy_scratch* s, const uint8_t* query, size_t qlen,
                   int32_t* scores);

/* hy_scan_scores without a scratch and without allocating: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same inputs, outputs and results as hy_scan_scores; the
 * working set (query copy, tile-boundary rows, systolic boundary rows of queries beyond 512
 * rows, the score array: about 4 bytes per database residue plus 4 per target) lives in one
 * static 64 MiB device array per GPU, held until the process ends, and the kernels run on the
 * null stream. A problem needing more returns HY_E_STATIC_LIMIT. Calls that overlap on one GPU
 * share that array and race: serialise them. */
int hy_scan_scores_static(const hy_db* db, const uint8_t* query, size_t qlen, int32_t* scores);

/* Alignment (src/align.rs:60-95): score, half-open spans [start, end) of query and target, and
 * n_ops op bytes (HY_OP_*, forward order) at ops + ops_offset of the buffer given to the scan. */
typedef struct hy_alignment {
    int32_t score;
    uint32_t reserved;
    uint64_t query_start, query_end, target_start, target_end;
    uint64_t n_ops, ops_offset;
} hy_alignment;

/* The best target with its canonical alignment (Database::scan_aligned): the winner of the
 * score pass (the hit hy_scan reports), traced back on the device with the reference's walk
 * preferences, so the alignment equals align::traceback of the query against that target.
 * ops needs ops_cap >= qlen + max_target_len bytes; the ops start at ops[out->ops_offset]. The
 * database must be built with HY_SEARCH_ALIGNMENT (else HY_E_NOT_ALIGNMENT). */
int hy_scan_aligned(const hy_db* db, hy_scratch* s, const uint8_t* query, size_t qlen,
                    uint64_t* db_index, hy_alignment* out, uint8_t* ops, size_t ops_cap);

/* Bytes of the ops buffer hy_scan_all_aligned needs for a query of qlen symbols:
 * n * qlen + the database's total residues, the most ops n alignments can have (an alignment
 * of the query against target k has at most qlen + len_k ops). */
int hy_db_aligned_ops_bytes(const hy_db* db, size_t qlen, uint64_t* bytes);

/* One alignment per target in database order (Database::scan_all_aligned): out has n entries,
 * out[k].ops_offset locates target k's ops in `ops` (ops_cap >= hy_db_aligned_ops_bytes). The
 * ops are packed in database order: out[0].ops_offset = 0 and target k + 1's ops follow target
 * k's, so only their total is written. The pairs run in chunks whose move codes fit the
 * scratch's traceback pool (8 GiB, or HYALITE_HIP_TRACE_POOL_BYTES when set). */
int hy_scan_all_aligned(const hy_db* db, hy_scratch* s, const uint8_t* query, size_t qlen,
                        hy_alignment* out, uint8_t* ops, size_t ops_cap);

/* ---- Pairwise alignment (free functions of src/kernel.rs, no database) ------------------------
 * Each call takes the scoring scheme as Scoring::new does and validates it in the same order
 * (HY_E_EMPTY_ALPHABET .. HY_E_GAP_OPEN_LESS_THAN_EXTEND), then every pair's symbols (query
 * first, then target: HY_E_SYMBOL_OUT_OF_RANGE, err_detail = the symbol) and its width proof
 * (HY_E_SCORE_RANGE_TOO_WIDE, err_detail = the bound). Results equal the reference's scalar
 * align_core (score, then the (target_end, query_end) tie-break; ends -1 for SearchType::Score).
 * A hy_pair_scratch holds a HIP stream and device buffers that grow to the largest problem seen;
 * one per thread, like PairScratch. */
int hy_pair_scratch_new(hy_pair_scratch** out);
void hy_pair_scratch_free(hy_pair_scratch* s);

/* One pair (align_pair_with): one wavefront aligns it in the systolic kernel. */
int hy_align_pair(hy_pair_scratch* s, const uint8_t* query, size_t qlen, const uint8_t* target,
                  size_t tlen, const int32_t* matrix, size_t matrix_len, uint32_t alphabet_len,
                  int32_t gap_open, int32_t gap_ext, uint32_t mode, uint32_t search_type,
                  int32_t* score, int64_t* query_end, int64_t* target_end, int64_t* err_detail);

/* A batch of independent pairs (align_pairs): hit k is align_pair of pair k (db_index k). Large
 * batches run one pair per lane (pairs sorted by length into 64-lane groups, a substitution
 * lookup per cell from LDS); small batches and long pairs run one pair per wavefront.
 * HYALITE_HIP_PAIRS_KERNEL=lanes|waves forces one of the two for every pair. On an error the
 * failing pair's index goes to *err_pair (may be NULL) and no hit is written. */
int hy_align_pairs(hy_pair_scratch* s, const uint8_t* const* queries, const size_t* qlens,
                   const uint8_t* const* targets, const size_t* tlens, size_t npairs,
                   const int32_t* matrix, size_t matrix_len, uint32_t alphabet_len,
                   int32_t gap_open, int32_t gap_ext, uint32_t mode, uint32_t search_type,
                   int32_t* scores, int64_t* query_ends, int64_t* target_ends,
                   int64_t* err_detail, uint64_t* err_pair);

/* hy_align_pairs over concatenated sequences: pair k is queries[query_offsets[k] ..
 * query_offsets[k + 1]) against targets[target_offsets[k] .. target_offsets[k + 1]) (n + 1
 * offsets per side). Same results and errors as hy_align_pairs; no per-pair pointers. */
int hy_align_pairs_flat(hy_pair_scratch* s, const uint8_t* queries,
                        const uint64_t* query_offsets, const uint8_t* targets,
                        const uint64_t* target_offsets, size_t npairs, const int32_t* matrix,
                        size_t matrix_len, uint32_t alphabet_len, int32_t gap_open,
                        int32_t gap_ext, uint32_t mode, uint32_t search_type, int32_t* scores,
                        int64_t* query_ends, int64_t* target_ends, int64_t* err_detail,
                        uint64_t* err_pair);

/* Device milliseconds of the batches run through this scratch (HIP events around the kernels,
 * after the uploads and before the copy back) and their count; reset zeroes them. No reference
 * counterpart: instrumentation. */
int hy_pair_scratch_profile_read(hy_pair_scratch* s, double* device_ms, uint64_t* calls,
                                 int reset);

/* Local-mode per-target-position maxima (align_pair_position_max_with): colmax[t] = the best
 * local score ending at target position t (tlen entries); *score = their maximum, *target_end =
 * its first position or -1 when the best is 0 (query end untracked). */
int hy_align_pair_position_max(hy_pair_scratch* s, const uint8_t* query, size_t qlen,
                               const uint8_t* target, size_t tlen, const int32_t* matrix,
                               size_t matrix_len, uint32_t alphabet_len, int32_t gap_open,
                               int32_t gap_ext, int32_t* colmax, int32_t* score,
                               int64_t* target_end, int64_t* err_detail);

/* bwa-mem's second-best score over a column-maxima array (score2): *found = 1 with
 * (*score2, *position) for the best secondary peak outside the primary's exclusion window,
 * *found = 0 for None. Host arithmetic on the caller's array. */
int hy_score2(const int32_t* colmax, size_t n, int32_t best_score, uint64_t best_target_end,
              int32_t matrix_max, int32_t min_score, int32_t* found, int32_t* score2,
              uint64_t* position);

/* Batched form: nq queries in one submission; hit k equals hy_scan of query k. The queries run
 * in one persistent kernel over a (target group, query) work queue — in local mode two queries
 * share each lane, a (target group, query pair) queue — the winners' end positions in one
 * launch, and the whole submission is captured once as a HIP graph and replayed. Batches larger
 * than the scratch's reservation (or than its batch_chunk) run as several submissions; a
 * submission of a single query takes the one-query path. No reference counterpart: the
 * reference loops Database::scan (src/database.rs:83). */
int hy_scan_batch(const hy_db* db, hy_scratch* s, const uint8_t* const* queries,
                  const size_t* qlens, size_t nq, int32_t* scores, uint64_t* db_indices,
                  int64_t* query_ends, int64_t* target_ends);

/* Reserve device and pinned buffers for batches of up to max_queries queries (done on first use
 * with min(nq, 256) otherwise). The persistent kernel's boundary rows need one slot of
 * (longest packed target + 1) x 512 bytes per resident workgroup. */
int hy_scratch_reserve_batch(hy_scratch* s, size_t max_queries);

/* Launch shapes of the packed 16-bit kernels, held per scratch (defaults come from the database:
 * 32-row tiles, one wavefront per workgroup, the occupancy limit of resident workgroups and
 * 256 queries per batched submission, up to 4,096 for databases of few packed groups). Shapes change speed only, never results. No reference
 * counterpart (the reference picks its SIMD layout in src/inter.rs:277-386). */
typedef struct hy_tuning {
    int32_t tile_rows;        /* one-query scans: query rows per register tile (16, 24 or 32) */
    int32_t waves;            /* one-query scans: wavefronts per workgroup (1, 4 or 8; 4 and 8
                                 share a residue-pair profile in LDS) */
    int32_t batch_tile_rows;  /* batched form: query rows per register tile (16, 24 or 32) */
    int32_t batch_wgs_per_cu; /* batched form: resident workgroups per CU (0 = occupancy limit) */
    int32_t batch_chunk;      /* batched form: queries per submission (1..65536) */
    int32_t batch_pairs;      /* batched form, local mode: 1 = two queries per lane (paired
                                 profile), 0 = two targets per lane; other modes ignore it */
} hy_tuning;

int hy_scratch_get_tuning(const hy_scratch* s, hy_tuning* out);
/* Validates every field (HY_E_INVALID_ARGUMENT otherwise); a change of the batched form's
 * shapes releases its reservation, which the next batched scan makes again. */
int hy_scratch_set_tuning(hy_scratch* s, const hy_tuning* t);

/* Autotuner: times candidate shapes on the given queries against the whole database and keeps
 * the fastest in the scratch. forms: bit 0 = one-query scans (tile rows x waves, timed on 3
 * queries spread over the length range), bit 1 = batched form (tile rows x resident workgroups
 * x, in local mode, query pairs or target pairs per lane, timed on 8 adjacent pairs of the
 * length-sorted queries, which pair in the lanes as the whole batch does; the submission size
 * becomes the whole call, up to 1,024 queries, and no less than the scratch's default). With HYALITE_HIP_TUNE_CACHE naming a file, results are
 * looked up there first (key: device, mode, alphabet, database shape, sample lengths) and
 * appended after tuning. `out` (optional) receives the chosen shapes. */
int hy_scratch_autotune(const hy_db* db, hy_scratch* s, const uint8_t* const* queries,
                        const size_t* qlens, size_t nq, uint32_t forms, hy_tuning* out);

/* Instrumentation (no counterpart in the reference; for benchmarks and profilers). When enabled,
 * every scan records HIP events on the scratch's own stream around the per-target kernel
 * launches, and hy_scratch_profile_read returns the accumulated device time: `inter` is the
 * inter-sequence kernel (the dominant one), `tail` the work beside it on the scratch's second
 * stream (the systolic kernel over long targets, and the 32-bit kernel over targets that fail the
 * packed width proof when the packed kernel runs too), and `launches` the number of scans timed. reset != 0 clears the counters after reading. */
int hy_scratch_profile(hy_scratch* s, int enable);
int hy_scratch_profile_read(hy_scratch* s, double* inter_ms, double* tail_ms, uint64_t* launches,
                            int reset);
/* Device time of the batched form's persistent kernel (profiling runs submit without a graph). */
int hy_scratch_batch_profile_read(hy_scratch* s, double* kernel_ms, uint64_t* launches, int reset);
/* Which packed 16-bit kernel the scratch's last one-query scan ran over the database's short packed
 * groups: HY_KERNEL_NONE (none: no such groups, or no scan yet), HY_KERNEL_PACKED (one group per
 * wavefront), HY_KERNEL_PIPELINED (one group per workgroup, a query tile per wavefront; score-only
 * or end-tracking), HY_KERNEL_CHAINED (a group's query tiles over several workgroups; score-only
 * or end-tracking) or
 * HY_KERNEL_ENDS (end tracking, one group per wavefront). Instrumentation for tests and
 * benchmarks; no counterpart in the reference. */
enum {
    HY_KERNEL_NONE = -1,
    HY_KERNEL_PACKED = 0,
    HY_KERNEL_PIPELINED = 1,
    HY_KERNEL_CHAINED = 2,
    HY_KERNEL_ENDS = 3
};
int hy_scratch_last_kernel(const hy_scratch* s, int32_t* kernel);

/* Serving form (no reference counterpart: the reference is a library without a request loop).
 * A server keeps the database resident and runs lanes, each a scratch with its own stream and
 * worker thread. Submitted queries wait in one FIFO; a free lane admits the queries at its head,
 * up to max_batch, and runs them as one batched scan (hy_scan_batch), so requests that arrive
 * while other batches are on the GPU are admitted as soon as a lane is free and their work
 * overlaps on the device. Admission is by device bytes: scratch_budget_bytes covers the lanes'
 * buffers (lane_bytes each) plus the requests in flight, each charged its own bytes (its query
 * copy, key and result slots, and for a query beyond 512 rows its winner pass's boundary row of
 * max_target_len + 1 entries); a request is admitted when its bytes fit what is left. The server
 * starts with one lane and adds one, up to max_lanes, when a request waits, every lane is busy
 * and the budget holds another lane plus that request (prealloc = 1 builds every lane the budget
 * holds at once instead). HY_E_OUT_OF_MEMORY if the budget does not hold one lane and one
 * request of max_query_len. Every hit equals Database::scan of that query alone
 * (src/database.rs:83). */
typedef struct hy_server hy_server;

typedef struct hy_server_config {
    uint64_t scratch_budget_bytes; /* device bytes the lanes and requests may hold; 0 = no limit */
    uint32_t max_lanes;            /* concurrent batches (1..64) */
    uint32_t max_batch;            /* queries per batch (1..65536) */
    uint32_t prealloc;             /* 1: build every lane at construction; 0: on demand */
    uint32_t reserved[3];          /* must be 0 */
} hy_server_config;

typedef struct hy_server_result {
    int32_t score;
    int32_t status;                /* HY_OK, or the scan's error code */
    uint64_t db_index;
    int64_t query_end;             /* -1 = None */
    int64_t target_end;
    uint64_t submit_ns, admit_ns, done_ns;  /* steady-clock timestamps */
    uint32_t batch_size;           /* queries in the batch that carried this one */
    uint32_t lane;
} hy_server_result;

typedef struct hy_server_stats {
    uint64_t submitted, completed, batches;
    uint64_t query_residues;       /* sum of completed query lengths */
    uint32_t lanes, max_batch;     /* lanes built so far */
    uint64_t lane_bytes;           /* device bytes one lane holds */
    uint64_t peak_bytes;           /* most bytes held at once (lanes + requests in flight) */
    uint64_t peak_queries;         /* most requests in flight at once */
    uint64_t inflight_bytes;       /* bytes of the requests in flight now */
} hy_server_stats;

int hy_server_new(const hy_db* db, const hy_server_config* cfg, hy_server** out);
/* Copies the query; validates it as hy_scan does (codes 100, 101) before queueing. */
int hy_server_submit(hy_server* srv, const uint8_t* query, size_t qlen, uint64_t* ticket);
/* Waits for a ticket (timeout_us < 0: no limit). HY_E_TIMEOUT if it has not completed in time
 * (the ticket stays valid); HY_E_INVALID_ARGUMENT for an unknown or already collected ticket. */
int hy_server_wait(hy_server* srv, uint64_t ticket, int64_t timeout_us, hy_server_result* out);
int hy_server_stats_read(hy_server* srv, hy_server_stats* out);
/* Finishes every queued query, then stops the lanes and frees the server. */
void hy_server_free(hy_server* srv);

/* Human-readable message for a status code (error.rs:98-157 wording where one exists). */
const char* hy_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif
