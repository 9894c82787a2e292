// Internal types shared by the host library (hy_host.cpp) and the kernels (hy_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>
#include <stdint.h>

namespace hy {

// Alignment modes (src/mode.rs:11-27); numbering matches the C ABI.
enum : int { kSW = 0, kNW = 1, kHW = 2, kOV = 3, kSHW = 4 };

// Sentinel of the 32-bit realisation (src/kernel.rs:39, src/inter.rs:80-98).
constexpr int kNeg32 = INT32_MIN / 4;
// Substitution value used for padded query rows and padded target columns in the 32-bit
// kernels: low enough that a diagonal step through it never beats a real cell, high enough that
// adding it to any real cell cannot wrap.
constexpr int kPad32 = -(1 << 29);
// Matrix entries are clamped to this range before they reach a 32-bit kernel. Within the width
// proof no non-local cell can use an entry beyond it, and in local mode a clamped entry still
// drives its diagonal below 0, so results are unchanged.
constexpr int kEntryMax32 = (1 << 29) - 1;

// Lanes per wavefront on CDNA4.
constexpr int kWave = 64;

// One wavefront's worth of targets for the inter-sequence kernels. Targets are sorted by length
// (descending) before grouping, so the lanes of a group sweep similar column counts.
struct Group {
    uint32_t first;      // first slot (index into the slot-ordered arrays)
    uint32_t count;      // live lanes, <= 64
    uint32_t ncols;      // longest target in the group = columns swept
    uint32_t pad;
    uint64_t word_off;   // residue words: words[word_off + (col/4)*64 + lane], byte col%4
    uint64_t spill_off;  // boundary rows: spill[spill_off + col*64 + lane], col in [0, ncols]
};

// Single-best result slot written by the device.
struct Result {
    int32_t score;
    uint32_t db_index;
    int64_t query_end;   // -1 = None
    int64_t target_end;  // -1 = None
    int32_t check;       // 0 when the device self-check passed
    int32_t pad;
};

struct QueryRef {
    const uint8_t* seq;  // device pointer to the encoded query
    int m;               // query length
};

struct InterArgs {
    const Group* groups;
    const uint32_t* words;
    const uint32_t* lens;    // per slot
    const uint32_t* gidx;    // per slot -> database index
    const uint8_t* query;
    int m;
    const int32_t* matrix;   // al x al, entries clamped to +-kEntryMax32
    int al;                  // matrix row stride
    int al_eff;              // symbols that can occur (min(al, 255)); symbol al_eff = padding
    int go, ge;
    int2* spill;
    int32_t* out_score;      // database order (may be null)
    int64_t* out_qend;       // database order (ends variant)
    int64_t* out_tend;
    unsigned long long* key; // fused argmax key (may be null)
};

// Packed 16-bit inter-sequence kernel: two targets per lane (low half = even slot, high half =
// odd slot of the lane's pair), 128 targets per group.
struct Inter16Args {
    const Group* groups;     // count <= 128
    uint32_t group_first;    // first group of the launch
    uint32_t ngroups;        // groups of the launch
    const uint2* words;      // words[word_off + (col/4)*64 + lane] = (low target, high target)
    const uint32_t* lens;    // per slot
    const uint32_t* gidx;    // per slot -> database index
    const uint8_t* query;
    int m;
    const int32_t* matrix;   // al x al (clamped to the 32-bit range; saturated to 16 bits here)
    int al, al_eff;
    int go, ge;
    uint2* spill;            // spill[spill_off + col*64 + lane] = (H - go pair, F pair)
    int32_t* out_score;      // database order (may be null)
    unsigned long long* key; // fused argmax (may be null)
    int64_t* out_qend;       // end-tracking launches (local mode): database order
    int64_t* out_tend;
};

// Batched form: queries concatenated in one device buffer, processed by the persistent kernel.
struct BatchArgs {
    const uint8_t* queries;     // concatenated encoded queries
    const uint32_t* qoff;       // query k starts at queries + qoff[k]
    const uint32_t* qlen;
    const uint32_t* qorder;     // query indices by descending length
    const uint32_t* nq;         // number of queries (device memory, so a graph can be replayed)
    unsigned int* counter;      // work-queue head, zeroed before each launch
    unsigned long long* keys;   // per-query argmax keys, zeroed before each launch
    uint2* slots;               // boundary rows: one slot of slot_stride entries per workgroup
    uint64_t slot_stride;
};

// Largest pair-profile footprint the packed kernel stages in LDS (bytes).
constexpr uint32_t kPairProfileMaxBytes = 150u * 1024u;

struct SysArgs {
    const uint8_t* flat;      // database residues in original order
    const uint64_t* off;      // n+1 offsets
    const uint32_t* list;     // database indices to align (list mode)
    const unsigned long long* winner_key;  // winner mode: the index comes from this key
    const uint8_t* query;
    int m;
    const int32_t* matrix;
    int al, al_eff;
    int go, ge;
    int2* spill;              // per block: spill[blockIdx.x * spill_stride + col]
    uint64_t spill_stride;
    int32_t* out_score;       // list mode: database order
    int64_t* out_qend;
    int64_t* out_tend;
    unsigned long long* key;  // list mode: fused argmax key (may be null)
    Result* result;           // winner mode output
    // Batched form (qoff != null): blockIdx.y selects the query; key, winner_key and result are
    // per-query arrays, and blocks with blockIdx.y >= *nq exit at once.
    const uint8_t* queries;
    const uint32_t* qoff;
    const uint32_t* qlen;
    const uint32_t* nq;
    int qsel;                 // >= 0: every block uses query qsel instead of blockIdx.y
    int pairs;                // nonzero: block b aligns pair list[b] (its own query and target)
};

// Traceback (src/align.rs:169-420), one (query, target) pair per item. The DP kernel writes a
// 4-bit move code per cell into the item's slab and the pair's end cell into ends[item]; the walk
// kernel reads them back, one lane per item, and writes the alignment's ops and a TraceOut.
//
// Slab layout (32-bit words): the query is cut into passes of 64*R rows (R rows per lane); pass
// p, step s, lane l holds the R nibbles of rows p*64R + l*R + 1 .. + R at column j = s - l + 1,
// at word (p * (n + 63) + s) * 64 + l. With more than one pass the boundary rows between passes
// (int2 per column, n + 1 of them) follow the codes.
struct TraceOut {          // same layout as hy_alignment
    int32_t score;
    int32_t check;         // 0 when the walk ended where the DP says it must
    uint64_t query_start, query_end, target_start, target_end;
    uint64_t n_ops, ops_offset;
};

struct TraceArgs {
    const uint8_t* flat;
    const uint64_t* off;
    const unsigned long long* winner_key;  // non-null: one item, the key's target
    uint32_t first;           // list mode: item k is database index first + k
    uint32_t nitems;
    const uint8_t* query;
    int m;
    const int32_t* matrix;
    int al, al_eff;
    int go, ge;
    int rows;                 // R (rows per lane) of the DP launch
    uint32_t* codes;          // slab pool
    const uint64_t* slab_off; // per database index (words); null: every item at offset 0
    uint64_t slab_base;       // slab_off of the first item of the launch
    int4* ends;               // per item: (score, end row, end column, database index)
    uint8_t* ops;             // winner: ops at [0, m + n); list: target g at g * m + off[g]
    TraceOut* out;            // winner: out[0]; list: out[database index]
    const uint32_t* list;     // non-null: item k is database index list[first + k] (the slab
                              // offsets then follow list positions)
    // Packed form (k_trace16): groups [group_first, group_first + ngroups) of the packed 16-bit
    // layout, item k = slot slot0 + k (its end cell and group at index slot0 + k of ends and
    // item_group); codes of group g at gslab[g] - gslab_base.
    const Group* groups16;
    const uint2* words16;
    const uint32_t* lens16;
    const uint32_t* gidx16;
    uint2* spill16;           // boundary rows between query tiles (the scan's layout)
    uint32_t group_first, ngroups, slot0;
    const uint64_t* gslab;
    uint64_t gslab_base;
    uint32_t* item_group;     // per item: its group (written by the DP, read by the walk)
};

// Rows per lane of the traceback DP for a query of m rows (2, 4, 6 or 8; 8 beyond 384 rows).
inline int trace_rows(int m) { return m <= 128 ? 2 : (m <= 256 ? 4 : (m <= 384 ? 6 : 8)); }
// Slab words of one pair, rounded to whole 64-word (256-byte) lines.
inline uint64_t trace_slab_words(int m, uint64_t n) {
    const uint64_t P = 64ull * (uint64_t)trace_rows(m);
    const uint64_t npass = ((uint64_t)m + P - 1) / P;
    uint64_t w = npass * (n + 63) * 64 + (npass > 1 ? 2 * (n + 1) : 0);
    return (w + 63) & ~63ull;
}

// Packed traceback DP (k_trace16): query rows per tile, and move-code words of one group for a
// query of m rows: per tile, column and 4-row block, one 32-bit word per lane (each 16-bit half
// holds its target's four 4-bit codes).
constexpr int kTrace16Rows = 32;
inline uint64_t trace16_group_words(int m, uint32_t ncols) {
    const uint64_t tiles = ((uint64_t)m + kTrace16Rows - 1) / kTrace16Rows;
    return tiles * (uint64_t)ncols * (kTrace16Rows / 4) * 64;
}
// Its workgroup of W wavefronts (at most kTrace16MaxWaves; twelve, one pass of a 360-row query,
// measured slower than eight) takes in LDS: per wavefront its tile profile, al_eff + 1 symbols
// of trace16_profile_stride() words (half the tile rows, padded so that a symbol's row starts a
// 16-byte line on alternate banks); W - 1 double-buffered row hand-offs of 64 lanes; W x 64
// lanes x 6 ints for the fold of the bests. trace16_waves gives the most wavefronts, up to
// `want` and the tile count, whose LDS fits a launch's 64 KB (one always does: al_eff <= 255 in
// the packed layout).
constexpr int kTrace16MaxWaves = 8;
constexpr int trace16_profile_stride() {
    return ((kTrace16Rows / 2) % 8 == 4) ? kTrace16Rows / 2 : kTrace16Rows / 2 + 4;
}
inline uint32_t trace16_pp_words(int al_eff) {
    return ((uint32_t)(al_eff + 1) * (uint32_t)trace16_profile_stride() + 3u) & ~3u;
}
inline size_t trace16_lds_bytes(int al_eff, int waves) {
    return (size_t)waves * trace16_pp_words(al_eff) * 4 + (size_t)(waves - 1) * 128 * 8 +
           (size_t)waves * 64 * 6 * 4;
}
inline int trace16_waves(int al_eff, int m, int want) {
    const int tiles = (m + kTrace16Rows - 1) / kTrace16Rows;
    int w = std::max(1, std::min(std::min(want, kTrace16MaxWaves), tiles));
    while (w > 1 && trace16_lds_bytes(al_eff, w) > 64u * 1024u) --w;
    return w;
}
template <int MODE>
hipError_t launch_trace16_m(const TraceArgs& a, hipStream_t st);
hipError_t launch_trace16(int mode, const TraceArgs& a, hipStream_t stream);

// Independent (query, target) pairs (align_pairs, src/kernel.rs:429-452), one pair per lane.
// Pairs are sorted by (query length, target length), descending, and cut into groups of 64;
// each group's queries and targets are lane-interleaved like the database layout (one 32-bit
// word carries four symbols of a lane; the padding symbol al_eff fills past each length).
struct PairGroup {
    uint32_t first;      // first slot of the group
    uint32_t count;      // live lanes, <= 64
    uint32_t ncols;      // longest target
    uint32_t nrows;      // longest query
    uint64_t tword_off;  // target words: twords[tword_off + (col/4)*64 + lane]
    uint64_t qword_off;  // query words: qwords[qword_off + (row/4)*64 + lane], whole tiles
    uint64_t spill_off;  // tile boundary rows: spill[spill_off + col*64 + lane], col <= ncols
};

struct PairArgs {
    const PairGroup* groups;
    const uint32_t* twords;
    const uint32_t* qwords;
    const uint32_t* tlens;    // per slot
    const uint32_t* qlens;    // per slot
    const uint32_t* pidx;     // per slot -> pair index
    const int32_t* matrix;    // al x al, entries clamped to +-kEntryMax32
    int al, al_eff;
    int go, ge;
    int2* spill;
    int32_t* out_score;       // per pair
    int64_t* out_qend;
    int64_t* out_tend;
};

// Local-mode pairs two per lane (k_pairs16, alphabets up to kPairs16MaxAlphabet): a group
// (PairGroup, count <= 128) puts slot first + 2l in the low 16-bit half of lane l and slot
// first + 2l + 1 in the high half. With S = al + 1 (symbol al pads), padding scoring -32768:
//  * alphabets up to kPairs16TableMaxAlphabet: a target word holds four columns as bytes
//    ta * S + tb, a query word two rows as 16-bit byte offsets 4 * (qa * S + qb) * S^2 into the
//    kernel's LDS table T[(qa*S + qb)*S^2 + ta*S + tb] = (sub(qa, ta), sub(qb, tb));
//  * larger ones, up to kPairs16MaxAlphabet (pairs16_wide): four columns per lane take a uint2
//    (twords[tword_off + (col/4) * 128 + 2 * lane + {0, 1}]), 16 bits per column holding the
//    symbols ta and tb; a query word holds one row as the 16-bit byte offsets 4 * C * S * qa and
//    4 * C * S * qb into the kernel's S x S table of scores, one lookup per half. The table keeps
//    each score in a 32-bit slot, kPairs16Copies = C copies interleaved (entry e, copy c at slot
//    e * C + c) and lane l reading copy l mod C, so a half-wavefront's 32 lookups meet at most
//    two to a bank instead of landing on random banks: at most 40,000 bytes (S <= 25).
// Every pair's width proof gives 16 bits, its query is at most 4,096 and its target at most
// 65,535 symbols.
constexpr uint32_t kPairs16MaxAlphabet = 24;
constexpr uint32_t kPairs16TableMaxAlphabet = 7;
constexpr uint32_t kPairs16Copies = 16;
__host__ __device__ inline bool pairs16_wide(int al) {
    return al > (int)kPairs16TableMaxAlphabet;
}
struct Pair16Args {
    const PairGroup* groups;
    const uint32_t* twords;
    const uint32_t* qwords;
    const uint32_t* tlens;    // per slot
    const uint32_t* qlens;    // per slot
    const uint32_t* pidx;     // per slot -> pair index
    const int32_t* matrix;    // al x al
    int al, go, ge;
    uint32_t ngroups;         // set by the launcher
    uint2* spill;             // tile boundary rows (H, F') per column
    int32_t* out_score;       // per pair
    int64_t* out_qend;
    int64_t* out_tend;
};

struct PairPackArgs {
    const PairGroup* groups;
    const uint8_t* qflat;
    const uint64_t* qoff;     // per pair
    const uint8_t* tflat;
    const uint64_t* toff;
    const uint32_t* pidx;     // per slot -> pair index
    const uint32_t* qlens;    // per slot
    const uint32_t* tlens;    // per slot
    uint32_t* qwords;
    uint32_t* twords;
    uint32_t pad_symbol;
    int qtile;                // rows per tile: each group's query words cover whole tiles
};

// Local-mode per-target-position maxima of one pair (align_pair_position_max,
// src/kernel.rs:474-575): the systolic shape with every lane folding its rows' maxima into
// colmax[j - 1] (zeroed beforehand), then the best and its first column.
struct PosMaxArgs {
    const uint8_t* query;
    int m;
    const uint8_t* target;
    int n;
    const int32_t* matrix;
    int al, al_eff;
    int go, ge;
    int2* spill;              // n + 1 entries (queries beyond one pass)
    int32_t* colmax;          // n entries
    int2* best;               // (score, target end or -1)
};

// Multi-wavefront pair kernel (k_band32): block b aligns pair list[b] with W wavefronts of one
// workgroup, wavefront w holding query rows w*64R+1 .. (w+1)*64R of each pass of 64*R*W rows and
// running a fixed lag behind wavefront w-1, whose bottom row it reads from an LDS ring. Outputs:
// the score (and ends) per pair, or the local-mode column maxima of a single pair (zeroed by the
// kernel) with the best and its first column.
enum { kBandScore = 0, kBandEnds = 1, kBandColmax = 2 };
struct BandArgs {
    const uint8_t* queries;
    const uint32_t* qoff;     // per pair
    const uint32_t* qlen;
    const uint8_t* targets;
    const uint64_t* toff;     // per pair, n + 1 entries
    const uint32_t* list;     // block b -> pair index
    const int32_t* matrix;
    int al, go, ge;
    int2* spill;              // per block: spill[blockIdx.x * spill_stride + col], longer queries
    uint64_t spill_stride;
    int32_t* out_score;       // per pair
    int64_t* out_qend;
    int64_t* out_tend;
    int32_t* colmax;          // kBandColmax: n entries of the one pair
    int2* best;               // kBandColmax: (score, target end or -1)
    // Winner mode (a database scan's single best, one block): the target is the database entry
    // the key names (targets/toff are the database's residues and offsets), the query is
    // `queries` with wm rows, and the result (ends, with a self-check of the key's score) goes to
    // *result.
    const unsigned long long* winner_key;
    Result* result;
    int wm;
};
// Alphabets the band kernel takes (its substitution matrix and rings fit 64 KB of LDS).
constexpr uint32_t kBandMaxAlphabet = 63;
// Launch shape of the band kernel for a longest query of m rows: r rows per lane, w wavefronts
// (one pass covers 64 * r * w rows). HYALITE_HIP_BAND_SHAPE="r,w" overrides it (shapes the
// launcher instantiates only).
struct BandShape {
    int r, w;
};
BandShape band_shape(int m);
inline int band_rows_per_pass(int m) {
    const BandShape s = band_shape(m);
    return 64 * s.r * s.w;
}

// Large copies between pageable host memory and the device (hy_xfer.cpp): two pinned slots per
// worker thread and the event of the DMA that last used each. Uploads return with their DMAs
// queued on `stream` (after the work already there); downloads return complete.
struct PinnedRing {
    static constexpr int kThreads = 8;
    static constexpr size_t kSlot = 4ull << 20;
    uint8_t* ring = nullptr;
    hipEvent_t ev[2 * kThreads] = {};
    // A DMA recorded by ev[k] may still read slot k: the slot is not refilled before it ends,
    // also when that DMA was queued by an earlier call. Slot k belongs to worker k / 2.
    bool busy[2 * kThreads] = {};
    PinnedRing() = default;
    PinnedRing(const PinnedRing&) = delete;
    PinnedRing& operator=(const PinnedRing&) = delete;
    ~PinnedRing();
};
// Below this a plain hipMemcpyAsync is as fast.
constexpr size_t kPinnedRingMin = 8ull << 20;
// max_byte (optional) receives the largest byte of src, found while copying it.
int upload_pageable(PinnedRing& r, int device, void* dst, const void* src, size_t bytes,
                    hipStream_t stream, uint8_t* max_byte = nullptr);
// An upload of several pageable buffers through the ring that runs on the worker threads while
// the caller goes on: start() returns at once; finish() joins the workers, after which every
// DMA is queued on `stream` and max_byte holds the largest byte of the segments marked as
// symbols. Work that needs the data is queued after finish(). The source buffers must stay
// valid until finish() or the destructor (which joins) returns; the ring is not used by
// anything else meanwhile.
struct UpSegment {
    void* dst;
    const void* src;
    size_t bytes;
    bool symbols;
};
class UploadJob {
public:
    UploadJob() = default;
    UploadJob(const UploadJob&) = delete;
    UploadJob& operator=(const UploadJob&) = delete;
    ~UploadJob();
    int start(PinnedRing& r, int device, const UpSegment* segs, int nseg, hipStream_t stream);
    int finish(uint8_t* max_byte);

private:
    struct Piece {
        uint8_t* dst;
        const uint8_t* src;
        size_t len;
        bool symbols;
    };
    std::vector<Piece> pieces_;
    std::vector<std::thread> pool_;
    std::vector<int> status_;
    std::vector<uint8_t> wmax_;
};
// The largest byte of p[0, n) (host memory), and a copy that finds it on the way.
uint8_t host_max_byte(const uint8_t* p, size_t n);
uint8_t copy_max_byte(uint8_t* dst, const uint8_t* src, size_t n);
struct Segment {
    void* dst;
    const void* src;
    size_t bytes;
};
// Every segment (device src -> pageable dst), after the work already queued on `stream`.
int download_pageable(PinnedRing& r, int device, const Segment* segs, int nseg,
                      hipStream_t stream);
// The segments back to pageable memory, complete on return: through the ring's worker threads
// when together they hold at least kPinnedRingMin bytes, else through its first two slots.
int download(PinnedRing& r, int device, const Segment* segs, int nseg, hipStream_t stream);

// Process-wide lock held across a graph capture (the batched form's first submission) and around
// the library's allocation and release work (databases, scratches and their reservations, pair
// scratches, the traceback buffers, the pinned transfer slots), so that no allocation, release
// or null-stream copy of one thread falls inside another thread's capture: the runtime may
// invalidate a capture in progress for such a call, whatever the capture mode. Recursive: the
// entry points nest (a server builds scratches, an image load builds a database).
std::recursive_mutex& api_lock();

// Launchers (hy_launch.hip). All launch on `stream` and never synchronise.
hipError_t launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v,
                           hipStream_t stream);
hipError_t launch_inter32(int mode, bool ends, const InterArgs& a, uint32_t ngroups,
                          hipStream_t stream);
// nq_grid > 1: batched form, a grid of nblocks x nq_grid blocks; a.m is then the longest query
// of the batch (it fixes the rows per lane).
// Database layouts on the device: lane-interleaved words of the 32-bit groups (64 targets) and
// of the packed groups (128 targets, two per lane), from the residues and n + 1 offsets.
hipError_t launch_db_pack32(const Group* groups, uint32_t ngroups, const uint32_t* gidx,
                           const uint64_t* off, const uint8_t* flat, uint32_t* words,
                           uint32_t pad, hipStream_t stream);
hipError_t launch_db_pack16(const Group* groups, uint32_t ngroups, const uint32_t* gidx,
                           const uint64_t* off, const uint8_t* flat, uint2* words, uint32_t pad,
                           hipStream_t stream);
// Pairs: lane-interleave a batch's queries and targets (one workgroup per group).
hipError_t launch_pair_pack(const PairPackArgs& a, uint32_t ngroups, hipStream_t stream);
// The two-pairs-per-lane words of k_pairs16 (a.pad_symbol = al, a.qtile its tile rows).
hipError_t launch_pair_pack16(const PairPackArgs& a, uint32_t ngroups, hipStream_t stream);
// Local-mode pairs two per lane; q: rows per tile (16 or 32).
hipError_t launch_pairs16_sw(bool ends, const Pair16Args& a, uint32_t ngroups, int q,
                             hipStream_t stream);
// Pairs: one launch over `ngroups` groups. q: rows per tile (16 or 32).
hipError_t launch_pairs32(int mode, bool ends, const PairArgs& a, uint32_t ngroups, int q,
                          hipStream_t stream);
// Column maxima of one local-mode pair (zeroes colmax first).
hipError_t launch_posmax(const PosMaxArgs& a, hipStream_t stream);
// Band kernel over nblocks pairs whose longest query has m rows (it fixes W and R); alphabets up
// to kBandMaxAlphabet. out: kBandScore, kBandEnds, or kBandColmax (local mode, one block).
hipError_t launch_band32(int mode, int out, const BandArgs& a, uint32_t nblocks, int m,
                         hipStream_t stream);
hipError_t launch_fill_i32(int32_t* p, uint64_t n, int32_t v, hipStream_t stream);
// Traceback: the DP over a.nitems items, then the walk (one lane per item).
hipError_t launch_trace(int mode, const TraceArgs& a, hipStream_t stream);
hipError_t launch_sys32(int mode, bool ends, bool winner, const SysArgs& a, uint32_t nblocks,
                        hipStream_t stream, uint32_t nq_grid = 1);
// q: query rows per register tile (16, 24 or 32); waves: 1 = one-wavefront workgroups with a
// per-symbol profile, 4 or 8 = wavefronts per workgroup sharing one residue-pair profile.
hipError_t launch_inter16(int mode, const Inter16Args& a, int q, int waves, hipStream_t stream);
// Persistent batch kernel (one-wavefront workgroups, per-symbol profile): `workgroups` resident
// workgroups drain nq * a.ngroups (group, query) items.
hipError_t launch_inter16_batch(int mode, const Inter16Args& a, const BatchArgs& b, int q,
                                uint32_t workgroups, hipStream_t stream);
// Tile rows the packed kernel actually uses for a requested q.
int inter16_tile_rows(int q);
// LDS bytes of the pair profile for an alphabet of al_eff symbols (plus the padding symbol).
uint32_t inter16_lds_bytes(int al_eff, int q);
// LDS bytes of the per-symbol profile (one-wavefront workgroups of larger alphabets).
uint32_t inter16p_lds_bytes(int al_eff, int q);
hipError_t launch_key_result(const unsigned long long* key, Result* r, hipStream_t stream);
// Op packing (scan_all_aligned) in two launches around a host read of the total:
// launch_ops_sums leaves in bsum (ops_sum_words(n) entries) each block of kOpsBlock targets'
// first packed offset and in bsum[ops_blocks(n)] the total; launch_ops_pack then moves target k's
// n_ops bytes from src + out[k].ops_offset to dst + (that packed offset) and rewrites
// out[k].ops_offset to it. dst holds at least the total and does not overlap src.
constexpr int kOpsBlock = 1024;
inline uint64_t ops_blocks(uint64_t n) { return (n + kOpsBlock - 1) / kOpsBlock; }
inline uint64_t ops_sum_words(uint64_t n) { return ops_blocks(n) + 1; }
hipError_t launch_ops_sums(const TraceOut* out, uint64_t n, uint64_t* bsum, hipStream_t stream);
hipError_t launch_ops_pack(TraceOut* out, uint64_t n, const uint64_t* bsum, const uint8_t* src,
                           uint8_t* dst, hipStream_t stream);
// Batched form: zero n keys and the queue counter; write a Result per query from its key.
hipError_t launch_batch_reset(unsigned long long* keys, uint32_t n, unsigned int* counter,
                              hipStream_t stream);
hipError_t launch_key_results(const unsigned long long* keys, Result* r, const uint32_t* nq,
                              uint32_t cap, hipStream_t stream);
// Resident one-wavefront workgroups of the batch kernel per CU at the occupancy limit.
uint32_t inter16_batch_occupancy(int mode, int q, int al_eff);
// CUs of the current device.
uint32_t device_cus();
// Batched form, local mode, two queries per lane (hy_kernels.hip compiled for SW): launch,
// per-CU occupancy and LDS bytes of the paired profile.
hipError_t launch_inter16q_batch(const Inter16Args& a, const BatchArgs& b, int q, uint32_t wgs,
                                 hipStream_t st);
uint32_t inter16q_batch_occupancy(int q, int al_eff);
uint32_t inter16q_lds_bytes(int al_eff, int q);

// Per-mode launchers, instantiated by hy_kernels.hip once per alignment mode.
template <int MODE>
hipError_t launch_inter32_m(bool ends, const InterArgs& a, uint32_t ngroups, hipStream_t st);
template <int MODE>
hipError_t launch_sys32_m(bool ends, bool winner, const SysArgs& a, uint32_t nblocks,
                          hipStream_t st, uint32_t nq);
template <int MODE>
hipError_t launch_inter16_m(const Inter16Args& a, int q, int waves, hipStream_t st);
template <int MODE>
hipError_t launch_inter16_batch_m(const Inter16Args& a, const BatchArgs& b, int q, uint32_t wgs,
                                  hipStream_t st);
template <int MODE>
uint32_t inter16_batch_occupancy_m(int q, int al_eff);
template <int MODE>
hipError_t launch_trace_m(const TraceArgs& a, hipStream_t st);
template <int MODE>
hipError_t launch_pairs32_m(bool ends, const PairArgs& a, uint32_t ngroups, int q,
                            hipStream_t st);
hipError_t launch_posmax_sw(const PosMaxArgs& a, hipStream_t st);
template <int MODE>
hipError_t launch_band32_m(int out, const BandArgs& a, uint32_t nblocks, int m, hipStream_t st);
// Pipelined packed one-query kernel: one workgroup per group, one query tile per wavefront, at
// most kPipeMaxWaves tiles; scores and the argmax key, or (ends) each target's score and end
// cell, folded across the wavefronts with the full tie-break. inter16_pipe_rows gives the tile rows
// for a query of m rows (0 when it needs more than kPipeMaxWaves tiles or only one): 16 rows up
// to 64, 24 up to 192, 32 beyond. Measured on a 100-row query (10,000 DNA targets): 16-row tiles
// (7 wavefronts) 0.205 ms, 24 (5) 0.191 ms, 32 (4) 0.228 ms per launch.
constexpr int kPipeMaxWaves = 8;
// The batched form gives each resident workgroup a boundary slot of (widest group + 1) x 64
// entries of 8 bytes; packed groups wider than this many columns (long targets) run in its
// systolic list instead, which bounds a slot at 4 MB (8.6 GB over 2,048 resident workgroups)
// whatever the database's longest packable target.
constexpr uint32_t kBatchSlotMaxCols = 8192;
inline int inter16_pipe_rows(int m) {
    if (m <= 16) return 0;
    if (m <= 64) return 16;
    if (m <= 192) return 24;
    if (m <= 32 * kPipeMaxWaves) return 32;
    return 0;
}
// The same for an end-tracking scan: local mode keeps tiles of at most 24 rows (its end tracking
// at 32 rows spills registers), so queries up to 24 x kPipeMaxWaves rows.
inline int inter16_pipe_rows_ends(int mode, int m) {
    const int q = inter16_pipe_rows(m);
    if (mode != kSW || q <= 24) return q;
    return m <= 24 * kPipeMaxWaves ? 24 : 0;
}
// One-wavefront workgroups of the packed kernel keep a residue-pair profile of their tile (the
// two targets' scores of a query row in one word, read with no per-lane pairing) when it takes at
// most kPairProfileSmallBytes of LDS, which holds for alphabets of up to 9 symbols (DNA, with
// its padding symbol: 5 x 5 pairs); larger alphabets keep the per-symbol profile and pair per lane.
constexpr uint32_t kPairProfileSmallBytes = 16u * 1024u;
inline bool inter16_pair_small(int al_eff, int q) {
    return inter16_lds_bytes(al_eff, q) <= kPairProfileSmallBytes;
}
// Words of one wavefront's tile profile for q-row tiles.
inline uint32_t inter16_profile_words(int al_eff, int q) {
    return (inter16_pair_small(al_eff, q) ? inter16_lds_bytes(al_eff, q)
                                          : inter16p_lds_bytes(al_eff, q)) / 4;
}
// LDS bytes of a pipelined workgroup of W wavefronts with q-row tiles (profiles, double buffers,
// the fold of the running maxima, or of the six end words per lane); the host uses the kernel
// only within 64 KB.
inline uint32_t inter16_pipe_lds_bytes(int al_eff, int q, int W, bool ends = false) {
    return (uint32_t)W * inter16_profile_words(al_eff, q) * 4u + (uint32_t)(W - 1) * 4096u +
           (uint32_t)W * 256u * (ends ? 6u : 1u);
}
template <int MODE>
hipError_t launch_inter16_pipe_m(const Inter16Args& a, int q, bool ends, hipStream_t st);
hipError_t launch_inter16_pipe(int mode, const Inter16Args& a, int q, bool ends,
                               hipStream_t stream);
// Chained pipelined kernel (k_inter16_chain): the query tiles of a group spread over `chunks`
// workgroups of W wavefronts, consecutive chunks linked through the group's boundary buffer and
// one progress word per link. At most kChainMaxChunks chunks; tile rows 8, 16, 24 or 32. The
// launch takes (ngroups rounded up to 8) x chunks workgroup tickets. With ends every chunk folds
// its wavefronts' best cells, joins those of the chunks before it and passes them on through
// `fold`; the last chunk writes each target's score and ends.
constexpr int kChainMaxChunks = 8;
constexpr int kChainMaxLinks = kChainMaxChunks - 1;
struct ChainArgs {
    unsigned long long* progress;  // [group * kChainMaxLinks + link], zeroed once at allocation
    unsigned long long epoch;      // launch number << 32 (every launch takes a new one)
    unsigned int* err;             // host-visible (pinned) flag, set when a wait gave up
    unsigned long long* ticket;    // workgroup start counter, never reset
    unsigned long long ticket_base;  // its value when this launch starts (tickets taken before)
    int tiles;                     // query tiles of q rows, the last one possibly shorter
    int chunks;                    // workgroups per group: ceil(tiles / W)
    int slack;                     // blocks a reader lets its writer publish before it starts
    int every;                     // a writer publishes its progress every `every` blocks
    uint32_t spin_limit;           // polls before a wait gives up (kChainSpinLimit)
    uint2* fold;  // ends: [(group * kChainMaxLinks + link) * 3 + k][64] best cells passed on
};
// uint2 entries of the ends fold buffer for `groups` packed groups.
inline uint64_t chain_fold_entries(uint64_t groups) {
    return (groups > 0 ? groups : 1) * (uint64_t)kChainMaxLinks * 3u * 64u;
}
// LDS bytes of a chained workgroup of W wavefronts with tiles of q rows.
inline uint32_t inter16_chain_lds_bytes(int al_eff, int q, int W, bool ends) {
    return inter16_pipe_lds_bytes(al_eff, q, W, ends);
}
// Local-mode end tracking keeps 16-row tiles or shorter in the chained kernel (the reader's
// prefetched rows on top of a 24- or 32-row tile's end tracking would spill).
inline bool inter16_chain_rows_ok(int mode, int q, bool ends) {
    return (q == 8 || q == 16 || q == 24 || q == 32) && !(ends && mode == kSW && q > 16);
}
// Polls of a chained wait (0.5-1 us each: an L2 round trip and a short sleep) before it gives up:
// seconds, far beyond any launch's wait, so that a reader always ends even if its writer never
// published. Tests set fewer (HYALITE_HIP_CHAIN_SPIN_LIMIT) to take the give-up path.
constexpr uint32_t kChainSpinLimit = 1u << 22;
template <int MODE>
hipError_t launch_inter16_chain_m(const Inter16Args& a, const ChainArgs& c, int q, int waves,
                                  bool ends, hipStream_t st);
hipError_t launch_inter16_chain(int mode, const Inter16Args& a, const ChainArgs& c, int q,
                                int waves, bool ends, hipStream_t stream);
// The packed one-query kernel with end tracking: score and ends of every target of the launch's
// groups into out_score / out_qend / out_tend.
template <int MODE>
hipError_t launch_inter16_ends_m(const Inter16Args& a, hipStream_t st);
hipError_t launch_inter16_ends(int mode, const Inter16Args& a, hipStream_t stream);

}  // namespace hy
