// Internal state of the C ABI objects: the device-resident database (hy_db) and a thread's
// scratch (hy_scratch). Shared by the host-side translation units of the library.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <vector>

#include "../../include/hyalite_hip.h"
#include "hy_internal.h"

namespace hy {
inline void dev_free(void* p) {
    if (p) (void)hipFree(p);
}
}  // namespace hy

struct hy_db {
    std::atomic<int> refs{1};
    int device = 0;
    // Validated scheme (src/scoring.rs:41-48) and builder fields.
    uint32_t al = 0, al_eff = 0;
    int32_t go = 0, ge = 0, min_e = 0, max_e = 0;
    uint32_t mode = 0, search_type = 0;
    uint64_t traceback_max_bytes = 0;  // SearchType::Alignment's budget (0 otherwise)
    std::vector<int32_t> h_matrix;     // the scheme's matrix as given (al x al)
    uint64_t n = 0, mql = 0, mtl = 0, total_residues = 0;
    int widest_bits = 8;
    uint64_t count_by_width[3] = {0, 0, 0};
    // Device copies.
    int32_t* d_matrix = nullptr;   // al x al, entries clamped to +-kEntryMax32
    uint8_t* d_flat = nullptr;     // residues in database order
    uint64_t* d_off = nullptr;     // n+1 offsets
    std::vector<uint64_t> h_off;   // host copy of the offsets (traceback slab and op layout)
    // Inter-sequence partition (one target per lane).
    std::vector<hy::Group> groups;
    hy::Group* d_groups = nullptr;
    uint32_t* d_words = nullptr;
    uint32_t* d_lens = nullptr;
    uint32_t* d_gidx = nullptr;
    uint64_t inter_spill = 0;      // int2 entries
    // Groups [0, groups_score32) hold the targets that are not eligible for the packed kernel;
    // scans run the 32-bit kernel over those and the packed kernel over the rest. ends32_all: a
    // local-mode database with a packed group wider than 65,535 columns, whose end-tracking scans
    // run every non-long target at 32 bits (the groups after groups_score32 hold them again).
    uint32_t groups_score32 = 0;
    bool ends32_all = false;
    uint64_t score32_spill = 0;    // int2 entries of groups [0, groups_score32)
    // Packed 16-bit partition (two targets per lane).
    std::vector<hy::Group> groups16;
    hy::Group* d_groups16 = nullptr;
    uint2* d_words16 = nullptr;
    uint32_t* d_lens16 = nullptr;
    uint32_t* d_gidx16 = nullptr;
    uint64_t inter16_spill = 0;    // uint2 entries
    int q16 = 32;        // query rows per register tile of the packed kernel
    int waves16 = 4;     // wavefronts per workgroup of the packed kernel
    // Groups [0, groups16_short_first) hold long targets; one-query scans start after them.
    uint32_t groups16_short_first = 0;
    // The batched form's persistent kernel runs groups [groups16_batch_first, end); the groups
    // before are wider than kBatchSlotMaxCols and their targets join sys_list_batch.
    uint32_t groups16_batch_first = 0;
    uint32_t batch_max_ncols16 = 0;  // widest group the persistent kernel runs
    uint32_t max_ncols16 = 0;
    uint64_t residues16 = 0;       // residues of the packed targets
    // Systolic list (long targets, or everything when the alphabet leaves no padding symbol).
    std::vector<uint32_t> sys_list;
    uint32_t* d_sys_list = nullptr;
    uint64_t sys_max_len = 0;
    // Targets the packed traceback (k_trace16) does not take, in database order: those outside
    // the packed layout and those in groups wider than 65,535 columns.
    std::vector<uint32_t> trace_rest;
    uint32_t* d_trace_rest = nullptr;
    // Long targets the packed kernel cannot take: the batched form's systolic list.
    std::vector<uint32_t> sys_list_batch;
    uint32_t* d_sys_list_batch = nullptr;
    uint64_t sys_batch_max_len = 0;

    ~hy_db() {
        (void)hipSetDevice(device);
        hy::dev_free(d_matrix); hy::dev_free(d_flat); hy::dev_free(d_off);
        hy::dev_free(d_groups); hy::dev_free(d_words); hy::dev_free(d_lens); hy::dev_free(d_gidx);
        hy::dev_free(d_groups16); hy::dev_free(d_words16); hy::dev_free(d_lens16); hy::dev_free(d_gidx16);
        hy::dev_free(d_sys_list);
        hy::dev_free(d_sys_list_batch);
        hy::dev_free(d_trace_rest);
    }
};

struct hy_scratch {
    int device = 0;
    const hy_db* db = nullptr;  // sized for this database (not owned)
    hipStream_t stream = nullptr;
    uint8_t* d_query = nullptr;
    uint8_t* h_query = nullptr;          // pinned staging
    int2* d_spill = nullptr;
    uint64_t spill_entries = 0;
    int2* d_sys_spill = nullptr;
    uint64_t sys_spill_stride = 0;
    uint64_t sys_spill_entries = 0;
    int32_t* d_scores = nullptr;
    int64_t* d_qend = nullptr;
    int64_t* d_tend = nullptr;
    unsigned long long* d_key = nullptr;
    hy::Result* h_result = nullptr;  // pinned; the one-query scan's last kernel writes it
    // Instrumentation: events around the per-target kernels of the last scan.
    bool profile = false;
    bool pending = false;
    bool pending_tail = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // Auxiliary stream for the long-target kernel, with its fork/join events.
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    double inter_ms = 0.0, tail_ms = 0.0;
    uint64_t launches = 0;
    // Chained pipelined kernel: progress words (one per link of each packed group) and the
    // workgroup ticket counter, zeroed at allocation; the launch epoch; the tickets earlier
    // launches took; and a pinned flag the kernel raises if a wait gave up.
    unsigned long long* d_chain = nullptr;  // progress words, then the ticket counter
    uint2* d_chain_fold = nullptr;  // end-tracking chained scans: best cells passed between chunks
    unsigned int* h_chain_err = nullptr;
    uint32_t chain_epoch = 0;
    uint64_t chain_tickets = 0;             // tickets taken by earlier launches
    int32_t last_kernel = -1;               // HY_KERNEL_* of the last one-query scan
    // Batched form (reserved on first use or by hy_scratch_reserve_batch).
    uint32_t bcap = 0;                   // queries per submission
    uint32_t bq_stride = 0;              // bytes reserved per query (max_query_len, >= 1)
    uint8_t* d_bq = nullptr;             // concatenated queries
    uint8_t* h_bq = nullptr;             // pinned staging
    uint32_t* d_bmeta = nullptr;         // qoff[cap] | qlen[cap] | qorder[cap] | nq | counter
    uint32_t* h_bmeta = nullptr;         // pinned
    unsigned long long* d_bkeys = nullptr;
    hy::Result* d_bres = nullptr;
    hy::Result* h_bres = nullptr;        // pinned
    uint2* d_slots = nullptr;            // persistent kernel boundary rows
    uint64_t slot_stride = 0;
    uint32_t nslots = 0;
    int2* d_bwin_spill = nullptr;        // winner pass, multi-pass queries
    int2* d_btail_spill = nullptr;       // long-target tail, one query at a time
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double batch_ms = 0.0;               // instrumentation: persistent kernel time
    uint64_t batch_launches = 0;
    hipEvent_t bev[2] = {nullptr, nullptr};
    // Launch shapes (hy_tuning), initialised from the database.
    int q16 = 32, waves16 = 1;           // one-query scans
    int bq16 = 32, bwpc = 0;             // batched form: tile rows, workgroups per CU (0 = max)
    uint32_t bchunk = 256;               // batched form: queries per submission
    int bpairs = 1;                      // batched form, local mode: two queries per lane
    uint64_t dev_bytes = 0;              // device memory held by the one-query buffers
    uint64_t batch_bytes = 0;            // device memory held by the batched-form reservation
    // Traceback (grown on first use): move-code slab pool, per-item ends, per-target outputs and
    // op bytes, slab offsets per database index for one rows-per-lane / pass count.
    uint32_t* d_codes = nullptr;
    uint64_t codes_words = 0;
    int4* d_tends = nullptr;
    uint64_t tends_cap = 0;
    uint32_t* d_titem = nullptr;         // packed traceback: each item's group
    uint64_t titem_cap = 0;
    uint64_t* d_gslab = nullptr;         // packed traceback: code words before each group
    std::vector<uint64_t> h_gslab;
    int gslab_m = -1;                    // query length the group offsets were laid out for
    int slab_list = -1;                  // slab offsets laid out over: 1 the rest list, 0 all
    hy::TraceOut* d_tout = nullptr;
    uint64_t tout_cap = 0;
    uint8_t* d_ops = nullptr;
    uint64_t ops_cap = 0;
    uint64_t* d_opsum = nullptr;         // op packing: per-block offsets and the total
    uint64_t opsum_cap = 0;
    hy::PinnedRing ring;  // large pageable copies back (scan_all_aligned)
    uint64_t* d_slab_off = nullptr;
    int slab_m = -1;                     // query length the slab offsets were laid out for
    std::vector<uint64_t> h_slab_off;
    hy::TraceOut* h_tout = nullptr;      // pinned, one entry (winner)
    uint8_t* h_ops = nullptr;            // pinned, max_query_len + max_target_len (winner)
    unsigned long long* h_key = nullptr; // pinned (winner)
    uint64_t trace_bytes = 0;            // device memory held by the traceback buffers

    ~hy_scratch() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        hy::dev_free(d_query); hy::dev_free(d_spill); hy::dev_free(d_sys_spill);
        hy::dev_free(d_scores); hy::dev_free(d_qend); hy::dev_free(d_tend);
        hy::dev_free(d_key);
        hy::dev_free(d_bq); hy::dev_free(d_bmeta); hy::dev_free(d_bkeys); hy::dev_free(d_bres);
        hy::dev_free(d_slots); hy::dev_free(d_bwin_spill); hy::dev_free(d_btail_spill);
        hy::dev_free(d_codes); hy::dev_free(d_tends); hy::dev_free(d_tout); hy::dev_free(d_ops);
        hy::dev_free(d_slab_off); hy::dev_free(d_titem); hy::dev_free(d_gslab); hy::dev_free(d_chain); hy::dev_free(d_chain_fold);
        hy::dev_free(d_opsum);
        if (h_chain_err) (void)hipHostFree(h_chain_err);
        if (h_tout) (void)hipHostFree(h_tout);
        if (h_ops) (void)hipHostFree(h_ops);
        if (h_key) (void)hipHostFree(h_key);
        if (h_bq) (void)hipHostFree(h_bq);
        if (h_bmeta) (void)hipHostFree(h_bmeta);
        if (h_bres) (void)hipHostFree(h_bres);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& e : bev)
            if (e) (void)hipEventDestroy(e);
        if (h_query) (void)hipHostFree(h_query);
        if (h_result) (void)hipHostFree(h_result);
        if (aux) (void)hipStreamSynchronize(aux);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (aux) (void)hipStreamDestroy(aux);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace hy {
// Scan plumbing shared by the host translation units (defined in hy_host.cpp).
// check_query: the scan contract (database.rs:79-89). stage_query: copy the query to the device
// on the scratch's stream. enqueue_targets: the per-target pass (see hy_host.cpp).
// collect_profile: fold the last scan's kernel times into the scratch's counters.
int check_query(const hy_db* db, const uint8_t* q, size_t qlen);
int stage_query(const hy_db* db, hy_scratch* s, const uint8_t* q, size_t qlen);
int enqueue_targets(const hy_db* db, hy_scratch* s, int m, bool ends, bool write_scores,
                    unsigned long long* key);
void collect_profile(hy_scratch* s);
// After a scan's stream synchronisation: collect_profile, then HY_E_INTERNAL if a chained
// kernel's wait gave up (its results are not trusted), else HY_OK.
int scan_done(hy_scratch* s);
// The width proof (src/width.rs:107-169): magnitude bound of a (mode, scheme, lengths) problem
// and the narrowest width holding it (8, 16, 32; 0 when not even i32 suffices).
__int128 magnitude_bound(uint32_t mode, int32_t min_e, int32_t max_e, int32_t go, int32_t ge,
                         uint64_t mql, uint64_t mtl);
int width_bits_for(__int128 bound, int64_t* bound_out);
// The first gfx950 device, or -1.
int first_gfx950_device();
// Scoring::new's checks (src/scoring.rs:65-100) and the matrix's entry bounds.
int validate_scoring(const int32_t* matrix, size_t matrix_len, uint32_t alphabet_len,
                     int32_t gap_open, int32_t gap_ext, int32_t* min_e, int32_t* max_e,
                     int64_t* err_detail);
}  // namespace hy
