// Aligned scans on the device (Database::scan_aligned / scan_all_aligned, src/database.rs:528-580).
// scan_aligned runs the score pass, whose fused argmax key names the winner on the device, then
// the traceback kernels (hy_trace.hip) on that one pair; scan_all_aligned runs the traceback
// kernels over every target in database order, in chunks whose move-code slabs fit the pool.
// Both give the alignment align::traceback gives for the pair (src/align.rs:192-226).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/hyalite_hip.h"
#include "hy_internal.h"
#include "hy_state.h"

static_assert(sizeof(hy::TraceOut) == sizeof(hy_alignment), "TraceOut mirrors hy_alignment");
static_assert(offsetof(hy::TraceOut, n_ops) == offsetof(hy_alignment, n_ops),
              "TraceOut mirrors hy_alignment");

namespace {

constexpr uint64_t kDefaultTracePool = 8ull << 30;  // bytes of move codes (two chunks' worth)

uint64_t trace_pool_bytes() {
    if (const char* v = std::getenv("HYALITE_HIP_TRACE_POOL_BYTES")) {
        char* end = nullptr;
        const unsigned long long b = std::strtoull(v, &end, 10);
        if (end != v) return (uint64_t)b;
    }
    return kDefaultTracePool;
}

// Grow a device buffer to at least `need` elements; the old contents are dropped. Every entry
// point synchronises its stream before returning, so nothing is in flight on the buffer here.
template <typename T>
int grow(T** p, uint64_t* cap, uint64_t need) {
    need = std::max<uint64_t>(need, 1);
    if (*cap >= need) return HY_OK;
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    hy::dev_free(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc((void**)p, need * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return HY_E_OUT_OF_MEMORY;
    }
    *cap = need;
    return HY_OK;
}

void account(hy_scratch* s) {
    s->trace_bytes = s->codes_words * 4 + s->tends_cap * sizeof(int4) +
                     s->titem_cap * sizeof(uint32_t) + s->tout_cap * sizeof(hy::TraceOut) +
                     s->opsum_cap * sizeof(uint64_t) +
                     s->ops_cap + (s->d_slab_off ? (s->db->n + 1) * sizeof(uint64_t) : 0) +
                     (s->d_gslab ? (s->db->groups16.size() + 1) * sizeof(uint64_t) : 0);
}

hy::TraceArgs trace_args(const hy_db* db, hy_scratch* s, int m) {
    hy::TraceArgs a{};
    a.flat = db->d_flat;
    a.off = db->d_off;
    a.query = s->d_query;
    a.m = m;
    a.matrix = db->d_matrix;
    a.al = (int)db->al;
    a.al_eff = (int)db->al_eff;
    a.go = db->go;
    a.ge = db->ge;
    a.rows = hy::trace_rows(m);
    a.codes = s->d_codes;
    a.ends = s->d_tends;
    a.ops = s->d_ops;
    a.out = s->d_tout;
    return a;
}

int aligned_preamble(const hy_db* db, hy_scratch* s, const uint8_t* query, size_t qlen) {
    if (s->db != db) return HY_E_INVALID_ARGUMENT;
    if (db->search_type != HY_SEARCH_ALIGNMENT) return HY_E_NOT_ALIGNMENT;
    return hy::check_query(db, query, qlen);
}

}  // namespace

extern "C" int hy_db_aligned_ops_bytes(const hy_db* db, size_t qlen, uint64_t* bytes) {
    if (!db || !bytes) return HY_E_NULL_ARGUMENT;
    *bytes = db->n * (uint64_t)qlen + db->total_residues;
    return HY_OK;
}

extern "C" int hy_scan_aligned(const hy_db* db, hy_scratch* s, const uint8_t* query, size_t qlen,
                               uint64_t* db_index, hy_alignment* out, uint8_t* ops,
                               size_t ops_cap) {
    if (!db || !s || !db_index || !out || !ops) return HY_E_NULL_ARGUMENT;
    int rc = aligned_preamble(db, s, query, qlen);
    if (rc != HY_OK) return rc;
    if ((uint64_t)ops_cap < (uint64_t)qlen + db->mtl) return HY_E_INVALID_ARGUMENT;
    const int m = (int)qlen;
    if (hipSetDevice(db->device) != hipSuccess) return HY_E_DEVICE;
    {
        std::lock_guard<std::recursive_mutex> api(hy::api_lock());
        if ((rc = grow(&s->d_codes, &s->codes_words, hy::trace_slab_words(m, db->mtl))) !=
                HY_OK ||
            (rc = grow(&s->d_tends, &s->tends_cap, 1)) != HY_OK ||
            (rc = grow(&s->d_tout, &s->tout_cap, 1)) != HY_OK ||
            (rc = grow(&s->d_ops, &s->ops_cap, db->mql + db->mtl)) != HY_OK)
            return rc;
        account(s);
        if (!s->h_tout && hipHostMalloc((void**)&s->h_tout, sizeof(hy::TraceOut)) != hipSuccess)
            return HY_E_OUT_OF_MEMORY;
        if (!s->h_key &&
            hipHostMalloc((void**)&s->h_key, sizeof(unsigned long long)) != hipSuccess)
            return HY_E_OUT_OF_MEMORY;
        if (!s->h_ops && hipHostMalloc((void**)&s->h_ops,
                                       std::max<uint64_t>(db->mql + db->mtl, 1)) != hipSuccess)
            return HY_E_OUT_OF_MEMORY;
    }

    if ((rc = hy::stage_query(db, s, query, qlen)) != HY_OK) return rc;
    if (hy::launch_fill_u64(s->d_key, 1, 0ull, s->stream) != hipSuccess) return HY_E_DEVICE;
    if ((rc = hy::enqueue_targets(db, s, m, false, false, s->d_key)) != HY_OK) return rc;
    hy::TraceArgs a = trace_args(db, s, m);
    a.winner_key = s->d_key;
    a.nitems = 1;
    if (hy::launch_trace((int)db->mode, a, s->stream) != hipSuccess) return HY_E_DEVICE;
    if (hipMemcpyAsync(s->h_tout, s->d_tout, sizeof(hy::TraceOut), hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess ||
        hipMemcpyAsync(s->h_key, s->d_key, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess ||
        hipMemcpyAsync(s->h_ops, s->d_ops, (size_t)(qlen + db->mtl), hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess)
        return HY_E_DEVICE;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
    if ((rc = hy::scan_done(s)) != HY_OK) return rc;
    const hy::TraceOut o = *s->h_tout;
    if (o.check != 0 || o.ops_offset + o.n_ops > (uint64_t)qlen + db->mtl) return HY_E_INTERNAL;
    std::memcpy(ops, s->h_ops + o.ops_offset, (size_t)o.n_ops);
    *db_index = 0xFFFFFFFFull - (*s->h_key & 0xFFFFFFFFull);
    out->score = o.score;
    out->reserved = 0;
    out->query_start = o.query_start;
    out->query_end = o.query_end;
    out->target_start = o.target_start;
    out->target_end = o.target_end;
    out->n_ops = o.n_ops;
    out->ops_offset = 0;
    return HY_OK;
}

// The packed traceback (k_trace16) takes the targets of the packed layout unless
// HYALITE_HIP_TRACE16=0 (then every target runs in the systolic k_trace_dp, as the single-pair
// scan_aligned does).
static bool trace16_enabled(const hy_db* db) {
    const char* e = std::getenv("HYALITE_HIP_TRACE16");
    return !(e && std::atoi(e) == 0) && !db->groups16.empty();
}

extern "C" int hy_scan_all_aligned(const hy_db* db, hy_scratch* s, const uint8_t* query,
                                   size_t qlen, hy_alignment* out, uint8_t* ops, size_t ops_cap) {
    if (!db || !s || !out || !ops) return HY_E_NULL_ARGUMENT;
    int rc = aligned_preamble(db, s, query, qlen);
    if (rc != HY_OK) return rc;
    const uint64_t n = db->n;
    const uint64_t ops_need = n * (uint64_t)qlen + db->total_residues;
    if ((uint64_t)ops_cap < ops_need) return HY_E_INVALID_ARGUMENT;
    const int m = (int)qlen;
    if (hipSetDevice(db->device) != hipSuccess) return HY_E_DEVICE;
    const bool packed = trace16_enabled(db);
    const std::vector<hy::Group>& g16 = db->groups16;
    const uint64_t ng = packed ? g16.size() : 0;
    auto traced16 = [&](uint64_t g) { return g16[g].ncols <= 65535u; };

    // Systolic part: the targets of trace_rest (or every target), slab offsets per list position
    // for this query length (R and the pass count follow m).
    const uint64_t nlist = packed ? db->trace_rest.size() : n;
    auto list_at = [&](uint64_t k) -> uint64_t { return packed ? db->trace_rest[k] : k; };
    if (s->slab_m != m || s->slab_list != (packed ? 1 : 0) || !s->d_slab_off) {
        s->h_slab_off.assign(nlist + 1, 0);
        for (uint64_t k = 0; k < nlist; ++k) {
            const uint64_t t = list_at(k);
            s->h_slab_off[k + 1] =
                s->h_slab_off[k] + hy::trace_slab_words(m, db->h_off[t + 1] - db->h_off[t]);
        }
        std::lock_guard<std::recursive_mutex> api(hy::api_lock());
        if (!s->d_slab_off &&
            hipMalloc((void**)&s->d_slab_off, (n + 1) * sizeof(uint64_t)) != hipSuccess) {
            s->d_slab_off = nullptr;
            return HY_E_OUT_OF_MEMORY;
        }
        if (hipMemcpy(s->d_slab_off, s->h_slab_off.data(), (nlist + 1) * sizeof(uint64_t),
                      hipMemcpyHostToDevice) != hipSuccess) {
            s->slab_m = -1;
            return HY_E_DEVICE;
        }
        s->slab_m = m;
        s->slab_list = packed ? 1 : 0;
    }
    // Packed part: code words before each group for this query length.
    if (packed && (s->gslab_m != m || !s->d_gslab)) {
        s->h_gslab.assign(ng + 1, 0);
        for (uint64_t g = 0; g < ng; ++g)
            s->h_gslab[g + 1] =
                s->h_gslab[g] + (traced16(g) ? hy::trace16_group_words(m, g16[g].ncols) : 0);
        std::lock_guard<std::recursive_mutex> api(hy::api_lock());
        if (!s->d_gslab &&
            hipMalloc((void**)&s->d_gslab, (ng + 1) * sizeof(uint64_t)) != hipSuccess) {
            s->d_gslab = nullptr;
            return HY_E_OUT_OF_MEMORY;
        }
        if (hipMemcpy(s->d_gslab, s->h_gslab.data(), (ng + 1) * sizeof(uint64_t),
                      hipMemcpyHostToDevice) != hipSuccess) {
            s->gslab_m = -1;
            return HY_E_DEVICE;
        }
        s->gslab_m = m;
    }
    const std::vector<uint64_t>& so = s->h_slab_off;
    const std::vector<uint64_t>& gs = s->h_gslab;
    // The pool holds the configured bytes, at least one target's slab or one group's codes, and
    // at most what the whole run needs.
    uint64_t one = nlist ? hy::trace_slab_words(m, db->mtl) : 0;
    uint64_t total = so[nlist];
    for (uint64_t g = 0; g < ng; ++g)
        if (traced16(g)) one = std::max<uint64_t>(one, gs[g + 1] - gs[g]);
    if (ng) total = std::max<uint64_t>(total, gs[ng]);
    const uint64_t pool = std::min<uint64_t>(total, std::max<uint64_t>(one, trace_pool_bytes() / 4));
    if ((rc = grow(&s->d_codes, &s->codes_words, std::max<uint64_t>(pool, 1))) != HY_OK ||
        (rc = grow(&s->d_tends, &s->tends_cap, n)) != HY_OK ||
        (rc = grow(&s->d_titem, &s->titem_cap, packed ? n : 1)) != HY_OK ||
        (rc = grow(&s->d_tout, &s->tout_cap, n)) != HY_OK ||
        (rc = grow(&s->d_ops, &s->ops_cap, std::max<uint64_t>(ops_need, 1))) != HY_OK ||
        (rc = grow(&s->d_opsum, &s->opsum_cap, hy::ops_sum_words(n))) != HY_OK)
        return rc;
    account(s);
    if (!s->h_key) {
        std::lock_guard<std::recursive_mutex> api(hy::api_lock());
        if (hipHostMalloc((void**)&s->h_key, sizeof(unsigned long long)) != hipSuccess)
            return HY_E_OUT_OF_MEMORY;
    }

    if ((rc = hy::stage_query(db, s, query, qlen)) != HY_OK) return rc;
    // Packed chunks run on the scratch's two streams, each stream with its own half of the pool,
    // so a chunk's DP starts while the other stream's last groups and walk finish; a half's next
    // chunk follows that half's walk in its stream's order. The systolic chunks run in stream
    // order after them, each DP after the previous chunk's walk.
    if (ng) {
        hy::TraceArgs a = trace_args(db, s, m);
        // Wavefronts per group: one query tile each, as many as kTrace16MaxWaves, the tiles and
        // the workgroup's LDS allow (HYALITE_HIP_TRACE16_WAVES=1..8 asks for fewer).
        int want = hy::kTrace16MaxWaves;
        if (const char* v = std::getenv("HYALITE_HIP_TRACE16_WAVES")) {
            const int x = std::atoi(v);
            if (x >= 1 && x <= hy::kTrace16MaxWaves) want = x;
        }
        a.rows = hy::trace16_waves((int)db->al_eff, m, want);
        a.groups16 = db->d_groups16;
        a.words16 = db->d_words16;
        a.lens16 = db->d_lens16;
        a.gidx16 = db->d_gidx16;
        a.spill16 = reinterpret_cast<uint2*>(s->d_spill);
        a.gslab = s->d_gslab;
        a.item_group = s->d_titem;
        // Chunks of consecutive groups whose codes fit the pool (at least one group each); a
        // group the packed traceback does not take ends a chunk.
        const uint64_t half = s->codes_words / 2;
        const bool two = half >= one;  // each half holds the largest group's codes
        const uint64_t cap = two ? half : s->codes_words;
        if (two && (hipEventRecord(s->fork, s->stream) != hipSuccess ||
                    hipStreamWaitEvent(s->aux, s->fork, 0) != hipSuccess))
            return HY_E_DEVICE;
        // Each chunk goes to the stream whose queued chunks are estimated to end first (its pool
        // half with it). The estimate, in wavefront column steps: the longer of the chunk's
        // longest group (passes of its columns plus the pipeline fill, times the wavefronts that
        // share a SIMD) and its tile columns spread over the device's SIMDs, which the two
        // streams share; plus the walk, about a fiftieth of a step per target. Groups are
        // sorted longest first, so the first chunk is the one its longest group bounds.
        const int W = a.rows;
        const uint64_t tiles = (uint64_t)(m + hy::kTrace16Rows - 1) / hy::kTrace16Rows;
        const uint64_t passes = (tiles + W - 1) / W;
        const double simds = 4.0 * hy::device_cus();
        double queued[2] = {0.0, 0.0};
        for (uint64_t g0 = 0; g0 < ng;) {
            if (!traced16(g0)) {
                ++g0;
                continue;
            }
            uint64_t g1 = g0 + 1;
            while (g1 < ng && traced16(g1) && gs[g1 + 1] - gs[g0] <= cap) ++g1;
            double tile_cols = 0.0;
            for (uint64_t g = g0; g < g1; ++g) tile_cols += (double)tiles * g16[g].ncols;
            const double longest =
                (double)passes * (g16[g0].ncols + W - 1) * std::max(1.0, W / 4.0);
            const double items = (double)(g16[g1 - 1].first + g16[g1 - 1].count - g16[g0].first);
            const double est = std::max(longest, 2.0 * tile_cols / simds) + items / 50.0;
            const bool odd = two && queued[1] < queued[0];
            queued[odd ? 1 : 0] += est;
            a.codes = s->d_codes + (odd ? half : 0);
            a.group_first = (uint32_t)g0;
            a.ngroups = (uint32_t)(g1 - g0);
            a.slot0 = g16[g0].first;
            a.nitems = g16[g1 - 1].first + g16[g1 - 1].count - a.slot0;
            a.gslab_base = gs[g0];
            if (hy::launch_trace16((int)db->mode, a, odd ? s->aux : s->stream) != hipSuccess)
                return HY_E_DEVICE;
            g0 = g1;
        }
        if (two && (hipEventRecord(s->join, s->aux) != hipSuccess ||
                    hipStreamWaitEvent(s->stream, s->join, 0) != hipSuccess))
            return HY_E_DEVICE;
    }
    if (nlist) {
        hy::TraceArgs a = trace_args(db, s, m);
        a.slab_off = s->d_slab_off;
        a.list = packed ? db->d_trace_rest : nullptr;
        // Chunks of consecutive list entries whose slabs fit the pool (at least one each).
        for (uint64_t k0 = 0; k0 < nlist;) {
            uint64_t k1 = k0 + 1;
            while (k1 < nlist && so[k1 + 1] - so[k0] <= s->codes_words) ++k1;
            a.first = (uint32_t)k0;
            a.nitems = (uint32_t)(k1 - k0);
            a.slab_base = so[k0];
            if (hy::launch_trace((int)db->mode, a, s->stream) != hipSuccess) return HY_E_DEVICE;
            k0 = k1;
        }
    }
    // The ops are laid end to end in database order in the code pool (free once the walks have
    // ended), so that only their bytes come back instead of every target's whole op region.
    // The total is read first: the pool grows when it cannot hold them.
    if (hy::launch_ops_sums(s->d_tout, n, s->d_opsum, s->stream) != hipSuccess ||
        hipMemcpyAsync(s->h_key, s->d_opsum + hy::ops_blocks(n), sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return HY_E_DEVICE;
    const uint64_t packed_bytes = *s->h_key;
    if (packed_bytes > ops_need) return HY_E_INTERNAL;
    if ((rc = grow(&s->d_codes, &s->codes_words, (packed_bytes + 3) / 4)) != HY_OK) return rc;
    account(s);
    uint8_t* d_packed = reinterpret_cast<uint8_t*>(s->d_codes);
    if (hy::launch_ops_pack(s->d_tout, n, s->d_opsum, s->d_ops, d_packed, s->stream) !=
        hipSuccess)
        return HY_E_DEVICE;
    // The caller's buffers are pageable: they come back through the pinned ring when large.
    const hy::Segment segs[2] = {{out, s->d_tout, n * sizeof(hy::TraceOut)},
                                 {ops, d_packed, (size_t)packed_bytes}};
    if ((rc = hy::download(s->ring, s->device, segs, 2, s->stream)) != HY_OK) return rc;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
    for (uint64_t k = 0; k < n; ++k) {
        if (out[k].reserved != 0) return HY_E_INTERNAL;
        if (out[k].ops_offset + out[k].n_ops > packed_bytes) return HY_E_INTERNAL;
    }
    return HY_OK;
}
