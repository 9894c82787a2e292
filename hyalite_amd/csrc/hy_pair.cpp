// Pairwise alignment on the device: align_pair, align_pairs, align_pair_position_max and score2
// (src/kernel.rs:285-575, 741-792). No database: every call carries its scoring scheme and
// sequences, stages them through a pinned buffer of the caller's hy_pair_scratch and runs
//   * one pair, or the long pairs of a batch: k_sys32 in pair mode (one wavefront per pair);
//   * the rest of a large batch: k_pair_pack + k_pairs32 (one pair per lane, hy_pair.hip);
//   * position maxima: k_posmax (one wavefront).
// Validation follows the reference: Scoring::new's checks, then per pair the symbols (query, then
// target) and the width proof for the pair's own lengths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/hyalite_hip.h"
#include "hy_internal.h"
#include "hy_state.h"

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct hy_pair_scratch {
    int device = 0;
    hipStream_t stream = nullptr, aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around a batch's device work
    double device_ms = 0.0;
    uint64_t calls = 0;
    DevBuf mat, qflat, tflat, qoff, toff, qoff32, qlen32, list, nq, groups, twords, qwords, tlens,
        qlens, pidx, spill, sys_spill, scores, qend, tend, colmax, best, arena;
    hy::PinnedRing ring;  // large pageable uploads
    uint8_t* stage = nullptr;  // pinned
    size_t stage_bytes = 0;

    ~hy_pair_scratch() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (aux) (void)hipStreamSynchronize(aux);
        for (DevBuf* b : {&mat, &qflat, &tflat, &qoff, &toff, &qoff32, &qlen32, &list, &nq,
                          &groups, &twords, &qwords, &tlens, &qlens, &pidx, &spill, &sys_spill,
                          &scores, &qend, &tend, &colmax, &best, &arena})
            hy::dev_free(b->p);
        if (stage) (void)hipHostFree(stage);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (aux) (void)hipStreamDestroy(aux);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Pairs past these lengths always run one per wavefront: a lane would sweep too many cells.
constexpr size_t kLaneMaxQuery = 4096;
constexpr size_t kLaneMaxTarget = 65536;
// A batch runs one pair per lane when it has at least this many short pairs (enough 64-lane
// groups to keep every SIMD busy); smaller batches run one pair per wavefront.
constexpr size_t kLaneMinPairs = 32768;
// Boundary rows of one launch of long multi-pass pairs.
constexpr size_t kSysSpillBudget = 512ull << 20;
// Returned by the exact per-pair checks when every pair passes.
constexpr int kChecksPassed = -1;

// Every device buffer is allocated this many bytes past its capacity: the pack kernels read
// sequences a word at a time and may touch up to 7 bytes past a sequence's last one.
constexpr size_t kDevSlack = 16;

// Grow a device buffer (contents dropped; every entry point synchronises before returning).
int grow(DevBuf& b, size_t bytes) {
    bytes = std::max<size_t>(bytes, 16);
    if (b.bytes >= bytes) return HY_OK;
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    hy::dev_free(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (hipMalloc(&b.p, bytes + kDevSlack) != hipSuccess) {
        b.p = nullptr;
        return HY_E_OUT_OF_MEMORY;
    }
    b.bytes = bytes;
    return HY_OK;
}

int grow_stage(hy_pair_scratch* s, size_t bytes) {
    bytes = std::max<size_t>(bytes, 64);
    if (s->stage_bytes >= bytes) return HY_OK;
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    if (s->stage) (void)hipHostFree(s->stage);
    s->stage = nullptr;
    s->stage_bytes = 0;
    if (hipHostMalloc((void**)&s->stage, bytes) != hipSuccess) {
        s->stage = nullptr;
        return HY_E_OUT_OF_MEMORY;
    }
    s->stage_bytes = bytes;
    return HY_OK;
}

// Sequential layout of one call's host-to-device data inside the pinned stage.
struct Stager {
    hy_pair_scratch* s;
    size_t used = 0;
    // symbols: the piece holds sequence symbols; commit() folds their largest byte into
    // max_symbol (on the host, while or before copying them).
    struct Piece { size_t off, bytes; DevBuf* dst; const void* ext; bool symbols; };
    std::vector<Piece> pieces;
    uint8_t max_symbol = 0;
    // Reserve `bytes` for `dst` (16-byte aligned) and return the host pointer to fill later.
    size_t add(DevBuf* dst, size_t bytes, bool symbols = false) {
        const size_t off = (used + 15) & ~size_t(15);
        pieces.push_back({off, bytes, dst, nullptr, symbols});
        used = off + bytes;
        return off;
    }
    // Upload `bytes` from caller memory straight to `dst`, without passing through the stage.
    void add_external(DevBuf* dst, const void* src, size_t bytes, bool symbols = false) {
        pieces.push_back({0, bytes, dst, src, symbols});
    }
    uint8_t* at(size_t off) { return s->stage + off; }
    int commit() {
        for (const Piece& p : pieces) {
            int rc = grow(*p.dst, p.bytes);
            if (rc != HY_OK) return rc;
        }
        for (const Piece& p : pieces) {
            if (p.ext && p.bytes >= hy::kPinnedRingMin) {
                uint8_t mx = 0;
                const int rc = hy::upload_pageable(s->ring, s->device, p.dst->p, p.ext, p.bytes,
                                                   s->stream, p.symbols ? &mx : nullptr);
                if (rc != HY_OK) return rc;
                max_symbol = std::max(max_symbol, mx);
                continue;
            }
            const void* from = p.ext ? p.ext : static_cast<const void*>(s->stage + p.off);
            if (p.symbols)
                max_symbol = std::max(
                    max_symbol, hy::host_max_byte(static_cast<const uint8_t*>(from), p.bytes));
            if (p.bytes && hipMemcpyAsync(p.dst->p, from, p.bytes, hipMemcpyHostToDevice,
                                          s->stream) != hipSuccess)
                return HY_E_DEVICE;
        }
        return HY_OK;
    }
};

struct Scheme {
    int32_t min_e = 0, max_e = 0;
};

int check_scheme(const int32_t* matrix, size_t matrix_len, uint32_t al, int32_t go, int32_t ge,
                 Scheme* sc, int64_t* err_detail) {
    return hy::validate_scoring(matrix, matrix_len, al, go, ge, &sc->min_e, &sc->max_e,
                                err_detail);
}

// The per-pair checks of align_pair_with (src/kernel.rs:362-374): symbols (query, then target),
// then the width proof for the pair's lengths.
int check_pair(const uint8_t* q, size_t m, const uint8_t* t, size_t n, uint32_t al, uint32_t mode,
               int32_t go, int32_t ge, const Scheme& sc, int64_t* err_detail) {
    if ((m && !q) || (n && !t)) return HY_E_NULL_ARGUMENT;
    for (size_t k = 0; k < m; ++k)
        if ((uint32_t)q[k] >= al) {
            if (err_detail) *err_detail = q[k];
            return HY_E_SYMBOL_OUT_OF_RANGE;
        }
    for (size_t k = 0; k < n; ++k)
        if ((uint32_t)t[k] >= al) {
            if (err_detail) *err_detail = t[k];
            return HY_E_SYMBOL_OUT_OF_RANGE;
        }
    int64_t bound = 0;
    if (hy::width_bits_for(hy::magnitude_bound(mode, sc.min_e, sc.max_e, go, ge, m, n), &bound) ==
        0) {
        if (err_detail) *err_detail = bound;
        return HY_E_SCORE_RANGE_TOO_WIDE;
    }
    if (m > 0x7FFFFFF0ull || n > 0x7FFFFFF0ull) return HY_E_INVALID_ARGUMENT;
    return HY_OK;
}

// The matrix as the 32-bit kernels take it: entries clamped to +-kEntryMax32.
void stage_matrix(Stager& st, hy_pair_scratch* s, const int32_t* matrix, size_t len) {
    const size_t off = st.add(&s->mat, len * sizeof(int32_t));
    int32_t* d = reinterpret_cast<int32_t*>(st.at(off));
    for (size_t k = 0; k < len; ++k)
        d[k] = std::max(-hy::kEntryMax32, std::min(hy::kEntryMax32, matrix[k]));
}

bool tracks_end(uint32_t search_type) { return search_type != HY_SEARCH_SCORE; }

// HYALITE_HIP_PAIRS_KERNEL: "lanes" puts every pair of a batch one per lane (k_pairs32),
// "waves" runs every pair one per wavefront (k_sys32 pair mode), "band" one per multi-wavefront
// workgroup (k_band32, alphabets up to kBandMaxAlphabet). Unset: the choice below.
enum { kPickAuto = 0, kPickLanes = 1, kPickWaves = 2, kPickBand = 3 };
int kernel_choice() {
    const char* v = std::getenv("HYALITE_HIP_PAIRS_KERNEL");
    if (!v) return kPickAuto;
    if (std::strcmp(v, "lanes") == 0) return kPickLanes;
    if (std::strcmp(v, "waves") == 0) return kPickWaves;
    if (std::strcmp(v, "band") == 0) return kPickBand;
    return kPickAuto;
}

bool band_eligible(uint32_t al) { return al >= 1 && al <= hy::kBandMaxAlphabet; }

// HYALITE_HIP_PAIRS16: 0 keeps lane pairs one per lane, 2 puts every eligible pair two per lane,
// anything else (the default) lets the batch size decide for alphabets beyond the pair table.
enum Pairs16Choice { kPairs16Off, kPairs16Auto, kPairs16Always };
Pairs16Choice pairs16_choice() {
    const char* v = std::getenv("HYALITE_HIP_PAIRS16");
    if (v && std::strcmp(v, "0") == 0) return kPairs16Off;
    if (v && std::strcmp(v, "2") == 0) return kPairs16Always;
    return kPairs16Auto;
}

// Pairs that do not run one per lane go to the band kernel when the alphabet fits it, except
// score-only batches of more than kBandMaxPairs: there one wavefront per pair fills the chip and
// its leaner score-only loop wins. Measured (device time, SW, DNA): 16,384 pairs of 150 x 1,400
// take 2.23 ms one per wavefront vs 2.71 ms in bands (Score) but 3.29 vs 2.94 ms with ends;
// 2,048 pairs of 1,000 x 5,000 take 5.6 ms in bands vs 6.2 ms (Score) and 6.0 vs 10.3 ms (ends).
constexpr size_t kBandMaxPairs = 4096;
bool use_band(int pick, uint32_t al, size_t npairs, bool ends) {
    if (!band_eligible(al) || pick == kPickWaves) return false;
    return pick == kPickBand || ends || npairs <= kBandMaxPairs;
}

// Band kernel for the pairs in `idx` (ordered by rows-per-pass class): one launch per class,
// the classes whose queries exceed one pass chunked so their boundary rows fit
// kSysSpillBudget. `base` carries the device sequences, matrix and outputs.
int enqueue_band(hy_pair_scratch* s, const std::vector<uint32_t>& idx,
                 const std::vector<size_t>& qlens, const std::vector<size_t>& tlens,
                 const hy::BandArgs& base, const uint32_t* d_list, uint32_t mode, bool ends,
                 hipStream_t stream) {
    size_t b = 0;
    while (b < idx.size()) {
        const int c = hy::band_rows_per_pass((int)qlens[idx[b]]);
        size_t e = b, maxm = 0, maxn = 0;
        while (e < idx.size() && hy::band_rows_per_pass((int)qlens[idx[e]]) == c) {
            maxm = std::max(maxm, qlens[idx[e]]);
            maxn = std::max(maxn, tlens[idx[e]]);
            ++e;
        }
        const bool multipass = maxm > (size_t)c;
        const size_t per = multipass ? (maxn + 1) * sizeof(int2) : 0;
        const size_t chunk = multipass ? std::max<size_t>(1, kSysSpillBudget / per) : (e - b);
        if (multipass) {
            const int rc = grow(s->sys_spill, std::min(chunk, e - b) * per);
            if (rc != HY_OK) return rc;
        }
        for (size_t c0 = b; c0 < e; c0 += chunk) {
            const size_t cnt = std::min(chunk, e - c0);
            hy::BandArgs a = base;
            a.list = d_list + c0;
            a.spill = static_cast<int2*>(s->sys_spill.p);
            a.spill_stride = multipass ? maxn + 1 : 0;
            if (hy::launch_band32((int)mode, ends ? hy::kBandEnds : hy::kBandScore, a,
                                  (uint32_t)cnt, (int)maxm, stream) != hipSuccess)
                return HY_E_DEVICE;
        }
        b = e;
    }
    return HY_OK;
}

// One pair through the band kernel with one upload and one copy back. The arena (device) and
// its image at the start of the pinned stage hold: target offsets {0, n} at 0; list {0}, query
// offset {0} and length {m} at 16; the matrix (clamped for the 32-bit kernels) at 32; the query
// and the target; then, 16-byte aligned, the output: out_bytes written by the kernel and copied
// back into the stage at the same offset.
struct Single {
    size_t q = 0, t = 0, out = 0, in_bytes = 0, total = 0;
};

int band_single(hy_pair_scratch* s, const uint8_t* query, size_t qlen, const uint8_t* target,
                size_t tlen, const int32_t* matrix, size_t matrix_len, uint32_t al, int32_t go,
                int32_t ge, uint32_t mode, int out, size_t out_bytes, Single* lay) {
    Single L;
    L.q = 32 + matrix_len * sizeof(int32_t);
    L.t = L.q + qlen;
    L.in_bytes = L.t + tlen;
    L.out = (L.in_bytes + 15) & ~size_t(15);
    L.total = L.out + out_bytes;
    int rc;
    if ((rc = grow_stage(s, L.total)) != HY_OK || (rc = grow(s->arena, L.total)) != HY_OK)
        return rc;
    const bool multipass = qlen > (size_t)hy::band_rows_per_pass((int)qlen);
    if (multipass && (rc = grow(s->sys_spill, (tlen + 1) * sizeof(int2))) != HY_OK) return rc;
    uint8_t* h = s->stage;
    const uint64_t off2[2] = {0, (uint64_t)tlen};
    const uint32_t meta[4] = {0, 0, (uint32_t)qlen, 0};
    std::memcpy(h, off2, sizeof(off2));
    std::memcpy(h + 16, meta, sizeof(meta));
    int32_t* hm = reinterpret_cast<int32_t*>(h + 32);
    for (size_t k = 0; k < matrix_len; ++k)
        hm[k] = std::max(-hy::kEntryMax32, std::min(hy::kEntryMax32, matrix[k]));
    if (qlen) std::memcpy(h + L.q, query, qlen);
    if (tlen) std::memcpy(h + L.t, target, tlen);
    uint8_t* d = static_cast<uint8_t*>(s->arena.p);
    if (hipMemcpyAsync(d, h, L.in_bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return HY_E_DEVICE;
    hy::BandArgs a{};
    a.queries = d + L.q;
    a.qoff = reinterpret_cast<const uint32_t*>(d + 20);
    a.qlen = reinterpret_cast<const uint32_t*>(d + 24);
    a.targets = d + L.t;
    a.toff = reinterpret_cast<const uint64_t*>(d);
    a.list = reinterpret_cast<const uint32_t*>(d + 16);
    a.matrix = reinterpret_cast<const int32_t*>(d + 32);
    a.al = (int)al;
    a.go = go;
    a.ge = ge;
    a.spill = static_cast<int2*>(s->sys_spill.p);
    a.spill_stride = 0;
    if (out == hy::kBandColmax) {
        a.best = reinterpret_cast<int2*>(d + L.out);
        a.colmax = reinterpret_cast<int32_t*>(d + L.out + 8);
    } else {
        a.out_score = reinterpret_cast<int32_t*>(d + L.out);
        a.out_qend = reinterpret_cast<int64_t*>(d + L.out + 8);
        a.out_tend = reinterpret_cast<int64_t*>(d + L.out + 16);
    }
    if (hy::launch_band32((int)mode, out, a, 1, (int)qlen, s->stream) != hipSuccess)
        return HY_E_DEVICE;
    if (hipMemcpyAsync(h + L.out, d + L.out, out_bytes, hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return HY_E_DEVICE;
    *lay = L;
    return HY_OK;
}

// One wavefront per pair for the pairs in `idx` (k_sys32 pair mode): one launch per rows-per-
// lane class (query up to 128, 256, 384 rows, and longer), the longest class in chunks whose
// boundary rows fit kSysSpillBudget.
int enqueue_waves(hy_pair_scratch* s, const std::vector<uint32_t>& idx,
                  const std::vector<size_t>& qlens, const std::vector<size_t>& tlens,
                  const uint32_t* d_list, uint32_t al, int32_t go, int32_t ge, uint32_t mode,
                  bool ends, hipStream_t stream) {
    // idx is ordered by class; find the class boundaries.
    auto cls = [&](uint32_t k) { const size_t m = qlens[k];
                                 return m <= 128 ? 0 : (m <= 256 ? 1 : (m <= 384 ? 2 : 3)); };
    size_t b = 0;
    while (b < idx.size()) {
        size_t e = b;
        const int c = cls(idx[b]);
        size_t maxm = 0, maxn = 0;
        while (e < idx.size() && cls(idx[e]) == c) {
            maxm = std::max(maxm, qlens[idx[e]]);
            maxn = std::max(maxn, tlens[idx[e]]);
            ++e;
        }
        const bool multipass = maxm > 512;
        const size_t per = multipass ? (maxn + 1) * sizeof(int2) : 0;
        const size_t chunk = multipass ? std::max<size_t>(1, kSysSpillBudget / per) : (e - b);
        if (multipass) {
            const int rc = grow(s->sys_spill, std::min(chunk, e - b) * per);
            if (rc != HY_OK) return rc;
        }
        for (size_t c0 = b; c0 < e; c0 += chunk) {
            const size_t cnt = std::min(chunk, e - c0);
            hy::SysArgs a{};
            a.flat = static_cast<const uint8_t*>(s->tflat.p);
            a.off = static_cast<const uint64_t*>(s->toff.p);
            a.list = d_list + c0;
            a.m = (int)maxm;
            a.matrix = static_cast<const int32_t*>(s->mat.p);
            a.al = (int)al;
            a.al_eff = (int)std::min<uint32_t>(al, 256);
            a.go = go;
            a.ge = ge;
            a.spill = static_cast<int2*>(s->sys_spill.p);
            a.spill_stride = maxn + 1;
            a.out_score = static_cast<int32_t*>(s->scores.p);
            a.out_qend = static_cast<int64_t*>(s->qend.p);
            a.out_tend = static_cast<int64_t*>(s->tend.p);
            a.queries = static_cast<const uint8_t*>(s->qflat.p);
            a.qoff = static_cast<const uint32_t*>(s->qoff32.p);
            a.qlen = static_cast<const uint32_t*>(s->qlen32.p);
            a.nq = static_cast<const uint32_t*>(s->nq.p);
            a.qsel = -1;
            a.pairs = 1;
            if (hy::launch_sys32((int)mode, ends, false, a, (uint32_t)cnt, stream) != hipSuccess)
                return HY_E_DEVICE;
        }
        b = e;
    }
    return HY_OK;
}

}  // namespace

extern "C" int hy_pair_scratch_new(hy_pair_scratch** out) {
    if (!out) return HY_E_NULL_ARGUMENT;
    *out = nullptr;
    const int device = hy::first_gfx950_device();
    if (device < 0) return HY_E_BACKEND_UNAVAILABLE;
    if (hipSetDevice(device) != hipSuccess) return HY_E_DEVICE;
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    hy_pair_scratch* s = new (std::nothrow) hy_pair_scratch();
    if (!s) return HY_E_OUT_OF_MEMORY;
    s->device = device;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
        delete s;
        return HY_E_DEVICE;
    }
    *out = s;
    return HY_OK;
}

extern "C" void hy_pair_scratch_free(hy_pair_scratch* s) {
    if (!s) return;
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    delete s;
}

extern "C" int hy_align_pair(hy_pair_scratch* s, const uint8_t* query, size_t qlen,
                             const uint8_t* target, size_t tlen, const int32_t* matrix,
                             size_t matrix_len, uint32_t alphabet_len, int32_t gap_open,
                             int32_t gap_ext, uint32_t mode, uint32_t search_type, int32_t* score,
                             int64_t* query_end, int64_t* target_end, int64_t* err_detail) {
    if (!s || !score || !query_end || !target_end) return HY_E_NULL_ARGUMENT;
    if (err_detail) *err_detail = 0;
    if (mode > HY_MODE_SHW || search_type > HY_SEARCH_ALIGNMENT) return HY_E_INVALID_ARGUMENT;
    Scheme sc;
    int rc = check_scheme(matrix, matrix_len, alphabet_len, gap_open, gap_ext, &sc, err_detail);
    if (rc != HY_OK) return rc;
    rc = check_pair(query, qlen, target, tlen, alphabet_len, mode, gap_open, gap_ext, sc,
                    err_detail);
    if (rc != HY_OK) return rc;
    if (hipSetDevice(s->device) != hipSuccess) return HY_E_DEVICE;
    const bool ends = tracks_end(search_type);
    if (band_eligible(alphabet_len) && kernel_choice() != kPickWaves) {
        Single L;
        if ((rc = band_single(s, query, qlen, target, tlen, matrix, matrix_len, alphabet_len,
                              gap_open, gap_ext, mode, ends ? hy::kBandEnds : hy::kBandScore, 24,
                              &L)) != HY_OK)
            return rc;
        const uint8_t* o = s->stage + L.out;
        int32_t sv = 0;
        int64_t ev[2] = {-1, -1};
        std::memcpy(&sv, o, sizeof(sv));
        std::memcpy(ev, o + 8, sizeof(ev));
        *score = sv;
        *query_end = ends ? ev[0] : -1;
        *target_end = ends ? ev[1] : -1;
        return HY_OK;
    }

    Stager st{s};
    const size_t o_q = st.add(&s->qflat, qlen);
    const size_t o_t = st.add(&s->tflat, tlen);
    const size_t o_off = st.add(&s->toff, 2 * sizeof(uint64_t));
    const size_t o_list = st.add(&s->list, sizeof(uint32_t));
    const size_t o_qoff = st.add(&s->qoff32, sizeof(uint32_t));
    const size_t o_qlen = st.add(&s->qlen32, sizeof(uint32_t));
    const size_t o_nq = st.add(&s->nq, sizeof(uint32_t));
    const size_t mat_at = st.used;
    (void)mat_at;
    if ((rc = grow_stage(s, st.used + matrix_len * sizeof(int32_t) + 64)) != HY_OK) return rc;
    stage_matrix(st, s, matrix, matrix_len);
    if (qlen) std::memcpy(st.at(o_q), query, qlen);
    if (tlen) std::memcpy(st.at(o_t), target, tlen);
    const uint64_t off2[2] = {0, (uint64_t)tlen};
    std::memcpy(st.at(o_off), off2, sizeof(off2));
    const uint32_t zero = 0, one = 1, m32 = (uint32_t)qlen;
    std::memcpy(st.at(o_list), &zero, 4);
    std::memcpy(st.at(o_qoff), &zero, 4);
    std::memcpy(st.at(o_qlen), &m32, 4);
    std::memcpy(st.at(o_nq), &one, 4);
    if ((rc = st.commit()) != HY_OK) return rc;
    if ((rc = grow(s->scores, sizeof(int32_t))) != HY_OK ||
        (rc = grow(s->qend, sizeof(int64_t))) != HY_OK ||
        (rc = grow(s->tend, sizeof(int64_t))) != HY_OK)
        return rc;
    std::vector<uint32_t> idx{0};
    std::vector<size_t> ql{qlen}, tl{tlen};
    if ((rc = enqueue_waves(s, idx, ql, tl, static_cast<const uint32_t*>(s->list.p),
                            alphabet_len, gap_open, gap_ext, mode, ends, s->stream)) != HY_OK)
        return rc;
    // Results come back through the stage (its inputs are already on the device in stream order).
    int32_t* h_s = reinterpret_cast<int32_t*>(s->stage);
    int64_t* h_e = reinterpret_cast<int64_t*>(s->stage + 16);
    if (hipMemcpyAsync(h_s, s->scores.p, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess ||
        hipMemcpyAsync(h_e, s->qend.p, sizeof(int64_t), hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess ||
        hipMemcpyAsync(h_e + 1, s->tend.p, sizeof(int64_t), hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess)
        return HY_E_DEVICE;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
    *score = *h_s;
    *query_end = ends ? h_e[0] : -1;
    *target_end = ends ? h_e[1] : -1;
    return HY_OK;
}

namespace {

// The sequences of a batch: per-pair pointers and lengths, or one buffer per side with n + 1
// offsets.
struct PairSource {
    const uint8_t* const* qp = nullptr;
    const size_t* qlen = nullptr;
    const uint8_t* const* tp = nullptr;
    const size_t* tlen = nullptr;
    const uint8_t* qflat = nullptr;
    const uint64_t* qoff = nullptr;
    const uint8_t* tflat = nullptr;
    const uint64_t* toff = nullptr;
    size_t m(size_t k) const { return qp ? qlen[k] : (size_t)(qoff[k + 1] - qoff[k]); }
    size_t n(size_t k) const { return tp ? tlen[k] : (size_t)(toff[k + 1] - toff[k]); }
    const uint8_t* q(size_t k) const { return qp ? qp[k] : qflat + qoff[k]; }
    const uint8_t* t(size_t k) const { return tp ? tp[k] : tflat + toff[k]; }
};

// Small batches (sequences up to kSmallBatchBytes) in the band kernel with one upload and one
// copy back: the arena holds target offsets, query offsets and lengths, the pair order (by rows
// class), the clamped matrix, both sequence buffers and then the outputs. Symbols were checked
// on the host.
constexpr uint64_t kSmallBatchBytes = 1ull << 20;

int small_band_batch(hy_pair_scratch* s, const PairSource& src, size_t npairs,
                     const std::vector<size_t>& ql, const std::vector<size_t>& tl, uint64_t qtot,
                     uint64_t ttot, const int32_t* matrix, size_t matrix_len, uint32_t al,
                     int32_t go, int32_t ge, uint32_t mode, bool ends, int32_t* scores,
                     int64_t* query_ends, int64_t* target_ends) {
    std::vector<uint32_t> idx(npairs);
    for (uint32_t k = 0; k < (uint32_t)npairs; ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        return hy::band_rows_per_pass((int)ql[a]) < hy::band_rows_per_pass((int)ql[b]);
    });
    size_t used = 0;
    auto take = [&](size_t bytes) {
        const size_t o = (used + 15) & ~size_t(15);
        used = o + bytes;
        return o;
    };
    const size_t o_toff = take((npairs + 1) * sizeof(uint64_t));
    const size_t o_qoff = take(npairs * sizeof(uint32_t));
    const size_t o_qlen = take(npairs * sizeof(uint32_t));
    const size_t o_list = take(npairs * sizeof(uint32_t));
    const size_t o_mat = take(matrix_len * sizeof(int32_t));
    const size_t o_q = take(qtot);
    const size_t o_t = take(ttot);
    const size_t in_bytes = used;
    const size_t o_sc = take(npairs * sizeof(int32_t));
    const size_t o_qe = take(npairs * sizeof(int64_t));
    const size_t o_te = take(npairs * sizeof(int64_t));
    int rc;
    if ((rc = grow_stage(s, used)) != HY_OK || (rc = grow(s->arena, used)) != HY_OK) return rc;
    uint8_t* h = s->stage;
    uint64_t* toff = reinterpret_cast<uint64_t*>(h + o_toff);
    uint32_t* qoff = reinterpret_cast<uint32_t*>(h + o_qoff);
    uint32_t* qlen = reinterpret_cast<uint32_t*>(h + o_qlen);
    uint64_t qa = 0, ta = 0;
    for (size_t k = 0; k < npairs; ++k) {
        qoff[k] = (uint32_t)qa;
        qlen[k] = (uint32_t)ql[k];
        toff[k] = ta;
        if (ql[k]) std::memcpy(h + o_q + qa, src.q(k), ql[k]);
        if (tl[k]) std::memcpy(h + o_t + ta, src.t(k), tl[k]);
        qa += ql[k];
        ta += tl[k];
    }
    toff[npairs] = ta;
    std::memcpy(h + o_list, idx.data(), npairs * sizeof(uint32_t));
    int32_t* hm = reinterpret_cast<int32_t*>(h + o_mat);
    for (size_t k = 0; k < matrix_len; ++k)
        hm[k] = std::max(-hy::kEntryMax32, std::min(hy::kEntryMax32, matrix[k]));
    uint8_t* d = static_cast<uint8_t*>(s->arena.p);
    if (hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipEventRecord(s->ev0, s->stream) != hipSuccess)
        return HY_E_DEVICE;
    hy::BandArgs ba{};
    ba.queries = d + o_q;
    ba.qoff = reinterpret_cast<const uint32_t*>(d + o_qoff);
    ba.qlen = reinterpret_cast<const uint32_t*>(d + o_qlen);
    ba.targets = d + o_t;
    ba.toff = reinterpret_cast<const uint64_t*>(d + o_toff);
    ba.matrix = reinterpret_cast<const int32_t*>(d + o_mat);
    ba.al = (int)al;
    ba.go = go;
    ba.ge = ge;
    ba.out_score = reinterpret_cast<int32_t*>(d + o_sc);
    ba.out_qend = reinterpret_cast<int64_t*>(d + o_qe);
    ba.out_tend = reinterpret_cast<int64_t*>(d + o_te);
    if ((rc = enqueue_band(s, idx, ql, tl, ba, reinterpret_cast<const uint32_t*>(d + o_list),
                           mode, ends, s->stream)) != HY_OK)
        return rc;
    const size_t back = (ends ? used : o_qe) - o_sc;
    if (hipEventRecord(s->ev1, s->stream) != hipSuccess ||
        hipMemcpyAsync(h + o_sc, d + o_sc, back, hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return HY_E_DEVICE;
    {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s->ev0, s->ev1) == hipSuccess) {
            s->device_ms += ms;
            s->calls += 1;
        }
    }
    std::memcpy(scores, h + o_sc, npairs * sizeof(int32_t));
    for (size_t k = 0; k < npairs; ++k) {
        int64_t qe = -1, te = -1;
        if (ends) {
            std::memcpy(&qe, h + o_qe + k * sizeof(int64_t), sizeof(qe));
            std::memcpy(&te, h + o_te + k * sizeof(int64_t), sizeof(te));
        }
        if (query_ends) query_ends[k] = qe;
        if (target_ends) target_ends[k] = te;
    }
    return HY_OK;
}

int align_pairs_core(hy_pair_scratch* s, const PairSource& src, size_t npairs,
                     const int32_t* matrix, size_t matrix_len, uint32_t alphabet_len,
                     int32_t gap_open, int32_t gap_ext, uint32_t mode, uint32_t search_type,
                     int32_t* scores, int64_t* query_ends, int64_t* target_ends,
                     int64_t* err_detail, uint64_t* err_pair) {
    if (!s || (npairs && !scores)) return HY_E_NULL_ARGUMENT;
    if (err_detail) *err_detail = 0;
    if (err_pair) *err_pair = 0;
    if (mode > HY_MODE_SHW || search_type > HY_SEARCH_ALIGNMENT) return HY_E_INVALID_ARGUMENT;
    Scheme sc;
    int rc = check_scheme(matrix, matrix_len, alphabet_len, gap_open, gap_ext, &sc, err_detail);
    if (rc != HY_OK) return rc;
    // The exact per-pair checks in the reference's order (pair by pair: symbols, then width)
    // run only when a whole-batch check fails: the largest symbol, and the width proof at the
    // largest lengths (the bound grows with both lengths, so it covers every pair).
    auto exact_checks = [&]() -> int {
        for (size_t k = 0; k < npairs; ++k) {
            const int r = check_pair(src.q(k), src.m(k), src.t(k), src.n(k), alphabet_len, mode,
                                     gap_open, gap_ext, sc, err_detail);
            if (r != HY_OK) {
                if (err_pair) *err_pair = k;
                return r;
            }
        }
        return kChecksPassed;
    };
    uint64_t qtot = 0, ttot = 0;
    size_t maxm = 0, maxn = 0;
    for (size_t k = 0; k < npairs; ++k) {
        const size_t m = src.m(k), n = src.n(k);
        if ((m && !src.q(k)) || (n && !src.t(k))) return HY_E_NULL_ARGUMENT;
        qtot += m;
        ttot += n;
        maxm = std::max(maxm, m);
        maxn = std::max(maxn, n);
    }
    {
        int64_t bound = 0;
        if (hy::width_bits_for(hy::magnitude_bound(mode, sc.min_e, sc.max_e, gap_open, gap_ext,
                                                   maxm, maxn), &bound) == 0 ||
            maxm > 0x7FFFFFF0ull || maxn > 0x7FFFFFF0ull) {
            rc = exact_checks();
            if (rc != kChecksPassed) return rc;
        }
    }
    if (npairs == 0) return HY_OK;
    if (npairs > 0xFFFFFFF0ull || qtot > 0xFFFFFFF0ull) return HY_E_INVALID_ARGUMENT;
    if (hipSetDevice(s->device) != hipSuccess) return HY_E_DEVICE;
    const bool ends = tracks_end(search_type);
    const uint32_t al_eff = std::min<uint32_t>(alphabet_len, 256);

    // Split: one pair per lane for the short pairs of a large batch, one per wavefront else.
    std::vector<size_t> ql(npairs), tl(npairs);
    for (size_t k = 0; k < npairs; ++k) {
        ql[k] = src.m(k);
        tl[k] = src.n(k);
    }
    const int pick = kernel_choice();
    // Small batches that would run in bands: symbols checked here, one upload, one copy back.
    if (qtot + ttot <= kSmallBatchBytes && pick != kPickLanes &&
        (pick == kPickBand || npairs < kLaneMinPairs) &&
        use_band(pick, alphabet_len, npairs, ends)) {
        uint8_t mx = 0;
        for (size_t k = 0; k < npairs; ++k) {
            const uint8_t* q = src.q(k);
            const uint8_t* t = src.t(k);
            for (size_t i = 0; i < ql[k]; ++i) mx = std::max(mx, q[i]);
            for (size_t i = 0; i < tl[k]; ++i) mx = std::max(mx, t[i]);
        }
        if (mx >= alphabet_len) {
            rc = exact_checks();
            if (rc != kChecksPassed) return rc;
            return HY_E_INTERNAL;
        }
        return small_band_batch(s, src, npairs, ql, tl, qtot, ttot, matrix, matrix_len,
                                alphabet_len, gap_open, gap_ext, mode, ends, scores, query_ends,
                                target_ends);
    }
    // Large flat sequence buffers start going up now, on the pinned-ring workers (which also
    // find their largest symbol), while the lanes are split, sorted and grouped below; the
    // kernels are queued only after the workers have joined.
    hy::UploadJob job;
    const bool early = !src.qp && !src.tp && qtot + ttot >= hy::kPinnedRingMin;
    if (early) {
        if ((rc = grow(s->qflat, qtot)) != HY_OK || (rc = grow(s->tflat, ttot)) != HY_OK)
            return rc;
        const hy::UpSegment segs[2] = {{s->qflat.p, src.qflat + src.qoff[0], qtot, true},
                                       {s->tflat.p, src.tflat + src.toff[0], ttot, true}};
        if ((rc = job.start(s->ring, s->device, segs, 2, s->stream)) != HY_OK) return rc;
    }
    std::vector<uint32_t> lanes, waves;
    size_t nshort = 0;
    for (size_t k = 0; k < npairs; ++k)
        nshort += ql[k] <= kLaneMaxQuery && tl[k] <= kLaneMaxTarget;
    const bool use_lanes =
        al_eff <= 255 && (pick == kPickLanes || (pick == kPickAuto && nshort >= kLaneMinPairs));
    for (uint32_t k = 0; k < (uint32_t)npairs; ++k) {
        const bool shortp = ql[k] <= kLaneMaxQuery && tl[k] <= kLaneMaxTarget;
        if (use_lanes && (shortp || pick == kPickLanes)) lanes.push_back(k);
        else waves.push_back(k);
    }
    const bool band = use_band(pick, alphabet_len, waves.size(), tracks_end(search_type));
    // Local-mode lane pairs of alphabets up to kPairs16MaxAlphabet whose width proof gives 16
    // bits (queries up to kLaneMaxQuery, targets below 65,536) run two per lane (k_pairs16: one
    // pair-table lookup per cell pair up to 7 symbols, two 16-bit lookups beyond).
    // HYALITE_HIP_PAIRS16=0 keeps them one per lane.
    std::vector<uint32_t> lanes16;
    const Pairs16Choice choice16 = pairs16_choice();
    if (mode == HY_MODE_SW && alphabet_len <= hy::kPairs16MaxAlphabet && choice16 != kPairs16Off &&
        !lanes.empty()) {
        auto bits16 = [&](size_t m, size_t n) {
            int64_t b = 0;
            const int bits = hy::width_bits_for(
                hy::magnitude_bound(HY_MODE_SW, sc.min_e, sc.max_e, gap_open, gap_ext, m, n), &b);
            return bits != 0 && bits <= 16;
        };
        size_t lm = 0, ln = 0;
        for (uint32_t k : lanes) {
            lm = std::max(lm, ql[k]);
            ln = std::max(ln, tl[k]);
        }
        // The bound grows with both lengths: one proof at the longest covers every pair.
        if (lm <= kLaneMaxQuery && ln <= 65535 && bits16(lm, ln)) {
            lanes16.swap(lanes);
        } else {
            std::vector<uint32_t> rest;
            rest.reserve(lanes.size());
            lanes16.reserve(lanes.size());
            for (uint32_t k : lanes) {
                const bool fits = ql[k] <= kLaneMaxQuery && tl[k] <= 65535 && bits16(ql[k], tl[k]);
                (fits ? lanes16 : rest).push_back(k);
            }
            lanes.swap(rest);
        }
    }
    // Alphabets beyond the pair table take two lookups per cell pair: two pairs per lane pay off
    // once their groups (128 pairs) cover half the SIMDs; fewer leave SIMDs idle that one pair
    // per lane (twice the groups) would use. 16-row tiles (two wavefronts per SIMD; the 32-row
    // form needs 256 registers, one) beyond one group per SIMD, 32-row tiles below. Measured on protein pairs of 250-350 x 300-400 (SW,
    // Score, 1,024 SIMDs): 50,000 pairs 1,496 GCUPS one per lane vs 1,165 two per lane (32-row
    // tiles); 100,000 pairs 2,045 vs 2,209 (32) and 2,078 (16); 200,000 pairs 2,469 vs 2,577 (32)
    // and 2,797 (16); 400,000 pairs 2,830 vs 2,898 and 3,278.
    const bool wide16 = hy::pairs16_wide((int)alphabet_len);
    const uint64_t simds = 4ull * hy::device_cus();
    const uint64_t ngroups16 = (lanes16.size() + 127) / 128;
    if (wide16 && choice16 == kPairs16Auto && !lanes16.empty() && 2 * ngroups16 < simds) {
        std::vector<uint32_t> merged(lanes.size() + lanes16.size());
        std::merge(lanes.begin(), lanes.end(), lanes16.begin(), lanes16.end(), merged.begin());
        lanes.swap(merged);
        lanes16.clear();
    }
    // Lanes: sorted by (query length, target length) descending, cut into groups. The sort key
    // packs both lengths (complemented, 13 + 17 bits: lane pairs are at most 4096 x 65536 unless
    // the lane kernel is forced) above the pair index.
    auto sort_lanes = [&](std::vector<uint32_t>& list) {
        std::vector<uint64_t> keys(list.size());
        for (size_t i = 0; i < list.size(); ++i) {
            const uint64_t k = list[i];
            const uint64_t qk = 0x1FFFull - std::min<uint64_t>(ql[k], 0x1FFF);
            const uint64_t tk = 0x1FFFFull - std::min<uint64_t>(tl[k], 0x1FFFF);
            keys[i] = (qk << 51) | (tk << 34) | k;
        }
        // Stable LSD radix sort on the 30 length bits (two 15-bit digits); the pairs come in
        // ascending index order, so equal lengths stay in index order, as a full sort of the
        // keys would leave them.
        std::vector<uint64_t> tmp(keys.size());
        std::vector<uint32_t> cnt(1u << 15);
        for (int pass = 0; pass < 2; ++pass) {
            const int shift = 34 + 15 * pass;
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (uint64_t k : keys) ++cnt[(k >> shift) & 0x7FFF];
            uint32_t sum = 0;
            for (uint32_t& c : cnt) {
                const uint32_t v = c;
                c = sum;
                sum += v;
            }
            for (uint64_t k : keys) tmp[cnt[(k >> shift) & 0x7FFF]++] = k;
            keys.swap(tmp);
        }
        for (size_t i = 0; i < list.size(); ++i) list[i] = (uint32_t)(keys[i] & 0x3FFFFFFFFull);
    };
    sort_lanes(lanes);
    sort_lanes(lanes16);
    const size_t nslots = lanes.size() + lanes16.size();
    // The rest: ordered by the kernel's rows class (rows per pass of the band kernel, rows per
    // lane of the one-wavefront kernel) for one launch per class.
    std::stable_sort(waves.begin(), waves.end(), [&](uint32_t a, uint32_t b) {
        auto c = [&](uint32_t k) {
            if (band) return hy::band_rows_per_pass((int)ql[k]);
            return ql[k] <= 128 ? 0 : (ql[k] <= 256 ? 1 : (ql[k] <= 384 ? 2 : 3));
        };
        return c(a) < c(b);
    });
    // Rows per register tile of the lane kernel. Score-only runs take 32 (fewer boundary-row
    // trips per cell); end tracking carries three more values per cell and fits twice the
    // resident wavefronts at 16. Measured on 200,000 150 x 500 DNA pairs: 2.37k vs 2.13k GCUPS
    // (Score), 1.78k vs 1.62k (ScoreEnd). HYALITE_HIP_PAIRS_TILE=16|32 overrides it.
    int qtile = ends ? 16 : 32;
    if (const char* v = std::getenv("HYALITE_HIP_PAIRS_TILE")) qtile = std::atoi(v) >= 32 ? 32 : 16;
    // Two pairs per lane: 32-row tiles for both search types with the pair table (200,000
    // 150 x 500 DNA pairs: 3.57k vs 3.34k GCUPS Score, 2.76k vs 2.63k ScoreEnd against 16-row
    // tiles); with two lookups as measured above. HYALITE_HIP_PAIRS16_TILE=16|32 overrides it.
    int qtile16 = (wide16 && ngroups16 > simds) ? 16 : 32;
    if (const char* v = std::getenv("HYALITE_HIP_PAIRS16_TILE"))
        qtile16 = std::atoi(v) >= 32 ? 32 : 16;
    // Groups of both lane kernels in one array (k_pairs32's first), their slots in one list
    // (lanes, then lanes16), their words and boundary rows in shared buffers.
    std::vector<hy::PairGroup> groups;
    uint64_t twords = 0, qwords = 0, spill = 0;
    // Per group: target words (words_per_step per four columns), query words (rows_per_word rows
    // each, whole tiles) and, past one tile, boundary rows.
    auto add_groups = [&](const std::vector<uint32_t>& list, size_t base, uint32_t per, int qt,
                          int rows_per_word, uint32_t words_per_step) {
        for (size_t g0 = 0; g0 < list.size(); g0 += per) {
            hy::PairGroup g{};
            g.first = (uint32_t)(base + g0);
            g.count = (uint32_t)std::min<size_t>(per, list.size() - g0);
            for (uint32_t l = 0; l < g.count; ++l) {
                g.ncols = std::max<uint32_t>(g.ncols, (uint32_t)tl[list[g0 + l]]);
                g.nrows = std::max<uint32_t>(g.nrows, (uint32_t)ql[list[g0 + l]]);
            }
            const uint64_t tiles = (g.nrows + qt - 1) / qt;
            g.tword_off = twords;
            twords += (uint64_t)((g.ncols + 3) / 4) * words_per_step;
            g.qword_off = qwords;
            qwords += tiles * (uint64_t)(qt / rows_per_word) * 64;
            g.spill_off = tiles > 1 ? spill : 0;
            if (tiles > 1) spill += ((uint64_t)g.ncols + 1) * 64;
            groups.push_back(g);
        }
    };
    add_groups(lanes, 0, 64, qtile, 4, 64);
    const size_t ng32 = groups.size();
    add_groups(lanes16, lanes.size(), 128, qtile16, wide16 ? 1 : 2, wide16 ? 128 : 64);
    const size_t ng16 = groups.size() - ng32;

    Stager st{s};
    // Flat sequences go to the device straight from the caller's buffers; per-pair pointers are
    // concatenated in the stage first.
    size_t o_q = 0, o_t = 0;
    if (!early) {
        if (src.qp) o_q = st.add(&s->qflat, qtot, true);
        else st.add_external(&s->qflat, src.qflat + src.qoff[0], qtot, true);
        if (src.tp) o_t = st.add(&s->tflat, ttot, true);
        else st.add_external(&s->tflat, src.tflat + src.toff[0], ttot, true);
    }
    const size_t o_qoff = st.add(&s->qoff, (npairs + 1) * sizeof(uint64_t));
    const size_t o_toff = st.add(&s->toff, (npairs + 1) * sizeof(uint64_t));
    size_t o_q32 = 0, o_l32 = 0, o_list = 0, o_nq = 0, o_groups = 0, o_tl = 0, o_ql = 0, o_pi = 0;
    if (!waves.empty()) {
        o_q32 = st.add(&s->qoff32, npairs * sizeof(uint32_t));
        o_l32 = st.add(&s->qlen32, npairs * sizeof(uint32_t));
        o_list = st.add(&s->list, waves.size() * sizeof(uint32_t));
        o_nq = st.add(&s->nq, sizeof(uint32_t));
    }
    if (nslots) {
        o_groups = st.add(&s->groups, groups.size() * sizeof(hy::PairGroup));
        o_tl = st.add(&s->tlens, nslots * sizeof(uint32_t));
        o_ql = st.add(&s->qlens, nslots * sizeof(uint32_t));
        o_pi = st.add(&s->pidx, nslots * sizeof(uint32_t));
    }
    if ((rc = grow_stage(s, st.used + matrix_len * sizeof(int32_t) + 64)) != HY_OK) return rc;
    stage_matrix(st, s, matrix, matrix_len);
    {
        uint8_t* qd = src.qp ? st.at(o_q) : nullptr;
        uint8_t* td = src.tp ? st.at(o_t) : nullptr;
        uint64_t* qo = reinterpret_cast<uint64_t*>(st.at(o_qoff));
        uint64_t* to = reinterpret_cast<uint64_t*>(st.at(o_toff));
        uint64_t qa = 0, ta = 0;
        if (src.qp) {
            for (size_t k = 0; k < npairs; ++k) {
                qo[k] = qa;
                if (ql[k]) std::memcpy(qd + qa, src.q(k), ql[k]);
                qa += ql[k];
            }
        } else {
            for (size_t k = 0; k < npairs; ++k) qo[k] = src.qoff[k] - src.qoff[0];
            qa = qtot;
        }
        if (src.tp) {
            for (size_t k = 0; k < npairs; ++k) {
                to[k] = ta;
                if (tl[k]) std::memcpy(td + ta, src.t(k), tl[k]);
                ta += tl[k];
            }
        } else {
            for (size_t k = 0; k < npairs; ++k) to[k] = src.toff[k] - src.toff[0];
            ta = ttot;
        }
        qo[npairs] = qa;
        to[npairs] = ta;

        if (!waves.empty()) {
            uint32_t* q32 = reinterpret_cast<uint32_t*>(st.at(o_q32));
            uint32_t* l32 = reinterpret_cast<uint32_t*>(st.at(o_l32));
            for (size_t k = 0; k < npairs; ++k) {
                q32[k] = (uint32_t)qo[k];
                l32[k] = (uint32_t)ql[k];
            }
            std::memcpy(st.at(o_list), waves.data(), waves.size() * sizeof(uint32_t));
            const uint32_t np = (uint32_t)npairs;
            std::memcpy(st.at(o_nq), &np, sizeof(np));
        }
        if (nslots) {
            std::memcpy(st.at(o_groups), groups.data(), groups.size() * sizeof(hy::PairGroup));
            uint32_t* tls = reinterpret_cast<uint32_t*>(st.at(o_tl));
            uint32_t* qls = reinterpret_cast<uint32_t*>(st.at(o_ql));
            uint32_t* pis = reinterpret_cast<uint32_t*>(st.at(o_pi));
            size_t i = 0;
            for (const std::vector<uint32_t>* list : {&lanes, &lanes16})
                for (uint32_t k : *list) {
                    tls[i] = (uint32_t)tl[k];
                    qls[i] = (uint32_t)ql[k];
                    pis[i] = k;
                    ++i;
                }
        }
    }
    if ((rc = st.commit()) != HY_OK) return rc;
    if (early) {
        uint8_t mx = 0;
        if ((rc = job.finish(&mx)) != HY_OK) return rc;
        st.max_symbol = std::max(st.max_symbol, mx);
    }
    // Symbols: the largest byte of both sides, found on the host while the sequences were copied
    // for upload and checked before any kernel indexes the matrix with them (their copies are
    // queued, nothing waits for them here); the exact per-pair order of the reference's errors
    // is recovered only when one is out of range.
    if (st.max_symbol >= alphabet_len) {
        if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
        rc = exact_checks();
        if (rc != kChecksPassed) return rc;
        return HY_E_INTERNAL;
    }
    if ((rc = grow(s->scores, npairs * sizeof(int32_t))) != HY_OK ||
        (rc = grow(s->qend, npairs * sizeof(int64_t))) != HY_OK ||
        (rc = grow(s->tend, npairs * sizeof(int64_t))) != HY_OK)
        return rc;

    if (hipEventRecord(s->ev0, s->stream) != hipSuccess) return HY_E_DEVICE;
    // Long pairs and small batches: the aux stream, forked after the uploads.
    const bool both = !waves.empty() && nslots > 0;
    hipStream_t wstream = both ? s->aux : s->stream;
    if (both && (hipEventRecord(s->fork, s->stream) != hipSuccess ||
                 hipStreamWaitEvent(s->aux, s->fork, 0) != hipSuccess))
        return HY_E_DEVICE;
    if (!waves.empty()) {
        const uint32_t* d_list = static_cast<const uint32_t*>(s->list.p);
        if (band) {
            hy::BandArgs ba{};
            ba.queries = static_cast<const uint8_t*>(s->qflat.p);
            ba.qoff = static_cast<const uint32_t*>(s->qoff32.p);
            ba.qlen = static_cast<const uint32_t*>(s->qlen32.p);
            ba.targets = static_cast<const uint8_t*>(s->tflat.p);
            ba.toff = static_cast<const uint64_t*>(s->toff.p);
            ba.matrix = static_cast<const int32_t*>(s->mat.p);
            ba.al = (int)alphabet_len;
            ba.go = gap_open;
            ba.ge = gap_ext;
            ba.out_score = static_cast<int32_t*>(s->scores.p);
            ba.out_qend = static_cast<int64_t*>(s->qend.p);
            ba.out_tend = static_cast<int64_t*>(s->tend.p);
            rc = enqueue_band(s, waves, ql, tl, ba, d_list, mode, ends, wstream);
        } else {
            rc = enqueue_waves(s, waves, ql, tl, d_list, alphabet_len, gap_open, gap_ext, mode,
                               ends, wstream);
        }
        if (rc != HY_OK) return rc;
    }
    if (nslots) {
        if ((rc = grow(s->twords, twords * 4)) != HY_OK ||
            (rc = grow(s->qwords, qwords * 4)) != HY_OK ||
            (rc = grow(s->spill, std::max<uint64_t>(spill, 1) * sizeof(int2))) != HY_OK)
            return rc;
        hy::PairPackArgs pk{};
        pk.groups = static_cast<const hy::PairGroup*>(s->groups.p);
        pk.qflat = static_cast<const uint8_t*>(s->qflat.p);
        pk.qoff = static_cast<const uint64_t*>(s->qoff.p);
        pk.tflat = static_cast<const uint8_t*>(s->tflat.p);
        pk.toff = static_cast<const uint64_t*>(s->toff.p);
        pk.pidx = static_cast<const uint32_t*>(s->pidx.p);
        pk.qlens = static_cast<const uint32_t*>(s->qlens.p);
        pk.tlens = static_cast<const uint32_t*>(s->tlens.p);
        pk.qwords = static_cast<uint32_t*>(s->qwords.p);
        pk.twords = static_cast<uint32_t*>(s->twords.p);
        pk.pad_symbol = al_eff;
        pk.qtile = qtile;
        if (hy::launch_pair_pack(pk, (uint32_t)ng32, s->stream) != hipSuccess)
            return HY_E_DEVICE;
        hy::PairArgs pa{};
        pa.groups = pk.groups;
        pa.twords = pk.twords;
        pa.qwords = pk.qwords;
        pa.tlens = pk.tlens;
        pa.qlens = pk.qlens;
        pa.pidx = pk.pidx;
        pa.matrix = static_cast<const int32_t*>(s->mat.p);
        pa.al = (int)alphabet_len;
        pa.al_eff = (int)al_eff;
        pa.go = gap_open;
        pa.ge = gap_ext;
        pa.spill = static_cast<int2*>(s->spill.p);
        pa.out_score = static_cast<int32_t*>(s->scores.p);
        pa.out_qend = static_cast<int64_t*>(s->qend.p);
        pa.out_tend = static_cast<int64_t*>(s->tend.p);
        if (hy::launch_pairs32((int)mode, ends, pa, (uint32_t)ng32, qtile, s->stream) !=
            hipSuccess)
            return HY_E_DEVICE;
        if (ng16) {
            hy::PairPackArgs pk16 = pk;
            pk16.groups = pk.groups + ng32;
            pk16.pad_symbol = alphabet_len;
            pk16.qtile = qtile16;
            hy::Pair16Args p16{};
            p16.groups = pk16.groups;
            p16.twords = pk.twords;
            p16.qwords = pk.qwords;
            p16.tlens = pk.tlens;
            p16.qlens = pk.qlens;
            p16.pidx = pk.pidx;
            p16.matrix = pa.matrix;
            p16.al = (int)alphabet_len;
            p16.go = gap_open;
            p16.ge = gap_ext;
            p16.spill = static_cast<uint2*>(s->spill.p);
            p16.out_score = pa.out_score;
            p16.out_qend = pa.out_qend;
            p16.out_tend = pa.out_tend;
            if (hy::launch_pair_pack16(pk16, (uint32_t)ng16, s->stream) != hipSuccess ||
                hy::launch_pairs16_sw(ends, p16, (uint32_t)ng16, qtile16, s->stream) !=
                    hipSuccess)
                return HY_E_DEVICE;
        }
    }
    if (both && (hipEventRecord(s->join, s->aux) != hipSuccess ||
                 hipStreamWaitEvent(s->stream, s->join, 0) != hipSuccess))
        return HY_E_DEVICE;
    if (hipEventRecord(s->ev1, s->stream) != hipSuccess) return HY_E_DEVICE;
    if (hipMemcpyAsync(scores, s->scores.p, npairs * sizeof(int32_t), hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess)
        return HY_E_DEVICE;
    if (ends) {
        if (query_ends && hipMemcpyAsync(query_ends, s->qend.p, npairs * sizeof(int64_t),
                                         hipMemcpyDeviceToHost, s->stream) != hipSuccess)
            return HY_E_DEVICE;
        if (target_ends && hipMemcpyAsync(target_ends, s->tend.p, npairs * sizeof(int64_t),
                                          hipMemcpyDeviceToHost, s->stream) != hipSuccess)
            return HY_E_DEVICE;
    }
    if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
    {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s->ev0, s->ev1) == hipSuccess) {
            s->device_ms += ms;
            s->calls += 1;
        }
    }
    if (!ends) {
        for (size_t k = 0; k < npairs; ++k) {
            if (query_ends) query_ends[k] = -1;
            if (target_ends) target_ends[k] = -1;
        }
    }
    return HY_OK;
}

}  // namespace

extern "C" int hy_align_pairs(hy_pair_scratch* s, const uint8_t* const* queries,
                              const size_t* qlens, const uint8_t* const* targets,
                              const size_t* tlens, size_t npairs, const int32_t* matrix,
                              size_t matrix_len, uint32_t alphabet_len, int32_t gap_open,
                              int32_t gap_ext, uint32_t mode, uint32_t search_type,
                              int32_t* scores, int64_t* query_ends, int64_t* target_ends,
                              int64_t* err_detail, uint64_t* err_pair) {
    if (npairs && (!queries || !qlens || !targets || !tlens)) return HY_E_NULL_ARGUMENT;
    PairSource src;
    src.qp = queries;
    src.qlen = qlens;
    src.tp = targets;
    src.tlen = tlens;
    return align_pairs_core(s, src, npairs, matrix, matrix_len, alphabet_len, gap_open, gap_ext,
                            mode, search_type, scores, query_ends, target_ends, err_detail,
                            err_pair);
}

extern "C" int hy_align_pairs_flat(hy_pair_scratch* s, const uint8_t* queries,
                                   const uint64_t* query_offsets, const uint8_t* targets,
                                   const uint64_t* target_offsets, size_t npairs,
                                   const int32_t* matrix, size_t matrix_len,
                                   uint32_t alphabet_len, int32_t gap_open, int32_t gap_ext,
                                   uint32_t mode, uint32_t search_type, int32_t* scores,
                                   int64_t* query_ends, int64_t* target_ends,
                                   int64_t* err_detail, uint64_t* err_pair) {
    if (npairs && (!query_offsets || !target_offsets)) return HY_E_NULL_ARGUMENT;
    for (size_t k = 0; k < npairs; ++k)
        if (query_offsets[k + 1] < query_offsets[k] || target_offsets[k + 1] < target_offsets[k])
            return HY_E_INVALID_ARGUMENT;
    PairSource src;
    src.qflat = queries;
    src.qoff = query_offsets;
    src.tflat = targets;
    src.toff = target_offsets;
    return align_pairs_core(s, src, npairs, matrix, matrix_len, alphabet_len, gap_open, gap_ext,
                            mode, search_type, scores, query_ends, target_ends, err_detail,
                            err_pair);
}

extern "C" int hy_pair_scratch_profile_read(hy_pair_scratch* s, double* device_ms,
                                            uint64_t* calls, int reset) {
    if (!s || !device_ms || !calls) return HY_E_NULL_ARGUMENT;
    *device_ms = s->device_ms;
    *calls = s->calls;
    if (reset) {
        s->device_ms = 0.0;
        s->calls = 0;
    }
    return HY_OK;
}

extern "C" int hy_align_pair_position_max(hy_pair_scratch* s, const uint8_t* query, size_t qlen,
                                          const uint8_t* target, size_t tlen,
                                          const int32_t* matrix, size_t matrix_len,
                                          uint32_t alphabet_len, int32_t gap_open,
                                          int32_t gap_ext, int32_t* colmax, int32_t* score,
                                          int64_t* target_end, int64_t* err_detail) {
    if (!s || !score || !target_end || (tlen && !colmax)) return HY_E_NULL_ARGUMENT;
    if (err_detail) *err_detail = 0;
    Scheme sc;
    int rc = check_scheme(matrix, matrix_len, alphabet_len, gap_open, gap_ext, &sc, err_detail);
    if (rc != HY_OK) return rc;
    rc = check_pair(query, qlen, target, tlen, alphabet_len, HY_MODE_SW, gap_open, gap_ext, sc,
                    err_detail);
    if (rc != HY_OK) return rc;
    if (hipSetDevice(s->device) != hipSuccess) return HY_E_DEVICE;
    if (band_eligible(alphabet_len) && kernel_choice() != kPickWaves) {
        Single L;
        if ((rc = band_single(s, query, qlen, target, tlen, matrix, matrix_len, alphabet_len,
                              gap_open, gap_ext, HY_MODE_SW, hy::kBandColmax,
                              8 + tlen * sizeof(int32_t), &L)) != HY_OK)
            return rc;
        const uint8_t* o = s->stage + L.out;
        int2 b;
        std::memcpy(&b, o, sizeof(b));
        if (tlen) std::memcpy(colmax, o + 8, tlen * sizeof(int32_t));
        *score = b.x;
        *target_end = b.y;
        return HY_OK;
    }
    Stager st{s};
    const size_t o_q = st.add(&s->qflat, qlen);
    const size_t o_t = st.add(&s->tflat, tlen);
    if ((rc = grow_stage(s, st.used + matrix_len * sizeof(int32_t) + 64)) != HY_OK) return rc;
    stage_matrix(st, s, matrix, matrix_len);
    if (qlen) std::memcpy(st.at(o_q), query, qlen);
    if (tlen) std::memcpy(st.at(o_t), target, tlen);
    if ((rc = st.commit()) != HY_OK) return rc;
    if ((rc = grow(s->colmax, tlen * sizeof(int32_t))) != HY_OK ||
        (rc = grow(s->best, sizeof(int2))) != HY_OK ||
        (rc = grow(s->sys_spill, qlen > 512 ? (tlen + 1) * sizeof(int2) : 16)) != HY_OK)
        return rc;
    hy::PosMaxArgs pa{};
    pa.query = static_cast<const uint8_t*>(s->qflat.p);
    pa.m = (int)qlen;
    pa.target = static_cast<const uint8_t*>(s->tflat.p);
    pa.n = (int)tlen;
    pa.matrix = static_cast<const int32_t*>(s->mat.p);
    pa.al = (int)alphabet_len;
    pa.al_eff = (int)std::min<uint32_t>(alphabet_len, 256);
    pa.go = gap_open;
    pa.ge = gap_ext;
    pa.spill = static_cast<int2*>(s->sys_spill.p);
    pa.colmax = static_cast<int32_t*>(s->colmax.p);
    pa.best = static_cast<int2*>(s->best.p);
    if (hy::launch_posmax(pa, s->stream) != hipSuccess) return HY_E_DEVICE;
    int2* h_best = reinterpret_cast<int2*>(s->stage);
    if ((tlen && hipMemcpyAsync(colmax, s->colmax.p, tlen * sizeof(int32_t),
                                hipMemcpyDeviceToHost, s->stream) != hipSuccess) ||
        hipMemcpyAsync(h_best, s->best.p, sizeof(int2), hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess)
        return HY_E_DEVICE;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
    *score = h_best->x;
    *target_end = h_best->y;
    return HY_OK;
}

// score2 (src/kernel.rs:741-792): w = ceil(best / matrix_max); columns at or above min_score
// that continue the open peak's maximum position extend it (keeping the higher score), any other
// such column closes it and opens a new one; a closed peak outside [te - w, te + w] becomes the
// runner-up only when strictly higher, so ties keep the smallest position.
extern "C" int hy_score2(const int32_t* colmax, size_t n, int32_t best_score,
                         uint64_t best_target_end, int32_t matrix_max, int32_t min_score,
                         int32_t* found, int32_t* score2, uint64_t* position) {
    if (!found || !score2 || !position || (n && !colmax)) return HY_E_NULL_ARGUMENT;
    *found = 0;
    if (matrix_max <= 0) return HY_OK;
    const int64_t w = std::max<int64_t>(
        ((int64_t)std::max(best_score, 0) + (int64_t)matrix_max - 1) / (int64_t)matrix_max, 0);
    const int64_t low = (int64_t)best_target_end - w, high = (int64_t)best_target_end + w;
    bool have = false, open = false;
    int32_t rs = 0, ps = 0;
    uint64_t rp = 0, pp = 0;
    auto close = [&]() {
        const int64_t p = (int64_t)pp;
        if ((p < low || p > high) && (!have || ps > rs)) {
            have = true;
            rs = ps;
            rp = pp;
        }
    };
    for (size_t i = 0; i < n; ++i) {
        const int32_t v = colmax[i];
        if (v < min_score) continue;
        if (open && pp + 1 == i) {
            if (ps < v) { ps = v; pp = i; }
        } else {
            if (open) close();
            open = true;
            ps = v;
            pp = i;
        }
    }
    if (open) close();
    if (have) {
        *found = 1;
        *score2 = rs;
        *position = rp;
    }
    return HY_OK;
}
