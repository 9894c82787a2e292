// Host side of the MI355X database-scan backend: the C ABI of include/hyalite_hip.h.
//
// Build (hy_db_build) mirrors DatabaseBuilder::build (src/database.rs:737-877): the same
// validation order and error codes, the per-sequence width proof (src/database.rs:884-948,
// src/width.rs:107-169), then a one-time upload of the database into HBM in the layouts the
// kernels read. Scans mirror Database::scan / scan_all / scan_scores (src/database.rs:83-515);
// they allocate nothing, run on the Scratch's own stream, and use no CPU fallback: every score
// and end position comes from a kernel in hy_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/hyalite_hip.h"
#include "hy_internal.h"
#include "hy_state.h"

using hy::Group;
using hy::Result;
using hy::dev_free;

namespace {

constexpr uint32_t kNoPadSymbol = 255;  // al_eff <= 255 leaves symbol value al_eff free for padding

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

template <typename T>
hipError_t dev_upload(T** dst, const std::vector<T>& src) {
    *dst = nullptr;
    if (src.empty()) return hipSuccess;
    hipError_t e = hipMalloc((void**)dst, src.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}


}  // namespace

namespace hy {

// ScoreWidth::max_abs (src/width.rs:48-55) and required_width (src/width.rs:145-169).
int width_bits_for(__int128 bound, int64_t* bound_out) {
    const int64_t b = bound > (__int128)INT64_MAX ? INT64_MAX : (int64_t)bound;
    if (bound_out) *bound_out = b;
    if (b <= 127) return 8;
    if (b <= 32767) return 16;
    if (b <= 536870911) return 32;
    return 0;
}

// magnitude_bound (src/width.rs:107-137), evaluated in 128-bit.
__int128 magnitude_bound(uint32_t mode, int32_t min_e, int32_t max_e, int32_t go, int32_t ge,
                         uint64_t mql, uint64_t mtl) {
    const __int128 m = (__int128)mql, n = (__int128)mtl;
    const __int128 min_ij = m < n ? m : n;
    const __int128 max_pos = max_e > 0 ? (__int128)max_e : 0;
    const __int128 max_neg = min_e < 0 ? -(__int128)min_e : 0;
    const __int128 positive = min_ij * max_pos;
    __int128 negative;
    if (mode == HY_MODE_SW) {
        negative = go;
    } else if (mode == HY_MODE_OV) {
        negative = min_ij * max_neg + go;
    } else {
        __int128 run = m + n - 1;
        if (run < 0) run = 0;
        negative = (m + n) * max_neg + go + run * ge;
    }
    return positive > negative ? positive : negative;
}

// Scoring::new (src/scoring.rs:65-100): alphabet, matrix shape, gap signs and order, in that
// order; on success the matrix's entry bounds.
int validate_scoring(const int32_t* matrix, size_t matrix_len, uint32_t alphabet_len,
                     int32_t gap_open, int32_t gap_ext, int32_t* min_e, int32_t* max_e,
                     int64_t* err_detail) {
    if (alphabet_len == 0) return HY_E_EMPTY_ALPHABET;
    if ((uint64_t)alphabet_len * (uint64_t)alphabet_len != (uint64_t)matrix_len) {
        if (err_detail) *err_detail = (int64_t)matrix_len;
        return HY_E_MATRIX_SHAPE;
    }
    if (!matrix) return HY_E_NULL_ARGUMENT;
    if (gap_open < 0 || gap_ext < 0) return HY_E_NEGATIVE_GAP_PENALTY;
    if (gap_open < gap_ext) return HY_E_GAP_OPEN_LESS_THAN_EXTEND;
    *min_e = *max_e = matrix[0];
    for (size_t k = 1; k < matrix_len; ++k) {
        *min_e = std::min(*min_e, matrix[k]);
        *max_e = std::max(*max_e, matrix[k]);
    }
    return HY_OK;
}

}  // namespace hy

namespace {
using hy::magnitude_bound;
using hy::width_bits_for;

// HYALITE_HIP_CHAIN=1: the chained kernel whenever a valid shape exists (tests).
bool chain_forced() {
    const char* e = std::getenv("HYALITE_HIP_CHAIN");
    return e && std::atoi(e) == 1;
}

// The traceback budget of a database (src/align.rs:228-268): the smallest max_bytes with which
// the reference's traceback serves every (m, n) in [0, max_m] x [0, max_n]. A full traceback
// holds three i32 matrices of (m+1)(n+1) cells; the checkpoint path, with strip height
// k = isqrt(m), holds H and F at every k-th row plus one (k+1)-row strip of H, E and F. The
// cheaper of the two at (max_m, max_n), raised to the single-row/column full footprints of an
// empty query or target, which take the full path unconditionally. Saturating u64 arithmetic.
uint64_t sat_mul(uint64_t a, uint64_t b) {
    return (a != 0 && b > UINT64_MAX / a) ? UINT64_MAX : a * b;
}

uint64_t full_traceback_bytes(uint64_t m, uint64_t n) {
    return sat_mul(sat_mul(sat_mul(m + 1, n + 1), 3), 4);
}

uint64_t isqrt64(uint64_t v) {
    uint64_t r = (uint64_t)std::sqrt((double)v);
    while (r > 0 && r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

uint64_t traceback_budget_for_database(uint64_t max_m, uint64_t max_n) {
    uint64_t need = full_traceback_bytes(max_m, max_n);
    if (max_m > 0 && max_n > 0) {
        const uint64_t k = std::max<uint64_t>(isqrt64(max_m), 1);
        const uint64_t ckpt = sat_mul(sat_mul(2 * (max_m / k + 1) + 3 * (k + 1), max_n + 1), 4);
        need = std::min(need, ckpt);
    }
    return std::max({need, full_traceback_bytes(max_m, 0), full_traceback_bytes(0, max_n)});
}

std::string lower_trim(const char* s) {
    std::string v(s ? s : "");
    size_t a = 0, b = v.size();
    while (a < b && std::isspace((unsigned char)v[a])) ++a;
    while (b > a && std::isspace((unsigned char)v[b - 1])) --b;
    v = v.substr(a, b - a);
    for (auto& c : v) c = (char)std::tolower((unsigned char)c);
    return v;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Database and Scratch
// ---------------------------------------------------------------------------------------------

// ---------------------------------------------------------------------------------------------
// Backend resolution (src/backend.rs)
// ---------------------------------------------------------------------------------------------

extern "C" int hy_backend_available(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 0;
    for (int d = 0; d < count; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) != hipSuccess) continue;
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) return 1;
    }
    return 0;
}

extern "C" int hy_backend_parse(const char* name, int32_t* out) {
    if (!out) return HY_E_NULL_ARGUMENT;
    const std::string v = lower_trim(name);
    if (v == "auto") { *out = HY_BACKEND_AUTO; return HY_OK; }
    if (v == "hip") { *out = HY_BACKEND_HIP; return HY_OK; }
    if (v == "scalar" || v == "sse4.1" || v == "sse41" || v == "avx2" || v == "neon") {
        *out = -1;  // a reference CPU backend: valid name, not served by this library
        return HY_OK;
    }
    return HY_E_INVALID_BACKEND_NAME;
}

int hy::first_gfx950_device() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return -1;
    for (int d = 0; d < count; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) != hipSuccess) continue;
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) return d;
    }
    return -1;
}

// ---------------------------------------------------------------------------------------------
// Build
// ---------------------------------------------------------------------------------------------

namespace hy {
std::recursive_mutex& api_lock() {
    static std::recursive_mutex m;
    return m;
}
}  // namespace hy

extern "C" int hy_db_build(const uint8_t* const* seqs, const size_t* lens, size_t n,
                           uint32_t have, const int32_t* matrix, size_t matrix_len,
                           uint32_t alphabet_len, int32_t gap_open, int32_t gap_ext, uint32_t mode,
                           uint32_t search_type, uint64_t traceback_max_bytes,
                           size_t max_query_len, uint32_t backend, hy_db** out,
                           int64_t* err_detail) {
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    if (!out) return HY_E_NULL_ARGUMENT;
    *out = nullptr;
    if (err_detail) *err_detail = 0;

    // Scoring::new (src/scoring.rs:65-100) runs before the builder sees the scheme.
    int32_t min_e = 0, max_e = 0;
    if (have & HY_HAVE_SCORING) {
        const int rc = hy::validate_scoring(matrix, matrix_len, alphabet_len, gap_open, gap_ext,
                                            &min_e, &max_e, err_detail);
        if (rc != HY_OK) return rc;
    }
    // IncompleteBuilder in field order (src/database.rs:738-749).
    const uint32_t order[4] = {HY_HAVE_SEQUENCES, HY_HAVE_SCORING, HY_HAVE_MODE,
                               HY_HAVE_MAX_QUERY_LEN};
    for (int k = 0; k < 4; ++k) {
        if (!(have & order[k])) {
            if (err_detail) *err_detail = k;
            return HY_E_INCOMPLETE_BUILDER;
        }
    }
    if (mode > HY_MODE_SHW || search_type > HY_SEARCH_ALIGNMENT) return HY_E_INVALID_ARGUMENT;
    if (n == 0) return HY_E_EMPTY_DATABASE;
    if (!seqs || !lens) return HY_E_NULL_ARGUMENT;
    // SymbolOutOfRange: the first bad symbol in database order (src/database.rs:756-766).
    // Each sequence's largest byte first (16 bytes per step); the exact first bad symbol is
    // looked for only in a sequence that has one.
    uint64_t mtl = 0, total = 0;
    for (size_t k = 0; k < n; ++k) {
        if (lens[k] && !seqs[k]) return HY_E_NULL_ARGUMENT;
        if (lens[k] && (uint32_t)hy::host_max_byte(seqs[k], lens[k]) >= alphabet_len) {
            for (size_t x = 0; x < lens[k]; ++x) {
                if ((uint32_t)seqs[k][x] >= alphabet_len) {
                    if (err_detail) *err_detail = seqs[k][x];
                    return HY_E_SYMBOL_OUT_OF_RANGE;
                }
            }
        }
        mtl = std::max<uint64_t>(mtl, lens[k]);
        total += lens[k];
    }
    if (mtl > 0x7FFFFFF0ull || max_query_len > 0x7FFFFFF0ull || n > 0xFFFFFFF0ull)
        return HY_E_INVALID_ARGUMENT;
    // Width proof at the declared maxima (src/database.rs:771).
    int64_t bound = 0;
    const int widest = width_bits_for(
        magnitude_bound(mode, min_e, max_e, gap_open, gap_ext, max_query_len, mtl), &bound);
    if (widest == 0) {
        if (err_detail) *err_detail = bound;
        return HY_E_SCORE_RANGE_TOO_WIDE;
    }
    // An Alignment search proves its traceback budget once for the whole problem box
    // (database.rs:773-787), so the aligned scans never fail on it later.
    if (search_type == HY_SEARCH_ALIGNMENT) {
        const uint64_t needed = traceback_budget_for_database(max_query_len, mtl);
        if (needed > traceback_max_bytes) {
            if (err_detail) std::memcpy(err_detail, &needed, sizeof(needed));
            return HY_E_TRACEBACK_BUDGET;
        }
    }
    // Backend: explicit choice, else HYALITE_BACKEND, else Auto (src/database.rs:789-794).
    if (backend != HY_BACKEND_AUTO && backend != HY_BACKEND_HIP) return HY_E_INVALID_ARGUMENT;
    if (backend == HY_BACKEND_AUTO) {
        const char* env = std::getenv("HYALITE_BACKEND");
        if (env && !lower_trim(env).empty()) {
            int32_t b = 0;
            if (hy_backend_parse(env, &b) != HY_OK) return HY_E_INVALID_BACKEND_NAME;
            if (b < 0) return HY_E_BACKEND_UNAVAILABLE;  // a CPU backend was requested
        }
    }
    const int device = hy::first_gfx950_device();
    if (device < 0) return HY_E_BACKEND_UNAVAILABLE;
    if (hipSetDevice(device) != hipSuccess) return HY_E_DEVICE;

    hy_db* db = new (std::nothrow) hy_db();
    if (!db) return HY_E_OUT_OF_MEMORY;
    db->device = device;
    db->al = alphabet_len;
    db->al_eff = std::min<uint32_t>(alphabet_len, 256);
    db->go = gap_open;
    db->ge = gap_ext;
    db->min_e = min_e;
    db->max_e = max_e;
    db->mode = mode;
    db->search_type = search_type;
    db->traceback_max_bytes = search_type == HY_SEARCH_ALIGNMENT ? traceback_max_bytes : 0;
    db->n = n;
    db->mql = max_query_len;
    db->mtl = mtl;
    db->total_residues = total;
    db->widest_bits = widest;

    // Per-sequence widths (src/database.rs:896-903). A target whose proof gives i8 or i16 AND
    // leaves room for one gap opening below the bound (bound + gap_open <= 32767, so H - go never
    // saturates) runs in the packed 16-bit kernel; every other target runs at 32 bits. Any width
    // the proof accepts is exact (DETERMINISM.md §4, §6), so the split changes speed only.
    // Tile shape of the packed kernel: 32 query rows per register tile, one wavefront per
    // workgroup with a per-symbol profile; HYALITE_HIP_TILE="rows,waves" overrides it (waves 4 or
    // 8 select workgroups sharing a residue-pair profile). A Scratch can change it (hy_tuning).
    db->q16 = 32;
    db->waves16 = 1;
    if (const char* t = std::getenv("HYALITE_HIP_TILE")) {
        int q = 0, w = 0;
        if (std::sscanf(t, "%d,%d", &q, &w) == 2 && q > 0 && w > 0) {
            db->q16 = hy::inter16_tile_rows(q);
            db->waves16 = w >= 8 ? 8 : (w >= 4 ? 4 : 1);
        }
    }
    if (db->waves16 > 1)
        while (db->q16 > 16 &&
               hy::inter16_lds_bytes((int)db->al_eff, db->q16) > hy::kPairProfileMaxBytes)
            db->q16 -= 8;
    const bool pair_profile_fits =
        db->al_eff <= kNoPadSymbol &&
        (db->waves16 == 1 ||
         hy::inter16_lds_bytes((int)db->al_eff, db->q16) <= hy::kPairProfileMaxBytes);
    // The bound grows with the target length, so each width (and the packed kernel's gap-open
    // headroom) holds up to a length threshold, found once by bisection.
    auto bound_at = [&](uint64_t len) -> int64_t {
        int64_t b = 0;
        (void)width_bits_for(
            magnitude_bound(mode, min_e, max_e, gap_open, gap_ext, max_query_len, len), &b);
        return b;
    };
    auto longest_within = [&](int64_t limit) -> int64_t {  // -1: not even length 0
        if (bound_at(0) > limit) return -1;
        uint64_t lo = 0, hi = mtl;
        if (bound_at(hi) <= limit) return (int64_t)hi;
        while (hi - lo > 1) {
            const uint64_t mid = lo + (hi - lo) / 2;
            (bound_at(mid) <= limit ? lo : hi) = mid;
        }
        return (int64_t)lo;
    };
    const int64_t len8 = longest_within(127), len16 = longest_within(32767);
    const int64_t len_packed = longest_within(32767 - (int64_t)gap_open);
    std::vector<uint8_t> elig16(n, 0);
    for (size_t k = 0; k < n; ++k) {
        const int64_t len = (int64_t)lens[k];
        db->count_by_width[len <= len8 ? 0 : (len <= len16 ? 1 : 2)]++;
        elig16[k] = pair_profile_fits && len <= len_packed;
    }

    int rc = HY_OK;
    auto fail = [&](int code) { delete db; return code; };

    // Matrix, clamped for the 32-bit kernels; the scheme as given stays on the host.
    db->h_matrix.assign(matrix, matrix + matrix_len);
    std::vector<int32_t> mat(matrix, matrix + matrix_len);
    for (auto& v : mat) v = std::max(-hy::kEntryMax32, std::min(hy::kEntryMax32, v));
    if (dev_upload(&db->d_matrix, mat) != hipSuccess) return fail(HY_E_DEVICE);

    // Database residues in original order.
    {
        std::vector<uint64_t> off(n + 1, 0);
        for (size_t k = 0; k < n; ++k) off[k + 1] = off[k] + lens[k];
        // Sequences that already lie back to back in one buffer (a database image, a flat
        // batch from the caller) upload from there; others are gathered into one first.
        bool contiguous = total > 0;
        for (size_t k = 0; contiguous && k + 1 < n; ++k)
            contiguous = reinterpret_cast<uintptr_t>(seqs[k]) + lens[k] ==
                         reinterpret_cast<uintptr_t>(seqs[k + 1]);
        std::unique_ptr<uint8_t[]> flat;
        const uint8_t* host = contiguous ? seqs[0] : nullptr;
        if (!contiguous) {
            // Every byte is written below, so the buffer is left uninitialised.
            flat.reset(new (std::nothrow) uint8_t[std::max<uint64_t>(total, 1)]);
            if (!flat) return fail(HY_E_OUT_OF_MEMORY);
            for (size_t k = 0; k < n; ++k)
                if (lens[k]) std::memcpy(flat.get() + off[k], seqs[k], lens[k]);
            host = flat.get();
        }
        if (dev_upload(&db->d_off, off) != hipSuccess) return fail(HY_E_DEVICE);
        if (hipMalloc(reinterpret_cast<void**>(&db->d_flat), std::max<uint64_t>(total, 1)) !=
                hipSuccess ||
            hipMemcpy(db->d_flat, host, std::max<uint64_t>(total, 1), hipMemcpyHostToDevice) !=
                hipSuccess)
            return fail(HY_E_DEVICE);
        db->h_off = std::move(off);
    }

    // Split: long targets go to the systolic kernel, the rest to the inter-sequence kernel. A
    // database whose alphabet uses every byte value has no padding symbol and runs systolic.
    // Longest first, ties in database order: a stable LSD radix sort on the complemented
    // length (three 11-bit digits cover lengths below 2^33).
    std::vector<uint32_t> order_idx(n);
    {
        std::vector<uint32_t> tmp(n);
        std::iota(order_idx.begin(), order_idx.end(), 0u);
        std::vector<uint32_t> cnt(1u << 11);
        const uint64_t top = mtl;
        for (int pass = 0; pass < 3; ++pass) {
            const int shift = 11 * pass;
            if ((top >> shift) == 0 && pass > 0) break;
            auto digit = [&](uint32_t k) { return (uint32_t)(((top - lens[k]) >> shift) & 0x7FF); };
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (uint32_t k : order_idx) ++cnt[digit(k)];
            uint32_t sum = 0;
            for (uint32_t& c : cnt) {
                const uint32_t v = c;
                c = sum;
                sum += v;
            }
            for (uint32_t k : order_idx) tmp[cnt[digit(k)]++] = k;
            order_idx.swap(tmp);
        }
    }
    // A one-query scan lasts as long as its slowest wavefront: a packed group's time grows with
    // its longest target, so long targets go to the systolic kernel on a second stream instead.
    // Long = beyond 1.6x the average column share of a resident wavefront (residues / 128 per
    // group / resident workgroups) or beyond 8x the mean length, whichever is shorter, and at
    // least 1,024. Measured: 1M protein targets (share 1.5k columns) 6.6k GCUPS at 2.4k vs
    // 5.3k at 3.2k; 2M targets (share 3.0k) 71 ms per query at 3.2k vs 85 ms at 4.9k.
    // HYALITE_HIP_LONG_THRESHOLD overrides it.
    const uint32_t resident =
        hy::inter16_batch_occupancy((int)mode, db->q16, (int)db->al_eff) * hy::device_cus();
    const double share = (double)total / 128.0 / (double)std::max<uint32_t>(resident, 1);
    const double mean = (double)total / (double)n;
    uint64_t long_threshold =
        std::max<uint64_t>(1024, (uint64_t)std::min(1.6 * share, 8.0 * mean));
    if (const char* t = std::getenv("HYALITE_HIP_LONG_THRESHOLD")) {
        const long long v = std::atoll(t);
        if (v > 0) long_threshold = (uint64_t)v;
    }
    const bool all_systolic = db->al_eff > kNoPadSymbol;
    // 32-bit inter layout order: [targets not eligible for the packed kernel][eligible ones],
    // each part length-sorted and starting a fresh group, so score-only scans can run the 32-bit
    // kernel over the first part alone; end-tracking scans run it over both.
    // The packed layout holds every eligible target, long ones first in groups of their own:
    // one-query scans skip those groups (the systolic kernel takes the long targets), while the
    // batched form runs them in the persistent queue, where many queries keep them balanced.
    std::vector<uint32_t> inter_idx, idx16;
    inter_idx.reserve(n);
    for (uint32_t k : order_idx) {
        const bool is_long = all_systolic || lens[k] > long_threshold;
        if (is_long) {
            db->sys_list.push_back(k);
            db->sys_max_len = std::max<uint64_t>(db->sys_max_len, lens[k]);
            if (elig16[k] && !all_systolic) {
                idx16.push_back(k);
            } else {
                db->sys_list_batch.push_back(k);
                db->sys_batch_max_len = std::max<uint64_t>(db->sys_batch_max_len, lens[k]);
            }
        } else if (!elig16[k]) {
            inter_idx.push_back(k);
            // The batched form has no 32-bit inter-sequence queue: these targets join its
            // systolic list.
            db->sys_list_batch.push_back(k);
            db->sys_batch_max_len = std::max<uint64_t>(db->sys_batch_max_len, lens[k]);
        }
    }
    const size_t n_not16 = inter_idx.size();
    const size_t n16_long = idx16.size();
    // End-tracking scans run the packed targets in the packed kernel too, except in local mode
    // when a packed group is wider than the 16-bit column its end tracking keeps (65,535): then
    // every non-long target runs at 32 bits, and only then does the 32-bit layout hold the
    // packed targets as well. The longest short packed target comes first in length order.
    for (uint32_t k : order_idx)
        if (!(all_systolic || lens[k] > long_threshold) && elig16[k]) {
            db->ends32_all = mode == HY_MODE_SW && ((lens[k] + 3) & ~(size_t)3) > 65535;
            break;
        }
    for (uint32_t k : order_idx)
        if (!(all_systolic || lens[k] > long_threshold) && elig16[k]) {
            if (db->ends32_all) inter_idx.push_back(k);
            idx16.push_back(k);
        }

    // Inter layout: groups of 64 consecutive (length-sorted) targets; residues interleaved so a
    // wavefront's load of column word w reads 64 consecutive 32-bit words. It holds the targets
    // the packed kernel cannot take (and the packed ones as well when ends32_all is set).
    {
        const uint32_t pad = db->al_eff;  // padding symbol
        const size_t ns = inter_idx.size();
        std::vector<uint32_t> slot_len(ns), slot_gidx(ns);
        for (size_t s = 0; s < ns; ++s) {
            slot_gidx[s] = inter_idx[s];
            slot_len[s] = (uint32_t)lens[inter_idx[s]];
        }
        uint64_t words = 0, spill = 0;
        auto add_groups = [&](size_t lo, size_t hi) {
            for (size_t s = lo; s < hi; s += 64) {
                Group g{};
                g.first = (uint32_t)s;
                g.count = (uint32_t)std::min<size_t>(64, hi - s);
                g.ncols = (slot_len[s] + 3u) & ~3u;  // whole four-column words
                g.word_off = words;
                g.spill_off = spill;
                words += (uint64_t)((g.ncols + 3) / 4) * 64;
                spill += (uint64_t)(g.ncols + 1) * 64;
                db->groups.push_back(g);
            }
        };
        add_groups(0, n_not16);
        db->groups_score32 = (uint32_t)db->groups.size();
        db->score32_spill = spill;
        add_groups(n_not16, ns);
        db->inter_spill = spill;
        if (dev_upload(&db->d_groups, db->groups) != hipSuccess) return fail(HY_E_DEVICE);
        if (dev_upload(&db->d_lens, slot_len) != hipSuccess) return fail(HY_E_DEVICE);
        if (dev_upload(&db->d_gidx, slot_gidx) != hipSuccess) return fail(HY_E_DEVICE);
        // The interleaved words are gathered on the device from the residues.
        if (hipMalloc((void**)&db->d_words, std::max<uint64_t>(words, 1) * sizeof(uint32_t)) !=
            hipSuccess) {
            db->d_words = nullptr;
            return fail(HY_E_OUT_OF_MEMORY);
        }
        if (hy::launch_db_pack32(db->d_groups, (uint32_t)db->groups.size(), db->d_gidx, db->d_off,
                                 db->d_flat, db->d_words, pad, nullptr) != hipSuccess)
            return fail(HY_E_DEVICE);
    }

    // Packed 16-bit layout: groups of 128 length-sorted targets, lane l holding slots 2l (low
    // half) and 2l+1 (high half); column word w of a lane is a uint2 (low target's four
    // residues, high target's four residues).
    {
        const uint32_t pad = db->al_eff;
        const size_t ns = idx16.size();
        std::vector<uint32_t> slot_len(ns), slot_gidx(ns);
        for (size_t s = 0; s < ns; ++s) {
            slot_gidx[s] = idx16[s];
            slot_len[s] = (uint32_t)lens[idx16[s]];
            db->residues16 += slot_len[s];
        }
        uint64_t words = 0, spill = 0;
        auto add_groups16 = [&](size_t lo, size_t hi) {
            for (size_t s = lo; s < hi; s += 128) {
                Group g{};
                g.first = (uint32_t)s;
                g.count = (uint32_t)std::min<size_t>(128, hi - s);
                g.ncols = (slot_len[s] + 3u) & ~3u;  // whole four-column words
                g.word_off = words;
                g.spill_off = spill;
                words += (uint64_t)((g.ncols + 3) / 4) * 64;
                spill += (uint64_t)(g.ncols + 1) * 64;
                db->groups16.push_back(g);
            }
        };
        add_groups16(0, n16_long);
        db->groups16_short_first = (uint32_t)db->groups16.size();
        add_groups16(n16_long, ns);
        db->max_ncols16 = db->groups16.empty() ? 0 : db->groups16.front().ncols;
        db->inter16_spill = spill;
        // Groups wider than a batched-form slot run in the batched form's systolic list.
        uint32_t bf = 0;
        while (bf < db->groups16.size() && db->groups16[bf].ncols > hy::kBatchSlotMaxCols) {
            const Group& g = db->groups16[bf];
            for (uint32_t k = 0; k < g.count; ++k) {
                const uint32_t t = slot_gidx[g.first + k];
                db->sys_list_batch.push_back(t);
                db->sys_batch_max_len = std::max<uint64_t>(db->sys_batch_max_len, lens[t]);
            }
            ++bf;
        }
        db->groups16_batch_first = bf;
        db->batch_max_ncols16 = bf < db->groups16.size() ? db->groups16[bf].ncols : 0;

        if (dev_upload(&db->d_groups16, db->groups16) != hipSuccess) return fail(HY_E_DEVICE);
        if (dev_upload(&db->d_lens16, slot_len) != hipSuccess) return fail(HY_E_DEVICE);
        if (dev_upload(&db->d_gidx16, slot_gidx) != hipSuccess) return fail(HY_E_DEVICE);
        if (hipMalloc((void**)&db->d_words16, std::max<uint64_t>(words, 1) * sizeof(uint2)) !=
            hipSuccess) {
            db->d_words16 = nullptr;
            return fail(HY_E_OUT_OF_MEMORY);
        }
        if (hy::launch_db_pack16(db->d_groups16, (uint32_t)db->groups16.size(), db->d_gidx16,
                                 db->d_off, db->d_flat, db->d_words16, pad, nullptr) !=
            hipSuccess)
            return fail(HY_E_DEVICE);
    }
    if (hipDeviceSynchronize() != hipSuccess) return fail(HY_E_DEVICE);
    if (dev_upload(&db->d_sys_list, db->sys_list) != hipSuccess) return fail(HY_E_DEVICE);
    if (dev_upload(&db->d_sys_list_batch, db->sys_list_batch) != hipSuccess)
        return fail(HY_E_DEVICE);
    // Targets outside the packed traceback: not in the packed layout, or in a group whose columns
    // exceed the 16-bit end tracking.
    {
        std::vector<char> packed_tb(n, 0);
        for (const Group& g : db->groups16)
            if (g.ncols <= 65535u)
                for (uint32_t k = 0; k < g.count; ++k) packed_tb[idx16[g.first + k]] = 1;
        for (uint32_t k = 0; k < n; ++k)
            if (!packed_tb[k]) db->trace_rest.push_back(k);
        if (dev_upload(&db->d_trace_rest, db->trace_rest) != hipSuccess) return fail(HY_E_DEVICE);
    }
    (void)rc;
    *out = db;
    return HY_OK;
}

extern "C" int hy_db_retain(hy_db* db) {
    if (!db) return HY_E_NULL_ARGUMENT;
    db->refs.fetch_add(1);
    return HY_OK;
}

extern "C" void hy_db_free(hy_db* db) {
    if (db && db->refs.fetch_sub(1) == 1) {
        std::lock_guard<std::recursive_mutex> api(hy::api_lock());
        delete db;
    }
}

extern "C" int hy_db_info(const hy_db* db, uint32_t* widest, uint64_t* n, uint64_t* mql,
                          uint64_t* mtl, uint32_t* mode, uint32_t* st) {
    if (!db) return HY_E_NULL_ARGUMENT;
    if (widest) *widest = (uint32_t)db->widest_bits;
    if (n) *n = db->n;
    if (mql) *mql = db->mql;
    if (mtl) *mtl = db->mtl;
    if (mode) *mode = db->mode;
    if (st) *st = db->search_type;
    return HY_OK;
}

extern "C" int hy_db_partitions(const hy_db* db, uint64_t* n8, uint64_t* n16, uint64_t* n32) {
    if (!db) return HY_E_NULL_ARGUMENT;
    if (n8) *n8 = db->count_by_width[0];
    if (n16) *n16 = db->count_by_width[1];
    if (n32) *n32 = db->count_by_width[2];
    return HY_OK;
}

extern "C" int hy_scratch_new(const hy_db* db, hy_scratch** out) {
    if (!db || !out) return HY_E_NULL_ARGUMENT;
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    *out = nullptr;
    if (hipSetDevice(db->device) != hipSuccess) return HY_E_DEVICE;
    hy_scratch* s = new (std::nothrow) hy_scratch();
    if (!s) return HY_E_OUT_OF_MEMORY;
    s->device = db->device;
    s->db = db;
    s->q16 = db->q16;
    s->waves16 = db->waves16;
    s->bq16 = db->q16;
    // Batched submissions: each query (pair) adds one item per packed group to the persistent
    // queue, so a database of few groups needs more queries per submission to give the resident
    // workgroups (about 8 per CU) several items each; 256 otherwise, at most 4,096, and no more
    // than keeps the per-query winner rows of queries beyond 512 rows within 256 MB.
    if (!db->groups16.empty()) {
        const uint64_t slots = 8ull * std::max<uint32_t>(hy::device_cus(), 1);
        const uint64_t groups = db->groups16.size();
        const uint64_t row_bytes = db->mql > 512 ? (db->mtl + 1) * sizeof(int2) : 0;
        uint64_t chunk = 256;
        while (chunk < 4096 && groups * chunk < 4 * slots &&
               2 * chunk * row_bytes <= (256ull << 20))
            chunk *= 2;
        s->bchunk = (uint32_t)chunk;
    }
    auto fail = [&](int code) { delete s; return code; };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->join, hipEventDisableTiming) != hipSuccess)
        return fail(HY_E_DEVICE);
    const size_t qcap = std::max<uint64_t>(db->mql, 1);
    if (hipMalloc((void**)&s->d_query, qcap) != hipSuccess) return fail(HY_E_OUT_OF_MEMORY);
    if (hipHostMalloc((void**)&s->h_query, qcap) != hipSuccess) return fail(HY_E_OUT_OF_MEMORY);
    // Boundary rows: the packed kernel's groups from entry 0; the 32-bit groups that run beside it
    // (targets that fail the packed width proof, on the auxiliary stream) after them. An
    // end-tracking scan that runs every 32-bit group alone uses the buffer from entry 0.
    s->spill_entries = std::max<uint64_t>(
        std::max(db->inter_spill, db->inter16_spill + db->score32_spill), 1);
    if (hipMalloc((void**)&s->d_spill, s->spill_entries * sizeof(int2)) != hipSuccess)
        return fail(HY_E_OUT_OF_MEMORY);
    // Systolic boundary rows are needed only when a query can exceed one pass (512 rows).
    const bool multipass = db->mql > 512;
    s->sys_spill_stride = db->sys_max_len + 1;
    const uint64_t sys_entries =
        std::max<uint64_t>(multipass ? (db->sys_list.size() * s->sys_spill_stride) : 1,
                           multipass ? db->mtl + 1 : 1);
    if (hipMalloc((void**)&s->d_sys_spill, sys_entries * sizeof(int2)) != hipSuccess)
        return fail(HY_E_OUT_OF_MEMORY);
    s->sys_spill_entries = sys_entries;
    const uint64_t n = db->n;
    if (hipMalloc((void**)&s->d_scores, n * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void**)&s->d_qend, n * sizeof(int64_t)) != hipSuccess ||
        hipMalloc((void**)&s->d_tend, n * sizeof(int64_t)) != hipSuccess)
        return fail(HY_E_OUT_OF_MEMORY);
    if (hipMalloc((void**)&s->d_key, sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc((void**)&s->h_result, sizeof(Result)) != hipSuccess)
        return fail(HY_E_OUT_OF_MEMORY);
    // Progress words and the ticket counter of the chained pipelined kernel (databases with few
    // packed groups), zeroed once: every launch counts in an epoch of its own and knows how many
    // tickets the launches before it took.
    const uint64_t chain_words =
        (uint64_t)std::max<size_t>(db->groups16.size(), 1) * hy::kChainMaxLinks + 1;
    if (hipMalloc((void**)&s->d_chain, chain_words * sizeof(unsigned long long)) != hipSuccess)
        return fail(HY_E_OUT_OF_MEMORY);
    if (hipMemset(s->d_chain, 0, chain_words * sizeof(unsigned long long)) != hipSuccess)
        return fail(HY_E_DEVICE);
    if (hipHostMalloc((void**)&s->h_chain_err, sizeof(unsigned int)) != hipSuccess)
        return fail(HY_E_OUT_OF_MEMORY);
    *s->h_chain_err = 0u;
    // The end-tracking chained kernel's fold slots (10.5 KB per group), for the databases of few
    // packed groups it serves; with more groups chain_shape never picks it.
    uint64_t fold_entries = 0;
    if (db->groups16.size() <= 2u * (size_t)hy::device_cus() || chain_forced()) {
        fold_entries = hy::chain_fold_entries(db->groups16.size());
        if (hipMalloc((void**)&s->d_chain_fold, fold_entries * sizeof(uint2)) != hipSuccess)
            return fail(HY_E_OUT_OF_MEMORY);
    }
    s->dev_bytes = qcap + s->spill_entries * sizeof(int2) + sys_entries * sizeof(int2) +
                   n * (sizeof(int32_t) + 2 * sizeof(int64_t)) + sizeof(unsigned long long) +
                   chain_words * sizeof(unsigned long long) +
                   fold_entries * sizeof(uint2);
    *out = s;
    return HY_OK;
}

extern "C" void hy_scratch_free(hy_scratch* s) {
    if (!s) return;
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    delete s;
}

// ---------------------------------------------------------------------------------------------
// Scans
// ---------------------------------------------------------------------------------------------

namespace hy {

// Contract checks the reference turns into panics (src/database.rs:79-89): an over-long query or
// a symbol outside the alphabet.
int check_query(const hy_db* db, const uint8_t* q, size_t qlen) {
    if (qlen > db->mql) return HY_E_QUERY_TOO_LONG;
    if (qlen && !q) return HY_E_NULL_ARGUMENT;
    for (size_t k = 0; k < qlen; ++k)
        if ((uint32_t)q[k] >= db->al) return HY_E_QUERY_SYMBOL;
    return HY_OK;
}

int stage_query(const hy_db* db, hy_scratch* s, const uint8_t* q, size_t qlen) {
    if (hipSetDevice(db->device) != hipSuccess) return HY_E_DEVICE;
    if (qlen) {
        std::memcpy(s->h_query, q, qlen);
        if (hipMemcpyAsync(s->d_query, s->h_query, qlen, hipMemcpyHostToDevice, s->stream) !=
            hipSuccess)
            return HY_E_DEVICE;
    }
    return HY_OK;
}

hy::InterArgs inter_args(const hy_db* db, hy_scratch* s, int m) {
    hy::InterArgs a{};
    a.groups = db->d_groups;
    a.words = db->d_words;
    a.lens = db->d_lens;
    a.gidx = db->d_gidx;
    a.query = s->d_query;
    a.m = m;
    a.matrix = db->d_matrix;
    a.al = (int)db->al;
    a.al_eff = (int)db->al_eff;
    a.go = db->go;
    a.ge = db->ge;
    a.spill = s->d_spill;
    return a;
}

hy::SysArgs sys_args(const hy_db* db, hy_scratch* s, int m) {
    hy::SysArgs a{};
    a.flat = db->d_flat;
    a.off = db->d_off;
    a.list = db->d_sys_list;
    a.query = s->d_query;
    a.m = m;
    a.matrix = db->d_matrix;
    a.al = (int)db->al;
    a.al_eff = (int)db->al_eff;
    a.go = db->go;
    a.ge = db->ge;
    a.spill = s->d_sys_spill;
    a.spill_stride = s->sys_spill_stride;
    return a;
}

// Tile rows of the pipelined packed kernel for a scan of `ngroups` packed groups with a query of
// m rows (score-only, or with ends), or 0 for the one-wavefront-per-group kernel. One wavefront
// per group leaves most SIMDs idle when the groups number fewer than the CUs (a 10,000-target
// database has 79); the pipelined kernel gives each group a workgroup whose wavefronts take one
// query tile each. It needs 2 to 8 tiles and its LDS within 64 KB. HYALITE_HIP_PIPE=0 turns it
// off, =1 uses it whenever the query and LDS allow (for tests).
int pipe_rows(const hy_db* db, int m, uint32_t ngroups, bool ends = false) {
    const char* e = std::getenv("HYALITE_HIP_PIPE");
    const int mode_env = e ? std::atoi(e) : -1;
    int q = ends ? hy::inter16_pipe_rows_ends((int)db->mode, m) : hy::inter16_pipe_rows(m);
    // HYALITE_HIP_PIPE_ROWS=16|24|32 asks for that tile height where the query fits its tiles
    // (at most 24 for local-mode ends).
    if (const char* r = std::getenv("HYALITE_HIP_PIPE_ROWS")) {
        const int want = std::atoi(r);
        if ((want == 16 || want == 24 || want == 32) && q != 0 &&
            !(ends && db->mode == HY_MODE_SW && want > 24) &&
            (m + want - 1) / want >= 2 && (m + want - 1) / want <= hy::kPipeMaxWaves)
            q = want;
    }
    if (mode_env == 0 || q == 0 || ngroups == 0) return 0;
    const int W = (m + q - 1) / q;
    if (hy::inter16_pipe_lds_bytes((int)db->al_eff, q, W, ends) > 64u * 1024u) return 0;
    if (mode_env != 1 && ngroups > hy::device_cus()) return 0;
    return q;
}

// Shape of the chained pipelined kernel for a scan of `ngroups` packed groups with a query of m
// rows (score-only, or with ends), or rows = 0 for none. The pipelined kernel keeps a group on one CU, and the
// one-wavefront kernel a group on one SIMD, so a database of few groups leaves most of the chip
// idle and its longest group sets the time; chaining spreads a group's query tiles over `chunks`
// workgroups (CUs) of `waves` wavefronts. The shape is the first of (tile rows x waves) 8x4, 8x8,
// 16x8, 16x4, 24x4, 24x8, 32x4, 32x8 that gives 2-8 chunks within 3,200 wavefronts (12.5 per
// CU); a query too long for the pipelined kernel takes the valid shape of fewest wavefronts when
// none fits that budget. Measured on DNA databases of 1,280-20,000 targets (10-157 groups) with
// queries of 40-400 rows, that rule picks the fastest shape or one within 5% of it; for 10,000
// targets and a 100-row query (8x4, 4 chunks) 0.169 ms per scan against 0.218 ms pipelined, for
// 400 rows (16x8, 4 chunks) 0.46 ms against 2.57 ms with one wavefront per group. Databases of
// more than 2 x CUs groups keep the other kernels. HYALITE_HIP_CHAIN=0 turns it off, =1 takes a
// valid shape whatever its size; HYALITE_HIP_CHAIN_ROWS (8, 16, 24, 32) and
// HYALITE_HIP_CHAIN_WAVES (1-8) fix the shape, HYALITE_HIP_CHAIN_SLACK (1-64) and
// HYALITE_HIP_CHAIN_EVERY (1-16) the link's reader lead and publication interval (tests and
// tuning); HYALITE_HIP_CHAIN_SPIN_LIMIT (0 up to the default) the polls before a wait gives up
// (tests of the error path: the scan then reports HY_E_INTERNAL).
struct ChainShape {
    int rows = 0, waves = 0, tiles = 0, chunks = 0, slack = 4, every = 1;
    uint32_t spin_limit = hy::kChainSpinLimit;
};
ChainShape chain_shape(const hy_db* db, int m, uint32_t ngroups, bool ends = false) {
    ChainShape c;
    const char* e = std::getenv("HYALITE_HIP_CHAIN");
    const int mode_env = e ? std::atoi(e) : -1;
    if (mode_env == 0 || ngroups == 0 || m <= 8) return c;
    const uint32_t cus = hy::device_cus();
    if (mode_env != 1 && ngroups > 2 * cus) return c;
    int want_rows = 0, want_waves = 0;
    if (const char* r = std::getenv("HYALITE_HIP_CHAIN_ROWS")) {
        const int v = std::atoi(r);
        if (v == 8 || v == 16 || v == 24 || v == 32) want_rows = v;
    }
    if (const char* w = std::getenv("HYALITE_HIP_CHAIN_WAVES")) {
        const int v = std::atoi(w);
        if (v >= 1 && v <= 8) want_waves = v;
    }
    if (const char* sl = std::getenv("HYALITE_HIP_CHAIN_SLACK")) {
        const int v = std::atoi(sl);
        if (v >= 1 && v <= 64) c.slack = v;
    }
    if (const char* ev = std::getenv("HYALITE_HIP_CHAIN_EVERY")) {
        const int v = std::atoi(ev);
        if (v >= 1 && v <= 16) c.every = v;
    }
    if (const char* sp = std::getenv("HYALITE_HIP_CHAIN_SPIN_LIMIT")) {
        const long v = std::atol(sp);
        if (v >= 0 && v <= (long)hy::kChainSpinLimit) c.spin_limit = (uint32_t)v;
    }
    static const int kShapes[][2] = {{8, 4},  {8, 8},  {16, 8}, {16, 4},
                                     {24, 4}, {24, 8}, {32, 4}, {32, 8}};
    const uint64_t g8 = (uint64_t)(ngroups + 7) / 8 * 8;
    const uint64_t budget = 3200;  // wavefronts
    const bool no_pipe =
        (ends ? hy::inter16_pipe_rows_ends((int)db->mode, m) : hy::inter16_pipe_rows(m)) == 0;
    struct Cand {
        int rows = 0, waves = 0, tiles = 0, chunks = 0;
        uint64_t total = 0;  // wavefronts of the launch
    };
    // A shape is valid when it gives 2-8 chunks and its workgroup's LDS fits 64 KB (and, with
    // ends, its tile height is one the end-tracking chained kernel has).
    auto make = [&](int rows, int waves, Cand& out) {
        if (!hy::inter16_chain_rows_ok((int)db->mode, rows, ends)) return false;
        const int tiles = (m + rows - 1) / rows;
        waves = std::min(waves, tiles);
        const int chunks = (tiles + waves - 1) / waves;
        if (chunks < 2 || chunks > hy::kChainMaxChunks) return false;
        if (hy::inter16_chain_lds_bytes((int)db->al_eff, rows, waves, ends) > 64u * 1024u)
            return false;
        out.rows = rows;
        out.waves = waves;
        out.tiles = tiles;
        out.chunks = chunks;
        out.total = g8 * (uint64_t)chunks * (uint64_t)waves;
        return true;
    };
    Cand pick, cd;
    if (want_rows || want_waves) {
        if (make(want_rows ? want_rows : 16, want_waves ? want_waves : 4, cd) &&
            (mode_env == 1 || cd.total <= budget || no_pipe))
            pick = cd;
    } else {
        for (const auto& sh : kShapes)
            if (make(sh[0], sh[1], cd) && (mode_env == 1 || cd.total <= budget)) {
                pick = cd;
                break;
            }
        if (pick.rows == 0 && no_pipe)
            for (const auto& sh : kShapes)
                if (make(sh[0], sh[1], cd) && (pick.rows == 0 || cd.total < pick.total)) pick = cd;
    }
    if (pick.rows == 0) return c;
    c.rows = pick.rows;
    c.waves = pick.waves;
    c.tiles = pick.tiles;
    c.chunks = pick.chunks;
    return c;
}

// Arguments of one chained launch of shape cs (a new epoch; the tickets it takes are counted
// after the launch).
hy::ChainArgs chain_args(const hy_db* db, hy_scratch* s, const ChainShape& cs) {
    hy::ChainArgs ca{};
    ca.progress = s->d_chain;
    ca.epoch = (unsigned long long)(++s->chain_epoch) << 32;
    ca.err = s->h_chain_err;
    ca.ticket = s->d_chain + std::max<size_t>(db->groups16.size(), 1) * hy::kChainMaxLinks;
    ca.ticket_base = s->chain_tickets;
    ca.tiles = cs.tiles;
    ca.chunks = cs.chunks;
    ca.slack = cs.slack;
    ca.every = cs.every;
    ca.spin_limit = cs.spin_limit;
    ca.fold = s->d_chain_fold;
    return ca;
}

// Enqueue the per-target pass. ends: write (score, qend, tend) per target; otherwise write scores
// (when `scores` is set) and/or fold the argmax key (when `key` is set).
int enqueue_targets(const hy_db* db, hy_scratch* s, int m, bool ends, bool write_scores,
                    unsigned long long* key) {
    hy::InterArgs ia = inter_args(db, s, m);
    ia.out_score = (ends || write_scores) ? s->d_scores : nullptr;
    ia.out_qend = s->d_qend;
    ia.out_tend = s->d_tend;
    ia.key = ends ? nullptr : key;

    // Scans run the targets the packed kernel accepts there and the rest at 32 bits, with or
    // without end tracking. The local-mode end tracking keeps a tile's best column in 16 bits, so
    // its groups must stay within 65,535 columns (else every non-long target runs at 32 bits);
    // the other modes keep their ends in 32 bits.
    const bool has16 = db->groups16.size() > db->groups16_short_first;
    const bool packed_ends = ends && has16 && !db->ends32_all;
    const bool run16 = (!ends || packed_ends) && has16;
    const uint32_t g32 =
        (ends && !packed_ends) ? (uint32_t)db->groups.size() : db->groups_score32;

    // Work beside the packed kernel goes to the auxiliary stream, forked from and joined back
    // into the main one, so its few long wavefronts overlap the packed kernel instead of
    // trailing or preceding it: the long targets (systolic kernel), and the 32-bit groups when
    // the packed kernel runs too. Those groups hold targets that fail the packed width proof,
    // the longest ones below the long-target threshold: a handful of groups, each lane running
    // m x (its length) cells. Measured on 2M protein targets with queries up to 3,000: 9 such
    // groups ahead of the packed kernel on one stream took 71 ms per query on average (4.4k
    // GCUPS); beside it, the packed kernel's rate.
    const bool tail = !db->sys_list.empty();
    const bool side32 = run16 && g32 > 0;
    const bool fork = tail || side32;
    if (fork) {
        if (hipEventRecord(s->fork, s->stream) != hipSuccess ||
            hipStreamWaitEvent(s->aux, s->fork, 0) != hipSuccess)
            return HY_E_DEVICE;
        if (s->profile && hipEventRecord(s->ev[2], s->aux) != hipSuccess) return HY_E_DEVICE;
        if (side32) {
            hy::InterArgs sa = ia;
            sa.spill = s->d_spill + db->inter16_spill;  // a region of their own
            if (hy::launch_inter32((int)db->mode, ends, sa, g32, s->aux) != hipSuccess)
                return HY_E_DEVICE;
        }
        if (tail) {
            hy::SysArgs sa = sys_args(db, s, m);
            sa.out_score = ia.out_score;
            sa.out_qend = s->d_qend;
            sa.out_tend = s->d_tend;
            sa.key = ia.key;
            if (hy::launch_sys32((int)db->mode, ends, false, sa, (uint32_t)db->sys_list.size(),
                                 s->aux) != hipSuccess)
                return HY_E_DEVICE;
        }
        if (s->profile && hipEventRecord(s->ev[3], s->aux) != hipSuccess) return HY_E_DEVICE;
        if (hipEventRecord(s->join, s->aux) != hipSuccess) return HY_E_DEVICE;
    }

    if (s->profile && hipEventRecord(s->ev[0], s->stream) != hipSuccess) return HY_E_DEVICE;
    if (!side32 && hy::launch_inter32((int)db->mode, ends, ia, g32, s->stream) != hipSuccess)
        return HY_E_DEVICE;
    s->last_kernel = HY_KERNEL_NONE;
    if (run16) {
        hy::Inter16Args pa{};
        pa.groups = db->d_groups16;
        pa.group_first = db->groups16_short_first;
        pa.ngroups = (uint32_t)db->groups16.size() - db->groups16_short_first;
        pa.words = db->d_words16;
        pa.lens = db->d_lens16;
        pa.gidx = db->d_gidx16;
        pa.query = s->d_query;
        pa.m = m;
        pa.matrix = db->d_matrix;
        pa.al = (int)db->al;
        pa.al_eff = (int)db->al_eff;
        pa.go = db->go;
        pa.ge = db->ge;
        pa.spill = reinterpret_cast<uint2*>(s->d_spill);
        pa.out_score = ia.out_score;
        pa.key = key;
        if (ends) {
            pa.key = nullptr;
            pa.out_qend = s->d_qend;
            pa.out_tend = s->d_tend;
            const ChainShape cs =
                s->d_chain_fold ? chain_shape(db, m, pa.ngroups, true) : ChainShape();
            if (cs.rows != 0) {
                // Few groups: the chained kernel, each chunk's end cells passed on to the next.
                hy::ChainArgs ca = chain_args(db, s, cs);
                if (hy::launch_inter16_chain((int)db->mode, pa, ca, cs.rows, cs.waves, true,
                                             s->stream) != hipSuccess)
                    return HY_E_DEVICE;
                s->chain_tickets += (uint64_t)(pa.ngroups + 7) / 8 * 8 * (uint64_t)cs.chunks;
                s->last_kernel = HY_KERNEL_CHAINED;
            } else if (const int pq = pipe_rows(db, m, pa.ngroups, true)) {
                // Few groups: the pipelined kernel, its wavefronts' end cells folded.
                if (hy::launch_inter16_pipe((int)db->mode, pa, pq, true, s->stream) !=
                    hipSuccess)
                    return HY_E_DEVICE;
                s->last_kernel = HY_KERNEL_PIPELINED;
            } else {
                if (hy::launch_inter16_ends((int)db->mode, pa, s->stream) != hipSuccess)
                    return HY_E_DEVICE;
                s->last_kernel = HY_KERNEL_ENDS;
            }
        } else if (const ChainShape cs = chain_shape(db, m, pa.ngroups); cs.rows != 0) {
            hy::ChainArgs ca = chain_args(db, s, cs);
            if (hy::launch_inter16_chain((int)db->mode, pa, ca, cs.rows, cs.waves, false,
                                         s->stream) != hipSuccess)
                return HY_E_DEVICE;
            s->chain_tickets += (uint64_t)(pa.ngroups + 7) / 8 * 8 * (uint64_t)cs.chunks;
            s->last_kernel = HY_KERNEL_CHAINED;
        } else if (const int pq = pipe_rows(db, m, pa.ngroups)) {
            if (hy::launch_inter16_pipe((int)db->mode, pa, pq, false, s->stream) != hipSuccess)
                return HY_E_DEVICE;
            s->last_kernel = HY_KERNEL_PIPELINED;
        } else {
            if (hy::launch_inter16((int)db->mode, pa, s->q16, s->waves16, s->stream) !=
                hipSuccess)
                return HY_E_DEVICE;
            s->last_kernel = HY_KERNEL_PACKED;
        }
    }
    if (s->profile && hipEventRecord(s->ev[1], s->stream) != hipSuccess) return HY_E_DEVICE;
    if (fork && hipStreamWaitEvent(s->stream, s->join, 0) != hipSuccess) return HY_E_DEVICE;
    if (s->profile) {
        s->pending = true;
        s->pending_tail = fork;
    }
    return HY_OK;
}

// After the scan's stream synchronisation: fold the recorded kernel times into the counters.
void collect_profile(hy_scratch* s) {
    if (!s->profile || !s->pending) return;
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, s->ev[0], s->ev[1]) == hipSuccess &&
        (!s->pending_tail || hipEventElapsedTime(&b, s->ev[2], s->ev[3]) == hipSuccess)) {
        s->inter_ms += a;
        s->tail_ms += b;
        s->launches += 1;
    }
    s->pending = false;
}

int scan_done(hy_scratch* s) {
    collect_profile(s);
    if (s->h_chain_err && __atomic_load_n(s->h_chain_err, __ATOMIC_ACQUIRE) != 0u) {
        *s->h_chain_err = 0u;
        return HY_E_INTERNAL;
    }
    return HY_OK;
}

}  // namespace hy

using hy::check_query;
using hy::collect_profile;
using hy::scan_done;
using hy::inter_args;
using hy::sys_args;
using hy::enqueue_targets;
using hy::stage_query;

extern "C" int hy_scan(const hy_db* db, hy_scratch* s, const uint8_t* query, size_t qlen,
                       int32_t* score, uint64_t* db_index, int64_t* query_end,
                       int64_t* target_end) {
    if (!db || !s || !score || !db_index || !query_end || !target_end) return HY_E_NULL_ARGUMENT;
    if (s->db != db) return HY_E_INVALID_ARGUMENT;
    int rc = check_query(db, query, qlen);
    if (rc != HY_OK) return rc;
    if ((rc = stage_query(db, s, query, qlen)) != HY_OK) return rc;
    const int m = (int)qlen;
    if (hy::launch_fill_u64(s->d_key, 1, 0ull, s->stream) != hipSuccess) return HY_E_DEVICE;
    if ((rc = enqueue_targets(db, s, m, false, false, s->d_key)) != HY_OK) return rc;
    // The result record goes straight to the pinned host record (device-accessible), with no
    // copy launch after it; marked unchecked until the last kernel writes it.
    Result* res = s->h_result;
    res->check = 1;
    const bool ends = db->search_type != HY_SEARCH_SCORE;
    if (ends) {
        // The winner's end positions: one pass over the winning target only (the role of
        // align_core on the winner in src/database.rs:183-197), chosen on the device. The band
        // kernel's wavefronts split the query rows between them (a long query's winner pass:
        // 2,168 rows against its winning target took 13 ms in the one-wavefront systolic kernel);
        // it takes alphabets up to 63 symbols, and boundary rows (queries beyond one of its
        // passes) only when the scratch holds a target's worth.
        const bool band = db->al <= hy::kBandMaxAlphabet &&
                          (m <= hy::band_rows_per_pass(m) || s->sys_spill_entries >= db->mtl + 1);
        if (band) {
            hy::BandArgs ba{};
            ba.queries = s->d_query;
            ba.wm = m;
            ba.targets = db->d_flat;
            ba.toff = db->d_off;
            ba.matrix = db->d_matrix;
            ba.al = (int)db->al;
            ba.go = db->go;
            ba.ge = db->ge;
            ba.spill = s->d_sys_spill;
            ba.spill_stride = 0;
            ba.winner_key = s->d_key;
            ba.result = res;
            if (hy::launch_band32((int)db->mode, hy::kBandEnds, ba, 1, m, s->stream) != hipSuccess)
                return HY_E_DEVICE;
        } else {
            hy::SysArgs sa = sys_args(db, s, m);
            sa.winner_key = s->d_key;
            sa.result = res;
            if (hy::launch_sys32((int)db->mode, true, true, sa, 1, s->stream) != hipSuccess)
                return HY_E_DEVICE;
        }
    } else {
        if (hy::launch_key_result(s->d_key, res, s->stream) != hipSuccess)
            return HY_E_DEVICE;
    }
    if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
    if ((rc = scan_done(s)) != HY_OK) return rc;
    const Result r = *res;
    if (r.check != 0) return HY_E_INTERNAL;
    *score = r.score;
    *db_index = r.db_index;
    *query_end = r.query_end;
    *target_end = r.target_end;
    return HY_OK;
}

extern "C" int hy_scan_all(const hy_db* db, hy_scratch* s, const uint8_t* query, size_t qlen,
                           int32_t* scores, int64_t* query_ends, int64_t* target_ends) {
    if (!db || !s || !scores) return HY_E_NULL_ARGUMENT;
    if (s->db != db) return HY_E_INVALID_ARGUMENT;
    int rc = check_query(db, query, qlen);
    if (rc != HY_OK) return rc;
    if ((rc = stage_query(db, s, query, qlen)) != HY_OK) return rc;
    const bool ends = db->search_type != HY_SEARCH_SCORE;
    if ((rc = enqueue_targets(db, s, (int)qlen, ends, true, nullptr)) != HY_OK) return rc;
    const size_t n = (size_t)db->n;
    hy::Segment segs[3];
    int nseg = 0;
    segs[nseg++] = {scores, s->d_scores, n * sizeof(int32_t)};
    if (ends && query_ends) segs[nseg++] = {query_ends, s->d_qend, n * sizeof(int64_t)};
    if (ends && target_ends) segs[nseg++] = {target_ends, s->d_tend, n * sizeof(int64_t)};
    if ((rc = hy::download(s->ring, s->device, segs, nseg, s->stream)) != HY_OK) return rc;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
    if ((rc = scan_done(s)) != HY_OK) return rc;
    if (!ends) {
        for (size_t k = 0; k < n; ++k) {
            if (query_ends) query_ends[k] = -1;
            if (target_ends) target_ends[k] = -1;
        }
    }
    return HY_OK;
}

extern "C" int hy_scan_scores(const hy_db* db, hy_scratch* s, const uint8_t* query, size_t qlen,
                              int32_t* scores) {
    if (!db || !s || !scores) return HY_E_NULL_ARGUMENT;
    if (s->db != db) return HY_E_INVALID_ARGUMENT;
    int rc = check_query(db, query, qlen);
    if (rc != HY_OK) return rc;
    if ((rc = stage_query(db, s, query, qlen)) != HY_OK) return rc;
    if ((rc = enqueue_targets(db, s, (int)qlen, false, true, nullptr)) != HY_OK) return rc;
    if (hipMemcpyAsync(scores, s->d_scores, (size_t)db->n * sizeof(int32_t),
                       hipMemcpyDeviceToHost, s->stream) != hipSuccess)
        return HY_E_DEVICE;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
    if ((rc = scan_done(s)) != HY_OK) return rc;
    return HY_OK;
}

// ---------------------------------------------------------------------------------------------
// Batched form
// ---------------------------------------------------------------------------------------------

namespace {

// The batched form pairs queries in the lanes (k_inter16q_batch) in local mode unless the
// scratch's launch shapes say otherwise.
bool batch_query_pairs(const hy_db* db, const hy_scratch* s) {
    return db->mode == HY_MODE_SW && s->bpairs != 0;
}

// Release the batched form's buffers and the graph captured for them.
void release_batch(hy_scratch* s) {
    dev_free(s->d_bq); dev_free(s->d_bmeta); dev_free(s->d_bkeys); dev_free(s->d_bres);
    dev_free(s->d_slots); dev_free(s->d_bwin_spill); dev_free(s->d_btail_spill);
    if (s->h_bq) (void)hipHostFree(s->h_bq);
    if (s->h_bmeta) (void)hipHostFree(s->h_bmeta);
    if (s->h_bres) (void)hipHostFree(s->h_bres);
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->d_bq = nullptr; s->d_bmeta = nullptr; s->d_bkeys = nullptr; s->d_bres = nullptr;
    s->d_slots = nullptr; s->d_bwin_spill = nullptr; s->d_btail_spill = nullptr;
    s->h_bq = nullptr; s->h_bmeta = nullptr; s->h_bres = nullptr;
    s->gexec = nullptr; s->graph = nullptr;
    s->bcap = 0;
    s->nslots = 0;
    s->batch_bytes = 0;
}

}  // namespace

extern "C" int hy_scratch_reserve_batch(hy_scratch* s, size_t max_queries) {
    if (!s) return HY_E_NULL_ARGUMENT;
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    if (max_queries == 0 || max_queries > 65536) return HY_E_INVALID_ARGUMENT;
    if (s->bcap >= max_queries) return HY_OK;
    const hy_db* db = s->db;
    if (hipSetDevice(s->device) != hipSuccess) return HY_E_DEVICE;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
    release_batch(s);  // a smaller reservation

    const uint32_t cap = (uint32_t)max_queries;
    const uint32_t stride = (uint32_t)std::max<uint64_t>(db->mql, 1);
    const size_t qbytes = (size_t)cap * stride;
    const size_t meta = (size_t)3 * cap + 2;
    if (hipMalloc((void**)&s->d_bq, qbytes) != hipSuccess ||
        hipHostMalloc((void**)&s->h_bq, qbytes) != hipSuccess ||
        hipMalloc((void**)&s->d_bmeta, meta * sizeof(uint32_t)) != hipSuccess ||
        hipHostMalloc((void**)&s->h_bmeta, meta * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&s->d_bkeys, cap * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void**)&s->d_bres, cap * sizeof(hy::Result)) != hipSuccess ||
        hipHostMalloc((void**)&s->h_bres, cap * sizeof(hy::Result)) != hipSuccess)
        return HY_E_OUT_OF_MEMORY;
    std::memset(s->h_bq, 0, qbytes);
    std::memset(s->h_bmeta, 0, meta * sizeof(uint32_t));
    if (db->groups16.size() > db->groups16_batch_first) {
        uint32_t per_cu =
            batch_query_pairs(db, s)
                ? hy::inter16q_batch_occupancy(s->bq16, (int)db->al_eff)
                : hy::inter16_batch_occupancy((int)db->mode, s->bq16, (int)db->al_eff);
        if (s->bwpc > 0) per_cu = std::min<uint32_t>(per_cu, (uint32_t)s->bwpc);
        s->nslots = per_cu * hy::device_cus();
        s->slot_stride = (uint64_t)(db->batch_max_ncols16 + 1) * 64;
        if (s->nslots == 0 ||
            hipMalloc((void**)&s->d_slots, (size_t)s->nslots * s->slot_stride * sizeof(uint2)) !=
                hipSuccess)
            return HY_E_OUT_OF_MEMORY;
    }
    const bool multipass = db->mql > 512;
    if (multipass) {
        if (hipMalloc((void**)&s->d_bwin_spill, (size_t)cap * (db->mtl + 1) * sizeof(int2)) !=
            hipSuccess)
            return HY_E_OUT_OF_MEMORY;
        if (!db->sys_list_batch.empty() &&
            hipMalloc((void**)&s->d_btail_spill, db->sys_list_batch.size() *
                                                     (db->sys_batch_max_len + 1) * sizeof(int2)) !=
                hipSuccess)
            return HY_E_OUT_OF_MEMORY;
    }
    for (auto& e : s->bev)
        if (!e && hipEventCreate(&e) != hipSuccess) return HY_E_DEVICE;
    s->bq_stride = stride;
    s->bcap = cap;
    s->batch_bytes = qbytes + meta * sizeof(uint32_t) +
                     cap * (sizeof(unsigned long long) + sizeof(hy::Result)) +
                     (uint64_t)s->nslots * s->slot_stride * sizeof(uint2);
    if (multipass)
        s->batch_bytes += (uint64_t)cap * (db->mtl + 1) * sizeof(int2) +
                          db->sys_list_batch.size() * (db->sys_batch_max_len + 1) * sizeof(int2);
    return HY_OK;
}

namespace {

// Enqueue one batched submission of up to bcap queries whose bytes and metadata are already in
// the pinned staging buffers. Every shape is fixed by the reservation, so the same sequence can
// be captured once as a graph and replayed: per-batch values live in device memory.
int enqueue_batch(const hy_db* db, hy_scratch* s, bool timed) {
    const uint32_t cap = s->bcap;
    uint32_t* d_qoff = s->d_bmeta;
    uint32_t* d_qlen = s->d_bmeta + cap;
    uint32_t* d_qorder = s->d_bmeta + 2 * cap;
    uint32_t* d_nq = s->d_bmeta + 3 * cap;
    unsigned int* d_counter = s->d_bmeta + 3 * cap + 1;
    const hipStream_t st = s->stream;
    if (hipMemcpyAsync(s->d_bq, s->h_bq, (size_t)cap * s->bq_stride, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(s->d_bmeta, s->h_bmeta, ((size_t)3 * cap + 1) * sizeof(uint32_t),
                       hipMemcpyHostToDevice, st) != hipSuccess)
        return HY_E_DEVICE;
    if (hy::launch_batch_reset(s->d_bkeys, cap, d_counter, st) != hipSuccess) return HY_E_DEVICE;

    // Long targets the packed kernel cannot take: one systolic launch per query slot on the
    // auxiliary stream (slots past the batch size exit at once).
    const bool tail = !db->sys_list_batch.empty();
    if (tail) {
        if (hipEventRecord(s->fork, st) != hipSuccess ||
            hipStreamWaitEvent(s->aux, s->fork, 0) != hipSuccess)
            return HY_E_DEVICE;
        hy::SysArgs sa{};
        sa.flat = db->d_flat;
        sa.off = db->d_off;
        sa.list = db->d_sys_list_batch;
        sa.m = (int)db->mql;
        sa.matrix = db->d_matrix;
        sa.al = (int)db->al;
        sa.al_eff = (int)db->al_eff;
        sa.go = db->go;
        sa.ge = db->ge;
        sa.spill = s->d_btail_spill;
        sa.spill_stride = db->sys_batch_max_len + 1;
        sa.key = s->d_bkeys;
        sa.queries = s->d_bq;
        sa.qoff = d_qoff;
        sa.qlen = d_qlen;
        sa.nq = d_nq;
        for (uint32_t k = 0; k < cap; ++k) {
            sa.qsel = (int)k;
            if (hy::launch_sys32((int)db->mode, false, false, sa,
                                 (uint32_t)db->sys_list_batch.size(), s->aux) != hipSuccess)
                return HY_E_DEVICE;
        }
        if (hipEventRecord(s->join, s->aux) != hipSuccess) return HY_E_DEVICE;
    }

    if (db->groups16.size() > db->groups16_batch_first) {
        hy::Inter16Args pa{};
        pa.groups = db->d_groups16;
        pa.group_first = db->groups16_batch_first;
        pa.ngroups = (uint32_t)db->groups16.size() - db->groups16_batch_first;
        pa.words = db->d_words16;
        pa.lens = db->d_lens16;
        pa.gidx = db->d_gidx16;
        pa.matrix = db->d_matrix;
        pa.al = (int)db->al;
        pa.al_eff = (int)db->al_eff;
        pa.go = db->go;
        pa.ge = db->ge;
        hy::BatchArgs b{};
        b.queries = s->d_bq;
        b.qoff = d_qoff;
        b.qlen = d_qlen;
        b.qorder = d_qorder;
        b.nq = d_nq;
        b.counter = d_counter;
        b.keys = s->d_bkeys;
        b.slots = s->d_slots;
        b.slot_stride = s->slot_stride;
        if (timed && hipEventRecord(s->bev[0], st) != hipSuccess) return HY_E_DEVICE;
        const hipError_t e =
            batch_query_pairs(db, s)
                ? hy::launch_inter16q_batch(pa, b, s->bq16, s->nslots, st)
                : hy::launch_inter16_batch((int)db->mode, pa, b, s->bq16, s->nslots, st);
        if (e != hipSuccess) return HY_E_DEVICE;
        if (timed && hipEventRecord(s->bev[1], st) != hipSuccess) return HY_E_DEVICE;
    }
    if (tail && hipStreamWaitEvent(st, s->join, 0) != hipSuccess) return HY_E_DEVICE;

    if (db->search_type != HY_SEARCH_SCORE) {
        hy::SysArgs sa{};
        sa.flat = db->d_flat;
        sa.off = db->d_off;
        sa.winner_key = s->d_bkeys;
        sa.m = (int)db->mql;
        sa.matrix = db->d_matrix;
        sa.al = (int)db->al;
        sa.al_eff = (int)db->al_eff;
        sa.go = db->go;
        sa.ge = db->ge;
        sa.spill = s->d_bwin_spill;
        sa.spill_stride = db->mtl + 1;
        sa.result = s->d_bres;
        sa.queries = s->d_bq;
        sa.qoff = d_qoff;
        sa.qlen = d_qlen;
        sa.nq = d_nq;
        sa.qsel = -1;
        if (hy::launch_sys32((int)db->mode, true, true, sa, 1, st, cap) != hipSuccess)
            return HY_E_DEVICE;
    } else {
        if (hy::launch_key_results(s->d_bkeys, s->d_bres, d_nq, cap, st) != hipSuccess)
            return HY_E_DEVICE;
    }
    if (hipMemcpyAsync(s->h_bres, s->d_bres, (size_t)cap * sizeof(hy::Result),
                       hipMemcpyDeviceToHost, st) != hipSuccess)
        return HY_E_DEVICE;
    return HY_OK;
}

}  // namespace

extern "C" int hy_scan_batch(const hy_db* db, hy_scratch* s, const uint8_t* const* queries,
                             const size_t* qlens, size_t nq, int32_t* scores,
                             uint64_t* db_indices, int64_t* query_ends, int64_t* target_ends) {
    if (!db || !s || (nq && (!queries || !qlens || !scores || !db_indices || !query_ends ||
                             !target_ends)))
        return HY_E_NULL_ARGUMENT;
    if (s->db != db) return HY_E_INVALID_ARGUMENT;
    for (size_t k = 0; k < nq; ++k) {
        const int rc = check_query(db, queries[k], qlens[k]);
        if (rc != HY_OK) return rc;
    }
    if (nq == 0) return HY_OK;
    int rc = HY_OK;
    const size_t want = std::min<size_t>(nq, s->bchunk);
    if (s->bcap < want && (rc = hy_scratch_reserve_batch(s, want)) != HY_OK) return rc;
    if (hipSetDevice(db->device) != hipSuccess) return HY_E_DEVICE;
    const uint32_t cap = s->bcap;
    const uint32_t per_sub = std::min<uint32_t>(cap, s->bchunk);
    std::vector<uint32_t> order;
    for (size_t base = 0; base < nq; base += per_sub) {
        const uint32_t cnt = (uint32_t)std::min<size_t>(per_sub, nq - base);
        // A submission of one query has nothing to share a launch with: the one-query path (with
        // its long-target split) is faster, and the query-pair kernel would idle half its lanes.
        if (cnt == 1) {
            if ((rc = hy_scan(db, s, queries[base], qlens[base], &scores[base], &db_indices[base],
                              &query_ends[base], &target_ends[base])) != HY_OK)
                return rc;
            continue;
        }
        uint32_t* qoff = s->h_bmeta;
        uint32_t* qlen = s->h_bmeta + cap;
        uint32_t* qord = s->h_bmeta + 2 * cap;
        for (uint32_t k = 0; k < cnt; ++k) {
            qoff[k] = k * s->bq_stride;
            qlen[k] = (uint32_t)qlens[base + k];
            if (qlen[k]) std::memcpy(s->h_bq + qoff[k], queries[base + k], qlen[k]);
        }
        order.resize(cnt);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return qlen[a] > qlen[b]; });
        for (uint32_t k = 0; k < cnt; ++k) qord[k] = order[k];
        s->h_bmeta[3 * cap] = cnt;
        // Profiling runs the sequence directly (with events); otherwise the first submission is
        // captured as a graph and every later one replays it.
        if (s->profile) {
            if ((rc = enqueue_batch(db, s, true)) != HY_OK) return rc;
        } else {
            if (!s->gexec) {
                // No other thread's allocation or release may fall inside the capture.
                std::lock_guard<std::recursive_mutex> api(hy::api_lock());
                if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                    return HY_E_DEVICE;
                rc = enqueue_batch(db, s, false);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(s->stream, &g);
                if (rc != HY_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (e != hipSuccess) return HY_E_DEVICE;
                s->graph = g;
                if (hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
                    s->gexec = nullptr;
                    return HY_E_DEVICE;
                }
            }
            if (hipGraphLaunch(s->gexec, s->stream) != hipSuccess) return HY_E_DEVICE;
        }
        if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
        if (s->profile && db->groups16.size() > db->groups16_batch_first) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, s->bev[0], s->bev[1]) == hipSuccess) {
                s->batch_ms += ms;
                s->batch_launches += 1;
            }
        }
        for (uint32_t k = 0; k < cnt; ++k) {
            const hy::Result& r = s->h_bres[k];
            if (r.check != 0) return HY_E_INTERNAL;
            scores[base + k] = r.score;
            db_indices[base + k] = r.db_index;
            query_ends[base + k] = r.query_end;
            target_ends[base + k] = r.target_end;
        }
    }
    return HY_OK;
}

extern "C" int hy_scratch_batch_profile_read(hy_scratch* s, double* kernel_ms, uint64_t* launches,
                                             int reset) {
    if (!s) return HY_E_NULL_ARGUMENT;
    if (kernel_ms) *kernel_ms = s->batch_ms;
    if (launches) *launches = s->batch_launches;
    if (reset) {
        s->batch_ms = 0.0;
        s->batch_launches = 0;
    }
    return HY_OK;
}

extern "C" int hy_scratch_get_tuning(const hy_scratch* s, hy_tuning* out) {
    if (!s || !out) return HY_E_NULL_ARGUMENT;
    out->tile_rows = s->q16;
    out->waves = s->waves16;
    out->batch_tile_rows = s->bq16;
    out->batch_wgs_per_cu = s->bwpc;
    out->batch_chunk = (int32_t)s->bchunk;
    out->batch_pairs = s->bpairs;
    return HY_OK;
}

extern "C" int hy_scratch_set_tuning(hy_scratch* s, const hy_tuning* t) {
    if (!s || !t) return HY_E_NULL_ARGUMENT;
    const hy_db* db = s->db;
    auto rows_ok = [](int q) { return q == 16 || q == 24 || q == 32; };
    if (!rows_ok(t->tile_rows) || !rows_ok(t->batch_tile_rows) ||
        !(t->batch_pairs == 0 || t->batch_pairs == 1) ||
        !(t->waves == 1 || t->waves == 4 || t->waves == 8) || t->batch_wgs_per_cu < 0 ||
        t->batch_wgs_per_cu > 64 || t->batch_chunk < 1 || t->batch_chunk > 65536)
        return HY_E_INVALID_ARGUMENT;
    // Shared-profile workgroups need the residue-pair profile to fit in LDS.
    if (t->waves > 1 &&
        hy::inter16_lds_bytes((int)db->al_eff, t->tile_rows) > hy::kPairProfileMaxBytes)
        return HY_E_INVALID_ARGUMENT;
    if (hipSetDevice(s->device) != hipSuccess) return HY_E_DEVICE;
    if (t->batch_tile_rows != s->bq16 || t->batch_wgs_per_cu != s->bwpc ||
        t->batch_pairs != s->bpairs) {
        std::lock_guard<std::recursive_mutex> api(hy::api_lock());
        if (hipStreamSynchronize(s->stream) != hipSuccess) return HY_E_DEVICE;
        release_batch(s);
    }
    s->q16 = t->tile_rows;
    s->waves16 = t->waves;
    s->bq16 = t->batch_tile_rows;
    s->bwpc = t->batch_wgs_per_cu;
    s->bchunk = (uint32_t)t->batch_chunk;
    s->bpairs = t->batch_pairs;
    return HY_OK;
}

extern "C" int hy_scratch_profile(hy_scratch* s, int enable) {
    if (!s) return HY_E_NULL_ARGUMENT;
    if (hipSetDevice(s->device) != hipSuccess) return HY_E_DEVICE;
    if (enable) {
        for (auto& e : s->ev)
            if (!e && hipEventCreate(&e) != hipSuccess) return HY_E_DEVICE;
    }
    s->profile = enable != 0;
    s->pending = false;
    return HY_OK;
}

extern "C" int hy_scratch_profile_read(hy_scratch* s, double* inter_ms, double* tail_ms,
                                       uint64_t* launches, int reset) {
    if (!s) return HY_E_NULL_ARGUMENT;
    if (inter_ms) *inter_ms = s->inter_ms;
    if (tail_ms) *tail_ms = s->tail_ms;
    if (launches) *launches = s->launches;
    if (reset) {
        s->inter_ms = s->tail_ms = 0.0;
        s->launches = 0;
    }
    return HY_OK;
}

extern "C" int hy_scratch_last_kernel(const hy_scratch* s, int32_t* kernel) {
    if (!s || !kernel) return HY_E_NULL_ARGUMENT;
    *kernel = s->last_kernel;
    return HY_OK;
}

extern "C" const char* hy_strerror(int code) {
    switch (code) {
    case HY_OK: return "ok";
    case HY_E_EMPTY_ALPHABET: return "alphabet length must be greater than zero";
    case HY_E_MATRIX_SHAPE: return "substitution matrix must have alphabet_len squared entries";
    case HY_E_NEGATIVE_GAP_PENALTY: return "gap penalties must be non-negative magnitudes";
    case HY_E_GAP_OPEN_LESS_THAN_EXTEND: return "gap_open must be >= gap_ext (Opal issue #28)";
    case HY_E_SCORE_RANGE_TOO_WIDE:
        return "reachable score magnitude bound exceeds the i32 range; reduce sequence lengths "
               "or penalty magnitudes";
    case HY_E_SYMBOL_OUT_OF_RANGE: return "encoded symbol is out of range for alphabet_len";
    case HY_E_INCOMPLETE_BUILDER: return "database builder is missing a required field";
    case HY_E_EMPTY_DATABASE: return "database must contain at least one sequence";
    case HY_E_BACKEND_UNAVAILABLE: return "backend hip is not available on this build/device";
    case HY_E_INVALID_BACKEND_NAME:
        return "unrecognised backend name; expected one of: auto, scalar, sse4.1, avx2, neon, hip";
    case HY_E_TRACEBACK_BUDGET: return "traceback budget exceeded";
    case HY_E_QUERY_TOO_LONG: return "query length exceeds the declared max_query_len";
    case HY_E_QUERY_SYMBOL: return "query symbol is out of range for alphabet_len";
    case HY_E_NULL_ARGUMENT: return "a required pointer argument was null";
    case HY_E_INVALID_ARGUMENT: return "invalid argument";
    case HY_E_DEVICE: return "a HIP runtime call failed";
    case HY_E_OUT_OF_MEMORY: return "out of memory";
    case HY_E_INTERNAL: return "device self-check failed";
    case HY_E_TIMEOUT: return "timed out waiting for a served query";
    case HY_E_BAD_IMAGE: return "not a database image, or a damaged one";
    case HY_E_STATIC_LIMIT:
        return "the problem needs more device memory than the static scratch array holds";
    case HY_E_NOT_ALIGNMENT:
        return "scan_aligned/scan_all_aligned require a database built with SearchType::Alignment";
    default: return "unknown status code";
    }
}
