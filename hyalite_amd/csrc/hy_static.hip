// hy_scan_scores_static: hy_scan_scores for callers that cannot allocate device memory (code
// under graph capture, tight serving loops, harnesses that forbid hipMalloc). It takes no
// hy_scratch: the query copy, the tile-boundary rows of the inter-sequence kernels, the systolic
// boundary rows of queries beyond one pass and the score array live in one file-scope device
// array of kStaticScratchBytes, and the kernels run on the null stream. The call makes no
// hipMalloc, hipFree or other allocating call; a problem whose working set exceeds the array is
// refused with HY_E_STATIC_LIMIT before anything is queued.
//
// Every call on a device shares that device's array, which keeps the last call's contents:
// nothing in it has to start at zero (every boundary row is written by one query tile before the
// next tile reads it, every score is written once), but two calls that overlap on one device
// race on it, so callers serialise them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/hyalite_hip.h"
#include "hy_internal.h"
#include "hy_state.h"

namespace {

constexpr uint64_t kStaticScratchBytes = 64ull << 20;
constexpr int kMaxDevices = 64;

// One copy per device; the host-side name is only a stand-in for the symbol, so every host use
// goes through the address hipGetSymbolAddress returns for the current device.
__device__ __align__(256) uint2 scratch[kStaticScratchBytes / sizeof(uint2)];

std::mutex g_mu;
uint8_t* g_ptr[kMaxDevices] = {};

// The array's address on `device` (the current device), looked up once per device.
int scratch_on(int device, uint8_t** out) {
    if (device < 0 || device >= kMaxDevices) return HY_E_DEVICE;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HY_E_DEVICE;
        g_ptr[device] = static_cast<uint8_t*>(p);
    }
    *out = g_ptr[device];
    return HY_OK;
}

uint64_t align256(uint64_t b) { return (b + 255) & ~255ull; }

}  // namespace

extern "C" int hy_scan_scores_static(const hy_db* db, const uint8_t* query, size_t qlen,
                                     int32_t* scores) {
    if (!db || !scores) return HY_E_NULL_ARGUMENT;
    int rc = hy::check_query(db, query, qlen);
    if (rc != HY_OK) return rc;
    const int m = (int)qlen;
    // Layout of the array for this call: query | boundary rows | systolic rows | scores.
    const bool has16 = db->groups16.size() > db->groups16_short_first;
    const uint64_t q_off = 0;
    const uint64_t spill_off = q_off + align256(std::max<uint64_t>(qlen, 1));
    const uint64_t spill_entries = db->inter16_spill + db->score32_spill;
    const uint64_t sys_off = spill_off + align256(std::max<uint64_t>(spill_entries, 1) * 8);
    const bool multipass = m > 512 && !db->sys_list.empty();
    const uint64_t sys_stride = db->sys_max_len + 1;
    const uint64_t sys_bytes = multipass ? db->sys_list.size() * sys_stride * 8 : 8;
    const uint64_t score_off = sys_off + align256(sys_bytes);
    const uint64_t total = score_off + align256(std::max<uint64_t>(db->n, 1) * 4);
    if (total > kStaticScratchBytes) return HY_E_STATIC_LIMIT;

    if (hipSetDevice(db->device) != hipSuccess) return HY_E_DEVICE;
    uint8_t* base = nullptr;
    if ((rc = scratch_on(db->device, &base)) != HY_OK) return rc;
    uint8_t* d_query = base + q_off;
    int2* d_spill = reinterpret_cast<int2*>(base + spill_off);
    int2* d_sys = reinterpret_cast<int2*>(base + sys_off);
    int32_t* d_scores = reinterpret_cast<int32_t*>(base + score_off);
    const hipStream_t st = nullptr;
    if (qlen && hipMemcpyAsync(d_query, query, qlen, hipMemcpyHostToDevice, st) != hipSuccess)
        return HY_E_DEVICE;

    // The score-only pass of hy_scan_scores on one stream: the 32-bit groups (targets outside
    // the packed width proof) with boundary rows after the packed kernel's, the long targets in
    // the systolic kernel, then the packed kernel over the rest.
    hy::InterArgs ia{};
    ia.groups = db->d_groups;
    ia.words = db->d_words;
    ia.lens = db->d_lens;
    ia.gidx = db->d_gidx;
    ia.query = d_query;
    ia.m = m;
    ia.matrix = db->d_matrix;
    ia.al = (int)db->al;
    ia.al_eff = (int)db->al_eff;
    ia.go = db->go;
    ia.ge = db->ge;
    ia.spill = d_spill + (has16 ? db->inter16_spill : 0);
    ia.out_score = d_scores;
    if (hy::launch_inter32((int)db->mode, false, ia, db->groups_score32, st) != hipSuccess)
        return HY_E_DEVICE;
    if (!db->sys_list.empty()) {
        hy::SysArgs sa{};
        sa.flat = db->d_flat;
        sa.off = db->d_off;
        sa.list = db->d_sys_list;
        sa.query = d_query;
        sa.m = m;
        sa.matrix = db->d_matrix;
        sa.al = (int)db->al;
        sa.al_eff = (int)db->al_eff;
        sa.go = db->go;
        sa.ge = db->ge;
        sa.spill = d_sys;
        sa.spill_stride = sys_stride;
        sa.out_score = d_scores;
        sa.qsel = -1;
        if (hy::launch_sys32((int)db->mode, false, false, sa, (uint32_t)db->sys_list.size(),
                             st) != hipSuccess)
            return HY_E_DEVICE;
    }
    if (has16) {
        hy::Inter16Args pa{};
        pa.groups = db->d_groups16;
        pa.group_first = db->groups16_short_first;
        pa.ngroups = (uint32_t)db->groups16.size() - db->groups16_short_first;
        pa.words = db->d_words16;
        pa.lens = db->d_lens16;
        pa.gidx = db->d_gidx16;
        pa.query = d_query;
        pa.m = m;
        pa.matrix = db->d_matrix;
        pa.al = (int)db->al;
        pa.al_eff = (int)db->al_eff;
        pa.go = db->go;
        pa.ge = db->ge;
        pa.spill = reinterpret_cast<uint2*>(d_spill);
        pa.out_score = d_scores;
        if (hy::launch_inter16((int)db->mode, pa, db->q16, db->waves16, st) != hipSuccess)
            return HY_E_DEVICE;
    }
    if (hipMemcpyAsync(scores, d_scores, (size_t)db->n * sizeof(int32_t), hipMemcpyDeviceToHost,
                       st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return HY_E_DEVICE;
    return HY_OK;
}
