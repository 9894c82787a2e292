// Large copies between pageable host memory and the device through pinned slots.
//
// The runtime stages a copy from or to pageable memory through pinned memory on one thread,
// which keeps the DMA engine idle about half the time (130 MB up in 6.5 ms, 3.1 ms of it
// transferring). Here worker threads each own two pinned slots of PinnedRing::kSlot bytes and
// take every kThreads-th piece of the buffer: on the way up a thread copies a piece into a free
// slot and queues its DMA; on the way down it queues the DMA of its next piece before copying
// the current one out, so host copies run in parallel and overlap the transfers.
#include <emmintrin.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/hyalite_hip.h"
#include "hy_internal.h"

namespace hy {

uint8_t host_max_byte(const uint8_t* p, size_t n) {
    __m128i m = _mm_setzero_si128();
    size_t i = 0;
    for (; i + 16 <= n; i += 16)
        m = _mm_max_epu8(m, _mm_loadu_si128(reinterpret_cast<const __m128i*>(p + i)));
    alignas(16) uint8_t lanes[16];
    _mm_store_si128(reinterpret_cast<__m128i*>(lanes), m);
    uint8_t mx = 0;
    for (uint8_t v : lanes) mx = std::max(mx, v);
    for (; i < n; ++i) mx = std::max(mx, p[i]);
    return mx;
}

uint8_t copy_max_byte(uint8_t* dst, const uint8_t* src, size_t n) {
    __m128i m = _mm_setzero_si128();
    size_t i = 0;
    for (; i + 16 <= n; i += 16) {
        const __m128i v = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + i));
        _mm_storeu_si128(reinterpret_cast<__m128i*>(dst + i), v);
        m = _mm_max_epu8(m, v);
    }
    alignas(16) uint8_t lanes[16];
    _mm_store_si128(reinterpret_cast<__m128i*>(lanes), m);
    uint8_t mx = 0;
    for (uint8_t v : lanes) mx = std::max(mx, v);
    for (; i < n; ++i) {
        dst[i] = src[i];
        mx = std::max(mx, src[i]);
    }
    return mx;
}

PinnedRing::~PinnedRing() {
    if (ring) (void)hipHostFree(ring);
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
}

namespace {

int ring_ready(PinnedRing& r) {
    if (r.ring) return HY_OK;
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    if (hipHostMalloc((void**)&r.ring, 2 * PinnedRing::kThreads * PinnedRing::kSlot) !=
        hipSuccess) {
        r.ring = nullptr;
        return HY_E_OUT_OF_MEMORY;
    }
    for (hipEvent_t& e : r.ev)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return HY_E_DEVICE;
    return HY_OK;
}

// Runs work(t) on min(kThreads, pieces) threads, each with `device` current; the first
// failing status, or HY_OK.
template <class F>
int run_workers(int device, size_t pieces, F work) {
    const int workers = (int)std::min<size_t>(PinnedRing::kThreads, pieces);
    std::vector<int> status(workers, HY_OK);
    std::vector<std::thread> pool;
    pool.reserve(workers);
    int spawn = HY_OK;
    try {
        for (int t = 0; t < workers; ++t)
            pool.emplace_back([&, t]() {
                // HIP's current device is per thread.
                status[t] = hipSetDevice(device) == hipSuccess ? work(t) : HY_E_DEVICE;
            });
    } catch (...) {  // no thread to run on: no exception crosses the C ABI
        spawn = HY_E_OUT_OF_MEMORY;
    }
    for (std::thread& th : pool) th.join();
    if (spawn != HY_OK) return spawn;
    for (int st : status)
        if (st != HY_OK) return st;
    return HY_OK;
}

}  // namespace

UploadJob::~UploadJob() {
    for (std::thread& th : pool_)
        if (th.joinable()) th.join();
}

int UploadJob::start(PinnedRing& r, int device, const UpSegment* segs, int nseg,
                     hipStream_t stream) {
    pieces_.clear();
    for (int k = 0; k < nseg; ++k)
        for (size_t off = 0; off < segs[k].bytes; off += PinnedRing::kSlot)
            pieces_.push_back({static_cast<uint8_t*>(segs[k].dst) + off,
                               static_cast<const uint8_t*>(segs[k].src) + off,
                               std::min(PinnedRing::kSlot, segs[k].bytes - off),
                               segs[k].symbols});
    if (pieces_.empty()) return HY_OK;
    const int rc = ring_ready(r);
    if (rc != HY_OK) return rc;
    const int workers = (int)std::min<size_t>(PinnedRing::kThreads, pieces_.size());
    status_.assign(workers, HY_OK);
    wmax_.assign(workers, 0);
    pool_.reserve(workers);
    try {
    for (int t = 0; t < workers; ++t)
        pool_.emplace_back([this, &r, device, stream, t]() {
            // HIP's current device is per thread.
            if (hipSetDevice(device) != hipSuccess) {
                status_[t] = HY_E_DEVICE;
                return;
            }
            size_t used = 0;
            for (size_t p = (size_t)t; p < pieces_.size(); p += PinnedRing::kThreads, ++used) {
                const Piece& pc = pieces_[p];
                const int slot = 2 * t + (int)(used & 1);
                uint8_t* h = r.ring + (size_t)slot * PinnedRing::kSlot;
                // The slot's previous piece, from this job or an earlier one, must have left
                // before the slot is refilled.
                if (r.busy[slot]) {
                    if (hipEventSynchronize(r.ev[slot]) != hipSuccess) {
                        status_[t] = HY_E_DEVICE;
                        return;
                    }
                    r.busy[slot] = false;
                }
                if (pc.symbols) wmax_[t] = std::max(wmax_[t], copy_max_byte(h, pc.src, pc.len));
                else std::memcpy(h, pc.src, pc.len);
                if (hipMemcpyAsync(pc.dst, h, pc.len, hipMemcpyHostToDevice, stream) !=
                        hipSuccess ||
                    hipEventRecord(r.ev[slot], stream) != hipSuccess) {
                    status_[t] = HY_E_DEVICE;
                    return;
                }
                r.busy[slot] = true;
            }
        });
    } catch (...) {  // no thread to run on: the started ones finish, nothing crosses the C ABI
        for (std::thread& th : pool_) th.join();
        pool_.clear();
        return HY_E_OUT_OF_MEMORY;
    }
    return HY_OK;
}

int UploadJob::finish(uint8_t* max_byte) {
    for (std::thread& th : pool_) th.join();
    pool_.clear();
    if (max_byte) *max_byte = 0;
    for (int st : status_)
        if (st != HY_OK) return st;
    if (max_byte)
        for (uint8_t v : wmax_) *max_byte = std::max(*max_byte, v);
    return HY_OK;
}

int upload_pageable(PinnedRing& r, int device, void* dst, const void* src, size_t bytes,
                    hipStream_t stream, uint8_t* max_byte) {
    UploadJob job;
    const UpSegment seg{dst, src, bytes, max_byte != nullptr};
    const int rc = job.start(r, device, &seg, 1, stream);
    const int fin = job.finish(max_byte);
    return rc != HY_OK ? rc : fin;
}

int download_pageable(PinnedRing& r, int device, const Segment* segs, int nseg,
                      hipStream_t stream) {
    // Pieces of kSlot bytes over all segments, in order.
    struct Piece {
        uint8_t* dst;
        const uint8_t* src;
        size_t len;
    };
    constexpr int T = PinnedRing::kThreads;
    constexpr size_t SLOT = PinnedRing::kSlot;
    std::vector<Piece> pieces;
    for (int k = 0; k < nseg; ++k)
        for (size_t off = 0; off < segs[k].bytes; off += SLOT)
            pieces.push_back({static_cast<uint8_t*>(segs[k].dst) + off,
                              static_cast<const uint8_t*>(segs[k].src) + off,
                              std::min(SLOT, segs[k].bytes - off)});
    if (pieces.empty()) return HY_OK;
    int rc = ring_ready(r);
    if (rc != HY_OK) return rc;
    const size_t np = pieces.size();
    return run_workers(device, np, [&](int t) -> int {
        auto issue = [&](size_t p, int slot) -> bool {
            // An upload's DMA still reading the slot precedes this one on the stream only if
            // it was queued on the same stream; wait for it regardless.
            if (r.busy[slot] && hipEventSynchronize(r.ev[slot]) != hipSuccess) return false;
            r.busy[slot] = true;
            return hipMemcpyAsync(r.ring + (size_t)slot * SLOT, pieces[p].src, pieces[p].len,
                                  hipMemcpyDeviceToHost, stream) == hipSuccess &&
                   hipEventRecord(r.ev[slot], stream) == hipSuccess;
        };
        if ((size_t)t < np && !issue((size_t)t, 2 * t)) return HY_E_DEVICE;
        size_t i = 0;
        for (size_t p = (size_t)t; p < np; p += T, ++i) {
            const int slot = 2 * t + (int)(i & 1);
            // The next piece goes into the other slot, whose piece was copied out last round.
            if (p + T < np && !issue(p + T, 2 * t + (int)((i + 1) & 1))) return HY_E_DEVICE;
            if (hipEventSynchronize(r.ev[slot]) != hipSuccess) return HY_E_DEVICE;
            r.busy[slot] = false;
            std::memcpy(pieces[p].dst, r.ring + (size_t)slot * SLOT, pieces[p].len);
        }
        return HY_OK;
    });
}

int download(PinnedRing& r, int device, const Segment* segs, int nseg, hipStream_t stream) {
    size_t total = 0;
    for (int k = 0; k < nseg; ++k) total += segs[k].bytes;
    if (total >= kPinnedRingMin) return download_pageable(r, device, segs, nseg, stream);
    if (total == 0) return HY_OK;
    // Smaller results: one DMA per segment into the ring's first two slots (kPinnedRingMin bytes,
    // contiguous) and host copies out, instead of the runtime's pageable copies. Those were slow
    // in some states of the process heap: 200 KB of scan_all results took 0.51 ms per call
    // instead of 0.20 after a phase that allocated and freed gigabytes, until the allocator's
    // trim and mmap thresholds were fixed.
    static_assert(kPinnedRingMin <= 2 * PinnedRing::kSlot, "two slots hold a small download");
    int rc = ring_ready(r);
    if (rc != HY_OK) return rc;
    for (int slot = 0; slot < 2; ++slot) {
        if (r.busy[slot] && hipEventSynchronize(r.ev[slot]) != hipSuccess) return HY_E_DEVICE;
        r.busy[slot] = false;
    }
    size_t off = 0;
    for (int k = 0; k < nseg; ++k) {
        if (segs[k].bytes && hipMemcpyAsync(r.ring + off, segs[k].src, segs[k].bytes,
                                            hipMemcpyDeviceToHost, stream) != hipSuccess)
            return HY_E_DEVICE;
        off += segs[k].bytes;
    }
    r.busy[0] = r.busy[1] = true;
    if (hipEventRecord(r.ev[0], stream) != hipSuccess ||
        hipEventRecord(r.ev[1], stream) != hipSuccess ||
        hipEventSynchronize(r.ev[0]) != hipSuccess)
        return HY_E_DEVICE;
    r.busy[0] = r.busy[1] = false;
    off = 0;
    for (int k = 0; k < nseg; ++k) {
        if (segs[k].bytes) std::memcpy(segs[k].dst, r.ring + off, segs[k].bytes);
        off += segs[k].bytes;
    }
    return HY_OK;
}

}  // namespace hy
