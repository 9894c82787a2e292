// Mode-independent kernels (argmax keys, results, fills) and the dispatch from a runtime
// alignment mode to the per-mode launchers that hy_kernels.hip exports (one object per mode).
#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "hy_device.h"

namespace hy {
namespace {

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int32_t* p, uint64_t n, int32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Four consecutive bytes p[c .. c + 3] as one little-endian word from two aligned dword loads.
// The caller guarantees c + 4 <= the sequence length; the second load may read up to 7 bytes
// past p[c], which the pair buffers' 8 bytes of slack keep inside the allocation.
__device__ __forceinline__ uint32_t load4_unaligned(const uint8_t* p, uint64_t c) {
    const uintptr_t addr = reinterpret_cast<uintptr_t>(p + c);
    const uint32_t* w = reinterpret_cast<const uint32_t*>(addr & ~(uintptr_t)3);
    return __builtin_amdgcn_alignbyte(w[1], w[0], (uint32_t)(addr & 3));
}

// One workgroup per pair group: lane l gathers four symbols of its pair per word (the padding
// symbol past the pair's length, and for lanes past the group's count) and the group's words
// are written lane-interleaved, so the pair kernel's per-column load is one coalesced line.
__global__ __launch_bounds__(64) void k_pair_pack(PairPackArgs a) {
    const PairGroup g = a.groups[blockIdx.x];
    const int lane = threadIdx.x;
    const bool live = lane < (int)g.count;
    const uint32_t slot = g.first + (live ? lane : 0);
    const uint32_t pair = a.pidx[slot];
    const uint32_t pad = a.pad_symbol & 0xffu;
    const uint64_t tn = live ? a.tlens[slot] : 0, qn = live ? a.qlens[slot] : 0;
    const uint8_t* t = a.tflat + a.toff[pair];
    const uint8_t* q = a.qflat + a.qoff[pair];
    const uint32_t tw = (g.ncols + 3) / 4;
    for (uint32_t w = 0; w < tw; ++w) {
        uint32_t v = 0;
        if ((uint64_t)w * 4 + 4 <= tn) {
            v = load4_unaligned(t, (uint64_t)w * 4);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t c = (uint64_t)w * 4 + k;
                v |= (c < tn ? (uint32_t)t[c] : pad) << (8 * k);
            }
        }
        a.twords[g.tword_off + (uint64_t)w * 64 + lane] = v;
    }
    const uint32_t qt = (uint32_t)a.qtile;
    const uint32_t qw = (g.nrows + qt - 1) / qt * qt / 4;
    for (uint32_t w = 0; w < qw; ++w) {
        uint32_t v = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t r = (uint64_t)w * 4 + k;
            v |= (r < qn ? (uint32_t)q[r] : pad) << (8 * k);
        }
        a.qwords[g.qword_off + (uint64_t)w * 64 + lane] = v;
    }
}

// The two-pairs-per-lane layout of k_pairs16: lane l packs slots 2l (low) and 2l + 1 (high) of
// its group; a target word holds four columns as bytes ta * S + tb, a query word two rows as
// 16-bit byte offsets 4 * (qa * S + qb) * S^2 (S = al + 1, symbol al pads; symbols past the
// alphabet become padding, as in k_pair_pack).
// Four threads per lane (256 per group), each writing every fourth word of the lane.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_pair_pack16(PairPackArgs a) {
    const PairGroup g = a.groups[blockIdx.x];
    const int lane = threadIdx.x & 63, part = threadIdx.x >> 6;
    const uint32_t pad = a.pad_symbol, S = pad + 1;
    const uint32_t sa = 2 * lane, sb = 2 * lane + 1;
    const bool la = sa < g.count, lb = sb < g.count;
    const uint32_t slota = g.first + (la ? sa : 0), slotb = g.first + (lb ? sb : 0);
    const uint32_t pa = a.pidx[slota], pb = a.pidx[slotb];
    const uint64_t tna = la ? a.tlens[slota] : 0, tnb = lb ? a.tlens[slotb] : 0;
    const uint64_t qna = la ? a.qlens[slota] : 0, qnb = lb ? a.qlens[slotb] : 0;
    const uint8_t* ta = a.tflat + a.toff[pa];
    const uint8_t* tb = a.tflat + a.toff[pb];
    const uint8_t* qa = a.qflat + a.qoff[pa];
    const uint8_t* qb = a.qflat + a.qoff[pb];
    const uint32_t tw = (g.ncols + 3) / 4;
    for (uint32_t w = part; w < tw; w += 4) {
        const uint64_t c0 = (uint64_t)w * 4;
        // Four symbols of each target: whole words where both lie inside their sequences.
        uint32_t xa = 0, xb = 0;
        if (c0 + 4 <= tna) {
            xa = load4_unaligned(ta, c0);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) xa |= (c0 + k < tna ? (uint32_t)ta[c0 + k] : pad) << (8 * k);
        }
        if (c0 + 4 <= tnb) {
            xb = load4_unaligned(tb, c0);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) xb |= (c0 + k < tnb ? (uint32_t)tb[c0 + k] : pad) << (8 * k);
        }
        if constexpr (WIDE) {
            // Column c: bytes ta, tb.
            uint32_t v[2] = {0u, 0u};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t x = min((xa >> (8 * k)) & 0xffu, pad);
                const uint32_t y = min((xb >> (8 * k)) & 0xffu, pad);
                v[k >> 1] |= (x | y << 8) << (16 * (k & 1));
            }
            reinterpret_cast<uint2*>(a.twords + g.tword_off)[(uint64_t)w * 64 + lane] =
                make_uint2(v[0], v[1]);
        } else {
            uint32_t v = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t x = min((xa >> (8 * k)) & 0xffu, pad);
                const uint32_t y = min((xb >> (8 * k)) & 0xffu, pad);
                v |= (x * S + y) << (8 * k);
            }
            a.twords[g.tword_off + (uint64_t)w * 64 + lane] = v;
        }
    }
    const uint32_t qt = (uint32_t)a.qtile;
    const uint32_t row_bytes = 4 * kPairs16Copies * S;  // WIDE table rows
    constexpr uint32_t rows_per_word = WIDE ? 1 : 2;
    const uint32_t qw = (g.nrows + qt - 1) / qt * qt / rows_per_word;
    for (uint32_t w = part; w < qw; w += 4) {
        uint32_t v = 0;
#pragma unroll
        for (uint32_t k = 0; k < rows_per_word; ++k) {
            const uint64_t r = (uint64_t)w * rows_per_word + k;
            const uint32_t x = r < qna ? min((uint32_t)qa[r], pad) : pad;
            const uint32_t y = r < qnb ? min((uint32_t)qb[r], pad) : pad;
            // Byte offsets of the rows: of a pair-table row (two rows per word), or of each
            // half's row of the score table (WIDE, one row per word).
            v |= WIDE ? (row_bytes * x) | (row_bytes * y) << 16
                      : (4 * (x * S + y) * S * S) << (16 * k);
        }
        a.qwords[g.qword_off + (uint64_t)w * 64 + lane] = v;
    }
}

// Database layouts built on the device from the residues in database order: one workgroup per
// group, lane l gathering its target's four residues per word (the padding symbol past its
// length, and for lanes past the group's count).
__global__ __launch_bounds__(64) void k_db_pack32(const Group* groups, const uint32_t* gidx,
                                                  const uint64_t* off, const uint8_t* flat,
                                                  uint32_t* words, uint32_t pad) {
    const Group g = groups[blockIdx.x];
    const int lane = threadIdx.x;
    const bool live = lane < (int)g.count;
    const uint32_t k = live ? gidx[g.first + lane] : 0u;
    const uint64_t len = live ? off[k + 1] - off[k] : 0;
    const uint8_t* t = flat + off[k];
    const uint32_t nw = (g.ncols + 3) / 4;
    for (uint32_t w = 0; w < nw; ++w) {
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint64_t c = (uint64_t)w * 4 + b;
            v |= (c < len ? (uint32_t)t[c] : pad) << (8 * b);
        }
        words[g.word_off + (uint64_t)w * 64 + lane] = v;
    }
}

// The packed layout: lane l holds slots 2l (low word) and 2l + 1 (high word) of its group.
__global__ __launch_bounds__(64) void k_db_pack16(const Group* groups, const uint32_t* gidx,
                                                  const uint64_t* off, const uint8_t* flat,
                                                  uint2* words, uint32_t pad) {
    const Group g = groups[blockIdx.x];
    const int lane = threadIdx.x;
    uint64_t len[2];
    const uint8_t* t[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t slot = 2 * lane + h;
        const bool live = slot < g.count;
        const uint32_t k = live ? gidx[g.first + slot] : 0u;
        len[h] = live ? off[k + 1] - off[k] : 0;
        t[h] = flat + off[k];
    }
    const uint32_t nw = (g.ncols + 3) / 4;
    for (uint32_t w = 0; w < nw; ++w) {
        uint32_t v[2] = {0, 0};
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint64_t c = (uint64_t)w * 4 + b;
                v[h] |= (c < len[h] ? (uint32_t)t[h][c] : pad) << (8 * b);
            }
        words[g.word_off + (uint64_t)w * 64 + lane] = make_uint2(v[0], v[1]);
    }
}

__global__ void k_batch_reset(unsigned long long* keys, uint32_t n, unsigned int* counter) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = 0ull;
    if (i == 0) *counter = 0u;
}

__global__ void k_key_results(const unsigned long long* keys, Result* r, const uint32_t* nq,
                              uint32_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap || i >= *nq) return;
    const unsigned long long k = keys[i];
    r[i].score = key_score(k);
    r[i].db_index = key_index(k);
    r[i].query_end = -1;
    r[i].target_end = -1;
    r[i].check = k == 0ull ? 1 : 0;
    r[i].pad = 0;
}

__global__ void k_key_result(const unsigned long long* key, Result* r) {
    if (threadIdx.x == 0) {
        const unsigned long long k = *key;
        r->score = key_score(k);
        r->db_index = key_index(k);
        r->query_end = -1;
        r->target_end = -1;
        r->check = k == 0ull ? 1 : 0;
        r->pad = 0;
    }
}

// Op packing (scan_all_aligned): each walk leaves target k's ops at the back of its region of
// the op buffer; these kernels lay them end to end in database order. Block b covers targets
// b * kOpsBlock .. + kOpsBlock - 1, four consecutive targets per thread.
static_assert(kOpsBlock == 4 * 256, "four targets per thread, 256 threads");

// Inclusive sum over the 256 threads of a block (x is this thread's value); tmp: 256 entries.
__device__ __forceinline__ uint64_t block_scan256(uint64_t x, uint64_t* tmp) {
    const int t = threadIdx.x;
    tmp[t] = x;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const uint64_t y = t >= d ? tmp[t - d] : 0ull;
        __syncthreads();
        tmp[t] += y;
        __syncthreads();
    }
    const uint64_t r = tmp[t];
    __syncthreads();  // tmp is free again for the caller's next use
    return r;
}

__global__ __launch_bounds__(256) void k_ops_blocksum(const TraceOut* out, uint64_t n,
                                                      uint64_t* bsum) {
    __shared__ uint64_t tmp[256];
    const uint64_t i0 = (uint64_t)blockIdx.x * kOpsBlock + 4u * threadIdx.x;
    uint64_t s = 0;
    for (int k = 0; k < 4; ++k)
        if (i0 + k < n) s += out[i0 + k].n_ops;
    const uint64_t incl = block_scan256(s, tmp);
    if (threadIdx.x == 255) bsum[blockIdx.x] = incl;
}

// One block: bsum[0 .. nb) becomes its exclusive prefix sums, bsum[nb] the total.
__global__ __launch_bounds__(256) void k_ops_scan_blocks(uint64_t* bsum, uint64_t nb) {
    __shared__ uint64_t tmp[256];
    uint64_t carry = 0;
    for (uint64_t c = 0; c < nb; c += 256) {
        const uint64_t i = c + threadIdx.x;
        const uint64_t x = i < nb ? bsum[i] : 0ull;
        const uint64_t incl = block_scan256(x, tmp);
        if (i < nb) bsum[i] = carry + incl - x;
        // Thread 255's inclusive sum is the chunk's total: every thread reads it from tmp.
        if (threadIdx.x == 255) tmp[0] = incl;
        __syncthreads();
        carry += tmp[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

// Targets' new offsets (block base bsum[b] plus the block's own prefix), written to
// out[k].ops_offset after its source offset is read, then the block's output range
// [noff[0], noff[cnt]) copied byte by byte: consecutive threads take consecutive output bytes,
// so loads and stores stay coalesced within a target, and each thread walks its target index
// forward as its byte advances.
__global__ __launch_bounds__(256) void k_ops_compact(TraceOut* out, uint64_t n,
                                                     const uint64_t* bsum, const uint8_t* src,
                                                     uint8_t* dst) {
    __shared__ uint64_t tmp[256];
    __shared__ uint64_t noff[kOpsBlock + 1];
    __shared__ uint64_t soff[kOpsBlock];
    const uint64_t b0 = (uint64_t)blockIdx.x * kOpsBlock;
    const int cnt = (int)min<uint64_t>((uint64_t)kOpsBlock, n - b0);
    const int t = threadIdx.x;
    uint64_t len[4], src_off[4], s = 0;
    for (int k = 0; k < 4; ++k) {
        const int j = 4 * t + k;
        len[k] = j < cnt ? out[b0 + j].n_ops : 0ull;
        src_off[k] = j < cnt ? out[b0 + j].ops_offset : 0ull;
        s += len[k];
    }
    uint64_t o = bsum[blockIdx.x] + block_scan256(s, tmp) - s;
    for (int k = 0; k < 4; ++k) {
        const int j = 4 * t + k;
        if (j < cnt) {
            noff[j] = o;
            soff[j] = src_off[k];
            out[b0 + j].ops_offset = o;
        }
        o += len[k];
    }
    // The thread holding the block's last target (cnt >= 1) ends the range.
    if (t == (cnt - 1) / 4) noff[cnt] = o;
    __syncthreads();
    const uint64_t lo = noff[0], hi = noff[cnt];
    int k = 0;
    for (uint64_t x = lo + (uint64_t)t; x < hi; x += 256) {
        while (k + 1 < cnt && noff[k + 1] <= x) ++k;
        dst[x] = src[soff[k] + (x - noff[k])];
    }
}

}  // namespace

hipError_t launch_inter16_batch(int mode, const Inter16Args& a, const BatchArgs& b, int q,
                                uint32_t workgroups, hipStream_t stream) {
    if (a.ngroups == 0 || workgroups == 0) return hipSuccess;
    switch (mode) {
    case kSW: return launch_inter16_batch_m<kSW>(a, b, q, workgroups, stream);
    case kNW: return launch_inter16_batch_m<kNW>(a, b, q, workgroups, stream);
    case kHW: return launch_inter16_batch_m<kHW>(a, b, q, workgroups, stream);
    case kOV: return launch_inter16_batch_m<kOV>(a, b, q, workgroups, stream);
    case kSHW: return launch_inter16_batch_m<kSHW>(a, b, q, workgroups, stream);
    default: return hipErrorInvalidValue;
    }
}

int inter16_tile_rows(int q) { return q <= 16 ? 16 : (q <= 24 ? 24 : 32); }

uint32_t inter16p_lds_bytes(int al_eff, int q) {
    const uint32_t rows = (uint32_t)inter16_tile_rows(q) / 2;
    const uint32_t rd = rows % 8 == 4 ? rows : rows + 4;
    return ((uint32_t)al_eff + 1) * rd * 4u;
}

uint32_t inter16_lds_bytes(int al_eff, int q) {
    const uint32_t A = (uint32_t)al_eff + 1;
    return A * A * (uint32_t)(inter16_tile_rows(q) + 4) * 4u;
}

hipError_t launch_inter16_pipe(int mode, const Inter16Args& a, int q, bool ends,
                               hipStream_t stream) {
    if (a.ngroups == 0) return hipSuccess;
    switch (mode) {
    case kSW: return launch_inter16_pipe_m<kSW>(a, q, ends, stream);
    case kNW: return launch_inter16_pipe_m<kNW>(a, q, ends, stream);
    case kHW: return launch_inter16_pipe_m<kHW>(a, q, ends, stream);
    case kOV: return launch_inter16_pipe_m<kOV>(a, q, ends, stream);
    case kSHW: return launch_inter16_pipe_m<kSHW>(a, q, ends, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_inter16_chain(int mode, const Inter16Args& a, const ChainArgs& c, int q,
                                int waves, bool ends, hipStream_t stream) {
    if (a.ngroups == 0) return hipSuccess;
    switch (mode) {
    case kSW: return launch_inter16_chain_m<kSW>(a, c, q, waves, ends, stream);
    case kNW: return launch_inter16_chain_m<kNW>(a, c, q, waves, ends, stream);
    case kHW: return launch_inter16_chain_m<kHW>(a, c, q, waves, ends, stream);
    case kOV: return launch_inter16_chain_m<kOV>(a, c, q, waves, ends, stream);
    case kSHW: return launch_inter16_chain_m<kSHW>(a, c, q, waves, ends, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_inter16_ends(int mode, const Inter16Args& a, hipStream_t stream) {
    if (a.ngroups == 0) return hipSuccess;
    switch (mode) {
    case kSW: return launch_inter16_ends_m<kSW>(a, stream);
    case kNW: return launch_inter16_ends_m<kNW>(a, stream);
    case kHW: return launch_inter16_ends_m<kHW>(a, stream);
    case kOV: return launch_inter16_ends_m<kOV>(a, stream);
    case kSHW: return launch_inter16_ends_m<kSHW>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_inter16(int mode, const Inter16Args& a, int q, int waves, hipStream_t stream) {
    if (a.ngroups == 0) return hipSuccess;
    switch (mode) {
    case kSW: return launch_inter16_m<kSW>(a, q, waves, stream);
    case kNW: return launch_inter16_m<kNW>(a, q, waves, stream);
    case kHW: return launch_inter16_m<kHW>(a, q, waves, stream);
    case kOV: return launch_inter16_m<kOV>(a, q, waves, stream);
    case kSHW: return launch_inter16_m<kSHW>(a, q, waves, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v,
                           hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

hipError_t launch_db_pack32(const Group* groups, uint32_t ngroups, const uint32_t* gidx,
                           const uint64_t* off, const uint8_t* flat, uint32_t* words,
                           uint32_t pad, hipStream_t stream) {
    if (ngroups == 0) return hipSuccess;
    hipLaunchKernelGGL(k_db_pack32, dim3(ngroups), dim3(64), 0, stream, groups, gidx, off, flat,
                       words, pad);
    return hipGetLastError();
}

hipError_t launch_db_pack16(const Group* groups, uint32_t ngroups, const uint32_t* gidx,
                           const uint64_t* off, const uint8_t* flat, uint2* words, uint32_t pad,
                           hipStream_t stream) {
    if (ngroups == 0) return hipSuccess;
    hipLaunchKernelGGL(k_db_pack16, dim3(ngroups), dim3(64), 0, stream, groups, gidx, off, flat,
                       words, pad);
    return hipGetLastError();
}


hipError_t launch_fill_i32(int32_t* p, uint64_t n, int32_t v, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

hipError_t launch_pair_pack(const PairPackArgs& a, uint32_t ngroups, hipStream_t stream) {
    if (ngroups == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_pack, dim3(ngroups), dim3(64), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_pair_pack16(const PairPackArgs& a, uint32_t ngroups, hipStream_t stream) {
    if (ngroups == 0) return hipSuccess;
    if (a.pad_symbol < 1 || a.pad_symbol > kPairs16MaxAlphabet) return hipErrorInvalidValue;
    if (pairs16_wide((int)a.pad_symbol))
        hipLaunchKernelGGL(k_pair_pack16<true>, dim3(ngroups), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(k_pair_pack16<false>, dim3(ngroups), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_pairs32(int mode, bool ends, const PairArgs& a, uint32_t ngroups, int q,
                          hipStream_t stream) {
    if (ngroups == 0) return hipSuccess;
    switch (mode) {
    case kSW: return launch_pairs32_m<kSW>(ends, a, ngroups, q, stream);
    case kNW: return launch_pairs32_m<kNW>(ends, a, ngroups, q, stream);
    case kHW: return launch_pairs32_m<kHW>(ends, a, ngroups, q, stream);
    case kOV: return launch_pairs32_m<kOV>(ends, a, ngroups, q, stream);
    case kSHW: return launch_pairs32_m<kSHW>(ends, a, ngroups, q, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_band32(int mode, int out, const BandArgs& a, uint32_t nblocks, int m,
                         hipStream_t stream) {
    if (nblocks == 0) return hipSuccess;
    if (out == kBandColmax && (mode != kSW || nblocks != 1)) return hipErrorInvalidValue;
    switch (mode) {
    case kSW: return launch_band32_m<kSW>(out, a, nblocks, m, stream);
    case kNW: return launch_band32_m<kNW>(out, a, nblocks, m, stream);
    case kHW: return launch_band32_m<kHW>(out, a, nblocks, m, stream);
    case kOV: return launch_band32_m<kOV>(out, a, nblocks, m, stream);
    case kSHW: return launch_band32_m<kSHW>(out, a, nblocks, m, stream);
    default: return hipErrorInvalidValue;
    }
}

namespace {
bool band_shape_known(int r, int w) {
    switch (r * 100 + w) {
    case 101: case 102: case 104: case 201: case 202: case 204: case 208: case 216:
    case 401: case 402: case 404: case 408: case 416:
        return true;
    default:
        return false;
    }
}
}  // namespace

BandShape band_shape(int m) {
    if (const char* v = std::getenv("HYALITE_HIP_BAND_SHAPE")) {
        int r = 0, w = 0;
        if (std::sscanf(v, "%d,%d", &r, &w) == 2 && band_shape_known(r, w)) return {r, w};
    }
    // Measured on single DNA pairs (SW; 30 x 30 to 3,000 x 10,000): up to 256 rows one
    // wavefront of up to four rows per lane is fastest (a lone wavefront is latency-bound, but
    // the band pipeline's lag costs more on short targets); longer queries split into bands of
    // 128-256 rows so that two to four wavefronts share each SIMD.
    if (m <= 64) return {1, 1};
    if (m <= 128) return {2, 1};
    if (m <= 256) return {4, 1};
    if (m <= 512) return {2, 4};
    if (m <= 1024) return {2, 8};
    if (m <= 2048) return {4, 8};
    return {4, 16};
}

hipError_t launch_posmax(const PosMaxArgs& a, hipStream_t stream) {
    const hipError_t e = launch_fill_i32(a.colmax, (uint64_t)a.n, 0, stream);
    if (e != hipSuccess) return e;
    return launch_posmax_sw(a, stream);
}

hipError_t launch_batch_reset(unsigned long long* keys, uint32_t n, unsigned int* counter,
                              hipStream_t stream) {
    const uint32_t blocks = (n + 255) / 256 + (n == 0);
    hipLaunchKernelGGL(k_batch_reset, dim3(blocks), dim3(256), 0, stream, keys, n, counter);
    return hipGetLastError();
}

hipError_t launch_key_results(const unsigned long long* keys, Result* r, const uint32_t* nq,
                              uint32_t cap, hipStream_t stream) {
    if (cap == 0) return hipSuccess;
    hipLaunchKernelGGL(k_key_results, dim3((cap + 255) / 256), dim3(256), 0, stream, keys, r, nq,
                       cap);
    return hipGetLastError();
}

uint32_t device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        return 0;
    return (uint32_t)cus;
}

uint32_t inter16_batch_occupancy(int mode, int q, int al_eff) {
    switch (mode) {
    case kSW: return inter16_batch_occupancy_m<kSW>(q, al_eff);
    case kNW: return inter16_batch_occupancy_m<kNW>(q, al_eff);
    case kHW: return inter16_batch_occupancy_m<kHW>(q, al_eff);
    case kOV: return inter16_batch_occupancy_m<kOV>(q, al_eff);
    case kSHW: return inter16_batch_occupancy_m<kSHW>(q, al_eff);
    default: return 2;
    }
}

hipError_t launch_key_result(const unsigned long long* key, Result* r, hipStream_t stream) {
    hipLaunchKernelGGL(k_key_result, dim3(1), dim3(64), 0, stream, key, r);
    return hipGetLastError();
}

hipError_t launch_ops_sums(const TraceOut* out, uint64_t n, uint64_t* bsum, hipStream_t stream) {
    const uint64_t nb = ops_blocks(n);
    if (nb > 0x7FFFFFFFull) return hipErrorInvalidValue;
    if (nb) {
        hipLaunchKernelGGL(k_ops_blocksum, dim3((uint32_t)nb), dim3(256), 0, stream, out, n, bsum);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_ops_scan_blocks, dim3(1), dim3(256), 0, stream, bsum, nb);
    return hipGetLastError();
}

hipError_t launch_ops_pack(TraceOut* out, uint64_t n, const uint64_t* bsum, const uint8_t* src,
                           uint8_t* dst, hipStream_t stream) {
    const uint64_t nb = ops_blocks(n);
    if (nb == 0) return hipSuccess;
    if (nb > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ops_compact, dim3((uint32_t)nb), dim3(256), 0, stream, out, n, bsum, src,
                       dst);
    return hipGetLastError();
}

hipError_t launch_inter32(int mode, bool ends, const InterArgs& a, uint32_t ngroups,
                          hipStream_t stream) {
    if (ngroups == 0) return hipSuccess;
    switch (mode) {
    case kSW: return launch_inter32_m<kSW>(ends, a, ngroups, stream);
    case kNW: return launch_inter32_m<kNW>(ends, a, ngroups, stream);
    case kHW: return launch_inter32_m<kHW>(ends, a, ngroups, stream);
    case kOV: return launch_inter32_m<kOV>(ends, a, ngroups, stream);
    case kSHW: return launch_inter32_m<kSHW>(ends, a, ngroups, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_trace(int mode, const TraceArgs& a, hipStream_t stream) {
    switch (mode) {
    case kSW: return launch_trace_m<kSW>(a, stream);
    case kNW: return launch_trace_m<kNW>(a, stream);
    case kHW: return launch_trace_m<kHW>(a, stream);
    case kOV: return launch_trace_m<kOV>(a, stream);
    case kSHW: return launch_trace_m<kSHW>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_trace16(int mode, const TraceArgs& a, hipStream_t stream) {
    switch (mode) {
    case kSW: return launch_trace16_m<kSW>(a, stream);
    case kNW: return launch_trace16_m<kNW>(a, stream);
    case kHW: return launch_trace16_m<kHW>(a, stream);
    case kOV: return launch_trace16_m<kOV>(a, stream);
    case kSHW: return launch_trace16_m<kSHW>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sys32(int mode, bool ends, bool winner, const SysArgs& a, uint32_t nblocks,
                        hipStream_t stream, uint32_t nq_grid) {
    if ((!winner && nblocks == 0) || nq_grid == 0) return hipSuccess;
    switch (mode) {
    case kSW: return launch_sys32_m<kSW>(ends, winner, a, nblocks, stream, nq_grid);
    case kNW: return launch_sys32_m<kNW>(ends, winner, a, nblocks, stream, nq_grid);
    case kHW: return launch_sys32_m<kHW>(ends, winner, a, nblocks, stream, nq_grid);
    case kOV: return launch_sys32_m<kOV>(ends, winner, a, nblocks, stream, nq_grid);
    case kSHW: return launch_sys32_m<kSHW>(ends, winner, a, nblocks, stream, nq_grid);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace hy
