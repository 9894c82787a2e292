// Database-scan kernels for CDNA4 (gfx950).
//
// Two families compute the affine-gap (Gotoh) recurrence of hyalite's scalar oracle
// (src/kernel.rs:168-217) with its answer regions and tie-break (src/kernel.rs:113-129, 223-260):
//
//  * k_inter32 — inter-sequence (Rognes) kernel: one target per lane, a tile of Q query rows held
//    in VGPRs, the query profile for the tile staged in LDS, the target columns streamed from a
//    lane-interleaved HBM layout (one 32-bit word carries four columns of a lane). The tile's
//    bottom row is spilled to a per-group boundary buffer and read back by the next tile. It
//    replaces scan_batch / scan_batch_ends (src/inter.rs:683-997).
//
//  * k_sys32 — intra-sequence systolic kernel: one (query, target) pair per wavefront, each lane
//    owning R consecutive query rows; lane l works on column s-l+1 at step s and receives the row
//    above from lane l-1 through a DPP wave shift. It serves long targets and recovers the end
//    positions of the single-best winner (the role align_core plays in src/database.rs:183-197).
//
// Arithmetic is the 32-bit realisation of DETERMINISM.md §1: sentinel i32::MIN/4, wrapping
// add/sub, exact for every input the width proof accepts.
//
// Every kernel and launcher here is a template over the alignment mode; this file is compiled
// once per mode (-DHY_KMODE=0..4, the numbering of hy::kSW..kSHW) and each object exports that
// mode's launchers. hy_launch.hip maps a runtime mode onto them.
#include "hy_device.h"

#ifndef HY_KMODE
#error "hy_kernels.hip is compiled once per alignment mode with -DHY_KMODE=0..4"
#endif

namespace hy {
namespace {

// ------------------------------------------------------------------------------------------------
// Inter-sequence kernel, 32-bit lanes.
// ------------------------------------------------------------------------------------------------
//
// Per cell (i, j) with HG = H - gap_open kept per row (so H - go is computed once and feeds both
// E of the next column and F of the next row) and the profile holding sub + go:
//     diag = HG[i-1][j-1] + (sub + go)       = H[i-1][j-1] + sub
//     E    = max(HG[i][j-1], E - ge)
//     F    = max(HG[i-1][j], F - ge)
//     H    = max(diag, E, F)   (local: max(H, 0))
// The padding symbol (al_eff) and padded query rows use kPad32, so in local mode padded cells
// never exceed a real cell that precedes them in (col, row) order; in the other modes the answer
// cells are selected explicitly.
template <int MODE, int Q, bool ENDS>
__global__ __launch_bounds__(64) void k_inter32(InterArgs a) {
    using F = MF<MODE>;
    extern __shared__ int prof[];
    constexpr int S = Q + 4;  // row stride: consecutive symbols start 4 banks apart
    const int lane = threadIdx.x;
    const Group g = a.groups[blockIdx.x];
    const bool live = lane < (int)g.count;
    const uint32_t slot = g.first + (live ? lane : 0);
    const int n = live ? (int)a.lens[slot] : 0;
    const int L = (int)g.ncols;
    const int m = a.m, go = a.go, ge = a.ge;

    int bs, bj, bi;
    border_best<MODE>(m, n, go, ge, bs, bj, bi);

    const int nchunks = (m + Q - 1) / Q;
    int2* sp = a.spill + g.spill_off + lane;
    const uint32_t* wp = a.words + g.word_off + lane;

    for (int c = 0; c < nchunks; ++c) {
        const int i0 = c * Q + 1;
        const bool lastc = c == nchunks - 1;
        const int rm = m - i0;  // register of row m in the last tile
        __syncthreads();
        for (int k = lane; k < (a.al_eff + 1) * Q; k += 64) {
            const int sym = k / Q, r = k - sym * Q, i = i0 + r;
            int v = kPad32;
            if (sym < a.al_eff && i <= m) v = a.matrix[(int)a.query[i - 1] * a.al + sym] + go;
            prof[sym * S + r] = v;
        }
        __syncthreads();

        int HG[Q], E[Q];
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            HG[r] = wsub(border<MODE>(i0 + r, 0, go, ge), go);
            E[r] = kNeg32;
        }
        int HGd = wsub(border<MODE>(i0 - 1, 0, go, ge), go);
        int cbs = INT_MIN, cbj = 0, cbi = 0;  // tile-local best (local mode, ends)

        // Column words and the previous tile's boundary values are fetched one four-column
        // block ahead, so their HBM latency overlaps the current block's compute.
        uint32_t w = L > 0 ? wp[0] : 0u;
        int2 spv[4];
#pragma unroll
        for (int b = 0; b < 4; ++b)
            spv[b] = (c > 0 && b + 1 <= L) ? sp[(size_t)(b + 1) * 64] : make_int2(0, 0);
        for (int jb = 0; jb < L; jb += 4) {
            uint32_t wn = 0u;
            int2 spn[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) spn[b] = make_int2(0, 0);
            if (jb + 4 < L) {
                wn = wp[(size_t)((jb >> 2) + 1) * 64];
                if (c > 0) {
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        if (jb + 5 + b <= L) spn[b] = sp[(size_t)(jb + 5 + b) * 64];
                }
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = jb + b + 1;
                if (j > L) break;
                int hgup, f;
                if (c == 0) {
                    hgup = wsub(border<MODE>(0, j, go, ge), go);
                    f = kNeg32;
                } else {
                    hgup = spv[b].x;
                    f = spv[b].y;
                }
                const int sym = (int)((w >> (8 * b)) & 0xffu);
                const int4* pr = reinterpret_cast<const int4*>(&prof[sym * S]);
                int sg[Q];
#pragma unroll
                for (int r4 = 0; r4 < Q / 4; ++r4) {
                    const int4 v = pr[r4];
                    sg[4 * r4 + 0] = v.x; sg[4 * r4 + 1] = v.y;
                    sg[4 * r4 + 2] = v.z; sg[4 * r4 + 3] = v.w;
                }
                int hgd = HGd;
                HGd = hgup;
                int cmax = INT_MIN;
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    const int diag = wadd(hgd, sg[r]);
                    const int e = max(HG[r], wsub(E[r], ge));
                    f = max(hgup, wsub(f, ge));
                    int h = max(max(diag, e), f);
                    if (F::local) h = max(h, 0);
                    hgd = HG[r];
                    E[r] = e;
                    HG[r] = wsub(h, go);
                    hgup = HG[r];
                    if (F::local) {
                        if (ENDS) {
                            if (h > cbs) { cbs = h; cbj = j; cbi = i0 + r; }
                        } else {
                            cmax = max(cmax, h);
                        }
                    }
                }
                if (F::local && !ENDS) bs = max(bs, cmax);
                if (c + 1 < nchunks) sp[(size_t)j * 64] = make_int2(HG[Q - 1], f);
                if (!F::local && lastc && (F::lrow || j == n) && j <= n) {
                    int v = HG[0];
#pragma unroll
                    for (int r = 1; r < Q; ++r)
                        if (r == rm) v = HG[r];
                    v = wadd(v, go);
                    if (ENDS) consider(bs, bj, bi, v, j, m);
                    else bs = max(bs, v);
                }
                if (F::lcol && j == n) {
#pragma unroll
                    for (int r = 0; r < Q; ++r) {
                        if (i0 + r <= m) {
                            const int v = wadd(HG[r], go);
                            if (ENDS) consider(bs, bj, bi, v, n, i0 + r);
                            else bs = max(bs, v);
                        }
                    }
                }
            }
            w = wn;
#pragma unroll
            for (int b = 0; b < 4; ++b) spv[b] = spn[b];
        }
        if (F::local && ENDS && cbs != INT_MIN) consider(bs, bj, bi, cbs, cbj, cbi);
    }

    const uint32_t gi = a.gidx[slot];
    if (ENDS) {
        if (live) {
            a.out_score[gi] = bs;
            a.out_qend[gi] = bi == 0 ? -1 : (int64_t)bi - 1;
            a.out_tend[gi] = bj == 0 ? -1 : (int64_t)bj - 1;
        }
    } else {
        if (live && a.out_score) a.out_score[gi] = bs;
        if (a.key) wave_argmax(live ? make_key(bs, gi) : 0ull, a.key);
    }
}

// ------------------------------------------------------------------------------------------------
// Inter-sequence kernel, packed 16-bit lanes: two targets per lane.
// ------------------------------------------------------------------------------------------------
//
// The same recurrence as k_inter32 on pairs of 16-bit values with saturating packed arithmetic
// (v_pk_add_i16/v_pk_sub_i16/v_pk_sub_u16 with clamp, v_pk_max_i16): the i8/i16 realisation of
// DETERMINISM.md §1, sentinel = i16::MIN. A lane's two targets read different residues; the
// profile of a tile holds sub(q_r, s) + go (local mode: sub(q_r, s)) saturated to 16 bits for
// every symbol s < A = al_eff + 1 (the padding symbol's entries are i16::MIN), either per symbol
// (paired per lane with v_perm_b32) or per residue pair (PAIR: pp[(a * A + b) * S + r]).
//
// Exactness: the host sends a target here only when its width proof gives |cell| <= bound with
// bound + gap_open <= 32767, so H - gap_open never saturates on a real cell and
// HG[i-1][j-1] + (sub + go) equals H[i-1][j-1] + sub whenever the sum is in range (a saturated
// diagonal is below every real cell and loses the max). Padding cells are bounded by saturation
// and never flow into real cells; in local mode they never exceed a real cell.


template <int MODE>
__device__ __forceinline__ int border16(int i, int j, int go, int ge) {
    if (i == 0) return (j == 0 || MF<MODE>::top) ? 0 : sat16(-(long long)gap32(go, ge, j));
    return MF<MODE>::left ? 0 : sat16(-(long long)gap32(go, ge, i));
}

// PAIR = true: the residue-pair profile above, shared by NW wavefronts. PAIR = false: a plain
// per-symbol profile of the tile (two query rows per 32-bit word), private to a one-wavefront
// workgroup; each lane reads its two targets' rows and pairs them with one v_perm_b32 per query
// row. The second form needs ~1 KB of LDS instead of ~70 KB (protein), so occupancy is bounded
// by registers only and no barrier couples wavefronts.
// Register budget of the packed kernels: at least two resident wavefronts per SIMD (at most 256
// VGPRs), measured faster than one wavefront with more registers for the 32-row local tile.
#ifndef HY_PACKED_WAVES_PER_EU
#define HY_PACKED_WAVES_PER_EU 2
#endif
// Query rows per tile of the packed kernel's end-tracking form in the non-local modes.
#ifndef HY_ENDS_TILE_NONLOCAL
#define HY_ENDS_TILE_NONLOCAL 32
#endif
// Running state of one wavefront's two-targets-per-lane pass over one group, carried across the
// query tiles.
struct PairRun {
    s2v run;
    // Local mode keeps four independent running maxima (by row mod 4) so the tracker is not one
    // long chain of dependent packed maxima.
    s2v runq[4];
    // End tracking (local mode): each target's best (score, column, row) so far, merged tile by
    // tile with the answer tie-break.
    int bsA, bjA, biA, bsB, bjB, biB;
};

// Pipelined tiles (k_inter16_pipe): the wavefronts of a workgroup hold consecutive query tiles
// of one group, wavefront `wave` running one four-column block behind wavefront wave - 1. A tile's
// bottom row goes to the next wavefront through an LDS double buffer (block b in half b & 1,
// entry (half * 4 + column in block) * 64 + lane) instead of HBM, and every wavefront takes the
// same nsteps workgroup barriers, one per block step, active or not.
//
// Chained workgroups (k_inter16_chain) extend the pipeline across CUs: the first wavefront of a
// workgroup after the first of its group reads its top rows from the group's boundary buffer
// (column j at gin[j * 64]), which the previous workgroup's last wavefront writes (gout). The
// writer publishes a count of completed four-column blocks in a 64-bit progress word,
// (launch epoch << 32) + count; the reader polls it, one poll per block issued before the block
// and consumed after it, and blocks only when the rows of the block after next are not yet
// published. Counts of earlier launches carry smaller epochs, so the words need no reset between
// launches. Every access to the rows and the word is agent-coherent (sc1: it bypasses the
// per-XCD L2, whose agent-scope release would write back every dirty line of the L2 at each
// publication); the writer waits for its row stores to complete (vmcnt(0)) before it stores
// the count, and the reader issues its row loads only after a poll returned that count.
// Workgroup-scope fences keep the compiler from moving those accesses across the protocol.
struct Pipe {
    int wave = 0, nsteps = 0;
    const uint2* rin = nullptr;  // the previous wavefront's bottom rows (not the first tile)
    uint2* rout = nullptr;       // this wavefront's bottom rows (not the last tile)
    const uint2* gin = nullptr;  // chained: the previous workgroup's bottom rows
    uint2* gout = nullptr;       // chained: this workgroup's bottom rows for the next one
    const unsigned long long* pin = nullptr;  // progress of gin
    unsigned long long* pout = nullptr;       // progress of gout
    unsigned long long epoch = 0;             // launch epoch << 32
    int slack = 2;                            // blocks published before the reader starts
    int every = 1;                            // the writer publishes every `every` blocks
    unsigned int* err = nullptr;              // host-visible flag: a wait gave up
    uint32_t spin_limit = kChainSpinLimit;    // polls before a wait gives up
};

// Agent-coherent row accesses of the chained link.
__device__ __forceinline__ uint2 chain_load(const uint2* p) {
    const unsigned long long v =
        __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
    return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
}
__device__ __forceinline__ void chain_store(uint2* p, uint2 v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                       (unsigned long long)v.x | ((unsigned long long)v.y << 32),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The count of this launch in a progress word value (0 for a value of an earlier launch).
__device__ __forceinline__ uint32_t chain_count(const Pipe& p, unsigned long long v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const unsigned long long u = ((unsigned long long)hi << 32) | lo;
    return u >= p.epoch ? (uint32_t)(u - p.epoch) : 0u;
}

// Wait until the progress word shows at least `need` blocks of this launch; returns the count
// seen. On giving up, raises the host-visible flag (the launch's results are then reported as an
// internal error) and returns `need`.
__device__ __forceinline__ uint32_t chain_wait(const Pipe& p, const unsigned long long* word,
                                               uint32_t need) {
    for (uint32_t spins = 0;; ++spins) {
        const uint32_t n = chain_count(
            p, __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (n >= need) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            return n;
        }
        if (spins >= p.spin_limit) {
            if ((threadIdx.x & 63) == 0)
                __hip_atomic_store(p.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            return need;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Publish `count` completed blocks: the wavefront's earlier row stores complete first.
__device__ __forceinline__ void chain_publish(const Pipe& p, unsigned long long* word,
                                              uint32_t count) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): every vector memory access of the wave
    if ((threadIdx.x & 63) == 0)
        __hip_atomic_store(word, p.epoch + count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One tile of QT query rows starting at row i0 over every column of the group: the profile of
// those rows is staged in LDS, the rows' H/E columns live in registers, and the tile's bottom row
// goes to `sp` (column j of this wavefront at sp[j * 64]) for the next tile unless this is the last.
// LINKIN: the chained reader (pipelined tiles only): the tile's top rows come from the previous
// workgroup through pipe.gin. A role fixed at compile time, so that the block code of the other
// wavefronts has no global-memory top rows to wait for.
template <int MODE, int QT, int NW, bool PAIR, bool FIRST, bool ENDS = false, bool PIPE = false,
          bool LINKIN = false>
__device__ __forceinline__ void inter16_tile(const Inter16Args& a, const Group& g, bool wlive,
                                             const uint8_t* query, int m, int i0, bool lastc,
                                             int nA, int nB, uint2* sp, const uint2* wp,
                                             uint32_t* pp, PairRun& st, Pipe pipe = Pipe()) {
    using F = MF<MODE>;
    static_assert(!LINKIN || (PIPE && !FIRST), "a chained reader is a pipelined later tile");
    constexpr int S = QT + 4;
    // Single profile: row stride in 32-bit words, an odd multiple of 4 so that the rows of
    // different symbols start in different 4-bank groups.
    constexpr int RD = ((QT / 2) % 8 == 4) ? QT / 2 : QT / 2 + 4;
    const int L = (int)g.ncols;
    const int go = a.go, ge = a.ge;
    const int A = a.al_eff + 1;
    const int rm = m - i0;
    const s2v go2 = splat2(go), ge2 = splat2(ge);
    // Local mode keeps H itself per row and the profile holds sub; the other modes keep
    // HG = H - go and the profile holds sub + go.
    const int gadd = F::local ? 0 : go;
    // Pipelined tiles: each wavefront stages its own tile's profile (its own LDS region).
    const int tid = PIPE ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    const int tstride = PIPE ? 64 : 64 * NW;
    __syncthreads();
    if (PAIR) {
        for (int k = tid; k < A * A * QT; k += tstride) {
            const int pr = k / QT, r = k - pr * QT, i = i0 + r;
            const int ca = pr / A, cb = pr - ca * A;
            int va = -32768, vb = -32768;
            if (i <= m) {
                const int* row = a.matrix + (int)query[i - 1] * a.al;
                if (ca < a.al_eff) va = sat16((long long)row[ca] + gadd);
                if (cb < a.al_eff) vb = sat16((long long)row[cb] + gadd);
            }
            pp[pr * S + r] =
                (uint32_t)(uint16_t)(short)va | ((uint32_t)(uint16_t)(short)vb << 16);
        }
    } else {
        for (int k = tid; k < A * (QT / 2); k += tstride) {
            const int sym = k / (QT / 2), rr = k - sym * (QT / 2);
            const int i = i0 + 2 * rr;  // rows i (low half) and i + 1 (high half)
            int v0 = -32768, v1 = -32768;
            if (sym < a.al_eff) {
                if (i <= m) v0 = sat16((long long)a.matrix[(int)query[i - 1] * a.al + sym] + gadd);
                if (i + 1 <= m) v1 = sat16((long long)a.matrix[(int)query[i] * a.al + sym] + gadd);
            }
            pp[sym * RD + rr] = (uint32_t)(uint16_t)(short)v0 | ((uint32_t)(uint16_t)(short)v1 << 16);
        }
    }
    __syncthreads();
    if (!wlive) return;

    // Per row: HG = H - go (local mode: H) of the last column, and E of the next column (local
    // mode: E floored at 0). Column 0 is the DP border.
    s2v HG[QT], E[QT];
#pragma unroll
    for (int r = 0; r < QT; ++r) {
        HG[r] = F::local ? splat2(0)
                         : splat2(sat16((long long)border16<MODE>(i0 + r, 0, go, ge) - go));
        E[r] = splat2(F::local ? 0 : -32768);
    }
    s2v HGd = F::local ? splat2(0)
                       : splat2(sat16((long long)border16<MODE>(i0 - 1, 0, go, ge) - go));
    s2v run = st.run;
    s2v runq[4] = {st.runq[0], st.runq[1], st.runq[2], st.runq[3]};
    // End tracking: the tile's best score per target, its first column and, in that column, its
    // first row (0-based in the tile). Cells are met column by column in increasing order and, in
    // a column, row by row, which is the answer tie-break's order, so strictly greater is the only
    // update at every level.
    s2v tb = splat2(0), tc = splat2(0), tr = splat2(0);
    const s2v one2 = splat2(1);

    // Column words and boundary inputs stream one four-column block ahead: the loads for block
    // k+1 are issued before block k's compute and pinned there by a scheduling barrier (the
    // register pressure of the tile would otherwise make the scheduler sink them next to their
    // use), so their latency overlaps a whole block. The first tile takes its top inputs from the
    // DP border instead. The host pads every group's column count to a multiple of four with the
    // inert padding symbol, so a block has no exits and the packed registers stay packed across
    // it; the prefetch after the last block re-reads that block's inputs (in bounds).
    auto fetch = [&](int jb, uint2& wq, uint2* top) HY_INLINE {
        wq = wp[(size_t)(jb >> 2) * 64];
        if (!FIRST && !PIPE) {
#pragma unroll
            for (int b = 0; b < 4; ++b) top[b] = sp[(size_t)(jb + b + 1) * 64];
        }
    };
    auto block = [&](int jb, const uint2 w, const uint2* topin) HY_INLINE {
        // Top inputs of the block's columns: row 0 of the DP (local mode: (H, F' of row 1) =
        // (0, 0); otherwise (H - go, F) of row 0) or the previous tile's boundary row.
        uint2 spv[4];
#pragma unroll
        for (int b = 0; b < 4; ++b)
            spv[b] = !FIRST ? topin[b]
                     : F::local
                         ? make_uint2(0u, 0u)
                         : make_uint2(as_u32(splat2(sat16(
                                          (long long)border16<MODE>(0, jb + b + 1, go, ge) - go))),
                                      0x80008000u);
        uint2 outb[4];
        // Substitution pairs of column b of the block: (sub(q_r, low target), sub(q_r, high
        // target)) + go for the tile's rows.
        auto load_sg = [&](int b, s2v* sg) HY_INLINE {
            const int ca = (int)((w.x >> (8 * b)) & 0xffu);
            const int cb = (int)((w.y >> (8 * b)) & 0xffu);
            if (PAIR) {
                const uint4* prow = reinterpret_cast<const uint4*>(&pp[(ca * A + cb) * S]);
#pragma unroll
                for (int r4 = 0; r4 < QT / 4; ++r4) {
                    const uint4 v = prow[r4];
                    sg[4 * r4 + 0] = as_s2(v.x); sg[4 * r4 + 1] = as_s2(v.y);
                    sg[4 * r4 + 2] = as_s2(v.z); sg[4 * r4 + 3] = as_s2(v.w);
                }
            } else {
                const uint4* ra = reinterpret_cast<const uint4*>(&pp[ca * RD]);
                const uint4* rb = reinterpret_cast<const uint4*>(&pp[cb * RD]);
#pragma unroll
                for (int r8 = 0; r8 < QT / 8; ++r8) {
                    const uint4 va = ra[r8], vb = rb[r8];
                    const uint32_t wa[4] = {va.x, va.y, va.z, va.w};
                    const uint32_t wb[4] = {vb.x, vb.y, vb.z, vb.w};
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        // (A_r, B_r) and (A_r+1, B_r+1) from (A_r, A_r+1) and (B_r, B_r+1).
                        sg[8 * r8 + 2 * k] =
                            as_s2(__builtin_amdgcn_perm(wb[k], wa[k], 0x05040100u));
                        sg[8 * r8 + 2 * k + 1] =
                            as_s2(__builtin_amdgcn_perm(wb[k], wa[k], 0x07060302u));
                    }
                }
            }
        };
        // Local mode. Every quantity is kept floored at 0: with H >= 0, E and F only matter
        // through max(H, 0), and max(a - b, 0) is one unsigned saturating subtraction, so
        //     H = max(H[i-1][j-1] + sub, E', F')        (>= 0: the local clamp comes free)
        //     E'[i][j+1] = max(H - go, E' - ge)+,  F'[i+1][j] = max(H - go, F' - ge)+
        // equals the reference's recurrence for every cell that can win (H is exact; E', F' are
        // max(E, 0), max(F, 0)). 9 packed operations per cell pair including the running max.
        // Columns are processed in pairs skewed by one row — row r of column j and row r-1 of
        // column j+1 are independent — so two dependency chains interleave. The boundary row of a
        // tile is (H of its last row, F' of the row below).
        if (F::local) {
            // With ENDS the running maximum becomes a per-column one that also keeps the first
            // row reaching it: cm = max(cm, h); cr = r where h > cm (rows only grow within a
            // column, so the latest strict improvement is the first row reaching the maximum).
            // The mask is one clamped v_pk_mad_u16 and the select one v_bfi_b32: four operations
            // per cell pair on top of the cell's eight.
            auto cell = [&](int r, s2v sgv, s2v& hd, s2v& f, s2v& cm, s2v& cr) HY_INLINE {
                const s2v h = p_max(p_max(p_add(hd, sgv), E[r]), f);
                hd = HG[r];
                HG[r] = h;
                const s2v hg = p_subu(h, go2);
                E[r] = p_max(hg, p_subu(E[r], ge2));
                f = p_max(hg, p_subu(f, ge2));
                if constexpr (ENDS) {
                    const uint32_t up = p_nonzero_mask(p_subu(h, cm));
                    cm = p_max(cm, h);
                    cr = as_s2((up & as_u32(splat2(r))) | (~up & as_u32(cr)));
                } else {
                    runq[r & 3] = p_max(runq[r & 3], h);
                }
            };
            // A column's (max, first row) folded into the tile's best: strictly greater only.
            auto close_column = [&](int j, s2v cm, s2v cr) HY_INLINE {
                const s2v up = p_minu_pk(p_subu(cm, tb), one2);
                tb = p_max(tb, cm);
                tc = p_maxu(tc, p_madu(up, splat2(j), splat2(0)));
                tr = p_madu(up, p_subw(cr, tr), tr);
            };
#pragma unroll
            for (int bp = 0; bp < 4; bp += 2) {
                s2v sgA[QT], sgB[QT];
                load_sg(bp, sgA);
                load_sg(bp + 1, sgB);
                s2v fA = as_s2(spv[bp].y), fB = as_s2(spv[bp + 1].y);
                s2v hdA = HGd, hdB = as_s2(spv[bp].x);
                HGd = as_s2(spv[bp + 1].x);
                s2v cmA = splat2(0), crA = splat2(0), cmB = splat2(0), crB = splat2(0);
#pragma unroll
                for (int r = 0; r <= QT; ++r) {
                    if (r < QT) cell(r, sgA[r], hdA, fA, cmA, crA);
                    if (r == QT - 1) outb[bp] = make_uint2(as_u32(HG[QT - 1]), as_u32(fA));
                    if (r >= 1) cell(r - 1, sgB[r - 1], hdB, fB, cmB, crB);
                }
                outb[bp + 1] = make_uint2(as_u32(HG[QT - 1]), as_u32(fB));
                if constexpr (ENDS) {
                    close_column(jb + bp + 1, cmA, crA);
                    close_column(jb + bp + 2, cmB, crB);
                }
                // One scheduling region per column pair: hoisting the next pair's E - ge for
                // every row above this one spilled the 32-row first tile (a wait per block).
                if (QT >= 32) __builtin_amdgcn_sched_barrier(0);
            }
        } else {
            // One column with its answer cells: the last row in the last tile (every column up
            // to the target's end, or only its end column), and the whole column at j == n.
            auto column = [&](int b) HY_INLINE {
                const int j = jb + b + 1;
                s2v hgup = as_s2(spv[b].x), f = as_s2(spv[b].y);
                s2v sg[QT];
                load_sg(b, sg);
                s2v hgd = HGd;
                HGd = hgup;
                // Row m of the query (last tile only) is taken as it is computed: selecting it
                // from the row array afterwards would index the array at run time and move it
                // to scratch memory.
                s2v v = splat2(-32768);
#pragma unroll
                for (int r = 0; r < QT; ++r) {
                    const s2v diag = p_add(hgd, sg[r]);
                    const s2v e = p_max(HG[r], p_sub(E[r], ge2));
                    f = p_max(hgup, p_sub(f, ge2));
                    const s2v h = p_max(p_max(diag, e), f);
                    hgd = HG[r];
                    E[r] = e;
                    HG[r] = p_sub(h, go2);
                    hgup = HG[r];
                    if (lastc && r == rm) v = hgup;
                }
                outb[b] = make_uint2(as_u32(HG[QT - 1]), as_u32(f));
                // End tracking: each answer cell of a target joins its best with the full
                // (score, column, row) tie-break, in 32 bits per target; the border cells are
                // already there (border_best), and one column can offer row m twice (last row and
                // last column), which the tie-break absorbs.
                if (lastc) {
                    const bool la = F::lrow ? j <= nA : j == nA;
                    const bool lb = F::lrow ? j <= nB : j == nB;
                    if constexpr (ENDS) {
                        const s2v h = p_add(v, go2);
                        if (la) consider(st.bsA, st.bjA, st.biA, (int)h.x, j, m);
                        if (lb) consider(st.bsB, st.bjB, st.biB, (int)h.y, j, m);
                    } else {
                        run = p_max(run, keep2(p_add(v, go2), la, lb));
                    }
                }
                if (F::lcol && (j == nA || j == nB)) {
#pragma unroll
                    for (int r = 0; r < QT; ++r)
                        if (i0 + r <= m) {
                            if constexpr (ENDS) {
                                const s2v h = p_add(HG[r], go2);
                                if (j == nA) consider(st.bsA, st.bjA, st.biA, (int)h.x, j, i0 + r);
                                if (j == nB) consider(st.bsB, st.bjB, st.biB, (int)h.y, j, i0 + r);
                            } else {
                                run = p_max(run, keep2(p_add(HG[r], go2), j == nA, j == nB));
                            }
                        }
                }
            };
            // Columns without answer cells run in skewed pairs (row r of column j with row r-1
            // of column j+1), as in local mode, so two dependency chains interleave; the last
            // tile and the few column pairs holding a target's end column run one column at a
            // time.
            auto cell = [&](int r, s2v sgv, s2v& hgd, s2v& hgup, s2v& f) HY_INLINE {
                const s2v diag = p_add(hgd, sgv);
                const s2v e = p_max(HG[r], p_sub(E[r], ge2));
                f = p_max(hgup, p_sub(f, ge2));
                const s2v h = p_max(p_max(diag, e), f);
                hgd = HG[r];
                E[r] = e;
                HG[r] = p_sub(h, go2);
                hgup = HG[r];
            };
#pragma unroll
            for (int bp = 0; bp < 4; bp += 2) {
                const int jA = jb + bp + 1, jB = jA + 1;
                const bool ends = F::lcol && (jA == nA || jA == nB || jB == nA || jB == nB);
                if (lastc || __builtin_amdgcn_ballot_w64(ends) != 0) {
                    column(bp);
                    column(bp + 1);
                } else {
                    s2v sgA[QT], sgB[QT];
                    load_sg(bp, sgA);
                    load_sg(bp + 1, sgB);
                    s2v hgupA = as_s2(spv[bp].x), fA = as_s2(spv[bp].y);
                    s2v hgupB = as_s2(spv[bp + 1].x), fB = as_s2(spv[bp + 1].y);
                    s2v hgdA = HGd, hgdB = hgupA;
                    HGd = hgupB;
#pragma unroll
                    for (int r = 0; r <= QT; ++r) {
                        if (r < QT) cell(r, sgA[r], hgdA, hgupA, fA);
                        if (r == QT - 1) outb[bp] = make_uint2(as_u32(HG[QT - 1]), as_u32(fA));
                        if (r >= 1) cell(r - 1, sgB[r - 1], hgdB, hgupB, fB);
                    }
                    outb[bp + 1] = make_uint2(as_u32(HG[QT - 1]), as_u32(fB));
                }
            }
        }
        if (!lastc) {
            if constexpr (PIPE) {
                if (pipe.gout) {
                    // The blocks before this one were stored a block ago: publishing them here,
                    // before this block's stores, keeps the wait for completion off those. Every
                    // pipe.every blocks: each publication waits for the row stores in flight.
                    if (jb > 0 && (jb >> 2) % pipe.every == 0)
                        chain_publish(pipe, pipe.pout, (uint32_t)(jb >> 2));
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        chain_store(pipe.gout + (size_t)(jb + b + 1) * 64, outb[b]);
                } else {
                    uint2* ro = pipe.rout + (size_t)((jb >> 2) & 1) * 256 + (threadIdx.x & 63);
#pragma unroll
                    for (int b = 0; b < 4; ++b) ro[b * 64] = outb[b];
                }
            } else {
#pragma unroll
                for (int b = 0; b < 4; ++b) sp[(size_t)(jb + b + 1) * 64] = outb[b];
            }
        }
    };
    uint2 w = make_uint2(0, 0), wn = make_uint2(0, 0);
    uint2 top[4], topn[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) top[b] = topn[b] = make_uint2(0, 0);
    if constexpr (PIPE) {
        // Block bi at step wave + bi; its top rows were written by the previous wavefront at the
        // step before, with a workgroup barrier between. Residue words stream three blocks ahead:
        // a short tile's block computes in about the latency of a load. The steps before and
        // after this wavefront's blocks take the barriers only, so the block loop issues the same
        // loads every step (clamped at the last block) and the compiler can keep them in flight
        // across steps instead of waiting for every load at the top of a block.
        const int nblk = L >> 2;
        const int lastb = max(nblk - 1, 0);
        auto word = [&](int b) HY_INLINE { return wp[(size_t)min(b, lastb) * 64]; };
        for (int s = 0; s < pipe.wave; ++s) __syncthreads();
        if (nblk > 0) {
            uint2 wn2 = word(2), wn3 = make_uint2(0, 0);
            w = word(0);
            wn = word(1);
            // Chained reader: block bi's top rows were loaded two blocks ahead, once a poll had
            // shown them published, and each poll is read a block after the one it was issued
            // before, so the latency of these accesses overlaps two blocks of compute.
            uint32_t avail = 0;
            uint2 topn2[4];
            unsigned long long poll_prev = 0;
            if constexpr (LINKIN) {
                avail = chain_wait(pipe, pipe.pin, (uint32_t)min(nblk, max(pipe.slack, 2)));
                const int c1 = min(1, lastb) * 4;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    topn[b] = chain_load(pipe.gin + (size_t)(b + 1) * 64);
                    topn2[b] = chain_load(pipe.gin + (size_t)(c1 + b + 1) * 64);
                }
            }
            for (int bi = 0; bi < nblk; ++bi) {
                const int jb = bi * 4;
                wn3 = word(bi + 3);
                unsigned long long poll = 0;
                if constexpr (LINKIN) {
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        top[b] = topn[b];
                        topn[b] = topn2[b];
                    }
                    if (bi + 2 < nblk && avail < (uint32_t)bi + 3)
                        avail = chain_wait(pipe, pipe.pin, (uint32_t)bi + 3);
                    // Block bi + 2's rows (past the last block: the last block's again, unused).
                    const int cb = min(bi + 2, lastb) * 4;
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        topn2[b] = chain_load(pipe.gin + (size_t)(cb + b + 1) * 64);
                    poll = __hip_atomic_load(pipe.pin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else if constexpr (!FIRST) {
                    const uint2* ri = pipe.rin + (size_t)(bi & 1) * 256 + (threadIdx.x & 63);
#pragma unroll
                    for (int b = 0; b < 4; ++b) top[b] = ri[b * 64];
                }
                __builtin_amdgcn_sched_barrier(0);
                block(jb, w, top);
                // The register rotation below copies this block's loads: kept after the block,
                // where they have arrived, instead of hoisted above it to wait for them.
                __builtin_amdgcn_sched_barrier(0);
                w = wn;
                wn = wn2;
                wn2 = wn3;
                if constexpr (LINKIN) {
                    // The poll issued before the previous block: rows up to its count may now be
                    // read.
                    avail = max(avail, chain_count(pipe, poll_prev));
                    poll_prev = poll;
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
                __syncthreads();
            }
        }
        for (int s = pipe.wave + nblk; s < pipe.nsteps; ++s) __syncthreads();
        if (pipe.gout) chain_publish(pipe, pipe.pout, (uint32_t)nblk);
    } else {
        if (L > 0) fetch(0, w, top);
        for (int jb = 0; jb < L; jb += 4) {
            fetch(jb + 4 < L ? jb + 4 : jb, wn, topn);
            __builtin_amdgcn_sched_barrier(0);
            block(jb, w, top);
            __builtin_amdgcn_sched_barrier(0);  // the copies below wait for the prefetch
            w = wn;
#pragma unroll
            for (int b = 0; b < 4; ++b) top[b] = topn[b];
        }
    }
    st.run = run;
#pragma unroll
    for (int k = 0; k < 4; ++k) st.runq[k] = runq[k];
    if constexpr (ENDS && F::local) {
        // The tile's best of each target joins the target's best with the full tie-break: a
        // later tile can hold an equal score in an earlier column.
        const uint32_t b = as_u32(tb), c = as_u32(tc), r = as_u32(tr);
        const int sa = (int)(b & 0xFFFFu), sb = (int)(b >> 16);
        if (sa > 0) consider(st.bsA, st.bjA, st.biA, sa, (int)(c & 0xFFFFu), i0 + (int)(r & 0xFFFFu));
        if (sb > 0) consider(st.bsB, st.bjB, st.biB, sb, (int)(c >> 16), i0 + (int)(r >> 16));
    }
}

// The tile above with FIRST chosen at run time (the first tile of a group starts from the DP
// border instead of a boundary row), and the chained reader role for a pipelined tile whose
// Pipe has a link in (compiled in for score-only pipelined tiles and for the chained kernel).
template <int MODE, int QT, int NW, bool PAIR, bool ENDS = false, bool PIPE = false,
          bool CHAIN = false>
__device__ __forceinline__ void inter16_tile_at(const Inter16Args& a, const Group& g, bool wlive,
                                                const uint8_t* query, int m, int i0, bool first,
                                                bool lastc, int nA, int nB, uint2* sp,
                                                const uint2* wp, uint32_t* pp, PairRun& st,
                                                Pipe pipe = Pipe()) {
    if (first) {
        inter16_tile<MODE, QT, NW, PAIR, true, ENDS, PIPE>(a, g, wlive, query, m, i0, lastc, nA,
                                                           nB, sp, wp, pp, st, pipe);
        return;
    }
    if constexpr (PIPE && (!ENDS || CHAIN)) {
        if (pipe.gin != nullptr) {
            inter16_tile<MODE, QT, NW, PAIR, false, ENDS, PIPE, PIPE>(a, g, wlive, query, m, i0,
                                                                      lastc, nA, nB, sp, wp, pp,
                                                                      st, pipe);
            return;
        }
    }
    inter16_tile<MODE, QT, NW, PAIR, false, ENDS, PIPE>(a, g, wlive, query, m, i0, lastc, nA, nB,
                                                        sp, wp, pp, st, pipe);
}

// One group of 128 targets against one query: full tiles of Q rows, then one tile of the fewest
// rows (8, 16, ... Q) that covers the remainder, so short queries waste at most 7 padded rows.
// The boundary rows go through `spill` (column j of this wavefront at spill[j * 64 + lane]); the
// two targets' scores go to out_score and their best key is folded into *key. Called by the
// one-shot kernel (one group per wavefront) and by the persistent batch kernel (a queue of
// (group, query) items).
template <int MODE, int Q, int NW, bool PAIR, bool ENDS = false>
__device__ __forceinline__ void inter16_group(const Inter16Args& a, const Group& g, bool wlive,
                                              const uint8_t* query, int m, uint2* spill,
                                              unsigned long long* key, uint32_t* pp) {
    using F = MF<MODE>;
    static_assert(Q % 8 == 0 && Q >= 8 && Q <= 32, "tile rows");
    const int lane = threadIdx.x & 63;
    const bool liveA = 2 * lane < (int)g.count, liveB = 2 * lane + 1 < (int)g.count;
    const uint32_t slotA = g.first + 2 * lane, slotB = slotA + 1;
    const int nA = liveA ? (int)a.lens[slotA] : 0;
    const int nB = liveB ? (int)a.lens[slotB] : 0;
    const int go = a.go, ge = a.ge;

    int bsA, bsB, dj, di;
    PairRun st;
    border_best<MODE>(m, nA, go, ge, bsA, dj, di);
    st.bsA = bsA; st.bjA = dj; st.biA = di;
    border_best<MODE>(m, nB, go, ge, bsB, dj, di);
    st.bsB = bsB; st.bjB = dj; st.biB = di;
    st.run = splat2(-32768);
#pragma unroll
    for (int k = 0; k < 4; ++k) st.runq[k] = st.run;

    uint2* sp = spill + lane;
    const uint2* wp = a.words + g.word_off + lane;
    const int nfull = m / Q, rem = m - nfull * Q;
    for (int c = 0; c < nfull; ++c)
        inter16_tile_at<MODE, Q, NW, PAIR, ENDS>(a, g, wlive, query, m, c * Q + 1, c == 0,
                                                 rem == 0 && c == nfull - 1, nA, nB, sp, wp, pp,
                                                 st);
    if (rem > 0) {
        const int i0 = nfull * Q + 1;
        const bool first = nfull == 0;
        if (rem <= 8)
            inter16_tile_at<MODE, 8, NW, PAIR, ENDS>(a, g, wlive, query, m, i0, first, true, nA,
                                                     nB, sp, wp, pp, st);
        else if (Q > 16 && rem <= 16)
            inter16_tile_at<MODE, 16, NW, PAIR, ENDS>(a, g, wlive, query, m, i0, first, true, nA,
                                                      nB, sp, wp, pp, st);
        else if (Q > 24 && rem <= 24)
            inter16_tile_at<MODE, 24, NW, PAIR, ENDS>(a, g, wlive, query, m, i0, first, true, nA,
                                                      nB, sp, wp, pp, st);
        else
            inter16_tile_at<MODE, Q, NW, PAIR, ENDS>(a, g, wlive, query, m, i0, first, true, nA,
                                                     nB, sp, wp, pp, st);
    }
    if (!wlive) return;
    if constexpr (ENDS) {
        // Per target: score, then the 0-based inclusive ends (-1 on the border).
        if (liveA) {
            const uint32_t gi = a.gidx[slotA];
            a.out_score[gi] = st.bsA;
            a.out_qend[gi] = st.biA == 0 ? -1 : (int64_t)st.biA - 1;
            a.out_tend[gi] = st.bjA == 0 ? -1 : (int64_t)st.bjA - 1;
        }
        if (liveB) {
            const uint32_t gi = a.gidx[slotB];
            a.out_score[gi] = st.bsB;
            a.out_qend[gi] = st.biB == 0 ? -1 : (int64_t)st.biB - 1;
            a.out_tend[gi] = st.bjB == 0 ? -1 : (int64_t)st.bjB - 1;
        }
        return;
    }
    const s2v run =
        p_max(st.run, p_max(p_max(st.runq[0], st.runq[1]), p_max(st.runq[2], st.runq[3])));
    const int sA = max(bsA, (int)run.x), sB = max(bsB, (int)run.y);
    unsigned long long k = 0ull;
    if (liveA) {
        const uint32_t gi = a.gidx[slotA];
        if (a.out_score) a.out_score[gi] = sA;
        k = make_key(sA, gi);
    }
    if (liveB) {
        const uint32_t gi = a.gidx[slotB];
        if (a.out_score) a.out_score[gi] = sB;
        const unsigned long long kb = make_key(sB, gi);
        k = kb > k ? kb : k;
    }
    if (key) wave_argmax(k, key);
}

template <int MODE, int Q, int NW, bool PAIR, bool ENDS = false>
__global__ __launch_bounds__(64 * NW)
__attribute__((amdgpu_waves_per_eu(HY_PACKED_WAVES_PER_EU))) void k_inter16(Inter16Args a) {
    extern __shared__ __attribute__((aligned(16))) uint32_t pp[];
    const uint32_t gix = a.group_first + blockIdx.x * NW + (threadIdx.x >> 6);
    const bool wlive = gix < a.group_first + a.ngroups;
    Group g{};
    if (wlive) g = a.groups[gix];
    inter16_group<MODE, Q, NW, PAIR, ENDS>(a, g, wlive, a.query, a.m, a.spill + g.spill_off,
                                           a.key, pp);
}

// Pipelined one-query kernel for databases of few groups (short queries against small
// databases, where one wavefront per group would leave most SIMDs idle): one workgroup per group,
// its W wavefronts holding the query's W tiles of Q rows (the last the fewest of 8/16/24/Q rows
// that cover the remainder), wavefront w one four-column block behind wavefront w - 1 and reading
// its bottom row from LDS (Pipe). Every wavefront keeps the running maxima of its own tile; they
// are folded through LDS and wavefront 0 writes the scores and the argmax key. With ENDS every
// wavefront instead keeps each target's best cell of its tile (the end tracking of the
// one-wavefront kernel, started from the border's best), and wavefront 0 folds them with the
// full tie-break and writes score and ends. LDS: W profiles of pp_words words, then (W - 1)
// double buffers of 512 uint2, then W x 64 words for the fold (6 words per lane with ENDS).
template <int MODE, int Q, bool PAIR, bool ENDS = false>
__global__ __launch_bounds__(512)
__attribute__((amdgpu_waves_per_eu(HY_PACKED_WAVES_PER_EU))) void k_inter16_pipe(
    Inter16Args a, uint32_t pp_words) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int W = (int)(blockDim.x >> 6);
    const Group g = a.groups[a.group_first + blockIdx.x];
    uint32_t* pp = lds + (size_t)wv * pp_words;
    uint2* ring = reinterpret_cast<uint2*>(lds + (size_t)W * pp_words);
    uint32_t* fold = lds + (size_t)W * pp_words + (size_t)(W - 1) * 1024;
    Pipe pipe;
    pipe.wave = wv;
    pipe.nsteps = (int)(g.ncols >> 2) + W - 1;
    pipe.rin = wv > 0 ? ring + (size_t)(wv - 1) * 512 : nullptr;
    pipe.rout = wv + 1 < W ? ring + (size_t)wv * 512 : nullptr;

    const int m = a.m, go = a.go, ge = a.ge;
    const bool liveA = 2 * lane < (int)g.count, liveB = 2 * lane + 1 < (int)g.count;
    const uint32_t slotA = g.first + 2 * lane, slotB = slotA + 1;
    const int nA = liveA ? (int)a.lens[slotA] : 0;
    const int nB = liveB ? (int)a.lens[slotB] : 0;
    PairRun st;
    if constexpr (ENDS) {
        int dj, di;
        border_best<MODE>(m, nA, go, ge, st.bsA, dj, di);
        st.bjA = dj;
        st.biA = di;
        border_best<MODE>(m, nB, go, ge, st.bsB, dj, di);
        st.bjB = dj;
        st.biB = di;
    }
    st.run = splat2(-32768);
#pragma unroll
    for (int k = 0; k < 4; ++k) st.runq[k] = st.run;
    const uint2* wp = a.words + g.word_off + lane;
    const int nfull = m / Q, rem = m - nfull * Q;
    if (wv < nfull) {
        inter16_tile_at<MODE, Q, 1, PAIR, ENDS, true>(a, g, true, a.query, m, wv * Q + 1, wv == 0,
                                                       rem == 0 && wv == nfull - 1, nA, nB,
                                                       nullptr, wp, pp, st, pipe);
    } else {
        const int i0 = nfull * Q + 1;
        const bool first = nfull == 0;
        if (rem <= 8)
            inter16_tile_at<MODE, 8, 1, PAIR, ENDS, true>(a, g, true, a.query, m, i0, first,
                                                           true, nA, nB, nullptr, wp, pp, st,
                                                           pipe);
        else if (Q > 16 && rem <= 16)
            inter16_tile_at<MODE, 16, 1, PAIR, ENDS, true>(a, g, true, a.query, m, i0, first,
                                                            true, nA, nB, nullptr, wp, pp, st,
                                                            pipe);
        else if (Q > 24 && rem <= 24)
            inter16_tile_at<MODE, 24, 1, PAIR, ENDS, true>(a, g, true, a.query, m, i0, first,
                                                            true, nA, nB, nullptr, wp, pp, st,
                                                            pipe);
        else
            inter16_tile_at<MODE, Q, 1, PAIR, ENDS, true>(a, g, true, a.query, m, i0, first,
                                                           true, nA, nB, nullptr, wp, pp, st,
                                                           pipe);
    }
    if constexpr (ENDS) {
        // Fold the wavefronts' best cells (a total order: any fold order gives the same answer)
        // and write each target's score and 0-based inclusive ends (-1 on the border).
        int* fw = reinterpret_cast<int*>(fold) + ((size_t)wv * 64 + lane) * 6;
        fw[0] = st.bsA; fw[1] = st.bjA; fw[2] = st.biA;
        fw[3] = st.bsB; fw[4] = st.bjB; fw[5] = st.biB;
        __syncthreads();
        if (wv != 0) return;
        for (int k = 1; k < W; ++k) {
            const int* fk = reinterpret_cast<const int*>(fold) + ((size_t)k * 64 + lane) * 6;
            consider(st.bsA, st.bjA, st.biA, fk[0], fk[1], fk[2]);
            consider(st.bsB, st.bjB, st.biB, fk[3], fk[4], fk[5]);
        }
        if (liveA) {
            const uint32_t gi = a.gidx[slotA];
            a.out_score[gi] = st.bsA;
            a.out_qend[gi] = st.biA == 0 ? -1 : (int64_t)st.biA - 1;
            a.out_tend[gi] = st.bjA == 0 ? -1 : (int64_t)st.bjA - 1;
        }
        if (liveB) {
            const uint32_t gi = a.gidx[slotB];
            a.out_score[gi] = st.bsB;
            a.out_qend[gi] = st.biB == 0 ? -1 : (int64_t)st.biB - 1;
            a.out_tend[gi] = st.bjB == 0 ? -1 : (int64_t)st.bjB - 1;
        }
        return;
    }
    const s2v mine =
        p_max(st.run, p_max(p_max(st.runq[0], st.runq[1]), p_max(st.runq[2], st.runq[3])));
    fold[wv * 64 + lane] = as_u32(mine);
    __syncthreads();
    if (wv != 0) return;
    s2v run = mine;
    for (int k = 1; k < W; ++k) run = p_max(run, as_s2(fold[k * 64 + lane]));
    int bsA, bsB, dj, di;
    border_best<MODE>(m, nA, go, ge, bsA, dj, di);
    border_best<MODE>(m, nB, go, ge, bsB, dj, di);
    const int sA = max(bsA, (int)run.x), sB = max(bsB, (int)run.y);
    unsigned long long k = 0ull;
    if (liveA) {
        const uint32_t gi = a.gidx[slotA];
        if (a.out_score) a.out_score[gi] = sA;
        k = make_key(sA, gi);
    }
    if (liveB) {
        const uint32_t gi = a.gidx[slotB];
        if (a.out_score) a.out_score[gi] = sB;
        const unsigned long long kb = make_key(sB, gi);
        k = kb > k ? kb : k;
    }
    if (a.key) wave_argmax(k, a.key);
}

// Chained pipelined kernel: the pipelined kernel's query tiles spread over c.chunks workgroups per
// group (chunk k holding tiles [k * W, k * W + W) of the c.tiles tiles, W = wavefronts per
// workgroup), each chunk reading the previous one's bottom rows through the group's boundary
// buffer (Pipe, chained). A database of few groups then keeps several CUs per group busy instead
// of one. Each workgroup takes a ticket v (the order in which the launch's workgroups start) and
// serves group (v / 8 / chunks) * 8 + v % 8 as its chunk (v / 8) % chunks. A chunk waits only on
// the chunk before it, whose ticket is smaller, so that workgroup has already started: the
// launch progresses whatever order the workgroups are dispatched in. The chunks of a group share
// v mod 8, which keeps them on one XCD as long as workgroups start in the order the dispatcher
// hands them to the XCDs in turn. The chunks fold their running maxima in order: chunk k waits
// for chunk k - 1 to publish nblk + 1 blocks (its maxima, already folded with the chunks before
// it, in column 0 of the buffer), folds its own, and passes them on; the last chunk writes the
// scores and the argmax key. Wavefronts of a short last chunk without a tile take the
// workgroup barriers only. ENDS: every wavefront keeps each target's best cell of its tile (from
// the border's best, as in k_inter16_pipe); wavefront 0 folds the workgroup's through LDS, joins
// those of the chunks before it (c.fold, slot of link k - 1, read once link k - 1 shows nblk + 1
// blocks) and passes the result on (slot of link k); the last chunk writes each target's score
// and 0-based ends. The fold is a total order, so the chunks' order of folding does not change
// the answer.
template <int MODE, int Q, bool PAIR, bool ENDS = false>
__global__ __launch_bounds__(512)
__attribute__((amdgpu_waves_per_eu(HY_PACKED_WAVES_PER_EU))) void k_inter16_chain(
    Inter16Args a, ChainArgs c, uint32_t pp_words) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int W = (int)(blockDim.x >> 6);
    __shared__ uint32_t ticket_s;
    if (threadIdx.x == 0) ticket_s = (uint32_t)(atomicAdd(c.ticket, 1ull) - c.ticket_base);
    __syncthreads();
    const uint32_t v = ticket_s;
    if (v >= gridDim.x) {
        // The counter disagrees with the tickets the host counted (only after a launch that did
        // not complete): report it instead of serving a wrong chunk.
        if (threadIdx.x == 0)
            __hip_atomic_store(c.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return;
    }
    const uint32_t rest = v >> 3;
    const int k = (int)(rest % (uint32_t)c.chunks);
    const uint32_t gi = (rest / (uint32_t)c.chunks) * 8u + (v & 7u);
    if (gi >= a.ngroups) return;  // the whole workgroup: a padding group of the grid
    const Group g = a.groups[a.group_first + gi];
    const int t0 = k * W;
    const int tn = min(c.tiles - t0, W);  // tiles of this chunk
    const int nblk = (int)(g.ncols >> 2);
    uint32_t* pp = lds + (size_t)wv * pp_words;
    uint2* ring = reinterpret_cast<uint2*>(lds + (size_t)W * pp_words);
    uint32_t* fold = lds + (size_t)W * pp_words + (size_t)(W - 1) * 1024;
    uint2* link = a.spill + g.spill_off + lane;
    unsigned long long* prog = c.progress + (size_t)gi * kChainMaxLinks;
    Pipe pipe;
    pipe.wave = wv;
    pipe.nsteps = nblk + W - 1;
    pipe.rin = wv > 0 ? ring + (size_t)(wv - 1) * 512 : nullptr;
    pipe.rout = wv + 1 < tn ? ring + (size_t)wv * 512 : nullptr;
    pipe.epoch = c.epoch;
    pipe.slack = c.slack;
    pipe.every = c.every;
    pipe.err = c.err;
    pipe.spin_limit = c.spin_limit;
    if (wv == 0 && k > 0) {
        pipe.gin = link;
        pipe.pin = prog + (k - 1);
    }
    if (wv == tn - 1 && k + 1 < c.chunks) {
        pipe.gout = link;
        pipe.pout = prog + k;
    }

    const int m = a.m, go = a.go, ge = a.ge;
    const bool liveA = 2 * lane < (int)g.count, liveB = 2 * lane + 1 < (int)g.count;
    const uint32_t slotA = g.first + 2 * lane, slotB = slotA + 1;
    const int nA = liveA ? (int)a.lens[slotA] : 0;
    const int nB = liveB ? (int)a.lens[slotB] : 0;
    PairRun st;
    if constexpr (ENDS) {
        int dj, di;
        border_best<MODE>(m, nA, go, ge, st.bsA, dj, di);
        st.bjA = dj;
        st.biA = di;
        border_best<MODE>(m, nB, go, ge, st.bsB, dj, di);
        st.bjB = dj;
        st.biB = di;
    }
    st.run = splat2(-32768);
#pragma unroll
    for (int q = 0; q < 4; ++q) st.runq[q] = st.run;
    const uint2* wp = a.words + g.word_off + lane;
    const int nfull = m / Q, rem = m - nfull * Q;
    const int t = t0 + wv;  // this wavefront's tile
    if (wv >= tn) {
        // No tile: the profile staging's two barriers and one per block step.
        for (int s = 0; s < 2 + pipe.nsteps; ++s) __syncthreads();
    } else if (t < nfull) {
        inter16_tile_at<MODE, Q, 1, PAIR, ENDS, true, true>(a, g, true, a.query, m, t * Q + 1, t == 0,
                                                        rem == 0 && t == nfull - 1, nA, nB,
                                                        nullptr, wp, pp, st, pipe);
    } else {
        const int i0 = nfull * Q + 1;
        const bool first = nfull == 0;
        if (rem <= 8)
            inter16_tile_at<MODE, 8, 1, PAIR, ENDS, true, true>(a, g, true, a.query, m, i0, first,
                                                            true, nA, nB, nullptr, wp, pp, st,
                                                            pipe);
        else if (Q > 16 && rem <= 16)
            inter16_tile_at<MODE, 16, 1, PAIR, ENDS, true, true>(a, g, true, a.query, m, i0, first,
                                                             true, nA, nB, nullptr, wp, pp, st,
                                                             pipe);
        else if (Q > 24 && rem <= 24)
            inter16_tile_at<MODE, 24, 1, PAIR, ENDS, true, true>(a, g, true, a.query, m, i0, first,
                                                             true, nA, nB, nullptr, wp, pp, st,
                                                             pipe);
        else
            inter16_tile_at<MODE, Q, 1, PAIR, ENDS, true, true>(a, g, true, a.query, m, i0, first,
                                                            true, nA, nB, nullptr, wp, pp, st,
                                                            pipe);
    }
    if constexpr (ENDS) {
        int* fw = reinterpret_cast<int*>(fold) + ((size_t)wv * 64 + lane) * 6;
        fw[0] = st.bsA; fw[1] = st.bjA; fw[2] = st.biA;
        fw[3] = st.bsB; fw[4] = st.bjB; fw[5] = st.biB;
        __syncthreads();
        if (wv != 0) return;
        for (int q = 1; q < W; ++q) {
            const int* fq = reinterpret_cast<const int*>(fold) + ((size_t)q * 64 + lane) * 6;
            consider(st.bsA, st.bjA, st.biA, fq[0], fq[1], fq[2]);
            consider(st.bsB, st.bjB, st.biB, fq[3], fq[4], fq[5]);
        }
        uint2* slots = c.fold + (size_t)gi * kChainMaxLinks * 192 + lane;
        if (k > 0) {
            (void)chain_wait(pipe, prog + (k - 1), (uint32_t)nblk + 1);
            const uint2* fin = slots + (size_t)(k - 1) * 192;
            const uint2 x = chain_load(fin), y = chain_load(fin + 64), z = chain_load(fin + 128);
            consider(st.bsA, st.bjA, st.biA, (int)x.x, (int)x.y, (int)y.x);
            consider(st.bsB, st.bjB, st.biB, (int)y.y, (int)z.x, (int)z.y);
        }
        if (k + 1 < c.chunks) {
            uint2* fout = slots + (size_t)k * 192;
            chain_store(fout, make_uint2((uint32_t)st.bsA, (uint32_t)st.bjA));
            chain_store(fout + 64, make_uint2((uint32_t)st.biA, (uint32_t)st.bsB));
            chain_store(fout + 128, make_uint2((uint32_t)st.bjB, (uint32_t)st.biB));
            chain_publish(pipe, prog + k, (uint32_t)nblk + 1);
            return;
        }
        if (liveA) {
            const uint32_t gx = a.gidx[slotA];
            a.out_score[gx] = st.bsA;
            a.out_qend[gx] = st.biA == 0 ? -1 : (int64_t)st.biA - 1;
            a.out_tend[gx] = st.bjA == 0 ? -1 : (int64_t)st.bjA - 1;
        }
        if (liveB) {
            const uint32_t gx = a.gidx[slotB];
            a.out_score[gx] = st.bsB;
            a.out_qend[gx] = st.biB == 0 ? -1 : (int64_t)st.biB - 1;
            a.out_tend[gx] = st.bjB == 0 ? -1 : (int64_t)st.bjB - 1;
        }
        return;
    }
    const s2v mine =
        p_max(st.run, p_max(p_max(st.runq[0], st.runq[1]), p_max(st.runq[2], st.runq[3])));
    fold[wv * 64 + lane] = as_u32(mine);
    __syncthreads();
    if (wv != 0) return;
    s2v run = mine;
    for (int q = 1; q < W; ++q) run = p_max(run, as_s2(fold[q * 64 + lane]));
    if (k > 0) {
        (void)chain_wait(pipe, prog + (k - 1), (uint32_t)nblk + 1);
        run = p_max(run, as_s2(chain_load(link).x));
    }
    if (k + 1 < c.chunks) {
        chain_store(link, make_uint2(as_u32(run), 0u));
        chain_publish(pipe, prog + k, (uint32_t)nblk + 1);
        return;
    }
    int bsA, bsB, dj, di;
    border_best<MODE>(m, nA, go, ge, bsA, dj, di);
    border_best<MODE>(m, nB, go, ge, bsB, dj, di);
    const int sA = max(bsA, (int)run.x), sB = max(bsB, (int)run.y);
    unsigned long long key = 0ull;
    if (liveA) {
        const uint32_t gx = a.gidx[slotA];
        if (a.out_score) a.out_score[gx] = sA;
        key = make_key(sA, gx);
    }
    if (liveB) {
        const uint32_t gx = a.gidx[slotB];
        if (a.out_score) a.out_score[gx] = sB;
        const unsigned long long kb = make_key(sB, gx);
        key = kb > key ? kb : key;
    }
    if (a.key) wave_argmax(key, a.key);
}

// Persistent batch kernel: each one-wavefront workgroup takes (group, query) items from a device
// counter until the queue is empty. Items run group-major over groups in descending column count
// and, within a group, queries in descending length, so the longest work starts first. Each
// workgroup owns one boundary-row slot of bt.slot_stride entries. PAIR: the residue-pair
// profile (small alphabets), else the per-symbol one.
template <int MODE, int Q, bool PAIR>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(HY_PACKED_WAVES_PER_EU))) void k_inter16_batch(Inter16Args a,
                                                                                   BatchArgs bt) {
    extern __shared__ __attribute__((aligned(16))) uint32_t pp[];
    __shared__ uint32_t item_s;
    uint2* slot = bt.slots + (size_t)blockIdx.x * bt.slot_stride;
    const uint32_t nq = *bt.nq;
    const uint32_t n_items = nq * a.ngroups;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) item_s = atomicAdd(bt.counter, 1u);
        __syncthreads();
        const uint32_t item = item_s;
        if (item >= n_items) break;
        const uint32_t gix = a.group_first + item / nq;
        const uint32_t qi = bt.qorder[item % nq];
        const Group g = a.groups[gix];
        inter16_group<MODE, Q, 1, PAIR>(a, g, true, bt.queries + bt.qoff[qi], (int)bt.qlen[qi],
                                        slot, bt.keys + qi, pp);
    }
}



// ------------------------------------------------------------------------------------------------
// Batched form, local mode: two QUERIES per lane.
// ------------------------------------------------------------------------------------------------
// Each lane takes one target at a time (the lane's two targets of a group in two passes) against
// a pair of queries held in the two 16-bit halves, so the tile profile is stored already paired,
// pq[s * S + r] = (sub(qa_r, s), sub(qb_r, s)), and a column's substitution pairs are read with
// no per-lane pairing: one v_perm_b32 per cell pair less than the two-targets kernel. The cell
// update is the local-mode one of inter16_tile; the rows of the shorter query past its end hold
// i16::MIN in the profile, so their cells only carry gap extensions of real cells and never
// exceed the real maximum.
template <int QT, bool FIRST>
__device__ __forceinline__ void inter16q_tile(const Inter16Args& a, const Group& g, int pass,
                                              const uint8_t* qa, int ma, const uint8_t* qb,
                                              int mb, int i0, bool lastc, uint2* sp,
                                              const uint2* wp, uint32_t* pp, s2v (&runq)[4]) {
    constexpr int S = QT + 4;  // words per symbol row: 16-byte aligned, rows 4 banks apart
    const int L = (int)g.ncols;
    const int A = a.al_eff + 1;
    const s2v go2 = splat2(a.go), ge2 = splat2(a.ge);
    __syncthreads();
    for (int k = threadIdx.x; k < A * QT; k += 64) {
        const int sym = k / QT, r = k - sym * QT, i = i0 + r;
        int va = -32768, vb = -32768;
        if (sym < a.al_eff) {
            if (i <= ma) va = sat16((long long)a.matrix[(int)qa[i - 1] * a.al + sym]);
            if (i <= mb) vb = sat16((long long)a.matrix[(int)qb[i - 1] * a.al + sym]);
        }
        pp[sym * S + r] = (uint32_t)(uint16_t)(short)va | ((uint32_t)(uint16_t)(short)vb << 16);
    }
    __syncthreads();

    s2v HG[QT], E[QT];  // per row: H of the last column and E' of the next one
#pragma unroll
    for (int r = 0; r < QT; ++r) HG[r] = E[r] = splat2(0);
    s2v HGd = splat2(0);

    // Prefetch of the next four-column block pinned ahead of the current block's compute, as in
    // inter16_tile; a lane's residues are the low (pass 0) or high (pass 1) word of the pair.
    auto fetch = [&](int jb, uint2& wq, uint2* top) HY_INLINE {
        wq = wp[(size_t)(jb >> 2) * 64];
        if (!FIRST) {
#pragma unroll
            for (int b = 0; b < 4; ++b) top[b] = sp[(size_t)(jb + b + 1) * 64];
        }
    };
    auto block = [&](int jb, const uint2 wv, const uint2* topin) HY_INLINE {
        const uint32_t w = pass ? wv.y : wv.x;
        uint2 spv[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) spv[b] = FIRST ? make_uint2(0u, 0u) : topin[b];
        uint2 outb[4];
        auto load_sg = [&](int b, s2v* sg) HY_INLINE {
            const int c = (int)((w >> (8 * b)) & 0xffu);
            const uint4* row = reinterpret_cast<const uint4*>(&pp[c * S]);
#pragma unroll
            for (int r4 = 0; r4 < QT / 4; ++r4) {
                const uint4 v = row[r4];
                sg[4 * r4 + 0] = as_s2(v.x); sg[4 * r4 + 1] = as_s2(v.y);
                sg[4 * r4 + 2] = as_s2(v.z); sg[4 * r4 + 3] = as_s2(v.w);
            }
        };
        auto cell = [&](int r, s2v sgv, s2v& hd, s2v& f) HY_INLINE {
            const s2v h = p_max(p_max(p_add(hd, sgv), E[r]), f);
            hd = HG[r];
            HG[r] = h;
            const s2v hg = p_subu(h, go2);
            E[r] = p_max(hg, p_subu(E[r], ge2));
            f = p_max(hg, p_subu(f, ge2));
            runq[r & 3] = p_max(runq[r & 3], h);
        };
#pragma unroll
        for (int bp = 0; bp < 4; bp += 2) {
            s2v sgA[QT], sgB[QT];
            load_sg(bp, sgA);
            load_sg(bp + 1, sgB);
            s2v fA = as_s2(spv[bp].y), fB = as_s2(spv[bp + 1].y);
            s2v hdA = HGd, hdB = as_s2(spv[bp].x);
            HGd = as_s2(spv[bp + 1].x);
#pragma unroll
            for (int r = 0; r <= QT; ++r) {
                if (r < QT) cell(r, sgA[r], hdA, fA);
                if (r == QT - 1) outb[bp] = make_uint2(as_u32(HG[QT - 1]), as_u32(fA));
                if (r >= 1) cell(r - 1, sgB[r - 1], hdB, fB);
            }
            outb[bp + 1] = make_uint2(as_u32(HG[QT - 1]), as_u32(fB));
        }
        if (!lastc) {
#pragma unroll
            for (int b = 0; b < 4; ++b) sp[(size_t)(jb + b + 1) * 64] = outb[b];
        }
    };
    uint2 w = make_uint2(0, 0), wn = make_uint2(0, 0);
    uint2 top[4], topn[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) top[b] = topn[b] = make_uint2(0, 0);
    if (L > 0) fetch(0, w, top);
    for (int jb = 0; jb < L; jb += 4) {
        fetch(jb + 4 < L ? jb + 4 : jb, wn, topn);
        __builtin_amdgcn_sched_barrier(0);
        block(jb, w, top);
        __builtin_amdgcn_sched_barrier(0);  // the copies below wait for the prefetch
        w = wn;
#pragma unroll
        for (int b = 0; b < 4; ++b) top[b] = topn[b];
    }
}

template <int QT>
__device__ __forceinline__ void inter16q_tile_at(const Inter16Args& a, const Group& g, int pass,
                                                 const uint8_t* qa, int ma, const uint8_t* qb,
                                                 int mb, int i0, bool first, bool lastc,
                                                 uint2* sp, const uint2* wp, uint32_t* pp,
                                                 s2v (&runq)[4]) {
    if (first)
        inter16q_tile<QT, true>(a, g, pass, qa, ma, qb, mb, i0, lastc, sp, wp, pp, runq);
    else
        inter16q_tile<QT, false>(a, g, pass, qa, ma, qb, mb, i0, lastc, sp, wp, pp, runq);
}

// One group of 128 targets against a query pair (ma >= mb; qb absent when kb is ~0u): two passes,
// each folding the pass's 64 targets into both queries' argmax keys.
template <int Q>
__device__ __forceinline__ void inter16q_group(const Inter16Args& a, const Group& g,
                                               const uint8_t* qa, int ma, const uint8_t* qb,
                                               int mb, uint32_t ka, uint32_t kb, uint2* spill,
                                               unsigned long long* keys, uint32_t* pp) {
    const int lane = threadIdx.x & 63;
    uint2* sp = spill + lane;
    const uint2* wp = a.words + g.word_off + lane;
    const int nfull = ma / Q, rem = ma - nfull * Q;
    for (int pass = 0; pass < 2; ++pass) {
        s2v runq[4] = {splat2(0), splat2(0), splat2(0), splat2(0)};
        for (int c = 0; c < nfull; ++c)
            inter16q_tile_at<Q>(a, g, pass, qa, ma, qb, mb, c * Q + 1, c == 0,
                                rem == 0 && c == nfull - 1, sp, wp, pp, runq);
        if (rem > 0) {
            const int i0 = nfull * Q + 1;
            const bool first = nfull == 0;
            if (rem <= 8)
                inter16q_tile_at<8>(a, g, pass, qa, ma, qb, mb, i0, first, true, sp, wp, pp, runq);
            else if (Q > 16 && rem <= 16)
                inter16q_tile_at<16>(a, g, pass, qa, ma, qb, mb, i0, first, true, sp, wp, pp,
                                     runq);
            else if (Q > 24 && rem <= 24)
                inter16q_tile_at<24>(a, g, pass, qa, ma, qb, mb, i0, first, true, sp, wp, pp,
                                     runq);
            else
                inter16q_tile_at<Q>(a, g, pass, qa, ma, qb, mb, i0, first, true, sp, wp, pp, runq);
        }
        const s2v run = p_max(p_max(runq[0], runq[1]), p_max(runq[2], runq[3]));
        const int t = 2 * lane + pass;
        unsigned long long key_a = 0ull, key_b = 0ull;
        if (t < (int)g.count) {
            const uint32_t gi = a.gidx[g.first + t];
            key_a = make_key((int)run.x, gi);
            key_b = make_key((int)run.y, gi);
        }
        wave_argmax(key_a, keys + ka);
        if (kb != ~0u) wave_argmax(key_b, keys + kb);
    }
}

// Persistent batch kernel over (group, query pair) items: queries pair up in descending length
// order (qorder[2p], qorder[2p + 1]); an odd last query pairs with an empty one.
template <int Q>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(HY_PACKED_WAVES_PER_EU))) void k_inter16q_batch(Inter16Args a,
                                                                                    BatchArgs bt) {
    extern __shared__ __attribute__((aligned(16))) uint32_t pp[];
    __shared__ uint32_t item_s;
    uint2* slot = bt.slots + (size_t)blockIdx.x * bt.slot_stride;
    const uint32_t nq = *bt.nq;
    const uint32_t npairs = (nq + 1) / 2;
    const uint32_t n_items = npairs * a.ngroups;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) item_s = atomicAdd(bt.counter, 1u);
        __syncthreads();
        const uint32_t item = item_s;
        if (item >= n_items) break;
        const uint32_t gix = a.group_first + item / npairs;
        const uint32_t p = item % npairs;
        const uint32_t ia = bt.qorder[2 * p];
        const bool hasb = 2 * p + 1 < nq;
        const uint32_t ib = hasb ? bt.qorder[2 * p + 1] : ia;
        const Group g = a.groups[gix];
        inter16q_group<Q>(a, g, bt.queries + bt.qoff[ia], (int)bt.qlen[ia],
                          bt.queries + bt.qoff[ib], hasb ? (int)bt.qlen[ib] : 0, ia,
                          hasb ? ib : ~0u, slot, bt.keys, pp);
    }
}

// ------------------------------------------------------------------------------------------------
// Intra-sequence systolic kernel, 32-bit lanes, one (query, target) pair per wavefront.
// ------------------------------------------------------------------------------------------------
// LDSMAT: the substitution matrix (al_eff^2 entries) is staged in LDS; otherwise it is read from
// global memory (alphabets above 128 symbols).
template <int MODE, int R, bool ENDS, bool WINNER, bool LDSMAT>
__global__ __launch_bounds__(64) void k_sys32(SysArgs a) {
    using F = MF<MODE>;
    extern __shared__ int smat[];
    __shared__ uint8_t tring[256];
    __shared__ int2 sring[256];
    const int lane = threadIdx.x;
    // Winner mode: an all-zero key means no hit was recorded; fall back to index 0 and let the
    // self-check below flag it.
    const bool batch = a.qoff != nullptr;
    // Pair mode: block b aligns pair list[b], whose query is qoff/qlen[list[b]].
    const uint32_t qi =
        batch ? (a.qsel >= 0 ? (uint32_t)a.qsel : (a.pairs ? a.list[blockIdx.x] : blockIdx.y)) : 0u;
    if (batch && qi >= *a.nq) return;
    const uint8_t* query = batch ? a.queries + a.qoff[qi] : a.query;
    const unsigned long long wk = WINNER ? a.winner_key[qi] : 0ull;
    const uint32_t gi = WINNER ? (wk != 0ull ? key_index(wk) : 0u) : a.list[blockIdx.x];
    const uint8_t* t = a.flat + a.off[gi];
    const int n = (int)(a.off[gi + 1] - a.off[gi]);
    const int m = batch ? (int)a.qlen[qi] : a.m, go = a.go, ge = a.ge, ale = a.al_eff;
    const int rowstride = LDSMAT ? ale : a.al;
    // LDS matrix rows 0..ale-1 hold the scores; row ale is the padding row (kPad32), used by query
    // rows past the end so the lookup needs no branch.
    if (LDSMAT) {
        for (int k = lane; k < (ale + 1) * ale; k += 64)
            smat[k] = k < ale * ale ? a.matrix[(k / ale) * a.al + (k % ale)] : kPad32;
    }
    __syncthreads();

    int bs = INT_MIN, bj = INT_MAX, bi = INT_MAX;
    if (lane == 0) border_best<MODE>(m, n, go, ge, bs, bj, bi);

    constexpr int P = 64 * R;
    const int npass = (m + P - 1) / P;
    const size_t spill_block =
        WINNER ? qi
               : ((a.qsel >= 0 || a.pairs) ? blockIdx.x : (size_t)qi * gridDim.x + blockIdx.x);
    int2* sp = a.spill + spill_block * a.spill_stride;
    for (int p = 0; p < npass; ++p) {
        const int r0 = p * P + lane * R + 1;
        int qb[R], HG[R], E[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r0 + r;
            qb[r] = i <= m ? (int)query[i - 1] * rowstride : (LDSMAT ? ale * ale : -1);
            HG[r] = wsub(border<MODE>(i, 0, go, ge), go);
            E[r] = kNeg32;
        }
        int HGd = wsub(border<MODE>(r0 - 1, 0, go, ge), go);
        int topHG = 0, topF = kNeg32;
        // Residues and the previous pass's boundary row are staged through LDS rings of four
        // 64-entry blocks: block e+1 is fetched into registers at the start of 64-step epoch e
        // and stored at its end, so global latency overlaps a whole epoch of compute. At step s
        // lane l reads residue s-l (blocks e-1, e) and lane 0 reads boundary column s+1.
        tring[lane] = lane < n ? t[lane] : (uint8_t)0;
        if (p > 0) sring[lane] = lane + 1 <= n ? sp[lane + 1] : make_int2(0, kNeg32);
        __syncthreads();
        uint8_t tnext = 0;
        int2 snext = make_int2(0, kNeg32);
        const int steps = n + 63;
        for (int s = 0; s < steps; ++s) {
            if ((s & 63) == 0) {
                const int idx = s + 64 + lane;
                tnext = idx < n ? t[idx] : (uint8_t)0;
                if (p > 0) {
                    const int col = s + 65 + lane;
                    snext = col <= n ? sp[col] : make_int2(0, kNeg32);
                }
            }
            const int j = s - lane + 1;
            if (lane == 0 && j <= n) {
                if (p == 0) {
                    topHG = wsub(border<MODE>(0, j, go, ge), go);
                    topF = kNeg32;
                } else {
                    const int2 v = sring[s & 255];
                    topHG = v.x;
                    topF = v.y;
                }
            }
            int outHG = 0, outF = 0;
            if (j >= 1 && j <= n) {
                const int c = tring[(j - 1) & 255];
                int f = topF, hgup = topHG, hgd = HGd;
                HGd = topHG;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int sub = LDSMAT ? smat[qb[r] + c]
                                           : (qb[r] >= 0 ? a.matrix[qb[r] + c] : kPad32);
                    const int diag = wadd(wadd(hgd, go), sub);
                    const int e = max(HG[r], wsub(E[r], ge));
                    f = max(hgup, wsub(f, ge));
                    int h = max(max(diag, e), f);
                    if (F::local) h = max(h, 0);
                    hgd = HG[r];
                    E[r] = e;
                    HG[r] = wsub(h, go);
                    hgup = HG[r];
                    // Candidate cells of the mode's answer region, selected without branches;
                    // a non-candidate offers INT_MIN, which never beats a real cell.
                    const int i = r0 + r;
                    const bool cand =
                        i <= m && (F::local || (F::lcol && j == n) || (i == m && (F::lrow || j == n)));
                    const int v = cand ? h : INT_MIN;
                    if (ENDS) consider(bs, bj, bi, v, j, i);
                    else bs = max(bs, v);
                }
                outHG = HG[R - 1];
                outF = f;
                if (lane == 63 && p + 1 < npass) sp[j] = make_int2(outHG, outF);
            }
            // Lane l receives lane l-1's bottom row (wave_shr:1); lane 0 keeps its own value.
            topHG = __builtin_amdgcn_update_dpp(topHG, outHG, 0x138, 0xf, 0xf, false);
            topF = __builtin_amdgcn_update_dpp(topF, outF, 0x138, 0xf, 0xf, false);
            if ((s & 63) == 63) {
                tring[(s + 1 + lane) & 255] = tnext;
                if (p > 0) sring[(s + 1 + lane) & 255] = snext;
                __syncthreads();
            }
        }
        __threadfence_block();
        __syncthreads();
    }

    if (ENDS) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const int os = __shfl_xor(bs, off), oj = __shfl_xor(bj, off), oi = __shfl_xor(bi, off);
            consider(bs, bj, bi, os, oj, oi);
        }
    } else {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) bs = max(bs, __shfl_xor(bs, off));
    }
    if (WINNER) {
        if (lane == 0) {
            const int ks = key_score(wk);
            Result* r = a.result + qi;
            r->score = bs;
            r->db_index = gi;
            r->query_end = ENDS ? (bi == 0 ? -1 : (int64_t)bi - 1) : -1;
            r->target_end = ENDS ? (bj == 0 ? -1 : (int64_t)bj - 1) : -1;
            r->check = (bs == ks && wk != 0ull) ? 0 : 1;
            r->pad = 0;
        }
    } else {
        if (lane == 0) {
            if (a.out_score) a.out_score[gi] = bs;
            if (ENDS) {
                a.out_qend[gi] = bi == 0 ? -1 : (int64_t)bi - 1;
                a.out_tend[gi] = bj == 0 ? -1 : (int64_t)bj - 1;
            }
            if (a.key) atomicMax(a.key + qi, make_key(bs, gi));
        }
    }
}

template <int MODE, int R, bool LDSMAT>
hipError_t launch_sys32_mrl(bool ends, bool winner, const SysArgs& a, uint32_t nblocks,
                            hipStream_t st, uint32_t nq) {
    const size_t lds = LDSMAT ? (size_t)(a.al_eff + 1) * a.al_eff * sizeof(int) : 0;
    if (winner) {
        if (ends)
            hipLaunchKernelGGL((k_sys32<MODE, R, true, true, LDSMAT>), dim3(1, nq), dim3(64), lds,
                               st, a);
        else
            hipLaunchKernelGGL((k_sys32<MODE, R, false, true, LDSMAT>), dim3(1, nq), dim3(64), lds,
                               st, a);
    } else {
        if (ends)
            hipLaunchKernelGGL((k_sys32<MODE, R, true, false, LDSMAT>), dim3(nblocks, nq), dim3(64),
                               lds, st, a);
        else
            hipLaunchKernelGGL((k_sys32<MODE, R, false, false, LDSMAT>), dim3(nblocks, nq),
                               dim3(64), lds, st, a);
    }
    return hipGetLastError();
}

template <int MODE, int R>
hipError_t launch_sys32_mr(bool ends, bool winner, const SysArgs& a, uint32_t nblocks,
                           hipStream_t st, uint32_t nq) {
    if ((a.al_eff + 1) * a.al_eff <= 16384)
        return launch_sys32_mrl<MODE, R, true>(ends, winner, a, nblocks, st, nq);
    return launch_sys32_mrl<MODE, R, false>(ends, winner, a, nblocks, st, nq);
}

template <int MODE, int Q, int NW, bool PAIR>
hipError_t launch_inter16_qw(const Inter16Args& a, hipStream_t st) {
    const uint32_t lds = PAIR ? inter16_lds_bytes(a.al_eff, Q) : inter16p_lds_bytes(a.al_eff, Q);
    const uint32_t blocks = (a.ngroups + NW - 1) / NW;
    auto kern = k_inter16<MODE, Q, NW, PAIR>;
    if (lds > 64u * 1024u) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * NW), lds, st, a);
    return hipGetLastError();
}

template <int MODE, int Q, bool ENDS>
hipError_t launch_inter16_pipe_q(const Inter16Args& a, hipStream_t st) {
    const int W = (a.m + Q - 1) / Q;
    if (W < 2 || W > kPipeMaxWaves) return hipErrorInvalidValue;
    if (ENDS && (!a.out_score || !a.out_qend || !a.out_tend)) return hipErrorInvalidValue;
    const uint32_t pp_words = inter16_profile_words(a.al_eff, Q);
    const size_t lds = inter16_pipe_lds_bytes(a.al_eff, Q, W, ENDS);
    if (lds > 64u * 1024u) return hipErrorInvalidValue;
    if (inter16_pair_small(a.al_eff, Q))
        hipLaunchKernelGGL((k_inter16_pipe<MODE, Q, true, ENDS>), dim3(a.ngroups), dim3(64 * W),
                           lds, st, a, pp_words);
    else
        hipLaunchKernelGGL((k_inter16_pipe<MODE, Q, false, ENDS>), dim3(a.ngroups), dim3(64 * W),
                           lds, st, a, pp_words);
    return hipGetLastError();
}

template <int MODE, int Q, bool ENDS>
hipError_t launch_inter16_chain_q(const Inter16Args& a, const ChainArgs& c, int W,
                                  hipStream_t st) {
    if (ENDS && (!a.out_score || !a.out_qend || !a.out_tend || !c.fold))
        return hipErrorInvalidValue;
    const uint32_t pp_words = inter16_profile_words(a.al_eff, Q);
    const size_t lds = inter16_chain_lds_bytes(a.al_eff, Q, W, ENDS);
    if (lds > 64u * 1024u) return hipErrorInvalidValue;
    const uint32_t blocks = (a.ngroups + 7u) / 8u * 8u * (uint32_t)c.chunks;
    if (inter16_pair_small(a.al_eff, Q))
        hipLaunchKernelGGL((k_inter16_chain<MODE, Q, true, ENDS>), dim3(blocks), dim3(64 * W),
                           lds, st, a, c, pp_words);
    else
        hipLaunchKernelGGL((k_inter16_chain<MODE, Q, false, ENDS>), dim3(blocks), dim3(64 * W),
                           lds, st, a, c, pp_words);
    return hipGetLastError();
}

template <int MODE, int Q>
hipError_t launch_inter16_batch_q(const Inter16Args& a, const BatchArgs& b, uint32_t wgs,
                                  hipStream_t st) {
    const uint32_t lds = inter16_profile_words(a.al_eff, Q) * 4u;
    if (inter16_pair_small(a.al_eff, Q))
        hipLaunchKernelGGL((k_inter16_batch<MODE, Q, true>), dim3(wgs), dim3(64), lds, st, a, b);
    else
        hipLaunchKernelGGL((k_inter16_batch<MODE, Q, false>), dim3(wgs), dim3(64), lds, st, a, b);
    return hipGetLastError();
}

}  // namespace

template <int MODE>
hipError_t launch_inter32_m(bool ends, const InterArgs& a, uint32_t ngroups, hipStream_t st) {
    constexpr int Q = 32;
    const size_t lds = (size_t)(a.al_eff + 1) * (Q + 4) * sizeof(int);
    if (ends)
        hipLaunchKernelGGL((k_inter32<MODE, Q, true>), dim3(ngroups), dim3(64), lds, st, a);
    else
        hipLaunchKernelGGL((k_inter32<MODE, Q, false>), dim3(ngroups), dim3(64), lds, st, a);
    return hipGetLastError();
}

// Rows per lane: the smallest of 2, 4, 6, 8 whose 64-lane pass covers the query (one pass up to
// 512 rows); longer queries take several passes of 8 rows per lane.
template <int MODE>
hipError_t launch_sys32_m(bool ends, bool winner, const SysArgs& a, uint32_t nblocks,
                          hipStream_t st, uint32_t nq) {
    if (a.m <= 128) return launch_sys32_mr<MODE, 2>(ends, winner, a, nblocks, st, nq);
    if (a.m <= 256) return launch_sys32_mr<MODE, 4>(ends, winner, a, nblocks, st, nq);
    if (a.m <= 384) return launch_sys32_mr<MODE, 6>(ends, winner, a, nblocks, st, nq);
    return launch_sys32_mr<MODE, 8>(ends, winner, a, nblocks, st, nq);
}

// waves: 1 = one-wavefront workgroups, each with its own profile (the residue-pair profile for
// small alphabets, the per-symbol one otherwise); 4 or 8 = workgroups sharing the residue-pair
// profile.
template <int MODE>
hipError_t launch_inter16_m(const Inter16Args& a, int q, int waves, hipStream_t st) {
    if (waves <= 1 && inter16_pair_small(a.al_eff, q)) {
        // Small alphabets: the tile's residue-pair profile fits a few KB of the wavefront's LDS.
        if (q <= 16) return launch_inter16_qw<MODE, 16, 1, true>(a, st);
        if (q <= 24) return launch_inter16_qw<MODE, 24, 1, true>(a, st);
        return launch_inter16_qw<MODE, 32, 1, true>(a, st);
    }
    if (waves <= 1) {
        if (q <= 16) return launch_inter16_qw<MODE, 16, 1, false>(a, st);
        if (q <= 24) return launch_inter16_qw<MODE, 24, 1, false>(a, st);
        return launch_inter16_qw<MODE, 32, 1, false>(a, st);
    }
    if (waves >= 8) {
        if (q <= 16) return launch_inter16_qw<MODE, 16, 8, true>(a, st);
        if (q <= 24) return launch_inter16_qw<MODE, 24, 8, true>(a, st);
        return launch_inter16_qw<MODE, 32, 8, true>(a, st);
    }
    if (q <= 16) return launch_inter16_qw<MODE, 16, 4, true>(a, st);
    if (q <= 24) return launch_inter16_qw<MODE, 24, 4, true>(a, st);
    return launch_inter16_qw<MODE, 32, 4, true>(a, st);
}

// Tile rows of the pipelined kernel: the smallest of 16, 24, 32 whose tiles fit kPipeMaxWaves
// wavefronts; q is that choice (inter16_pipe_rows).
template <int MODE>
hipError_t launch_inter16_pipe_m(const Inter16Args& a, int q, bool ends, hipStream_t st) {
    if (a.ngroups == 0) return hipSuccess;
    if (ends) {
        // Local-mode end tracking keeps tiles of at most 24 rows (inter16_pipe_rows_ends).
        if (q == 16) return launch_inter16_pipe_q<MODE, 16, true>(a, st);
        if (q == 24) return launch_inter16_pipe_q<MODE, 24, true>(a, st);
        if constexpr (MODE != kSW) {
            if (q == 32) return launch_inter16_pipe_q<MODE, 32, true>(a, st);
        }
        return hipErrorInvalidValue;
    }
    if (q == 16) return launch_inter16_pipe_q<MODE, 16, false>(a, st);
    if (q == 24) return launch_inter16_pipe_q<MODE, 24, false>(a, st);
    if (q == 32) return launch_inter16_pipe_q<MODE, 32, false>(a, st);
    return hipErrorInvalidValue;
}

// Chained pipelined kernel: q-row tiles, `waves` wavefronts (tiles) per workgroup. The shape
// must cover the query exactly: tiles = ceil(m / q), chunks = ceil(tiles / waves), so every
// chunk holds at least one tile and its last wavefront holding one is the chunk's writer.
template <int MODE>
hipError_t launch_inter16_chain_m(const Inter16Args& a, const ChainArgs& c, int q, int waves,
                                  bool ends, hipStream_t st) {
    if (a.ngroups == 0) return hipSuccess;
    if (q <= 0 || waves < 1 || waves > 8 || c.tiles != (a.m + q - 1) / q ||
        c.chunks != (c.tiles + waves - 1) / waves || c.chunks < 1 ||
        c.chunks > kChainMaxChunks || c.every < 1 || !c.progress || !c.err || !c.ticket ||
        !inter16_chain_rows_ok(MODE, q, ends))
        return hipErrorInvalidValue;
    if (ends) {
        if (q == 8) return launch_inter16_chain_q<MODE, 8, true>(a, c, waves, st);
        if (q == 16) return launch_inter16_chain_q<MODE, 16, true>(a, c, waves, st);
        if constexpr (MODE != kSW) {
            if (q == 24) return launch_inter16_chain_q<MODE, 24, true>(a, c, waves, st);
            if (q == 32) return launch_inter16_chain_q<MODE, 32, true>(a, c, waves, st);
        }
        return hipErrorInvalidValue;
    }
    if (q == 8) return launch_inter16_chain_q<MODE, 8, false>(a, c, waves, st);
    if (q == 16) return launch_inter16_chain_q<MODE, 16, false>(a, c, waves, st);
    if (q == 24) return launch_inter16_chain_q<MODE, 24, false>(a, c, waves, st);
    if (q == 32) return launch_inter16_chain_q<MODE, 32, false>(a, c, waves, st);
    return hipErrorInvalidValue;
}

template <int MODE>
hipError_t launch_inter16_batch_m(const Inter16Args& a, const BatchArgs& b, int q, uint32_t wgs,
                                  hipStream_t st) {
    if (q <= 16) return launch_inter16_batch_q<MODE, 16>(a, b, wgs, st);
    if (q <= 24) return launch_inter16_batch_q<MODE, 24>(a, b, wgs, st);
    return launch_inter16_batch_q<MODE, 32>(a, b, wgs, st);
}

// Resident one-wavefront workgroups of the batch kernel per CU at the occupancy limit.
template <int MODE>
uint32_t inter16_batch_occupancy_m(int q, int al_eff) {
    int per_cu = 0;
    const int rows = inter16_tile_rows(q);
    const size_t lds = inter16_profile_words(al_eff, rows) * 4u + sizeof(uint32_t);
    hipError_t e = hipErrorInvalidValue;
    auto occ = [&](const void* f) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64, lds);
    };
    if (inter16_pair_small(al_eff, rows)) {
        if (rows == 16) occ(reinterpret_cast<const void*>(k_inter16_batch<MODE, 16, true>));
        else if (rows == 24) occ(reinterpret_cast<const void*>(k_inter16_batch<MODE, 24, true>));
        else occ(reinterpret_cast<const void*>(k_inter16_batch<MODE, 32, true>));
    } else {
        if (rows == 16) occ(reinterpret_cast<const void*>(k_inter16_batch<MODE, 16, false>));
        else if (rows == 24) occ(reinterpret_cast<const void*>(k_inter16_batch<MODE, 24, false>));
        else occ(reinterpret_cast<const void*>(k_inter16_batch<MODE, 32, false>));
    }
    if (e != hipSuccess || per_cu <= 0) per_cu = 2;
    return (uint32_t)per_cu;
}

// End tracking in the packed one-query kernel (scan_all with ScoreEnd): one wavefront per
// workgroup with its own profile (the one-query default shape).
template <int MODE>
hipError_t launch_inter16_ends_m(const Inter16Args& a, hipStream_t st) {
    if (a.ngroups == 0) return hipSuccess;
    // Local mode: 24-row tiles fit the end tracking without register spills (32-row tiles spill
    // 160 bytes per lane). Measured, scan_all with ends, 360 aa x 1M protein targets: 30.9 ms
    // against 31.3 ms at 32 rows and 32.3 ms at 24 rows with the compiler's compare-and-select
    // minimum. The other modes track ends at their answer cells only (the last row of the last
    // tile, the end column), so their tile keeps the score-only shape.
    constexpr int Q = MODE == kSW ? 24 : HY_ENDS_TILE_NONLOCAL;
    if (inter16_pair_small(a.al_eff, Q)) {
        const uint32_t lds = inter16_lds_bytes(a.al_eff, Q);
        hipLaunchKernelGGL((k_inter16<MODE, Q, 1, true, true>), dim3(a.ngroups), dim3(64), lds,
                           st, a);
    } else {
        const uint32_t lds = inter16p_lds_bytes(a.al_eff, Q);
        hipLaunchKernelGGL((k_inter16<MODE, Q, 1, false, true>), dim3(a.ngroups), dim3(64), lds,
                           st, a);
    }
    return hipGetLastError();
}

#if HY_KMODE == 0

uint32_t inter16q_lds_bytes(int al_eff, int q) {
    return ((uint32_t)al_eff + 1) * ((uint32_t)inter16_tile_rows(q) + 4) * 4u;
}

hipError_t launch_inter16q_batch(const Inter16Args& a, const BatchArgs& b, int q, uint32_t wgs,
                                 hipStream_t st) {
    if (a.ngroups == 0 || wgs == 0) return hipSuccess;
    const uint32_t lds = inter16q_lds_bytes(a.al_eff, q);
    const int rows = inter16_tile_rows(q);
    if (rows == 16)
        hipLaunchKernelGGL((k_inter16q_batch<16>), dim3(wgs), dim3(64), lds, st, a, b);
    else if (rows == 24)
        hipLaunchKernelGGL((k_inter16q_batch<24>), dim3(wgs), dim3(64), lds, st, a, b);
    else
        hipLaunchKernelGGL((k_inter16q_batch<32>), dim3(wgs), dim3(64), lds, st, a, b);
    return hipGetLastError();
}

uint32_t inter16q_batch_occupancy(int q, int al_eff) {
    int per_cu = 0;
    const int rows = inter16_tile_rows(q);
    const size_t lds = inter16q_lds_bytes(al_eff, rows);
    hipError_t e = hipErrorInvalidValue;
    auto occ = [&](const void* f) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64, lds);
    };
    if (rows == 16) occ(reinterpret_cast<const void*>(k_inter16q_batch<16>));
    else if (rows == 24) occ(reinterpret_cast<const void*>(k_inter16q_batch<24>));
    else occ(reinterpret_cast<const void*>(k_inter16q_batch<32>));
    if (e != hipSuccess || per_cu <= 0) per_cu = 2;
    return (uint32_t)per_cu;
}
#endif

template hipError_t launch_inter32_m<HY_KMODE>(bool, const InterArgs&, uint32_t, hipStream_t);
template hipError_t launch_sys32_m<HY_KMODE>(bool, bool, const SysArgs&, uint32_t, hipStream_t,
                                             uint32_t);
template hipError_t launch_inter16_m<HY_KMODE>(const Inter16Args&, int, int, hipStream_t);
template hipError_t launch_inter16_batch_m<HY_KMODE>(const Inter16Args&, const BatchArgs&, int,
                                                     uint32_t, hipStream_t);
template uint32_t inter16_batch_occupancy_m<HY_KMODE>(int, int);
template hipError_t launch_inter16_ends_m<HY_KMODE>(const Inter16Args&, hipStream_t);
template hipError_t launch_inter16_pipe_m<HY_KMODE>(const Inter16Args&, int, bool, hipStream_t);
template hipError_t launch_inter16_chain_m<HY_KMODE>(const Inter16Args&, const ChainArgs&, int,
                                                     int, bool, hipStream_t);

}  // namespace hy
