// Traceback kernels (src/align.rs:169-420), compiled once per alignment mode (HY_KMODE).
//
//  * k_trace_dp — the full affine DP of one (query, target) pair per wavefront, in the systolic
//    shape of the long-target scan kernel: lane l owns R consecutive query rows, step s feeds it
//    column s - l + 1, and each lane passes its bottom row to the next lane with wave_shr:1. Per
//    cell it keeps no score, only the 4-bit code of the move the canonical backward walk takes
//    there, so one 32-bit word per lane and step holds R <= 8 cells and the store is one
//    coalesced 256-byte line per wavefront and step. It also finds the end cell with the
//    answer-region tie-break (score, then smallest column, then smallest row).
//  * k_trace_walk — the backward walk, one lane per pair: from the end cell it follows the codes
//    (a local alignment stops at its first zero cell; a diagonal step is preferred to a gap, a
//    target-consuming gap to a query-consuming one, closing a gap to extending it) and writes the
//    ops from the back of the pair's op region towards the front, so they end up in forward
//    order without a reversal pass.
//
// Code nibble of cell (i, j), i, j >= 1: bits 0-1 the walk's move out of H (0 stop at a local
// zero, 1 diagonal, 2 into the horizontal gap E, 3 into the vertical gap F); bit 2 set when
// E(i, j) = H(i, j-1) - go, i.e. the walk in E closes the gap here; bit 3 the same for F with
// H(i-1, j). The tests are the walk's own equalities (align.rs:386-415) on the same values.
#include "hy_device.h"

#ifndef HY_KMODE
#error "compile hy_trace.hip with -DHY_KMODE=<mode>"
#endif

namespace hy {
namespace {

// AlignOp (src/align.rs:47-57).
constexpr uint8_t kOpMatch = 0, kOpMismatch = 1, kOpIns = 2, kOpDel = 3;

template <int MODE, int R, bool LDSMAT>
__global__ __launch_bounds__(64) void k_trace_dp(TraceArgs a) {
    using F = MF<MODE>;
    extern __shared__ int smat[];
    __shared__ uint8_t tring[256];
    __shared__ int2 sring[256];
    const int lane = threadIdx.x;
    const uint32_t item = blockIdx.x;
    const uint32_t pos = a.first + item;  // list position (the slab offsets follow it)
    uint32_t gi = a.list ? a.list[pos] : pos;
    if (a.winner_key) {
        const unsigned long long wk = *a.winner_key;
        gi = wk != 0ull ? key_index(wk) : 0u;
    }
    const uint8_t* t = a.flat + a.off[gi];
    const int n = (int)(a.off[gi + 1] - a.off[gi]);
    const int m = a.m, go = a.go, ge = a.ge, ale = a.al_eff;
    const int rowstride = LDSMAT ? ale : a.al;
    // The LDS copy of the matrix carries +go, so the diagonal candidate H(i-1, j-1) + sub is one
    // add from the stored H - go. Matrix entries are clamped to +-(2^29 - 1) and go < 2^29, so
    // the biased entries cannot wrap.
    if (LDSMAT) {
        for (int k = lane; k < (ale + 1) * ale; k += 64)
            smat[k] = k < ale * ale ? a.matrix[(k / ale) * a.al + (k % ale)] + go : kPad32;
    }
    __syncthreads();

    int bs = INT_MIN, bj = INT_MAX, bi = INT_MAX;
    if (lane == 0) border_best<MODE>(m, n, go, ge, bs, bj, bi);

    constexpr int P = 64 * R;
    const int npass = (m + P - 1) / P;
    const int S = n + 63;
    uint32_t* codes = a.codes + (a.slab_off ? a.slab_off[pos] - a.slab_base : 0ull);
    int2* sp = reinterpret_cast<int2*>(codes + (size_t)npass * S * 64);
    for (int p = 0; p < npass; ++p) {
        const int r0 = p * P + lane * R + 1;
        int qb[R], HG[R], E[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r0 + r;
            qb[r] = i <= m ? (int)a.query[i - 1] * rowstride : (LDSMAT ? ale * ale : -1);
            HG[r] = wsub(border<MODE>(i, 0, go, ge), go);
            E[r] = kNeg32;
        }
        int HGd = wsub(border<MODE>(r0 - 1, 0, go, ge), go);
        int topHG = 0, topF = kNeg32;
        // Best answer-region cell of this lane in this pass. A lane meets its cells in
        // (column, row) order, the tie-break's order, so a strictly greater score is the only
        // update; passes are merged with the full tie-break below.
        int pbs = INT_MIN, pbj = INT_MAX, pbi = INT_MAX;
        // Residues and the previous pass's boundary row go through LDS rings of four 64-entry
        // blocks, fetched one 64-step epoch ahead (the scheme of the long-target scan kernel).
        tring[lane] = lane < n ? t[lane] : (uint8_t)0;
        if (p > 0) sring[lane] = lane + 1 <= n ? sp[lane + 1] : make_int2(0, kNeg32);
        __syncthreads();
        uint8_t tnext = 0;
        int2 snext = make_int2(0, kNeg32);
        uint32_t* cw = codes + (size_t)p * S * 64 + lane;
        for (int s = 0; s < S; ++s) {
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
            uint32_t code = 0;
            if (j >= 1 && j <= n) {
                const int c = tring[(j - 1) & 255];
                int f = topF, hgup = topHG, hgd = HGd;
                HGd = topHG;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int sub = LDSMAT ? smat[qb[r] + c]
                                           : (qb[r] >= 0 ? a.matrix[qb[r] + c] + go : kPad32);
                    const int diag = wadd(hgd, sub);
                    const int e = max(HG[r], wsub(E[r], ge));
                    const int fv = max(hgup, wsub(f, ge));
                    int h = max(max(diag, e), fv);
                    if (F::local) h = max(h, 0);
                    const uint32_t mv = (F::local && h == 0) ? 0u
                                        : (h == diag ? 1u : (h == e ? 2u : 3u));
                    code |= (mv | (e == HG[r] ? 4u : 0u) | (fv == hgup ? 8u : 0u)) << (4 * r);
                    f = fv;
                    hgd = HG[r];
                    E[r] = e;
                    HG[r] = wsub(h, go);
                    hgup = HG[r];
                    const int i = r0 + r;
                    const bool cand =
                        i <= m && (F::local || (F::lcol && j == n) || (i == m && (F::lrow || j == n)));
                    const bool take = cand & (h > pbs);
                    pbs = take ? h : pbs;
                    pbj = take ? j : pbj;
                    pbi = take ? i : pbi;
                }
                outHG = HG[R - 1];
                outF = f;
                if (lane == 63 && p + 1 < npass) sp[j] = make_int2(outHG, outF);
            }
            cw[(size_t)s * 64] = code;
            topHG = __builtin_amdgcn_update_dpp(topHG, outHG, 0x138, 0xf, 0xf, false);
            topF = __builtin_amdgcn_update_dpp(topF, outF, 0x138, 0xf, 0xf, false);
            if ((s & 63) == 63) {
                tring[(s + 1 + lane) & 255] = tnext;
                if (p > 0) sring[(s + 1 + lane) & 255] = snext;
                __syncthreads();
            }
        }
        consider(bs, bj, bi, pbs, pbj, pbi);
        __threadfence_block();
        __syncthreads();
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int os = __shfl_xor(bs, off), oj = __shfl_xor(bj, off), oi = __shfl_xor(bi, off);
        consider(bs, bj, bi, os, oj, oi);
    }
    if (lane == 0) a.ends[item] = make_int4(bs, bi, bj, (int)gi);
}

// PACKED: the codes of k_trace16 (item k is slot slot0 + k of its group: lane (slot - first) / 2,
// 16-bit half (slot - first) % 2; tile c, column j, rows 4b..4b+3 of the tile in word
// ((c * ncols + j - 1) * Q / 4 + b) * 64 + lane). Otherwise the systolic slabs of k_trace_dp.
// The walk of item k (one lane). k_trace_walk runs it one lane per item; k_trace16 runs it for its
// own group's targets once their end cells are folded.
template <int MODE, bool PACKED>
__device__ __forceinline__ void walk_item(const TraceArgs& a, uint32_t k) {
    using F = MF<MODE>;
    // Packed items are indexed by their slot (chunks may overlap in time), others by item.
    const int4 en = a.ends[PACKED ? a.slot0 + k : k];
    const uint32_t gi = (uint32_t)en.w;
    int i = en.y, j = en.z;
    const uint8_t* t = a.flat + a.off[gi];
    const int n = (int)(a.off[gi + 1] - a.off[gi]);
    const int m = a.m, R = a.rows, P = 64 * a.rows, S = n + 63;
    const uint32_t* codes;
    uint32_t L = 0, shift = 0;
    if constexpr (PACKED) {
        const uint32_t gix = a.item_group[a.slot0 + k];
        const Group g = a.groups16[gix];
        const uint32_t rel = a.slot0 + k - g.first;
        L = g.ncols;
        shift = 16u * (rel & 1u);
        codes = a.codes + (a.gslab[gix] - a.gslab_base) + (rel >> 1);
    } else {
        codes = a.codes + (a.slab_off ? a.slab_off[a.first + k] - a.slab_base : 0ull);
    }
    const uint64_t obase = a.winner_key ? 0ull : (uint64_t)gi * (uint64_t)m + a.off[gi];
    uint8_t* ops = a.ops + obase;
    uint64_t w = (uint64_t)m + (uint64_t)n;
    auto nib = [&](int ci, int cj) -> uint32_t {
        const int ii = ci - 1;
        if constexpr (PACKED) {
            constexpr int Q = kTrace16Rows;
            const int c = ii / Q, r = ii - c * Q;
            const uint32_t word =
                codes[(((size_t)c * L + (size_t)(cj - 1)) * (Q / 4) + (size_t)(r >> 2)) * 64];
            return (word >> (shift + 4u * (uint32_t)(r & 3))) & 15u;
        } else {
            const int p = ii / P, rem = ii - p * P, l = rem / R, r = rem - l * R;
            const uint32_t word = codes[((size_t)p * S + (size_t)(cj - 1 + l)) * 64 + l];
            return (word >> (4 * r)) & 15u;
        }
    };
    // 0 = H, 1 = E (horizontal gap, Del), 2 = F (vertical gap, Ins). Every iteration emits an op
    // or moves from H into a gap state that emits next, so 2 (m + n) + 2 iterations always end
    // the walk; running out of them flags the pair as failed instead of looping.
    int st = 0;
    uint32_t code = 0;
    bool ok = false;
    for (int guard = 2 * (m + n) + 2; guard > 0; --guard) {
        if (st == 0) {
            if (i == 0 || j == 0) {
                // Borders: a local alignment or a free border stops; a penalised border is one
                // gap run to the origin (align.rs:364-383).
                if (!F::local) {
                    if (i == 0 && !F::top)
                        for (; j > 0; --j) ops[--w] = kOpDel;
                    if (j == 0 && !F::left)
                        for (; i > 0; --i) ops[--w] = kOpIns;
                }
                ok = true;
                break;
            }
            code = nib(i, j);
            const uint32_t mv = code & 3u;
            if (mv == 0u) {
                ok = true;
                break;
            }
            if (mv == 1u) {
                ops[--w] = a.query[i - 1] == t[j - 1] ? kOpMatch : kOpMismatch;
                --i;
                --j;
            } else {
                st = (int)mv - 1;
            }
        } else if (st == 1) {
            ops[--w] = kOpDel;
            // E(i, 1) always closes (E(i, 0) is the sentinel); forcing it keeps j >= 1 in E.
            const bool close = (code & 4u) != 0u || j == 1;
            --j;
            if (close) st = 0;
            else code = nib(i, j);
        } else {
            ops[--w] = kOpIns;
            const bool close = (code & 8u) != 0u || i == 1;
            --i;
            if (close) st = 0;
            else code = nib(i, j);
        }
    }
    TraceOut o;
    o.score = en.x;
    o.check = ok ? 0 : 1;
    if (a.winner_key) {
        const unsigned long long wk = *a.winner_key;
        if (wk == 0ull || key_score(wk) != en.x) o.check = 1;
    }
    o.query_start = (uint64_t)i;
    o.query_end = (uint64_t)en.y;
    o.target_start = (uint64_t)j;
    o.target_end = (uint64_t)en.z;
    o.n_ops = (uint64_t)m + (uint64_t)n - w;
    o.ops_offset = obase + w;
    a.out[a.winner_key ? 0u : gi] = o;
}

template <int MODE, bool PACKED>
__global__ __launch_bounds__(256) void k_trace_walk(TraceArgs a) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k < a.nitems) walk_item<MODE, PACKED>(a, k);
}

// Packed traceback DP (scan_all_aligned over the packed 16-bit layout): one workgroup of W
// wavefronts per group, two targets per lane in the 16-bit halves. The query's tiles of Q rows go
// through the workgroup in passes of W: in a pass wavefront w holds tile pass * W + w and runs one
// column behind wavefront w - 1, whose bottom row it reads from an LDS double buffer (every
// wavefront takes one workgroup barrier per column step); the first wavefront of a pass reads
// the previous pass's bottom row from the scan's boundary buffer, written there by the last
// wavefront. A group's latency is then about its columns times the passes, not times the tiles,
// and a chunk of groups whose codes fit the pool keeps W times as many wavefronts busy. Each
// wavefront stages its own tile's per-symbol profile (sub + go, paired per lane with
// v_perm_b32). Cell values are kept as in the non-local packed scan, HG = H - go and true E, F
// (local mode adds the clamp at 0): the host sends a target here only when its width proof keeps
// every real cell and H - go within 16 bits, so every equality the code tests is exact. Per cell
// it computes k_trace_dp's 4-bit code with packed operations (an equality is one xor and an
// opaque v_pk_min_u16 with 1: written as the vector builtin, the compiler turns that minimum
// into a compare and a select per half and a repack) and adds it into one word per 4-row block
// (four codes per 16-bit half), stored per tile, column and block: word
// ((t * L + j - 1) * Q / 4 + block) * 64 + lane. The end cell is found as the packed
// end-tracking scan finds it (local: the first row of each column's maximum, columns merged in
// order; other modes: the answer cells), each wavefront over its tiles, folded with the full
// tie-break at the end. Codes of padded rows and columns are computed and never read: the walk
// stays inside the target.

// The DP border saturated to 16 bits.
template <int MODE>
__device__ __forceinline__ int tborder16(int i, int j, int go, int ge) {
    if (i == 0) return (j == 0 || MF<MODE>::top) ? 0 : sat16(-(long long)gap32(go, ge, j));
    return MF<MODE>::left ? 0 : sat16(-(long long)gap32(go, ge, i));
}

template <int MODE>
__global__ __launch_bounds__(64 * kTrace16MaxWaves) void k_trace16(TraceArgs a, uint32_t pp_words) {
    using F = MF<MODE>;
    constexpr int Q = kTrace16Rows;
    static_assert(Q % 8 == 0, "tile rows: whole profile words and 4-row code blocks");
    constexpr int RD = trace16_profile_stride();
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    typedef unsigned short u2 __attribute__((ext_vector_type(2)));
    const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int W = (int)(blockDim.x >> 6);
    uint32_t* pp = lds + (size_t)wv * pp_words;
    uint2* ring = reinterpret_cast<uint2*>(lds + (size_t)W * pp_words);  // (W - 1) x 2 x 64
    int* fold = reinterpret_cast<int*>(ring + (size_t)(W - 1) * 128);    // W x 64 x 6
    const uint32_t gix = a.group_first + blockIdx.x;
    const Group g = a.groups16[gix];
    const bool liveA = 2 * lane < (int)g.count, liveB = 2 * lane + 1 < (int)g.count;
    const uint32_t slotA = g.first + 2 * lane, slotB = slotA + 1;
    const int nA = liveA ? (int)a.lens16[slotA] : 0;
    const int nB = liveB ? (int)a.lens16[slotB] : 0;
    const int L = (int)g.ncols;
    const int m = a.m, go = a.go, ge = a.ge;
    const int A = a.al_eff + 1;
    const s2v go2 = splat2(go), ge2 = splat2(ge), one2 = splat2(1), zero2 = splat2(0);
    const s2v neg2 = splat2(-32768);
    int bsA, bjA, biA, bsB, bjB, biB;
    border_best<MODE>(m, nA, go, ge, bsA, bjA, biA);
    border_best<MODE>(m, nB, go, ge, bsB, bjB, biB);
    const uint2* wp = a.words16 + g.word_off + lane;
    uint2* sp = a.spill16 + g.spill_off + lane;
    uint32_t* cw = a.codes + (a.gslab[gix] - a.gslab_base) + lane;
    const int tiles = (m + Q - 1) / Q;
    const int passes = (tiles + W - 1) / W;
    const uint2* rin = wv > 0 ? ring + (size_t)(wv - 1) * 128 + lane : nullptr;
    uint2* rout = wv + 1 < W ? ring + (size_t)wv * 128 + lane : nullptr;
    for (int p = 0; p < passes; ++p) {
        const int t = p * W + wv;  // this wavefront's tile in the pass
        const bool active = t < tiles;
        const int i0 = t * Q + 1;
        const bool lastc = t == tiles - 1;
        const int rm = m - i0;  // register of row m in the last tile
        __syncthreads();
        if (active)
            for (int k = lane; k < A * (Q / 2); k += 64) {
                const int sym = k / (Q / 2), rr = k - sym * (Q / 2);
                const int i = i0 + 2 * rr;  // rows i (low half) and i + 1 (high half)
                int v0 = -32768, v1 = -32768;
                if (sym < a.al_eff) {
                    if (i <= m)
                        v0 = sat16((long long)a.matrix[(int)a.query[i - 1] * a.al + sym] + go);
                    if (i + 1 <= m)
                        v1 = sat16((long long)a.matrix[(int)a.query[i] * a.al + sym] + go);
                }
                pp[sym * RD + rr] =
                    (uint32_t)(uint16_t)(short)v0 | ((uint32_t)(uint16_t)(short)v1 << 16);
            }
        __syncthreads();

        s2v HG[Q], E[Q];
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            HG[r] = splat2(sat16((long long)tborder16<MODE>(i0 + r, 0, go, ge) - go));
            E[r] = neg2;
        }
        s2v HGd = splat2(sat16((long long)tborder16<MODE>(i0 - 1, 0, go, ge) - go));
        s2v tb = zero2, tc = zero2, tr = zero2;  // local mode: tile best, its column, its row
        // Steps before this wavefront's first column take the barrier only; then one column per
        // step, residue words and (first wavefront of a later pass) the boundary row loaded one
        // column ahead.
        for (int s = 0; s < wv; ++s) __syncthreads();
        uint2 w = make_uint2(0, 0), top = make_uint2(0, 0);
        if (active && L > 0) {
            w = wp[0];
            if (t > 0 && wv == 0) top = sp[64];
        }
        for (int j = 1; j <= L; ++j) {
            if (active) {
                const int b = (j - 1) & 3;
                uint2 wn = w, topn = top;
                if (j < L) {
                    if (b == 3) wn = wp[(size_t)(j >> 2) * 64];
                    if (t > 0 && wv == 0) topn = sp[(size_t)(j + 1) * 64];
                }
                s2v hgup, f;
                if (t == 0) {
                    hgup = splat2(sat16((long long)tborder16<MODE>(0, j, go, ge) - go));
                    f = neg2;
                } else if (wv == 0) {
                    hgup = as_s2(top.x);
                    f = as_s2(top.y);
                } else {
                    const uint2 v = rin[(size_t)(j & 1) * 64];
                    hgup = as_s2(v.x);
                    f = as_s2(v.y);
                }
                const int ca = (int)((w.x >> (8 * b)) & 0xffu);
                const int cb = (int)((w.y >> (8 * b)) & 0xffu);
                s2v sg[Q];
                {
                    const uint4* ra = reinterpret_cast<const uint4*>(&pp[ca * RD]);
                    const uint4* rb = reinterpret_cast<const uint4*>(&pp[cb * RD]);
#pragma unroll
                    for (int r8 = 0; r8 < Q / 8; ++r8) {
                        const uint4 va = ra[r8], vb = rb[r8];
                        const uint32_t wa[4] = {va.x, va.y, va.z, va.w};
                        const uint32_t wb[4] = {vb.x, vb.y, vb.z, vb.w};
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            sg[8 * r8 + 2 * k] =
                                as_s2(__builtin_amdgcn_perm(wb[k], wa[k], 0x05040100u));
                            sg[8 * r8 + 2 * k + 1] =
                                as_s2(__builtin_amdgcn_perm(wb[k], wa[k], 0x07060302u));
                        }
                    }
                }
                s2v hgd = HGd;
                HGd = hgup;
                s2v cm = zero2, cr = zero2;  // local mode: the column's maximum and first row
                s2v vrow = neg2;             // other modes: row m's H - go (last tile)
                uint32_t acc[Q / 4];
#pragma unroll
                for (int k = 0; k < Q / 4; ++k) acc[k] = 0u;
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    const s2v hgold = HG[r];  // H(i, j-1) - go
                    const s2v fup = hgup;     // H(i-1, j) - go
                    const s2v diag = p_add(hgd, sg[r]);
                    const s2v e = p_max(HG[r], p_sub(E[r], ge2));
                    f = p_max(hgup, p_sub(f, ge2));
                    s2v h = p_max(p_max(diag, e), f);
                    if (F::local) h = p_max(h, zero2);
                    hgd = HG[r];
                    E[r] = e;
                    HG[r] = p_sub(h, go2);
                    hgup = HG[r];
                    // Move code: 0 local zero, 1 diagonal, 2 into E, 3 into F; +4 where E closes
                    // (E == H(i, j-1) - go), +8 where F closes (F == H(i-1, j) - go).
                    const s2v dnz = p_minu_pk(as_s2(as_u32(h) ^ as_u32(diag)), one2);
                    const s2v enz = p_minu_pk(as_s2(as_u32(h) ^ as_u32(e)), one2);
                    const s2v ecl = p_minu_pk(as_s2(as_u32(e) ^ as_u32(hgold)), one2);
                    const s2v fcl = p_minu_pk(as_s2(as_u32(f) ^ as_u32(fup)), one2);
                    u2 mv = __builtin_bit_cast(u2, p_madu(dnz, enz, dnz)) + (u2)1;  // 1, 2, 3
                    if (F::local) mv = mv * __builtin_bit_cast(u2, p_minu_pk(h, one2));
                    const u2 code = mv + (u2)4 - __builtin_bit_cast(u2, ecl) * (u2)4 + (u2)8 -
                                    __builtin_bit_cast(u2, fcl) * (u2)8;
                    acc[r >> 2] = as_u32(p_madu(__builtin_bit_cast(s2v, code),
                                                splat2(1 << (4 * (r & 3))), as_s2(acc[r >> 2])));
                    if constexpr (F::local) {
                        const uint32_t up = p_nonzero_mask(p_subu(h, cm));
                        cm = p_max(cm, h);
                        cr = as_s2((up & as_u32(splat2(r))) | (~up & as_u32(cr)));
                    } else {
                        if (lastc && r == rm) vrow = hgup;
                    }
                }
#pragma unroll
                for (int k = 0; k < Q / 4; ++k)
                    cw[(((size_t)t * L + (size_t)(j - 1)) * (Q / 4) + (size_t)k) * 64] = acc[k];
                if (!lastc) {
                    const uint2 out = make_uint2(as_u32(HG[Q - 1]), as_u32(f));
                    if (rout)
                        rout[(size_t)(j & 1) * 64] = out;
                    else
                        sp[(size_t)j * 64] = out;
                }
                if constexpr (F::local) {
                    // The column's (max, first row) joins the tile's best: strictly greater only.
                    const s2v up = p_minu_pk(p_subu(cm, tb), one2);
                    tb = p_max(tb, cm);
                    tc = p_maxu(tc, p_madu(up, splat2(j), zero2));
                    tr = p_madu(up, p_subw(cr, tr), tr);
                } else {
                    if (lastc) {
                        const bool la = F::lrow ? j <= nA : j == nA;
                        const bool lb = F::lrow ? j <= nB : j == nB;
                        const s2v hh = p_add(vrow, go2);
                        if (la) consider(bsA, bjA, biA, (int)hh.x, j, m);
                        if (lb) consider(bsB, bjB, biB, (int)hh.y, j, m);
                    }
                    if (F::lcol && (j == nA || j == nB)) {
#pragma unroll
                        for (int r = 0; r < Q; ++r)
                            if (i0 + r <= m) {
                                const s2v hh = p_add(HG[r], go2);
                                if (j == nA) consider(bsA, bjA, biA, (int)hh.x, j, i0 + r);
                                if (j == nB) consider(bsB, bjB, biB, (int)hh.y, j, i0 + r);
                            }
                    }
                }
                w = wn;
                top = topn;
            }
            __syncthreads();
        }
        for (int s = wv; s < W - 1; ++s) __syncthreads();
        if constexpr (F::local) {
            if (active) {
                const uint32_t bb = as_u32(tb), cc = as_u32(tc), rr = as_u32(tr);
                const int sa = (int)(bb & 0xFFFFu), sb = (int)(bb >> 16);
                if (sa > 0)
                    consider(bsA, bjA, biA, sa, (int)(cc & 0xFFFFu), i0 + (int)(rr & 0xFFFFu));
                if (sb > 0)
                    consider(bsB, bjB, biB, sb, (int)(cc >> 16), i0 + (int)(rr >> 16));
            }
        }
    }
    // Fold the wavefronts' bests (the tie-break is a total order, so any fold order gives the
    // same answer) and write each target's end cell and group.
    int* fw = fold + ((size_t)wv * 64 + lane) * 6;
    fw[0] = bsA; fw[1] = bjA; fw[2] = biA; fw[3] = bsB; fw[4] = bjB; fw[5] = biB;
    __syncthreads();
    if (wv == 0) {
        for (int k = 1; k < W; ++k) {
            const int* fk = fold + ((size_t)k * 64 + lane) * 6;
            consider(bsA, bjA, biA, fk[0], fk[1], fk[2]);
            consider(bsB, bjB, biB, fk[3], fk[4], fk[5]);
        }
        if (liveA) {
            a.ends[slotA] = make_int4(bsA, biA, bjA, (int)a.gidx16[slotA]);
            a.item_group[slotA] = gix;
        }
        if (liveB) {
            a.ends[slotB] = make_int4(bsB, biB, bjB, (int)a.gidx16[slotB]);
            a.item_group[slotB] = gix;
        }
    }
    // The group's walks, while its codes are freshest in the caches: wavefront 0 walks the low
    // halves' targets and wavefront 1 the high halves' (wavefront 0 both when it is alone); the
    // others end here, so the CU can start another group beside the walks. The barrier makes
    // wavefront 0's end cells visible to wavefront 1.
    __syncthreads();
    if (wv >= 2) return;
    for (int h = W == 1 ? 0 : wv; h < 2; h += W == 1 ? 1 : 2) {
        if (h == 0 ? liveA : liveB)
            walk_item<MODE, true>(a, (h == 0 ? slotA : slotB) - a.slot0);
    }
}

template <int R, bool LDSMAT>
hipError_t launch_trace_dp(const TraceArgs& a, hipStream_t st) {
    const size_t lds = LDSMAT ? (size_t)(a.al_eff + 1) * a.al_eff * sizeof(int) : 0;
    hipLaunchKernelGGL((k_trace_dp<HY_KMODE, R, LDSMAT>), dim3(a.nitems), dim3(64), lds, st, a);
    return hipGetLastError();
}

template <int R>
hipError_t launch_trace_dp_r(const TraceArgs& a, hipStream_t st) {
    if ((a.al_eff + 1) * a.al_eff <= 16384) return launch_trace_dp<R, true>(a, st);
    return launch_trace_dp<R, false>(a, st);
}

}  // namespace

template <int MODE>
hipError_t launch_trace_m(const TraceArgs& a, hipStream_t st) {
    static_assert(MODE == HY_KMODE, "one mode per translation unit");
    if (a.nitems == 0) return hipSuccess;
    if (a.rows != trace_rows(a.m)) return hipErrorInvalidValue;
    hipError_t e;
    switch (a.rows) {
    case 2: e = launch_trace_dp_r<2>(a, st); break;
    case 4: e = launch_trace_dp_r<4>(a, st); break;
    case 6: e = launch_trace_dp_r<6>(a, st); break;
    default: e = launch_trace_dp_r<8>(a, st); break;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_trace_walk<HY_KMODE, false>), dim3((a.nitems + 255) / 256), dim3(256),
                       0, st, a);
    return hipGetLastError();
}

template hipError_t launch_trace_m<HY_KMODE>(const TraceArgs&, hipStream_t);

// Packed traceback: the DP over groups [group_first, group_first + ngroups) of the packed layout
// (a.rows wavefronts per workgroup), each workgroup then walking its group's targets (slots
// slot0 .. slot0 + nitems - 1 over the launch).
template <int MODE>
hipError_t launch_trace16_m(const TraceArgs& a, hipStream_t st) {
    static_assert(MODE == HY_KMODE, "one mode per translation unit");
    if (a.ngroups == 0) return hipSuccess;
    const int W = a.rows;
    if (W < 1 || W > kTrace16MaxWaves) return hipErrorInvalidValue;
    const uint32_t pp_words = trace16_pp_words(a.al_eff);
    const size_t lds = trace16_lds_bytes(a.al_eff, W);
    if (lds > 64u * 1024u) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_trace16<MODE>), dim3(a.ngroups), dim3(64 * W), lds, st, a, pp_words);
    return hipGetLastError();
}

template hipError_t launch_trace16_m<HY_KMODE>(const TraceArgs&, hipStream_t);

}  // namespace hy
