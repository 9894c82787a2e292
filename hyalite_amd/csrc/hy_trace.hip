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
    uint32_t gi = a.first + item;
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
    uint32_t* codes = a.codes + (a.slab_off ? a.slab_off[gi] - a.slab_base : 0ull);
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

template <int MODE>
__global__ __launch_bounds__(256) void k_trace_walk(TraceArgs a) {
    using F = MF<MODE>;
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= a.nitems) return;
    const int4 en = a.ends[k];
    const uint32_t gi = (uint32_t)en.w;
    int i = en.y, j = en.z;
    const uint8_t* t = a.flat + a.off[gi];
    const int n = (int)(a.off[gi + 1] - a.off[gi]);
    const int m = a.m, R = a.rows, P = 64 * a.rows, S = n + 63;
    const uint32_t* codes = a.codes + (a.slab_off ? a.slab_off[gi] - a.slab_base : 0ull);
    const uint64_t obase = a.winner_key ? 0ull : (uint64_t)gi * (uint64_t)m + a.off[gi];
    uint8_t* ops = a.ops + obase;
    uint64_t w = (uint64_t)m + (uint64_t)n;
    auto nib = [&](int ci, int cj) -> uint32_t {
        const int ii = ci - 1;
        const int p = ii / P, rem = ii - p * P, l = rem / R, r = rem - l * R;
        const uint32_t word = codes[((size_t)p * S + (size_t)(cj - 1 + l)) * 64 + l];
        return (word >> (4 * r)) & 15u;
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
    hipLaunchKernelGGL((k_trace_walk<HY_KMODE>), dim3((a.nitems + 255) / 256), dim3(256), 0, st,
                       a);
    return hipGetLastError();
}

template hipError_t launch_trace_m<HY_KMODE>(const TraceArgs&, hipStream_t);

}  // namespace hy
