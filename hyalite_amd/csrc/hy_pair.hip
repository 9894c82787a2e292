// Pairwise kernels (src/kernel.rs:331-575), compiled once per alignment mode (HY_KMODE).
//
//  * k_pairs32 — align_pairs: one independent (query, target) pair per lane. It is the
//    inter-sequence kernel's recurrence with the query no longer shared by the wavefront: each
//    lane holds its own query symbols for a tile of Q rows in registers (as row offsets into the
//    substitution matrix staged in LDS) and streams its own target from the group's
//    lane-interleaved words, so every cell costs one LDS lookup of (query symbol, target symbol)
//    instead of a shared profile row. Lengths differ per lane: the answer cells of each lane
//    (its last row m, its last column n) are selected per lane.
//  * k_posmax (local mode only) — align_pair_position_max: one pair per wavefront in the systolic
//    shape of the long-target kernel; each lane folds the maximum of its real rows at a column
//    into colmax[j - 1] with an atomic max, and after the sweep the wavefront reduces the array to
//    the best score and its first column.
//
// Single pairs of every mode (align_pair) and the long pairs of a batch run in k_sys32's pair
// mode (one wavefront per pair).
#include "hy_device.h"

#ifndef HY_KMODE
#error "compile hy_pair.hip with -DHY_KMODE=<mode>"
#endif

namespace hy {
namespace {

// Per cell (i, j) of lane l's own pair, with HG = H - go kept per row and the LDS matrix holding
// sub + go (padding symbol row and column kPad32):
//     diag = HG[i-1][j-1] + smat[q_i][t_j],  E = max(HG[i][j-1], E - ge),
//     F = max(HG[i-1][j], F - ge),  H = max(diag, E, F)  (local: max(H, 0)).
// In local mode a padded cell (past the lane's m or n) never exceeds a real cell that precedes it
// in (column, row) order in the same lane, so the strict-greater tile best needs no mask; in the
// other modes the answer cells are selected explicitly.
template <int MODE, int Q, bool ENDS, bool LDSMAT>
__global__ __launch_bounds__(64) void k_pairs32(PairArgs a) {
    using F = MF<MODE>;
    extern __shared__ int smat[];
    const int lane = threadIdx.x;
    const PairGroup g = a.groups[blockIdx.x];
    const bool live = lane < (int)g.count;
    const uint32_t slot = g.first + (live ? lane : 0);
    const int n = live ? (int)a.tlens[slot] : 0;
    const int m = live ? (int)a.qlens[slot] : 0;
    const int L = (int)g.ncols;
    const int go = a.go, ge = a.ge, ale = a.al_eff;
    const int stride = ale + 1;
    if (LDSMAT) {
        for (int k = lane; k < stride * stride; k += 64) {
            const int qs = k / stride, ts = k - qs * stride;
            smat[k] = (qs < ale && ts < ale) ? a.matrix[qs * a.al + ts] + go : kPad32;
        }
        __syncthreads();
    }

    int bs, bj, bi;
    border_best<MODE>(m, n, go, ge, bs, bj, bi);

    const int nchunks = ((int)g.nrows + Q - 1) / Q;
    int2* sp = a.spill + g.spill_off + lane;
    const uint32_t* wp = a.twords + g.tword_off + lane;
    const uint32_t* qp = a.qwords + g.qword_off + lane;

    for (int c = 0; c < nchunks; ++c) {
        const int i0 = c * Q + 1;
        const int rm = m - i0;                   // this lane's row m, when 0 <= rm < Q
        const bool mhere = rm >= 0 && rm < Q;
        int qb[Q];
#pragma unroll
        for (int r4 = 0; r4 < Q / 4; ++r4) {
            const uint32_t qw = qp[(size_t)(c * (Q / 4) + r4) * 64];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int sym = (int)((qw >> (8 * k)) & 0xffu);
                qb[4 * r4 + k] = LDSMAT ? sym * stride : sym;
            }
        }
        int HG[Q], E[Q];
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            HG[r] = wsub(border<MODE>(i0 + r, 0, go, ge), go);
            E[r] = kNeg32;
        }
        int HGd = wsub(border<MODE>(i0 - 1, 0, go, ge), go);
        int cbs = INT_MIN, cbj = 0, cbi = 0;  // tile-local best (local mode, ends)

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
                const int tsym = (int)((w >> (8 * b)) & 0xffu);
                int hgd = HGd;
                HGd = hgup;
                int cmax = INT_MIN;
                int vrow = INT_MIN;
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    int sub;
                    if (LDSMAT) {
                        sub = smat[qb[r] + tsym];
                    } else {
                        sub = (qb[r] < ale && tsym < ale) ? a.matrix[qb[r] * a.al + tsym] + go
                                                          : kPad32;
                    }
                    const int diag = wadd(hgd, sub);
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
                    } else if (F::lrow) {
                        vrow = r == rm ? h : vrow;
                    }
                }
                if (F::local && !ENDS) bs = max(bs, cmax);
                if (c + 1 < nchunks) sp[(size_t)j * 64] = make_int2(HG[Q - 1], f);
                if (!F::local && mhere && j <= n) {
                    if (F::lrow) {
                        if (ENDS) consider(bs, bj, bi, vrow, j, m);
                        else bs = max(bs, vrow);
                    } else if (j == n) {
                        int v = HG[0];
#pragma unroll
                        for (int r = 1; r < Q; ++r) v = r == rm ? HG[r] : v;
                        v = wadd(v, go);
                        if (ENDS) consider(bs, bj, bi, v, n, m);
                        else bs = max(bs, v);
                    }
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

    if (live) {
        const uint32_t pi = a.pidx[slot];
        a.out_score[pi] = bs;
        if (ENDS) {
            a.out_qend[pi] = bi == 0 ? -1 : (int64_t)bi - 1;
            a.out_tend[pi] = bj == 0 ? -1 : (int64_t)bj - 1;
        }
    }
}

template <int Q, bool ENDS, bool LDSMAT>
hipError_t launch_pairs_q(const PairArgs& a, uint32_t ngroups, hipStream_t st) {
    const size_t lds = LDSMAT ? (size_t)(a.al_eff + 1) * (a.al_eff + 1) * sizeof(int) : 0;
    hipLaunchKernelGGL((k_pairs32<HY_KMODE, Q, ENDS, LDSMAT>), dim3(ngroups), dim3(64), lds, st,
                       a);
    return hipGetLastError();
}

template <int Q>
hipError_t launch_pairs_qe(bool ends, const PairArgs& a, uint32_t ngroups, hipStream_t st) {
    const bool ldsmat = (a.al_eff + 1) * (a.al_eff + 1) <= 16384;
    if (ends) {
        return ldsmat ? launch_pairs_q<Q, true, true>(a, ngroups, st)
                      : launch_pairs_q<Q, true, false>(a, ngroups, st);
    }
    return ldsmat ? launch_pairs_q<Q, false, true>(a, ngroups, st)
                  : launch_pairs_q<Q, false, false>(a, ngroups, st);
}

// ------------------------------------------------------------------------------------------------
// Band kernel: one pair per workgroup of W wavefronts.
// ------------------------------------------------------------------------------------------------
// A single wavefront sweeping a pair is latency-bound: the R rows of a lane form one dependent
// chain per step (F runs down the rows), and every step waits on LDS. Here W wavefronts split
// each pass of 64*R*W rows into bands of 64*R rows. Wavefront w runs the systolic sweep of its
// band kBandLag steps behind wavefront w-1; the bands land on different SIMDs, and with several
// per SIMD their chains interleave. Inside a step nothing waits on memory: a lane's residue is
// its left neighbour's from the step before (DPP), lane 0's comes from registers loaded two
// 64-step epochs ahead (readlane), and the matrix entries and lane 0's top value for step s + 1
// are fetched during step s.
//
// Lane 0's top values come from an LDS ring per wavefront: ring[w] for w >= 1 holds the bottom
// row of wavefront w-1 (written by its lane 63), ring[0] the pass's top row (the DP border, or
// the previous pass's bottom row), which wavefront 0 refills once per epoch, an epoch ahead.
//
// Timing: at local step s lane l works column s - l + 1, so the bottom lane of wavefront w-1
// writes column j at its step j + 62, and lane 0 of wavefront w fetches it one step early, at
// its step j - 2. On a global clock G wavefront w runs local step G - w*kBandLag, from local
// step -1 (fetch only), and every wavefront takes a workgroup barrier after each G = 63 mod 64,
// idle or not, so the barriers pair up. With kBandLag = 128 the write (global (w-1)*128 + j +
// 62) and the fetch ((w-1)*128 + j + 126) are 64 steps apart, so a barrier always falls between
// them, and the slot is rewritten (column j + 256) only after two more barriers. kBandLag is a
// multiple of 64, so every wavefront's epoch ends (local s = 63 mod 64) are barrier steps.
constexpr int kBandLag = 128;

__device__ __forceinline__ int residue_or0(const uint8_t* t, int idx, int n) {
    return idx < n ? (int)t[idx] : 0;
}

// Top row of a pass at column col: the DP border (first pass) or the previous pass's bottom row.
template <int MODE>
__device__ __forceinline__ int2 pass_top(const int2* sp, bool first, int col, int n, int go,
                                         int ge) {
    if (first) return make_int2(wsub(border<MODE>(0, col, go, ge), go), kNeg32);
    return col <= n ? sp[col] : make_int2(0, kNeg32);
}

// Lean single-pass sweep of a local-mode pair by one wavefront (the query's rows fit one pass of
// 64 R rows; a winner pass of a short query is its usual case). Lane l holds rows lR+1 .. lR+R and
// at step s works column j = s - l + 1. Against the general sweep below it drops every per-step
// test and hand-off it can:
//  * no column-range mask: columns j < 1 read the padding symbol (sub = kPad32), for which the
//    local recurrence recomputes the left border (H = 0, H - go = E = F = -go) from the border
//    tops every lane starts with, and columns j > n read it too, so their cells come from real
//    cells through gaps only, never above them (strict > keeps the first real column on ties);
//  * lane 0's top input is the DP border (H = 0, F = -inf), held in the lane-0 slot of the
//    top registers, which the lane shift never writes; no ring, no barriers;
//  * the residue of column j comes from a 256-entry ring in LDS (slot x & 255 holds residue x,
//    the padding symbol outside 0..n-1), refilled three epochs of 64 steps ahead, read two
//    steps ahead; the substitution scores of step s + 1 are read during step s.
// smat: (al + 1) x (al + 1) entries, the matrix plus go with row and column al the padding.
// Rows past m record nothing (their row best starts at INT_MAX), and a row best starts at 0:
// score-0 ends are the border's (0, 0, 0), which wins every tie.
//
// KNOWN (a winner pass, whose best score S > 0 the scan has already found): a row records the
// first column where H reaches S, which needs no running maximum, and the sweep stops early.
// A first hit in lane l at step s is at column s - l + 1; the answer's column is at most that
// of any hit, so once the last lane holding query rows (L) has passed the smallest column hit
// so far, every cell that can be the answer has been swept. The hitting lane with the largest
// index holds that column (the ballot's last set bit), so the bound needs no reduction.
template <int MODE, int R, bool ENDS, bool KNOWN>
__device__ __forceinline__ void band_lean(const uint8_t* query, const uint8_t* t, int m, int n,
                                          int al, int go, int ge, const int* smat, uint8_t* tres,
                                          int S, int& bs, int& bj, int& bi) {
    static_assert(MF<MODE>::local, "the lean sweep relies on the local clamp at 0");
    static_assert(!KNOWN || ENDS, "a known score only shortens an end search");
    const int lane = threadIdx.x & 63;
    const int A1 = al + 1;
    const int r0 = lane * R + 1;
    const int ngo = wsub(0, go);
    int qb[R], HG[R], E[R], rb[R], rj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = r0 + r;
        qb[r] = (i <= m ? (int)query[i - 1] : al) * A1;
        HG[r] = ngo;
        E[r] = kNeg32;
        // KNOWN: the value a row looks for (rows past m look for none: H is never negative).
        rb[r] = KNOWN ? (i <= m ? S : -1) : (i <= m ? 0 : INT_MAX);
        rj[r] = INT_MAX;
    }
    auto res = [&](int x) HY_INLINE { return (x >= 0 && x < n) ? (int)t[x] : al; };
    // Epochs 0-2 and the padding before column 1 (slots 192-255 stand for x = -64 .. -1).
    tres[lane] = (uint8_t)res(lane);
    tres[64 + lane] = (uint8_t)res(64 + lane);
    tres[128 + lane] = (uint8_t)res(128 + lane);
    tres[192 + lane] = (uint8_t)al;
    int rf = res(192 + lane);  // epoch 3, stored at the end of epoch 0
    const int last = (m + R - 1) / R - 1;  // the last lane holding query rows
    const int steps = n + last;
    int stop = steps;
    // Registers alternate between even and odd steps so that nothing is copied between steps
    // (a copy of a register with an LDS read in flight would wait for the read): the tops of
    // step s sit in T[s & 1] and its diagonal input in T[(s + 1) & 1], the lane shift writes
    // step s + 1's tops over the latter; substitution scores and residues ping-pong the same way.
    int TA = ngo, TB = ngo, topF = kNeg32;
    int jv = 1 - lane;
    int subA[R], subB[R];
    {
        const int c0 = tres[(0 - lane) & 255];
#pragma unroll
        for (int r = 0; r < R; ++r) subA[r] = smat[qb[r] + c0];
    }
    int cA = 0, cB = tres[(1 - lane) & 255];
    // One step s: sub holds its scores, cuse the residue of step s + 1 (its scores go to subn),
    // cload receives the residue of step s + 2; top and diag as above.
    auto step = [&](int s, const int* sub, int* subn, int cuse, int& cload, int& top,
                    int& diagin) HY_INLINE {
        cload = tres[(s + 2 - lane) & 255];
#pragma unroll
        for (int r = 0; r < R; ++r) subn[r] = smat[qb[r] + cuse];
        int f = topF, hgup = top, dg = diagin;
        int hv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int diag = wadd(dg, sub[r]);
            const int e = max(HG[r], wsub(E[r], ge));
            f = max(hgup, wsub(f, ge));
            const int h = max(max(max(diag, e), f), 0);
            dg = HG[r];
            E[r] = e;
            HG[r] = wsub(h, go);
            hgup = HG[r];
            if (KNOWN) {
                hv[r] = h;
            } else if (ENDS) {
                const bool up = h > rb[r];
                rb[r] = up ? h : rb[r];
                rj[r] = up ? jv : rj[r];
            } else {
                bs = max(bs, h);
            }
        }
        if (KNOWN) {
            // Lanes with a row at S: one wavefront-wide compare mask per row (v_cmp into SGPRs).
            unsigned long long hits = 0ull;
#pragma unroll
            for (int r = 0; r < R; ++r) hits |= __builtin_amdgcn_sicmp(hv[r], rb[r], 32);
            if (hits != 0ull) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    rj[r] = (hv[r] == rb[r] && rj[r] == INT_MAX) ? jv : rj[r];
                const int lmax = 63 - __builtin_clzll(hits);
                stop = min(stop, s + 1 - lmax + last);
            }
        }
        // Step s + 1's tops: lane l takes lane l-1's bottom row; lane 0 keeps the border.
        diagin = __builtin_amdgcn_update_dpp(diagin, HG[R - 1], 0x138, 0xf, 0xf, false);
        topF = __builtin_amdgcn_update_dpp(topF, f, 0x138, 0xf, 0xf, false);
        ++jv;
        if ((s & 63) == 63) {
            // Epoch e = s >> 6 ends: epoch e + 3 replaces epoch e - 1, the next is loaded.
            tres[(s + 129 + lane) & 255] = (uint8_t)rf;
            rf = res(s + 193 + lane);
        }
    };
    int s = 0;
    for (; s + 1 < stop; s += 2) {
        step(s, subA, subB, cB, cA, TA, TB);
        step(s + 1, subB, subA, cA, cB, TB, TA);
    }
    if (s < stop) step(s, subA, subB, cB, cA, TA, TB);
    if (ENDS) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (rj[r] != INT_MAX) consider(bs, bj, bi, rb[r], rj[r], r0 + r);
    }
}

template <int MODE, int R, int W, int OUT>
__global__ __launch_bounds__(64 * W) void k_band32(BandArgs a) {
    using F = MF<MODE>;
    constexpr bool ENDS = OUT == kBandEnds;
    constexpr bool COLMAX = OUT == kBandColmax;
    // Local-mode ends: each row keeps its first best column (ties in a row go to the smaller
    // column, the lane's rows are merged with the tie-break at the end).
    constexpr bool ROWBEST = ENDS && F::local;
    static_assert(!COLMAX || MODE == kSW, "column maxima are a local-mode output");
    extern __shared__ int smat[];
    __shared__ int2 ring[W][256];
    __shared__ int red[3][W];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool winner = a.winner_key != nullptr;
    const unsigned long long wk = winner ? *a.winner_key : 0ull;
    const uint32_t pid = winner ? (wk != 0ull ? key_index(wk) : 0u) : a.list[blockIdx.x];
    const uint8_t* query = winner ? a.queries : a.queries + a.qoff[pid];
    const uint8_t* t = a.targets + a.toff[pid];
    const int n = (int)(a.toff[pid + 1] - a.toff[pid]);
    const int m = winner ? a.wm : (int)a.qlen[pid], go = a.go, ge = a.ge, al = a.al;
    constexpr int P = 64 * R * W;
    const int npass = n > 0 ? (m + P - 1) / P : 0;
    // One wavefront with the whole query in one pass, local mode: the lean sweep (band_lean).
    constexpr bool LEAN_OK = W == 1 && F::local && !COLMAX;
    const bool lean = LEAN_OK && npass == 1;
    if (lean) {
        // The matrix plus go, row and column al the padding symbol.
        for (int k = threadIdx.x; k < (al + 1) * (al + 1); k += 64 * W) {
            const int r = k / (al + 1), c = k - r * (al + 1);
            smat[k] = r < al && c < al ? a.matrix[r * al + c] + go : kPad32;
        }
    } else {
        // The matrix plus go, row al the padding row for query rows past m.
        for (int k = threadIdx.x; k < (al + 1) * al; k += 64 * W)
            smat[k] = k < al * al ? a.matrix[k] + go : kPad32;
    }
    if (COLMAX) {
        for (int k = threadIdx.x; k < n; k += 64 * W) a.colmax[k] = 0;
        __threadfence();  // the zeros reach the device-scope atomics below
    }
    __syncthreads();

    int bs = INT_MIN, bj = INT_MAX, bi = INT_MAX;
    if (!COLMAX && threadIdx.x == 0) border_best<MODE>(m, n, go, ge, bs, bj, bi);

    if constexpr (LEAN_OK) {
        uint8_t* tres = reinterpret_cast<uint8_t*>(&ring[0][0]);
        const int S = (ENDS && winner && wk != 0ull) ? key_score(wk) : 0;
        if (lean && S > 0) {
            if constexpr (ENDS)
                band_lean<MODE, R, true, true>(query, t, m, n, al, go, ge, smat, tres, S, bs, bj,
                                               bi);
        } else if (lean) {
            band_lean<MODE, R, ENDS, false>(query, t, m, n, al, go, ge, smat, tres, 0, bs, bj,
                                            bi);
        }
    }
    int2* sp = a.spill + (size_t)blockIdx.x * a.spill_stride;
    for (int p = 0; p < (lean ? 0 : npass); ++p) {
        const bool first = p == 0;
        // Lane l takes column n at step n + l - 1. A lone wavefront's last pass ends when its
        // last lane holding query rows has (no other wavefront shares its barriers, and no later
        // pass needs lane 63's bottom row); otherwise every lane runs to the end.
        int steps = n + 63;
        if (W == 1 && p + 1 == npass) steps = n + (min(m - p * P, P) + R - 1) / R - 1;
        const int r0 = p * P + (w * 64 + lane) * R + 1;
        int qb[R], HG[R], E[R], sub[R], rb[R], rj[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r0 + r;
            qb[r] = (i <= m ? (int)query[i - 1] : al) * al;
            HG[r] = wsub(border<MODE>(i, 0, go, ge), go);
            E[r] = kNeg32;
            sub[r] = 0;
            rb[r] = i <= m ? INT_MIN : INT_MAX;  // rows past m never record a best
            rj[r] = INT_MAX;
        }
        int HGd = wsub(border<MODE>(r0 - 1, 0, go, ge), go);
        int topHG = 0, topF = kNeg32, c = 0;
        const bool anyrow = r0 <= m;
        // Wavefront 0 seeds its ring with the top row's columns 1-192 (read up to epoch 1);
        // afterwards the epoch ending at local step s = 64e + 63 stores the columns of epoch
        // e + 2 from tf, loaded at the end of the epoch before.
        int2 tf = make_int2(0, kNeg32);
        if (w == 0) {
            for (int k = lane; k < 192; k += 64)
                ring[0][(k + 1) & 255] = pass_top<MODE>(sp, first, k + 1, n, go, ge);
            tf = pass_top<MODE>(sp, first, 66 + lane, n, go, ge);
        }
        // Residue epoch registers: lane k holds residue 64e + k of the current epoch (rc), the
        // next (rn) and the one after (rf). Local step -1 closes epoch -1.
        int rc = 0, rn = residue_or0(t, lane, n), rf = residue_or0(t, 64 + lane, n);
        __syncthreads();

        // Idle lead-in: the barriers of global steps -1 .. w*kBandLag - 2.
        for (int k = 0; k < 2 * w; ++k) __syncthreads();
        for (int s = -1; s < steps; ++s) {
            // Fetch for step s + 1: residues (lane 0 from the epoch registers, the others from
            // their left neighbour), matrix entries, and lane 0's top (column s + 2).
            const int k1 = (s + 1) & 63;
            const int rl = __builtin_amdgcn_readlane(k1 == 0 ? rn : rc, k1);
            const int cn = __builtin_amdgcn_update_dpp(rl, c, 0x138, 0xf, 0xf, false);
            int subn[R];
#pragma unroll
            for (int r = 0; r < R; ++r) subn[r] = smat[qb[r] + cn];
            const int2 tv = ring[w][(s + 2) & 255];

            const int j = s - lane + 1;
            int outHG = 0, outF = 0;
            if (j >= 1 && j <= n) {
                int f = topF, hgup = topHG, hgd = HGd;
                HGd = topHG;
                int cm = 0;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int diag = wadd(hgd, sub[r]);
                    const int e = max(HG[r], wsub(E[r], ge));
                    f = max(hgup, wsub(f, ge));
                    int h = max(max(diag, e), f);
                    if (F::local) h = max(h, 0);
                    hgd = HG[r];
                    E[r] = e;
                    HG[r] = wsub(h, go);
                    hgup = HG[r];
                    const int i = r0 + r;
                    if (COLMAX) {
                        // Rows past m would carry real values forward along the gap chains.
                        cm = i <= m ? max(cm, h) : cm;
                    } else if (ROWBEST) {
                        const bool up = h > rb[r];
                        rb[r] = up ? h : rb[r];
                        rj[r] = up ? j : rj[r];
                    } else {
                        const bool cand = i <= m && (F::local || (F::lcol && j == n) ||
                                                     (i == m && (F::lrow || j == n)));
                        const int v = cand ? h : INT_MIN;
                        if (ENDS) consider(bs, bj, bi, v, j, i);
                        else bs = max(bs, v);
                    }
                }
                outHG = HG[R - 1];
                outF = f;
                if (lane == 63) {
                    if (w + 1 < W) ring[w + 1][j & 255] = make_int2(outHG, outF);
                    else if (p + 1 < npass) sp[j] = make_int2(outHG, outF);
                }
                if (COLMAX && anyrow && cm > 0) atomicMax(a.colmax + (j - 1), cm);
            }
            // Step s + 1's inputs: lane l takes lane l-1's bottom row (wave_shr:1), lane 0 the
            // fetched top.
            topHG = __builtin_amdgcn_update_dpp(tv.x, outHG, 0x138, 0xf, 0xf, false);
            topF = __builtin_amdgcn_update_dpp(tv.y, outF, 0x138, 0xf, 0xf, false);
            c = cn;
#pragma unroll
            for (int r = 0; r < R; ++r) sub[r] = subn[r];
            if ((s & 63) == 63) {
                // Epoch e = s >> 6 ends: s + 1 = 64(e + 1).
                const int s1 = s + 1;
                rc = rn;
                rn = rf;
                rf = residue_or0(t, s1 + 128 + lane, n);  // epoch e + 3
                if (w == 0) {
                    // The top row read during epoch e + 2 (columns 64(e+2) + 2 .. + 65).
                    ring[0][(s1 + 66 + lane) & 255] = tf;
                    tf = pass_top<MODE>(sp, first, s1 + 130 + lane, n, go, ge);
                }
                __syncthreads();
            }
        }
        // Idle tail: the barriers of the global steps after this wavefront's last.
        for (int k = 0; k < 2 * (W - 1 - w); ++k) __syncthreads();
        if (ROWBEST) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (rb[r] != INT_MAX) consider(bs, bj, bi, rb[r], rj[r], r0 + r);
        }
        __threadfence_block();
        __syncthreads();
    }

    if (COLMAX) {
        // Every atomic of the workgroup is done; the best score (>= 0) and its smallest column.
        __threadfence();
        __syncthreads();
        int best = 0, bc = INT_MAX;
        for (int k = threadIdx.x; k < n; k += 64 * W) {
            const int v =
                __hip_atomic_load(a.colmax + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (v > best) { best = v; bc = k; }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const int ob = __shfl_xor(best, off), oc = __shfl_xor(bc, off);
            const bool take = ob > best || (ob == best && oc < bc);
            best = take ? ob : best;
            bc = take ? oc : bc;
        }
        if (lane == 0) { red[0][w] = best; red[1][w] = bc; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 1; k < W; ++k) {
                const bool take = red[0][k] > best || (red[0][k] == best && red[1][k] < bc);
                best = take ? red[0][k] : best;
                bc = take ? red[1][k] : bc;
            }
            *a.best = make_int2(best, best > 0 ? bc : -1);
        }
        return;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int os = __shfl_xor(bs, off);
        if (ENDS) consider(bs, bj, bi, os, __shfl_xor(bj, off), __shfl_xor(bi, off));
        else bs = max(bs, os);
    }
    if (lane == 0) { red[0][w] = bs; red[1][w] = bj; red[2][w] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < W; ++k) {
            if (ENDS) consider(bs, bj, bi, red[0][k], red[1][k], red[2][k]);
            else bs = max(bs, red[0][k]);
        }
        if (winner) {
            Result* r = a.result;
            r->score = bs;
            r->db_index = pid;
            r->query_end = ENDS ? (bi == 0 ? -1 : (int64_t)bi - 1) : -1;
            r->target_end = ENDS ? (bj == 0 ? -1 : (int64_t)bj - 1) : -1;
            r->check = (bs == key_score(wk) && wk != 0ull) ? 0 : 1;
            r->pad = 0;
            return;
        }
        a.out_score[pid] = bs;
        if (ENDS) {
            a.out_qend[pid] = bi == 0 ? -1 : (int64_t)bi - 1;
            a.out_tend[pid] = bj == 0 ? -1 : (int64_t)bj - 1;
        }
    }
}

template <int R, int W, int OUT>
hipError_t launch_band_rwo(const BandArgs& a, uint32_t nblocks, hipStream_t st) {
    const size_t lds = (size_t)(a.al + 1) * (a.al + 1) * sizeof(int);
    hipLaunchKernelGGL((k_band32<HY_KMODE, R, W, OUT>), dim3(nblocks), dim3(64 * W), lds, st, a);
    return hipGetLastError();
}

// Shapes (rows per lane R, wavefronts W) the launcher can pick; band_shape() chooses.
template <int OUT>
hipError_t launch_band_o(const BandArgs& a, uint32_t nblocks, int m, hipStream_t st) {
    const BandShape sh = band_shape(m);
    switch (sh.r * 100 + sh.w) {
    case 101: return launch_band_rwo<1, 1, OUT>(a, nblocks, st);
    case 102: return launch_band_rwo<1, 2, OUT>(a, nblocks, st);
    case 104: return launch_band_rwo<1, 4, OUT>(a, nblocks, st);
    case 201: return launch_band_rwo<2, 1, OUT>(a, nblocks, st);
    case 202: return launch_band_rwo<2, 2, OUT>(a, nblocks, st);
    case 204: return launch_band_rwo<2, 4, OUT>(a, nblocks, st);
    case 208: return launch_band_rwo<2, 8, OUT>(a, nblocks, st);
    case 216: return launch_band_rwo<2, 16, OUT>(a, nblocks, st);
    case 401: return launch_band_rwo<4, 1, OUT>(a, nblocks, st);
    case 402: return launch_band_rwo<4, 2, OUT>(a, nblocks, st);
    case 404: return launch_band_rwo<4, 4, OUT>(a, nblocks, st);
    case 408: return launch_band_rwo<4, 8, OUT>(a, nblocks, st);
    case 416: return launch_band_rwo<4, 16, OUT>(a, nblocks, st);
    default: return hipErrorInvalidValue;
    }
}

#if HY_KMODE == 0
// ------------------------------------------------------------------------------------------------
// Local-mode pairs, two per lane in packed 16-bit arithmetic.
// ------------------------------------------------------------------------------------------------
// Lane l aligns slot first + 2l in the low half and first + 2l + 1 in the high half of every
// register. With S = al + 1 (symbol al pads), small alphabets (WIDE = false) take both halves'
// substitution scores from one LDS lookup: the table T[(qa*S + qb)*S^2 + ta*S + tb] holds
// (sub(qa, ta), sub(qb, tb)); a lane keeps the byte offsets of its tile's query-row pairs in
// registers and reads the column's target pair from its words. Larger alphabets (WIDE = true)
// look each half up in an S x S score table kept in C = kPairs16Copies interleaved copies (slot
// e*C + c), lane l reading copy l mod C so that lookups spread over the LDS banks: a query word
// holds the two rows' byte offsets 4*C*S*qa and 4*C*S*qb, a target column the symbols ta and
// tb. WIDE workgroups hold four wavefronts (four groups) sharing one table.
// The recurrence is the packed local one of the scan kernels, every quantity
// floored at 0 (H exact, E' = max(E, 0), F' = max(F, 0)):
//     H = max(H[i-1][j-1] + sub, E', F'),  E' = max(H - go, E' - ge)+,  F' = max(H - go, F' - ge)+
// so a cell pair costs the lookup and 9 packed operations. Cells past a pair's own query or
// target (padding, -32768) never exceed a real cell that precedes them in (column, row) order,
// so neither the running maximum nor the strictly-greater end tracking needs a mask. With ends,
// each column keeps its maximum and the first row reaching it; a column that strictly beats the
// tile's best replaces it; each tile's best is folded into the lane's with the full tie-break.
template <int Q, bool ENDS, bool WIDE>
__global__ __launch_bounds__(WIDE ? 256 : 64) void k_pairs16(Pair16Args a) {
    extern __shared__ uint32_t tab[];
    const int lane = threadIdx.x & 63;
    const uint32_t gi = WIDE ? blockIdx.x * 4 + (threadIdx.x >> 6) : blockIdx.x;
    const int al = a.al, S = al + 1, S2 = S * S, go = a.go, ge = a.ge;
    constexpr int C = kPairs16Copies;
    if constexpr (WIDE) {
        for (int k = threadIdx.x; k < S2 * C; k += 256) {
            const int q = k / C / S, t = k / C % S;
            tab[k] = (uint16_t)(q < al && t < al ? sat16(a.matrix[q * al + t]) : -32768);
        }
    } else {
        for (int k = lane; k < S2 * S2; k += 64) {
            const int qa = k / (S2 * S), qb = (k / S2) % S, ta = (k / S) % S, tb = k % S;
            const int lo = (qa < al && ta < al) ? sat16(a.matrix[qa * al + ta]) : -32768;
            const int hi = (qb < al && tb < al) ? sat16(a.matrix[qb * al + tb]) : -32768;
            tab[k] = as_u32(pair2(lo, hi));
        }
    }
    __syncthreads();
    if (gi >= a.ngroups) return;  // no barrier below
    const PairGroup g = a.groups[gi];
    const uint8_t* tb8 = reinterpret_cast<const uint8_t*>(tab);
    // WIDE: the lane's copy, added to a column's symbol offsets.
    const uint32_t copy4 = 4u * (uint32_t)(lane % C);

    const s2v go2 = splat2(go), ge2 = splat2(ge), zero2 = splat2(0), one2 = splat2(1);
    const int L = (int)g.ncols;
    const int ntiles = ((int)g.nrows + Q - 1) / Q;
    uint2* sp = a.spill + g.spill_off + lane;
    // Target words: four columns per lane and step, one word (bytes) or a uint2 (WIDE: 16 bits
    // per column).
    const uint32_t* wp = a.twords + g.tword_off + lane;
    const uint2* wp2 = reinterpret_cast<const uint2*>(a.twords + g.tword_off) + lane;
    const uint32_t* qp = a.qwords + g.qword_off + lane;
    // Lane bests per half (score, column, row); (0, 0, 0) is the empty alignment.
    int bs[2] = {0, 0}, bj[2] = {0, 0}, bi[2] = {0, 0};
    s2v runq[4] = {zero2, zero2, zero2, zero2};

    for (int c = 0; c < ntiles; ++c) {
        const int i0 = c * Q + 1;
        uint32_t qrow[Q];
        s2v HG[Q], E[Q];
        s2v tb = zero2, tc = zero2, tr = zero2;  // tile best, its column and row in the tile
        auto cell = [&](int r, s2v sg, s2v& hd, s2v& f, s2v& cm, s2v& cr) HY_INLINE {
            const s2v h = p_max(p_max(p_add(hd, sg), E[r]), f);
            hd = HG[r];
            HG[r] = h;
            const s2v hg = p_subu(h, go2);
            E[r] = p_max(hg, p_subu(E[r], ge2));
            f = p_max(hg, p_subu(f, ge2));
            if constexpr (ENDS) {
                // cr = r where h beats the column's maximum so far: a clamped-multiply mask and
                // a bit select (the latest strict improvement is the first row reaching it).
                const uint32_t up = p_nonzero_mask(p_subu(h, cm));
                cm = p_max(cm, h);
                cr = as_s2((up & as_u32(splat2(r))) | (~up & as_u32(cr)));
            } else {
                runq[r & 3] = p_max(runq[r & 3], h);
            }
        };
        auto close_column = [&](int j, s2v cm, s2v cr) HY_INLINE {
            const s2v up = p_minu_pk(p_subu(cm, tb), one2);
            tb = p_max(tb, cm);
            tc = p_maxu(tc, p_madu(up, splat2(j), zero2));
            tr = p_madu(up, p_subw(cr, tr), tr);
        };
        if constexpr (WIDE) {
#pragma unroll
            for (int r = 0; r < Q; ++r) qrow[r] = qp[(size_t)(c * Q + r) * 64];
        } else {
#pragma unroll
            for (int r2 = 0; r2 < Q / 2; ++r2) {
                const uint32_t qw = qp[(size_t)(c * (Q / 2) + r2) * 64];
                qrow[2 * r2] = qw & 0xffffu;
                qrow[2 * r2 + 1] = qw >> 16;
            }
        }
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            HG[r] = zero2;
            E[r] = zero2;
        }
        s2v HGd = zero2;

        uint2 w = make_uint2(0u, 0u);
        if (L > 0) w = WIDE ? wp2[0] : make_uint2(wp[0], 0u);
        uint2 spv[4];
#pragma unroll
        for (int b = 0; b < 4; ++b)
            spv[b] = (c > 0 && b + 1 <= L) ? sp[(size_t)(b + 1) * 64] : make_uint2(0u, 0u);
        for (int jb = 0; jb < L; jb += 4) {
            uint2 wn = make_uint2(0u, 0u);
            uint2 spn[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) spn[b] = make_uint2(0u, 0u);
            if (jb + 4 < L) {
                if constexpr (WIDE) wn = wp2[(size_t)((jb >> 2) + 1) * 64];
                else wn.x = wp[(size_t)((jb >> 2) + 1) * 64];
                if (c > 0) {
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        if (jb + 5 + b <= L) spn[b] = sp[(size_t)(jb + 5 + b) * 64];
                }
            }
            // Columns in pairs (j, j + 1), skewed by one row: row r of column j and row r - 1
            // of column j + 1 are independent, so two dependency chains interleave. A column
            // past the group's end (odd L) computes padding that never wins and is not stored.
#pragma unroll
            for (int bp = 0; bp < 4; bp += 2) {
                const int jA = jb + bp + 1, jB = jA + 1;
                if (jA > L) break;
                s2v sgA[Q], sgB[Q];
                if constexpr (WIDE) {
                    // Column jA's symbols in w.x (bp = 0) or w.y (bp = 2), column jB's above.
                    const uint32_t wv = bp == 0 ? w.x : w.y;
                    const uint32_t aA = (wv & 0xffu) * (4 * C) + copy4;
                    const uint32_t bA = ((wv >> 8) & 0xffu) * (4 * C) + copy4;
                    const uint32_t aB = ((wv >> 16) & 0xffu) * (4 * C) + copy4;
                    const uint32_t bB = (wv >> 24) * (4 * C) + copy4;
                    auto look = [&](uint32_t q, uint32_t ta, uint32_t tb) HY_INLINE {
                        s2v v;
                        v.x = *reinterpret_cast<const short*>(tb8 + (q & 0xffffu) + ta);
                        v.y = *reinterpret_cast<const short*>(tb8 + (q >> 16) + tb);
                        return v;
                    };
#pragma unroll
                    for (int r = 0; r < Q; ++r) {
                        sgA[r] = look(qrow[r], aA, bA);
                        sgB[r] = look(qrow[r], aB, bB);
                    }
                } else {
                    const uint32_t tpA = 4u * ((w.x >> (8 * bp)) & 0xffu);
                    const uint32_t tpB = 4u * ((w.x >> (8 * bp + 8)) & 0xffu);
#pragma unroll
                    for (int r = 0; r < Q; ++r) {
                        sgA[r] = as_s2(*reinterpret_cast<const uint32_t*>(tb8 + qrow[r] + tpA));
                        sgB[r] = as_s2(*reinterpret_cast<const uint32_t*>(tb8 + qrow[r] + tpB));
                    }
                }
                s2v fA = as_s2(spv[bp].y), fB = as_s2(spv[bp + 1].y);
                s2v hdA = HGd, hdB = as_s2(spv[bp].x);
                HGd = as_s2(spv[bp + 1].x);
                s2v cmA = zero2, crA = zero2, cmB = zero2, crB = zero2;
                uint2 outA = make_uint2(0u, 0u);
#pragma unroll
                for (int r = 0; r <= Q; ++r) {
                    if (r < Q) cell(r, sgA[r], hdA, fA, cmA, crA);
                    if (r == Q - 1) outA = make_uint2(as_u32(HG[Q - 1]), as_u32(fA));
                    if (r >= 1) cell(r - 1, sgB[r - 1], hdB, fB, cmB, crB);
                }
                if (c + 1 < ntiles) {
                    sp[(size_t)jA * 64] = outA;
                    if (jB <= L) sp[(size_t)jB * 64] = make_uint2(as_u32(HG[Q - 1]), as_u32(fB));
                }
                if constexpr (ENDS) {
                    close_column(jA, cmA, crA);
                    close_column(jB, cmB, crB);
                }
            }
            w = wn;
#pragma unroll
            for (int b = 0; b < 4; ++b) spv[b] = spn[b];
        }
        if constexpr (ENDS) {
            const uint32_t ub = as_u32(tb), uc = as_u32(tc), ur = as_u32(tr);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int sc = (int)(short)(uint16_t)(ub >> (16 * h));
                if (sc > 0)
                    consider(bs[h], bj[h], bi[h], sc, (int)((uc >> (16 * h)) & 0xffffu),
                             i0 + (int)((ur >> (16 * h)) & 0xffffu));
            }
        }
    }

    if constexpr (!ENDS) {
        const uint32_t u = as_u32(p_max(p_max(runq[0], runq[1]), p_max(runq[2], runq[3])));
        bs[0] = (int)(short)(uint16_t)u;
        bs[1] = (int)(short)(uint16_t)(u >> 16);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t sl = 2 * lane + h;
        if (sl < g.count) {
            const uint32_t pi = a.pidx[g.first + sl];
            a.out_score[pi] = bs[h];
            if (ENDS) {
                a.out_qend[pi] = bi[h] == 0 ? -1 : (int64_t)bi[h] - 1;
                a.out_tend[pi] = bj[h] == 0 ? -1 : (int64_t)bj[h] - 1;
            }
        }
    }
}

template <int Q, bool ENDS, bool WIDE>
hipError_t launch_pairs16_q(const Pair16Args& a, uint32_t ngroups, hipStream_t st) {
    const size_t S = (size_t)a.al + 1;
    const size_t lds = WIDE ? S * S * kPairs16Copies * 4 : S * S * S * S * 4;
    Pair16Args b = a;
    b.ngroups = ngroups;
    hipLaunchKernelGGL((k_pairs16<Q, ENDS, WIDE>), dim3(WIDE ? (ngroups + 3) / 4 : ngroups),
                       dim3(WIDE ? 256 : 64), lds, st, b);
    return hipGetLastError();
}

// Local mode, one pair per wavefront, R query rows per lane (the long-target kernel's shape).
template <int R, bool LDSMAT>
__global__ __launch_bounds__(64) void k_posmax(PosMaxArgs a) {
    extern __shared__ int smat[];
    __shared__ uint8_t tring[256];
    __shared__ int2 sring[256];
    const int lane = threadIdx.x;
    const uint8_t* t = a.target;
    const int n = a.n, m = a.m, go = a.go, ge = a.ge, ale = a.al_eff;
    const int rowstride = LDSMAT ? ale : a.al;
    if (LDSMAT) {
        for (int k = lane; k < (ale + 1) * ale; k += 64)
            smat[k] = k < ale * ale ? a.matrix[(k / ale) * a.al + (k % ale)] + go : kPad32;
    }
    __syncthreads();
    constexpr int P = 64 * R;
    const int npass = (m + P - 1) / P;
    int2* sp = a.spill;
    for (int p = 0; p < npass; ++p) {
        const int r0 = p * P + lane * R + 1;
        int qb[R], HG[R], E[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r0 + r;
            qb[r] = i <= m ? (int)a.query[i - 1] * rowstride : (LDSMAT ? ale * ale : -1);
            HG[r] = -go;  // H = 0 on the free left border
            E[r] = kNeg32;
        }
        int HGd = -go;
        int topHG = 0, topF = kNeg32;
        const bool anyrow = r0 <= m;
        tring[lane] = lane < n ? t[lane] : (uint8_t)0;
        if (p > 0) sring[lane] = lane + 1 <= n ? sp[lane + 1] : make_int2(0, kNeg32);
        __syncthreads();
        uint8_t tnext = 0;
        int2 snext = make_int2(0, kNeg32);
        const int S = n + 63;
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
                    topHG = -go;  // H = 0 on the free top border
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
                int cm = 0;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int sub = LDSMAT ? smat[qb[r] + c]
                                           : (qb[r] >= 0 ? a.matrix[qb[r] + c] + go : kPad32);
                    const int diag = wadd(hgd, sub);
                    const int e = max(HG[r], wsub(E[r], ge));
                    f = max(hgup, wsub(f, ge));
                    const int h = max(max(max(diag, e), f), 0);
                    hgd = HG[r];
                    E[r] = e;
                    HG[r] = wsub(h, go);
                    hgup = HG[r];
                    // Rows past m would carry real values forward along the gap chains, so they
                    // are kept out of the column maximum.
                    cm = (r0 + r <= m) ? max(cm, h) : cm;
                }
                outHG = HG[R - 1];
                outF = f;
                if (lane == 63 && p + 1 < npass) sp[j] = make_int2(outHG, outF);
                if (anyrow && cm > 0) atomicMax(a.colmax + (j - 1), cm);
            }
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
    // Every atomic of this workgroup is done; reduce the array: the best score (>= 0) and the
    // smallest column holding it (none when the best is 0).
    __threadfence();
    __syncthreads();
    int best = 0, bj = INT_MAX;
    for (int k = lane; k < n; k += 64) {
        const int v = __hip_atomic_load(a.colmax + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v > best) { best = v; bj = k; }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int ob = __shfl_xor(best, off), oj = __shfl_xor(bj, off);
        const bool take = ob > best || (ob == best && oj < bj);
        best = take ? ob : best;
        bj = take ? oj : bj;
    }
    if (lane == 0) *a.best = make_int2(best, best > 0 ? bj : -1);
}

template <int R>
hipError_t launch_posmax_r(const PosMaxArgs& a, hipStream_t st) {
    if ((a.al_eff + 1) * a.al_eff <= 16384) {
        const size_t lds = (size_t)(a.al_eff + 1) * a.al_eff * sizeof(int);
        hipLaunchKernelGGL((k_posmax<R, true>), dim3(1), dim3(64), lds, st, a);
    } else {
        hipLaunchKernelGGL((k_posmax<R, false>), dim3(1), dim3(64), 0, st, a);
    }
    return hipGetLastError();
}
#endif

}  // namespace

template <int MODE>
hipError_t launch_pairs32_m(bool ends, const PairArgs& a, uint32_t ngroups, int q,
                            hipStream_t st) {
    static_assert(MODE == HY_KMODE, "one mode per translation unit");
    if (q <= 16) return launch_pairs_qe<16>(ends, a, ngroups, st);
    return launch_pairs_qe<32>(ends, a, ngroups, st);
}

template hipError_t launch_pairs32_m<HY_KMODE>(bool, const PairArgs&, uint32_t, int, hipStream_t);

template <int MODE>
hipError_t launch_band32_m(int out, const BandArgs& a, uint32_t nblocks, int m, hipStream_t st) {
    static_assert(MODE == HY_KMODE, "one mode per translation unit");
    if (a.al < 1 || a.al > (int)kBandMaxAlphabet) return hipErrorInvalidValue;
    if (out == kBandEnds) return launch_band_o<kBandEnds>(a, nblocks, m, st);
    if (out == kBandScore) return launch_band_o<kBandScore>(a, nblocks, m, st);
#if HY_KMODE == 0
    if (out == kBandColmax) return launch_band_o<kBandColmax>(a, nblocks, m, st);
#endif
    return hipErrorInvalidValue;
}

template hipError_t launch_band32_m<HY_KMODE>(int, const BandArgs&, uint32_t, int, hipStream_t);

#if HY_KMODE == 0
hipError_t launch_pairs16_sw(bool ends, const Pair16Args& a, uint32_t ngroups, int q,
                             hipStream_t st) {
    if (ngroups == 0) return hipSuccess;
    if (a.al < 1 || a.al > (int)kPairs16MaxAlphabet) return hipErrorInvalidValue;
    if (pairs16_wide(a.al)) {
        if (ends) return q <= 16 ? launch_pairs16_q<16, true, true>(a, ngroups, st)
                                 : launch_pairs16_q<32, true, true>(a, ngroups, st);
        return q <= 16 ? launch_pairs16_q<16, false, true>(a, ngroups, st)
                       : launch_pairs16_q<32, false, true>(a, ngroups, st);
    }
    if (ends) return q <= 16 ? launch_pairs16_q<16, true, false>(a, ngroups, st)
                             : launch_pairs16_q<32, true, false>(a, ngroups, st);
    return q <= 16 ? launch_pairs16_q<16, false, false>(a, ngroups, st)
                   : launch_pairs16_q<32, false, false>(a, ngroups, st);
}

// Rows per lane as for the long-target kernel: 2, 4, 6 or 8 (several passes beyond 512 rows).
hipError_t launch_posmax_sw(const PosMaxArgs& a, hipStream_t st) {
    if (a.m <= 128) return launch_posmax_r<2>(a, st);
    if (a.m <= 256) return launch_posmax_r<4>(a, st);
    if (a.m <= 384) return launch_posmax_r<6>(a, st);
    return launch_posmax_r<8>(a, st);
}
#endif

}  // namespace hy
