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

#if HY_KMODE == 0
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

#if HY_KMODE == 0
// Rows per lane as for the long-target kernel: 2, 4, 6 or 8 (several passes beyond 512 rows).
hipError_t launch_posmax_sw(const PosMaxArgs& a, hipStream_t st) {
    if (a.m <= 128) return launch_posmax_r<2>(a, st);
    if (a.m <= 256) return launch_posmax_r<4>(a, st);
    if (a.m <= 384) return launch_posmax_r<6>(a, st);
    return launch_posmax_r<8>(a, st);
}
#endif

}  // namespace hy
