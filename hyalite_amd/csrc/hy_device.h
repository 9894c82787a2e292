// Device helpers shared by the kernel translation units: wrapping 32-bit arithmetic, the mode
// flags, DP borders, the answer tie-break and the 64-bit argmax key.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>

#include "hy_internal.h"

namespace hy {
namespace {

__device__ __forceinline__ int wadd(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
__device__ __forceinline__ int wsub(int a, int b) { return (int)((unsigned)a - (unsigned)b); }

// Flags::for_mode (src/kernel.rs:43-98).
template <int MODE>
struct MF {
    static constexpr bool top = MODE == kHW || MODE == kOV || MODE == kSW;
    static constexpr bool left = MODE == kOV || MODE == kSW || MODE == kSHW;
    static constexpr bool lrow = MODE == kHW || MODE == kOV;
    static constexpr bool lcol = MODE == kOV || MODE == kSHW;
    static constexpr bool local = MODE == kSW;
};

// gap_penalty (src/kernel.rs:103-109).
__device__ __forceinline__ int gap32(int go, int ge, int len) {
    return len == 0 ? 0 : (int)((long long)go + (long long)(len - 1) * (long long)ge);
}

// H on the DP border: row 0 (top) or column 0 (left).
template <int MODE>
__device__ __forceinline__ int border(int i, int j, int go, int ge) {
    if (i == 0) return (j == 0 || MF<MODE>::top) ? 0 : -gap32(go, ge, j);
    return MF<MODE>::left ? 0 : -gap32(go, ge, i);
}

// Best::consider (src/kernel.rs:113-129): maximise score, then minimise (col, row).
__device__ __forceinline__ void consider(int& bs, int& bj, int& bi, int s, int j, int i) {
    // Bitwise, not short-circuit, so the comparison stays branch-free in the cell loops.
    const bool take = (s > bs) | ((s == bs) & ((j < bj) | ((j == bj) & (i < bi))));
    bs = take ? s : bs;
    bj = take ? j : bj;
    bi = take ? i : bi;
}

// The answer-region cells that lie on the DP border, reduced in closed form. Borders never
// increase away from the origin (gaps are non-negative), so on each border segment of the region
// the cell nearest the origin dominates or ties-and-wins.
template <int MODE>
__device__ __forceinline__ void border_best(int m, int n, int go, int ge, int& bs, int& bj,
                                            int& bi) {
    bs = INT_MIN; bj = INT_MAX; bi = INT_MAX;
    if (MF<MODE>::local) { consider(bs, bj, bi, 0, 0, 0); return; }
    if (m == 0 || n == 0) consider(bs, bj, bi, border<MODE>(m, n, go, ge), n, m);
    if (MF<MODE>::lrow) consider(bs, bj, bi, border<MODE>(m, 0, go, ge), 0, m);
    if (MF<MODE>::lcol) consider(bs, bj, bi, border<MODE>(0, n, go, ge), n, 0);
}

// Tile code is written with lambdas over the tile's register arrays; every one must be
// inlined, or the arrays they capture are moved to scratch memory.
#define HY_INLINE __attribute__((always_inline))

// Packed 16-bit arithmetic: two lanes' worth of values per 32-bit register (v_pk_* forms).
typedef short s2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s2v as_s2(uint32_t u) { return __builtin_bit_cast(s2v, u); }
__device__ __forceinline__ uint32_t as_u32(s2v v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s2v p_add(s2v a, s2v b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ s2v p_sub(s2v a, s2v b) { return __builtin_elementwise_sub_sat(a, b); }
__device__ __forceinline__ s2v p_max(s2v a, s2v b) { return __builtin_elementwise_max(a, b); }
// Unsigned saturating subtraction (v_pk_sub_u16 with clamp): max(a - b, 0) for a, b in [0, 32767].
typedef unsigned short u2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s2v p_subu(s2v a, s2v b) {
    return __builtin_bit_cast(s2v, __builtin_elementwise_sub_sat(__builtin_bit_cast(u2v, a),
                                                                 __builtin_bit_cast(u2v, b)));
}
__device__ __forceinline__ s2v p_minu(s2v a, s2v b) {
    return __builtin_bit_cast(s2v, __builtin_elementwise_min(__builtin_bit_cast(u2v, a),
                                                             __builtin_bit_cast(u2v, b)));
}
// The same as one v_pk_min_u16 the compiler cannot see through. Written as the vector builtin,
// min(a -sat b, 1) is recognised as the comparison a > b and lowered to a compare and a select
// per half (with wait states between them). Which form is faster depends on the kernel's register
// pressure: measured, this one in the two-pairs-per-lane kernel (protein ScoreEnd 1,707 -> 1,793
// GCUPS) and in the one-query scan's end tracking at 24-row tiles; at 32-row tiles that kernel
// spills more with it than with the builtin.
__device__ __forceinline__ s2v p_minu_pk(s2v a, s2v b) {
    uint32_t d;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(d) : "v"(as_u32(a)), "v"(as_u32(b)));
    return as_s2(d);
}
// 0xFFFF in each half where d > 0, else 0, in one instruction: d x 0xFFFF with the unsigned
// clamp (v_pk_mad_u16 ... clamp), for a select by v_bfi_b32.
__device__ __forceinline__ uint32_t p_nonzero_mask(s2v d) {
    uint32_t m;
    asm("v_pk_mad_u16 %0, %1, %2, 0 clamp" : "=v"(m) : "v"(as_u32(d)), "v"(0xFFFFFFFFu));
    return m;
}
__device__ __forceinline__ s2v p_maxu(s2v a, s2v b) {
    return __builtin_bit_cast(s2v, __builtin_elementwise_max(__builtin_bit_cast(u2v, a),
                                                             __builtin_bit_cast(u2v, b)));
}
// a - b per 16-bit half, wrapping (v_pk_sub_u16 without clamp).
__device__ __forceinline__ s2v p_subw(s2v a, s2v b) {
    return __builtin_bit_cast(s2v, __builtin_bit_cast(u2v, a) - __builtin_bit_cast(u2v, b));
}
// a * b + c per 16-bit half, wrapping (v_pk_mul_lo_u16 / v_pk_mad_u16).
__device__ __forceinline__ s2v p_madu(s2v a, s2v b, s2v c) {
    return __builtin_bit_cast(s2v, __builtin_bit_cast(u2v, a) * __builtin_bit_cast(u2v, b) +
                                       __builtin_bit_cast(u2v, c));
}
__device__ __forceinline__ int sat16(long long v) {
    return v < -32768 ? -32768 : (v > 32767 ? 32767 : (int)v);
}
__device__ __forceinline__ s2v splat2(int v) {
    const uint32_t u = (uint32_t)(uint16_t)(short)v;
    return as_s2(u | (u << 16));
}
__device__ __forceinline__ s2v pair2(int lo, int hi) {
    return as_s2((uint32_t)(uint16_t)(short)lo | ((uint32_t)(uint16_t)(short)hi << 16));
}
// Keep v in the halves selected by (lo, hi), the 16-bit sentinel elsewhere.
__device__ __forceinline__ s2v keep2(s2v v, bool lo, bool hi) {
    const uint32_t m = (lo ? 0x0000FFFFu : 0u) | (hi ? 0xFFFF0000u : 0u);
    return as_s2((as_u32(v) & m) | (0x80008000u & ~m));
}

// 64-bit argmax key: higher score first, then the smaller database index.
__device__ __forceinline__ unsigned long long make_key(int score, uint32_t gi) {
    return ((unsigned long long)((uint32_t)score ^ 0x80000000u) << 32) |
           (unsigned long long)(0xFFFFFFFFu - gi);
}
__device__ __forceinline__ int key_score(unsigned long long k) {
    return (int)((uint32_t)(k >> 32) ^ 0x80000000u);
}
__device__ __forceinline__ uint32_t key_index(unsigned long long k) {
    return 0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFull);
}

__device__ __forceinline__ void wave_argmax(unsigned long long k, unsigned long long* dst) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(k, off);
        k = o > k ? o : k;
    }
    // Lane 0 of every wavefront (workgroups may hold several).
    if ((threadIdx.x & 63) == 0 && k != 0ull) atomicMax(dst, k);
}

}  // namespace
}  // namespace hy
