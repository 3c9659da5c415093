// pairhmm_device.hpp -- device helpers shared by the gfx950 PairHMM kernels (forward and backward).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pairhmm_kernels.hpp"

namespace gatk_pairhmm {
namespace dev {

// Fixed-point scale of the state matrices: D(0,.) = 2^kExp / H.  Every cell is <= 2^kExp (the
// recurrences are sub-stochastic), so FP32 keeps headroom below FLT_MAX = 2^128.
template <typename T>
struct Scale;
template <>
struct Scale<float> {
    static constexpr int kExp = 120;
};
template <>
struct Scale<double> {
    static constexpr int kExp = 1020;
};

constexpr int kDppWaveShr1 = 0x138;  // DPP_WF_SR1: lane i reads lane i-1; lane 0 has no source

// Lane k receives `src` from lane k-1; lane 0 keeps `old`.
__device__ __forceinline__ float wave_shr1(float old, float src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, src), kDppWaveShr1, 0xF,
                                                                 0xF, false));
}
__device__ __forceinline__ double wave_shr1(double old, double src) {
    const uint64_t o = __builtin_bit_cast(uint64_t, old), s = __builtin_bit_cast(uint64_t, src);
    const int lo = __builtin_amdgcn_update_dpp(static_cast<int>(o), static_cast<int>(s), kDppWaveShr1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(static_cast<int>(o >> 32), static_cast<int>(s >> 32), kDppWaveShr1,
                                               0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                          static_cast<uint32_t>(lo));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__host__ __device__ constexpr int round_up(int x, int m) { return (x + m - 1) / m * m; }

constexpr int kDppWaveShl1 = 0x130;  // DPP_WF_SL1: lane i reads lane i+1; lane 63 has no source

// Lane k receives `src` from lane k+1; lane 63 keeps `old`.
__device__ __forceinline__ float wave_shl1(float old, float src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, src), kDppWaveShl1, 0xF,
                                                                 0xF, false));
}
__device__ __forceinline__ double wave_shl1(double old, double src) {
    const uint64_t o = __builtin_bit_cast(uint64_t, old), s = __builtin_bit_cast(uint64_t, src);
    const int lo = __builtin_amdgcn_update_dpp(static_cast<int>(o), static_cast<int>(s), kDppWaveShl1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(static_cast<int>(o >> 32), static_cast<int>(s >> 32), kDppWaveShl1,
                                               0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                          static_cast<uint32_t>(lo));
}

// Lane k reads `src` of lane k-1; lane 0 reads 0.  With a zero fill the compiler folds the move
// into the consuming v_mul/v_fmac as a DPP operand (no separate v_mov_dpp).
__device__ __forceinline__ float from_prev_lane(float src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, src), kDppWaveShr1, 0xF, 0xF,
                                                              true));
}
__device__ __forceinline__ double from_prev_lane(double src) {
    const uint64_t s = __builtin_bit_cast(uint64_t, src);
    const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(s), kDppWaveShr1, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(s >> 32), kDppWaveShr1, 0xF, 0xF, true);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                          static_cast<uint32_t>(lo));
}

// a += src_a(lane k-1) * mul_a and b += src_b(lane k-1) * mul_b, lane 0 reading 0 for both sources:
// two v_fmac_f32 with a DPP wave_shr:1 source operand.  The compiler folds a DPP move into v_mul but
// not into v_fmac (tied accumulator), which would leave one v_mov_b32_dpp per term; writing the
// instructions out saves both.  The compiler's hazard recognizer does not look inside inline
// assembly, so the leading s_nop 4 provides the wait states a DPP read needs after a VALU write of
// its source (2) and after an EXEC write (5).
__device__ __forceinline__ void fmac_prev_lane2(float &a, float src_a, float mul_a, float &b, float src_b,
                                                float mul_b) {
    asm("s_nop 4\n\t"
        "v_fmac_f32_dpp %0, %2, %3 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %1, %4, %5 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : "+&v"(a), "+&v"(b)
        : "v"(src_a), "v"(mul_a), "v"(src_b), "v"(mul_b));
}

// Lane 0's value, wave-uniform (v_readfirstlane).
__device__ __forceinline__ float rfl(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double rfl(double x) {
    const uint64_t v = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<int>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<int>(v >> 32));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// Lanes 1..63: a = a0(lane k-1) * ma0 + a1(lane k-1) * ma1 and b = b0(lane k-1) * mb0 + b1(lane k-1) * mb1.
// Lane 0 keeps the a and b it brings in: a DPP wave_shr:1 without bound_ctrl leaves a lane whose
// source lies outside the wave unwritten.  Four VALU instructions, the lane exchange included; the
// leading s_nop 4 provides the DPP wait states the compiler does not track across inline assembly
// (2 after a VALU write of a DPP source, 5 after an EXEC write).
__device__ __forceinline__ void dpp_upper2(float &a, float a0, float ma0, float a1, float ma1, float &b, float b0,
                                           float mb0, float b1, float mb1) {
    asm("s_nop 4\n\t"
        "v_mul_f32_dpp %0, %2, %3 wave_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %6, %7 wave_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %0, %4, %5 wave_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %1, %8, %9 wave_shr:1 row_mask:0xf bank_mask:0xf"
        : "+&v"(a), "+&v"(b)
        : "v"(a0), "v"(ma0), "v"(a1), "v"(ma1), "v"(b0), "v"(mb0), "v"(b1), "v"(mb1));
}

// a^k for 0 <= k < 128 by repeated squaring (exact to a few ulp, unlike exp2(k log2 a)).
__device__ __forceinline__ float ipow(float a, int k) {
    float r = 1.f, b = a;
#pragma unroll
    for (int bit = 0; bit < 7; ++bit) {
        r = (k & (1 << bit)) ? r * b : r;
        b *= b;
    }
    return r;
}

// Lane 63's value, wave-uniform (v_readlane).
__device__ __forceinline__ float lane63(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), 63));
}

// Inclusive linear-recurrence scan over the wave: y_l = sum_{k <= l} x_k a^(l-k).  Hillis-Steele
// steps y_l += a^o y_(l-o) as v_fmac_f32 with a DPP source: row_shr 1, 2, 4, 8 within rows of 16
// (lanes without a source add 0), then row_bcast:15 into rows 1 and 3 (times a^((l mod 16) + 1))
// and row_bcast:31 into rows 2 and 3 (times a^(l - 31)).  s_nop 1 gives the two wait states a DPP
// read needs after the previous VALU write of its source; the leading s_nop 4 also covers an EXEC
// write.
__device__ __forceinline__ float linear_scan64(float x, float a1, float a2, float a4, float a8, float p_row,
                                               float p_31) {
    float y = x;
    asm("s_nop 4\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %2 row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %3 row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %4 row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %5 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %6 row_bcast:31 row_mask:0xc bank_mask:0xf"
        : "+&v"(y)
        : "v"(a1), "v"(a2), "v"(a4), "v"(a8), "v"(p_row), "v"(p_31));
    return y;
}

}  // namespace dev
}  // namespace gatk_pairhmm
