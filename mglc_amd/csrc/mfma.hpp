// mfma.hpp -- the forward MRT moment transform of FP16-stored cells on the
// matrix cores (lbm_params.transform = MFMA), for whole 64-cell wavefronts.
//
// m = M f: v_mfma_f32_32x32x16_f16 with A = M (moments x populations,
// integer entries, exact in FP16) and B = the stored FP16 populations, exact
// as loaded (scaled deviations (f - w) 2^k): every product is exact and the
// sums are FP32.  The inverse stays on the VALU: the relaxed moments are FP32
// and must stay so (in FP16 the moments divided by |row|^2 fall into the
// subnormal range), and on v_mfma_f32_32x32x2_f32 it took 20 MFMAs per wave
// and ran slower than the VALU form.
//
// Lane layout (checked on gfx950 by a probe): lane l holds
// A[m = l % 32][k = 8 (l / 32) + q] and B[k = 8 (l / 32) + q][n = l % 32],
// q = 0..7; result register r of lane l is D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// A wave holds 64 cells, one per lane: the columns n are cells 0..31 (group
// 0) and 32..63 (group 1), and v_permlane32_swap (exchanging the upper half
// of one register with the lower half of another) moves values between a
// lane-per-cell layout and these.
#pragma once
#include <hip/hip_runtime.h>

#include "lattice.hpp"

namespace lbm {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// IEEE binary16 bits of a small integer (|v| < 2048): exact.
constexpr unsigned short half_of_int(int v) {
    if (v == 0) return 0;
    const unsigned short sign = v < 0 ? 0x8000 : 0;
    unsigned a = (unsigned)(v < 0 ? -v : v);
    int e = 0;
    while ((a >> e) > 1) ++e;                             // 2^e <= a < 2^(e+1)
    const unsigned mant = (a << (10 - e)) & 0x3FF;        // e <= 10
    return (unsigned short)(sign | ((unsigned)(e + 15) << 10) | mant);
}

// The A operand, lane-major so that a lane reads its own as whole 16-byte
// words: per K block (populations 0..15; 16..18 and zeros) 8 FP16.
struct MfmaTables {
    unsigned short fwd[64][2][8];
};

constexpr MfmaTables make_mfma_tables() {
    MfmaTables t{};
    for (int l = 0; l < 64; ++l) {
        for (int kb = 0; kb < 2; ++kb)
            for (int q = 0; q < 8; ++q) {
                const int m = l % 32, k = 16 * kb + 8 * (l / 32) + q;
                t.fwd[l][kb][q] = (m < Q && k < Q) ? half_of_int(basis(m, k)) : 0;
            }
    }
    return t;
}

static __constant__ MfmaTables MFMA_TAB = make_mfma_tables();

__device__ __forceinline__ void swap32(unsigned& x, unsigned& y) {
    const auto p = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    x = p[0];
    y = p[1];
}

__device__ __forceinline__ void swap32(float& x, float& y) {
    unsigned a = __builtin_bit_cast(unsigned, x), b = __builtin_bit_cast(unsigned, y);
    swap32(a, b);
    x = __builtin_bit_cast(float, a);
    y = __builtin_bit_cast(float, b);
}

// Results of the two cell groups (d0: cells 0..31, d1: cells 32..63) to the
// lane-per-cell layout: rows 0..18 of each lane's own cell into out[].
__device__ __forceinline__ void gather_rows(f32x16& d0, f32x16& d1, float (&out)[Q]) {
#pragma unroll
    for (int r = 0; r < 11; ++r) {
        float x = d0[r], y = d1[r];
        swap32(x, y);
        d0[r] = x;
        d1[r] = y;
    }
    // d0[r] now holds row 8 (r / 4) + r % 4, d1[r] that row + 4, of the lane's cell
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        out[r] = d0[r];
        out[4 + r] = d1[r];
        out[8 + r] = d0[4 + r];
        out[12 + r] = d1[4 + r];
    }
    out[16] = d0[8];
    out[17] = d0[9];
    out[18] = d0[10];
}

// m = M h for the lane's cell; h: the 19 populations as stored (FP16), m in
// the same (scaled) units, FP32.  All 64 lanes must be active.
__device__ __forceinline__ void mrt_forward_mfma(const _Float16 (&h)[Q], float (&m)[Q], int lane) {
    half8 b0, b1, b2 = {};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        b0[q] = h[q];
        b1[q] = h[8 + q];
    }
    b2[0] = h[16];
    b2[1] = h[17];
    b2[2] = h[18];
    u32x4 B0 = __builtin_bit_cast(u32x4, b0), B1 = __builtin_bit_cast(u32x4, b1), B2 = __builtin_bit_cast(u32x4, b2);
    u32x4 Z = {0u, 0u, 0u, 0u};
    // K block 0 (populations 0..15): group 0's operand = own 0..7 below,
    // 8..15 of lane l - 32 above; group 1's = 0..7 of lane l + 32 below, own
    // 8..15 above.  K block 1 (16..18): the same with zeros for 24..31.
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        unsigned x = B0[r], y = B1[r];
        swap32(x, y);
        B0[r] = x;
        B1[r] = y;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        unsigned x = B2[r], y = Z[r];
        swap32(x, y);
        B2[r] = x;
        Z[r] = y;
    }
    const u32x4* tab = reinterpret_cast<const u32x4*>(&MFMA_TAB.fwd[lane][0][0]);
    const half8 a0 = __builtin_bit_cast(half8, tab[0]);
    const half8 a1 = __builtin_bit_cast(half8, tab[1]);
    f32x16 d0 = {}, d1 = {};
    d0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, __builtin_bit_cast(half8, B0), d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, __builtin_bit_cast(half8, B1), d1, 0, 0, 0);
    d0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, __builtin_bit_cast(half8, B2), d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, __builtin_bit_cast(half8, Z), d1, 0, 0, 0);
    gather_rows(d0, d1, m);
}

}  // namespace lbm
