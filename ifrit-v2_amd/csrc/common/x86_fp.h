// x86 SSE/AVX floating-point semantics needed to reproduce the tile
// rasterizer's CPU arithmetic bit for bit on gfx950.
//
// The CPU path (projects/softgraphics/src/engine/tileraster/TileRasterWorker.cpp)
// is written with SSE/AVX2 intrinsics whose results differ from plain IEEE
// C math in four places, each emulated here:
//   * RCPPS / RCPSS, the 12-bit reciprocal approximation
//     (include/ifrit/core/math/fastutil/FastUtilSimd.h:27-33, 87-97), whose
//     outputs differ between CPU vendors: a table for Intel and one for AMD;
//   * MINPS / MAXPS, which return the second operand when the comparison is
//     false (NaN or equal operands), unlike fminf/fmaxf;
//   * CVTTSS2SI, which returns INT_MIN for NaN and out-of-range inputs
//     where the GPU conversion saturates (TileRasterWorker.cpp:319-322);
//   * DPPS, which sums ((p0 + p1) + (p2 + p3)) with masked products as +0
//     (include/ifrit/core/math/simd/SimdVectors.h:497-510).
// Everything else is IEEE single precision with round-to-nearest-even; the
// kernels are compiled with -ffp-contract=off and call fmaf where the CPU
// code calls an FMA intrinsic.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define IFX_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#include <string.h>
#define IFX_HD static inline
#endif

namespace ifx {

// Whose RCPSS to reproduce.  The instruction is implementation-defined: on
// the Intel CPU its table was captured on the result depends on the top 11
// mantissa bits, on the AMD EPYC its other table was captured on, on the top
// 12; both scale exactly with the exponent and share the special cases.
enum RcpModel : int { kRcpIntel = 0, kRcpAmd = 1 };

#if defined(__HIPCC__)
__device__ __constant__ static const uint32_t kRcpTable[2048] = {
#include "x86_rcp_table.inc"
};
__device__ __constant__ static const uint32_t kRcpTableAmd[4096] = {
#include "x86_rcp_table_amd.inc"
};
#endif
// Host copies of the same data for host-side setup code.
static const uint32_t kRcpTableHost[2048] = {
#include "x86_rcp_table.inc"
};
static const uint32_t kRcpTableAmdHost[4096] = {
#include "x86_rcp_table_amd.inc"
};

IFX_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
IFX_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// RCPSS/RCPPS of the given model: table lookup on the top 11 (Intel) or 12
// (AMD) mantissa bits, exact exponent scaling, sub-normal results flushed to
// signed zero, zero and sub-normal inputs give signed infinity, infinities
// give signed zero, NaNs come back quieted.
IFX_HD float rcp_ps(float x, int model) {
    uint32_t b = f2u(x);
    uint32_t sign = b & 0x80000000u;
    uint32_t e = (b >> 23) & 0xffu;
    if (e == 0u) return u2f(sign | 0x7f800000u);
    if (e == 255u) return (b & 0x7fffffu) ? u2f(b | 0x400000u) : u2f(sign);
    const bool amd = model == kRcpAmd;
    const uint32_t idx = amd ? (b >> 11) & 0xfffu : (b >> 12) & 0x7ffu;
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t* tab = amd ? kRcpTableAmd : kRcpTable;
#else
    const uint32_t* tab = amd ? kRcpTableAmdHost : kRcpTableHost;
#endif
    uint32_t t = tab[idx];
    int re = (int)((t >> 23) & 0xffu) + 127 - (int)e;
    if (re <= 0) return u2f(sign);
    return u2f(sign | (t & 0x7fffffu) | ((uint32_t)re << 23));
}

// ApproxReciprocalNR256: r = rcp(x); (r + r) - x * (r * r).
IFX_HD float rcp_nr_ps(float x, int model) {
    float r = rcp_ps(x, model);
    float rr = r * r;
    float muls = x * rr;
    float twice = r + r;
    return twice - muls;
}

// ApproxReciprocal (scalar): r = rcp(x); r * (2 - r * x).
IFX_HD float rcp_nr_ss(float x, int model) {
    float r = rcp_ps(x, model);
    float rx = r * x;
    float t = 2.0f - rx;
    return r * t;
}

IFX_HD float min_ps(float a, float b) { return a < b ? a : b; }
IFX_HD float max_ps(float a, float b) { return a > b ? a : b; }

// CVTTSS2SI: truncation toward zero; 0x80000000 when NaN or out of range.
IFX_HD int cvtt_i32(float x) {
    if (!(x > -2147483904.0f && x < 2147483648.0f)) return (int)0x80000000u;
    return (int)x;
}

IFX_HD float neg(float x) { return u2f(f2u(x) ^ 0x80000000u); }

}  // namespace ifx
