// Device helpers shared by the rasterizer kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "../common/x86_fp.h"
#include "pipeline.h"

namespace ifrit_hip {
namespace dev {

constexpr float kClearDepth = 3.3961514e38f;  // bit pattern 0x7f7f7f7f (TileRasterWorker.cpp:40)
constexpr float kEps = 1e-8f;                  // TileRasterWorker.h:85

inline unsigned blocks(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ TriRec load_rec(const TriRecs& R, unsigned slot) {
    TriRec r;
    r.q0 = R.q0[slot];
    r.q1 = R.q1[slot];
    r.q2 = R.q2[slot];
    r.f1 = R.f1[slot];
    r.f2 = R.f2[slot];
    r.f3 = R.f3[slot];
    return r;
}

// Pixel box of a record: x0, y0, x1, y1 and the big flag.
struct PixBox {
    int x0, y0, x1, y1;
    bool big;
};
__device__ __forceinline__ PixBox pix_box(const TriRec& r) {
    const unsigned lo = __float_as_uint(r.q2.z), hi = __float_as_uint(r.q2.w);
    PixBox b;
    b.x0 = (int)(lo & 0xffffu);
    b.y0 = (int)(lo >> 16);
    b.x1 = (int)(hi & 0xffffu);
    b.y1 = (int)((hi >> 16) & 0x7fffu);
    b.big = (hi >> 31) != 0;
    return b;
}

// Edge coefficients of a TriRec: inside when fma(ey, y, ex * x) < -ez.
struct Edges {
    float x[3], y[3], z[3];
};

__device__ __forceinline__ Edges load_edges(const TriRec& r) {
    Edges E;
    E.x[0] = r.q0.x; E.y[0] = r.q0.y; E.z[0] = r.q0.z;
    E.x[1] = r.q0.w; E.y[1] = r.q1.x; E.z[1] = r.q1.y;
    E.x[2] = r.q1.z; E.y[2] = r.q1.w; E.z[2] = r.q2.x;
    return E;
}

__device__ __forceinline__ Edges load_edges(const TriRecs& R, unsigned slot) {
    const float4 a = R.q0[slot], b = R.q1[slot];
    const float c = R.q2[slot].x;
    Edges E;
    E.x[0] = a.x; E.y[0] = a.y; E.z[0] = a.z;
    E.x[1] = a.w; E.y[1] = b.x; E.z[1] = b.y;
    E.x[2] = b.z; E.y[2] = b.w; E.z[2] = c;
    return E;
}

// Tile classification of the binner (TileRasterWorker.cpp:341-395, SIMD128
// form): 0 rejected, 1 partially covered, 2 fully covered.
__device__ __forceinline__ int tile_class(const Edges& E, int tx, int ty) {
    const float x0 = (float)tx * 16.0f, y0 = (float)ty * 16.0f;
    const float x1 = x0 + 16.0f, y1 = y0 + 16.0f;
    int ntr = 0, nta = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const bool right = E.x[k] < 0.0f, down = E.y[k] < 0.0f;
        const float xr = right ? x1 : x0, yr = down ? y1 : y0;
        const float xa = right ? x0 : x1, ya = down ? y0 : y1;
        ntr += fmaf(E.x[k], xr, fmaf(E.y[k], yr, E.z[k])) < 0.0f;
        nta += fmaf(E.x[k], xa, fmaf(E.y[k], ya, E.z[k])) < 0.0f;
    }
    return ntr != 3 ? 0 : (nta == 3 ? 2 : 1);
}

// 4x4 block trivial reject / accept (TileRasterWorker.cpp:887-934): block
// corners chosen from the edge normals, test fma(ey, Y, ex * X) < -ez.
__device__ __forceinline__ void block_test(const Edges& E, float bMinX, float bMinY, int& trn, int& tan) {
    const float bMaxX = bMinX + 4.0f, bMaxY = bMinY + 4.0f;
    trn = 0;
    tan = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const bool right = E.x[k] < 0.0f, down = E.y[k] < 0.0f;
        const float xr = right ? bMaxX : bMinX, yr = down ? bMaxY : bMinY;
        const float xa = right ? bMinX : bMaxX, ya = down ? bMinY : bMaxY;
        const float nz = ifx::neg(E.z[k]);
        trn += fmaf(E.y[k], yr, E.x[k] * xr) < nz;
        tan += fmaf(E.y[k], ya, E.x[k] * xa) < nz;
    }
}

// Pixel / sample coverage test (TileRasterWorker.cpp:1042-1066).
__device__ __forceinline__ bool inside(const Edges& E, float x, float y) {
    int n = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) n += fmaf(E.y[k], y, E.x[k] * x) < ifx::neg(E.z[k]);
    return n == 3;
}

// Rejects the whole unit square [x, x+1] x [y, y+1] when an edge excludes
// the corner that minimises it.  Rounding is monotone, so no point of the
// square can pass inside() when this returns false.
__device__ __forceinline__ bool square_may_cover(const Edges& E, float x, float y) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float cx = E.x[k] < 0.0f ? x + 1.0f : x;
        const float cy = E.y[k] < 0.0f ? y + 1.0f : y;
        ok = ok && (fmaf(E.y[k], cy, E.x[k] * cx) < ifx::neg(E.z[k]));
    }
    return ok;
}

// Depth expressions of the three proposal classes.
// SIMD256 path (TileRasterWorker.cpp:2758-2769), also used for MSAA samples.
__device__ __forceinline__ float depth256(const float4& F1, const float4& F2, const float4& F3, float x, float y) {
    const float b0 = fmaf(F1.x, x, fmaf(F1.y, y, F1.z));
    const float b1 = fmaf(F2.x, x, fmaf(F2.y, y, F2.z));
    const float b2 = fmaf(F3.x, x, fmaf(F3.y, y, F3.z));
    return fmaf(b0, F1.w, fmaf(b1, F2.w, b2 * F3.w));
}
// SIMD128 path (TileRasterWorker.cpp:2423-2434).
__device__ __forceinline__ float depth128(const float4& F1, const float4& F2, const float4& F3, float x, float y) {
    const float b0 = fmaf(F1.x, x, fmaf(F1.y, y, F1.z));
    const float b1 = fmaf(F2.x, x, fmaf(F2.y, y, F2.z));
    const float b2 = fmaf(F3.x, x, fmaf(F3.y, y, F3.z));
    return fmaf(b2, F3.w, fmaf(b0, F1.w, b1 * F2.w));
}
// Scalar path with DPPS dot products (TileRasterWorker.cpp:2261-2270).
__device__ __forceinline__ float depth_scalar(const float4& F1, const float4& F2, const float4& F3, float x, float y) {
    const float b0 = (x * F1.x + y * F1.y) + (1.0f * F1.z + 0.0f * F1.w);
    const float b1 = (x * F2.x + y * F2.y) + (1.0f * F2.z + 0.0f * F2.w);
    const float b2 = (x * F3.x + y * F3.y) + (1.0f * F3.z + 0.0f * F3.w);
    return (F1.w * b0 + F2.w * b1) + (F3.w * b2 + 0.0f);
}

// Tagging mode, tile- and block-accepted fragments (pixelShadingSIMD256Grouped,
// TileRasterWorker.cpp:2672-2684): the x term enters first, then y.
__device__ __forceinline__ float depth_grouped(const float4& F1, const float4& F2, const float4& F3, float x, float y) {
    const float b0 = fmaf(F1.y, y, fmaf(F1.x, x, F1.z));
    const float b1 = fmaf(F2.y, y, fmaf(F2.x, x, F2.z));
    const float b2 = fmaf(F3.y, y, fmaf(F3.x, x, F3.z));
    return fmaf(b0, F1.w, fmaf(b1, F2.w, b2 * F3.w));
}

// Order-preserving map of FP32 depth to uint32 (-0 and +0 compare equal, as
// they do under the CPU path's float comparison).
__device__ __forceinline__ unsigned order_key(float d) {
    unsigned u = __float_as_uint(d);
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float order_key_inv(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Fragment key for the depth test: lexicographic (depth, prim*4 + fan).
__device__ __forceinline__ unsigned long long frag_key(float d, unsigned tie) {
    return ((unsigned long long)order_key(d) << 32) | tie;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

// Wave-aggregated counter increment: every active lane adds one to
// ctr[key] and receives a unique position; lanes sharing a key share one
// atomic.  The lanes are grouped by key with ballots first, so all groups'
// atomics go out as one instruction and the wave waits for memory once.
// Must be called by the whole (converged) wave.
__device__ __forceinline__ unsigned wave_inc(unsigned* ctr, int key, bool active) {
    const int lane = lane_id();
    unsigned long long pending = __ballot(active);
    int leader = lane;
    unsigned rank = 0, size = 0;
    while (pending) {
        const int l = __ffsll((unsigned long long)pending) - 1;
        const int lk = __shfl(key, l, 64);
        const unsigned long long grp = __ballot(active && key == lk) & pending;
        if ((grp >> lane) & 1ull) {
            leader = l;
            rank = (unsigned)__popcll(grp & ((1ull << lane) - 1ull));
        }
        if (lane == l) size = (unsigned)__popcll(grp);
        pending &= ~grp;
    }
    unsigned base = 0;
    if (active && leader == lane) base = atomicAdd(&ctr[key], size);
    base = __shfl(base, leader, 64);
    return base + rank;
}

// Sum over a 256-thread workgroup (all threads must call it).
__device__ __forceinline__ unsigned long long block_sum256(unsigned long long v) {
    __shared__ unsigned long long part[4];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    const unsigned long long r = part[0] + part[1] + part[2] + part[3];
    __syncthreads();
    return r;
}

__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Inclusive prefix sum over the 64 lanes of a wave with DPP moves: row
// shifts by 1, 2, 4 and 8 (a lane whose source lies outside its row of 16
// adds 0), then the last lane of rows 0 and 2 added to rows 1 and 3
// (row_bcast:15) and the last lane of row 1 to rows 2 and 3 (row_bcast:31).
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    static_assert(sizeof(T) == 4, "32-bit values");
    v += (T)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);
    v += (T)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);
    v += (T)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);
    v += (T)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);
    v += (T)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
    v += (T)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
    return v;
}

}  // namespace dev
}  // namespace ifrit_hip
