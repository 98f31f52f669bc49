// Raster helpers shared by the raster and shading kernels: the view of a
// tile, per-pixel context, the class depth, and the ordered traversal of a
// tile's list (depth functions other than LESS, blending).
#pragma once
#include "device_common.h"

namespace ifrit_hip {
namespace rc {

using namespace dev;

// Depth expression codes of a 1-sample fragment, kept in the low bits of its
// key: 1 SIMD256, 2 SIMD128, 3 scalar DPPS, or in the tagging mode
// (optsetForceDeterministic(false), TileRasterWorker.cpp:1581-1675) 3 the
// grouped form.  The deterministic mode draws tile- and block-accepted and
// 4x2-group fragments with SIMD256, 2x2 quads with SIMD128 and single pixels
// with the scalar form; the tagging mode draws accepted fragments grouped,
// 4x2 groups with SIMD256, and quads and single pixels with SIMD128
// (pixelShadingSingleQuad, TileRasterWorker.cpp:2123-2165).
__device__ __forceinline__ float class_depth(unsigned c, const float4& F1, const float4& F2, const float4& F3, float x,
                                             float y, bool tagging = false) {
    if (c == 1) return depth256(F1, F2, F3, x, y);
    if (c == 2) return depth128(F1, F2, F3, x, y);
    return tagging ? depth_grouped(F1, F2, F3, x, y) : depth_scalar(F1, F2, F3, x, y);
}

// class_depth with every expression evaluated and the class's selected:
// SIMD256 and SIMD128 share their three plane values.
__device__ __forceinline__ float class_depth_all(unsigned c, const float4& F1, const float4& F2, const float4& F3,
                                                 float x, float y, bool tagging) {
    const float b0 = fmaf(F1.x, x, fmaf(F1.y, y, F1.z));
    const float b1 = fmaf(F2.x, x, fmaf(F2.y, y, F2.z));
    const float b2 = fmaf(F3.x, x, fmaf(F3.y, y, F3.z));
    const float d256 = fmaf(b0, F1.w, fmaf(b1, F2.w, b2 * F3.w));
    const float d128 = fmaf(b2, F3.w, fmaf(b0, F1.w, b1 * F2.w));
    const float d3 = tagging ? depth_grouped(F1, F2, F3, x, y) : depth_scalar(F1, F2, F3, x, y);
    return c == 1 ? d256 : (c == 2 ? d128 : d3);
}

// Expression code of a fragment from how it was covered: accepted (tile or
// 4x4 block), its 4x2 group all inside, its 2x2 quad all inside, or alone.
__device__ __forceinline__ unsigned expr_code(bool accepted, bool group, bool quad, bool tagging) {
    if (accepted) return tagging ? 3u : 1u;
    if (group) return 1u;
    return quad || tagging ? 2u : 3u;
}

// The class byte a pixel reports: the expression code, with the tagging
// mode's grouped form as 4.
__device__ __forceinline__ uint8_t class_out(unsigned c, bool tagging) {
    return (uint8_t)(tagging && c == 3u ? 4u : c);
}

// The view a tile belongs to: its size and where its outputs and slotOf
// entries start.
struct ViewGeom {
    int width, height, tilesX;
    int tile;  // view-local tile index
    size_t pixBase;
    unsigned primBase;
    unsigned slotBase;  // record slot of the view's triangle 0
    int view;
};

__device__ __forceinline__ ViewGeom view_of_tile(const Views& V, int gtile) {
    const ViewDesc& D = V.desc[V.tileView[gtile]];
    ViewGeom G;
    G.width = D.width;
    G.height = D.height;
    G.tilesX = D.tilesX;
    G.tile = gtile - (int)D.tileBase;
    G.pixBase = (size_t)D.pixBase;
    G.primBase = D.primBase;
    G.slotBase = D.slotBase;
    G.view = V.tileView[gtile];
    return G;
}

// The rectangles are read from device memory at a uniform address (scalar
// loads), so that the test holds no vector registers.
__device__ __forceinline__ bool scissor_pass(const Scissor& sc, float x, float y) {
    if (sc.count == 0) return true;
    for (int i = 0; i < sc.count; i++) {
        const float4 r = sc.rects[i];
        if (x >= r.x && x <= r.x + r.z && y >= r.y && y <= r.y + r.w) return true;
    }
    return false;
}

// Every pixel of the rectangle [x0, x1] x [y0, y1] passes (one scissor holds it whole).
__device__ __forceinline__ bool scissor_all(const Scissor& sc, int x0, int y0, int x1, int y1) {
    if (sc.count == 0) return true;
    for (int i = 0; i < sc.count; i++) {
        const float4 r = sc.rects[i];
        if ((float)x0 >= r.x && (float)x1 <= r.x + r.z && (float)y0 >= r.y && (float)y1 <= r.y + r.w) return true;
    }
    return false;
}

struct PixelCtx {
    int tx, ty, px, py;
    bool inFb;
    size_t pix;
    float fpx, fpy, bMinX, bMinY;
};

__device__ __forceinline__ PixelCtx pixel_ctx(const ViewGeom& G, int lx, int ly) {
    PixelCtx C;
    C.tx = G.tile % G.tilesX;
    C.ty = G.tile / G.tilesX;
    C.px = C.tx * 16 + lx;
    C.py = C.ty * 16 + ly;
    C.inFb = C.px < G.width && C.py < G.height;
    C.pix = (size_t)C.py * G.width + C.px;
    C.fpx = (float)C.px;
    C.fpy = (float)C.py;
    C.bMinX = fmaf((float)(lx >> 2), 4.0f, (float)(C.tx * 16) * 1.0f);
    C.bMinY = fmaf((float)(ly >> 2), 4.0f, (float)(C.ty * 16) * 1.0f);
    return C;
}

// ---------------------------------------------------------------------------
// Ordered traversal.  A tile's list is sorted by (prim * 4 + fan, entry), the
// order of the reference's per-tile sort (TileRasterWorker.cpp:1225-1259);
// lists longer than the LDS buffer are taken in sorted windows of kOrderCap
// entries, each the smallest keys above the previous window.
// ---------------------------------------------------------------------------
constexpr int kOrderCap = 2048;

__device__ __forceinline__ bool depth_pass(int func, float d, float cur, bool scalarPath) {
    switch (func) {
        case 0: return false;                           // NEVER
        case 1: return !(d >= cur);                     // LESS
        case 2: return !(d != cur);                     // EQUAL
        case 3: return !(d > cur);                      // LESS_OR_EQUAL
        case 4: return scalarPath ? true : !(d <= cur); // GREATER
        case 5: return !(d == cur);                     // NOT_EQUAL
        case 6: return !(d < cur);                      // GREATER_OR_EQUAL
        default: return true;                           // ALWAYS
    }
}

// Coverage of one sample position for one triangle with the decisions of
// the raster phases: the tile accepted, its 4x4 block accepted, or the block
// not rejected and the position inside.
__device__ __forceinline__ bool covers(const Edges& E, bool full, int tan, int trn, float x, float y) {
    return full || tan == 3 || (trn == 3 && inside(E, x, y));
}

// Expression code of a covered 1-sample pixel (class_depth): accepted, its
// 4x2 group all inside, its 2x2 quad all inside, or a single pixel.
__device__ __forceinline__ unsigned pixel_class(const Edges& E, bool accepted, float fpx, float fpy, float bMinX,
                                                bool tagging) {
    if (accepted) return expr_code(true, false, false, tagging);
    const float gy = floorf(fpy * 0.5f) * 2.0f;
    bool g = true;
#pragma unroll
    for (int l = 0; l < 8; l++) g = g && inside(E, bMinX + (float)(l & 3), gy + (float)(l >> 2));
    if (g) return 1u;
    const float qx = floorf(fpx * 0.5f) * 2.0f;
    bool q = true;
#pragma unroll
    for (int l = 0; l < 4; l++) q = q && inside(E, qx + (float)(l & 1), gy + (float)(l >> 1));
    return expr_code(false, false, q, tagging);
}

// Ascending bitonic sort of n 64-bit keys in LDS (n a power of two).
__device__ __forceinline__ void bitonic_sort(unsigned long long* a, int n) {
    for (int k = 2; k <= n; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += 256) {
                const int l = i ^ j;
                if (l > i) {
                    const unsigned long long x = a[i], y = a[l];
                    if ((x > y) == ((i & k) == 0)) {
                        a[i] = y;
                        a[l] = x;
                    }
                }
            }
            __syncthreads();
        }
}

// Sorts a bitonic sequence of n keys ascending.
__device__ __forceinline__ void bitonic_merge(unsigned long long* a, int n) {
    for (int j = n >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < n; i += 256) {
            const int l = i ^ j;
            if (l > i) {
                const unsigned long long x = a[i], y = a[l];
                if (x > y) {
                    a[i] = y;
                    a[l] = x;
                }
            }
        }
        __syncthreads();
    }
}

// Sort key of a list entry: the record's primitive key, then the entry.
__device__ __forceinline__ unsigned long long order_entry_key(const TriRecs& recs, unsigned ent) {
    return ((unsigned long long)__float_as_uint(recs.q2[ent & 0x7fffffffu].y) << 32) | ent;
}

// Calls visit(entry) for every entry of a tile's list in sorted order, with
// the entry uniform across the workgroup (256 threads, all must call it).
// sA and sB are LDS buffers of kOrderCap keys.
template <class F>
__device__ __forceinline__ void for_each_ordered(const TriRecs& recs, const unsigned* __restrict__ binList,
                                                 unsigned start, unsigned cnt, unsigned long long* sA,
                                                 unsigned long long* sB, F&& visit) {
    const int t = threadIdx.x;
    unsigned long long last = 0;
    bool haveLast = false;
    while (cnt > 0u) {
        // the window: the (up to) kOrderCap smallest keys above `last`, sorted in sA
        int m;
        if (cnt <= (unsigned)kOrderCap) {
            int n2 = 64;
            while (n2 < (int)cnt) n2 <<= 1;
            for (int i = t; i < n2; i += 256) sA[i] = i < (int)cnt ? order_entry_key(recs, binList[start + i]) : ~0ull;
            __syncthreads();
            bitonic_sort(sA, n2);
            m = (int)cnt;
        } else {
            for (int i = t; i < kOrderCap; i += 256) sA[i] = ~0ull;
            for (unsigned b0 = 0; b0 < cnt; b0 += kOrderCap) {
                for (int i = t; i < kOrderCap; i += 256) {
                    unsigned long long k = ~0ull;
                    if (b0 + i < cnt) {
                        k = order_entry_key(recs, binList[start + b0 + i]);
                        if (haveLast && k <= last) k = ~0ull;
                    }
                    sB[i] = k;
                }
                __syncthreads();
                bitonic_sort(sB, kOrderCap);
                // the lower halves of two sorted runs form a bitonic sequence
                for (int i = t; i < kOrderCap; i += 256) sA[i] = min(sA[i], sB[kOrderCap - 1 - i]);
                __syncthreads();
                bitonic_merge(sA, kOrderCap);
            }
            // entries of the window: the keys below ~0 (sorted, so a count)
            unsigned mine = 0;
            for (int i = t; i < kOrderCap; i += 256) mine += sA[i] != ~0ull;
            m = (int)block_sum256((unsigned long long)mine);
        }
        for (int i = 0; i < m; i++) visit((unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)sA[i]));
        const bool done = cnt <= (unsigned)kOrderCap || m < kOrderCap;
        if (!done) last = sA[kOrderCap - 1];
        __syncthreads();  // sA is reloaded by the next window or tile
        if (done) break;
        haveLast = true;
    }
}

}  // namespace rc
}  // namespace ifrit_hip
