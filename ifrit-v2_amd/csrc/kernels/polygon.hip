// Point and line polygon modes (setRasterizerPolygonMode,
// TileRasterRendererCuda.h:106), after the CUDA spec's PointPipeline and
// LinePipeline (TileRasterCoreInvocationCuda.cu:2780-3335), at one sample per
// pixel.
//
// Every candidate (pixel, primitive) carries its own depth and meets the
// others in one 64-bit atomic minimum per pixel, order_key(depth) << 32 |
// primitive id, instead of the spec's per-pixel lists: the smallest depth
// below 1e9 wins and equal depths go to the lowest id (one of the outcomes the
// spec's atomics can produce).  The resolve keeps a winner below the stored
// depth (CUI:2915-2918, 3278-3281).  Each accepted point or line also writes a
// shading record whose planes interpolate its vertex weights, so the deferred
// shader (k_shade) shades it like a triangle fragment.
#include "pipeline.h"
#include "device_common.h"

namespace ifrit_hip {
namespace {

using namespace dev;

constexpr float kPolyEps = 1e-8f;  // CU_EPS (TileRasterConstantsCuda.h:23)
constexpr float kPolyFar = 1e9f;   // the fragment stage's initial nearest depth (CUI:2867, 3180)

// Core Lerp, a (1 - t) + b t (include/ifrit/core/math/VectorOps.h:204-215).
__device__ __forceinline__ float lerp1(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// int(v * (size - 1)) (CUI:2808-2809, 3067-3070), clamped where the product leaves the int range.
__device__ __forceinline__ int to_pixel(float v, int size) {
    const float p = v * (float)(size - 1);
    return (int)fminf(fmaxf(p, -1073741824.0f), 1073741824.0f);
}

__device__ __forceinline__ void offer(const PolyParams& P, int px, int py, float z, unsigned id,
                                      unsigned long long& frags) {
    frags++;
    if (!(z < kPolyFar)) return;
    atomicMin(&P.keys[(size_t)py * P.width + px], ((unsigned long long)order_key(z) << 32) | id);
}

// A shading record: planes f1, f2 (f3 = 0) whose normalised values weight the
// record's two rebased vertices (bx, by) and the primitive key tri * 4 + sub.
__device__ __forceinline__ void store_record(TriRecs recs, float4* bary, unsigned slot, float4 f1, float4 f2,
                                             const float* bx, const float* by, unsigned key) {
    recs.f1[slot] = f1;
    recs.f2[slot] = f2;
    recs.f3[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    recs.q2[slot] = make_float4(0.0f, __uint_as_float(key), 0.0f, 0.0f);
    bary[2 * slot] = make_float4(bx[0], bx[1], bx[2], by[0]);
    bary[2 * slot + 1] = make_float4(by[1], by[2], 0.0f, 0.0f);
}

// The workgroup's candidate count into the raster partials and its accepted
// primitives straight into the counters (the setup partials are zeroed: the
// grid covers both partial arrays).
__device__ __forceinline__ void block_stats(unsigned long long frags, unsigned long long acc, StatPartials st,
                                            DeviceCounters* ctr) {
    __shared__ unsigned long long red[2][4];
    for (int off = 32; off > 0; off >>= 1) {
        frags += __shfl_down(frags, off, 64);
        acc += __shfl_down(acc, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = frags;
        red[1][threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long f = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        const unsigned long long a = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        if ((int)blockIdx.x < st.rasterBlocks) {
            st.rasterFrags[blockIdx.x] = f;
            st.rasterCovs[blockIdx.x] = f;
        }
        if ((int)blockIdx.x < st.setupBlocks) st.setupAssembled[blockIdx.x] = 0;
        if (a) atomicAdd(&ctr->assembled, a);
    }
}

// Clears the per-pixel candidate keys.
__global__ void __launch_bounds__(256) k_poly_clear(unsigned long long* keys, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) keys[i] = ~0ull;
}

// pointRasterizationInsertKernel (CUI:2782-2816): one point per index (of
// whole triangles: trailing indices are ignored, as the fill mode ignores them).
__global__ void __launch_bounds__(256) k_points(PolyParams P, const float4* __restrict__ clip,
                                                const int* __restrict__ idx, TriRecs recs, float4* __restrict__ bary,
                                                StatPartials st, DeviceCounters* ctr) {
    unsigned long long frags = 0, acc = 0;
    // indices of whole triangles only: a point's record names its triangle's vertices
    const int n = P.nidx / 3 * 3;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int vi = idx[i];
        if (vi < 0 || vi >= P.nverts) {
            atomicOr(&ctr->status, kStatusBadIndex);
            continue;
        }
        const float4 v = clip[vi];
        if (!(v.w >= kPolyEps)) continue;
        float x = v.x / v.w, y = v.y / v.w;
        const float z = v.z / v.w;
        x = x * 0.5f + 0.5f;
        y = y * 0.5f + 0.5f;
        if (!(x >= 0.0f && x <= 1.0f && y >= 0.0f && y <= 1.0f)) continue;
        acc++;
        const int k = i % 3;
        const float bx[3] = {k == 0 ? 1.0f : 0.0f, 0.0f, 0.0f}, by[3] = {k == 1 ? 1.0f : 0.0f, 0.0f, 0.0f};
        store_record(recs, bary, (unsigned)i, make_float4(0.0f, 0.0f, 1.0f, z), make_float4(0.0f, 0.0f, 0.0f, 0.0f),
                     bx, by, (unsigned)(i / 3) * 4u + (unsigned)k);
        offer(P, to_pixel(x, P.width), to_pixel(y, P.height), z, (unsigned)i, frags);
    }
    block_stats(frags, acc, st, ctr);
}

// devLineCullClip (CUI:2976-3018): an end behind w = 0 is cut at w = 1; the
// divide and the [0, 1] mapping; the spec's second bounds test checks only the
// far end's lower bounds.  Non-finite ends are dropped.
__device__ bool clip_line(float* a, float* b, float* ba, float* bb) {
    if (a[3] < 0.0f && b[3] < 0.0f) return false;
    if (a[3] < 0.0f) {
        const float t = (-1.0f + b[3]) / (b[3] - a[3]);
        for (int k = 0; k < 4; k++) a[k] = lerp1(b[k], a[k], t);
        for (int k = 0; k < 3; k++) ba[k] = lerp1(bb[k], ba[k], t);
    }
    if (b[3] < 0.0f) {
        const float t = (-1.0f + a[3]) / (a[3] - b[3]);
        for (int k = 0; k < 4; k++) b[k] = lerp1(a[k], b[k], t);
        for (int k = 0; k < 3; k++) bb[k] = lerp1(ba[k], bb[k], t);
    }
    for (int e = 0; e < 2; e++) {
        float* v = e == 0 ? a : b;
        for (int k = 0; k < 3; k++) v[k] = v[k] / v[3];
        v[0] = v[0] * 0.5f + 0.5f;
        v[1] = v[1] * 0.5f + 0.5f;
        for (int k = 0; k < 4; k++)
            if (!isfinite(v[k])) return false;
    }
    if (a[0] < 0.0f || a[1] < 0.0f || a[0] > 1.0f || a[1] > 1.0f) return false;
    if (b[0] < 0.0f || b[1] < 0.0f) return false;
    return true;
}

// lineGeometryAssemblyKernel + bresenhamRasterizationKernel (CUI:3020-3126):
// one lane per triangle edge; the pixels of the edge in the spec's order with
// its float error accumulator, each with getPercent's depth (CUI:3181-3195,
// 3257-3266).  Pixels outside the frame are skipped and the walk stops once
// its advancing coordinate leaves the frame.
__global__ void __launch_bounds__(256) k_lines(PolyParams P, const float4* __restrict__ clip,
                                               const int* __restrict__ idx, TriRecs recs, float4* __restrict__ bary,
                                               StatPartials st, DeviceCounters* ctr) {
    unsigned long long frags = 0, acc = 0;
    const int nlines = P.nidx / 3 * 3;
    const float W = (float)P.width, H = (float)P.height;
    for (int line = blockIdx.x * 256 + threadIdx.x; line < nlines; line += gridDim.x * 256) {
        const int t = line / 3, e = line % 3;
        const int i0 = idx[3 * t + e], i1 = idx[3 * t + (e + 1) % 3];
        if (i0 < 0 || i1 < 0 || i0 >= P.nverts || i1 >= P.nverts) {
            atomicOr(&ctr->status, kStatusBadIndex);
            continue;
        }
        const float4 c0 = clip[i0], c1 = clip[i1];
        float a[4] = {c0.x, c0.y, c0.z, c0.w}, b[4] = {c1.x, c1.y, c1.z, c1.w};
        float ba[3] = {0.0f, 0.0f, 0.0f}, bb[3] = {0.0f, 0.0f, 0.0f};
        ba[e] = 1.0f;
        bb[(e + 1) % 3] = 1.0f;
        if (!clip_line(a, b, ba, bb)) continue;
        acc++;
        // shading planes: the weights of the two ends, perspective-correct along the major axis
        const float ax = a[0] * W, ay = a[1] * H, bxs = b[0] * W, bys = b[1] * H;
        const bool xMajorP = fabsf(bxs - ax) >= fabsf(bys - ay);
        {
            const float o = xMajorP ? ax : ay;
            const float invD = 1.0f / (xMajorP ? bxs - ax : bys - ay);
            const float iw1 = 1.0f / a[3], iw2 = 1.0f / b[3];
            const float s1 = -(invD * iw1), c1_ = (1.0f + o * invD) * iw1;
            const float s2 = invD * iw2, c2_ = -(o * invD) * iw2;
            const float4 f1 = xMajorP ? make_float4(s1, 0.0f, c1_, 0.0f) : make_float4(0.0f, s1, c1_, 0.0f);
            const float4 f2 = xMajorP ? make_float4(s2, 0.0f, c2_, 0.0f) : make_float4(0.0f, s2, c2_, 0.0f);
            const float rbx[3] = {ba[0], bb[0], 0.0f}, rby[3] = {ba[1], bb[1], 0.0f};
            store_record(recs, bary, (unsigned)line, f1, f2, rbx, rby, (unsigned)t * 4u + (unsigned)e);
        }
        int x1 = to_pixel(a[0], P.width), y1 = to_pixel(a[1], P.height);
        int x2 = to_pixel(b[0], P.width), y2 = to_pixel(b[1], P.height);
        const bool xMajor = abs(x1 - x2) > abs(y1 - y2);
        if (xMajor ? x2 < x1 : y2 < y1) {
            const int tx = x1, ty = y1;
            x1 = x2;
            y1 = y2;
            x2 = tx;
            y2 = ty;
        }
        int cx = x1, cy = y1;
        const int len = xMajor ? x2 - x1 : y2 - y1;
        const float slope = xMajor ? fabsf(1.0f * (float)(y2 - y1) / (float)(x2 - x1))
                                   : fabsf(1.0f * (float)(x2 - x1) / (float)(y2 - y1));
        const int step = xMajor ? (y2 - y1 > 0 ? 1 : -1) : (x2 - x1 > 0 ? 1 : -1);
        float d = 0.0f;
        for (int i = 0; i <= len; i++) {
            if (xMajor ? cx >= P.width : cy >= P.height) break;
            if (cx >= 0 && cy >= 0 && cx < P.width && cy < P.height) {
                const float p = xMajorP ? ((float)cx - ax) / (bxs - ax) : ((float)cy - ay) / (bys - ay);
                offer(P, cx, cy, lerp1(a[2], b[2], p), (unsigned)line, frags);
            }
            d += slope;
            if (d > 0.5f) {
                if (xMajor) cy += step;
                else cx += step;
                d -= 1.0f;
            }
            if (xMajor) cx++;
            else cy++;
        }
    }
    block_stats(frags, acc, st, ctr);
}

// Per pixel: the winning candidate when it is below the stored depth.
__global__ void __launch_bounds__(256) k_poly_resolve(PolyParams P, float* __restrict__ visDepth,
                                                      uint32_t* __restrict__ visSlot, uint8_t* __restrict__ visClass) {
    const size_t n = (size_t)P.width * P.height;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const unsigned long long k = P.keys[i];
        const float cur = P.clear ? P.clearDepth : visDepth[i];
        float d = cur;
        unsigned slot = P.clear ? 0xffffffffu : visSlot[i];
        if (k != ~0ull) {
            const float z = order_key_inv((unsigned)(k >> 32));
            if (z < cur) {
                d = z;
                slot = (unsigned)k;
            }
        }
        visDepth[i] = d;
        visSlot[i] = slot;
        if (visClass) visClass[i] = 0;
    }
}

}  // namespace

void launch_polygons(const PolyParams& p, const float4* clip, const int* idx, TriRecs recs, float4* bary,
                     float* visDepth, uint32_t* visSlot, uint8_t* visClass, StatPartials st, DeviceCounters* ctr,
                     int grid, hipStream_t s) {
    const size_t px = (size_t)p.width * p.height;
    hipLaunchKernelGGL(k_poly_clear, dim3(2048), dim3(256), 0, s, p.keys, px);
    if (p.mode == 2)
        hipLaunchKernelGGL(k_points, dim3(grid), dim3(256), 0, s, p, clip, idx, recs, bary, st, ctr);
    else
        hipLaunchKernelGGL(k_lines, dim3(grid), dim3(256), 0, s, p, clip, idx, recs, bary, st, ctr);
    hipLaunchKernelGGL(k_poly_resolve, dim3(2048), dim3(256), 0, s, p, visDepth, visSlot, visClass);
}

}  // namespace ifrit_hip
