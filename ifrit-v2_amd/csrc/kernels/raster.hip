// Fine raster and early Z on gfx950: one 256-thread workgroup per 16x16 tile.
//
// Coverage follows the reference's hierarchy exactly
// (projects/softgraphics/src/engine/tileraster/TileRasterWorker.cpp:787-1223):
// fully covered tiles from the binner, then 4x4 blocks by trivial reject /
// accept, then 4x2 groups, 2x2 quads and single pixels at 1 sample per pixel,
// each class with its own depth expression (TRW:2244-2918).  With MSAA every
// sample of a partially covered block is tested individually.
//
// Depth test.  The reference shades a tile's fragments in primitive order and
// keeps a fragment when !(depth >= stored) (TRW:2286-2290, 2482-2486,
// 2828-2832).  For finite depths the survivor is the lexicographic minimum of
// (depth, prim * 4 + fan), which is order-independent: the tile's visibility
// lives in LDS as 64-bit keys (ordered depth << 32 | prim * 4 + fan) and
// fragments are merged with 64-bit LDS atomic min in any order.
//
// Work split per tile list entry: triangles whose screen extent (a size hint
// from setup) is at most 64 pixels are rasterised one per lane; larger ones
// one pixel per lane, where each wave holds one row of 4x4 blocks so the 4x2
// group and quad classification of a pixel comes from one ballot.  Both paths
// make the same exact decisions; the split only balances the work.
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace ifrit_hip {

namespace {

using namespace dev;

constexpr int kChunk = 256;

__device__ __forceinline__ void merge(unsigned long long* cell, float d, unsigned tie) {
    if (d == d) atomicMin(cell, frag_key(d, tie));
}

// One lane rasterises one small triangle into the tile's LDS visibility.
// The lane first collects the 4x4 blocks that pass the reference's trivial
// reject test (TRW:887-934) and then, per candidate block, the pixels whose
// square may hold a covered point; it iterates only over set bits, so the
// lanes of a wave stay busy with their own triangles instead of stepping
// through the union of all lanes' blocks.  Both prunings are exact: the
// block test is the reference's own, and square_may_cover() is a monotone
// bound of the pixel test.
template <int S>
__device__ void raster_small(const RasterParams& P, const TriRec& R, int tx, int ty, unsigned long long* vis,
                             unsigned long long& frags, unsigned long long& covs) {
    const Edges E = load_edges(R);
    const float4 F1 = R.f1, F2 = R.f2, F3 = R.f3;
    const unsigned tie = __float_as_uint(R.q2.y);
    const float tileMinX = (float)(tx * 16) * 1.0f, tileMinY = (float)(ty * 16) * 1.0f;
    const int tx0 = tx * 16, ty0 = ty * 16;
    unsigned trMask = 0, taMask = 0;
#pragma unroll
    for (int bi = 0; bi < 16; bi++) {
        int trn, tan;
        block_test(E, fmaf((float)(bi & 3), 4.0f, tileMinX), fmaf((float)(bi >> 2), 4.0f, tileMinY), trn, tan);
        trMask |= (unsigned)(trn == 3) << bi;
        taMask |= (unsigned)(tan == 3) << bi;
    }
    while (trMask) {
        const int bi = __ffs(trMask) - 1;
        trMask &= trMask - 1u;
        const bool accept = (taMask >> bi) & 1u;
        const int px0 = tx0 + (bi & 3) * 4, py0 = ty0 + (bi >> 2) * 4;
        if (S == 1) {
            for (int sk = 0; sk < 2; sk++) {
                const int gy = py0 + 2 * sk;
                // lanes of the 4x2 group: (0,0) (1,0) (0,1) (1,1) (2,0) (3,0) (2,1) (3,1)
                unsigned acc = 0xFFu;
                if (!accept) {
                    acc = 0;
#pragma unroll
                    for (int l = 0; l < 8; l++) {
                        const int ox = (l & 1) + ((l >> 2) << 1), oy = (l >> 1) & 1;
                        acc |= (unsigned)inside(E, (float)(px0 + ox), (float)(gy + oy)) << l;
                    }
                }
                unsigned todo = acc;
                while (todo) {
                    const int l = __ffs(todo) - 1;
                    todo &= todo - 1u;
                    const int px = px0 + (l & 1) + ((l >> 2) << 1), py = gy + ((l >> 1) & 1);
                    if (px >= P.width || py >= P.height) continue;
                    const unsigned quad = (acc >> (l & 4)) & 0xFu;
                    const float fx = (float)px, fy = (float)py;
                    float d;
                    if (acc == 0xFFu)
                        d = depth256(F1, F2, F3, fx, fy);      // BLOCK or PIXEL_PACK4X2
                    else if (quad == 0xFu)
                        d = depth128(F1, F2, F3, fx, fy);      // PIXEL_PACK2X2
                    else
                        d = depth_scalar(F1, F2, F3, fx, fy);  // PIXEL
                    frags++;
                    covs++;
                    merge(&vis[(py - ty0) * 16 + (px - tx0)], d, tie);
                }
            }
        } else {
            unsigned pm = 0xFFFFu;
            if (!accept) {
                pm = 0;
#pragma unroll
                for (int q = 0; q < 16; q++)
                    pm |= (unsigned)square_may_cover(E, (float)(px0 + (q & 3)), (float)(py0 + (q >> 2))) << q;
            }
            while (pm) {
                const int q = __ffs(pm) - 1;
                pm &= pm - 1u;
                const int px = px0 + (q & 3), py = py0 + (q >> 2);
                if (px >= P.width || py >= P.height) continue;
                const float fx = (float)px, fy = (float)py;
                bool any = false;
                unsigned long long* cell = &vis[((py - ty0) * 16 + (px - tx0)) * S];
#pragma unroll
                for (int s = 0; s < S; s++) {
                    const float sx = fx + P.sx[s], sy = fy + P.sy[s];
                    if (!accept && !inside(E, sx, sy)) continue;
                    any = true;
                    covs++;
                    merge(&cell[s], depth256(F1, F2, F3, sx, sy), tie);
                }
                frags += any;
            }
        }
    }
}

// All 256 lanes rasterise one large triangle, one pixel each.  Each lane owns
// its pixel's LDS cells in this phase.
template <int S>
__device__ __forceinline__ void raster_big(const RasterParams& P, const TriRec& R, bool full, float fpx, float fpy,
                                           float bMinX, float bMinY, int gShift0, int qShift0, bool inFb,
                                           unsigned long long* cell, unsigned long long& frags,
                                           unsigned long long& covs) {
    const Edges E = load_edges(R);
    int trn = 3, tan = 3;
    if (!full) block_test(E, bMinX, bMinY, trn, tan);
    const float4 F1 = R.f1, F2 = R.f2, F3 = R.f3;
    const unsigned tie = __float_as_uint(R.q2.y);
    if constexpr (S == 1) {
        const bool pin = inside(E, fpx, fpy);
        const unsigned long long M = __ballot(pin);
        unsigned c = 0;
        if (trn == 3 && tan == 3) {
            c = 1;
        } else if (trn == 3) {
            const bool g = ((M >> gShift0) & 0xFull) == 0xFull && ((M >> (gShift0 + 16)) & 0xFull) == 0xFull;
            const bool qd = ((M >> qShift0) & 0x3ull) == 0x3ull && ((M >> (qShift0 + 16)) & 0x3ull) == 0x3ull;
            c = g ? 1u : (qd ? 2u : (pin ? 3u : 0u));
        }
        if (c && inFb) {
            const float d = c == 1 ? depth256(F1, F2, F3, fpx, fpy)
                                   : (c == 2 ? depth128(F1, F2, F3, fpx, fpy) : depth_scalar(F1, F2, F3, fpx, fpy));
            frags++;
            covs++;
            if (d == d) {
                const unsigned long long k = frag_key(d, tie);
                if (k < cell[0]) cell[0] = k;
            }
        }
    } else {
        if (trn != 3 || !inFb) return;
        if (tan != 3 && !square_may_cover(E, fpx, fpy)) return;
        bool any = false;
#pragma unroll
        for (int s = 0; s < S; s++) {
            const float sx = fpx + P.sx[s], sy = fpy + P.sy[s];
            if (tan != 3 && !inside(E, sx, sy)) continue;
            any = true;
            covs++;
            const float d = depth256(F1, F2, F3, sx, sy);
            if (d == d) {
                const unsigned long long k = frag_key(d, tie);
                if (k < cell[s]) cell[s] = k;
            }
        }
        frags += any;
    }
}

// Coverage class and depth of the winning triangle at a 1-sample pixel,
// recomputed with the same decisions as the raster phases.
__device__ __forceinline__ unsigned winner_class(const TriRec& R, int tx, int ty, float fpx, float fpy, float bMinX,
                                                 float bMinY, float& d) {
    const Edges E = load_edges(R);
    unsigned c;
    if (tile_class(E, tx, ty) == 2) {
        c = 1;
    } else {
        int trn, tan;
        block_test(E, bMinX, bMinY, trn, tan);
        if (tan == 3) {
            c = 1;
        } else {
            const float gx = bMinX, gy = floorf(fpy * 0.5f) * 2.0f;
            bool g = true;
#pragma unroll
            for (int l = 0; l < 8; l++) g = g && inside(E, gx + (float)(l & 3), gy + (float)(l >> 2));
            if (g) {
                c = 1;
            } else {
                const float qx = floorf(fpx * 0.5f) * 2.0f;
                bool q = true;
#pragma unroll
                for (int l = 0; l < 4; l++) q = q && inside(E, qx + (float)(l & 1), gy + (float)(l >> 1));
                c = q ? 2u : 3u;
            }
        }
    }
    d = c == 1 ? depth256(R.f1, R.f2, R.f3, fpx, fpy)
               : (c == 2 ? depth128(R.f1, R.f2, R.f3, fpx, fpy) : depth_scalar(R.f1, R.f2, R.f3, fpx, fpy));
    return c;
}

template <int S>
__global__ void __launch_bounds__(256) k_raster(RasterParams P, TriRecs recs,
                                                const unsigned* __restrict__ slotOf,
                                                const unsigned* __restrict__ binStart,
                                                const unsigned* __restrict__ binCount,
                                                const unsigned* __restrict__ binList, uint32_t* __restrict__ visId,
                                                float* __restrict__ visDepth, uint32_t* __restrict__ visSlot,
                                                uint8_t* __restrict__ visClass, StatPartials st,
                                                DeviceCounters* ctr) {
    __shared__ unsigned long long sVis[256 * S];
    __shared__ TriRec sRec[kChunk];
    __shared__ unsigned sEnt[kChunk];
    __shared__ unsigned short sBig[kChunk];
    __shared__ int sNumBig;

    const int tile = blockIdx.x;
    const int tx = tile % P.tilesX, ty = tile / P.tilesX;
    const int t = threadIdx.x;
    const int lx = t & 15, ly = t >> 4;
    const int px = tx * 16 + lx, py = ty * 16 + ly;
    const bool inFb = px < P.width && py < P.height;
    const size_t pix = (size_t)py * P.width + px;
    const float fpx = (float)px, fpy = (float)py;
    const float bMinX = fmaf((float)(lx >> 2), 4.0f, (float)(tx * 16) * 1.0f);
    const float bMinY = fmaf((float)(ly >> 2), 4.0f, (float)(ty * 16) * 1.0f);
    const int r = ly & 3;
    const int gShift0 = (r & ~1) * 16 + (lx & ~3);
    const int qShift0 = (r & ~1) * 16 + (lx & ~1);

#pragma unroll
    for (int s = 0; s < S; s++) sVis[t * S + s] = frag_key(P.clear || !inFb ? kClearDepth : visDepth[pix * S + s], 0u);

    unsigned long long frags = 0, covs = 0;
    const unsigned start = binStart[tile];
    const unsigned n = start >= P.binCap ? 0u : min(binCount[tile], P.binCap - start);
    for (unsigned base = 0; base < n; base += kChunk) {
        const int cnt = (int)min((unsigned)kChunk, n - base);
        __syncthreads();
        if (t == 0) sNumBig = 0;
        TriRec R;
        unsigned ent = 0;
        if (t < cnt) {
            ent = binList[start + base + t];
            R = load_rec(recs, ent & 0x7fffffffu);
            sRec[t] = R;
            sEnt[t] = ent;
        }
        __syncthreads();
        if (t < cnt) {
            const bool full = (ent >> 31) != 0;
            if (!full && !pix_box(R).big) {
                raster_small<S>(P, R, tx, ty, sVis, frags, covs);
            } else {
                const int k = atomicAdd(&sNumBig, 1);
                sBig[k] = (unsigned short)t;
            }
        }
        __syncthreads();
        const int nb = sNumBig;
        for (int b = 0; b < nb; b++) {
            const int j = sBig[b];
            raster_big<S>(P, sRec[j], (sEnt[j] >> 31) != 0, fpx, fpy, bMinX, bMinY, gShift0, qShift0, inFb,
                          &sVis[t * S], frags, covs);
        }
    }
    __syncthreads();

    if (inFb) {
        unsigned cls = 0;
#pragma unroll
        for (int s = 0; s < S; s++) {
            const unsigned long long k = sVis[t * S + s];
            if (k == frag_key(P.clear ? kClearDepth : visDepth[pix * S + s], 0u)) {
                if (P.clear) {
                    visDepth[pix * S + s] = kClearDepth;
                    visSlot[pix * S + s] = 0xffffffffu;
                    visId[pix * S + s] = 0xffffffffu;
                }
                continue;
            }
            const unsigned tie = (unsigned)k;
            const unsigned prim = tie >> 2;
            const unsigned slot = slotOf[prim] + (tie & 3u);
            float d = order_key_inv((unsigned)(k >> 32));
            if (S == 1) {
                cls = winner_class(load_rec(recs, slot), tx, ty, fpx, fpy, bMinX, bMinY, d);
            } else if ((unsigned)(k >> 32) == 0x80000000u) {
                // +0 and -0 share a key; recompute to return the winner's own bits
                d = depth256(recs.f1[slot], recs.f2[slot], recs.f3[slot], fpx + P.sx[s], fpy + P.sy[s]);
            }
            visDepth[pix * S + s] = d;
            visSlot[pix * S + s] = slot;
            visId[pix * S + s] = prim;
        }
        if (S == 1 && visClass && (cls || P.clear)) visClass[pix] = (uint8_t)cls;
    }
    const unsigned long long tf = block_sum256(frags);
    const unsigned long long tc = block_sum256(covs);
    if (t == 0) {
        st.rasterFrags[tile] = tf;
        st.rasterCovs[tile] = tc;
    }
}

}  // namespace

void launch_raster(const RasterParams& p, int samples, TriRecs recs, const unsigned* slotOf,
                   const unsigned* binStart, const unsigned* binCount, const unsigned* binList, uint32_t* visId,
                   float* visDepth, uint32_t* visSlot, uint8_t* visClass, StatPartials st, DeviceCounters* ctr,
                   hipStream_t s) {
    const dim3 grid(p.tilesX * p.tilesY);
    switch (samples) {
        case 1:
            hipLaunchKernelGGL(k_raster<1>, grid, dim3(256), 0, s, p, recs, slotOf, binStart, binCount, binList, visId,
                               visDepth, visSlot, visClass, st, ctr);
            break;
        case 4:
            hipLaunchKernelGGL(k_raster<4>, grid, dim3(256), 0, s, p, recs, slotOf, binStart, binCount, binList, visId,
                               visDepth, visSlot, visClass, st, ctr);
            break;
        default:
            hipLaunchKernelGGL(k_raster<8>, grid, dim3(256), 0, s, p, recs, slotOf, binStart, binCount, binList, visId,
                               visDepth, visSlot, visClass, st, ctr);
            break;
    }
}

}  // namespace ifrit_hip
