// Tile binning on gfx950 (the reference's executeBinner,
// projects/softgraphics/src/engine/tileraster/TileRasterWorker.cpp:314-397):
// every assembled triangle is tested against the 16x16 tiles of its tile
// rectangle; a tile whose trivial-reject corner is inside all three edges
// receives the triangle, flagged as fully covered when the trivial-accept
// corner is inside too.
//
// Two passes over the triangles (count, then scatter after an exclusive scan
// of the per-tile counts) build contiguous per-tile lists.  Lanes of a wave
// that hit the same tile share one atomic (wave_inc).  Triangles spanning
// more than bigThreshold tiles are deferred to a workgroup-per-triangle pass
// so that a few screen-sized triangles do not serialise one lane.
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace ifrit_hip {

namespace {

using dev::Edges;
using dev::load_edges;
using dev::tile_class;

// Shared traversal of the small triangles, one thread per (slot, view) with
// blockIdx.y = view.  SCATTER = false counts, SCATTER = true writes the lists.
template <bool SCATTER>
__global__ void __launch_bounds__(256) k_bin_small(BinParams P, Views V, TriRecs recs,
                                                   const int4* __restrict__ tileBox, unsigned* __restrict__ binCount,
                                                   const unsigned* __restrict__ binStart,
                                                   unsigned* __restrict__ binCursor, unsigned* __restrict__ binList,
                                                   unsigned* __restrict__ bigList, DeviceCounters* ctr) {
    const ViewDesc& D = V.desc[blockIdx.y];
    const unsigned tileBase = D.tileBase;
    const int tilesX = D.tilesX;
    const unsigned local = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned s = D.slotBase + local;
    bool valid = local < (unsigned)P.ntri + min(V.ctr[blockIdx.y].clipCount, P.clipCap);
    int4 b = make_int4(1, 1, 0, 0);
    if (valid) b = tileBox[s];
    valid = valid && b.x <= b.z && b.y <= b.w;
    const int w = b.z - b.x + 1;
    const int area = valid ? w * (b.w - b.y + 1) : 0;
    const bool big = valid && area > P.bigThreshold;
    if (!SCATTER) {
        // deferred triangles go to the per-workgroup pass
        const unsigned k = dev::wave_inc(&ctr->bigCount, 0, big);
        if (big) {
            if (k < P.bigCap)
                bigList[k] = s;
            else
                atomicOr(&ctr->status, kStatusBigOverflow);
        }
    }
    const int n = (valid && !big) ? area : 0;
    Edges E;
    if (n) E = load_edges(recs, s);
    const int iters = dev::wave_max(n);
    for (int i = 0; i < iters; i++) {
        const bool active = i < n;
        int t = 0, c = 0;
        if (active) {
            const int tx = b.x + i % w, ty = b.y + i / w;
            t = (int)tileBase + ty * tilesX + tx;
            c = tile_class(E, tx, ty);
        }
        const bool hit = active && c != 0;
        if (!SCATTER) {
            (void)dev::wave_inc(binCount, t, hit);
        } else {
            const unsigned pos = dev::wave_inc(binCursor, t, hit);
            if (hit) {
                const unsigned at = binStart[t] + pos;
                if (at < P.binCap) binList[at] = s | (c == 2 ? 0x80000000u : 0u);
            }
        }
    }
}

template <bool SCATTER>
__global__ void __launch_bounds__(256) k_bin_big(BinParams P, Views V, TriRecs recs,
                                                 const int4* __restrict__ tileBox, unsigned* __restrict__ binCount,
                                                 const unsigned* __restrict__ binStart, unsigned* __restrict__ binCursor,
                                                 unsigned* __restrict__ binList, const unsigned* __restrict__ bigList,
                                                 const DeviceCounters* ctr) {
    const unsigned nbig = min(ctr->bigCount, P.bigCap);
    for (unsigned i = blockIdx.x; i < nbig; i += gridDim.x) {
        const unsigned s = bigList[i];
        const ViewDesc& D = V.desc[s / V.slotsPerView];
        const unsigned tileBase = D.tileBase;
        const int tilesX = D.tilesX;
        const int4 b = tileBox[s];
        const Edges E = load_edges(recs, s);
        const int w = b.z - b.x + 1;
        const int area = w * (b.w - b.y + 1);
        for (int q = threadIdx.x; q < area; q += blockDim.x) {
            const int tx = b.x + q % w, ty = b.y + q / w;
            const int c = tile_class(E, tx, ty);
            if (!c) continue;
            const int t = (int)tileBase + ty * tilesX + tx;
            if (!SCATTER) {
                atomicAdd(&binCount[t], 1u);
            } else {
                const unsigned at = binStart[t] + atomicAdd(&binCursor[t], 1u);
                if (at < P.binCap) binList[at] = s | (c == 2 ? 0x80000000u : 0u);
            }
        }
    }
}

// Workgroup-wide inclusive scan over 1024 threads (16 waves).
__device__ __forceinline__ unsigned scan1024(unsigned v, unsigned* sWave, unsigned& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    __syncthreads();  // sWave is free
    if (lane == 63) sWave[w] = x;
    __syncthreads();
    unsigned base = 0;
    total = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const unsigned t = sWave[i];
        base += i < w ? t : 0u;
        total += t;
    }
    return base + x;
}

// One 1024-thread workgroup, in chunks of 8192 tiles: the exclusive scan of
// the per-tile counts (list starts), then the raster work of each tile: parts
// of at most 1 << q.partShift entries, the first partial key block of a tile with
// several parts, the list of such tiles (multiList[0, numMulti)) and the list
// of empty tiles (multiList[tiles, tiles + numEmpty)).  Parts of more than
// kRasterPass entries are listed from the front of the item array and the
// others from its back, so that the queue hands out the long ones first.
// Counts are staged in LDS so that global loads are coalesced.
__global__ void __launch_bounds__(1024) k_tile_scan(const unsigned* __restrict__ cnt, unsigned* __restrict__ start,
                                                    int n, unsigned cap, RasterQueue q, DeviceCounters* ctr) {
    constexpr int kPer = 8, kChunk = 1024 * kPer;
    __shared__ unsigned sCnt[kChunk];
    __shared__ unsigned sWave[16];
    const int t = threadIdx.x;
    const int sh = q.partShift;
    const unsigned part = 1u << sh;
    unsigned carryS = 0, carryA = 0, carryL = 0, carryB = 0, carryM = 0, carryE = 0;
    for (int base = 0; base < n; base += kChunk) {
        __syncthreads();
        for (int i = t; i < kChunk; i += 1024) sCnt[i] = base + i < n ? cnt[base + i] : 0u;
        __syncthreads();
        unsigned sum = 0;
#pragma unroll
        for (int i = 0; i < kPer; i++) sum += sCnt[t * kPer + i];
        unsigned totS;
        unsigned run = carryS + scan1024(sum, sWave, totS) - sum;
        // starts out; the tile's list length within the capacity replaces its count in LDS
        unsigned a = 0, l = 0, b = 0, m = 0, e = 0;  // long parts, short parts, partials, multi, empty
        for (int i = 0; i < kPer; i++) {
            const int tile = base + t * kPer + i;
            const unsigned c = sCnt[t * kPer + i];
            const unsigned st = run;
            run += c;
            if (tile < n) start[tile] = st;
            const unsigned total = st >= cap ? 0u : min(c, cap - st);
            const unsigned p = tile < n ? (total + part - 1u) >> sh : 0u;
            sCnt[t * kPer + i] = total;
            const unsigned last = total - (p ? p - 1u : 0u) * part;  // entries of the last part
            const unsigned shortParts = p && last <= (unsigned)kRasterPass ? 1u : 0u;
            a += p - shortParts;
            l += shortParts;
            b += p > 1u ? p : 0u;
            m += p > 1u;
            e += tile < n && p == 0u;
        }
        unsigned totA, totL, totB, totM, totE;
        unsigned ra = carryA + scan1024(a, sWave, totA) - a;
        unsigned rl = carryL + scan1024(l, sWave, totL) - l;
        unsigned rb = carryB + scan1024(b, sWave, totB) - b;
        unsigned rm = carryM + scan1024(m, sWave, totM) - m;
        unsigned re = carryE + scan1024(e, sWave, totE) - e;
        for (int i = 0; i < kPer; i++) {
            const int tile = base + t * kPer + i;
            if (tile >= n) break;
            const unsigned total = sCnt[t * kPer + i];
            const unsigned p = (total + part - 1u) >> sh;
            const unsigned last = total - (p ? p - 1u : 0u) * part;
            const unsigned longParts = p && last <= (unsigned)kRasterPass ? p - 1u : p;
            for (unsigned k = 0; k < longParts; k++)
                if (ra + k < q.itemCap) q.items[ra + k] = make_uint2((unsigned)tile, k);
            if (longParts < p && rl < q.itemCap) q.items[q.itemCap - 1u - rl] = make_uint2((unsigned)tile, p - 1u);
            if (p > 1u) q.multiList[rm] = (unsigned)tile;
            if (p == 0u) q.multiList[n + re] = (unsigned)tile;
            q.tilePartial[tile] = p > 1u ? (int)rb : -1;
            ra += longParts;
            rl += p - longParts;
            rb += p > 1u ? p : 0u;
            rm += p > 1u;
            re += p == 0u;
        }
        carryS += totS;
        carryA += totA;
        carryL += totL;
        carryB += totB;
        carryM += totM;
        carryE += totE;
    }
    if (t == 0) {
        ctr->binTotal = carryS;
        if (carryS > cap) atomicOr(&ctr->status, kStatusBinOverflow);
        // on overflow nothing is rasterised; the host grows the buffers and redraws
        const bool fits = carryA + carryL <= q.itemCap && carryB <= q.partialCap;
        ctr->numItems = fits ? carryA + carryL : 0u;
        ctr->numLongItems = carryA;
        ctr->numMulti = fits ? carryM : 0u;
        ctr->numEmpty = fits ? carryE : 0u;
        ctr->numPartials = carryB;
        ctr->itemHead = 0u;
        if (!fits) atomicOr(&ctr->status, kStatusRasterOverflow);
    }
}

}  // namespace

void launch_bin(const BinParams& p, const Views& views, TriRecs recs, const int4* tileBox, unsigned* binCount,
                unsigned* binStart, unsigned* binCursor, unsigned* binList, unsigned* bigList, const RasterQueue& q,
                DeviceCounters* ctr, hipStream_t s) {
    const size_t slots = (size_t)p.ntri + p.clipCap;
    const dim3 g(dev::blocks(slots > 0 ? slots : 1, 256), views.count);
    hipLaunchKernelGGL(k_bin_small<false>, g, dim3(256), 0, s, p, views, recs, tileBox, binCount, binStart, binCursor,
                       binList, bigList, ctr);
    hipLaunchKernelGGL(k_bin_big<false>, dim3(256), dim3(256), 0, s, p, views, recs, tileBox, binCount, binStart,
                       binCursor, binList, bigList, ctr);
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, binCount, binStart, p.tiles, p.binCap, q, ctr);
    hipLaunchKernelGGL(k_bin_small<true>, g, dim3(256), 0, s, p, views, recs, tileBox, binCount, binStart, binCursor,
                       binList, bigList, ctr);
    hipLaunchKernelGGL(k_bin_big<true>, dim3(256), dim3(256), 0, s, p, views, recs, tileBox, binCount, binStart,
                       binCursor, binList, bigList, ctr);
}

}  // namespace ifrit_hip
