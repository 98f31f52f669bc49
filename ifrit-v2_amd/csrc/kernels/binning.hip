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

__device__ __forceinline__ unsigned num_slots(const BinParams& P, const DeviceCounters* ctr) {
    return (unsigned)P.ntri + min(ctr->clipCount, P.clipCap);
}

// Shared traversal of the small triangles.  SCATTER = false counts,
// SCATTER = true writes the lists.
template <bool SCATTER>
__global__ void __launch_bounds__(256) k_bin_small(BinParams P, TriRecs recs,
                                                   const int4* __restrict__ tileBox, unsigned* __restrict__ binCount,
                                                   const unsigned* __restrict__ binStart,
                                                   unsigned* __restrict__ binCursor, unsigned* __restrict__ binList,
                                                   unsigned* __restrict__ bigList, DeviceCounters* ctr) {
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = s < num_slots(P, ctr);
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
            t = ty * P.tilesX + tx;
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
__global__ void __launch_bounds__(256) k_bin_big(BinParams P, TriRecs recs,
                                                 const int4* __restrict__ tileBox, unsigned* __restrict__ binCount,
                                                 const unsigned* __restrict__ binStart, unsigned* __restrict__ binCursor,
                                                 unsigned* __restrict__ binList, const unsigned* __restrict__ bigList,
                                                 const DeviceCounters* ctr) {
    const unsigned nbig = min(ctr->bigCount, P.bigCap);
    for (unsigned i = blockIdx.x; i < nbig; i += gridDim.x) {
        const unsigned s = bigList[i];
        const int4 b = tileBox[s];
        const Edges E = load_edges(recs, s);
        const int w = b.z - b.x + 1;
        const int area = w * (b.w - b.y + 1);
        for (int q = threadIdx.x; q < area; q += blockDim.x) {
            const int tx = b.x + q % w, ty = b.y + q / w;
            const int c = tile_class(E, tx, ty);
            if (!c) continue;
            const int t = ty * P.tilesX + tx;
            if (!SCATTER) {
                atomicAdd(&binCount[t], 1u);
            } else {
                const unsigned at = binStart[t] + atomicAdd(&binCursor[t], 1u);
                if (at < P.binCap) binList[at] = s | (c == 2 ? 0x80000000u : 0u);
            }
        }
    }
}

// Exclusive scan of the per-tile counts by one 1024-thread workgroup.
__global__ void __launch_bounds__(1024) k_bin_scan(const unsigned* __restrict__ cnt, unsigned* __restrict__ start,
                                                   int n, unsigned cap, DeviceCounters* ctr) {
    __shared__ unsigned part[1024];
    const int per = (n + 1023) / 1024;
    const int b0 = threadIdx.x * per;
    unsigned sum = 0;
    for (int i = 0; i < per; i++)
        if (b0 + i < n) sum += cnt[b0 + i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const unsigned v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned run = part[threadIdx.x] - sum;
    for (int i = 0; i < per; i++) {
        if (b0 + i < n) {
            start[b0 + i] = run;
            run += cnt[b0 + i];
        }
    }
    if (threadIdx.x == 1023) {
        ctr->binTotal = part[1023];
        if (part[1023] > cap) atomicOr(&ctr->status, kStatusBinOverflow);
    }
}

}  // namespace

void launch_bin(const BinParams& p, TriRecs recs, const int4* tileBox, unsigned* binCount, unsigned* binStart,
                unsigned* binCursor, unsigned* binList, unsigned* bigList, DeviceCounters* ctr, hipStream_t s) {
    const size_t slots = (size_t)p.ntri + p.clipCap;
    const int tiles = p.tilesX * p.tilesY;
    const dim3 g(dev::blocks(slots > 0 ? slots : 1, 256));
    hipLaunchKernelGGL(k_bin_small<false>, g, dim3(256), 0, s, p, recs, tileBox, binCount, binStart, binCursor,
                       binList, bigList, ctr);
    hipLaunchKernelGGL(k_bin_big<false>, dim3(256), dim3(256), 0, s, p, recs, tileBox, binCount, binStart, binCursor,
                       binList, bigList, ctr);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, binCount, binStart, tiles, p.binCap, ctr);
    hipLaunchKernelGGL(k_bin_small<true>, g, dim3(256), 0, s, p, recs, tileBox, binCount, binStart, binCursor,
                       binList, bigList, ctr);
    hipLaunchKernelGGL(k_bin_big<true>, dim3(256), dim3(256), 0, s, p, recs, tileBox, binCount, binStart, binCursor,
                       binList, bigList, ctr);
}

}  // namespace ifrit_hip
