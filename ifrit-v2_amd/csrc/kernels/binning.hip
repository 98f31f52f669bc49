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

__device__ __forceinline__ int4 unpack_box(const uint4& e) {
    return make_int4((int)(e.y & 0xFFFFu), (int)(e.y >> 16), (int)(e.z & 0xFFFFu), (int)(e.z >> 16));
}

// Shared traversal of the small triangles over the bin queues of all views.
// Triangles deferred by the small-triangle pass (more tiles than the
// threshold): one workgroup per triangle from `first` in steps of `stride`,
// its threads over the tile rectangle.
template <bool SCATTER>
__device__ __forceinline__ void bin_big(const BinParams& P, const Views& V, const TriRecs& recs,
                                        const uint4* __restrict__ queueAll, unsigned* __restrict__ binCount,
                                        const unsigned* __restrict__ binStart, unsigned* __restrict__ binCursor,
                                        unsigned* __restrict__ binList, const unsigned* __restrict__ bigList,
                                        const DeviceCounters* ctr, unsigned first, unsigned stride) {
    const unsigned nbig = min(ctr->bigCount, P.bigCap);
    for (unsigned i = first; i < nbig; i += stride) {
        const unsigned qi = bigList[i];
        const ViewDesc& D = V.desc[qi / V.slotsPerView];
        const unsigned tileBase = D.tileBase;
        const int tilesX = D.tilesX;
        const uint4 e = queueAll[qi];
        const unsigned s = e.x;
        const int4 b = unpack_box(e);
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

// The queue entries of view v are numbered in chunks of 256 that continue
// from the previous view's, and chunk c goes to workgroup c mod gridDim.x, so
// the work spreads evenly whatever the views' sizes.  SCATTER = false counts
// (and defers the big triangles), SCATTER = true writes the lists.
template <bool SCATTER>
__global__ void __launch_bounds__(256) k_bin_small(BinParams P, Views V, TriRecs recs,
                                                   const uint4* __restrict__ queueAll, unsigned* __restrict__ binCount,
                                                   const unsigned* __restrict__ binStart,
                                                   unsigned* __restrict__ binCursor, unsigned* __restrict__ binList,
                                                   unsigned* __restrict__ bigList, DeviceCounters* ctr) {
    const unsigned G = gridDim.x;
    unsigned chunkBase = 0;
    for (int v = 0; v < V.count; v++) {
        const unsigned nq = min(V.ctr[v].queued, V.slotsPerView);
        const unsigned nch = (nq + 255u) / 256u;
        const unsigned c0 = (blockIdx.x + G - chunkBase % G) % G;
        chunkBase += nch;
        if (c0 >= nch) continue;
        const ViewDesc& D = V.desc[v];
        const unsigned tileBase = D.tileBase;
        const int tilesX = D.tilesX;
        const uint4* queue = queueAll + (size_t)v * V.slotsPerView;
        for (unsigned c = c0; c < nch; c += G) {
            const unsigned qi = c * 256u + threadIdx.x;
            const bool valid = qi < nq;
            uint4 e = make_uint4(0u, 0u, 0u, 0u);
            if (valid) e = queue[qi];
            const unsigned s = e.x;
            const int4 b = unpack_box(e);
            const int w = b.z - b.x + 1;
            const int area = valid ? w * (b.w - b.y + 1) : 0;
            const bool big = valid && area > P.bigThreshold;
            if (!SCATTER) {
                // deferred triangles go to the per-workgroup pass (as their queue position)
                const unsigned k = dev::wave_inc(&ctr->bigCount, 0, big);
                if (big) {
                    if (k < P.bigCap)
                        bigList[k] = (unsigned)v * V.slotsPerView + qi;
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
                int t = 0, cls = 0;
                if (active) {
                    const int tx = b.x + i % w, ty = b.y + i / w;
                    t = (int)tileBase + ty * tilesX + tx;
                    cls = tile_class(E, tx, ty);
                }
                const bool hit = active && cls != 0;
                if (!SCATTER) {
                    (void)dev::wave_inc(binCount, t, hit);
                } else {
                    // the list start is loaded before the atomic, so that the two round trips overlap
                    const unsigned st0 = hit ? binStart[t] : 0u;
                    const unsigned pos = dev::wave_inc(binCursor, t, hit);
                    if (hit) {
                        const unsigned at = st0 + pos;
                        if (at < P.binCap) binList[at] = s | (cls == 2 ? 0x80000000u : 0u);
                    }
                }
            }
        }
    }
    // the scatter pass also takes the deferred triangles, whose list the count pass completed
    if constexpr (SCATTER)
        bin_big<true>(P, V, recs, queueAll, binCount, binStart, binCursor, binList, bigList, ctr, blockIdx.x, G);
}

// The count pass's deferred triangles (the scatter pass takes its own at the
// end of k_bin_small: the list is complete by then).
__global__ void __launch_bounds__(256) k_bin_big_count(BinParams P, Views V, TriRecs recs,
                                                       const uint4* __restrict__ queueAll,
                                                       unsigned* __restrict__ binCount,
                                                       const unsigned* __restrict__ bigList, const DeviceCounters* ctr) {
    bin_big<false>(P, V, recs, queueAll, binCount, nullptr, nullptr, nullptr, bigList, ctr, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Raster planning over the tiles of all views, in three launches so that
// every phase is spread over the chip: (A) per block of 2048 tiles the sums
// of six per-tile quantities, (B) one workgroup's exclusive scan of the block
// sums and the totals, (C) per block the list starts and the work lists.
// A tile's list is cut into items of at most 1 << partShift entries; items
// of more than kRasterPass entries are listed from the front of the item
// array and the others from its back, so that the queue hands out the long
// ones first.  Tiles with several items get a partial key block per item and
// go to multiList[0, numMulti); empty tiles go to multiList[tiles, ...).
// Lists need not be in tile order: every tile's range is contiguous and the
// ranges are disjoint.
// ---------------------------------------------------------------------------
constexpr int kPlanBlock = kPlanTiles;       // tiles per block
constexpr int kPlanPer = kPlanBlock / 256;  // tiles per thread
constexpr int kPlanQ = 6;                   // entries, long items, short items, partial blocks, multi, empty

struct TilePlan {
    unsigned q[kPlanQ];
};

__device__ __forceinline__ TilePlan tile_plan(unsigned c, bool valid, int sh) {
    TilePlan w;
    const unsigned part = 1u << sh;
    const unsigned p = valid ? (c + part - 1u) >> sh : 0u;
    const unsigned last = c - (p ? p - 1u : 0u) * part;  // entries of the last item
    const unsigned shortItems = p && last <= (unsigned)kRasterPass ? 1u : 0u;
    w.q[0] = valid ? c : 0u;
    w.q[1] = p - shortItems;
    w.q[2] = shortItems;
    w.q[3] = p > 1u ? p : 0u;
    w.q[4] = p > 1u ? 1u : 0u;
    w.q[5] = valid && p == 0u ? 1u : 0u;
    return w;
}

// Tile of thread t's i-th slot in block b: a thread takes kPlanPer
// consecutive tiles, so the work lists follow the tile order (neighbouring
// tiles share triangles, and the raster runs neighbouring items together).
__device__ __forceinline__ int plan_tile(int b, int i, int t) { return b * kPlanBlock + t * kPlanPer + i; }

// Exclusive scan of kPlanQ values over a 256-thread workgroup; every thread must call it.
__device__ __forceinline__ void scan256x6(unsigned* v, unsigned (*sWave)[4], unsigned* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kPlanQ; k++) {
        const unsigned x = dev::wave_incl_scan(v[k]);
        if (lane == 63) sWave[k][w] = x;
        v[k] = x - v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPlanQ; k++) {
        unsigned base = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) base += i < w ? sWave[k][i] : 0u;
        v[k] += base;
        if (total) total[k] = sWave[k][0] + sWave[k][1] + sWave[k][2] + sWave[k][3];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_plan_sums(const unsigned* __restrict__ cnt, int n, int sh,
                                                   unsigned* __restrict__ blockSums) {
    __shared__ unsigned sWave[kPlanQ][4];
    const int t = threadIdx.x;
    unsigned v[kPlanQ] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < kPlanPer; i++) {
        const int tile = plan_tile(blockIdx.x, i, t);
        const TilePlan w = tile_plan(tile < n ? cnt[tile] : 0u, tile < n, sh);
#pragma unroll
        for (int k = 0; k < kPlanQ; k++) v[k] += w.q[k];
    }
    unsigned total[kPlanQ];
    scan256x6(v, sWave, total);
    if (t < kPlanQ) blockSums[blockIdx.x * kPlanQ + t] = total[t];
}

// The plan's totals: list entries, items, partial blocks, multi-part and
// empty tiles; the overflow flags; the item streams' heads reset.
__device__ __forceinline__ void plan_totals(const unsigned* carry, unsigned cap, const RasterQueue& q,
                                            DeviceCounters* ctr) {
    ctr->binTotal = carry[0];
    const bool binFits = carry[0] <= cap;
    if (!binFits) atomicOr(&ctr->status, kStatusBinOverflow);
    const bool fits = carry[1] + carry[2] <= q.itemCap && carry[3] <= q.partialCap;
    if (!fits) atomicOr(&ctr->status, kStatusRasterOverflow);
    const bool go = binFits && fits;
    ctr->numItems = go ? carry[1] + carry[2] : 0u;
    ctr->numLongItems = carry[1];
    ctr->numPartials = carry[3];
    ctr->numMulti = go ? carry[4] : 0u;
    ctr->numEmpty = go ? carry[5] : 0u;
    for (int k = 0; k < kItemStreams; k++) ctr->itemHeads[k][0] = 0u;
}

// One 1024-thread workgroup: exclusive offsets of the block sums (in place)
// and the totals.  On a bin or work-list overflow nothing is rasterised; the
// host grows the buffers and draws again.
__global__ void __launch_bounds__(1024) k_plan_scan(unsigned* __restrict__ blockSums, int nblk, unsigned cap,
                                                    RasterQueue q, DeviceCounters* ctr) {
    __shared__ unsigned sWave[kPlanQ][16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    unsigned carry[kPlanQ] = {0, 0, 0, 0, 0, 0};
    for (int base = 0; base < nblk; base += 1024) {
        // the six quantities of block b scanned together: two barriers per 1024 blocks
        const int b = base + t;
        unsigned v[kPlanQ], x[kPlanQ];
#pragma unroll
        for (int k = 0; k < kPlanQ; k++) {
            v[k] = b < nblk ? blockSums[b * kPlanQ + k] : 0u;
            x[k] = v[k];
        }
#pragma unroll
        for (int k = 0; k < kPlanQ; k++) x[k] = dev::wave_incl_scan(x[k]);
        __syncthreads();  // sWave is free
        if (lane == 63)
#pragma unroll
            for (int k = 0; k < kPlanQ; k++) sWave[k][w] = x[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPlanQ; k++) {
            unsigned pre = 0, tot = 0;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                pre += i < w ? sWave[k][i] : 0u;
                tot += sWave[k][i];
            }
            if (b < nblk) blockSums[b * kPlanQ + k] = carry[k] + pre + x[k] - v[k];
            carry[k] += tot;
        }
    }
    if (t == 0) plan_totals(carry, cap, q, ctr);
}

// Writes the tiles' list starts and work items.  SCAN: blockOffs holds the
// blocks' own sums (at most 256 blocks), which every block scans itself, and
// block 0 also stores the totals, in place of k_plan_scan; otherwise
// blockOffs holds the blocks' offsets from k_plan_scan.
template <bool SCAN>
__global__ void __launch_bounds__(256) k_plan_write(const unsigned* __restrict__ cnt, unsigned* __restrict__ start,
                                                    int n, int sh, const unsigned* __restrict__ blockOffs, int nblk,
                                                    unsigned cap, RasterQueue q, DeviceCounters* ctr) {
    __shared__ unsigned sWave[kPlanQ][4];
    __shared__ unsigned sOff[kPlanQ];
    const int t = threadIdx.x;
    if constexpr (SCAN) {
        unsigned b[kPlanQ], tot[kPlanQ];
#pragma unroll
        for (int k = 0; k < kPlanQ; k++) b[k] = t < nblk ? blockOffs[t * kPlanQ + k] : 0u;
        scan256x6(b, sWave, tot);
        if (t == (int)blockIdx.x)
#pragma unroll
            for (int k = 0; k < kPlanQ; k++) sOff[k] = b[k];
        if (blockIdx.x == 0 && t == 0) plan_totals(tot, cap, q, ctr);
        __syncthreads();
    }
    unsigned c[kPlanPer];
    unsigned v[kPlanQ] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < kPlanPer; i++) {
        const int tile = plan_tile(blockIdx.x, i, t);
        c[i] = tile < n ? cnt[tile] : 0u;
        const TilePlan w = tile_plan(c[i], tile < n, sh);
#pragma unroll
        for (int k = 0; k < kPlanQ; k++) v[k] += w.q[k];
    }
    scan256x6(v, sWave, nullptr);
#pragma unroll
    for (int k = 0; k < kPlanQ; k++) v[k] += SCAN ? sOff[k] : blockOffs[blockIdx.x * kPlanQ + k];
    const unsigned part = 1u << sh;
    for (int i = 0; i < kPlanPer; i++) {
        const int tile = plan_tile(blockIdx.x, i, t);
        if (tile >= n) break;
        const TilePlan w = tile_plan(c[i], true, sh);
        const unsigned st = v[0], p = w.q[1] + w.q[2];
        start[tile] = st;
        for (unsigned k = 0; k < p; k++) {
            const unsigned first = st + k * part;
            const uint4 item = make_uint4((unsigned)tile, first, min(part, c[i] - k * part),
                                          p > 1u ? v[3] + k : 0xFFFFFFFFu);
            if (k < w.q[1]) {
                if (v[1] + k < q.itemCap) q.items[v[1] + k] = item;
            } else if (v[2] < q.itemCap) {
                q.items[q.itemCap - 1u - v[2]] = item;
            }
        }
        if (p > 1u) q.multiList[v[4]] = (unsigned)tile;
        if (p == 0u) q.multiList[n + v[5]] = (unsigned)tile;
        q.tilePartial[tile] = p > 1u ? (int)v[3] : -1;
#pragma unroll
        for (int k = 0; k < kPlanQ; k++) v[k] += w.q[k];
    }
}

}  // namespace

void launch_bin(const BinParams& p, const Views& views, TriRecs recs, const uint4* binQueue, unsigned* binCount,
                unsigned* binStart, unsigned* binCursor, unsigned* binList, unsigned* bigList, const RasterQueue& q,
                DeviceCounters* ctr, hipStream_t s) {
    const dim3 g(p.grid);
    hipLaunchKernelGGL(k_bin_small<false>, g, dim3(256), 0, s, p, views, recs, binQueue, binCount, binStart, binCursor,
                       binList, bigList, ctr);
    hipLaunchKernelGGL(k_bin_big_count, dim3(256), dim3(256), 0, s, p, views, recs, binQueue, binCount, bigList,
                       ctr);
    const int nblk = (p.tiles + kPlanBlock - 1) / kPlanBlock;
    hipLaunchKernelGGL(k_plan_sums, dim3(nblk), dim3(256), 0, s, binCount, p.tiles, q.partShift, q.planSums);
    if (nblk <= 256) {
        // few blocks: each block scans the block sums itself (no third launch)
        hipLaunchKernelGGL(k_plan_write<true>, dim3(nblk), dim3(256), 0, s, binCount, binStart, p.tiles,
                           q.partShift, q.planSums, nblk, p.binCap, q, ctr);
    } else {
        hipLaunchKernelGGL(k_plan_scan, dim3(1), dim3(1024), 0, s, q.planSums, nblk, p.binCap, q, ctr);
        hipLaunchKernelGGL(k_plan_write<false>, dim3(nblk), dim3(256), 0, s, binCount, binStart, p.tiles,
                           q.partShift, q.planSums, nblk, p.binCap, q, ctr);
    }
    hipLaunchKernelGGL(k_bin_small<true>, g, dim3(256), 0, s, p, views, recs, binQueue, binCount, binStart, binCursor,
                       binList, bigList, ctr);
}

}  // namespace ifrit_hip
