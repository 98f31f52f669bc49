// Fine raster and early Z on gfx950.
//
// Coverage follows the reference's hierarchy exactly
// (projects/softgraphics/src/engine/tileraster/TileRasterWorker.cpp:787-1223,
// "TRW"): fully covered tiles from the binner, then 4x4 blocks by trivial
// reject / accept, then 4x2 groups, 2x2 quads and single pixels at 1 sample
// per pixel, each class with its own depth expression (TRW:2244-2918).  With
// MSAA every sample of a partially covered block is tested individually.
//
// Depth test.  The reference shades a tile's fragments in primitive order and
// keeps a fragment when !(depth >= stored) (TRW:2286-2290, 2482-2486,
// 2828-2832).  For finite depths the survivor is the lexicographic minimum of
// (depth, prim * 4 + fan), which is order-independent: visibility is held as
// 64-bit keys (ordered depth << 32 | prim * 4 + fan) merged with atomic min.
//
// Work expansion.  Two lanes per list entry (one per half tile) run the block
// tests and build the entry's candidate mask: pixels (MSAA) or 4x2 groups (1
// sample) of the blocks that pass trivial reject, pruned by a monotone corner
// bound of the pixel test, which is exact.  A workgroup scan turns the masks into a flat
// list of (entry, pixel or group) units in LDS, and every lane then takes
// one unit: its sample tests, depths and LDS atomic-min merges.  Lanes stay
// busy whatever the mix of triangle sizes, which one-triangle-per-lane loops
// do not achieve.
//
// Scheduling.  A tile's list is cut into parts of at most 1 << partShift entries and
// all parts form one work queue that a persistent grid drains, so a few
// tiles holding thousands of sub-pixel triangles do not serialise on one
// compute unit.  A part goes through LDS in passes of kPass entries, all
// merging into the tile's keys.  Each part rasterises into LDS; a tile with one part writes
// its outputs from LDS, each part of a longer list stores its LDS keys to a
// private partial block and a second kernel takes the minimum over the parts
// and writes the outputs (it also clears tiles with an empty list).  No
// workgroup waits on another and no device-scope fence is issued (on the
// multi-XCD MI355X each one writes back an L2).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "raster_common.h"

namespace ifrit_hip {

namespace {

using namespace dev;
using namespace rc;

constexpr int kPass = kRasterPass;  // the most list entries per pass through LDS

// inside() with the three edge results combined by bitwise and: the same
// arithmetic, and the compiler keeps the sample loop free of branches.
__device__ __forceinline__ bool inside_and(const Edges& E, float x, float y) {
    const bool a = fmaf(E.y[0], y, E.x[0] * x) < ifx::neg(E.z[0]);
    const bool b = fmaf(E.y[1], y, E.x[1] * x) < ifx::neg(E.z[1]);
    const bool c = fmaf(E.y[2], y, E.x[2] * x) < ifx::neg(E.z[2]);
    return a & b & c;
}

__device__ __forceinline__ void merge(unsigned long long* cell, float d, unsigned tie) {
    if (d == d) atomicMin(cell, frag_key(d, tie));
}

// Rejects every integer point of [x, x + w] x [y, y + h] when an edge
// excludes the corner that minimises it; rounding is monotone, so no such
// point can pass inside() when this returns false.
__device__ __forceinline__ bool rect_may_cover(const Edges& E, float x, float y, float w, float h) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float cx = E.x[k] < 0.0f ? x + w : x;
        const float cy = E.y[k] < 0.0f ? y + h : y;
        ok &= fmaf(E.y[k], cy, E.x[k] * cx) < ifx::neg(E.z[k]);
    }
    return ok;
}

// rect_may_cover of the 16 unit squares of the 4x4 block at (x0, y0), bit
// 4 * row + column: the same products and fused sums per square (corners
// are small integers, so x0 + column + 1 is exact either way), with each
// edge's column products and row corners shared and no branches.
__device__ __forceinline__ unsigned block_square_mask(const Edges& E, float x0, float y0) {
    float ax[3][4], cy[3][4], nz[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float ox = E.x[k] < 0.0f ? 1.0f : 0.0f, oy = E.y[k] < 0.0f ? 1.0f : 0.0f;
        nz[k] = ifx::neg(E.z[k]);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ax[k][i] = E.x[k] * ((x0 + (float)i) + ox);
            cy[k][i] = (y0 + (float)i) + oy;
        }
    }
    unsigned m = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int i = q & 3, j = q >> 2;
        const unsigned c = (unsigned)(fmaf(E.y[0], cy[0][j], ax[0][i]) < nz[0]) &
                           (unsigned)(fmaf(E.y[1], cy[1][j], ax[1][i]) < nz[1]) &
                           (unsigned)(fmaf(E.y[2], cy[2][j], ax[2][i]) < nz[2]);
        m |= c << q;
    }
    return m;
}

// Exclusive scan over an NT-thread workgroup (64 or 256); every thread must call it.
template <int NT>
__device__ __forceinline__ unsigned scan_nt(unsigned v, unsigned* sWave, unsigned& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned x = wave_incl_scan(v);
    if constexpr (NT == 64) {
        total = __shfl(x, 63, 64);
        return x - v;
    } else {
        if (lane == 63) sWave[w] = x;
        __syncthreads();
        unsigned base = 0;
#pragma unroll
        for (int i = 0; i < NT / 64; i++) base += i < w ? sWave[i] : 0u;
        total = 0;
#pragma unroll
        for (int i = 0; i < NT / 64; i++) total += sWave[i];
        return base + x - v;
    }
}

// Sum over an NT-thread workgroup (all threads must call it).
template <int NT>
__device__ __forceinline__ unsigned long long sum_nt(unsigned long long v) {
    if constexpr (NT == 64) {
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        return v;
    } else {
        return block_sum256(v);
    }
}

// block_test of the eight 4x4 blocks of tile half h (blocks 8h .. 8h + 7,
// block rows 2h and 2h + 1): bit j of tr when block 8h + j is not rejected,
// of ta when it is trivially accepted; a tile-accepted entry (full) takes
// every block.
__device__ __forceinline__ void half_blocks(const Edges& E, bool full, int tx, int ty, int h, unsigned& tr,
                                            unsigned& ta) {
    if (full) {
        tr = ta = 0xFFu;
        return;
    }
    const float tileMinX = (float)(tx * 16) * 1.0f, tileMinY = (float)(ty * 16) * 1.0f;
    tr = 0;
    ta = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int bi = 8 * h + j;
        int trn, tan;
        block_test(E, fmaf((float)(bi & 3), 4.0f, tileMinX), fmaf((float)(bi >> 2), 4.0f, tileMinY), trn, tan);
        tr |= (unsigned)(trn == 3) << j;
        ta |= (unsigned)(tan == 3) << j;
    }
}

// Pixels of a 4x4 block inside the frame (bit 4 * row + column), given the
// frame's columns and rows left from the block's corner.
__device__ __forceinline__ unsigned block_fb_mask(int ncols, int nrows) {
    const unsigned c = ncols >= 4 ? 0xFu : (ncols <= 0 ? 0u : (1u << ncols) - 1u);
    const unsigned r = nrows >= 4 ? 0xFFFFu : (nrows <= 0 ? 0u : (1u << (4 * nrows)) - 1u);
    return (c * 0x1111u) & r;
}

// Candidate 4x2 groups of one half of a tile at 1 sample per pixel for one
// list entry (two lanes share an entry): 16 bits (block - 8h) * 2 + row
// pair, masked by fbMask.  Also returns the trivially accepted blocks of the
// half (bit block - 8h).
__device__ __forceinline__ void group_mask_half(const Edges& E, bool full, int tx, int ty, int h, unsigned fbMask,
                                                unsigned& m, unsigned& ta) {
    unsigned tr;
    half_blocks(E, full, tx, ty, h, tr, ta);
    if (full) {
        m = fbMask;
        return;
    }
    m = 0;
    while (tr) {
        const int j = __ffs(tr) - 1;
        tr &= tr - 1u;
        const int bi = 8 * h + j;
        const float x0 = (float)(tx * 16 + (bi & 3) * 4), y0 = (float)(ty * 16 + (bi >> 2) * 4);
        unsigned g = 3u;
        if (!((ta >> j) & 1u))
            g = (unsigned)rect_may_cover(E, x0, y0, 3.0f, 1.0f) |
                ((unsigned)rect_may_cover(E, x0, y0 + 2.0f, 3.0f, 1.0f) << 1);
        m |= g << (j * 2);
    }
    m &= fbMask;
}

// Coverage class and depth of the winning triangle at a 1-sample pixel,
// recomputed with the same decisions as the raster phases.
__device__ __forceinline__ unsigned winner_class(const TriRec& R, int tx, int ty, float fpx, float fpy, float bMinX,
                                                 float bMinY, float& d, bool tagging) {
    const Edges E = load_edges(R);
    bool accepted = tile_class(E, tx, ty) == 2;
    if (!accepted) {
        int trn, tan;
        block_test(E, bMinX, bMinY, trn, tan);
        accepted = tan == 3;
    }
    const unsigned c = pixel_class(E, accepted, fpx, fpy, bMinX, tagging);
    d = class_depth(c, R.f1, R.f2, R.f3, fpx, fpy, tagging);
    return c;
}

// Fragment keys: the order key of the depth in the high word; in the low
// word the primitive key prim * 4 + fan (S > 1), or that key times 4 plus
// the fragment's coverage class at 1 sample per pixel (a pixel never holds
// two fragments of one (prim, fan), so the class bits never decide a
// minimum; primitives stay below 2^28 there).
template <int S>
__device__ __forceinline__ unsigned key_tie(unsigned long long k) {
    return S == 1 ? (unsigned)k >> 2 : (unsigned)k;
}

// The record slot of a key's triangle: ptr[prim] + fan from the slot map, or,
// in a view where no triangle crossed w = 0 (direct != all ones), direct +
// prim: the setup kernel puts triangle p in slot slotBase + p, and only the
// clip kernel moves a triangle's slots.
struct SlotMap {
    const unsigned* ptr;
    unsigned direct;
};

// Record slots of a pixel's S final keys (slotOf[prim] + fan), all S lookups
// issued before the first is used; a key that holds no fragment looks up
// primitive 0 and its slot is not used.
template <int S>
__device__ __forceinline__ void key_slots(const SlotMap& slotOf, const unsigned long long* keys,
                                          unsigned long long none, unsigned* slots) {
    if (slotOf.direct != 0xffffffffu) {
        // no triangle of the view was clipped: triangle p sits in slot slotBase + p
#pragma unroll
        for (int s = 0; s < S; s++) slots[s] = slotOf.direct + (key_tie<S>(keys[s]) >> 2);
        return;
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        const unsigned tie = key_tie<S>(keys[s]);
        slots[s] = slotOf.ptr[keys[s] != none ? tie >> 2 : 0u] + (tie & 3u);
    }
}

// Writes the outputs of one pixel from its S final keys: depth and record
// slot per sample (the primitive ID is the slot's key >> 2, expanded only on
// readback).
template <int S>
__device__ __forceinline__ void write_pixel(const RasterParams& P, const TriRecs& recs, const SlotMap& slotOf,
                                            const PixelCtx& C, const unsigned long long* keys, float* visDepth,
                                            uint32_t* visSlot, uint8_t* visClass) {
    if (!C.inFb || !scissor_pass(P.scissor, C.fpx, C.fpy)) return;
    unsigned long long prev[S];
#pragma unroll
    for (int s = 0; s < S; s++) prev[s] = frag_key(P.clear ? P.clearDepth : visDepth[C.pix * S + s], 0u);
    unsigned slots[S];
    key_slots<S>(slotOf, keys, P.clear ? frag_key(P.clearDepth, 0u) : ~0ull, slots);
    unsigned cls = 0;
    bool won = false;
#pragma unroll
    for (int s = 0; s < S; s++) {
        const unsigned long long k = keys[s];
        if (k == prev[s]) {
            if (P.clear) {
                visDepth[C.pix * S + s] = P.clearDepth;
                visSlot[C.pix * S + s] = 0xffffffffu;
            }
            continue;
        }
        won = true;
        const unsigned slot = slots[s];
        float d = order_key_inv((unsigned)(k >> 32));
        if (S == 1 && (unsigned)(k >> 32) == 0x80000000u) {
            // +0 and -0 share a key: recompute the winner's own depth bits from its record
            cls = winner_class(load_rec(recs, slot), C.tx, C.ty, C.fpx, C.fpy, C.bMinX, C.bMinY, d, P.tagging);
        } else if (S == 1) {
            cls = (unsigned)k & 3u;
        } else if (S > 1 && (unsigned)(k >> 32) == 0x80000000u) {
            // +0 and -0 share a key; recompute to return the winner's own bits
            d = depth256(recs.f1[slot], recs.f2[slot], recs.f3[slot], C.fpx + P.sx[s], C.fpy + P.sy[s]);
        }
        visDepth[C.pix * S + s] = d;
        visSlot[C.pix * S + s] = slot;
    }
    if (S == 1 && visClass && (won || P.clear)) visClass[C.pix] = class_out(cls, P.tagging);
}

// One tile's outputs staged in LDS in global order (pixel-major, sample-
// minor) so that they leave with 16-byte stores of consecutive addresses;
// per-lane stores of one pixel's S samples would touch a partial line per
// sample.
template <int S>
struct TileOut {
    float d[256 * S];
    unsigned slot[256 * S];
};

// Outputs of one pixel of a cleared draw from its S final keys: depth and
// record slot per sample (the clear values where no fragment won) and, at 1
// sample per pixel, the coverage class.
template <int S>
__device__ __forceinline__ unsigned pixel_outputs(const RasterParams& P, const TriRecs& recs, const SlotMap& slotOf,
                                                  const PixelCtx& C, const unsigned long long* keys, float* d,
                                                  unsigned* slot) {
    unsigned cls = 0;
    const unsigned long long clearKey = frag_key(P.clearDepth, 0u);
    unsigned slots[S];
    key_slots<S>(slotOf, keys, clearKey, slots);
#pragma unroll
    for (int s = 0; s < S; s++) {
        const unsigned long long k = keys[s];
        d[s] = P.clearDepth;
        slot[s] = 0xffffffffu;
        if (k != clearKey) {
            slot[s] = slots[s];
            d[s] = order_key_inv((unsigned)(k >> 32));
            if (S == 1 && (unsigned)(k >> 32) == 0x80000000u) {
                // +0 and -0 share a key: recompute the winner's own depth bits from its record
                cls = winner_class(load_rec(recs, slot[s]), C.tx, C.ty, C.fpx, C.fpy, C.bMinX, C.bMinY, d[s],
                                   P.tagging);
            } else if (S == 1) {
                cls = (unsigned)k & 3u;
            } else if (S > 1 && (unsigned)(k >> 32) == 0x80000000u) {
                // +0 and -0 share a key; recompute to return the winner's own bits
                d[s] = depth256(recs.f1[slot[s]], recs.f2[slot[s]], recs.f3[slot[s]], C.fpx + P.sx[s],
                                C.fpy + P.sy[s]);
            }
        }
    }
    return cls;
}

// Outputs of one pixel of a cleared draw into the staging buffer; the 1-spp
// class byte goes straight out.
template <int S>
__device__ __forceinline__ void stage_pixel(const RasterParams& P, const TriRecs& recs, const SlotMap& slotOf,
                                            const PixelCtx& C, int t, const unsigned long long* keys, TileOut<S>& O,
                                            uint8_t* visClass, const float* visDepth, const uint32_t* visSlot) {
    if (!C.inFb) return;
    if (!scissor_pass(P.scissor, C.fpx, C.fpy)) {
        // outside the scissor: the tile's line stores write the pixel's current contents back
#pragma unroll
        for (int s = 0; s < S; s++) {
            O.d[t * S + s] = visDepth[C.pix * S + s];
            O.slot[t * S + s] = visSlot[C.pix * S + s];
        }
        return;
    }
    float d[S];
    unsigned slot[S];
    const unsigned cls = pixel_outputs<S>(P, recs, slotOf, C, keys, d, slot);
#pragma unroll
    for (int s = 0; s < S; s++) {
        O.d[t * S + s] = d[s];
        O.slot[t * S + s] = slot[s];
    }
    if (S == 1 && visClass) visClass[C.pix] = class_out(cls, P.tagging);
}

// Outputs of one pixel of a cleared draw straight to memory with 16-byte
// stores (S a multiple of 4: a pixel's samples fill whole 16-byte words, and
// a row of pixels whole lines).
template <int S>
__device__ __forceinline__ void emit_pixel(const RasterParams& P, const TriRecs& recs, const SlotMap& slotOf,
                                           const PixelCtx& C, const unsigned long long* keys, float* visDepth,
                                           uint32_t* visSlot) {
    static_assert(S % 4 == 0, "whole 16-byte words per pixel");
    if (!C.inFb || !scissor_pass(P.scissor, C.fpx, C.fpy)) return;
    float d[S];
    unsigned slot[S];
    pixel_outputs<S>(P, recs, slotOf, C, keys, d, slot);
    float4* const dd = reinterpret_cast<float4*>(visDepth + C.pix * S);
    uint4* const ds = reinterpret_cast<uint4*>(visSlot + C.pix * S);
#pragma unroll
    for (int v = 0; v < S / 4; v++) {
        dd[v] = make_float4(d[4 * v], d[4 * v + 1], d[4 * v + 2], d[4 * v + 3]);
        ds[v] = make_uint4(slot[4 * v], slot[4 * v + 1], slot[4 * v + 2], slot[4 * v + 3]);
    }
}

// Stores a staged tile with 16-byte stores; partial columns at the right
// edge fall back to scalars.  visDepth, visSlot and visClass point at the
// view's outputs; a view's first pixel starts a 16-byte line when S > 1 or its
// pixel base is a multiple of 4.
template <int S, int NT = 256>
__device__ __forceinline__ void store_tile(const ViewGeom& G, int tx, int ty, const TileOut<S>* O,
                                           float* visDepth, uint32_t* visSlot) {
    constexpr int perRow = 16 * S / 4;  // float4 per tile row
    const int cols = min(16, G.width - tx * 16), rows = min(16, G.height - ty * 16);
    const bool vec = cols == 16 && (S > 1 || ((G.width & 3) == 0 && (G.pixBase & 3) == 0));
    for (int i = threadIdx.x; i < 64 * S; i += NT) {
        const int row = i / perRow, q = i - row * perRow;
        if (row >= rows) continue;
        const size_t g = ((size_t)(ty * 16 + row) * G.width + tx * 16) * S + q * 4;
        const int l = row * 16 * S + q * 4;
        const float4 d = *reinterpret_cast<const float4*>(&O->d[l]);
        const uint4 sl = *reinterpret_cast<const uint4*>(&O->slot[l]);
        if (vec) {
            *reinterpret_cast<float4*>(&visDepth[g]) = d;
            *reinterpret_cast<uint4*>(&visSlot[g]) = sl;
        } else {
            const float dv[4] = {d.x, d.y, d.z, d.w};
            const unsigned sv[4] = {sl.x, sl.y, sl.z, sl.w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if ((q * 4 + e) / S >= cols) continue;
                visDepth[g + e] = dv[e];
                visSlot[g + e] = sv[e];
            }
        }
    }
}

// Clear values of one tile whose list is empty (depth clearDepth, slot
// 0xffffffff and, at 1 sample per pixel, class 0) by NT threads, with 16-byte
// stores where the tile's rows are whole.
template <int S, int NT>
__device__ __forceinline__ void clear_tile(const ViewGeom& G, float clearDepth, const Scissor& sc,
                                           float* visDepthAll, uint32_t* visSlotAll, uint8_t* visClassAll) {
    const int tx = G.tile % G.tilesX, ty = G.tile / G.tilesX;
    const int cols = min(16, G.width - tx * 16), rows = min(16, G.height - ty * 16);
    const bool whole = scissor_all(sc, tx * 16, ty * 16, tx * 16 + cols - 1, ty * 16 + rows - 1);
    const bool vec = whole && cols == 16 && (S > 1 || ((G.width & 3) == 0 && (G.pixBase & 3) == 0));
    float* const visDepth = visDepthAll + G.pixBase * S;
    uint32_t* const visSlot = visSlotAll + G.pixBase * S;
    constexpr int perRow = 16 * S / 4;  // 16-byte words per tile row
    for (int i = threadIdx.x; i < 64 * S; i += NT) {
        const int row = i / perRow, q = i - row * perRow;
        if (row >= rows) continue;
        const size_t g = ((size_t)(ty * 16 + row) * G.width + tx * 16) * S + q * 4;
        if (vec) {
            *reinterpret_cast<float4*>(&visDepth[g]) = make_float4(clearDepth, clearDepth, clearDepth, clearDepth);
            *reinterpret_cast<uint4*>(&visSlot[g]) = make_uint4(~0u, ~0u, ~0u, ~0u);
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int col = (q * 4 + e) / S;
                if (col >= cols) continue;
                if (!whole && !scissor_pass(sc, (float)(tx * 16 + col), (float)(ty * 16 + row))) continue;
                visDepth[g + e] = clearDepth;
                visSlot[g + e] = 0xffffffffu;
            }
        }
    }
    if (S == 1 && visClassAll) {
        uint8_t* const visClass = visClassAll + G.pixBase;
        for (int i = threadIdx.x; i < 256; i += NT) {
            const int row = i >> 4, col = i & 15;
            if (row < rows && col < cols &&
                (whole || scissor_pass(sc, (float)(tx * 16 + col), (float)(ty * 16 + row))))
                visClass[(size_t)(ty * 16 + row) * G.width + tx * 16 + col] = 0;
        }
    }
}

// LDS layout of the tile's keys: sample-major, 256 pixels per sample.
// Padded layouts lowered the bank-conflict share of LDS-active cycles (one
// key after every 32: 40 % to 35 %; four after every 32: 48 % to 31 %) but
// not the kernel time (unchanged, then 3 % slower from the extra address
// arithmetic): conflicts cost about 4 % of its cycles, and the kernel is
// bound by VALU issue.
constexpr int kVisStride = 256;
__device__ __forceinline__ int vis_at(int s, int p) { return s * kVisStride + p; }

// NT threads per tile: 256 (one pixel per thread) or, at 1 sample per pixel,
// 64 (one wave per tile, four pixels per lane: more tiles in flight per
// compute unit, for many small tiles).  Passes hold NT / 2 entries.  Held to
// 128 VGPRs so that four workgroups of 256 share a compute unit.
template <int S, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_raster(RasterParams P, Views V, TriRecs recs, const unsigned* __restrict__ slotOfAll,
         const unsigned* __restrict__ binList, const uint4* __restrict__ items, unsigned itemCap,
         const unsigned* __restrict__ emptyList,
         unsigned long long* __restrict__ partials, float* __restrict__ visDepthAll, uint32_t* __restrict__ visSlotAll,
         uint8_t* __restrict__ visClassAll, StatPartials st, DeviceCounters* ctr) {
    constexpr int PASS = NT / 2;            // list entries per pass through LDS
    constexpr int UNITS = S == 1 ? NT * 16 : NT * 8;  // units expanded into LDS per round
    constexpr int PAIRS = S == 1 ? 1 : 16 * PASS;     // (entry, block) pairs of a pass (S > 1)
    constexpr int PPT = 256 / NT;           // tile pixels per thread
    static_assert(PASS <= kPass, "units encode the entry in 7 bits");
    // key of sample s of tile pixel p at vis_at(s, p)
    __shared__ unsigned long long sVis[S * kVisStride];
    static_assert(sizeof(TileOut<S>) <= sizeof(sVis), "output staging reuses sVis");
    __shared__ TriRec sRec[PASS];
    __shared__ unsigned char sTa[2 * PASS];  // accepted blocks of each half
    __shared__ unsigned short sUnit[UNITS];
    __shared__ unsigned short sPair[PAIRS];
    __shared__ unsigned sWave[4];
    __shared__ unsigned sItem;

    const int t = threadIdx.x;
    const unsigned numItems = ctr->numItems, numLong = ctr->numLongItems;
    // tiles with an empty list follow the items: their clear values, KE tiles
    // per claim at 1 sample per pixel (a small write each)
    constexpr unsigned KE = S == 1 ? 8u : 1u;
    const unsigned numEmpty = P.clear ? ctr->numEmpty : 0u;
    const unsigned numEmptyItems = (numEmpty + KE - 1u) / KE;
    unsigned long long frags = 0, covs = 0;

    // thread 0 claims the next item of the workgroup's stream before the
    // output stage of the current one, so that the claim's latency overlaps
    // the output stage.  The counter's address is moved to vector registers:
    // a uniform address would turn the claim into a wave-aggregated atomic
    // whose result is awaited at once.
    const unsigned stream = blockIdx.x % kItemStreams;
    __attribute__((address_space(1))) unsigned* head =
        (__attribute__((address_space(1))) unsigned*)&ctr->itemHeads[stream][0];
    asm volatile("" : "+v"(head));
    unsigned claimed = 0;
    if (t == 0) claimed = __hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        if (t == 0) sItem = claimed;
        __syncthreads();
        const unsigned item = sItem * kItemStreams + stream;
        if (item >= numItems) {
            const unsigned e = item - numItems;
            if (e >= numEmptyItems) break;
            if (t == 0) claimed = __hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (unsigned k = e * KE; k < min(e * KE + KE, numEmpty); k++)
                clear_tile<S, NT>(view_of_tile(V, (int)emptyList[k]), P.clearDepth, P.scissor, visDepthAll,
                                  visSlotAll, visClassAll);
            __syncthreads();  // every lane has read sItem
            continue;
        }
        // long items from the front of the array, then short ones from its back
        const uint4 it = items[item < numLong ? item : itemCap - 1u - (item - numLong)];
        const int tile = __builtin_amdgcn_readfirstlane((int)it.x);
        const unsigned first = __builtin_amdgcn_readfirstlane(it.y);
        const int cnt = __builtin_amdgcn_readfirstlane((int)it.z);
        const unsigned pslot = __builtin_amdgcn_readfirstlane(it.w);
        const ViewGeom G = view_of_tile(V, tile);
        float* const visDepth = visDepthAll + G.pixBase * S;
        uint32_t* const visSlot = visSlotAll + G.pixBase * S;
        uint8_t* const visClass = visClassAll ? visClassAll + G.pixBase : nullptr;
        const SlotMap slotOf{slotOfAll + G.primBase, V.ctr[G.view].clipInCount == 0u ? G.slotBase : 0xffffffffu};
        const int ttx = G.tile % G.tilesX, tty = G.tile / G.tilesX;
        const int tx0 = ttx * 16, ty0 = tty * 16;

#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const int p = t + k * NT;
            const PixelCtx C = pixel_ctx(G, p & 15, p >> 4);
#pragma unroll
            for (int s = 0; s < S; s++)
                sVis[vis_at(s, p)] = frag_key(P.clear || !C.inFb ? P.clearDepth : visDepth[C.pix * S + s], 0u);
        }

        // two lanes per list entry: lane h = t & 1 takes block rows 2h, 2h + 1
        const int e0 = t >> 1, h = t & 1;
        // frame mask of this half tile: 4x2 groups with a pixel inside (S == 1;
        // pixels are checked again per unit), blocks with a pixel inside (S > 1)
        const int cols = min(16, G.width - tx0), rows = min(16, G.height - ty0);
        unsigned fbMask = 0, blkFb = 0;
        if constexpr (S == 1) {
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const int g = 16 * h + j;
                fbMask |= (unsigned)(((g >> 1) & 3) * 4 < cols && ((g >> 3) * 4 + (g & 1) * 2) < rows) << j;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) blkFb |= (unsigned)((j & 3) * 4 < cols && (2 * h + (j >> 2)) * 4 < rows) << j;
        }

        // merges units sUnit[0, nu) (entry << 8 | tile pixel or group) into sVis
        auto merge_units = [&](int nu) {
            for (int i = t; i < nu; i += NT) {
                const unsigned u = sUnit[i];
                const int e = (int)(u >> 8), c = (int)(u & 255u);
                const TriRec R = sRec[e];
                const Edges E = load_edges(R);
                const unsigned tie = __float_as_uint(R.q2.y);
                if constexpr (S == 1) {
                    const int bi = c >> 1;
                    const unsigned ta = sTa[2 * e + (bi >> 3)];
                    const int px0 = tx0 + (bi & 3) * 4, gy = ty0 + (bi >> 2) * 4 + (c & 1) * 2;
                    // pixels of the 4x2 group: (0,0) (1,0) (0,1) (1,1) (2,0) (3,0) (2,1) (3,1)
                    const bool accepted = (ta >> (bi & 7)) & 1u;
                    unsigned acc = 0xFFu;
                    if (!accepted) {
                        acc = 0;
#pragma unroll
                        for (int l = 0; l < 8; l++) {
                            const int ox = (l & 1) + ((l >> 2) << 1), oy = (l >> 1) & 1;
                            acc |= (unsigned)inside(E, (float)(px0 + ox), (float)(gy + oy)) << l;
                        }
                    }
                    // as with MSAA: every pixel's class depth is computed, only the merge is
                    // conditional
                    unsigned non = 0;
#pragma unroll
                    for (int l = 0; l < 8; l++) {
                        const int px = px0 + (l & 1) + ((l >> 2) << 1), py = gy + ((l >> 1) & 1);
                        const bool on = ((acc >> l) & 1u) & (px < G.width) & (py < G.height);
                        const unsigned quad = (acc >> (l & 4)) & 0xFu;
                        const unsigned cls = expr_code(accepted, acc == 0xFFu, quad == 0xFu, P.tagging);
                        const float d = class_depth_all(cls, R.f1, R.f2, R.f3, (float)px, (float)py, P.tagging);
                        non += on;
                        if (on & (d == d))
                            atomicMin(&sVis[vis_at(0, (py - ty0) * 16 + (px - tx0))], frag_key(d, (tie << 2) | cls));
                    }
                    frags += non;
                    covs += non;
                } else {
                    const int lx = c & 15, ly = c >> 4;
                    const int bi = (ly >> 2) * 4 + (lx >> 2);
                    const bool accept = (sTa[2 * e + (bi >> 3)] >> (bi & 7)) & 1u;
                    const float fx = (float)(tx0 + lx), fy = (float)(ty0 + ly);
                    // every sample's coverage and depth are computed (the lanes of a
                    // wave cover different pixels, so branches would run both ways);
                    // only the merge is conditional
                    unsigned ncov = 0;
#pragma unroll
                    for (int s = 0; s < S; s++) {
                        const float sx = fx + P.sx[s], sy = fy + P.sy[s];
                        const bool cov = accept | inside_and(E, sx, sy);
                        const float d = depth256(R.f1, R.f2, R.f3, sx, sy);
                        ncov += cov;
                        if (cov & (d == d)) atomicMin(&sVis[vis_at(s, c)], frag_key(d, tie));
                    }
                    covs += ncov;
                    frags += ncov != 0u;
                }
            }
        };

        // passes of PASS entries through LDS, all merging into sVis
        for (int pb = 0; pb < cnt; pb += PASS) {
            const int pc = min(PASS, cnt - pb);
            if constexpr (S == 1) {
                unsigned m = 0;
                if (e0 < pc) {
                    const unsigned ent = binList[first + pb + e0];
                    const TriRec R = load_rec(recs, ent & 0x7fffffffu);
                    if (h == 0) sRec[e0] = R;
                    unsigned ta;
                    group_mask_half(load_edges(R), (ent >> 31) != 0, ttx, tty, h, fbMask, m, ta);
                    sTa[t] = (unsigned char)ta;
                }
                const unsigned n = (unsigned)__popc(m);
                unsigned T;
                const unsigned off = scan_nt<NT>(n, sWave, T);

                for (unsigned w0 = 0; w0 < T; w0 += UNITS) {
                    // write this lane's units that fall in [w0, w0 + UNITS): entry << 8 | group
                    int pos = (int)off - (int)w0;
                    if (n && pos < UNITS && pos + (int)n > 0) {
                        unsigned bits = m;
                        const int cb = (e0 << 8) | (16 * h);
                        if (pos >= 0 && pos + (int)n <= UNITS) {
                            // the lane's units all fall in this window
                            while (bits) {
                                const int b = __ffs(bits) - 1;
                                bits &= bits - 1u;
                                sUnit[pos++] = (unsigned short)(cb + b);
                            }
                        } else {
                            while (bits) {
                                const int b = __ffs(bits) - 1;
                                bits &= bits - 1u;
                                if (pos >= 0 && pos < UNITS) sUnit[pos] = (unsigned short)(cb + b);
                                pos++;
                            }
                        }
                    }
                    __syncthreads();
                    merge_units((int)min((unsigned)UNITS, T - w0));
                    __syncthreads();
                }
            } else {
                // entry lanes find the blocks their half may cover; every
                // (entry, block) pair then takes one lane for its pixel mask
                unsigned tr = 0;
                if (e0 < pc) {
                    const unsigned ent = binList[first + pb + e0];
                    const TriRec R = load_rec(recs, ent & 0x7fffffffu);
                    if (h == 0) sRec[e0] = R;
                    unsigned ta;
                    half_blocks(load_edges(R), (ent >> 31) != 0, ttx, tty, h, tr, ta);
                    tr &= blkFb;
                    sTa[t] = (unsigned char)ta;
                }
                unsigned TP;
                int pp = (int)scan_nt<NT>((unsigned)__popc(tr), sWave, TP);
                while (tr) {
                    const int j = __ffs(tr) - 1;
                    tr &= tr - 1u;
                    sPair[pp++] = (unsigned short)((e0 << 4) | (8 * h + j));
                }
                __syncthreads();
                for (unsigned p0 = 0; p0 < TP; p0 += NT) {
                    unsigned pm = 0;
                    int pe = 0, bi = 0;
                    if (p0 + t < TP) {
                        const unsigned pr = sPair[p0 + t];
                        pe = (int)(pr >> 4);
                        bi = (int)(pr & 15u);
                        pm = block_fb_mask(cols - (bi & 3) * 4, rows - (bi >> 2) * 4);
                        if (!((sTa[2 * pe + (bi >> 3)] >> (bi & 7)) & 1u)) {
                            const Edges E = load_edges(sRec[pe]);
                            pm &= block_square_mask(E, (float)(tx0 + (bi & 3) * 4), (float)(ty0 + (bi >> 2) * 4));
                        }
                    }
                    const int n = __popc(pm);
                    unsigned T;
                    const unsigned off = scan_nt<NT>((unsigned)n, sWave, T);
                    for (unsigned w0 = 0; w0 < T; w0 += UNITS) {
                        // write this lane's units that fall in [w0, w0 + UNITS): tile pixel
                        // (4 (bi >> 2) + (b >> 2)) * 16 + 4 (bi & 3) + (b & 3) of block bit b
                        int pos = (int)off - (int)w0;
                        if (n && pos < UNITS && pos + n > 0) {
                            unsigned bits = pm;
                            const int cb = (pe << 8) | ((bi >> 2) * 64 + (bi & 3) * 4);
                            if (pos >= 0 && pos + n <= UNITS) {
                                // the lane's units all fall in this window
                                while (bits) {
                                    const int b = __ffs(bits) - 1;
                                    bits &= bits - 1u;
                                    sUnit[pos++] = (unsigned short)(cb + ((b >> 2) << 4) + (b & 3));
                                }
                            } else {
                                while (bits) {
                                    const int b = __ffs(bits) - 1;
                                    bits &= bits - 1u;
                                    if (pos >= 0 && pos < UNITS)
                                        sUnit[pos] = (unsigned short)(cb + ((b >> 2) << 4) + (b & 3));
                                    pos++;
                                }
                            }
                        }
                        __syncthreads();
                        merge_units((int)min((unsigned)UNITS, T - w0));
                        __syncthreads();
                    }
                    if (T == 0) __syncthreads();  // sWave is read by the next round's scan
                }
            }
            __syncthreads();  // sRec, sTa and sWave are reused by the next pass
        }

        if (t == 0) claimed = __hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (pslot != 0xFFFFFFFFu) {
            // one item of a longer list: its keys, sample-major like sVis
            unsigned long long* g = partials + (size_t)pslot * 256 * S;
#pragma unroll
            for (int k = 0; k < PPT; k++)
#pragma unroll
                for (int s = 0; s < S; s++) g[s * 256 + t + k * NT] = sVis[vis_at(s, t + k * NT)];
        } else {
            unsigned long long keys[PPT][S];
#pragma unroll
            for (int k = 0; k < PPT; k++)
#pragma unroll
                for (int s = 0; s < S; s++) keys[k][s] = sVis[vis_at(s, t + k * NT)];
            if (P.clear) {
                if constexpr (S > 1) {
                    // a pixel's samples are whole 16-byte words: straight out
#pragma unroll
                    for (int k = 0; k < PPT; k++) {
                        const int p = t + k * NT;
                        emit_pixel<S>(P, recs, slotOf, pixel_ctx(G, p & 15, p >> 4), keys[k], visDepth, visSlot);
                    }
                } else {
                    // every key is in registers: stage the outputs over sVis
                    TileOut<S>* O = reinterpret_cast<TileOut<S>*>(sVis);
                    __syncthreads();
#pragma unroll
                    for (int k = 0; k < PPT; k++) {
                        const int p = t + k * NT;
                        stage_pixel<S>(P, recs, slotOf, pixel_ctx(G, p & 15, p >> 4), p, keys[k], *O, visClass,
                                       visDepth, visSlot);
                    }
                    __syncthreads();
                    store_tile<S, NT>(G, ttx, tty, O, visDepth, visSlot);
                }
            } else {
#pragma unroll
                for (int k = 0; k < PPT; k++) {
                    const int p = t + k * NT;
                    write_pixel<S>(P, recs, slotOf, pixel_ctx(G, p & 15, p >> 4), keys[k], visDepth, visSlot,
                                   visClass);
                }
            }
        }
        __syncthreads();  // every lane has read sItem and the LDS of this item
    }
    const unsigned long long tf = sum_nt<NT>(frags);
    const unsigned long long tc = sum_nt<NT>(covs);
    if (t == 0) {
        st.rasterFrags[blockIdx.x] = tf;
        st.rasterCovs[blockIdx.x] = tc;
    }
}

// Minimum over the parts of every tile whose list was cut, then the outputs
// (the empty tiles' clear values come from k_raster).
template <int S>
__global__ void __launch_bounds__(256) k_raster_merge(RasterParams P, Views V, TriRecs recs,
                                                      const unsigned* __restrict__ slotOfAll,
                                                      const unsigned* __restrict__ binStart,
                                                      const unsigned* __restrict__ binCount,
                                                      const unsigned* __restrict__ multiList,
                                                      const int* __restrict__ tilePartial,
                                                      const unsigned long long* __restrict__ partials,
                                                      float* __restrict__ visDepthAll, uint32_t* __restrict__ visSlotAll,
                                                      uint8_t* __restrict__ visClassAll, int tiles, int partShift,
                                                      const DeviceCounters* ctr) {
    const unsigned numMulti = ctr->numMulti;
    const int t = threadIdx.x;
    // tiles whose list was cut into several items: the minimum over their partial blocks
    for (unsigned m = blockIdx.x; m < numMulti; m += gridDim.x) {
        const int tile = (int)multiList[m];
        const ViewGeom G = view_of_tile(V, tile);
        float* const visDepth = visDepthAll + G.pixBase * S;
        uint32_t* const visSlot = visSlotAll + G.pixBase * S;
        uint8_t* const visClass = visClassAll ? visClassAll + G.pixBase : nullptr;
        const SlotMap slotOf{slotOfAll + G.primBase, V.ctr[G.view].clipInCount == 0u ? G.slotBase : 0xffffffffu};
        const PixelCtx C = pixel_ctx(G, t & 15, t >> 4);
        const unsigned start = binStart[tile];
        const unsigned total = start >= P.binCap ? 0u : min(binCount[tile], P.binCap - start);
        const unsigned nparts = (total + (1u << partShift) - 1u) >> partShift;
        const unsigned long long* g = partials + (size_t)tilePartial[tile] * 256 * S + t;
        unsigned long long keys[S];
#pragma unroll
        for (int s = 0; s < S; s++) keys[s] = g[s * 256];
        for (unsigned p = 1; p < nparts; p++) {
            g += 256 * S;
#pragma unroll
            for (int s = 0; s < S; s++) keys[s] = min(keys[s], g[s * 256]);
        }
        if (P.clear) {
            if constexpr (S > 1) {
                emit_pixel<S>(P, recs, slotOf, C, keys, visDepth, visSlot);
            } else {
                __shared__ TileOut<S> O;  // output staging at 1 sample per pixel
                __syncthreads();          // the previous tile's staging has been stored
                stage_pixel<S>(P, recs, slotOf, C, t, keys, O, visClass, visDepth, visSlot);
                __syncthreads();
                store_tile<S>(G, C.tx, C.ty, &O, visDepth, visSlot);
            }
        } else {
            write_pixel<S>(P, recs, slotOf, C, keys, visDepth, visSlot, visClass);
        }
    }
}

// ---------------------------------------------------------------------------
// Ordered raster: depth functions other than LESS (and the depth test off,
// which is ALWAYS).  Their result depends on the order of the fragments, so
// each tile's list is taken in the order of the reference's per-tile sort
// (TileRasterWorker.cpp:1225-1259) and every thread folds the fragments of
// its own pixel in that order (TileRasterWorker.cpp:2272-2311 and the SIMD
// forms); the scalar path's GREATER tests nothing
// (TileRasterWorker.cpp:2278-2280).
// ---------------------------------------------------------------------------
// One workgroup per tile (persistent over the tiles of all views), one
// thread per pixel holding its S samples' depth and winning slot.
template <int S>
__global__ void __launch_bounds__(256) k_raster_ordered(RasterParams P, Views V, TriRecs recs,
                                                        const unsigned* __restrict__ binStart,
                                                        const unsigned* __restrict__ binCount,
                                                        const unsigned* __restrict__ binList, int tiles,
                                                        float* __restrict__ visDepthAll,
                                                        uint32_t* __restrict__ visSlotAll,
                                                        uint8_t* __restrict__ visClassAll, StatPartials st) {
    __shared__ unsigned long long sA[kOrderCap], sB[kOrderCap];
    const int t = threadIdx.x;
    unsigned long long frags = 0, covs = 0;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const ViewGeom G = view_of_tile(V, tile);
        const PixelCtx C = pixel_ctx(G, t & 15, t >> 4);
        float* const visDepth = visDepthAll + G.pixBase * S;
        uint32_t* const visSlot = visSlotAll + G.pixBase * S;
        uint8_t* const visClass = visClassAll ? visClassAll + G.pixBase : nullptr;
        float dep[S];
        unsigned slot[S];
        unsigned cls = 0;
#pragma unroll
        for (int s = 0; s < S; s++) {
            dep[s] = P.clear || !C.inFb ? P.clearDepth : visDepth[C.pix * S + s];
            slot[s] = P.clear || !C.inFb ? 0xffffffffu : visSlot[C.pix * S + s];
        }
        if (!P.clear && S == 1 && visClass && C.inFb) cls = visClass[C.pix];
        const unsigned start = binStart[tile];
        const unsigned cnt = start >= P.binCap ? 0u : min(binCount[tile], P.binCap - start);
        for_each_ordered(recs, binList, start, cnt, sA, sB, [&](unsigned ent) {
            const unsigned sl = ent & 0x7fffffffu;
            const bool full = (ent >> 31) != 0;
            if (!C.inFb) return;
            const TriRec R = load_rec(recs, sl);
            const Edges E = load_edges(R);
            int trn, tan;
            block_test(E, C.bMinX, C.bMinY, trn, tan);
            if constexpr (S == 1) {
                if (!covers(E, full, tan, trn, C.fpx, C.fpy)) return;
                const unsigned c = pixel_class(E, full || tan == 3, C.fpx, C.fpy, C.bMinX, P.tagging);
                const float d = class_depth(c, R.f1, R.f2, R.f3, C.fpx, C.fpy, P.tagging);
                frags++;
                covs++;
                // the scalar path's GREATER quirk belongs to the deterministic mode's single pixels
                if (!depth_pass(P.depthFunc, d, dep[0], c == 3u && !P.tagging)) return;
                dep[0] = P.depthWrite ? d * P.depthScale + P.depthBias : d;
                slot[0] = sl;
                cls = class_out(c, P.tagging);
            } else {
                bool any = false;
#pragma unroll
                for (int s = 0; s < S; s++) {
                    const float sx = C.fpx + P.sx[s], sy = C.fpy + P.sy[s];
                    if (!covers(E, full, tan, trn, sx, sy)) continue;
                    any = true;
                    covs++;
                    const float d = depth256(R.f1, R.f2, R.f3, sx, sy);
                    if (!depth_pass(P.depthFunc, d, dep[s], false)) continue;
                    dep[s] = P.depthWrite ? d * P.depthScale + P.depthBias : d;
                    slot[s] = sl;
                }
                frags += any;
            }
        });
        if (C.inFb && scissor_pass(P.scissor, C.fpx, C.fpy)) {
#pragma unroll
            for (int s = 0; s < S; s++) {
                visDepth[C.pix * S + s] = dep[s];
                visSlot[C.pix * S + s] = slot[s];
            }
            if (S == 1 && visClass) visClass[C.pix] = (uint8_t)cls;
        }
    }
    const unsigned long long tf = block_sum256(frags);
    const unsigned long long tc = block_sum256(covs);
    if (t == 0) {
        st.rasterFrags[blockIdx.x] = tf;
        st.rasterCovs[blockIdx.x] = tc;
    }
}

// Primitive ID per sample from the record slot (readback only).
__global__ void __launch_bounds__(256) k_slot_to_prim(const uint32_t* __restrict__ visSlot, const float4* __restrict__ q2,
                                                      size_t n, uint32_t* __restrict__ visId, int polygon) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const unsigned slot = visSlot[i];
        const unsigned key = slot == 0xffffffffu ? 0u : __float_as_uint(q2[slot].y);
        visId[i] = slot == 0xffffffffu ? 0xffffffffu : (polygon ? (key >> 2) * 3u + (key & 3u) : key >> 2);
    }
}

// Order-independent digest: the sum over i of SplitMix64(i << 32 ^ data[i])
// (the finaliser of Steele, Lea and Flood, 2014), reduced per workgroup and
// added into *out.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) k_digest(const uint32_t* __restrict__ data, size_t n,
                                                unsigned long long* out) {
    unsigned long long acc = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        acc += splitmix64(((unsigned long long)i << 32) ^ data[i]);
    acc = block_sum256(acc);
    if (threadIdx.x == 0) atomicAdd(out, acc);
}

}  // namespace

void launch_digest(const uint32_t* data, size_t n, unsigned long long* out, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_digest, dim3((unsigned)std::min<size_t>(2048, (n + 255) / 256)), dim3(256), 0, s, data, n,
                       out);
}

void launch_slot_to_prim(const uint32_t* visSlot, TriRecs recs, size_t n, uint32_t* visId, hipStream_t s,
                         bool polygon) {
    hipLaunchKernelGGL(k_slot_to_prim, dim3(2048), dim3(256), 0, s, visSlot, recs.q2, n, visId, polygon ? 1 : 0);
}

// The raster and merge launches are separate so that the host can time them.
void launch_raster(const RasterParams& p, const Views& views, int samples, TriRecs recs, const unsigned* slotOf,
                   const unsigned* binStart, const unsigned* binCount, const unsigned* binList, const RasterQueue& q,
                   int tiles, float* visDepth, uint32_t* visSlot, uint8_t* visClass, StatPartials st,
                   DeviceCounters* ctr, hipStream_t s) {
    (void)binStart;
    (void)binCount;
#define IFRIT_RASTER_LAUNCH(SN, NT)                                                                                 \
    hipLaunchKernelGGL((k_raster<SN, NT>), dim3(q.grid), dim3(NT), 0, s, p, views, recs, slotOf, binList, q.items, \
                       q.itemCap, q.multiList + tiles, q.partials, visDepth, visSlot, visClass, st, ctr)
    switch (samples) {
        case 1:
            if (q.tileThreads == 64)
                IFRIT_RASTER_LAUNCH(1, 64);
            else
                IFRIT_RASTER_LAUNCH(1, 256);
            break;
        case 4: IFRIT_RASTER_LAUNCH(4, 256); break;
        default: IFRIT_RASTER_LAUNCH(8, 256); break;
    }
#undef IFRIT_RASTER_LAUNCH
}

void launch_raster_ordered(const RasterParams& p, const Views& views, int samples, TriRecs recs,
                           const unsigned* binStart, const unsigned* binCount, const unsigned* binList, int tiles,
                           int grid, float* visDepth, uint32_t* visSlot, uint8_t* visClass, StatPartials st,
                           hipStream_t s) {
#define IFRIT_ORDERED_LAUNCH(SN)                                                                                  \
    hipLaunchKernelGGL(k_raster_ordered<SN>, dim3(grid), dim3(256), 0, s, p, views, recs, binStart, binCount,     \
                       binList, tiles, visDepth, visSlot, visClass, st)
    switch (samples) {
        case 1: IFRIT_ORDERED_LAUNCH(1); break;
        case 4: IFRIT_ORDERED_LAUNCH(4); break;
        default: IFRIT_ORDERED_LAUNCH(8); break;
    }
#undef IFRIT_ORDERED_LAUNCH
}

void launch_raster_merge(const RasterParams& p, const Views& views, int samples, TriRecs recs, const unsigned* slotOf,
                         const unsigned* binStart, const unsigned* binCount, const RasterQueue& q, float* visDepth,
                         uint32_t* visSlot, uint8_t* visClass, int tiles, DeviceCounters* ctr, hipStream_t s) {
#define IFRIT_MERGE_LAUNCH(SN)                                                                                     \
    hipLaunchKernelGGL(k_raster_merge<SN>, dim3(2 * q.mergeGrid), dim3(256), 0, s, p, views, recs, slotOf, binStart, \
                       binCount, q.multiList, q.tilePartial, q.partials, visDepth, visSlot, visClass, tiles,       \
                       q.partShift, ctr)
    switch (samples) {
        case 1: IFRIT_MERGE_LAUNCH(1); break;
        case 4: IFRIT_MERGE_LAUNCH(4); break;
        default: IFRIT_MERGE_LAUNCH(8); break;
    }
#undef IFRIT_MERGE_LAUNCH
}

}  // namespace ifrit_hip
