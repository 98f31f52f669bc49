// Device-side data layout and launch interface of the MI355X tile rasterizer.
//
// HBM layout per draw (V vertices, T input triangles, W x H pixels, S samples):
//   clip        float4[V]                 vertex stage output
//   recs        TriRecs, T + clipCap slots assembled triangles, slot = primitive for
//                                         unclipped triangles, T + k for clip fans
//   bary        float4[2 * (T+clipCap)]   vertex-weight bases bx, by of clip fans (shading only)
//   tileBox     int4[T + clipCap]         tile rectangle of each slot (x0 > x1: empty)
//   slotOf      uint[T]                   first slot of each primitive (its fans follow)
//   binCount/binStart/binCursor  uint[tiles]
//   binList     uint[binCap]              slot | 0x80000000 when the tile is fully covered
//   visId/visDepth/visSlot  [W*H*S]       per-sample winner, pixels row-major, samples innermost
//   rgba8       uchar4[W*H]               resolved colour
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "neural.h"

namespace ifrit_hip {

// One assembled triangle for coverage and depth (6 x 16 bytes):
//   q0 = e1.xyz, e2.x   q1 = e2.yz, e3.xy
//   q2 = e3.z, key, pixel box (x0 | y0 << 16), (x1 | y1 << 16 | big << 31)
//   f1, f2, f3 = interpolation bases (x, y, z) and vz = (z/w) * w.
// key = prim * 4 + fan orders fragments the way the reference's per-tile
// sort by originalPrimitive does (TileRasterWorker.cpp:1244-1258).  The
// pixel box is the padded screen extent, a hint that selects the fine-raster
// path (triangles flagged big are rasterised one pixel per lane).
struct __align__(16) TriRec {
    float4 q0, q1, q2;
    float4 f1, f2, f3;
};

// Records are stored as six float4 streams (structure of arrays) so that a
// wave's stores and the tile loader's gathers are 16-byte accesses.
struct TriRecs {
    float4 *q0, *q1, *q2, *f1, *f2, *f3;
};

struct SetupParams {
    int width, height;
    int tilesX, tilesY;
    int ntri, nverts;
    int ccw;
    int msaa;           // 1 disables the small-triangle cull
    unsigned clipCap;   // capacity for clip fan triangles and for triangles to clip
    float invW, invH;   // 1 / width, 1 / height (TileRasterRenderer.cpp:59-60)
    float invTileWidth; // ApproxReciprocal(16.0f) (TileRasterWorker.cpp:499)
};

struct BinParams {
    int tilesX, tilesY;
    int ntri;
    unsigned clipCap;
    unsigned binCap;
    unsigned bigCap;
    int bigThreshold;   // triangles spanning more tiles go to the per-workgroup binner
};

struct RasterParams {
    int width, height;
    int tilesX, tilesY;
    int depthFunc;
    int clear;
    unsigned binCap;
    float sx[8], sy[8];  // sample offsets from the pixel corner
};

struct TexDesc {
    unsigned long long base;  // byte offset of level 0 in the texture pool
    int width, height, levels, pad;
    unsigned levelOffset[16];  // texel offsets of each level relative to base
};

struct ShadeParams {
    int width, height, samples;
    int program;
    int stride, offWorld, offNormal, offUv, offTangent;
    int trisPerMaterial;
    int sampler, maxAniso;
    int numTextures;
    float camera[3];
    float ambient, kd, ks, shininess;
    int numLights;
    float lightPos[4][4];
    float lightColor[4][4];
    int writeFloat;          // also write RGBA32F
    int ntri;                // slots >= ntri are clip fans with their own bx, by
    NeuralShadeParams neural;  // programs 4 (neural material) and 5 (training features)
};

// Status word bits raised by the kernels; checked by the host after a frame.
enum StatusBits : unsigned {
    kStatusClipOverflow = 1u,
    kStatusBinOverflow = 2u,
    kStatusBigOverflow = 4u,
    kStatusBadIndex = 8u,
    kStatusRasterOverflow = 16u,
};

// Per-workgroup statistics (no same-address atomics in the hot kernels);
// k_finalize reduces them into DeviceCounters.
struct StatPartials {
    unsigned long long* setupAssembled;  // one per setup workgroup
    unsigned long long* rasterFrags;     // one per raster workgroup
    unsigned long long* rasterCovs;      // one per raster workgroup
    int setupBlocks, rasterBlocks;
};

// Counters written by the kernels (device memory, zeroed each draw).
struct DeviceCounters {
    unsigned status;
    unsigned clipCount;
    unsigned bigCount;
    unsigned binTotal;
    unsigned long long fragments;
    unsigned long long coveredSamples;
    unsigned long long assembled;
    unsigned clipInCount;
    unsigned itemHead;   // next raster work item
    unsigned numItems;
    unsigned numMulti;     // tiles cut into several parts
    unsigned numPartials;  // partial blocks those need
    unsigned numEmpty;     // tiles with an empty list
    unsigned numLongItems; // items of more than kRasterPass entries (front of the item array)
};

// Raster work: a work item covers at most kRasterPart list entries of one
// tile, taken in passes of kRasterPass entries (the LDS capacity).
constexpr int kRasterPass = 128;
constexpr int kRasterPart = 512;

// Raster work: (tile, part) items drained through DeviceCounters::itemHead;
// per tile the first partial block of a tile cut into several parts (-1
// otherwise); the list of such tiles followed, from index `tiles`, by the
// empty tiles; partial key blocks of 256 * samples keys each.
struct RasterQueue {
    uint2* items;
    int* tilePartial;
    unsigned* multiList;
    unsigned long long* partials;
    unsigned itemCap, partialCap;
    int grid;
};

// Launch wrappers (pipeline.hip).
void launch_vertex(const float* vtx, int stride, int nverts, const float* mvp16, int passthrough, float4* clip,
                   unsigned* binCount, unsigned* binCursor, int tiles, DeviceCounters* ctr, hipStream_t s);
void launch_reset(unsigned* binCount, unsigned* binCursor, int tiles, DeviceCounters* ctr, hipStream_t s);
void launch_setup(const SetupParams& p, const float4* clip, const int* idx, TriRecs recs, float4* bary, int4* tileBox,
                  unsigned* slotOf, unsigned* clipList, StatPartials st, DeviceCounters* ctr, hipStream_t s);
void launch_finalize(StatPartials st, DeviceCounters* ctr, hipStream_t s);
void launch_bin(const BinParams& p, TriRecs recs, const int4* tileBox, unsigned* binCount, unsigned* binStart,
                unsigned* binCursor, unsigned* binList, unsigned* bigList, const RasterQueue& q, DeviceCounters* ctr,
                hipStream_t s);
void launch_raster(const RasterParams& p, int samples, TriRecs recs, const unsigned* slotOf,
                   const unsigned* binStart, const unsigned* binCount, const unsigned* binList, const RasterQueue& q,
                   float* visDepth, uint32_t* visSlot, uint8_t* visClass, StatPartials st, DeviceCounters* ctr,
                   hipStream_t s);
void launch_raster_merge(const RasterParams& p, int samples, TriRecs recs, const unsigned* slotOf,
                         const unsigned* binStart, const unsigned* binCount, const RasterQueue& q, float* visDepth,
                         uint32_t* visSlot, uint8_t* visClass, DeviceCounters* ctr, hipStream_t s);
void launch_slot_to_prim(const uint32_t* visSlot, TriRecs recs, size_t n, uint32_t* visId, hipStream_t s);
void launch_shade(const ShadeParams& p, TriRecs recs, const float4* bary, const uint32_t* visSlot,
                  const float* vtx, const int* idx, const TexDesc* tex, const uint8_t* texPool, uchar4* rgba8,
                  float4* rgba32f, hipStream_t s);

// Texture helpers (texture.hip).
void launch_fbm_texture(uint8_t* level0, int size, unsigned seed, hipStream_t s);
void launch_mip_level(const uint8_t* src, int sw, int sh, uint8_t* dst, int dw, int dh, hipStream_t s);

}  // namespace ifrit_hip
