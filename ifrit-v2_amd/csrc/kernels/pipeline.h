// Device-side data layout and launch interface of the MI355X tile rasterizer.
//
// A draw renders a batch of views (one or more) of the same vertex and index
// data; every stage is one launch over all views.  HBM layout per draw
// (V vertices, T input triangles, N views, SV = T + clipCap record slots per
// view, S samples):
//   views       ViewDesc[N]               per-view camera, size and bases
//   tileView    ushort[tiles]             view of each 16x16 tile (tiles of all views)
//   shadeView   ushort[shadeTiles]        view of each shading tile
//   clip        float4[N][V]              vertex stage output
//   recs        TriRecs, N * SV slots     assembled triangles; in view v, slot v*SV + prim
//                                         for unclipped triangles, v*SV + T + k for clip fans
//   bary        float4[2 * N * SV]        vertex-weight bases bx, by of clip fans (shading only)
//   binQueue    uint4[N * SV]             per view, the assembled triangles to bin: (slot,
//                                         x0 | y0 << 16, x1 | y1 << 16, 0), view-local tiles
//   slotOf      uint[N][T]                first slot of each primitive (its fans follow)
//   binCount/binStart/binCursor  uint[tiles]
//   binList     uint[binCap]              slot | 0x80000000 when the tile is fully covered
//   visId/visDepth/visSlot  [pixels * S]  per-sample winner; views one after another, pixels
//                                         row-major within a view, samples innermost
//   rgba8       uchar4[pixels]            resolved colour
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

// One view of a draw.  View v owns tiles [tileBase, tileBase + tilesX *
// tilesY), record slots [slotBase, slotBase + SV), slotOf entries
// [primBase, primBase + T) and pixels [pixBase, pixBase + width * height).
struct ViewDesc {
    float mvp[16];
    float camera[3];
    int shade;                   // 1: the fragment program runs for this view
    int width, height, tilesX, tilesY;
    float invW, invH;            // 1 / width, 1 / height (TileRasterRenderer.cpp:59-60)
    unsigned tileBase;           // first 16x16 tile
    unsigned shadeBase;          // first shading tile (shaded views only)
    unsigned slotBase;           // first record slot (v * SV)
    unsigned primBase;           // first slotOf entry (v * T)
    unsigned long long pixBase;  // first pixel of the view's outputs
};

// Per-view counters (zeroed each draw), one 128-byte line each so
// that the views' atomics do not share a line.
struct ViewCounters {
    unsigned clipCount;    // clip fan slots taken
    unsigned clipInCount;  // triangles crossing w = 0
    unsigned queued;       // assembled triangles queued for binning
    unsigned pad[29];
};

struct Views {
    const ViewDesc* desc;
    const unsigned short* tileView;   // view of each 16x16 tile
    const unsigned short* shadeView;  // view of each shading tile
    ViewCounters* ctr;
    int count;
    unsigned slotsPerView;            // SV = T + clipCap
};

// Scissor rectangles (SetScissors / setScissorTestEnable of the CUDA renderer,
// TileRasterRendererCuda.h:101-102): a pixel (x, y) of a view passes when
// count is 0 or it lies in one rectangle, r.x <= x <= r.x + r.z and
// r.y <= y <= r.y + r.w (TileRasterCoreInvocationCuda.cu:1699-1711).  A
// draw leaves pixels that fail untouched: visibility, depth and colour
// (its clear included, as the CUDA clear kernels are scissored,
// TileRasterCoreInvocationCuda.cu:3336-3424).
constexpr int kMaxScissors = 8;
struct Scissor {
    int count;            // 0: every pixel passes
    const float4* rects;  // device memory: (x, y, width, height) per rectangle
};

// IfritCullMode (include/ifrit/softgraphics/engine/base/Constants.h:29-35).
enum CullMode { kCullNone = 0, kCullFront = 1, kCullBack = 2, kCullFrontAndBack = 3 };

// Per-view fields (width, height, tiles, invW, invH) are filled in by the
// kernels from the view's ViewDesc.
struct SetupParams {
    int width, height;
    int tilesX, tilesY;
    int ntri, nverts;
    int cullMode;       // IfritCullMode: 0 none, 1 front, 2 back (the CPU path), 3 front and back
    unsigned char* swapped;  // per record slot, cull modes none and front: 1 when the triangle was set
                             // up with its first and third vertex exchanged (null for back)
    int msaa;           // 1 disables the small-triangle cull
    unsigned clipCap;   // capacity for clip fan triangles and for triangles to clip
    float invW, invH;   // 1 / width, 1 / height (TileRasterRenderer.cpp:59-60)
    float invTileWidth; // ApproxReciprocal(16.0f) (TileRasterWorker.cpp:499)
    int rcpModel;       // ifx::RcpModel: whose RCPPS the reciprocals of w and of the area reproduce
    // per view, per group of kGroupTris triangles: 1 when the group's bounds lie
    // outside a frustum plane with a margin (null: no group culling)
    const unsigned char* groupOut;
    int groups;
};

// Triangles per culling group (a multiple of the 64-lane wave).
constexpr int kGroupTris = 256;

struct BinParams {
    int tiles;          // tiles of all views
    int grid;           // workgroups of the binning launches
    int ntri;
    unsigned clipCap;
    unsigned binCap;
    unsigned bigCap;
    int bigThreshold;   // triangles spanning more tiles go to the per-workgroup binner
};

struct RasterParams {
    int depthFunc;
    int clear;
    float clearDepth;    // depth a cleared sample starts from (0x7f7f7f7f in the CPU path)
    int tagging;         // 1 spp: the tagging mode's depth expressions (class_depth)
    int depthWrite;      // the fragment program stores depth * depthScale + depthBias after the test
    float depthScale, depthBias;
    Scissor scissor;
    unsigned binCap;
    float sx[8], sy[8];  // sample offsets from the pixel corner
};

struct TexDesc {
    unsigned long long base;  // byte offset of level 0 in the texture pool
    int width, height, levels, layers;
    // levels follow level 0 packed, each max(1, width >> l) x max(1, height >> l) texels per
    // layer, the layers of a level one after another (a cube map has six, +x -x +y -y +z -z);
    // the sampler computes the offsets from width, height and layers
};

// Sampler state of a program (IfritSamplerT, include/ifrit/softgraphics/engine/base/Structures.h:52-61)
// beyond the level-of-detail choice: addressing per axis, the filter within a level and the border.
enum AddressMode { kAddrRepeat = 0, kAddrMirroredRepeat = 1, kAddrClampToEdge = 2, kAddrClampToBorder = 3,
                   kAddrMirrorClampToEdge = 4 };
struct SamplerMode {
    int addrU, addrV;  // AddressMode
    int nearest;       // 1: nearest texel within a level (IF_FILTER_NEAREST), 0: bilinear
    unsigned border;   // RGBA8 border colour: 0xff000000 opaque black, 0xffffffff opaque white
};

struct ShadeParams {
    int samples;
    int program;
    int stride, offWorld, offNormal, offUv, offTangent;
    int trisPerMaterial;
    int sampler, maxAniso;
    int numTextures;
    float ambient, kd, ks, shininess;
    int numLights;
    float lightPos[4][4];
    float lightColor[4][4];
    int writeFloat;          // also write RGBA32F
    int ntri;                // view-local slots >= ntri are clip fans with their own bx, by
    int shadeTile;           // pixels per side of a shading tile (8, 16 or 32): one wave shades
                             // its 8x8 sub-tiles in turn
    NeuralShadeParams neural;  // programs 4 (neural material) and 5 (training features)
    const unsigned char* swapped;  // per record slot: vertices 0 and 2 exchanged by setup (null: none)
    float clearColor[4];     // colour of samples no triangle covers (the resolve weighs them)
    float clearDepth;        // depth a cleared pixel starts from (blending draws)
    int depthWrite;          // blending draws: the program stores depth * depthScale + depthBias
    float depthScale, depthBias;
    Scissor scissor;
    SamplerMode samp;        // addressing and filter of the program's lookups
    int sampGeneric;         // 1: samp differs from repeat + bilinear (the sampled-kernel variant runs)
    float lod;               // cube-map program: the level of detail of its lookups
};

// Blend factors, each encoded (c0, c1, c2, c3) as the reference does
// (TileRasterRenderer.cpp:212-333): colour source, colour destination, alpha
// source, alpha destination.
struct BlendCoefs {
    float sc[4], dc[4], sa[4], da[4];
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

// Counters written by the kernels (device memory, zeroed each draw by the vertex or reset kernel).
// Raster work items are drained through this many counters, one per XCD:
// stream k holds items k, k + kItemStreams, ... and is taken by the
// workgroups with blockIdx.x % kItemStreams == k (one XCD's), so the claims
// of the persistent raster spread over separate lines instead of one.
constexpr int kItemStreams = 8;

struct DeviceCounters {
    unsigned status;
    unsigned clipCount;  // clip fan slots of all views (reduced with the statistics)
    unsigned bigCount;
    unsigned binTotal;
    unsigned long long fragments;
    unsigned long long coveredSamples;
    unsigned long long assembled;
    unsigned clipMax;    // largest per-view clipCount or clipInCount (sizes clipCap on overflow)
    unsigned numItems;
    unsigned numMulti;     // tiles cut into several parts
    unsigned numPartials;  // partial blocks those need
    unsigned numEmpty;     // tiles with an empty list
    unsigned numLongItems; // items of more than kRasterPass entries (front of the item array)
    alignas(128) unsigned itemHeads[kItemStreams][32];  // next item of each stream, one 128-byte line each
};

// Raster work: a work item covers at most 1 << partShift list entries of
// one tile (256 to 1024), taken in passes of kRasterPass entries (the LDS
// capacity).
constexpr int kRasterPass = 128;
// Tiles per workgroup of the tile-plan kernels (one per thread): their block
// sums take (tiles + kPlanTiles - 1) / kPlanTiles + 1 groups of six words.
constexpr int kPlanTiles = 256;
constexpr int kRasterPartShift = 9;

// Raster work: items {tile, first list entry, entries, partial block or
// ~0u} drained through DeviceCounters::itemHeads; per tile the first partial
// block of a tile cut into several items (-1 otherwise); the list of such
// tiles followed, from index `tiles`, by the empty tiles; partial key blocks
// of 256 * samples keys each; planning scratch (6 words per 2048 tiles).
struct RasterQueue {
    uint4* items;
    int* tilePartial;
    unsigned* multiList;
    unsigned long long* partials;
    unsigned* planSums;
    unsigned itemCap, partialCap;
    int grid;         // raster workgroups
    int mergeGrid;    // half the merge workgroups
    int partShift;
    int tileThreads;  // threads per tile at 1 sample per pixel: 64 or 256
};

// Launch wrappers.  clipAt: -1 transforms attribute 0 (xyz) by each view's
// matrix; otherwise the record holds the clip position at float clipAt.
void launch_vertex(const float* vtx, int stride, int nverts, const Views& views, int clipAt, float4* clip,
                   unsigned* binCount, unsigned* binCursor, int tiles, DeviceCounters* ctr,
                   const float4* groupBounds, int groups, unsigned char* groupOut, hipStream_t s);
void launch_group_bounds(const float* vtx, int stride, int nverts, const int* idx, int ntri, float4* bounds,
                         hipStream_t s);
void launch_reset(unsigned* binCount, unsigned* binCursor, int tiles, const Views& views, DeviceCounters* ctr,
                  hipStream_t s);
void launch_setup(const SetupParams& p, const Views& views, const float4* clip, const int* idx, TriRecs recs, float4* bary, uint4* binQueue,
                  unsigned* slotOf, unsigned* clipList, StatPartials st, DeviceCounters* ctr, hipStream_t s);
void launch_finalize(StatPartials st, const Views& views, DeviceCounters* ctr, hipStream_t s);
void launch_bin(const BinParams& p, const Views& views, TriRecs recs, const uint4* binQueue, unsigned* binCount, unsigned* binStart,
                unsigned* binCursor, unsigned* binList, unsigned* bigList, const RasterQueue& q, DeviceCounters* ctr,
                hipStream_t s);
void launch_raster(const RasterParams& p, const Views& views, int samples, TriRecs recs, const unsigned* slotOf,
                   const unsigned* binStart, const unsigned* binCount, const unsigned* binList, const RasterQueue& q,
                   int tiles, float* visDepth, uint32_t* visSlot, uint8_t* visClass, StatPartials st,
                   DeviceCounters* ctr, hipStream_t s);
void launch_raster_merge(const RasterParams& p, const Views& views, int samples, TriRecs recs, const unsigned* slotOf,
                         const unsigned* binStart, const unsigned* binCount, const RasterQueue& q, float* visDepth,
                         uint32_t* visSlot, uint8_t* visClass, int tiles, DeviceCounters* ctr, hipStream_t s);
// Depth functions other than LESS: each tile's list in primitive order, folded
// per pixel (replaces launch_raster and launch_raster_merge).
void launch_raster_ordered(const RasterParams& p, const Views& views, int samples, TriRecs recs,
                           const unsigned* binStart, const unsigned* binCount, const unsigned* binList, int tiles,
                           int grid, float* visDepth, uint32_t* visSlot, uint8_t* visClass, StatPartials st,
                           hipStream_t s);
// IDs from record slots: the primitive key >> 2, or for point and line records (polygon) the key's
// triangle * 3 + its point or edge.
void launch_slot_to_prim(const uint32_t* visSlot, TriRecs recs, size_t n, uint32_t* visId, hipStream_t s,
                         bool polygon = false);
void launch_shade(const ShadeParams& p, const Views& views, int shadeTiles, TriRecs recs, const float4* bary, const uint32_t* visSlot,
                  const float* vtx, const int* idx, const TexDesc* tex, const uint8_t* texPool, uchar4* rgba8,
                  float4* rgba32f, hipStream_t s);

// Blending, 1 sample per pixel: the ordered traversal with the depth test, each
// passing fragment shaded and blended in primitive order (replaces the raster,
// merge and shade launches).
void launch_blend_ordered(const ShadeParams& p, const BlendCoefs& b, int depthFunc, int clear, unsigned binCap,
                          const Views& views, TriRecs recs, const float4* bary, const float* vtx, const int* idx,
                          const TexDesc* tex, const uint8_t* texPool, const unsigned* binStart,
                          const unsigned* binCount, const unsigned* binList, int tiles, int grid, float* visDepth,
                          uint32_t* visSlot, uint8_t* visClass, uchar4* rgba8, float4* rgba32f, StatPartials st,
                          hipStream_t s);

// Point (mode 2) and line (mode 1) polygon modes, one sample per pixel (polygon.hip):
// candidates per pixel meet in keys (order_key(depth) << 32 | primitive id),
// then the resolve writes depth and record slot where the nearest one is below
// the stored depth.  Record slots: the index position of a point, 3 * triangle
// + edge of a line.
struct PolyParams {
    int mode;
    int width, height;
    int nidx, nverts;
    int clear;
    float clearDepth;
    unsigned long long* keys;  // width * height
};
void launch_polygons(const PolyParams& p, const float4* clip, const int* idx, TriRecs recs, float4* bary,
                     float* visDepth, uint32_t* visSlot, uint8_t* visClass, StatPartials st, DeviceCounters* ctr,
                     int grid, hipStream_t s);

// The built-in meshlet program of DrawMeshTasks (mesh.hip): work groups first .. first + groups - 1,
// meshlet m owning vertices vo[m] .. vo[m + 1] - 1 and indices io[m] .. io[m + 1] - 1 (local to the
// meshlet); emitted vertices are (clip xyzw, record) at vo[m] - vertexBase, indices at
// io[m] - indexBase.
struct MeshParams {
    float mvp[16];
    int recordFloats;
    int first, groups;
    int vertexBase, indexBase;
};
void launch_mesh_expand(const MeshParams& p, const float* records, const int* indices, const int* vo, const int* io,
                        float* outVtx, int* outIdx, DeviceCounters* ctr, hipStream_t s);
// The built-in geometry program: triangle t's vertices, transformed (or passed through) and moved
// by offset, emitted as records (input record, then clip xyzw) 3t .. 3t + 2, indices 3t .. 3t + 2.
struct GeometryParams {
    float mvp[16];
    float offset[4];
    int passthrough;
    int stride;        // floats per input vertex record
    int nverts, ntri;
};
void launch_geometry_expand(const GeometryParams& p, const float* vtx, const int* idx, float* outVtx, int* outIdx,
                            DeviceCounters* ctr, hipStream_t s);

// Texture helpers (texture.hip).
// Order-independent 64-bit digest of n words (equal buffers give equal digests).
void launch_digest(const uint32_t* data, size_t n, unsigned long long* out, hipStream_t s);
void launch_fbm_texture(uint8_t* level0, int size, unsigned seed, hipStream_t s);
void launch_mip_level(const uint8_t* src, int sw, int sh, uint8_t* dst, int dw, int dh, hipStream_t s);
// blitImage (TileRasterImageOpInvocationsCuda.cu:36-131) between RGBA8 levels: destination texel
// (dx0 + x, dy0 + y) for x < dx1 - dx0, y < dy1 - dy0 reads the source at
// (x / (dx1 - dx0) (sx1 - sx0) + sx0, likewise y); linear: bilinear with the right and lower
// neighbours clamped to the source level, on the 0..255 values, rounded half up; nearest:
// the texel at the rounded coordinate, clamped.
struct BlitArgs {
    int sw, sh, dw, dh;               // level sizes
    unsigned sx0, sy0, sx1, sy1;      // source region
    unsigned dx0, dy0, dx1, dy1;      // destination region
    int linear;
};
void launch_blit(const uint8_t* src, uint8_t* dst, const BlitArgs& a, hipStream_t s);

}  // namespace ifrit_hip
