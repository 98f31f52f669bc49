// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the Ifrit-v2 soft renderer's deterministic tile
// rasterization path (projects/softgraphics/src/engine/tileraster).  It is the
// checker for the MI355X kernels: only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and never as the thing measured or
// shipped.  The product library does not link it.
//
// Arithmetic contract (pinned against the compiled reference in oracle/_ref):
//   * compiled with -ffp-contract=off; every fused multiply-add below is an
//     explicit std::fmaf at a place where the reference calls an FMA intrinsic;
//   * RCPPS/RCPSS come from the 2048-entry table captured from an Intel CPU
//     (ifrit-v2_amd/csrc/common/x86_rcp_table.inc) unless live mode is set,
//     in which case the host's own instruction is used;
//   * MINPS/MAXPS, CVTTSS2SI and DPPS follow their x86 definitions.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Coverage classes of the deterministic CPU path (TileRasterCommon.h:153-161),
// i.e. which depth expression produced a pixel's depth.
enum OracleCoverClass {
    ORC_NONE = 0,
    ORC_SIMD256 = 1,  // TILE, BLOCK or PIXEL_PACK4X2 proposals (TRW:2758-2769)
    ORC_SIMD128 = 2,  // PIXEL_PACK2X2 proposals (TRW:2423-2434)
    ORC_SCALAR = 3,   // PIXEL proposals (TRW:2261-2270)
    ORC_GROUPED = 4   // tagging mode: TILE and BLOCK proposals (pixelShadingSIMD256Grouped, TRW:2614-2725)
};

typedef struct {
    int width;
    int height;
    int samples;        // 1 (reference semantics), 4 or 8 (MSAA extension)
    int ccw;            // TileRasterFrontFace::COUNTER_CLOCKWISE swaps indices 0 and 2
    int depth_func;     // IfritCompareOp (Constants.h); LESS = 1
    int clear;          // 1: depth starts at 0x7f7f7f7f; 0: depth starts from *_in buffers
    int num_threads;    // tile-parallel worker threads (>= 1)
    int live_rcp;       // RCPSS/RCPPS: 0 the captured Intel table, 1 this CPU's own instruction, 2 the AMD table
    int cull_mode;      // IfritCullMode: 2 back (the CPU path); 1 front and 0 none set triangles up with
                        // vertices 0 and 2 exchanged (front: all; none: those back would cull); 3 all
    float clear_depth;  // depth a clear starts from: 0x7f7f7f7f in the CPU path
    int tagging;        // 1: the tagging (non-deterministic, optsetForceDeterministic(false)) mode's depth
                        // expressions at 1 sample per pixel: TILE and BLOCK proposals grouped
                        // (TRW:2672-2684), PIXEL proposals as a quad (pixelShadingSingleQuad, TRW:2123-2165);
                        // ignored with blending, which the reference always draws deterministically (TRW:410)
    int depth_write;    // 1: a fragment that passes the depth test stores depth * depth_scale + depth_bias, as
                        // a fragment shader writing its depth out-parameter does in the CPU path
                        // (TRW:2272-2384); ignored in the tagging mode, whose shading pass discards it
    float depth_scale, depth_bias;
} OracleRasterConfig;

typedef struct {
    int64_t input_triangles;
    int64_t assembled_triangles;   // after cull / clip
    int64_t clipped_inputs;        // inputs that went through homogeneous clipping
    int64_t bin_entries;           // (tile, triangle) pairs emitted by the binner
    int64_t fragments;             // (pixel, triangle) pairs with >= 1 covered sample
    int64_t covered_samples;       // (sample, triangle) coverage events, pre depth test
} OracleStats;

// One assembled triangle as the reference stores it (TileRasterContext.h:37-46).
typedef struct {
    float f1[4], f2[4], f3[4];
    float bx[3], by[3];
    float e1[3], e2[3], e3[3];
    float bbox[4];      // tile-space bbox used by the binner (minx, miny, maxx, maxy)
    int32_t prim;       // originalPrimitive = index offset / 3
    int32_t fan;        // fan index inside a clipped primitive (0 for unclipped)
    int32_t swapped;    // set up with vertices 0 and 2 exchanged (cull modes none and front)
} OracleTri;

// Sample positions (fractions of a pixel, relative to the integer pixel
// corner the CPU path samples at).  1x is the corner itself.
void oracle_sample_positions(int samples, float* xy_out);

// Vertex stage of the built-in transform program: clip = M * (p, 1) with the
// reference's MatMul summation order (include/ifrit/core/math/LinalgOps.h:40-48).
void oracle_vertex_transform(const float* vertices, int stride_floats, int nverts, const float* m_rowmajor,
                             float* clip_out);

typedef struct OracleFrame OracleFrame;

OracleFrame* oracle_frame_create(void);
void oracle_frame_destroy(OracleFrame* f);

// Runs setup, binning, coverage and the depth test.  Outputs per sample
// (row-major pixels, samples innermost): prim id (0xffffffff = empty),
// depth (FP32), assembled-triangle index (for shading) and, per pixel, the
// coverage class of the winning fragment (1 spp only; MSAA writes SIMD256).
// Point (mode 2) and line (mode 1) polygon modes at 1 sample per pixel
// (TileRasterCoreInvocationCuda.cu:2780-3335): per pixel the primitive id (the
// index position of a point, 3 * triangle + edge of a line) and depth of the
// nearest candidate below the stored depth; tri_out names the shading record
// (oracle_get_tris) of the winner.
int oracle_render_polygons(OracleFrame* f, const OracleRasterConfig* cfg, int mode, const float* clip, int nverts,
                           const int32_t* idx, int nidx, const float* depth_in, uint32_t* id_out, float* depth_out,
                           uint32_t* tri_out, OracleStats* stats);

int oracle_render(OracleFrame* f, const OracleRasterConfig* cfg, const float* clip_pos, int nverts,
                  const int32_t* indices, int nindices, const float* depth_in, uint32_t* id_out, float* depth_out,
                  uint32_t* tri_out, uint8_t* class_out, OracleStats* stats);

// Blending (TileRasterRenderer.cpp:212-333, TileRasterWorker.cpp:2346-2373) of
// the translucent flat colour (a hash of the primitive index) in primitive
// order, one sample per pixel; factors are IfritBlendFactor values.
typedef struct {
    int src_color, dst_color, src_alpha, dst_alpha;
} OracleBlend;
int oracle_render_blend(OracleFrame* f, const OracleRasterConfig* cfg, const float* clip_pos, int nverts,
                        const int32_t* indices, int nindices, const OracleBlend* blend, const float* depth_in,
                        const float* rgba_in, uint32_t* id_out, float* depth_out, float* rgba_out);

int oracle_num_tris(const OracleFrame* f);
// Assembled triangles sorted by (prim, fan); index = position in that order.
void oracle_get_tris(const OracleFrame* f, OracleTri* out);

#ifdef __cplusplus
}
#endif
