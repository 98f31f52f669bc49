// ORACLE — TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
// FP64 shading oracle for the built-in fragment programs.  The reference CPU
// path runs user C++ shaders and has no texture sampling; the CUDA path's
// sampler (include/ifrit/softgraphics/engine/math/ShaderBuiltinCuda.cuh:
// 36-295) is the behavioural spec.  This oracle evaluates the programs in
// double precision from the FP32 setup records of oracle_render, so that the
// FP32 kernels can be compared with it within one RGBA8 step.
//
// Discontinuous decisions (anisotropic tap count, major-axis choice, the
// N.L > 0 branch) are taken in double precision; pixels where the deciding
// quantity lies within a relative 1e-4 of a switch point are flagged, because
// an FP32 evaluation may legitimately land on the other side.
#pragma once
#include <stdint.h>

#include "oracle_core.h"

#ifdef __cplusplus
extern "C" {
#endif

enum OracleProgram {
    ORP_FLAT = 0,         // colour = bytes of hash(prim), alpha 255
    ORP_BLINN_PHONG = 1,  // textured Blinn-Phong, directional and point lights
    ORP_DEPTH_ONLY = 2,   // no colour output
    ORP_GGX_LAYERED = 3,  // layered GGX: texture 0 albedo, 1 tangent-space normal, 2 roughness
    ORP_NORMAL_LAMBERT = 7, // untextured: d = max(0, n.l / |n|), rgb = ambient + light rgb d + d^shininess
    ORP_CUBE_MAP = 8        // the six-layer texture along the interpolated normal at level `lod`
};
enum OracleSampler { ORS_BILINEAR = 0, ORS_TRILINEAR = 1, ORS_ANISO = 2 };

typedef struct {
    int program;
    int stride_floats;      // vertex record size
    int off_world;          // float offset of the world position (xyz)
    int off_normal;         // float offset of the normal (xyz)
    int off_uv;             // float offset of the texture coordinate (uv)
    int tris_per_material;  // texture index = prim / tris_per_material
    int sampler;            // OracleSampler
    int max_aniso;
    float camera[3];
    float ambient, kd, ks, shininess;
    int num_lights;         // <= 4
    float light_pos[4][4];  // w = 0: direction towards the light; w = 1: point light position
    float light_color[4][4];  // rgb, attenuation coefficient k in 1 / (1 + k d^2)
    int off_tangent;        // float offset of the tangent (xyz), ORP_GGX_LAYERED
    float clear_color[4];   // colour of samples no triangle covers (0 in the CPU path)
    int addr_u, addr_v;     // IfritSamplerAddressMode of the lookups (0 repeat ... 4 mirror clamp to edge)
    int nearest;            // 1: nearest texel within a level
    float border[4];        // clamp-to-border colour, 0..1
    float lod;              // ORP_CUBE_MAP: level of detail
} OracleShadeParams;

typedef struct {
    const uint8_t* texels;  // RGBA8, levels stored consecutively, level l is (w >> l) x (h >> l) per layer,
                            // its layers one after another
    int width, height, levels;
    int layers;             // 1, or 6 for a cube map (+x -x +y -y +z -z)
} OracleTexture;

// 32-bit integer hash used for flat colours (shared definition with the product).
uint32_t oracle_flat_hash(uint32_t prim);

// Shades every pixel from the per-sample winners tri_ids (indices into the
// frame's assembled triangles, 0xffffffff = empty; each distinct triangle is
// shaded once per pixel), clamps each sample's colour to [0, 1], averages the
// samples (empty samples count as the clear colour) and rounds to RGBA8.  Returns the number of flagged pixels.
int oracle_shade(const OracleFrame* f, const OracleShadeParams* p, int W, int H, int S, const uint32_t* tri_ids,
                 const int32_t* indices, const float* vertices, const OracleTexture* textures, int ntex,
                 uint8_t* rgba8_out, uint8_t* flag_out, int num_threads);

// Test hook: plants a known mistake in the shading oracle (0 restores it);
// see g_plant in oracle_shade.cpp.  Used only by the closed-form tests.
void oracle_set_plant(int plant);

// Box-filter mip generation, (a + b + c + d + 2) >> 2 per channel.
void oracle_build_mips(uint8_t* texels, int width, int height, int levels);

// One texture lookup of the sampler (0 bilinear on level 0, 1 trilinear, 2
// anisotropic) at (u, v) with the screen derivatives of (u, v), on an RGBA8
// texture with its packed mip chain; out = RGBA in 0..1.  Returns 1 when the
// tap count was too close to call (a whole-number anisotropy ratio).
int oracle_sample_tex(const uint8_t* texels, int width, int height, int levels, int sampler, int maxAniso, double u,
                      double v, double dudx, double dvdx, double dudy, double dvdy, double* out);
// The same lookup under a sampler's addressing (addr_u, addr_v), filter within a level (nearest) and
// border colour (0..1).
int oracle_sample_tex_mode(const uint8_t* texels, int width, int height, int levels, int sampler, int maxAniso,
                           int addr_u, int addr_v, int nearest, const float* border, double u, double v, double dudx,
                           double dvdx, double dudy, double dvdy, double* out);
// Cube-map lookup (six-layer texture) along direction d at a level of detail; returns 1 when the face
// choice was too close to call (two direction components within a relative 1e-6).
int oracle_cube_lookup(const uint8_t* texels, int size, int levels, int addr_u, int addr_v, int nearest,
                       const float* border, const double* d, double lod, double* out);

#ifdef __cplusplus
}
#endif
