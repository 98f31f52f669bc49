// Deferred shading and MSAA resolve on gfx950: one thread per pixel; each
// distinct triangle among the pixel's samples is shaded once at the pixel's
// shading position (the integer pixel position at 1 sample, the pixel centre
// with MSAA), weighted by its sample count, clamped and box-resolved to RGBA8.
// Texture sampling follows the CUDA path's sampler spec
// (include/ifrit/softgraphics/engine/math/ShaderBuiltinCuda.cuh:36-295) with
// texel centres at (i + 0.5) / size, footprints |d(uv)/dx| and |d(uv)/dy|,
// and symmetric anisotropic taps; derivatives are fine quad differences of
// the same triangle evaluated at the quad's other pixels.
#include <hip/hip_runtime.h>

#include "neural_device.h"
#include "raster_common.h"

namespace ifrit_hip {

namespace {

// ---------------------------------------------------------------------------
// Deferred shading and resolve
// ---------------------------------------------------------------------------
// Hardware transcendental and reciprocal instructions (about 1 ulp): the
// colour is compared at 1/255, so the IEEE-exact library sequences buy
// nothing here.  Visibility never goes through this file.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float flog2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }

struct ShadeTri {
    float4 f1, f2, f3;
    float bx[3], by[3];
    int prim;
    int swapped;  // set up with its first and third vertex exchanged (cull modes none and front)
};

// The ntri slots below fanStart (the view's first clip fan slot) are the
// unclipped triangles, triangle p in slot fanStart - ntri + p: their
// primitive comes from the slot itself, so the vertex indices can be read in
// the same round trip as the record (a fan slot reads its key first).  The
// record's planes are read last: a load still in flight when the two kinds
// of slot rejoin would otherwise be waited for there.
__device__ __forceinline__ void load_shade_tri(const TriRecs& recs, const float4* bary, unsigned fanStart,
                                               unsigned ntri, unsigned slot, ShadeTri& T,
                                               const unsigned char* swapped) {
    if (slot < fanStart) {
        T.prim = (int)(slot - (fanStart - ntri));
        // unclipped triangles keep bx = (0, 0, 1), by = (1, 0, 0) (TileRasterWorker.cpp:711-712)
        T.bx[0] = 0.0f; T.bx[1] = 0.0f; T.bx[2] = 1.0f;
        T.by[0] = 1.0f; T.by[1] = 0.0f; T.by[2] = 0.0f;
    } else {
        T.prim = (int)(__float_as_uint(recs.q2[slot].y) >> 2);
        const float4 a = bary[2 * slot], b = bary[2 * slot + 1];
        T.bx[0] = a.x; T.bx[1] = a.y; T.bx[2] = a.z;
        T.by[0] = a.w; T.by[1] = b.x; T.by[2] = b.y;
    }
    T.swapped = swapped ? swapped[slot] : 0;
    T.f1 = recs.f1[slot];
    T.f2 = recs.f2[slot];
    T.f3 = recs.f3[slot];
}

// Perspective-correct weights of the three original vertices at (x, y).
__device__ __forceinline__ void vertex_weights(const ShadeTri& T, float x, float y, float w[3]) {
    const float b1 = fmaf(T.f1.x, x, fmaf(T.f1.y, y, T.f1.z));
    const float b2 = fmaf(T.f2.x, x, fmaf(T.f2.y, y, T.f2.z));
    const float b3 = fmaf(T.f3.x, x, fmaf(T.f3.y, y, T.f3.z));
    const float inv = frcp(b1 + b2 + b3);
    const float B1 = b1 * inv, B2 = b2 * inv, B3 = b3 * inv;
    w[0] = fmaf(B1, T.bx[0], fmaf(B2, T.bx[1], B3 * T.bx[2]));
    w[1] = fmaf(B1, T.by[0], fmaf(B2, T.by[1], B3 * T.by[2]));
    w[2] = 1.0f - w[0] - w[1];
}

// The two mip levels a trilinear lookup at lod blends: level l0 = floor of
// the clamped lod and l1 = l0 + 1 (l0 at the last level or where the lod is
// a whole number, i.e. the blend fraction is 0), as the address
// of level l0's first texel, the texel distance to level l1, the levels'
// log2 sizes (textures are powers of two) and the blend fraction.
struct TriLevels {
    const uchar4* t0;
    unsigned d1;    // level l1 starts d1 texels after level l0
    unsigned dims;  // log2 width, log2 height of l0, then of l1, 5 bits each
    float f;
};

// What the sampler needs of a texture: level 0's byte offset in the pool,
// log2 of its width and height, and the number of levels.
struct TexHead {
    unsigned long long base;
    int lw, lh, levels;
};

__device__ __forceinline__ TexHead tex_head(const TexDesc& D) {
    return TexHead{D.base, __builtin_ctz((unsigned)D.width), __builtin_ctz((unsigned)D.height), D.levels};
}

// Texel offset of level l in the packed chain of a 2^a x 2^b texture whose
// level k holds max(1, 2^(a-k)) x max(1, 2^(b-k)) texels (the host's layout
// in createTexture).  Levels k <= min(a, b) hold 2^(a+b-2k)
// texels, a geometric sum; then the long side halves alone down to 1, and
// any further levels hold one texel.
__device__ __forceinline__ unsigned level_offset(int a, int b, int l) {
    const int m = min(a, b), M = max(a, b);
    const unsigned long long S4 = 1ull << (a + b + 2);
    unsigned long long off = (S4 - (S4 >> (2 * min(l, m + 1)))) / 3u;
    if (l > m + 1) {
        const int lc = min(l, M + 1);
        off += (1ull << (M - m)) - (1ull << (M - lc + 1)) + (unsigned long long)(l - lc);
    }
    return (unsigned)off;
}

// The level addresses come from the header and level_offset, so preparing
// them needs no memory access beyond the header itself.
__device__ __forceinline__ TriLevels tri_levels(const uint8_t* pool, const TexHead& H, float lod) {
    lod = fminf(fmaxf(lod, 0.0f), (float)(H.levels - 1));
    const int l0 = (int)floorf(lod);
    const int l1 = lod > (float)l0 ? min(l0 + 1, H.levels - 1) : l0;
    const int w0 = max(H.lw - l0, 0), h0 = max(H.lh - l0, 0);
    TriLevels L;
    L.t0 = reinterpret_cast<const uchar4*>(pool + H.base) + level_offset(H.lw, H.lh, l0);
    L.d1 = l1 > l0 ? 1u << (w0 + h0) : 0u;  // level l0's texel count
    L.dims = (unsigned)w0 | ((unsigned)h0 << 5) | ((unsigned)max(H.lw - l1, 0) << 10) |
             ((unsigned)max(H.lh - l1, 0) << 15);
    L.f = lod - (float)l0;
    return L;
}

// Bilinear lookup in one level of 2^lw x 2^lh texels with repeat
// addressing, texel centres at (i + 0.5) / size, split into the fetch of the
// four texels and their blend, so that a caller can put the loads of several
// lookups in flight before it blends any of them.
struct BilinearTexels {
    unsigned t00, t10, t01, t11;  // RGBA8 texels
    float ax, ay;
};

__device__ __forceinline__ BilinearTexels bilinear_fetch(const uchar4* __restrict__ T, int lw, int lh, float u,
                                                         float v) {
    const int w = 1 << lw, h = 1 << lh;
    const float s = fmaf(u, (float)w, -0.5f), r = fmaf(v, (float)h, -0.5f);
    const float fs = floorf(s), fr = floorf(r);
    const int x0 = (int)fs & (w - 1), y0 = (int)fr & (h - 1);
    const int x1 = (x0 + 1) & (w - 1), y1 = (y0 + 1) & (h - 1);
    const unsigned* const U = reinterpret_cast<const unsigned*>(T);
    BilinearTexels B;
    B.t00 = U[(y0 << lw) + x0];
    B.t10 = U[(y0 << lw) + x1];
    B.t01 = U[(y1 << lw) + x0];
    B.t11 = U[(y1 << lw) + x1];
    B.ax = s - fs;
    B.ay = r - fr;
    return B;
}

// Texel index i of a level n texels wide under an addressing mode (n a power
// of two), or -1 where clamp-to-border gives the border colour.  Mirrored
// repeat reflects every other period; mirror-clamp-to-edge reflects once
// about 0 and then clamps (the Vulkan definitions; the CUDA spec's
// mirrorclamp, ShaderOpsCuda.cuh:36-37, swaps the ends instead of mirroring).
__device__ __forceinline__ int texel_index(int i, int n, int mode) {
    switch (mode) {
        case kAddrMirroredRepeat: {
            const int m = i & (2 * n - 1);
            return m >= n ? 2 * n - 1 - m : m;
        }
        case kAddrClampToEdge: return min(max(i, 0), n - 1);
        case kAddrClampToBorder: return (i < 0 || i >= n) ? -1 : i;
        case kAddrMirrorClampToEdge: return min(i < 0 ? -1 - i : i, n - 1);
        default: return i & (n - 1);
    }
}

// bilinear_fetch under a sampler's addressing and filter: nearest takes the
// texel holding (u, v), floor(u w); border texels are the border colour.
__device__ __forceinline__ BilinearTexels bilinear_fetch_mode(const uchar4* __restrict__ T, int lw, int lh, float u,
                                                              float v, const SamplerMode& m) {
    const int w = 1 << lw, h = 1 << lh;
    const unsigned* const U = reinterpret_cast<const unsigned*>(T);
    auto at = [&](int x, int y) {
        const int ix = texel_index(x, w, m.addrU), iy = texel_index(y, h, m.addrV);
        return (ix < 0 || iy < 0) ? m.border : U[(iy << lw) + ix];
    };
    BilinearTexels B;
    if (m.nearest) {
        B.t00 = at((int)floorf(u * (float)w), (int)floorf(v * (float)h));
        B.t10 = B.t01 = B.t11 = B.t00;
        B.ax = B.ay = 0.0f;
        return B;
    }
    const float s = fmaf(u, (float)w, -0.5f), r = fmaf(v, (float)h, -0.5f);
    const float fs = floorf(s), fr = floorf(r);
    const int x0 = (int)fs, y0 = (int)fr;
    B.t00 = at(x0, y0);
    B.t10 = at(x0 + 1, y0);
    B.t01 = at(x0, y0 + 1);
    B.t11 = at(x0 + 1, y0 + 1);
    B.ax = s - fs;
    B.ay = r - fr;
    return B;
}

// Trilinear lookup with prepared levels.  Both levels are always fetched
// before either is blended, so all eight texel loads are in flight together
// (one memory round trip per tap).
struct TrilinearTexels {
    BilinearTexels b0, b1;
    float f;
};

__device__ __forceinline__ TrilinearTexels trilinear_fetch(const TriLevels& L, float u, float v) {
    TrilinearTexels T;
    T.b0 = bilinear_fetch(L.t0, (int)(L.dims & 31u), (int)((L.dims >> 5) & 31u), u, v);
    T.b1 = bilinear_fetch(L.t0 + L.d1, (int)((L.dims >> 10) & 31u), (int)((L.dims >> 15) & 31u), u, v);
    T.f = L.f;
    return T;
}

__device__ __forceinline__ TrilinearTexels trilinear_fetch_mode(const TriLevels& L, float u, float v,
                                                                const SamplerMode& m) {
    TrilinearTexels T;
    T.b0 = bilinear_fetch_mode(L.t0, (int)(L.dims & 31u), (int)((L.dims >> 5) & 31u), u, v, m);
    T.b1 = bilinear_fetch_mode(L.t0 + L.d1, (int)((L.dims >> 10) & 31u), (int)((L.dims >> 15) & 31u), u, v, m);
    T.f = L.f;
    return T;
}

// The blend as one weighted sum of the eight texels: a texel's weight is its
// bilinear weight times its level's weight (1 - f or f), with the scale of
// the stored bytes (1/255) folded into the level weights.  With f = 0 the
// second level's weights are 0.
__device__ __forceinline__ void bilinear_weights(const BilinearTexels& B, float g, float w[4]) {
    const float gy1 = g * B.ay, gy0 = g - gy1;
    w[1] = gy0 * B.ax;
    w[0] = gy0 - w[1];
    w[3] = gy1 * B.ax;
    w[2] = gy1 - w[3];
}

__device__ __forceinline__ float4 trilinear_blend(const TrilinearTexels& T) {
    constexpr float k = 1.0f / 255.0f;
    const float g1 = T.f * k;
    float w[8];
    bilinear_weights(T.b0, k - g1, w);
    bilinear_weights(T.b1, g1, w + 4);
    const unsigned t[8] = {T.b0.t00, T.b0.t10, T.b0.t01, T.b0.t11, T.b1.t00, T.b1.t10, T.b1.t01, T.b1.t11};
    float c[4];
#pragma unroll
    for (int ch = 0; ch < 4; ch++) {
        float r = w[0] * (float)((t[0] >> (8 * ch)) & 255u);
#pragma unroll
        for (int i = 1; i < 8; i++) r = fmaf(w[i], (float)((t[i] >> (8 * ch)) & 255u), r);
        c[ch] = r;
    }
    return make_float4(c[0], c[1], c[2], c[3]);
}

__device__ __forceinline__ float4 trilinear_levels(const TriLevels& L, float u, float v) {
    return trilinear_blend(trilinear_fetch(L, u, v));
}

__device__ float4 trilinear(const uint8_t* pool, const TexDesc& D, float u, float v, float lod) {
    return trilinear_levels(tri_levels(pool, tex_head(D), lod), u, v);
}

// Cube-map lookup (textureCubeLod, ShaderBuiltinCuda.cuh:184-231): the face
// and its (u, v) from the direction's major axis (0 +x, 1 -x, 2 +y, 3 -y,
// 4 +z, 5 -z, the Vulkan face table), then a trilinear lookup in that face at
// the given level of detail under the sampler's addressing and filter.  The
// texture holds six layers per level.
__device__ float4 cube_lookup(const uint8_t* pool, const TexDesc& D, const SamplerMode& m, float3 d, float lod) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int face;
    float ma, sc, tc;
    if (az >= ax && az >= ay) {
        face = d.z < 0.0f ? 5 : 4;
        ma = az;
        sc = d.z < 0.0f ? -d.x : d.x;
        tc = -d.y;
    } else if (ay >= ax) {
        face = d.y < 0.0f ? 3 : 2;
        ma = ay;
        sc = d.x;
        tc = d.y < 0.0f ? -d.z : d.z;
    } else {
        face = d.x < 0.0f ? 1 : 0;
        ma = ax;
        sc = d.x < 0.0f ? d.z : -d.z;
        tc = -d.y;
    }
    const float norm = 0.5f / ma;
    const float u = sc * norm + 0.5f, v = tc * norm + 0.5f;
    const TexHead H = tex_head(D);
    lod = fminf(fmaxf(lod, 0.0f), (float)(H.levels - 1));
    const int l0 = (int)floorf(lod);
    const int l1 = lod > (float)l0 ? min(l0 + 1, H.levels - 1) : l0;
    const int w0 = max(H.lw - l0, 0), h0 = max(H.lh - l0, 0), w1 = max(H.lw - l1, 0), h1 = max(H.lh - l1, 0);
    const uchar4* base = reinterpret_cast<const uchar4*>(pool + H.base);
    const uchar4* t0 = base + (size_t)D.layers * level_offset(H.lw, H.lh, l0) + ((size_t)face << (w0 + h0));
    const uchar4* t1 = base + (size_t)D.layers * level_offset(H.lw, H.lh, l1) + ((size_t)face << (w1 + h1));
    TrilinearTexels T;
    T.b0 = bilinear_fetch_mode(t0, w0, h0, u, v, m);
    T.b1 = bilinear_fetch_mode(t1, w1, h1, u, v, m);
    T.f = lod - (float)l0;
    return trilinear_blend(T);
}

// Footprint of a textured fragment: number of taps along the major axis,
// the level of detail and the tap step (SBC:244-272; trilinear and bilinear
// are the one-tap cases).
struct TexFootprint {
    float au, av, lod;
    int n;
};

__device__ __forceinline__ TexFootprint footprint(const TexHead& D, int sampler, int maxAniso, float dudx, float dvdx,
                                                  float dudy, float dvdy) {
    TexFootprint F{0.0f, 0.0f, 0.0f, 1};
    if (sampler == 0) return F;  // bilinear on level 0
    const float W = (float)(1 << D.lw), H = (float)(1 << D.lh);
    const float px = fsqrt((dudx * W) * (dudx * W) + (dvdx * H) * (dvdx * H));
    const float py = fsqrt((dudy * W) * (dudy * W) + (dvdy * H) * (dvdy * H));
    const float pmax = fmaxf(px, py), pmin = fminf(px, py);
    if (sampler == 1 || !(pmax > 0.0f)) {
        F.lod = pmax > 0.0f ? flog2(pmax) : 0.0f;
        return F;
    }
    const float ratio = pmin > 0.0f ? pmax * frcp(pmin) : (float)maxAniso;
    int n = ratio >= (float)maxAniso ? maxAniso : (int)ceilf(ratio);
    n = min(max(1, n), 16);  // the host clamps maxAniso to 16; the tap list holds 4-bit tap indices
    F.n = n;
    F.lod = flog2(pmax * frcp((float)n));
    F.au = px >= py ? dudx : dudy;
    F.av = px >= py ? dvdx : dvdy;
    return F;
}

__device__ __forceinline__ float sat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__device__ __forceinline__ uint32_t flat_hash(uint32_t prim) {
    uint32_t h = prim * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    return h;
}

// Dot products and lengths with fused multiply-adds (the shading is held
// to within one step of an FP64 computation, not to the CPU path's bits).
__device__ __forceinline__ float dot3(float3 a, float3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

__device__ __forceinline__ float3 norm3(float3 a) {
    const float l2 = dot3(a, a);
    const float inv = l2 > 0.0f ? frsq(l2) : 1.0f;
    return make_float3(a.x * inv, a.y * inv, a.z * inv);
}

// Everything of a fragment's colour except its albedo: colour = albedo.rgb *
// A + B, alpha = albedo.a (Blinn-Phong: A = ambient + sum att * light * kd *
// n.l, B = sum att * light * ks * (n.h)^shininess).  Also the texture
// coordinates and footprint when the fragment samples a texture.
struct FragPrep {
    float3 A, B;
    float alpha;  // translucent flat colour: alpha from the hash
    float u, v;
    TexFootprint fp;
    int mat;     // texture index, -1 when untextured
    TexHead th;  // its header
    // layered GGX: shading frame, view direction and position
    float3 N, Tg, Bg, V, Pw;
};


// Layered GGX of one light (Trowbridge-Reitz with alpha = roughness,
// Smith-Schlick k = alpha / 2, Schlick Fresnel F0 = 0.04, Lambert diffuse
// weighted by 1 - F), times n.l.
__device__ __forceinline__ float3 ggx_light(float3 n, float3 V, float3 L, float4 alb, float alpha, float kd, float ks) {
    const float ndl = dot3(n, L), ndv = dot3(n, V);
    if (!(ndl > 0.0f && ndv > 0.0f)) return make_float3(0, 0, 0);
    const float3 H = norm3(make_float3(L.x + V.x, L.y + V.y, L.z + V.z));
    const float ndh = fmaxf(dot3(n, H), 0.0f), vdh = fmaxf(dot3(V, H), 0.0f);
    const float a2 = alpha * alpha;
    const float den = ndh * ndh * (a2 - 1.0f) + 1.0f;
    const float D = a2 * frcp(3.14159265f * den * den);
    const float k = alpha * 0.5f;
    const float G = ndv * frcp(ndv * (1.0f - k) + k) * (ndl * frcp(ndl * (1.0f - k) + k));
    const float f = 1.0f - vdh, f2 = f * f;
    const float F = 0.04f + 0.96f * (f2 * f2 * f);
    const float spec = ks * (D * G * F * frcp(4.0f * ndl * ndv));
    const float diff = kd * (1.0f - F) * (1.0f / 3.14159265f);
    return make_float3((alb.x * diff + spec) * ndl, (alb.y * diff + spec) * ndl, (alb.z * diff + spec) * ndl);
}

// frame = false for callers that never run the programs from 3 on (layered GGX
// and the neural material): their shading-frame paths then fold away.
__device__ FragPrep prep_fragment(const ShadeParams& P, float3 cam, const ShadeTri& T, const float* vtx,
                                  const int* idx, const TexDesc* tex, const uint8_t* pool, int px, int py, float o,
                                  bool frame, bool generic = false) {
    const bool framed = frame && P.program >= 3;
    FragPrep R;
    R.mat = -1;
    R.th = TexHead{0ull, 0, 0, 1};
    R.u = R.v = 0.0f;
    R.fp = TexFootprint{0.0f, 0.0f, 0.0f, 0};
    R.alpha = 1.0f;
    if (P.program == 0 || P.program == 6) {
        // flat colour: bytes 0-2 of a hash of the primitive; the translucent flat colour also
        // takes alpha = (64 + (byte 3 & 127)) / 255
        const uint32_t h = flat_hash((uint32_t)T.prim);
        R.A = make_float3(0, 0, 0);
        R.B = make_float3((h & 0xff) * (1.0f / 255.0f), ((h >> 8) & 0xff) * (1.0f / 255.0f),
                          ((h >> 16) & 0xff) * (1.0f / 255.0f));
        if (P.program == 6) R.alpha = (float)(64u + ((h >> 24) & 0x7fu)) * (1.0f / 255.0f);
        return R;
    }
    const float x = (float)px + o, y = (float)py + o;
    // the planes weigh the vertices in the order setup took them: exchanged triangles
    // take their first and third vertex the other way round (chosen after the loads, so
    // that the index loads need not wait for the flag)
    const int i0 = idx[3 * T.prim], i1 = idx[3 * T.prim + 1], i2 = idx[3 * T.prim + 2];
    const float* v0 = vtx + (size_t)(T.swapped ? i2 : i0) * P.stride;
    const float* v1 = vtx + (size_t)i1 * P.stride;
    const float* v2 = vtx + (size_t)(T.swapped ? i0 : i2) * P.stride;
    if (P.numTextures > 0) {
        // the texture header depends on the primitive only: it is read in the same round
        // trip as the vertex indices (issued after them, so that they are in flight together)
        const int mat = P.trisPerMaterial > 0 ? T.prim / P.trisPerMaterial : 0;
        R.mat = framed ? 0 : min(mat, P.numTextures - 1);  // the layers share one footprint
        R.th = tex_head(tex[R.mat]);
    }
    // texture coordinates are read with the other vertex attributes (always: an untextured
    // vertex reads its first two floats)
    const int ou = P.numTextures > 0 ? P.offUv : 0;
    const float t0u = v0[ou], t0v = v0[ou + 1], t1u = v1[ou], t1v = v1[ou + 1], t2u = v2[ou], t2v = v2[ou + 1];
    float w[3];
    vertex_weights(T, x, y, w);
    auto interp = [&](int off) { return fmaf(w[0], v0[off], fmaf(w[1], v1[off], w[2] * v2[off])); };
    const float3 Pw = make_float3(interp(P.offWorld), interp(P.offWorld + 1), interp(P.offWorld + 2));
    float3 N = make_float3(interp(P.offNormal), interp(P.offNormal + 1), interp(P.offNormal + 2));
    if (framed) {
        // shading frame: interpolated normal, tangent made orthogonal to it, bitangent N x T
        N = norm3(N);
        float3 Tg = make_float3(interp(P.offTangent), interp(P.offTangent + 1), interp(P.offTangent + 2));
        const float nt = dot3(N, Tg);
        Tg = norm3(make_float3(Tg.x - N.x * nt, Tg.y - N.y * nt, Tg.z - N.z * nt));
        R.N = N;
        R.Tg = Tg;
        R.Bg = make_float3(N.y * Tg.z - N.z * Tg.y, N.z * Tg.x - N.x * Tg.z, N.x * Tg.y - N.y * Tg.x);
        R.Pw = Pw;
        R.V = norm3(make_float3(cam.x - Pw.x, cam.y - Pw.y, cam.z - Pw.z));
        R.A = R.B = make_float3(0, 0, 0);
    }
    if (P.numTextures > 0) {
        R.u = fmaf(w[0], t0u, fmaf(w[1], t1u, w[2] * t2u));
        R.v = fmaf(w[0], t0v, fmaf(w[1], t1v, w[2] * t2v));
        // fine quad derivatives (SBC:36-76): u(qx + 1, y) - u(qx, y) and u(x, qy + 1) - u(x, qy),
        // the same triangle at the quad's other shading positions.  The pixel is qx or qx + 1
        // (exactly, in float) and qy or qy + 1, so its own (u, v) is one end of each difference
        // and only the other column and the other row are evaluated.
        const float qx = (float)(px & ~1) + o, qy = (float)(py & ~1) + o;
        const bool oddx = (px & 1) != 0, oddy = (py & 1) != 0;
        float wx[3], wy[3];
        vertex_weights(T, oddx ? qx : qx + 1.0f, y, wx);
        vertex_weights(T, x, oddy ? qy : qy + 1.0f, wy);
        auto iu = [&](const float* ww) { return fmaf(ww[0], t0u, fmaf(ww[1], t1u, ww[2] * t2u)); };
        auto iv = [&](const float* ww) { return fmaf(ww[0], t0v, fmaf(ww[1], t1v, ww[2] * t2v)); };
        const float ux = iu(wx), vx = iv(wx), uy = iu(wy), vy = iv(wy);
        const float dudx = oddx ? R.u - ux : ux - R.u, dvdx = oddx ? R.v - vx : vx - R.v;
        const float dudy = oddy ? R.u - uy : uy - R.u, dvdy = oddy ? R.v - vy : vy - R.v;
        R.fp = footprint(R.th, P.sampler, P.maxAniso, dudx, dvdx, dudy, dvdy);
    }
    if (frame && P.program == 4 && P.neural.singleTap && R.mat >= 0) {
        // the neural material's latent: one trilinear tap at the isotropic footprint log2(pmax)
        R.fp.lod += flog2((float)R.fp.n);
        R.fp.n = 1;
        R.fp.au = R.fp.av = 0.0f;
    }
    if (framed) return R;  // lighting needs the sampled layers (or the network)
    if (generic && P.program == 8 && R.mat >= 0) {
        // cube-map lookup along the interpolated normal at the program's level of detail
        const float4 c = cube_lookup(pool, tex[R.mat], P.samp, N, P.lod);
        R.A = make_float3(0, 0, 0);
        R.B = make_float3(c.x, c.y, c.z);
        R.alpha = sat(c.w);
        R.mat = -1;  // no anisotropic taps
        return R;
    }
    if (P.program == 7) {
        // Lambert plus a (n.l)^shininess highlight from the interpolated normal and one directional
        // light given unnormalised, untextured: the reference CPU harness's shader
        // (d = max(0, n.l / |n|), rgb = ambient + light.rgb * d + d^shininess)
        const float inv = frsq(N.x * N.x + N.y * N.y + N.z * N.z + 1e-20f);
        float d = (N.x * P.lightPos[0][0] + N.y * P.lightPos[0][1] + N.z * P.lightPos[0][2]) * inv;
        d = d > 0.0f ? d : 0.0f;
        const float sp = d > 0.0f ? fexp2(P.shininess * flog2(d)) : 0.0f;
        R.A = make_float3(0, 0, 0);
        R.B = make_float3(P.ambient + P.lightColor[0][0] * d + sp, P.ambient + P.lightColor[0][1] * d + sp,
                          P.ambient + P.lightColor[0][2] * d + sp);
        return R;
    }
    N = norm3(N);
    const float3 V = norm3(make_float3(cam.x - Pw.x, cam.y - Pw.y, cam.z - Pw.z));
    float3 A = make_float3(P.ambient, P.ambient, P.ambient), B = make_float3(0, 0, 0);
    for (int l = 0; l < P.numLights; l++) {
        // n.l and the unnormalised half vector l + v, with l = Lv / |Lv|
        float att = 1.0f, ndl;
        float3 Hr;
        if (P.lightPos[l][3] == 0.0f) {
            const float3 L = norm3(make_float3(P.lightPos[l][0], P.lightPos[l][1], P.lightPos[l][2]));
            ndl = dot3(N, L);
            Hr = make_float3(L.x + V.x, L.y + V.y, L.z + V.z);
        } else {
            const float3 Lv = make_float3(P.lightPos[l][0] - Pw.x, P.lightPos[l][1] - Pw.y, P.lightPos[l][2] - Pw.z);
            const float d2 = dot3(Lv, Lv);
            const float inv = frsq(d2);
            ndl = dot3(N, Lv) * inv;
            Hr = make_float3(fmaf(Lv.x, inv, V.x), fmaf(Lv.y, inv, V.y), fmaf(Lv.z, inv, V.z));
            att = frcp(fmaf(P.lightColor[l][3], d2, 1.0f));
        }
        if (ndl > 0.0f) {
            const float h2 = dot3(Hr, Hr);
            const float ndh = fmaxf(dot3(N, Hr) * (h2 > 0.0f ? frsq(h2) : 1.0f), 0.0f);
            // ndh^shininess (powf's value at ndh = 0: 1 for shininess 0, else 0)
            const float pw = ndh > 0.0f ? fexp2(P.shininess * flog2(ndh)) : (P.shininess == 0.0f ? 1.0f : 0.0f);
            const float spec = P.ks * pw;
            const float d = att * P.kd * ndl;
            A.x = fmaf(d, P.lightColor[l][0], A.x);
            A.y = fmaf(d, P.lightColor[l][1], A.y);
            A.z = fmaf(d, P.lightColor[l][2], A.z);
            B.x = fmaf(att * spec, P.lightColor[l][0], B.x);
            B.y = fmaf(att * spec, P.lightColor[l][1], B.y);
            B.z = fmaf(att * spec, P.lightColor[l][2], B.z);
        }
    }
    R.A = A;
    R.B = B;
    return R;
}


// Clamped colour of a layered-GGX fragment from its filtered layers: albedo,
// tangent-space normal (xyz, encoded as 0.5 n + 0.5) and roughness (w).
__device__ float4 ggx_color(const ShadeParams& P, const FragPrep& R, float4 alb, float4 nm) {
    const float3 nts = norm3(make_float3(2.0f * nm.x - 1.0f, 2.0f * nm.y - 1.0f, 2.0f * nm.z - 1.0f));
    const float3 n = norm3(make_float3(R.Tg.x * nts.x + R.Bg.x * nts.y + R.N.x * nts.z,
                                       R.Tg.y * nts.x + R.Bg.y * nts.y + R.N.y * nts.z,
                                       R.Tg.z * nts.x + R.Bg.z * nts.y + R.N.z * nts.z));
    const float alpha = 0.05f + 0.95f * nm.w;
    float3 col = make_float3(alb.x * P.ambient, alb.y * P.ambient, alb.z * P.ambient);
    for (int l = 0; l < P.numLights; l++) {
        float3 L;
        float att = 1.0f;
        if (P.lightPos[l][3] == 0.0f) {
            L = norm3(make_float3(P.lightPos[l][0], P.lightPos[l][1], P.lightPos[l][2]));
        } else {
            const float3 Lv = make_float3(P.lightPos[l][0] - R.Pw.x, P.lightPos[l][1] - R.Pw.y,
                                          P.lightPos[l][2] - R.Pw.z);
            const float d2 = dot3(Lv, Lv);
            const float inv = frsq(d2);
            L = make_float3(Lv.x * inv, Lv.y * inv, Lv.z * inv);
            att = frcp(1.0f + P.lightColor[l][3] * d2);
        }
        const float3 c = ggx_light(n, R.V, L, alb, alpha, P.kd, P.ks);
        col.x += att * P.lightColor[l][0] * c.x;
        col.y += att * P.lightColor[l][1] * c.y;
        col.z += att * P.lightColor[l][2] * c.z;
    }
    return make_float4(sat(col.x), sat(col.y), sat(col.z), sat(alb.w));
}

// View and light (light 0) directions in the fragment's tangent frame: the
// neural material's direction inputs.
__device__ __forceinline__ void frame_dirs(const ShadeParams& P, const FragPrep& R, float3& vt, float3& lt) {
    vt = make_float3(dot3(R.V, R.Tg), dot3(R.V, R.Bg), dot3(R.V, R.N));
    float3 L = make_float3(0, 0, 1);
    if (P.numLights > 0) {
        if (P.lightPos[0][3] == 0.0f)
            L = norm3(make_float3(P.lightPos[0][0], P.lightPos[0][1], P.lightPos[0][2]));
        else
            L = norm3(make_float3(P.lightPos[0][0] - R.Pw.x, P.lightPos[0][1] - R.Pw.y, P.lightPos[0][2] - R.Pw.z));
    }
    lt = make_float3(dot3(L, R.Tg), dot3(L, R.Bg), dot3(L, R.N));
}

constexpr int kTapWindow = 128;

// One wave per 8x8 pixel sub-tile.  Work is expanded twice so that lanes stay
// busy whatever the mix of edge pixels and sampling footprints:
//   1. each pixel lists its distinct triangles (first-occurrence order, with
//      their sample counts) into one flat fragment list;
//   2. each lane prepares one fragment (interpolation, derivatives, lighting)
//      and its texture taps go into one flat tap list, one tap per lane;
//   3. the fragment lane sums its taps in tap order, the pixel lane sums its
//      fragments in sample order, so results do not depend on scheduling.
// (ox, oy) is the sub-tile's first pixel in its view; visSlot, rgba8 and
// rgba32f point at the view's outputs and pixBase is the view's first pixel.
template <int S, bool NEURAL, bool FP8, bool LAYERED, bool GENERIC = false>
__device__ __forceinline__ void shade_subtile(const ShadeParams& P, float3 cam, int width, int height,
                                              unsigned fanStart, int ox, int oy, size_t pixBase, const TriRecs& recs,
                                              const float4* __restrict__ bary, const uint32_t* __restrict__ visSlot,
                                              const float* __restrict__ vtx, const int* __restrict__ idx,
                                              const TexDesc* __restrict__ tex, const uint8_t* __restrict__ pool,
                                              uchar4* __restrict__ rgba8, float4* __restrict__ rgba32f) {
    __shared__ unsigned fSlot[64 * S];
    __shared__ unsigned char fCnt[64 * S];
    __shared__ unsigned char fPix[64 * S];
    __shared__ float4 cUV[64];      // u, v, au, av of the chunk's fragments
    __shared__ float cLod[64];
    __shared__ int cN[64];
    // The layered and neural programs keep a second value per tap; a single
    // texture needs only the fragment's prepared levels there, and its fragment
    // colours reuse the tap colours' space (the taps are summed first).  Less
    // LDS per wave lets more waves share a CU.
    constexpr bool kTwo = NEURAL || LAYERED;
    __shared__ int cMat[LAYERED ? 64 : 1];
    __shared__ float4 cColOwn[kTwo ? 64 : 1];
    __shared__ float4 xdir[64];     // network inputs 10..13 (view z, light xyz)
    __shared__ unsigned short tList[kTapWindow];
    // per tap: colour or latent channels 0-3 (tCol); layered GGX normal xyz
    // and roughness, or latent channels 4-7 (tCol2), or per fragment its two
    // mip levels (one texture); network inputs after the taps
    __shared__ float4 tBuf[kTapWindow + (kTwo ? kTapWindow : 64)];
    float4* const tCol = tBuf;
    float4* const tCol2 = tBuf + kTapWindow;
    float4* const cCol = kTwo ? cColOwn : tBuf;  // clamped colour of the chunk's fragments

    const int lane = threadIdx.x;
    const int px = ox + (lane & 7);
    const int py = oy + (lane >> 3);
    const bool inFb = px < width && py < height;
    const size_t pix = (size_t)py * width + px;

    // 1. distinct triangles of the pixel
    unsigned ids[S];
#pragma unroll
    for (int s = 0; s < S; s++) ids[s] = 0xffffffffu;
    if (inFb) {
        if constexpr (S % 4 == 0) {
            // a pixel's samples are contiguous and 16-byte aligned
#pragma unroll
            for (int v = 0; v < S / 4; v++) {
                const uint4 x = reinterpret_cast<const uint4*>(visSlot + pix * S)[v];
                ids[4 * v] = x.x;
                ids[4 * v + 1] = x.y;
                ids[4 * v + 2] = x.z;
                ids[4 * v + 3] = x.w;
            }
        } else {
#pragma unroll
            for (int s = 0; s < S; s++) ids[s] = visSlot[pix * S + s];
        }
    }
    // one comparison per pair of samples: sample s is a first occurrence when no
    // earlier sample holds its triangle, and each equal pair adds one to the
    // earlier sample's count (so a first occurrence counts all its samples)
    unsigned firstMask = 0;
    int cnt[S];
#pragma unroll
    for (int s = 0; s < S; s++) cnt[s] = 1;
#pragma unroll
    for (int s = 0; s < S; s++) {
        bool first = ids[s] != 0xffffffffu;
#pragma unroll
        for (int q = 0; q < s; q++) {
            const bool e = ids[q] == ids[s];
            first = first && !e;
            cnt[q] += e ? 1 : 0;
        }
        firstMask |= first ? 1u << s : 0u;
    }
    const int nf = __builtin_popcount(firstMask);
    const int incl = dev::wave_incl_scan(nf);
    const int poff = incl - nf;
    const int F = __shfl(incl, 63, 64);
    {
        int k = poff;
#pragma unroll
        for (int s = 0; s < S; s++) {
            if (firstMask & (1u << s)) {
                fSlot[k] = ids[s];
                fCnt[k] = (unsigned char)cnt[s];
                fPix[k] = (unsigned char)lane;
                k++;
            }
        }
    }
    __syncthreads();

    if (LAYERED && P.program == 5 && S == 1 && inFb && nf == 0) P.neural.featOut[4 * (pixBase + pix) + 1] = make_float4(0, 0, 0, 0);
    constexpr bool layered = LAYERED;  // programs 3 and 5: three textures per tap
    const float o = S == 1 ? 0.0f : 0.5f;
    // every sample starts at the clear colour; a fragment replaces its samples' share
    // (the clear colour 0 of the CPU path leaves this the plain sum of the fragments)
    float4 sum = make_float4(P.clearColor[0] * (float)S, P.clearColor[1] * (float)S, P.clearColor[2] * (float)S,
                             P.clearColor[3] * (float)S);
    for (int c0 = 0; c0 < F; c0 += 64) {
        // 2. prepare one fragment per lane
        const int fi = c0 + lane;
        const bool act = fi < F;
        FragPrep R;
        int n = 0;
        if (act) {
            const unsigned slot = fSlot[fi];
            const int fl = fPix[fi];
            ShadeTri T;
            load_shade_tri(recs, bary, fanStart, (unsigned)P.ntri, slot, T, P.swapped);
            const int qx = ox + (fl & 7), qy = oy + (fl >> 3);
            R = prep_fragment(P, cam, T, vtx, idx, tex, pool, qx, qy, o, NEURAL || LAYERED, GENERIC);
            if (R.mat >= 0) {
                n = R.fp.n;
                cUV[lane] = make_float4(R.u, R.v, R.fp.au, R.fp.av);
                cLod[lane] = R.fp.lod;
                cN[lane] = n;
                if constexpr (LAYERED) cMat[lane] = R.mat;
                if (!NEURAL && !layered) {
                    // the material's two levels, once per fragment, over tCol2 (only the layered and
                    // neural programs use it)
                    const TriLevels L = tri_levels(pool, R.th, R.fp.lod);
                    // kept as a byte offset into the pool: a pointer rebuilt from the pool
                    // argument stays in the global address space (global, not flat, loads)
                    const unsigned long long a = (unsigned long long)(reinterpret_cast<const uint8_t*>(L.t0) - pool);
                    reinterpret_cast<uint4*>(tCol2)[lane] = make_uint4((unsigned)a, (unsigned)(a >> 32), L.d1, L.dims);
                    cLod[lane] = L.f;
                }
            }
        }
        const int tincl = dev::wave_incl_scan(n);
        const int toff = tincl - n;
        const int TT = __shfl(tincl, 63, 64);
        float4 acc = make_float4(0, 0, 0, 0), acc2 = make_float4(0, 0, 0, 0);
        for (int w0 = 0; w0 < TT; w0 += kTapWindow) {
            __syncthreads();
            for (int i = 0; i < n; i++) {
                const int pos = toff + i - w0;
                if (pos >= 0 && pos < kTapWindow) tList[pos] = (unsigned short)((lane << 4) | i);
            }
            __syncthreads();
            const int nt = min(kTapWindow, TT - w0);
            if (!NEURAL && !layered) {
                // one texture, levels prepared per fragment
                auto tap_texels = [&](int j) {
                    const unsigned e = tList[j];
                    const int fl = (int)(e >> 4), ti = (int)(e & 15u);
                    const float4 uv = cUV[fl];
                    const float invn = frcp((float)cN[fl]);
                    const float to = ((float)ti + 0.5f) * invn - 0.5f;
                    const float tu = fmaf(uv.z, to, uv.x), tv = fmaf(uv.w, to, uv.y);
                    const uint4 q = reinterpret_cast<const uint4*>(tCol2)[fl];
                    TriLevels L;
                    L.t0 = reinterpret_cast<const uchar4*>(pool + (((unsigned long long)q.y << 32) | q.x));
                    L.d1 = q.z;
                    L.dims = q.w;
                    L.f = cLod[fl];
                    if constexpr (GENERIC) return trilinear_fetch_mode(L, tu, tv, P.samp);
                    return trilinear_fetch(L, tu, tv);
                };
                for (int j = lane; j < nt; j += 64) tCol[j] = trilinear_blend(tap_texels(j));
            } else {
                for (int j = lane; j < nt; j += 64) {
                    const unsigned e = tList[j];
                    const int fl = (int)(e >> 4), ti = (int)(e & 15u);
                    const float4 uv = cUV[fl];
                    const int tn = cN[fl];
                    const float invn = frcp((float)tn);
                    const float to = ((float)ti + 0.5f) * invn - 0.5f;
                    const float tu = fmaf(uv.z, to, uv.x), tv = fmaf(uv.w, to, uv.y);
                    if (NEURAL) {
                        float z[kLatentCh];
                        nn::latent_tap_h(P.neural.latent, tu, tv, cLod[fl] - (float)P.neural.lodShift, z);
                        tCol[j] = make_float4(z[0], z[1], z[2], z[3]);
                        tCol2[j] = make_float4(z[4], z[5], z[6], z[7]);
                        continue;
                    }
                    // layered GGX: three textures per tap
                    tCol[j] = trilinear(pool, tex[cMat[fl]], tu, tv, cLod[fl]);
                    const float4 nm = trilinear(pool, tex[1], tu, tv, cLod[fl]);
                    const float4 rg = trilinear(pool, tex[2], tu, tv, cLod[fl]);
                    tCol2[j] = make_float4(nm.x, nm.y, nm.z, rg.x);
                }
            }
            __syncthreads();
            // the lane's own taps of this window, added in tap order; four LDS reads are
            // issued before their adds
            const int b0 = max(toff - w0, 0), b1 = min(toff + n - w0, kTapWindow);
            auto add4 = [](float4& a, const float4 c) {
                a.x += c.x;
                a.y += c.y;
                a.z += c.z;
                a.w += c.w;
            };
            auto sum_taps = [&](const float4* buf, float4& a) {
                int j = b0;
                for (; j + 4 <= b1; j += 4) {
                    const float4 c0 = buf[j], c1 = buf[j + 1], c2 = buf[j + 2], c3 = buf[j + 3];
                    add4(a, c0);
                    add4(a, c1);
                    add4(a, c2);
                    add4(a, c3);
                }
                for (; j < b1; j++) add4(a, buf[j]);
            };
            sum_taps(tCol, acc);
            if (NEURAL || layered) sum_taps(tCol2, acc2);
        }
        if (act) {
            float4 alb = make_float4(1, 1, 1, 1);
            if (R.mat >= 0) {
                const float invn = frcp((float)n);
                alb = make_float4(acc.x * invn, acc.y * invn, acc.z * invn, acc.w * invn);
            }
            if (LAYERED) {
                float4 nm = make_float4(0.5f, 0.5f, 1.0f, 0.0f);
                if (R.mat >= 0) {
                    const float invn = frcp((float)n);
                    nm = make_float4(acc2.x * invn, acc2.y * invn, acc2.z * invn, acc2.w * invn);
                }
                cCol[lane] = ggx_color(P, R, alb, nm);
                if (P.program == 5 && S == 1 && R.mat >= 0) {
                    // training record of this pixel: latent footprint, view and light in the frame, target
                    const int fl = fPix[fi];
                    const size_t fpix = pixBase + (size_t)(oy + (fl >> 3)) * width + ox + (fl & 7);
                    float3 vt, lt;
                    frame_dirs(P, R, vt, lt);
                    const float4 cc = cCol[lane];
                    float4* rec = P.neural.featOut + 4 * fpix;
                    // the latent footprint the neural material will sample (see prep_fragment)
                    const bool one = P.neural.singleTap != 0;
                    rec[0] = make_float4(R.u, R.v, one ? 0.0f : R.fp.au, one ? 0.0f : R.fp.av);
                    rec[1] = make_float4(R.fp.lod + (one ? flog2((float)n) : 0.0f) - (float)P.neural.lodShift,
                                         one ? 1.0f : (float)n, vt.x, vt.y);
                    rec[2] = make_float4(vt.z, lt.x, lt.y, lt.z);
                    rec[3] = make_float4(cc.x, cc.y, cc.z, 0.0f);
                }
            } else if (NEURAL) {
                // latent and directions of this fragment; the network runs for the whole wave below
                const float invn = n > 0 ? frcp((float)n) : 0.0f;
                float3 vt, lt;
                frame_dirs(P, R, vt, lt);
                cUV[lane] = make_float4(acc.x * invn, acc.y * invn, acc.z * invn, acc.w * invn);
                cCol[lane] = make_float4(acc2.x * invn, acc2.y * invn, acc2.z * invn, acc2.w * invn);
                cN[lane] = __float_as_int(vt.x);
                cLod[lane] = vt.y;
                xdir[lane] = make_float4(vt.z, lt.x, lt.y, lt.z);
            } else {
                cCol[lane] = P.program == 0 || P.program == 6 || (GENERIC && P.program == 8)
                                 ? make_float4(sat(R.B.x), sat(R.B.y), sat(R.B.z), R.alpha)
                                 : make_float4(sat(fmaf(alb.x, R.A.x, R.B.x)),
                                               sat(fmaf(alb.y, R.A.y, R.B.y)), sat(fmaf(alb.z, R.A.z, R.B.z)),
                                               sat(alb.w));
            }
        }
        if constexpr (NEURAL) {
            // network inputs [64][16] over the tap buffer, outputs [64][4] over cUV
            __syncthreads();
            float* xs = reinterpret_cast<float*>(tBuf);
            float* ys = reinterpret_cast<float*>(cUV);
            {
                const float4 z0 = cUV[lane], z1 = cCol[lane], d = xdir[lane];
                const float v[16] = {z0.x, z0.y, z0.z, z0.w, z1.x, z1.y, z1.z, z1.w,
                                     __int_as_float(cN[lane]), cLod[lane], d.x, d.y, d.z, d.w, 0.0f, 0.0f};
                __syncthreads();
#pragma unroll
                for (int k = 0; k < 16; k++) xs[lane * 16 + k] = act ? v[k] : 0.0f;
            }
            __syncthreads();
            nn::mlp_forward<FP8>(P.neural.packed, xs, nullptr, nullptr, ys);
            __syncthreads();
            if (act) cCol[lane] = make_float4(sat(ys[lane * 4]), sat(ys[lane * 4 + 1]), sat(ys[lane * 4 + 2]), 1.0f);
        }
        __syncthreads();
        // 3. the pixel's fragments of this chunk, in sample order
        const int k0 = max(poff, c0), k1 = min(poff + nf, c0 + 64);
        for (int k = k0; k < k1; k++) {
            const float4 c = cCol[k - c0];
            const float cnt = (float)fCnt[k];
            sum.x += (c.x - P.clearColor[0]) * cnt;
            sum.y += (c.y - P.clearColor[1]) * cnt;
            sum.z += (c.z - P.clearColor[2]) * cnt;
            sum.w += (c.w - P.clearColor[3]) * cnt;
        }
        __syncthreads();
    }
    if (!inFb || !rc::scissor_pass(P.scissor, (float)px, (float)py)) return;
    const float inv = 1.0f / (float)S;
    const float4 r = make_float4(sum.x * inv, sum.y * inv, sum.z * inv, sum.w * inv);
    rgba8[pix] = make_uchar4((unsigned char)floorf(r.x * 255.0f + 0.5f), (unsigned char)floorf(r.y * 255.0f + 0.5f),
                             (unsigned char)floorf(r.z * 255.0f + 0.5f), (unsigned char)floorf(r.w * 255.0f + 0.5f));
    if (rgba32f) rgba32f[pix] = r;
}

// One wave per shading tile of 8N x 8N pixels (N = 1, 2 or 4: P.shadeTile is
// 8N), its 8x8 sub-tiles in row order.  Workgroups go to the XCDs in turn;
// an XCD takes runs of kShadeRun consecutive tiles (texture and vertex reuse
// in its L2), the runs of the eight XCDs interleaved, so that every XCD
// shades a share of every part of the frame (one band per XCD leaves the
// XCDs of the busy bands last to finish).  Tiles of all shaded views are
// numbered one after another.
constexpr int kShadeRun = 64;

template <int S, bool NEURAL, bool FP8, int N, bool LAYERED, bool GENERIC = false>
__device__ __forceinline__ void shade_tiles(const ShadeParams& P, const Views& V, const TriRecs& recs,
                                            const float4* __restrict__ bary, const uint32_t* __restrict__ visSlot,
                                            const float* __restrict__ vtx, const int* __restrict__ idx,
                                            const TexDesc* __restrict__ tex, const uint8_t* __restrict__ pool,
                                            uchar4* __restrict__ rgba8, float4* __restrict__ rgba32f, int tiles) {
    constexpr int T = 8 * N;
    const int g = (int)(blockIdx.x >> 3);
    const int tile = (g / kShadeRun) * (8 * kShadeRun) + (int)(blockIdx.x & 7u) * kShadeRun + g % kShadeRun;
    if (tile >= tiles) return;
    const ViewDesc& D = V.desc[V.shadeView[tile]];
    const int width = D.width, height = D.height;
    const int local = tile - (int)D.shadeBase, perRow = (width + T - 1) / T;
    const int ox0 = (local % perRow) * T, oy0 = (local / perRow) * T;
    const size_t pb = (size_t)D.pixBase;
    const float3 cam = make_float3(D.camera[0], D.camera[1], D.camera[2]);
    const unsigned fanStart = D.slotBase + (unsigned)P.ntri;
    if constexpr (N == 1) {
        shade_subtile<S, NEURAL, FP8, LAYERED, GENERIC>(P, cam, width, height, fanStart, ox0, oy0, pb, recs, bary,
                                                        visSlot + pb * S,
                                      vtx, idx, tex, pool, rgba8 + pb, rgba32f ? rgba32f + pb : nullptr);
    } else {
#pragma unroll 1
        for (int k = 0; k < N * N; k++) {
            const int ox = ox0 + (k % N) * 8, oy = oy0 + (k / N) * 8;
            if (ox >= width || oy >= height) continue;
            shade_subtile<S, NEURAL, FP8, LAYERED, GENERIC>(P, cam, width, height, fanStart, ox, oy, pb, recs, bary,
                                                            visSlot + pb * S,
                                          vtx, idx, tex, pool, rgba8 + pb, rgba32f ? rgba32f + pb : nullptr);
        }
    }
}

// One fragment's colour at a 1-sample pixel, for the ordered blending path:
// the arithmetic of the deferred shading in shade_subtile (taps summed in
// order), so an opaque fragment gets the same colour either way.
__device__ float4 shade_fragment(const ShadeParams& P, float3 cam, const ShadeTri& T, const float* vtx, const int* idx,
                                 const TexDesc* tex, const uint8_t* pool, int px, int py) {
    const FragPrep R = prep_fragment(P, cam, T, vtx, idx, tex, pool, px, py, 0.0f, true);
    if (P.program == 0 || P.program == 6) return make_float4(sat(R.B.x), sat(R.B.y), sat(R.B.z), R.alpha);
    float4 acc = make_float4(0, 0, 0, 0), acc2 = make_float4(0, 0, 0, 0);
    const int n = R.mat >= 0 ? R.fp.n : 0;
    for (int i = 0; i < n; i++) {
        const float invn = frcp((float)n);
        const float to = ((float)i + 0.5f) * invn - 0.5f;
        const float tu = fmaf(R.fp.au, to, R.u), tv = fmaf(R.fp.av, to, R.v);
        const float4 c = trilinear(pool, tex[R.mat], tu, tv, R.fp.lod);
        acc.x += c.x;
        acc.y += c.y;
        acc.z += c.z;
        acc.w += c.w;
        if (P.program == 3) {
            const float4 nm = trilinear(pool, tex[1], tu, tv, R.fp.lod);
            const float4 rg = trilinear(pool, tex[2], tu, tv, R.fp.lod);
            acc2.x += nm.x;
            acc2.y += nm.y;
            acc2.z += nm.z;
            acc2.w += rg.x;
        }
    }
    float4 alb = make_float4(1, 1, 1, 1);
    if (R.mat >= 0) {
        const float invn = frcp((float)n);
        alb = make_float4(acc.x * invn, acc.y * invn, acc.z * invn, acc.w * invn);
    }
    if (P.program == 3) {
        float4 nm = make_float4(0.5f, 0.5f, 1.0f, 0.0f);
        if (R.mat >= 0) {
            const float invn = frcp((float)n);
            nm = make_float4(acc2.x * invn, acc2.y * invn, acc2.z * invn, acc2.w * invn);
        }
        return ggx_color(P, R, alb, nm);
    }
    return make_float4(sat(fmaf(alb.x, R.A.x, R.B.x)), sat(fmaf(alb.y, R.A.y, R.B.y)), sat(fmaf(alb.z, R.A.z, R.B.z)),
                       sat(alb.w));
}

// The blend of TileRasterWorker.cpp:2346-2373 in FP32, in the same operation
// order; factors come encoded as (c0, c1, c2, c3) meaning
// (c0 + c1 (1 - src alpha) + c2 (1 - dst alpha)) (1 - c3)
// (TileRasterRenderer.cpp:212-333).
__device__ __forceinline__ float blend_factor(const float* k, float srcA, float dstA) {
    return (k[0] + k[1] * (1 - srcA) + k[2] * (1 - dstA)) * (1 - k[3]);
}

__device__ __forceinline__ float4 blend_pixel(const BlendCoefs& B, float4 src, float4 dst) {
    const float mxSrcX = blend_factor(B.sc, src.w, dst.w), mxDstX = blend_factor(B.dc, src.w, dst.w);
    const float mxSrcA = blend_factor(B.sa, src.w, dst.w), mxDstA = blend_factor(B.da, src.w, dst.w);
    return make_float4(dst.x * mxDstX + src.x * mxSrcX, dst.y * mxDstX + src.y * mxSrcX,
                       dst.z * mxDstX + src.z * mxSrcX, dst.w * mxDstA + src.w * mxSrcA);
}

using dev::Edges;
using dev::block_test;
using dev::block_sum256;
using dev::load_edges;
using dev::load_rec;

// Blending: the ordered traversal of each tile's list with the depth test of
// the ordered raster, and every fragment that passes shaded and blended into
// the pixel's FP32 colour in primitive order (1 sample per pixel).  Writes the
// visibility (last writer), the FP32 colour and its RGBA8 rounding.
__global__ void __launch_bounds__(256) k_blend_ordered(ShadeParams P, BlendCoefs B, int depthFunc, int clear,
                                                       unsigned binCap, Views V, TriRecs recs,
                                                       const float4* __restrict__ bary, const float* __restrict__ vtx,
                                                       const int* __restrict__ idx, const TexDesc* __restrict__ tex,
                                                       const uint8_t* __restrict__ pool,
                                                       const unsigned* __restrict__ binStart,
                                                       const unsigned* __restrict__ binCount,
                                                       const unsigned* __restrict__ binList, int tiles,
                                                       float* __restrict__ visDepthAll,
                                                       uint32_t* __restrict__ visSlotAll,
                                                       uint8_t* __restrict__ visClassAll, uchar4* __restrict__ rgba8All,
                                                       float4* __restrict__ rgba32fAll, StatPartials st) {
    __shared__ unsigned long long sA[rc::kOrderCap], sB[rc::kOrderCap];
    const int t = threadIdx.x;
    unsigned long long frags = 0;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const ViewDesc& D = V.desc[V.tileView[tile]];
        const rc::ViewGeom G = rc::view_of_tile(V, tile);
        const rc::PixelCtx C = rc::pixel_ctx(G, t & 15, t >> 4);
        const size_t pix = G.pixBase + C.pix;
        const float3 cam = make_float3(D.camera[0], D.camera[1], D.camera[2]);
        const unsigned fanStart = D.slotBase + (unsigned)P.ntri;
        const bool sciss = rc::scissor_pass(P.scissor, C.fpx, C.fpy);  // outside: the pixel is left as it is
        float dep = P.clearDepth;
        unsigned slot = 0xffffffffu, cls = 0;
        // the colour attachment is cleared to the clear colour (0 in the CPU path, TileRasterWorker.cpp:113-118)
        float4 col = make_float4(P.clearColor[0], P.clearColor[1], P.clearColor[2], P.clearColor[3]);
        if (!clear && C.inFb) {
            dep = visDepthAll[pix];
            slot = visSlotAll[pix];
            if (visClassAll) cls = visClassAll[pix];
            col = rgba32fAll[pix];
        }
        const unsigned start = binStart[tile];
        const unsigned cnt = start >= binCap ? 0u : min(binCount[tile], binCap - start);
        rc::for_each_ordered(recs, binList, start, cnt, sA, sB, [&](unsigned ent) {
            const unsigned sl = ent & 0x7fffffffu;
            const bool full = (ent >> 31) != 0;
            if (!C.inFb || !sciss) return;
            const TriRec R = load_rec(recs, sl);
            const Edges E = load_edges(R);
            int trn, tan;
            block_test(E, C.bMinX, C.bMinY, trn, tan);
            if (!rc::covers(E, full, tan, trn, C.fpx, C.fpy)) return;
            const unsigned c = rc::pixel_class(E, full || tan == 3, C.fpx, C.fpy, C.bMinX, false);
            const float d = rc::class_depth(c, R.f1, R.f2, R.f3, C.fpx, C.fpy);
            frags++;
            if (!rc::depth_pass(depthFunc, d, dep, c == 3u)) return;
            ShadeTri T;
            load_shade_tri(recs, bary, fanStart, (unsigned)P.ntri, sl, T, P.swapped);
            col = blend_pixel(B, shade_fragment(P, cam, T, vtx, idx, tex, pool, C.px, C.py), col);
            dep = P.depthWrite ? d * P.depthScale + P.depthBias : d;
            slot = sl;
            cls = c;
        });
        if (C.inFb && sciss) {
            visDepthAll[pix] = dep;
            visSlotAll[pix] = slot;
            if (visClassAll) visClassAll[pix] = (uint8_t)cls;
            rgba32fAll[pix] = col;
            rgba8All[pix] = make_uchar4((unsigned char)floorf(sat(col.x) * 255.0f + 0.5f),
                                        (unsigned char)floorf(sat(col.y) * 255.0f + 0.5f),
                                        (unsigned char)floorf(sat(col.z) * 255.0f + 0.5f),
                                        (unsigned char)floorf(sat(col.w) * 255.0f + 0.5f));
        }
    }
    const unsigned long long tf = block_sum256(frags);
    if (t == 0) {
        st.rasterFrags[blockIdx.x] = tf;
        st.rasterCovs[blockIdx.x] = tf;
    }
}

// One texture (Blinn-Phong) or flat colour: five waves per SIMD (at most 96
// VGPRs, with no spills, and under 8 KB of LDS per wave).  Fetching two taps
// per lane at once needs more registers: at four waves it is slower.
template <int S, int N>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_shade(ShadeParams P, Views V, TriRecs recs, const float4* __restrict__ bary, const uint32_t* __restrict__ visSlot,
        const float* __restrict__ vtx, const int* __restrict__ idx, const TexDesc* __restrict__ tex,
        const uint8_t* __restrict__ pool, uchar4* __restrict__ rgba8, float4* __restrict__ rgba32f, int tiles) {
    shade_tiles<S, false, false, N, false>(P, V, recs, bary, visSlot, vtx, idx, tex, pool, rgba8, rgba32f, tiles);
}

// The single-texture and cube-map programs under a sampler other than
// repeat + bilinear (address modes, border colours, nearest texels): a
// variant of k_shade, so that the default sampler's kernel carries none of it.
template <int S, int N>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_shade_sampled(ShadeParams P, Views V, TriRecs recs, const float4* __restrict__ bary,
                const uint32_t* __restrict__ visSlot, const float* __restrict__ vtx, const int* __restrict__ idx,
                const TexDesc* __restrict__ tex, const uint8_t* __restrict__ pool, uchar4* __restrict__ rgba8,
                float4* __restrict__ rgba32f, int tiles) {
    shade_tiles<S, false, false, N, false, true>(P, V, recs, bary, visSlot, vtx, idx, tex, pool, rgba8, rgba32f, tiles);
}

// Layered GGX (programs 3 and 5): four waves per SIMD (at most 128 VGPRs), as
// the neural program.
template <int S, int N>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_shade_layered(ShadeParams P, Views V, TriRecs recs, const float4* __restrict__ bary,
                const uint32_t* __restrict__ visSlot, const float* __restrict__ vtx, const int* __restrict__ idx,
                const TexDesc* __restrict__ tex, const uint8_t* __restrict__ pool, uchar4* __restrict__ rgba8,
                float4* __restrict__ rgba32f, int tiles) {
    shade_tiles<S, false, false, N, true>(P, V, recs, bary, visSlot, vtx, idx, tex, pool, rgba8, rgba32f, tiles);
}

// The neural material, held to four waves per SIMD like the other programs so
// that the latent fetches stay overlapped.
template <int S, bool FP8, int N>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_shade_neural(ShadeParams P, Views V, TriRecs recs, const float4* __restrict__ bary,
               const uint32_t* __restrict__ visSlot, const float* __restrict__ vtx, const int* __restrict__ idx,
               const TexDesc* __restrict__ tex, const uint8_t* __restrict__ pool, uchar4* __restrict__ rgba8,
               float4* __restrict__ rgba32f, int tiles) {
    shade_tiles<S, true, FP8, N, false>(P, V, recs, bary, visSlot, vtx, idx, tex, pool, rgba8, rgba32f, tiles);
}

}  // namespace

void launch_blend_ordered(const ShadeParams& p, const BlendCoefs& b, int depthFunc, int clear, unsigned binCap,
                          const Views& views, TriRecs recs, const float4* bary, const float* vtx, const int* idx,
                          const TexDesc* tex, const uint8_t* texPool, const unsigned* binStart,
                          const unsigned* binCount, const unsigned* binList, int tiles, int grid, float* visDepth,
                          uint32_t* visSlot, uint8_t* visClass, uchar4* rgba8, float4* rgba32f, StatPartials st,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_blend_ordered, dim3(grid), dim3(256), 0, s, p, b, depthFunc, clear, binCap, views, recs, bary,
                       vtx, idx, tex, texPool, binStart, binCount, binList, tiles, visDepth, visSlot, visClass, rgba8,
                       rgba32f, st);
}

void launch_shade(const ShadeParams& p, const Views& views, int shadeTiles, TriRecs recs, const float4* bary,
                  const uint32_t* visSlot, const float* vtx, const int* idx, const TexDesc* tex,
                  const uint8_t* texPool, uchar4* rgba8, float4* rgba32f, hipStream_t s) {
    if (shadeTiles <= 0) return;
    // runs of kShadeRun tiles per XCD: the grid covers whole rounds of eight runs
    const int perXcd = ((shadeTiles + 7) / 8 + kShadeRun - 1) / kShadeRun * kShadeRun;
    const dim3 grid(8 * perXcd);
#define IFRIT_ARGS p, views, recs, bary, visSlot, vtx, idx, tex, texPool, rgba8, rgba32f, shadeTiles
#define IFRIT_SHADE_N(SN, NN)                                                                                      \
    do {                                                                                                            \
        if (p.program == 3 || p.program == 5)                                                                       \
            hipLaunchKernelGGL((k_shade_layered<SN, NN>), grid, dim3(64), 0, s, IFRIT_ARGS);                        \
        else if (p.sampGeneric || p.program == 8)                                                                   \
            hipLaunchKernelGGL((k_shade_sampled<SN, NN>), grid, dim3(64), 0, s, IFRIT_ARGS);                        \
        else if (p.program != 4)                                                                                    \
            hipLaunchKernelGGL((k_shade<SN, NN>), grid, dim3(64), 0, s, IFRIT_ARGS);                                \
        else if (p.neural.precision == 1)                                                                           \
            hipLaunchKernelGGL((k_shade_neural<SN, true, NN>), grid, dim3(64), 0, s, IFRIT_ARGS);                   \
        else                                                                                                        \
            hipLaunchKernelGGL((k_shade_neural<SN, false, NN>), grid, dim3(64), 0, s, IFRIT_ARGS);                  \
    } while (0)
#define IFRIT_SHADE(SN)                                                                                             \
    do {                                                                                                            \
        if (p.shadeTile == 32)                                                                                      \
            IFRIT_SHADE_N(SN, 4);                                                                                   \
        else if (p.shadeTile == 16)                                                                                 \
            IFRIT_SHADE_N(SN, 2);                                                                                   \
        else                                                                                                        \
            IFRIT_SHADE_N(SN, 1);                                                                                   \
    } while (0)
    switch (p.samples) {
        case 1: IFRIT_SHADE(1); break;
        case 4: IFRIT_SHADE(4); break;
        default: IFRIT_SHADE(8); break;
    }
#undef IFRIT_SHADE
#undef IFRIT_SHADE_N
#undef IFRIT_ARGS
}

}  // namespace ifrit_hip
