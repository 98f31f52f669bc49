// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// Drives the reference CPU tile rasterizer, compiled from its own sources
// under /root/reference by oracle/ref/Makefile, so that the restatement in
// oracle/oracle_core.cpp can be checked against the real thing.  Nothing of
// the reference is copied here: this file only calls its public C++ API
// (TileRasterRenderer.h:86-104, FrameBuffer.h, VertexBuffer.h,
// BufferManager.h, Shaders.h).
//
// Exposed entry points (ctypes):
//   ref_render_visibility  per-pixel primitive ID and depth of the winning
//                          fragment for given clip-space vertices;
//   ref_render_timed       a full frame (vertex transform + Blinn-Phong-like
//                          shading) timed with the reference's own 17 threads;
//   ref_render_blend       the FP32 colour attachment of a draw with blending
//                          and a given depth function, each triangle coloured
//                          by a hash of its index (colour and alpha);
//   ref_render_visibility_tagging / _depthwrite
//                          ref_render_visibility in the tagging mode, or with a
//                          fragment shader that writes depth * scale + bias.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "ifrit/softgraphics/engine/base/FrameBuffer.h"
#include "ifrit/softgraphics/engine/base/Shaders.h"
#include "ifrit/softgraphics/engine/base/TypeDescriptor.h"
#include "ifrit/softgraphics/engine/base/VaryingDescriptor.h"
#include "ifrit/softgraphics/engine/base/VertexBuffer.h"
#include "ifrit/softgraphics/engine/bufferman/BufferManager.h"
#include "ifrit/softgraphics/engine/tileraster/TileRasterRenderer.h"

using namespace Ifrit;
using namespace Ifrit::Graphics::SoftGraphics;
using namespace Ifrit::Graphics::SoftGraphics::TileRaster;
using Ifrit::Graphics::SoftGraphics::Core::Data::ImageF32;

namespace {

// Outputs the clip position stored in attribute 0 and passes attribute 1
// through as varying 0.
class PassVS : public VertexShader {
public:
    void execute(const void* const* in, Vector4f* outPos, Vector4f* const* outVary) override {
        *outPos = *reinterpret_cast<const Vector4f*>(in[0]);
        *outVary[0] = *reinterpret_cast<const Vector4f*>(in[1]);
    }
};

// Writes (interpolated varying 0.x, fragment depth, 1, 1): with a constant
// per-triangle varying this records the winning primitive and its depth.
class IdDepthFS : public FragmentShader {
public:
    void execute(const void* vary, void* color, float* depth) override {
        const Vector4f v = reinterpret_cast<const Vector4f*>(vary)[0];
        Vector4f* c = reinterpret_cast<Vector4f*>(color);
        c->x = v.x;
        c->y = *depth;
        c->z = 1.0f;
        c->w = 1.0f;
    }
};

// Colour and alpha from a hash of the primitive index carried in varying 0
// (the same hash and byte layout as the GPU's translucent flat program):
// rgb = bytes 0-2 / 255, alpha = (64 + (byte 3 & 127)) / 255.
class FlatAlphaFS : public FragmentShader {
public:
    void execute(const void* vary, void* color, float* depth) override {
        (void)depth;
        const Vector4f v = reinterpret_cast<const Vector4f*>(vary)[0];
        const uint32_t prim = (uint32_t)std::nearbyint(v.x);
        uint32_t h = prim * 0x9E3779B1u;
        h ^= h >> 15;
        h *= 0x85EBCA77u;
        h ^= h >> 13;
        Vector4f* c = reinterpret_cast<Vector4f*>(color);
        c->x = (float)(h & 0xffu) * (1.0f / 255.0f);
        c->y = (float)((h >> 8) & 0xffu) * (1.0f / 255.0f);
        c->z = (float)((h >> 16) & 0xffu) * (1.0f / 255.0f);
        c->w = (float)(64u + ((h >> 24) & 0x7fu)) * (1.0f / 255.0f);
    }
};

// clip = M * (p, 1) in the MatMul order of LinalgOps.h:40-48; varying 0 is
// the normal.
class TransformVS : public VertexShader {
public:
    float m[16];
    void execute(const void* const* in, Vector4f* outPos, Vector4f* const* outVary) override {
        const float* p = reinterpret_cast<const float*>(in[0]);
        const float* n = reinterpret_cast<const float*>(in[1]);
        outPos->x = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3] * 1.0f;
        outPos->y = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7] * 1.0f;
        outPos->z = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11] * 1.0f;
        outPos->w = m[12] * p[0] + m[13] * p[1] + m[14] * p[2] + m[15] * 1.0f;
        outVary[0]->x = n[0];
        outVary[0]->y = n[1];
        outVary[0]->z = n[2];
        outVary[0]->w = 0.0f;
    }
};

// Lambert + Blinn-style specular with one directional light.
class LambertFS : public FragmentShader {
public:
    void execute(const void* vary, void* color, float* depth) override {
        const Vector4f n = reinterpret_cast<const Vector4f*>(vary)[0];
        const float l = 1.0f / std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z + 1e-20f);
        float d = (n.x * 0.3f + n.y * 0.8f - n.z * 0.52f) * l;
        d = d > 0.0f ? d : 0.0f;
        const float s = std::pow(d, 32.0f);
        Vector4f* c = reinterpret_cast<Vector4f*>(color);
        c->x = 0.1f + 0.8f * d + s;
        c->y = 0.1f + 0.6f * d + s;
        c->z = 0.1f + 0.4f * d + s;
        c->w = 1.0f;
    }
};

// Writes depth * scale + bias through the depth out-parameter and records
// (interpolated varying 0.x, written depth, 1, 1).
class DepthAffineFS : public FragmentShader {
public:
    float scale = 1.0f, bias = 0.0f;
    void execute(const void* vary, void* color, float* depth) override {
        const Vector4f v = reinterpret_cast<const Vector4f*>(vary)[0];
        const float d = *depth * scale + bias;
        *depth = d;
        Vector4f* c = reinterpret_cast<Vector4f*>(color);
        c->x = v.x;
        c->y = d;
        c->z = 1.0f;
        c->w = 1.0f;
    }
};

struct Target {
    std::unique_ptr<ImageF32> color, depth;
    FrameBuffer fb;
    Target(int w, int h) {
        color = std::make_unique<ImageF32>(w, h, 4);
        depth = std::make_unique<ImageF32>(w, h, 1);
        fb.SetColorAttachmentsCompatible(reinterpret_cast<ImageF32* const*>(&color), 1);
        fb.SetDepthAttachmentCompatible(depth.get());
    }
};

}  // namespace

// tagging: draws in the non-deterministic tagging mode (optsetForceDeterministic(false),
// TRW:1581-1675), whose depth write-back uses the tile's pixel origin, so depth is read
// from the depth attachment rather than from the fragment shader's output.
static int render_visibility(int W, int H, int depthFunc, bool tagging, const float* clip, const int32_t* idx,
                             int nidx, uint32_t* id_out, float* depth_out, const float* depthAffine = nullptr) {
    const int ntri = nidx / 3;
    if (ntri <= 0) return -1;
    Target tgt(W, H);
    VertexBuffer vb;
    vb.setLayout({TypeDescriptors.FLOAT4, TypeDescriptors.FLOAT4});
    vb.setVertexCount(3 * ntri);
    vb.allocateBuffer(3 * ntri);
    for (int p = 0; p < ntri; p++) {
        for (int k = 0; k < 3; k++) {
            const float* c = clip + 4 * idx[3 * p + k];
            vb.setValue<Vector4f>(3 * p + k, 0, Vector4f{c[0], c[1], c[2], c[3]});
            vb.setValue<Vector4f>(3 * p + k, 1, Vector4f{(float)p, 0.0f, 0.0f, 0.0f});
        }
    }
    std::vector<int> seq(3 * ntri);
    for (int i = 0; i < 3 * ntri; i++) seq[i] = i;
    auto mgr = std::make_shared<BufferManager::TrivialBufferManager>();
    mgr->Init();
    auto ib = mgr->CreateBuffer({seq.size() * sizeof(int)});
    mgr->bufferData(ib, seq.data(), 0, seq.size() * sizeof(int));

    PassVS vs;
    IdDepthFS idfs;
    DepthAffineFS affine;
    if (depthAffine) {
        affine.scale = depthAffine[0];
        affine.bias = depthAffine[1];
    }
    FragmentShader& fs = depthAffine ? static_cast<FragmentShader&>(affine) : static_cast<FragmentShader&>(idfs);
    VaryingDescriptor vd;
    vd.setVaryingDescriptors({TypeDescriptors.FLOAT4});
    auto r = std::make_shared<TileRasterRenderer>();
    r->Init();
    r->bindFrameBuffer(tgt.fb);
    r->bindVertexBuffer(vb);
    r->bindIndexBuffer(ib);
    r->bindVertexShaderLegacy(vs, vd);
    r->bindFragmentShader(fs);
    r->SetDepthFunc((IfritCompareOp)depthFunc);
    r->optsetForceDeterministic(!tagging);
    r->drawElements(3 * ntri, true);
    const float* c = tgt.color->getData();
    const float* z = tgt.depth->getData();
    int bad = 0;
    for (size_t i = 0; i < (size_t)W * H; i++) {
        if (c[4 * i + 3] == 1.0f) {
            const float v = c[4 * i];
            const float rv = std::nearbyint(v);
            if (std::fabs(v - rv) > 0.01f) bad++;
            id_out[i] = (uint32_t)rv;
            depth_out[i] = tagging ? z[i] : c[4 * i + 1];
        } else {
            id_out[i] = 0xffffffffu;
            depth_out[i] = 0.0f;
        }
    }
    return bad;
}

extern "C" int ref_render_visibility(int W, int H, int ccw, int depthFunc, const float* clip, int nverts,
                                     const int32_t* idx, int nidx, uint32_t* id_out, float* depth_out) {
    (void)ccw;
    (void)nverts;
    return render_visibility(W, H, depthFunc, false, clip, idx, nidx, id_out, depth_out);
}

extern "C" int ref_render_visibility_depthwrite(int W, int H, int depthFunc, float scale, float bias,
                                                const float* clip, const int32_t* idx, int nidx, uint32_t* id_out,
                                                float* depth_out) {
    const float affine[2] = {scale, bias};
    return render_visibility(W, H, depthFunc, false, clip, idx, nidx, id_out, depth_out, affine);
}

extern "C" int ref_render_visibility_tagging(int W, int H, int depthFunc, const float* clip, const int32_t* idx,
                                             int nidx, uint32_t* id_out, float* depth_out) {
    return render_visibility(W, H, depthFunc, true, clip, idx, nidx, id_out, depth_out);
}

extern "C" int ref_render_blend(int W, int H, int depthFunc, int srcColor, int dstColor, int srcAlpha, int dstAlpha,
                                const float* clip, const int32_t* idx, int nidx, float* rgba_out) {
    const int ntri = nidx / 3;
    if (ntri <= 0) return -1;
    Target tgt(W, H);
    VertexBuffer vb;
    vb.setLayout({TypeDescriptors.FLOAT4, TypeDescriptors.FLOAT4});
    vb.setVertexCount(3 * ntri);
    vb.allocateBuffer(3 * ntri);
    for (int p = 0; p < ntri; p++) {
        for (int k = 0; k < 3; k++) {
            const float* c = clip + 4 * idx[3 * p + k];
            vb.setValue<Vector4f>(3 * p + k, 0, Vector4f{c[0], c[1], c[2], c[3]});
            vb.setValue<Vector4f>(3 * p + k, 1, Vector4f{(float)p, 0.0f, 0.0f, 0.0f});
        }
    }
    std::vector<int> seq(3 * ntri);
    for (int i = 0; i < 3 * ntri; i++) seq[i] = i;
    auto mgr = std::make_shared<BufferManager::TrivialBufferManager>();
    mgr->Init();
    auto ib = mgr->CreateBuffer({seq.size() * sizeof(int)});
    mgr->bufferData(ib, seq.data(), 0, seq.size() * sizeof(int));
    PassVS vs;
    FlatAlphaFS fs;
    VaryingDescriptor vd;
    vd.setVaryingDescriptors({TypeDescriptors.FLOAT4});
    auto r = std::make_shared<TileRasterRenderer>();
    r->Init();
    r->bindFrameBuffer(tgt.fb);
    r->bindVertexBuffer(vb);
    r->bindIndexBuffer(ib);
    r->bindVertexShaderLegacy(vs, vd);
    r->bindFragmentShader(fs);
    r->SetDepthFunc((IfritCompareOp)depthFunc);
    IfritColorAttachmentBlendState bs{};
    bs.blendEnable = true;
    bs.srcColorBlendFactor = (IfritBlendFactor)srcColor;
    bs.dstColorBlendFactor = (IfritBlendFactor)dstColor;
    bs.srcAlphaBlendFactor = (IfritBlendFactor)srcAlpha;
    bs.dstAlphaBlendFactor = (IfritBlendFactor)dstAlpha;
    r->setBlendFunc(bs);
    r->drawElements(3 * ntri, true);
    std::memcpy(rgba_out, tgt.color->getData(), sizeof(float) * 4 * (size_t)W * H);
    return 0;
}

extern "C" double ref_render_timed(int W, int H, const float* verts, int stride, int nverts, const int32_t* idx,
                                   int nidx, const float* mvp, int frames, float* color_out) {
    Target tgt(W, H);
    VertexBuffer vb;
    vb.setLayout({TypeDescriptors.FLOAT3, TypeDescriptors.FLOAT3});
    vb.setVertexCount(nverts);
    vb.allocateBuffer(nverts);
    for (int i = 0; i < nverts; i++) {
        const float* v = verts + (size_t)i * stride;
        vb.setValue<Vector3f>(i, 0, Vector3f{v[0], v[1], v[2]});
        vb.setValue<Vector3f>(i, 1, Vector3f{v[3], v[4], v[5]});
    }
    auto mgr = std::make_shared<BufferManager::TrivialBufferManager>();
    mgr->Init();
    auto ib = mgr->CreateBuffer({(size_t)nidx * sizeof(int)});
    mgr->bufferData(ib, idx, 0, (size_t)nidx * sizeof(int));
    TransformVS vs;
    std::memcpy(vs.m, mvp, sizeof(vs.m));
    LambertFS fs;
    VaryingDescriptor vd;
    vd.setVaryingDescriptors({TypeDescriptors.FLOAT4});
    auto r = std::make_shared<TileRasterRenderer>();
    r->Init();
    r->bindFrameBuffer(tgt.fb);
    r->bindVertexBuffer(vb);
    r->bindIndexBuffer(ib);
    r->bindVertexShaderLegacy(vs, vd);
    r->bindFragmentShader(fs);
    r->drawElements(nidx, true);  // warm-up
    std::vector<double> ms;
    for (int f = 0; f < frames; f++) {
        auto t0 = std::chrono::steady_clock::now();
        r->drawElements(nidx, true);
        auto t1 = std::chrono::steady_clock::now();
        ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
    if (color_out) std::memcpy(color_out, tgt.color->getData(), (size_t)W * H * 4 * sizeof(float));
    std::sort(ms.begin(), ms.end());
    return ms[ms.size() / 2];
}
