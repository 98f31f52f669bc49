// extern "C" entry points: the reference's exported ABI
// (projects/softgraphics/src/engine/export/TileRasterRendererExport.cpp,
// EngineCommonExport.cpp, BufferManagerExport.cpp,
// projects/softgraphics/src/core/export/CommonExport.cpp) with the same
// symbols and argument meaning, backed by the MI355X renderer, plus the
// backend additions declared in include/ifrit_hip.h.
#include <cstring>
#include <new>

#include "renderer.hpp"

using namespace ifrit_hip;

struct IfritRendererHandle {
    TileRasterRenderer renderer;
};
struct IfritFrameBuffer {
    FrameBuffer fb;
};
struct IfritImageF32 {
    ImageF32 img;
};
struct IfritVertexBuffer {
    VertexBuffer vb;
};
struct IfritVaryingDescriptor {
    VaryingDescriptor vd;
};

namespace {
// TypeDescriptor sizes (TypeDescriptor.h:49-63); IFTP_UNDEFINED maps to FLOAT4
// as in EngineCommonExport.cpp:51-52.
int type_size(int t) {
    switch (t) {
        case IFTP_FLOAT1: case IFTP_INT1: case IFTP_UINT1: return 4;
        case IFTP_FLOAT2: case IFTP_INT2: case IFTP_UINT2: return 8;
        case IFTP_FLOAT3: case IFTP_INT3: case IFTP_UINT3: return 12;
        default: return 16;
    }
}
template <class F>
void guarded(IfritRendererHandle* h, F f) {
    if (!h) return;
    try {
        f(h->renderer);
    } catch (const std::bad_alloc&) {
        h->renderer.setError("host allocation failed");
    } catch (...) {
        h->renderer.setError("unexpected exception");
    }
}
}  // namespace

extern "C" {

IfritRendererHandle* iftrCreateInstance(void) { return new (std::nothrow) IfritRendererHandle(); }
void iftrDestroyInstance(IfritRendererHandle* h) { delete h; }
void iftrInit(IfritRendererHandle* h) {
    guarded(h, [](TileRasterRenderer& r) { r.Init(); });
}
void iftrBindFrameBuffer(IfritRendererHandle* h, IfritFrameBuffer* fb) {
    if (fb) guarded(h, [&](TileRasterRenderer& r) { r.bindFrameBuffer(fb->fb); });
}
void iftrBindVertexBuffer(IfritRendererHandle* h, const IfritVertexBuffer* vb) {
    if (vb) guarded(h, [&](TileRasterRenderer& r) { r.bindVertexBuffer(vb->vb); });
}
void iftrBindIndexBuffer(IfritRendererHandle* h, void* indexBuffer) {
    if (indexBuffer) guarded(h, [&](TileRasterRenderer& r) { r.bindIndexBuffer(*static_cast<Buffer*>(indexBuffer)); });
}
void iftrBindVertexShaderFunc(IfritRendererHandle* h, IfritVertexShaderFunctionalPtr func,
                              IfritVaryingDescriptor* vd) {
    guarded(h, [&](TileRasterRenderer& r) { r.bindVertexShaderFunc(func, vd ? &vd->vd : nullptr); });
}
void iftrBindFragmentShaderFunc(IfritRendererHandle* h, IfritFragmentShaderFunctionalPtr func) {
    guarded(h, [&](TileRasterRenderer& r) { r.bindFragmentShaderFunc(func); });
}
void iftrSetBlendFunc(IfritRendererHandle* h, IfritColorAttachmentBlendState* state) {
    if (state) guarded(h, [&](TileRasterRenderer& r) { r.setBlendFunc(*state); });
}
void iftrSetDepthFunc(IfritRendererHandle* h, int32_t state) {
    guarded(h, [&](TileRasterRenderer& r) { r.SetDepthFunc(state); });
}
void iftrOptsetForceDeterministic(IfritRendererHandle* h, int opt) {
    guarded(h, [&](TileRasterRenderer& r) { r.optsetForceDeterministic(opt != 0); });
}
void iftrOptSetDepthTestEnable(IfritRendererHandle* h, int opt) {
    guarded(h, [&](TileRasterRenderer& r) { r.optSetDepthTestEnable(opt != 0); });
}
void iftrOptsetDepthTestEnable(IfritRendererHandle* h, int opt) { iftrOptSetDepthTestEnable(h, opt); }
void iftrDrawLegacy(IfritRendererHandle* h, int numVertices, int clearFramebuffer) {
    guarded(h, [&](TileRasterRenderer& r) { r.drawElements(numVertices, clearFramebuffer != 0); });
}
void iftrClear(IfritRendererHandle* h) {
    guarded(h, [](TileRasterRenderer& r) { r.clear(); });
}
void iftrTest(void (*p)(int)) {
    if (p) p(114514);  // TileRasterRendererExport.cpp:151-155 calls back with this constant
}
void iftrBindVertexShader(IfritRendererHandle* h, void* p) {
    if (!p) return;
    const GpuVertexProgram* prog = static_cast<const GpuVertexProgram*>(p);
    guarded(h, [&](TileRasterRenderer& r) {
        if (prog->tag != kVertexProgramTag)
            r.setError("iftrBindVertexShader expects an object from iftrCreateGpuVertexProgram");
        else
            r.bindVertexShader(*prog);
    });
}
void iftrBindFragmentShader(IfritRendererHandle* h, void* p) {
    if (!p) return;
    const GpuFragmentProgram* prog = static_cast<const GpuFragmentProgram*>(p);
    guarded(h, [&](TileRasterRenderer& r) {
        if (prog->tag != kFragmentProgramTag)
            r.setError("iftrBindFragmentShader expects an object from iftrCreateGpuFragmentProgram");
        else
            r.bindFragmentShader(*prog);
    });
}

IfritFrameBuffer* iftrCreateFrameBuffer(void) { return new (std::nothrow) IfritFrameBuffer(); }
void iftrDestroyFrameBuffer(IfritFrameBuffer* p) { delete p; }
void iftrSetFrameBufferColorAttachmentFP32(IfritFrameBuffer* p, IfritImageF32* const* img, size_t n) {
    if (!p) return;
    p->fb.color.clear();
    for (size_t i = 0; i < n; i++) p->fb.color.push_back(img[i] ? &img[i]->img : nullptr);
    if (n && img[0]) {
        p->fb.width = (int)img[0]->img.width;
        p->fb.height = (int)img[0]->img.height;
    }
}
void iftrSetFrameBufferDepthAttachmentFP32(IfritFrameBuffer* p, IfritImageF32* img) {
    if (p) p->fb.depth = img ? &img->img : nullptr;
}
IfritVaryingDescriptor* iftrCreateVaryingDescriptor(void) { return new (std::nothrow) IfritVaryingDescriptor(); }
void iftrDestroyVaryingDescriptor(IfritVaryingDescriptor* p) { delete p; }
void iftrWriteVaryingDescriptor(IfritVaryingDescriptor* p, const int* d, size_t n) {
    if (!p) return;
    p->vd.types.assign(d, d + n);
}
IfritVertexBuffer* iftrCreateVertexBuffer(void) { return new (std::nothrow) IfritVertexBuffer(); }
void iftrDestroyVertexBuffer(IfritVertexBuffer* p) { delete p; }
void iftrSetVertexBufferLayout(IfritVertexBuffer* p, const int* d, size_t n) {
    if (!p) return;
    p->vb.layout.assign(d, d + n);
    p->vb.offsets.clear();
    int off = 0;
    for (size_t i = 0; i < n; i++) {
        p->vb.offsets.push_back(off);
        off += type_size(d[i]);
    }
    p->vb.elementSize = off;
}
void iftrSetVertexBufferSize(IfritVertexBuffer* p, size_t n) {
    if (p) p->vb.vertexCount = (int)n;
}
void iftrAllocateVertexBuffer(IfritVertexBuffer* p) {
    if (p) p->vb.buffer.assign((size_t)p->vb.vertexCount * p->vb.elementSize, 0);
}
void iftrSetVertexBufferValueFloat4(IfritVertexBuffer* p, int index, int attr, void* value) {
    // VertexBuffer::setValueFloat4Compatible writes only the attribute's own size
    if (!p || !value || attr < 0 || attr >= (int)p->vb.layout.size() || index < 0 || index >= p->vb.vertexCount) return;
    const size_t off = (size_t)index * p->vb.elementSize + p->vb.offsets[attr];
    std::memcpy(p->vb.buffer.data() + off, value, (size_t)type_size(p->vb.layout[attr]));
}

IfritImageF32* ifcrCreateImageFP32(size_t w, size_t h, size_t c) {
    IfritImageF32* p = new (std::nothrow) IfritImageF32();
    if (!p) return nullptr;
    p->img.width = w;
    p->img.height = h;
    p->img.channel = c;
    p->img.data.assign(w * h * c, 0.0f);
    return p;
}
void ifcrDestroyImageFP32(IfritImageF32* p) { delete p; }
float* ifcrGetImageRawDataFP32(IfritImageF32* p) { return p ? p->img.data.data() : nullptr; }

void* ifbufCreateBufferManager(void) { return new (std::nothrow) BufferManager(); }
void ifbufDestroyBufferManager(void* p) {
    BufferManager* m = static_cast<BufferManager*>(p);
    if (!m) return;
    for (Buffer* b : m->buffers) b->manager = nullptr;
    delete m;
}
void* ifbufCreateBuffer(void* pm, size_t n) {
    BufferManager* m = static_cast<BufferManager*>(pm);
    if (!m) return nullptr;
    Buffer* b = new (std::nothrow) Buffer();
    if (!b) return nullptr;
    b->manager = m;
    b->data.assign(n, 0);
    m->buffers.push_back(b);
    return b;
}
void ifbufDestroyBuffer(void* p) {
    Buffer* b = static_cast<Buffer*>(p);
    if (!b) return;
    if (b->manager) {
        auto& v = b->manager->buffers;
        for (size_t i = 0; i < v.size(); i++)
            if (v[i] == b) {
                v.erase(v.begin() + (long)i);
                break;
            }
    }
    delete b;
}
void ifbufBufferData(void* p, const void* data, size_t offset, size_t size) {
    Buffer* b = static_cast<Buffer*>(p);
    if (!b || !data || offset + size > b->data.size()) return;  // BufferManager.cpp:70-91 rejects out-of-range writes
    std::memcpy(b->data.data() + offset, data, size);
}

void* iftrCreateGpuVertexProgram(const IfritGpuVertexProgramDesc* d) {
    if (!d) return nullptr;
    GpuVertexProgram* p = new (std::nothrow) GpuVertexProgram();
    if (p) p->desc = *d;
    return p;
}
void* iftrCreateGpuFragmentProgram(const IfritGpuFragmentProgramDesc* d) {
    if (!d) return nullptr;
    GpuFragmentProgram* p = new (std::nothrow) GpuFragmentProgram();
    if (p) p->desc = *d;
    return p;
}
void* iftrCreateGpuMeshProgram(const IfritGpuMeshProgramDesc* d) {
    if (!d) return nullptr;
    GpuMeshProgram* p = new (std::nothrow) GpuMeshProgram();
    if (p) p->desc = *d;
    return p;
}
void* iftrCreateGpuTaskProgram(const IfritGpuTaskProgramDesc* d) {
    if (!d) return nullptr;
    GpuTaskProgram* p = new (std::nothrow) GpuTaskProgram();
    if (p) p->desc = *d;
    return p;
}
void* iftrCreateGpuGeometryProgram(const IfritGpuGeometryProgramDesc* d) {
    if (!d) return nullptr;
    GpuGeometryProgram* p = new (std::nothrow) GpuGeometryProgram();
    if (p) p->desc = *d;
    return p;
}
void iftrDestroyGpuProgram(void* p) {
    if (!p) return;
    const uint32_t tag = *static_cast<const uint32_t*>(p);
    if (tag == kVertexProgramTag)
        delete static_cast<GpuVertexProgram*>(p);
    else if (tag == kFragmentProgramTag)
        delete static_cast<GpuFragmentProgram*>(p);
    else if (tag == kMeshProgramTag)
        delete static_cast<GpuMeshProgram*>(p);
    else if (tag == kTaskProgramTag)
        delete static_cast<GpuTaskProgram*>(p);
    else if (tag == kGeometryProgramTag)
        delete static_cast<GpuGeometryProgram*>(p);
}
void iftrBindGeometryShader(IfritRendererHandle* h, void* p) {
    const GpuGeometryProgram* prog = static_cast<const GpuGeometryProgram*>(p);
    guarded(h, [&](TileRasterRenderer& r) {
        if (prog && prog->tag != kGeometryProgramTag)
            r.setError("iftrBindGeometryShader expects an object from iftrCreateGpuGeometryProgram");
        else
            r.bindGeometryShader(prog);
    });
}
void iftrBindMeshShader(IfritRendererHandle* h, void* p) {
    const GpuMeshProgram* prog = static_cast<const GpuMeshProgram*>(p);
    guarded(h, [&](TileRasterRenderer& r) {
        if (prog && prog->tag != kMeshProgramTag)
            r.setError("iftrBindMeshShader expects an object from iftrCreateGpuMeshProgram");
        else
            r.bindMeshShader(prog);
    });
}
void iftrBindTaskShader(IfritRendererHandle* h, void* p) {
    const GpuTaskProgram* prog = static_cast<const GpuTaskProgram*>(p);
    guarded(h, [&](TileRasterRenderer& r) {
        if (prog && prog->tag != kTaskProgramTag)
            r.setError("iftrBindTaskShader expects an object from iftrCreateGpuTaskProgram");
        else
            r.bindTaskShader(prog);
    });
}
int iftrCreateBuffer(IfritRendererHandle* h, int slot, int size) {
    int rc = -1;
    guarded(h, [&](TileRasterRenderer& r) { rc = r.createBuffer(slot, size); });
    return rc;
}
void iftrCopyHostBufferToBuffer(IfritRendererHandle* h, const void* src, int slot, int size) {
    guarded(h, [&](TileRasterRenderer& r) { r.copyHostBufferToBuffer(src, slot, size); });
}
void iftrDrawMeshTasks(IfritRendererHandle* h, int numWorkGroups, int firstWorkGroup, int clear) {
    guarded(h, [&](TileRasterRenderer& r) { r.drawMeshTasks(numWorkGroups, firstWorkGroup, clear != 0); });
}
void iftrSetClearValues(IfritRendererHandle* h, const IfritVector4f* colors, size_t n, float depth) {
    guarded(h, [&](TileRasterRenderer& r) { r.setClearValues(colors, n, depth); });
}
void iftrSetScissors(IfritRendererHandle* h, const IfritVector4f* rects, size_t n) {
    guarded(h, [&](TileRasterRenderer& r) { r.SetScissors(rects, n); });
}
void iftrSetScissorTestEnable(IfritRendererHandle* h, int enable) {
    guarded(h, [&](TileRasterRenderer& r) { r.setScissorTestEnable(enable != 0); });
}
void iftrSetCullMode(IfritRendererHandle* h, int mode) {
    guarded(h, [&](TileRasterRenderer& r) { r.SetCullMode(mode); });
}

void iftrSetSetupReciprocal(IfritRendererHandle* h, int model) {
    guarded(h, [&](TileRasterRenderer& r) { r.setSetupReciprocal(model); });
}

void iftrSetRasterizerPolygonMode(IfritRendererHandle* h, int mode) {
    guarded(h, [&](TileRasterRenderer& r) { r.setRasterizerPolygonMode(mode); });
}
void iftrSetMsaaSamples(IfritRendererHandle* h, int s) {
    guarded(h, [&](TileRasterRenderer& r) { r.setMsaaSamples(s); });
}
void iftrBindVertexDataFloat(IfritRendererHandle* h, const float* data, int stride, int n) {
    if (data) guarded(h, [&](TileRasterRenderer& r) { r.bindVertexDataFloat(data, stride, n); });
}
void iftrBindIndexData(IfritRendererHandle* h, const int32_t* data, int n) {
    if (data) guarded(h, [&](TileRasterRenderer& r) { r.bindIndexData(data, n); });
}
int iftrCreateTexture(IfritRendererHandle* h, int w, int hh, int levels) {
    int id = -1;
    guarded(h, [&](TileRasterRenderer& r) { id = r.createTexture(w, hh, levels); });
    return id;
}
int iftrCreateTextureLayers(IfritRendererHandle* h, int w, int hh, int levels, int layers) {
    int id = -1;
    guarded(h, [&](TileRasterRenderer& r) { id = r.createTexture(w, hh, levels, layers); });
    return id;
}
void iftrCopyHostBufferToImageLayer(IfritRendererHandle* h, int t, int layer, const void* p) {
    if (p) guarded(h, [&](TileRasterRenderer& r) { r.copyHostBufferToImageLayer(t, layer, p); });
}
int iftrCreateSampler(IfritRendererHandle* h, const IfritSamplerDesc* d) {
    int id = -1;
    if (d) guarded(h, [&](TileRasterRenderer& r) { id = r.CreateSampler(*d); });
    return id;
}
void iftrBlitImage(IfritRendererHandle* h, int src, int dst, const IfritImageBlit* region, int filter) {
    if (region) guarded(h, [&](TileRasterRenderer& r) { r.blitImage(src, dst, *region, filter); });
}
void iftrCopyHostBufferToImage(IfritRendererHandle* h, int t, const void* p) {
    if (p) guarded(h, [&](TileRasterRenderer& r) { r.copyHostBufferToImage(t, p); });
}
void iftrUploadTextureLevels(IfritRendererHandle* h, int t, const void* p) {
    if (p) guarded(h, [&](TileRasterRenderer& r) { r.uploadTextureLevels(t, p); });
}
void iftrGenerateFbmTexture(IfritRendererHandle* h, int t, uint32_t seed) {
    guarded(h, [&](TileRasterRenderer& r) { r.generateFbmTexture(t, seed); });
}
void iftrGenerateMipmap(IfritRendererHandle* h, int t) {
    guarded(h, [&](TileRasterRenderer& r) { r.generateMipmap(t); });
}
void iftrDownloadTexture(IfritRendererHandle* h, int t, void* p) {
    if (p) guarded(h, [&](TileRasterRenderer& r) { r.downloadTexture(t, p); });
}
void iftrSetRenderTarget(IfritRendererHandle* h, int w, int hh) {
    guarded(h, [&](TileRasterRenderer& r) { r.setRenderTarget(w, hh); });
}
void iftrGetVisibility(IfritRendererHandle* h, uint32_t* ids, float* depth) {
    guarded(h, [&](TileRasterRenderer& r) { r.getVisibility(ids, depth); });
}
void iftrGetCoverageClass(IfritRendererHandle* h, uint8_t* c) {
    if (c) guarded(h, [&](TileRasterRenderer& r) { r.getCoverageClass(c); });
}
void iftrGetColorRGBA8(IfritRendererHandle* h, uint8_t* c) {
    if (c) guarded(h, [&](TileRasterRenderer& r) { r.getColorRGBA8(c); });
}
void iftrGetFrameStats(IfritRendererHandle* h, IfritFrameStats* out) {
    if (out) guarded(h, [&](TileRasterRenderer& r) { r.getFrameStats(out); });
}
void iftrSetStageTiming(IfritRendererHandle* h, int on) {
    guarded(h, [&](TileRasterRenderer& r) { r.setStageTiming(on != 0); });
}
void iftrDrawAsync(IfritRendererHandle* h, int n, int clear) {
    guarded(h, [&](TileRasterRenderer& r) { r.drawAsync(n, clear != 0); });
}
void iftrSynchronize(IfritRendererHandle* h) {
    guarded(h, [](TileRasterRenderer& r) { r.synchronize(); });
}
const char* iftrGetLastError(IfritRendererHandle* h) { return h ? h->renderer.lastError() : "null renderer handle"; }

void iftrNeuralTrain(IfritRendererHandle* h, const IfritNeuralTrainDesc* d, IfritNeuralTrainReport* report) {
    if (!d) return;
    guarded(h, [&](TileRasterRenderer& r) { r.trainNeural(*d, report); });
}
void iftrNeuralSetParams(IfritRendererHandle* h, const float* params) {
    if (!params) return;
    guarded(h, [&](TileRasterRenderer& r) { r.setNeuralParams(params); });
}
void iftrNeuralGetParams(IfritRendererHandle* h, float* params) {
    if (!params) return;
    guarded(h, [&](TileRasterRenderer& r) { r.getNeuralParams(params); });
}
float iftrNeuralValidate(IfritRendererHandle* h, int precision, float minPsnrDb, int numIndices) {
    float r = 0.0f;
    guarded(h, [&](TileRasterRenderer& rr) { r = rr.validateNeural(precision, minPsnrDb, numIndices); });
    return r;
}
void iftrNeuralValidateEx(IfritRendererHandle* h, int precision, float minPsnrDb, int maxErrLsb, int numIndices,
                          IfritNeuralValidation* out) {
    guarded(h, [&](TileRasterRenderer& r) { r.validateNeuralEx(precision, minPsnrDb, maxErrLsb, numIndices, out); });
}
void iftrNeuralEval(IfritRendererHandle* h, const float* in, float* out, int n, int precision) {
    if (!in || !out) return;
    guarded(h, [&](TileRasterRenderer& r) { r.evalNeural(in, out, n, precision); });
}

void iftrSetColorFloatOutput(IfritRendererHandle* h, int enable) {
    guarded(h, [&](TileRasterRenderer& r) { r.setColorFloatOutput(enable != 0); });
}
void iftrGetColorRGBA32F(IfritRendererHandle* h, float* rgba) {
    if (rgba) guarded(h, [&](TileRasterRenderer& r) { r.getColorRGBA32F(rgba); });
}
void iftrSetViews(IfritRendererHandle* h, const IfritViewDesc* views, int numViews) {
    guarded(h, [&](TileRasterRenderer& r) { r.setViews(views, numViews); });
}
void iftrGetViewVisibility(IfritRendererHandle* h, int view, uint32_t* ids, float* depth) {
    guarded(h, [&](TileRasterRenderer& r) { r.getViewVisibility(view, ids, depth); });
}
void iftrGetViewColorRGBA8(IfritRendererHandle* h, int view, uint8_t* rgba) {
    if (rgba) guarded(h, [&](TileRasterRenderer& r) { r.getViewColorRGBA8(view, rgba); });
}
void iftrSetGraphMode(IfritRendererHandle* h, int enable) {
    guarded(h, [&](TileRasterRenderer& r) { r.setGraphMode(enable != 0); });
}
void iftrSetTuning(IfritRendererHandle* h, const IfritTuning* t) {
    if (t) guarded(h, [&](TileRasterRenderer& r) { r.setTuning(*t); });
}
void iftrGetTuning(IfritRendererHandle* h, IfritTuning* t) {
    if (t) guarded(h, [&](TileRasterRenderer& r) { r.getTuning(t); });
}
void iftrAutotune(IfritRendererHandle* h, int numIndices, int frames, IfritAutotuneReport* report) {
    guarded(h, [&](TileRasterRenderer& r) { r.autotune(numIndices, frames, report); });
}
void iftrDrawNoAlloc(IfritRendererHandle* h, const IfritNoAllocDraw* draw) {
    if (!draw) return;
    guarded(h, [&](TileRasterRenderer& r) { r.drawNoAlloc(*draw); });
}
uint64_t iftrGetDigest(IfritRendererHandle* h) {
    uint64_t d = 0;
    guarded(h, [&](TileRasterRenderer& r) { d = r.digest(); });
    return d;
}

}  // extern "C"
