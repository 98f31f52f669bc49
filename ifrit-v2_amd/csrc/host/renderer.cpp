// TileRasterRenderer for MI355X: buffer management and frame orchestration.
// The stage order follows the reference's draw call
// (projects/softgraphics/src/engine/tileraster/TileRasterRenderer.cpp:436-457,
// TileRasterWorker.cpp:108-135): vertex processing, geometry processing and
// binning, then per-tile rasterization, depth test and shading, with each
// stage a kernel on the renderer's stream instead of a thread barrier.
#include "renderer.hpp"

#include "../common/x86_fp.h"

#include <cpuid.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ifrit_hip {

namespace {
constexpr int kTile = 16;
constexpr int kMaxViews = 65535;  // view indices are 16-bit

// Sample offsets from the pixel corner: 1x samples the corner itself, like
// the CPU path; 4x and 8x are the standard patterns of
// projects/softgraphics/src/engine/tilerastercuda/TileRasterCoreInvocationCuda.cu:210-216.
const float kS1[2] = {0.0f, 0.0f};
const float kS4[8] = {0.375f, 0.125f, 0.875f, 0.375f, 0.125f, 0.625f, 0.625f, 0.875f};
const float kS8[16] = {0.5625f, 0.3125f, 0.4375f, 0.6875f, 0.8125f, 0.5625f, 0.3125f, 0.1875f,
                       0.1875f, 0.8125f, 0.0625f, 0.4375f, 0.6875f, 0.9375f, 0.9375f, 0.0625f};
}  // namespace

unsigned long long DevBuf::generation = 0;

bool DevBuf::ensure(size_t n) {
    if (n <= bytes && ptr) return true;
    release();
    generation++;
    size_t cap = std::max<size_t>(n, 256);
    if (hipMalloc(&ptr, cap) != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return false;
    }
    bytes = cap;
    return true;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

TileRasterRenderer::TileRasterRenderer() {}

TileRasterRenderer::~TileRasterRenderer() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    dropGraph();
    for (DevBuf* b : {&vtx_, &idx_, &clip_, &recs_[0], &recs_[1], &recs_[2], &recs_[3], &recs_[4], &recs_[5], &bary_, &binQueue_, &slotOf_, &binCount_, &binStart_, &binCursor_, &binList_,
                      &bigList_, &ctr_, &clipList_, &statPart_, &items_, &tilePartial_, &multiList_, &partials_, &visId_, &visDepth_, &visSlot_, &visClass_, &rgba8_, &rgba32f_, &texDesc_, &nnParams_, &nnGrads_, &nnM_, &nnV_,
                      &nnPacked_, &nnQinfo_, &lat32_, &latGrad_, &latM_, &latV_, &lat16_, &nnData_, &nnFeat_, &nnCount_,
                      &nnLoss_, &nnEvalIn_, &nnEvalOut_, &nnGradPart_, &viewDev_, &tileView_, &shadeView_,
                      &groupBounds_, &groupOut_, &swapped_, &scissorDev_,
                      &viewCtr_, &digest_, &planSums_, &polyKeys_, &meshVtx_, &meshIdx_})
        b->release();
    for (DevBuf& b : bufSlots_) b.release();
    for (auto& t : textures_)
        if (t.dev) (void)hipFree(t.dev);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void TileRasterRenderer::setError(const std::string& e) {
    if (error_.empty()) error_ = e;
}

bool TileRasterRenderer::ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    setError(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

const char* TileRasterRenderer::lastError() {
    if (error_.empty()) return nullptr;
    errorOut_ = error_;
    error_.clear();
    return errorOut_.c_str();
}

void TileRasterRenderer::Init() {
    if (initialized_) return;
    int n = 0;
    if (!ok(hipGetDeviceCount(&n), "hipGetDeviceCount")) return;
    if (n <= 0) {
        setError("no HIP device available");
        return;
    }
    if (!ok(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate")) return;
    for (auto& e : ev_)
        if (!ok(hipEventCreate(&e), "hipEventCreate")) return;
    if (!ctr_.ensure(sizeof(DeviceCounters))) setError("hipMalloc counters");
    std::string err;
    if (!noalloc_register_device(&err)) setError(err);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
        numCUs_ = cus;
    initialized_ = true;
}

void TileRasterRenderer::bindFrameBuffer(FrameBuffer& fb) {
    fb_ = &fb;
    if (!fb.color.empty() && fb.color[0]) {
        width_ = (int)fb.color[0]->width;
        height_ = (int)fb.color[0]->height;
    } else if (fb.depth) {
        width_ = (int)fb.depth->width;
        height_ = (int)fb.depth->height;
    }
}

void TileRasterRenderer::setRenderTarget(int w, int h) {
    fb_ = nullptr;
    width_ = w;
    height_ = h;
}

void TileRasterRenderer::bindVertexBuffer(const VertexBuffer& vb) {
    hostVb_ = &vb;
    vtxDirty_ = true;
}

void TileRasterRenderer::bindIndexBuffer(const Buffer& ib) {
    hostIb_ = &ib;
    idxDirty_ = true;
}

void TileRasterRenderer::bindVertexDataFloat(const float* data, int stride, int n) {
    hostVb_ = nullptr;
    Init();
    if (!initialized_) return;
    if (!vtx_.ensure((size_t)n * stride * sizeof(float))) {
        setError("hipMalloc vertex buffer");
        return;
    }
    if (!ok(hipMemcpy(vtx_.ptr, data, (size_t)n * stride * sizeof(float), hipMemcpyHostToDevice), "vertex upload"))
        return;
    nverts_ = n;
    stride_ = stride;
    vtxDirty_ = false;
    boundsDirty_ = true;
}

void TileRasterRenderer::bindIndexData(const int32_t* data, int n) {
    hostIb_ = nullptr;
    Init();
    if (!initialized_) return;
    if (!idx_.ensure((size_t)n * sizeof(int32_t))) {
        setError("hipMalloc index buffer");
        return;
    }
    if (!ok(hipMemcpy(idx_.ptr, data, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice), "index upload")) return;
    nidx_ = n;
    idxDirty_ = false;
    boundsDirty_ = true;
}

void TileRasterRenderer::bindVertexShaderFunc(IfritVertexShaderFunctionalPtr f, VaryingDescriptor* vd) {
    hostVs_ = f;
    hostVsVary_ = vd;
    vsBound_ = false;
    vtxDirty_ = true;
}

void TileRasterRenderer::bindFragmentShaderFunc(IfritFragmentShaderFunctionalPtr f) {
    hostFs_ = f != nullptr;
    fsBound_ = false;
    if (hostFs_)
        setError("host fragment shader functions cannot run on the GPU backend; bind a GPU program from "
                 "iftrCreateGpuFragmentProgram with iftrBindFragmentShader");
}

void TileRasterRenderer::bindVertexShader(const GpuVertexProgram& p) {
    vs_ = p;
    vsBound_ = true;
    hostVs_ = nullptr;
}

void TileRasterRenderer::bindFragmentShader(const GpuFragmentProgram& p) {
    fs_ = p;
    fsBound_ = true;
    hostFs_ = false;
}

namespace {
// A blend factor as the reference encodes it (TileRasterRenderer.cpp:212-333):
// (c0 + c1 (1 - src alpha) + c2 (1 - dst alpha)) (1 - c3).  The encoding makes
// SRC_ALPHA 1 - src alpha and ONE_MINUS_SRC_ALPHA src alpha (likewise DST);
// it is kept so that results match the reference.  False for the factors the
// reference rejects.
bool encode_blend_factor(int32_t f, float* c) {
    switch (f) {
        case 0: c[0] = 0; c[1] = 0; c[2] = 0; c[3] = 1; return true;   // ZERO
        case 1: c[0] = 1; c[1] = 0; c[2] = 0; c[3] = 0; return true;   // ONE
        case 6: c[0] = 0; c[1] = 1; c[2] = 0; c[3] = 0; return true;   // SRC_ALPHA
        case 7: c[0] = 1; c[1] = -1; c[2] = 0; c[3] = 0; return true;  // ONE_MINUS_SRC_ALPHA
        case 8: c[0] = 0; c[1] = 0; c[2] = 1; c[3] = 0; return true;   // DST_ALPHA
        case 9: c[0] = 1; c[1] = 0; c[2] = -1; c[3] = 0; return true;  // ONE_MINUS_DST_ALPHA
        default: return false;
    }
}
}  // namespace

void TileRasterRenderer::setBlendFunc(const IfritColorAttachmentBlendState& s) {
    BlendCoefs b{};
    if (s.blendEnable && (!encode_blend_factor(s.srcColorBlendFactor, b.sc) ||
                          !encode_blend_factor(s.dstColorBlendFactor, b.dc) ||
                          !encode_blend_factor(s.srcAlphaBlendFactor, b.sa) ||
                          !encode_blend_factor(s.dstAlphaBlendFactor, b.da))) {
        // the reference stops with "Unsupported blend factor" (TileRasterRenderer.cpp:241)
        setError("unsupported blend factor: ZERO, ONE, SRC_ALPHA, ONE_MINUS_SRC_ALPHA, DST_ALPHA and "
                 "ONE_MINUS_DST_ALPHA are supported");
        return;
    }
    blend_ = s;
    blendCoefs_ = b;
}

void TileRasterRenderer::SetDepthFunc(int f) {
    depthFuncSaved_ = f;
    if (depthTest_) depthFunc_ = f;
}

void TileRasterRenderer::optSetDepthTestEnable(bool opt) {
    depthTest_ = opt;
    depthFunc_ = opt ? depthFuncSaved_ : IF_COMPARE_OP_ALWAYS;
}

void TileRasterRenderer::setClearValues(const IfritVector4f* colors, size_t n, float depth) {
    if (std::isnan(depth)) return setError("the clear depth must not be NaN");
    // one colour attachment: the first clear colour applies, the others are kept by the caller
    float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (colors && n > 0) {
        c[0] = colors[0].x;
        c[1] = colors[0].y;
        c[2] = colors[0].z;
        c[3] = colors[0].w;
    }
    std::memcpy(clearColor_, c, sizeof(c));
    clearDepth_ = depth;
}

void TileRasterRenderer::SetScissors(const IfritVector4f* rects, size_t n) {
    if (n > (size_t)kMaxScissors) return setError("at most 8 scissor rectangles");
    if (n && !rects) return setError("null scissor rectangles");
    Init();
    if (!initialized_) return;
    // the kernels read the rectangles from device memory; a draw in flight may still read them
    if (!scissorDev_.ensure(kMaxScissors * sizeof(float4))) return setError("hipMalloc scissors");
    if (!ok(hipStreamSynchronize(stream_), "scissors") ||
        (n && !ok(hipMemcpy(scissorDev_.ptr, rects, n * sizeof(float4), hipMemcpyHostToDevice), "scissors")))
        return;
    scissors_.assign(rects, rects + n);
}

Scissor TileRasterRenderer::activeScissor() const {
    // the test off, or no rectangle, passes every pixel (TileRasterCoreInvocationCuda.cu:3688-3693)
    Scissor sc{};
    sc.count = scissorTest_ && scissorDev_.ptr ? (int)scissors_.size() : 0;
    sc.rects = scissorDev_.as<float4>();
    return sc;
}

void TileRasterRenderer::SetCullMode(int mode) {
    if (mode < kCullNone || mode > kCullFrontAndBack) return setError("unknown cull mode");
    cullMode_ = mode;
}

void TileRasterRenderer::setSetupReciprocal(int model) {
    if (model == IFHIP_RCP_HOST) {
        unsigned a = 0, b = 0, c = 0, d = 0;
        char vendor[13] = {0};
        if (__get_cpuid(0, &a, &b, &c, &d)) {
            std::memcpy(vendor, &b, 4);
            std::memcpy(vendor + 4, &d, 4);
            std::memcpy(vendor + 8, &c, 4);
        }
        if (std::strcmp(vendor, "GenuineIntel") == 0) model = IFHIP_RCP_INTEL;
        else if (std::strcmp(vendor, "AuthenticAMD") == 0) model = IFHIP_RCP_AMD;
        else return setError(std::string("no RCPPS table for this CPU vendor: ") + vendor);
    }
    if (model != IFHIP_RCP_INTEL && model != IFHIP_RCP_AMD) return setError("unknown reciprocal model");
    rcpModel_ = model == IFHIP_RCP_AMD ? ifx::kRcpAmd : ifx::kRcpIntel;
}

int TileRasterRenderer::createBuffer(int slot, int size) {
    Init();
    if (!initialized_) return -1;
    if (slot < 0 || slot >= kBufferSlots) return setError("buffer slot out of range (0 .. 63)"), -1;
    if (size <= 0) return setError("buffer size must be positive"), -1;
    if (!bufSlots_[slot].ensure((size_t)size)) return setError("hipMalloc buffer slot"), -1;
    bufSize_[slot] = size;
    return 0;
}

void TileRasterRenderer::copyHostBufferToBuffer(const void* src, int slot, int size) {
    if (slot < 0 || slot >= kBufferSlots || !bufSlots_[slot].ptr) return setError("copy into a buffer slot not created");
    if (!src || size < 0 || size > bufSize_[slot]) return setError("copy larger than the buffer slot");
    if (stream_) (void)hipStreamSynchronize(stream_);
    ok(hipMemcpy(bufSlots_[slot].ptr, src, (size_t)size, hipMemcpyHostToDevice), "buffer upload");
}

void TileRasterRenderer::bindMeshShader(const GpuMeshProgram* p) {
    msBound_ = p != nullptr;
    if (p) ms_ = *p;
}

void TileRasterRenderer::bindTaskShader(const GpuTaskProgram* p) {
    tsBound_ = p != nullptr;
    if (p) ts_ = *p;
}

// The meshlet program emits its work groups' meshlets into meshVtx_ /
// meshIdx_ (clip position then record per vertex, rebased indices), which are
// then drawn in place of the bound vertex and index buffers by the clip
// pass-through vertex program; the bound inputs are restored afterwards.
void TileRasterRenderer::drawMeshTasks(int numGroups, int firstGroup, bool clear) {
    Init();
    if (!initialized_) return;
    if (!msBound_) return setError("DrawMeshTasks: no mesh program bound");
    const IfritGpuMeshProgramDesc& m = ms_.desc;
    if (m.kind != IFHIP_MS_MESHLET) return setError("DrawMeshTasks: unknown mesh program");
    if (tsBound_ && ts_.desc.kind != IFHIP_TS_ONE_MESH_GROUP) return setError("DrawMeshTasks: unknown task program");
    if (numGroups < 0 || firstGroup < 0) return setError("DrawMeshTasks: negative work group count or start");
    // the meshlet program transforms by its own matrix: there is no per-view transform to apply
    if (!batch_.empty()) return setError("DrawMeshTasks draws a single view");
    if (m.recordFloats < 3 || m.recordFloats > 256) return setError("DrawMeshTasks: recordFloats must be 3 .. 256");
    const int slots[4] = {m.vertexSlot, m.indexSlot, m.vertexOffsetSlot, m.indexOffsetSlot};
    for (int sl : slots)
        if (sl < 0 || sl >= kBufferSlots || !bufSlots_[sl].ptr)
            return setError("DrawMeshTasks: a buffer slot of the mesh program was not created");
    const long long lastOff = (long long)firstGroup + numGroups;
    if ((lastOff + 1) * 4 > bufSize_[m.vertexOffsetSlot] || (lastOff + 1) * 4 > bufSize_[m.indexOffsetSlot])
        return setError("DrawMeshTasks: work groups beyond the offset buffers");
    if (stream_) (void)hipStreamSynchronize(stream_);
    int vo[2] = {0, 0}, io[2] = {0, 0};
    const int* vob = bufSlots_[m.vertexOffsetSlot].as<int>();
    const int* iob = bufSlots_[m.indexOffsetSlot].as<int>();
    if (!ok(hipMemcpy(&vo[0], vob + firstGroup, 4, hipMemcpyDeviceToHost), "mesh offsets") ||
        !ok(hipMemcpy(&vo[1], vob + lastOff, 4, hipMemcpyDeviceToHost), "mesh offsets") ||
        !ok(hipMemcpy(&io[0], iob + firstGroup, 4, hipMemcpyDeviceToHost), "mesh offsets") ||
        !ok(hipMemcpy(&io[1], iob + lastOff, 4, hipMemcpyDeviceToHost), "mesh offsets"))
        return;
    const long long nv = (long long)vo[1] - vo[0], ni = (long long)io[1] - io[0];
    if (vo[0] < 0 || io[0] < 0 || nv < 0 || ni < 0 || ni % 3)
        return setError("DrawMeshTasks: offsets must grow, and the index count must be a multiple of 3");
    if ((long long)vo[1] * m.recordFloats * 4 > bufSize_[m.vertexSlot] || (long long)io[1] * 4 > bufSize_[m.indexSlot])
        return setError("DrawMeshTasks: meshlets beyond the vertex or index buffer");
    // the offsets of every work group in between must grow too (checked on the device per meshlet is
    // not needed: a meshlet's bounds come from consecutive offsets, so a shrinking pair emits nothing)
    const int stride = 4 + m.recordFloats;
    if (nv * stride >= (1ll << 31) || ni >= (1ll << 31)) return setError("DrawMeshTasks: meshlets too large");
    if (!meshVtx_.ensure((size_t)std::max<long long>(nv, 1) * stride * 4) ||
        !meshIdx_.ensure((size_t)std::max<long long>(ni, 1) * 4))
        return setError("hipMalloc mesh outputs");
    MeshParams mp{};
    std::memcpy(mp.mvp, m.mvp, sizeof(mp.mvp));
    mp.recordFloats = m.recordFloats;
    mp.first = firstGroup;
    mp.groups = numGroups;
    mp.vertexBase = vo[0];
    mp.indexBase = io[0];
    if (!ctr_.ensure(sizeof(DeviceCounters))) return setError("hipMalloc counters");
    if (!ok(hipMemsetAsync(ctr_.ptr, 0, sizeof(unsigned), stream_), "mesh status")) return;
    launch_mesh_expand(mp, bufSlots_[m.vertexSlot].as<float>(), bufSlots_[m.indexSlot].as<int>(), vob, iob,
                       meshVtx_.as<float>(), meshIdx_.as<int>(), ctr_.as<DeviceCounters>(), stream_);
    if (!ok(hipGetLastError(), "mesh launch") || !ok(hipStreamSynchronize(stream_), "mesh work groups")) return;
    unsigned status = 0;
    if (!ok(hipMemcpy(&status, ctr_.ptr, sizeof(status), hipMemcpyDeviceToHost), "mesh status")) return;
    if (status & kStatusBadIndex) return setError("DrawMeshTasks: a meshlet index lies outside its meshlet");
    drawEmitted((int)nv, stride, m.recordFloats, (int)ni, clear, false);
}

void TileRasterRenderer::drawEmitted(int nv, int stride, int clipAt, int ni, bool clear, bool async) {
    std::swap(vtx_, meshVtx_);
    std::swap(idx_, meshIdx_);
    const GpuVertexProgram vsSaved = vs_;
    const bool vsBoundSaved = vsBound_, vtxDirtySaved = vtxDirty_, idxDirtySaved = idxDirty_;
    const bool boundsDirtySaved = boundsDirty_;
    const int nvertsSaved = nverts_, strideSaved = stride_, nidxSaved = nidx_, boundsTrisSaved = boundsTris_;
    auto hostVsSaved = hostVs_;
    auto hostVbSaved = hostVb_;
    auto hostIbSaved = hostIb_;
    nverts_ = nv;
    stride_ = stride;
    nidx_ = ni;
    vs_.desc.kind = IFHIP_VS_CLIP_PASSTHROUGH;
    vsBound_ = true;
    clipAt_ = clipAt;
    hostVs_ = nullptr;
    hostVb_ = nullptr;
    hostIb_ = nullptr;
    vtxDirty_ = idxDirty_ = false;
    if (async)
        drawAsync(ni, clear);
    else
        drawElements(ni, clear);
    std::swap(vtx_, meshVtx_);
    std::swap(idx_, meshIdx_);
    clipAt_ = -1;
    vs_ = vsSaved;
    vsBound_ = vsBoundSaved;
    vtxDirty_ = vtxDirtySaved;
    idxDirty_ = idxDirtySaved;
    boundsDirty_ = boundsDirtySaved;
    nverts_ = nvertsSaved;
    stride_ = strideSaved;
    nidx_ = nidxSaved;
    boundsTris_ = boundsTrisSaved;
    hostVs_ = hostVsSaved;
    hostVb_ = hostVbSaved;
    hostIb_ = hostIbSaved;
}

void TileRasterRenderer::bindGeometryShader(const GpuGeometryProgram* p) {
    gsBound_ = p != nullptr;
    if (p) gs_ = *p;
}

// The geometry stage (TileRasterCoreInvocationCuda.cu:500-560) with the
// built-in clip-offset program: every triangle of the draw is emitted once,
// as three unshared vertices in triangle order (the order the reference's
// deterministic sort keys give, :539-542), then drawn by the ordinary
// pipeline; the primitive id is the triangle's index.
void TileRasterRenderer::drawGeometry(int indexCount, bool clear, bool async) {
    Init();
    if (!initialized_) return;
    if (gs_.desc.kind != IFHIP_GS_CLIP_OFFSET) return setError("unknown geometry program");
    if (hostVs_) return setError("geometry programs follow a GPU vertex program, not a host vertex function");
    if (!batch_.empty()) return setError("geometry programs draw a single view");
    if (!vsBound_) return setError("no vertex program bound");
    if (!uploadInputs()) return;
    if (nverts_ <= 0 || !vtx_.ptr) return setError("no vertex buffer bound");
    if (indexCount < 0 || indexCount > nidx_) return setError("index count exceeds the bound index buffer");
    const bool pass = vs_.desc.kind == IFHIP_VS_CLIP_PASSTHROUGH;
    if (stride_ < (pass ? 4 : 3)) return setError("vertex records too short for the vertex program");
    const long long ntri = indexCount / 3, stride = stride_ + 4;
    if (3 * ntri * stride >= (1ll << 31)) return setError("geometry program output too large");
    if (stream_) (void)hipStreamSynchronize(stream_);  // an asynchronous draw may still read the last list
    if (!meshVtx_.ensure((size_t)std::max<long long>(3 * ntri, 1) * stride * 4) ||
        !meshIdx_.ensure((size_t)std::max<long long>(3 * ntri, 1) * 4))
        return setError("hipMalloc geometry outputs");
    GeometryParams gp{};
    std::memcpy(gp.mvp, vs_.desc.mvp, sizeof(gp.mvp));
    std::memcpy(gp.offset, gs_.desc.offset, sizeof(gp.offset));
    gp.passthrough = pass ? 1 : 0;
    gp.stride = stride_;
    gp.nverts = nverts_;
    gp.ntri = (int)ntri;
    if (!ctr_.ensure(sizeof(DeviceCounters))) return setError("hipMalloc counters");
    if (!ok(hipMemsetAsync(ctr_.ptr, 0, sizeof(unsigned), stream_), "geometry status")) return;
    launch_geometry_expand(gp, vtx_.as<float>(), idx_.as<int>(), meshVtx_.as<float>(), meshIdx_.as<int>(),
                           ctr_.as<DeviceCounters>(), stream_);
    if (!ok(hipGetLastError(), "geometry launch") || !ok(hipStreamSynchronize(stream_), "geometry program")) return;
    unsigned status = 0;
    if (!ok(hipMemcpy(&status, ctr_.ptr, sizeof(status), hipMemcpyDeviceToHost), "geometry status")) return;
    if (status & kStatusBadIndex) return setError("an index lies outside the bound vertex buffer");
    drawEmitted((int)(3 * ntri), (int)stride, stride_, (int)(3 * ntri), clear, async);
}

void TileRasterRenderer::setRasterizerPolygonMode(int mode) {
    if (mode < 0 || mode > 2) return setError("unknown polygon mode");
    polygonMode_ = mode;
}

void TileRasterRenderer::setMsaaSamples(int s) {
    if (s != 1 && s != 4 && s != 8) {
        setError("MSAA sample count must be 1, 4 or 8");
        return;
    }
    samples_ = s;
}

int TileRasterRenderer::createTexture(int w, int h, int levels, int layers) {
    Init();
    if ((w & (w - 1)) || (h & (h - 1)) || w <= 0 || h <= 0 || levels <= 0 || levels > 16) {
        setError("textures must have power-of-two sizes and 1..16 levels");
        return -1;
    }
    if (layers < 1 || layers > 64) {
        setError("textures have 1..64 layers (a cube map has 6)");
        return -1;
    }
    if ((long long)w * h * layers >= (1ll << 28)) {
        // the sampler addresses a texture's texels (all levels and layers) with 32-bit offsets
        setError("textures must have fewer than 2^28 texels in level 0 over all layers");
        return -1;
    }
    TextureSlot t;
    t.width = w;
    t.height = h;
    t.levels = levels;
    t.layers = layers;
    size_t off = 0;
    for (int l = 0; l < levels; l++) {
        t.levelOffset.push_back(off);
        off += t.levelTexels(l) * (size_t)layers;
    }
    t.bytes = off * 4;
    if (!ok(hipMalloc(&t.dev, t.bytes), "hipMalloc texture")) return -1;
    textures_.push_back(t);
    texDescDirty_ = true;
    return (int)textures_.size() - 1;
}

void TileRasterRenderer::copyHostBufferToImage(int tex, const void* rgba8) { copyHostBufferToImageLayer(tex, 0, rgba8); }

// copyHostBufferToImage (TileRasterImageOpInvocationsCuda.cu:157-181): level 0 of one layer.
void TileRasterRenderer::copyHostBufferToImageLayer(int tex, int layer, const void* rgba8) {
    if (tex < 0 || tex >= (int)textures_.size()) return setError("bad texture id");
    const TextureSlot& t = textures_[tex];
    if (layer < 0 || layer >= t.layers) return setError("bad texture layer");
    const size_t n = t.levelTexels(0) * 4;
    ok(hipMemcpy(t.dev + n * (size_t)layer, rgba8, n, hipMemcpyHostToDevice), "texture upload");
}

void TileRasterRenderer::uploadTextureLevels(int tex, const void* rgba8) {
    if (tex < 0 || tex >= (int)textures_.size()) return setError("bad texture id");
    const TextureSlot& t = textures_[tex];
    ok(hipMemcpy(t.dev, rgba8, t.bytes, hipMemcpyHostToDevice), "texture upload");
}

void TileRasterRenderer::downloadTexture(int tex, void* rgba8) {
    if (tex < 0 || tex >= (int)textures_.size()) return setError("bad texture id");
    const TextureSlot& t = textures_[tex];
    ok(hipMemcpy(rgba8, t.dev, t.bytes, hipMemcpyDeviceToHost), "texture download");
}

void TileRasterRenderer::generateFbmTexture(int tex, uint32_t seed) {
    if (tex < 0 || tex >= (int)textures_.size()) return setError("bad texture id");
    const TextureSlot& t = textures_[tex];
    if (t.width != t.height) return setError("fBm textures are square");
    launch_fbm_texture(t.dev, t.width, seed, stream_);
    ok(hipGetLastError(), "fbm launch");
}

void TileRasterRenderer::generateMipmap(int tex) {
    if (tex < 0 || tex >= (int)textures_.size()) return setError("bad texture id");
    const TextureSlot& t = textures_[tex];
    for (int l = 1; l < t.levels; l++) {
        const int sw = std::max(1, t.width >> (l - 1)), sh = std::max(1, t.height >> (l - 1));
        const int dw = std::max(1, t.width >> l), dh = std::max(1, t.height >> l);
        for (int f = 0; f < t.layers; f++)
            launch_mip_level(t.dev + (t.levelOffset[l - 1] + t.levelTexels(l - 1) * f) * 4, sw, sh,
                             t.dev + (t.levelOffset[l] + t.levelTexels(l) * f) * 4, dw, dh, stream_);
    }
    ok(hipGetLastError(), "mip launch");
    ok(hipStreamSynchronize(stream_), "mip sync");
}

// blitImage (TileRasterRendererCuda.h:95; TileRasterImageOpInvocationsCuda.cu:87-131) between levels and
// layers of RGBA8 textures.  The source layer is the source subresource's (the CUDA code reads the
// destination's baseArrayLayer for both, TileRasterImageOpInvocationsCuda.cu:117).
void TileRasterRenderer::blitImage(int src, int dst, const IfritImageBlit& r, int filter) {
    if (src < 0 || src >= (int)textures_.size() || dst < 0 || dst >= (int)textures_.size())
        return setError("bad texture id");
    if (filter != 0 && filter != 1) return setError("blit filter is IF_FILTER_NEAREST (0) or IF_FILTER_LINEAR (1)");
    const TextureSlot& S = textures_[src];
    const TextureSlot& D = textures_[dst];
    const int sl = (int)r.srcSubresource.mipLevel, dl = (int)r.dstSubresource.mipLevel;
    if (sl >= S.levels || dl >= D.levels || (int)r.srcSubresource.baseArrayLayer >= S.layers ||
        (int)r.dstSubresource.baseArrayLayer >= D.layers)
        return setError("blit level or layer out of range");
    BlitArgs a{};
    a.sw = std::max(1, S.width >> sl);
    a.sh = std::max(1, S.height >> sl);
    a.dw = std::max(1, D.width >> dl);
    a.dh = std::max(1, D.height >> dl);
    a.sx0 = r.srcExtentSt.width;
    a.sy0 = r.srcExtentSt.height;
    a.sx1 = r.srcExtentEd.width;
    a.sy1 = r.srcExtentEd.height;
    a.dx0 = r.dstExtentSt.width;
    a.dy0 = r.dstExtentSt.height;
    a.dx1 = r.dstExtentEd.width;
    a.dy1 = r.dstExtentEd.height;
    a.linear = filter;
    if (a.sx1 > (unsigned)a.sw || a.sy1 > (unsigned)a.sh || a.dx1 > (unsigned)a.dw || a.dy1 > (unsigned)a.dh ||
        a.sx0 >= a.sx1 || a.sy0 >= a.sy1 || a.dx0 >= a.dx1 || a.dy0 >= a.dy1)
        return setError("blit regions must be non-empty and inside their levels");
    if (src == dst && sl == dl && r.srcSubresource.baseArrayLayer == r.dstSubresource.baseArrayLayer)
        return setError("blit source and destination must be different levels or layers");
    const uint8_t* sp = S.dev + (S.levelOffset[sl] + S.levelTexels(sl) * r.srcSubresource.baseArrayLayer) * 4;
    uint8_t* dp = D.dev + (D.levelOffset[dl] + D.levelTexels(dl) * r.dstSubresource.baseArrayLayer) * 4;
    launch_blit(sp, dp, a, stream_);
    ok(hipGetLastError(), "blit launch");
    ok(hipStreamSynchronize(stream_), "blit");
}

// CreateSampler (TileRasterRendererCuda.h:92): the sampler objects a fragment program selects by id.
int TileRasterRenderer::CreateSampler(const IfritSamplerDesc& d) {
    auto mode_ok = [](int m) { return m >= kAddrRepeat && m <= kAddrMirrorClampToEdge; };
    if ((d.filterMode != 0 && d.filterMode != 1) || !mode_ok(d.addressModeU) || !mode_ok(d.addressModeV) ||
        !mode_ok(d.addressModeW) || (d.borderColor != 0 && d.borderColor != 1)) {
        setError("sampler: filter NEAREST/LINEAR, address modes 0..4, border colour BLACK/WHITE");
        return -1;
    }
    SamplerObject o;
    o.mode.addrU = d.addressModeU;
    o.mode.addrV = d.addressModeV;
    o.mode.nearest = d.filterMode == 0 ? 1 : 0;
    o.mode.border = d.borderColor == 1 ? 0xffffffffu : 0xff000000u;
    o.anisotropy = d.anisotropyEnable != 0;
    o.maxAnisotropy = d.maxAnisotropy;
    samplers_.push_back(o);
    return (int)samplers_.size();  // ids start at 1: a zero-initialised program description names none
}

void TileRasterRenderer::runHostVertexShader() {
    // The user's vertex function runs on the host (it is host code); its clip
    // positions and varyings are then uploaded and the GPU pipeline continues.
    const VertexBuffer* vb = hostVb_;
    const int nvary = hostVsVary_ ? (int)hostVsVary_->types.size() : 0;
    const int n = vb->vertexCount;
    std::vector<IfritVector4f> pos((size_t)n), vary((size_t)n * std::max(1, nvary));
    std::vector<const void*> in(vb->layout.size());
    std::vector<IfritVector4f*> out((size_t)std::max(1, nvary));
    for (int i = 0; i < n; i++) {
        for (size_t a = 0; a < vb->layout.size(); a++)
            in[a] = vb->buffer.data() + (size_t)i * vb->elementSize + vb->offsets[a];
        for (int k = 0; k < nvary; k++) out[k] = &vary[(size_t)i * nvary + k];
        hostVs_(in.data(), &pos[i], out.data());
    }
    clip_.ensure((size_t)n * sizeof(float4));
    ok(hipMemcpy(clip_.ptr, pos.data(), (size_t)n * 16, hipMemcpyHostToDevice), "clip upload");
    vtx_.ensure((size_t)n * std::max(1, nvary) * 16);
    ok(hipMemcpy(vtx_.ptr, vary.data(), (size_t)n * std::max(1, nvary) * 16, hipMemcpyHostToDevice), "varying upload");
    nverts_ = n;
    stride_ = 4 * std::max(1, nvary);
}

// Uploads host vertex and index buffers that changed (running a host vertex
// function over the vertices first, if one is bound).
bool TileRasterRenderer::uploadInputs() {
    if (hostVs_ && hostVb_) {
        runHostVertexShader();
        vtxDirty_ = false;
        boundsDirty_ = true;
    } else if (vtxDirty_ && hostVb_) {
        if (hostVb_->elementSize % 4) {
            setError("vertex records must be a multiple of 4 bytes");
            return false;
        }
        if (!vtx_.ensure(hostVb_->buffer.size())) return setError("hipMalloc vertex buffer"), false;
        if (!ok(hipMemcpy(vtx_.ptr, hostVb_->buffer.data(), hostVb_->buffer.size(), hipMemcpyHostToDevice),
                "vertex upload"))
            return false;
        nverts_ = hostVb_->vertexCount;
        stride_ = hostVb_->elementSize / 4;
        vtxDirty_ = false;
        boundsDirty_ = true;
    }
    if (idxDirty_ && hostIb_) {
        if (!idx_.ensure(hostIb_->data.size())) return setError("hipMalloc index buffer"), false;
        if (!ok(hipMemcpy(idx_.ptr, hostIb_->data.data(), hostIb_->data.size(), hipMemcpyHostToDevice),
                "index upload"))
            return false;
        nidx_ = (int)(hostIb_->data.size() / 4);
        idxDirty_ = false;
        boundsDirty_ = true;
    }
    return true;
}

bool TileRasterRenderer::prepare(int indexCount) {
    Init();
    if (!initialized_) return false;
    if (batch_.empty() && (width_ <= 0 || height_ <= 0)) {
        setError("no frame buffer bound");
        return false;
    }
    if (batch_.empty() && (width_ > 16384 || height_ > 16384)) {
        setError("frame buffer larger than 16384 x 16384");
        return false;
    }
    if (!batch_.empty() && (hostVs_ || !vsBound_)) {
        setError("multi-view batches need a GPU vertex program");
        return false;
    }
    if (depthFunc_ < IF_COMPARE_OP_NEVER || depthFunc_ > IF_COMPARE_OP_ALWAYS) {
        setError("unknown depth function");
        return false;
    }
    if (polygonMode_ != 0 && (samples_ != 1 || !batch_.empty() || blend_.blendEnable || fsWritesDepth())) {
        // the CUDA point and line pipelines draw one sample per pixel into one target (CUI:2931-3335)
        setError("point and line modes draw a single target at 1 sample per pixel, without blending or depth writes");
        return false;
    }
    const int fk = fs_.desc.kind;
    if (blend_.blendEnable && (samples_ != 1 || !batch_.empty() || !fsBound_ ||
                               (fk != IFHIP_FS_FLAT && fk != IFHIP_FS_FLAT_ALPHA && fk != IFHIP_FS_BLINN_PHONG &&
                                fk != IFHIP_FS_GGX_LAYERED && fk != IFHIP_FS_NORMAL_LAMBERT))) {
        setError("blending draws one target at 1 sample per pixel with the flat, translucent flat, Blinn-Phong or "
                 "layered-GGX program");
        return false;
    }
    if (!uploadInputs()) return false;
    if (nverts_ <= 0 || !vtx_.ptr) {
        setError("no vertex buffer bound");
        return false;
    }
    if (indexCount < 0 || indexCount > nidx_) {
        setError("index count exceeds the bound index buffer");
        return false;
    }
    if (!hostVs_ && !vsBound_) {
        setError("no vertex program bound");
        return false;
    }
    if (fsBound_ && fs_.desc.kind == IFHIP_FS_NEURAL && (!nnReady_ || textures_.empty())) {
        setError("the neural material has not been trained (iftrNeuralTrain) or its textures are not bound");
        return false;
    }
    if (fsBound_ && fs_.desc.samplerId > (int)samplers_.size()) {
        setError("the fragment program names a sampler that does not exist");
        return false;
    }
    if (fsBound_ && fs_.desc.kind != IFHIP_FS_FLAT && fs_.desc.kind != IFHIP_FS_FLAT_ALPHA &&
        fs_.desc.kind != IFHIP_FS_DEPTH_ONLY && fs_.desc.kind != IFHIP_FS_NORMAL_LAMBERT) {
        // the cube-map program reads six-layer textures, every other textured program one-layer ones
        const int want = fs_.desc.kind == IFHIP_FS_CUBE_MAP ? 6 : 1;
        for (const TextureSlot& t : textures_)
            if (t.layers != want) {
                setError(want == 6 ? "the cube-map program needs textures of six layers"
                                   : "textured programs other than the cube map need one-layer textures");
                return false;
            }
        if (fs_.desc.kind == IFHIP_FS_CUBE_MAP && textures_.empty()) {
            setError("the cube-map program needs a texture");
            return false;
        }
    }
    if (fsBound_ && (fs_.desc.kind == IFHIP_FS_GGX_LAYERED || fs_.desc.kind == IFHIP_FS_NEURAL) &&
        textures_.size() < 3) {
        // the layered kernels read the descriptors of textures 0, 1 and 2
        setError("the layered-GGX and neural programs need three textures (albedo, normal, roughness)");
        return false;
    }
    if (fsBound_ && (fs_.desc.kind == IFHIP_FS_GGX_LAYERED || fs_.desc.kind == IFHIP_FS_NEURAL) &&
        (stride_ < 11 || fs_.desc.offTangent + 3 > stride_)) {
        setError("the layered-GGX and neural programs need a tangent in the vertex record");
        return false;
    }
    if (!hostVs_ && vs_.desc.kind == IFHIP_VS_CLIP_PASSTHROUGH && stride_ < 4) {
        setError("the clip pass-through vertex program needs 4 floats per vertex");
        return false;
    }
    // object-space bounds of the triangle groups, for culling whole groups per view
    if (!hostVs_ && vs_.desc.kind == IFHIP_VS_TRANSFORM) {
        const int all = nidx_ / 3, groups = (all + kGroupTris - 1) / kGroupTris;
        if (boundsDirty_ || boundsTris_ != all) {
            if (!groupBounds_.ensure((size_t)std::max(groups, 1) * 32)) return setError("hipMalloc group bounds"), false;
            launch_group_bounds(vtx_.as<float>(), stride_, nverts_, idx_.as<int>(), all, groupBounds_.as<float4>(),
                                stream_);
            boundsDirty_ = false;
            boundsTris_ = all;
        }
    }
    const int ntri = indexCount / 3;
    curTris_ = ntri;
    curIdx_ = indexCount;
    clipCap_ = std::max<unsigned>(clipCap_, std::max(4096u, (unsigned)(ntri / 8)));
    if (!layoutViews(ntri)) return false;
    const size_t nv = viewDesc_.size();
    const size_t tiles = totalTiles_;
    const size_t px = totalPixels_;
    const size_t ps = px * samples_;
    const unsigned part = 1u << partShift_;
    if (!binCapSet_)  // default: room for six tiles per triangle and view (grown on overflow)
        binCap_ = std::max<unsigned>(
            binCap_, std::max(1u << 20, (unsigned)std::min<size_t>(0xFFFFFFF0ull, (size_t)ntri * 6 * nv)));
    bigCap_ = std::max<unsigned>(bigCap_, std::max(1024u, (unsigned)std::min<size_t>(0x7FFFFFFFull, ntri * nv / 16)));
    itemCap_ = std::max<unsigned>(itemCap_, (unsigned)tiles + binCap_ / part + 1u);
    partialCap_ = std::max<unsigned>(partialCap_, std::max(256u, binCap_ / (8u * part)));
    // point and line records take one slot per index (3 per triangle)
    const size_t slots = std::max(((size_t)ntri + clipCap_) * nv, polygonMode_ != 0 ? (size_t)indexCount : 0);
    const bool wantFloat = batch_.empty() && ((fb_ && !fb_->color.empty() && fb_->color[0]) || floatOut_ ||
                                              blend_.blendEnable);
    bool okAll = clip_.ensure((size_t)nverts_ * 16 * nv) &&
                 groupOut_.ensure(std::max<size_t>(1, (size_t)(boundsTris_ + kGroupTris - 1) / kGroupTris) * nv) &&
                 recs_[0].ensure(slots * 16) &&
                 recs_[1].ensure(slots * 16) && recs_[2].ensure(slots * 16) && recs_[3].ensure(slots * 16) &&
                 recs_[4].ensure(slots * 16) && recs_[5].ensure(slots * 16) && bary_.ensure(slots * 32) &&
                 binQueue_.ensure(slots * 16) && slotOf_.ensure((size_t)std::max(1, ntri) * 4 * nv) &&
                 (!needSwapFlags() || swapped_.ensure(slots)) &&
                 clipList_.ensure((size_t)clipCap_ * 4 * nv) &&
                 statPart_.ensure(((size_t)(ntri + 255) / 256 * nv + 2 * (size_t)rasterGrid() + 2) * 8) &&
                 items_.ensure((size_t)itemCap_ * 16) && tilePartial_.ensure(tiles * 4) &&
                 planSums_.ensure(((tiles + kPlanTiles - 1) / kPlanTiles + 1) * 6 * 4) &&
                 multiList_.ensure(tiles * 8) && partials_.ensure((size_t)partialCap_ * 256 * samples_ * 8) &&
                 binCount_.ensure(tiles * 4) && binStart_.ensure(tiles * 4) && binCursor_.ensure(tiles * 4) &&
                 binList_.ensure((size_t)binCap_ * 4) && bigList_.ensure((size_t)bigCap_ * 4) &&
                 visId_.ensure(ps * 4) && visDepth_.ensure(ps * 4) && visSlot_.ensure(ps * 4) &&
                 visClass_.ensure(px) && rgba8_.ensure(px * 4) && (!wantFloat || rgba32f_.ensure(px * 16)) &&
                 (polygonMode_ == 0 || polyKeys_.ensure(px * 8));
    if (!okAll) {
        setError("device allocation failed");
        return false;
    }
    if (texDescDirty_) {
        // the pool pointer the kernels offset is the lowest texture allocation, so that every texture's
        // base is a non-negative byte offset from a real device address
        texPool_ = nullptr;
        for (const TextureSlot& t : textures_)
            if (!texPool_ || t.dev < texPool_) texPool_ = t.dev;
        std::vector<TexDesc> d(std::max<size_t>(1, textures_.size()));
        for (size_t i = 0; i < textures_.size(); i++) {
            const TextureSlot& t = textures_[i];
            TexDesc td{};
            td.base = (unsigned long long)(t.dev - texPool_);
            td.width = t.width;
            td.height = t.height;
            td.levels = t.levels;
            td.layers = t.layers;
            d[i] = td;
        }
        if (!texDesc_.ensure(d.size() * sizeof(TexDesc))) return setError("hipMalloc textures"), false;
        if (!ok(hipMemcpy(texDesc_.ptr, d.data(), d.size() * sizeof(TexDesc), hipMemcpyHostToDevice), "tex desc"))
            return false;
        texDescDirty_ = false;
    }
    return true;
}

// Builds the view table of the next draw: the batch, or one view from the
// bound target and programs.  Tiles, record slots and pixels of the views are
// laid out one after another; each view's pixels start on a 64-pixel boundary.
bool TileRasterRenderer::layoutViews(int ntri) {
    std::vector<HostView> hv = batch_;
    if (hv.empty()) {
        HostView v{};
        v.width = width_;
        v.height = height_;
        std::memcpy(v.mvp, vs_.desc.mvp, sizeof(v.mvp));
        for (int k = 0; k < 3; k++) v.camera[k] = fs_.desc.camera[k];
        v.shade = fsBound_ && fs_.desc.kind != IFHIP_FS_DEPTH_ONLY ? 1 : 0;
        hv.push_back(v);
    }
    const unsigned sv = (unsigned)ntri + clipCap_;
    if ((unsigned long long)sv * hv.size() >= 0x80000000ull) {
        setError("too many triangles times views: record slots must stay below 2^31");
        return false;
    }
    if (samples_ == 1 && ntri >= (1 << 28)) {
        // 1-spp fragment keys carry the coverage class below the primitive key
        setError("too many triangles: at 1 sample per pixel a draw takes fewer than 2^28");
        return false;
    }
    std::vector<ViewDesc> d(hv.size());
    std::vector<int> layout;
    layout.reserve(3 * hv.size() + 1);
    unsigned tiles = 0, shade = 0;
    size_t pix = 0;
    for (size_t v = 0; v < hv.size(); v++) {
        const HostView& h = hv[v];
        ViewDesc& D = d[v];
        std::memset(&D, 0, sizeof(D));
        std::memcpy(D.mvp, h.mvp, sizeof(D.mvp));
        for (int k = 0; k < 3; k++) D.camera[k] = h.camera[k];
        D.shade = h.shade;
        D.width = h.width;
        D.height = h.height;
        D.tilesX = (h.width + kTile - 1) / kTile;
        D.tilesY = (h.height + kTile - 1) / kTile;
        D.invW = 1.0f / (float)h.width;  // TileRasterRenderer.cpp:59-60
        D.invH = 1.0f / (float)h.height;
        D.tileBase = tiles;
        tiles += (unsigned)(D.tilesX * D.tilesY);
        D.shadeBase = shade;
        if (h.shade)
            shade += (unsigned)(((h.width + shadeTile_ - 1) / shadeTile_) * ((h.height + shadeTile_ - 1) / shadeTile_));
        D.slotBase = (unsigned)v * sv;
        D.primBase = (unsigned)v * (unsigned)ntri;
        D.pixBase = pix;
        pix += ((size_t)h.width * h.height + 63) & ~(size_t)63;
        layout.push_back(h.width);
        layout.push_back(h.height);
        layout.push_back(h.shade);
    }
    layout.push_back(shadeTile_);
    totalTiles_ = tiles;
    shadeTiles_ = shade;
    totalPixels_ = pix;
    const bool descChanged = d.size() != viewDesc_.size() ||
                             std::memcmp(d.data(), viewDesc_.data(), d.size() * sizeof(ViewDesc)) != 0;
    if (descChanged) {
        if (!viewDev_.ensure(d.size() * sizeof(ViewDesc)) || !viewCtr_.ensure(d.size() * sizeof(ViewCounters)))
            return setError("hipMalloc views"), false;
        // the previous draw may still read the table
        if (!ok(hipStreamSynchronize(stream_), "view table") ||
            !ok(hipMemcpy(viewDev_.ptr, d.data(), d.size() * sizeof(ViewDesc), hipMemcpyHostToDevice), "view table"))
            return false;
        viewDesc_ = d;
    }
    if (layout != viewLayout_) {
        std::vector<unsigned short> tv(std::max(1u, tiles)), sh(std::max(1u, shade));
        for (size_t v = 0; v < d.size(); v++) {
            const unsigned n = (unsigned)(d[v].tilesX * d[v].tilesY);
            std::fill(tv.begin() + d[v].tileBase, tv.begin() + d[v].tileBase + n, (unsigned short)v);
            if (hv[v].shade) {
                const unsigned m = (unsigned)(((hv[v].width + shadeTile_ - 1) / shadeTile_) *
                                              ((hv[v].height + shadeTile_ - 1) / shadeTile_));
                std::fill(sh.begin() + d[v].shadeBase, sh.begin() + d[v].shadeBase + m, (unsigned short)v);
            }
        }
        if (!tileView_.ensure(tv.size() * 2) || !shadeView_.ensure(sh.size() * 2))
            return setError("hipMalloc view maps"), false;
        if (!ok(hipStreamSynchronize(stream_), "view maps") ||
            !ok(hipMemcpy(tileView_.ptr, tv.data(), tv.size() * 2, hipMemcpyHostToDevice), "view maps") ||
            !ok(hipMemcpy(shadeView_.ptr, sh.data(), sh.size() * 2, hipMemcpyHostToDevice), "view maps"))
            return false;
        viewLayout_ = layout;
    }
    return true;
}

Views TileRasterRenderer::deviceViews() const {
    Views V{};
    V.desc = viewDev_.as<ViewDesc>();
    V.tileView = tileView_.as<unsigned short>();
    V.shadeView = shadeView_.as<unsigned short>();
    V.ctr = viewCtr_.as<ViewCounters>();
    V.count = (int)viewDesc_.size();
    V.slotsPerView = (unsigned)curTris_ + clipCap_;
    return V;
}

TriRecs TileRasterRenderer::recPtrs() const {
    return TriRecs{recs_[0].as<float4>(), recs_[1].as<float4>(), recs_[2].as<float4>(),
                   recs_[3].as<float4>(), recs_[4].as<float4>(), recs_[5].as<float4>()};
}

StatPartials TileRasterRenderer::statPtrs() const {
    StatPartials st{};
    // setup workgroups: 1024 threads for up to four views, 256 beyond
    const int nv = (int)std::max<size_t>(1, viewDesc_.size()), per = nv <= 4 ? 1024 : 256;
    st.setupBlocks = (curTris_ + per - 1) / per * nv;
    st.rasterBlocks = rasterGrid();
    unsigned long long* base = statPart_.as<unsigned long long>();
    st.setupAssembled = base;
    st.rasterFrags = base + st.setupBlocks;
    st.rasterCovs = base + st.setupBlocks + st.rasterBlocks;
    return st;
}

// Raster workgroups: rasterGroupsPerCU_ per compute unit, four times as many
// when a 1-sample tile takes one wave.
int TileRasterRenderer::rasterGrid() const {
    return numCUs_ * rasterGroupsPerCU_ * (samples_ == 1 && tileThreads_ == 64 ? 4 : 1);
}

// Uniform block of the bound fragment program for the shading kernels.
ShadeParams TileRasterRenderer::shadeParams(bool writeFloat) const {
    ShadeParams sh{};
    const IfritGpuFragmentProgramDesc& d = fs_.desc;
    sh.samples = samples_;
    // the neural program falls back to the layered-GGX material when validation found it below its bound
    sh.program = (d.kind == IFHIP_FS_NEURAL && nnFallback_[d.precision == 1 ? 1 : 0]) ? IFHIP_FS_GGX_LAYERED : d.kind;
    sh.stride = stride_;
    sh.offWorld = d.offWorld;
    sh.offNormal = d.offNormal;
    sh.offUv = d.offUv;
    sh.offTangent = d.offTangent;
    sh.trisPerMaterial = d.trisPerMaterial;
    sh.sampler = d.sampler;
    sh.maxAniso = std::min(16, std::max(1, d.maxAniso));  // the tap list encodes up to 16 taps
    sh.samp = SamplerMode{kAddrRepeat, kAddrRepeat, 0, 0xff000000u};
    sh.lod = d.lod;
    if (d.samplerId >= 1 && d.samplerId <= (int)samplers_.size()) {
        // a sampler object: its addressing, filter and border; anisotropic filtering when it enables
        // it (maxAnisotropy taps at most), else the trilinear level of detail (SBC:357-377)
        const SamplerObject& o = samplers_[d.samplerId - 1];
        sh.samp = o.mode;
        sh.sampler = o.anisotropy ? 2 : 1;
        sh.maxAniso = std::min(16, std::max(1, (int)std::ceil(o.maxAnisotropy)));
    }
    sh.sampGeneric = (sh.samp.addrU != kAddrRepeat || sh.samp.addrV != kAddrRepeat || sh.samp.nearest) ? 1 : 0;
    // the normal-Lambert program samples no texture
    sh.numTextures = d.kind == IFHIP_FS_NORMAL_LAMBERT ? 0 : (int)textures_.size();
    sh.ambient = d.ambient;
    sh.kd = d.kd;
    sh.ks = d.ks;
    sh.shininess = d.shininess;
    sh.numLights = std::min(4, std::max(0, d.numLights));
    std::memcpy(sh.lightPos, d.lightPos, sizeof(sh.lightPos));
    std::memcpy(sh.lightColor, d.lightColor, sizeof(sh.lightColor));
    sh.writeFloat = writeFloat ? 1 : 0;
    sh.ntri = curTris_;
    sh.shadeTile = shadeTile_;
    sh.neural = neuralShadeParams();
    sh.swapped = needSwapFlags() ? swapped_.as<unsigned char>() : nullptr;
    std::memcpy(sh.clearColor, clearColor_, sizeof(sh.clearColor));
    sh.clearDepth = clearDepth_;
    sh.depthWrite = fsWritesDepth() ? 1 : 0;  // blending is always drawn deterministically
    sh.depthScale = d.depthScale;
    sh.depthBias = d.depthBias;
    sh.scissor = activeScissor();
    return sh;
}

void TileRasterRenderer::enqueue(bool clear) {
    const int tiles = (int)totalTiles_;
    const Views V = deviceViews();
    hipStream_t s = stream_;
    auto mark = [&](int i) {
        if (stageTiming_ && !capturing_) (void)hipEventRecord(ev_[i], s);
    };
    mark(0);
    // triangle groups outside a view are marked by the vertex launch and skipped by setup
    const bool cullGroups = !hostVs_ && vs_.desc.kind == IFHIP_VS_TRANSFORM && groupOut_.ptr;
    const int groups = (boundsTris_ + kGroupTris - 1) / kGroupTris;
    if (hostVs_)
        launch_reset(binCount_.as<unsigned>(), binCursor_.as<unsigned>(), tiles, V, ctr_.as<DeviceCounters>(), s);
    else  // the vertex launch also clears the counters
        launch_vertex(vtx_.as<float>(), stride_, nverts_, V,
                      clipAt_ >= 0 ? clipAt_ : vs_.desc.kind == IFHIP_VS_CLIP_PASSTHROUGH ? 0 : -1,
                      clip_.as<float4>(), binCount_.as<unsigned>(), binCursor_.as<unsigned>(), tiles,
                      ctr_.as<DeviceCounters>(), groupBounds_.as<float4>(), groups,
                      cullGroups ? groupOut_.as<unsigned char>() : nullptr, s);
    mark(1);
    if (polygonMode_ != 0) {
        enqueuePolygons(clear, V);
        return;
    }
    SetupParams sp{};
    sp.ntri = curTris_;
    sp.nverts = nverts_;
    sp.cullMode = cullMode_;
    sp.swapped = needSwapFlags() ? swapped_.as<unsigned char>() : nullptr;
    sp.msaa = samples_ > 1 ? 1 : 0;
    sp.clipCap = clipCap_;
    sp.invTileWidth = ifx::rcp_nr_ss(16.0f, rcpModel_);
    sp.rcpModel = rcpModel_;
    sp.groupOut = cullGroups ? groupOut_.as<unsigned char>() : nullptr;
    sp.groups = groups;
    launch_setup(sp, V, clip_.as<float4>(), idx_.as<int>(), recPtrs(), bary_.as<float4>(), binQueue_.as<uint4>(),
                 slotOf_.as<unsigned>(), clipList_.as<unsigned>(), statPtrs(), ctr_.as<DeviceCounters>(), s);
    mark(2);
    RasterQueue q{items_.as<uint4>(), tilePartial_.as<int>(), multiList_.as<unsigned>(),
                  partials_.as<unsigned long long>(), planSums_.as<unsigned>(), itemCap_, partialCap_, rasterGrid(),
                  numCUs_ * rasterGroupsPerCU_, partShift_, samples_ == 1 ? tileThreads_ : 256};
    // enough workgroups for every view's queue, at most 16 per compute unit
    const size_t queueChunks = ((size_t)curTris_ + clipCap_ + 255) / 256 * viewDesc_.size();
    BinParams bp{tiles, (int)std::max<size_t>(1, std::min<size_t>(queueChunks, (size_t)numCUs_ * 16)), curTris_,
                 clipCap_, binCap_, bigCap_, bigThreshold_};
    launch_bin(bp, V, recPtrs(), binQueue_.as<uint4>(), binCount_.as<unsigned>(), binStart_.as<unsigned>(),
               binCursor_.as<unsigned>(), binList_.as<unsigned>(), bigList_.as<unsigned>(), q,
               ctr_.as<DeviceCounters>(), s);
    mark(3);
    RasterParams rp{};
    rp.depthFunc = depthFunc_;
    rp.clear = clear ? 1 : 0;
    rp.clearDepth = clearDepth_;
    // optsetForceDeterministic(false) selects the tagging mode's depth expressions (TRW:1581-1675);
    // blending keeps the deterministic path, as the reference's does (TRW:410)
    rp.tagging = !deterministic_ && samples_ == 1 ? 1 : 0;
    // a depth-writing program stores its depth after the test; the tagging mode discards it
    rp.depthWrite = fsWritesDepth() && !rp.tagging ? 1 : 0;
    rp.depthScale = fs_.desc.depthScale;
    rp.depthBias = fs_.desc.depthBias;
    rp.scissor = activeScissor();
    rp.binCap = binCap_;
    const float* st = samples_ == 8 ? kS8 : (samples_ == 4 ? kS4 : kS1);
    for (int i = 0; i < samples_; i++) {
        rp.sx[i] = st[2 * i];
        rp.sy[i] = st[2 * i + 1];
    }
    // the 1-sample coverage class is a single-view diagnostic
    uint8_t* cls = batch_.empty() ? visClass_.as<uint8_t>() : nullptr;
    if (blend_.blendEnable) {
        // blending: fragments shaded and blended in primitive order (the shade stage is part of it)
        launch_blend_ordered(shadeParams(true), blendCoefs_, depthFunc_, clear ? 1 : 0, binCap_, V, recPtrs(),
                             bary_.as<float4>(), vtx_.as<float>(), idx_.as<int>(), texDesc_.as<TexDesc>(), texPool_,
                             binStart_.as<unsigned>(), binCount_.as<unsigned>(), binList_.as<unsigned>(), tiles,
                             rasterGrid(), visDepth_.as<float>(), visSlot_.as<uint32_t>(), cls, rgba8_.as<uchar4>(),
                             rgba32f_.as<float4>(), statPtrs(), s);
        mark(4);
    } else if (depthFunc_ == IF_COMPARE_OP_LESS && !rp.depthWrite) {
        // order-independent: the lexicographic minimum of (depth, primitive)
        launch_raster(rp, V, samples_, recPtrs(), slotOf_.as<unsigned>(), binStart_.as<unsigned>(),
                      binCount_.as<unsigned>(), binList_.as<unsigned>(), q, tiles, visDepth_.as<float>(),
                      visSlot_.as<uint32_t>(), cls, statPtrs(), ctr_.as<DeviceCounters>(), s);
        mark(4);
        launch_raster_merge(rp, V, samples_, recPtrs(), slotOf_.as<unsigned>(), binStart_.as<unsigned>(),
                            binCount_.as<unsigned>(), q, visDepth_.as<float>(), visSlot_.as<uint32_t>(), cls, tiles,
                            ctr_.as<DeviceCounters>(), s);
    } else {
        // every other depth function folds each pixel's fragments in primitive order
        launch_raster_ordered(rp, V, samples_, recPtrs(), binStart_.as<unsigned>(), binCount_.as<unsigned>(),
                              binList_.as<unsigned>(), tiles, rasterGrid(), visDepth_.as<float>(),
                              visSlot_.as<uint32_t>(), cls, statPtrs(), s);
        mark(4);
    }
    mark(5);
    if (fsBound_ && fs_.desc.kind != IFHIP_FS_DEPTH_ONLY && shadeTiles_ > 0 && !blend_.blendEnable) {
        const bool wantFloat = batch_.empty() && ((fb_ && !fb_->color.empty() && fb_->color[0]) || floatOut_);
        const ShadeParams sh = shadeParams(wantFloat);
        launch_shade(sh, V, (int)shadeTiles_, recPtrs(), bary_.as<float4>(), visSlot_.as<uint32_t>(),
                     vtx_.as<float>(), idx_.as<int>(), texDesc_.as<TexDesc>(), texPool_, rgba8_.as<uchar4>(),
                     wantFloat ? rgba32f_.as<float4>() : nullptr, s);
    }
    mark(6);
}

// Point and line modes: candidates per pixel, the resolve, then the deferred
// shading of the winners' records (their planes weight the point's vertex or
// the line's two ends).  Every workgroup of the candidate launch writes its
// statistics partials, so the grid covers the setup and raster partials.
void TileRasterRenderer::enqueuePolygons(bool clear, const Views& V) {
    hipStream_t s = stream_;
    auto mark = [&](int i) {
        if (stageTiming_ && !capturing_) (void)hipEventRecord(ev_[i], s);
    };
    PolyParams pp{};
    pp.mode = polygonMode_;
    pp.width = width_;
    pp.height = height_;
    pp.nidx = curIdx_;
    pp.nverts = nverts_;
    pp.clear = clear ? 1 : 0;
    pp.clearDepth = clearDepth_;
    pp.keys = polyKeys_.as<unsigned long long>();
    const StatPartials st = statPtrs();
    const int grid = std::max(std::max(st.setupBlocks, st.rasterBlocks), 1);
    mark(2);
    mark(3);
    launch_polygons(pp, clip_.as<float4>(), idx_.as<int>(), recPtrs(), bary_.as<float4>(), visDepth_.as<float>(),
                    visSlot_.as<uint32_t>(), visClass_.as<uint8_t>(), st, ctr_.as<DeviceCounters>(), grid, s);
    mark(4);
    mark(5);
    if (fsBound_ && fs_.desc.kind != IFHIP_FS_DEPTH_ONLY && shadeTiles_ > 0) {
        const bool wantFloat = (fb_ && !fb_->color.empty() && fb_->color[0]) || floatOut_;
        ShadeParams sh = shadeParams(wantFloat);
        sh.ntri = 0;  // every record carries its own vertex weights (bx, by)
        sh.swapped = nullptr;
        launch_shade(sh, V, (int)shadeTiles_, recPtrs(), bary_.as<float4>(), visSlot_.as<uint32_t>(),
                     vtx_.as<float>(), idx_.as<int>(), texDesc_.as<TexDesc>(), texPool_, rgba8_.as<uchar4>(),
                     wantFloat ? rgba32f_.as<float4>() : nullptr, s);
    }
    mark(6);
}

// Everything a captured graph bakes into its launches.  View tables live in
// device memory, so cameras and transforms change without a new capture.
std::vector<uint8_t> TileRasterRenderer::graphKey(bool clear) const {
    std::vector<uint8_t> k;
    auto put = [&](const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        k.insert(k.end(), b, b + n);
    };
    const int ints[] = {clear ? 1 : 0, curTris_, nverts_, stride_, samples_, depthFunc_, (int)viewDesc_.size(),
                        (int)totalTiles_, (int)shadeTiles_, vs_.desc.kind, fsBound_ ? 1 : 0, hostVs_ ? 1 : 0,
                        partShift_, rasterGroupsPerCU_, shadeTile_, bigThreshold_, nnFallback_[0] ? 1 : 0,
                        nnFallback_[1] ? 1 : 0, (int)textures_.size(), stageTiming_ ? 1 : 0,
                        fb_ && !fb_->color.empty() && fb_->color[0] ? 1 : 0, batch_.empty() ? 1 : 0,
                        floatOut_ ? 1 : 0, blend_.blendEnable ? 1 : 0, cullMode_, deterministic_ ? 1 : 0,
                        polygonMode_, curIdx_, clipAt_, rcpModel_};
    put(ints, sizeof(ints));
    // emitted lists swap buffers without reallocating them
    const void* inputs[] = {vtx_.ptr, idx_.ptr};
    put(inputs, sizeof(inputs));
    const unsigned caps[] = {clipCap_, binCap_, bigCap_, itemCap_, partialCap_};
    put(caps, sizeof(caps));
    put(&totalPixels_, sizeof(totalPixels_));
    put(&blendCoefs_, sizeof(blendCoefs_));
    put(clearColor_, sizeof(clearColor_));
    put(&clearDepth_, sizeof(clearDepth_));
    const Scissor sc = activeScissor();
    put(&sc, sizeof(sc));
    put(&DevBuf::generation, sizeof(DevBuf::generation));
    IfritGpuFragmentProgramDesc fd = fs_.desc;
    std::memset(fd.camera, 0, sizeof(fd.camera));  // cameras live in the view table
    put(&fd, sizeof(fd));
    const NeuralShadeParams np = neuralShadeParams();
    put(&np, sizeof(np));
    return k;
}

void TileRasterRenderer::dropGraph() {
    if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graphExec_ = nullptr;
    graph_ = nullptr;
    graphKey_.clear();
}

// Issues one draw: directly, or as a replay of the captured graph of this
// configuration (captured on first use).
void TileRasterRenderer::submit(bool clear) {
    if (!graphMode_) {
        enqueue(clear);
        return;
    }
    std::vector<uint8_t> key = graphKey(clear);
    if (!graphExec_ || key != graphKey_) {
        dropGraph();
        if (!ok(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "graph capture")) return;
        capturing_ = true;
        enqueue(clear);
        capturing_ = false;
        hipGraph_t g = nullptr;
        if (!ok(hipStreamEndCapture(stream_, &g), "graph capture")) return;
        graph_ = g;
        if (!ok(hipGraphInstantiate(&graphExec_, graph_, nullptr, nullptr, 0), "graph instantiate")) {
            dropGraph();
            return;
        }
        graphKey_ = std::move(key);
    }
    ok(hipGraphLaunch(graphExec_, stream_), "graph launch");
}

bool TileRasterRenderer::finish(bool clear, int indexCount) {
    (void)clear;
    (void)indexCount;
    // frame statistics are reduced from the per-workgroup partials only when read
    launch_finalize(statPtrs(), deviceViews(), ctr_.as<DeviceCounters>(), stream_);
    if (!ok(hipStreamSynchronize(stream_), "frame")) return false;
    DeviceCounters c{};
    if (!ok(hipMemcpy(&c, ctr_.ptr, sizeof(c), hipMemcpyDeviceToHost), "counters")) return false;
    stats_.inputTriangles = (int64_t)curTris_ * (int64_t)viewDesc_.size();
    stats_.assembledTriangles = (int64_t)c.assembled;
    stats_.clippedTriangles = (int64_t)c.clipCount;
    stats_.binEntries = c.binTotal;
    stats_.fragments = (int64_t)c.fragments;
    stats_.coveredSamples = (int64_t)c.coveredSamples;
    stats_.status = c.status;
    if (stageTiming_ && !graphMode_) {
        for (int i = 0; i < 6; i++) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev_[i], ev_[i + 1]);
            stats_.stageMs[i] = ms;
        }
    }
    // capacities grow in 64-bit arithmetic and stay below 2^32 - 16
    auto grow = [](unsigned cap, unsigned need) {
        const unsigned long long g = std::max<unsigned long long>(2ull * cap, need);
        return (unsigned)std::min<unsigned long long>(g, 0xFFFFFFF0ull);
    };
    if (c.status & kStatusClipOverflow) clipCap_ = grow(clipCap_, c.clipMax + 1024u);
    if (c.status & kStatusBinOverflow) binCap_ = grow(binCap_, (unsigned)std::min<unsigned long long>(
                                                                   (unsigned long long)c.binTotal + (1u << 20), 0xFFFFFFF0ull));
    if (c.status & kStatusBigOverflow) bigCap_ = grow(bigCap_, c.bigCount + 1024u);
    if (c.status & kStatusRasterOverflow) {
        itemCap_ = grow(itemCap_, c.numItems + 1024u);
        partialCap_ = grow(partialCap_, c.numPartials + 256u);
    }
    if (c.status & kStatusBadIndex) setError("index buffer references a vertex beyond the vertex buffer");
    return (c.status & (kStatusClipOverflow | kStatusBinOverflow | kStatusBigOverflow | kStatusRasterOverflow)) == 0;
}

void TileRasterRenderer::drawElements(int indexCount, bool clear) {
    if (gsBound_ && clipAt_ < 0) return drawGeometry(indexCount, clear, false);
    stats_.retries = 0;
    constexpr int kAttempts = 4;
    for (int attempt = 0; attempt < kAttempts; attempt++) {
        if (!prepare(indexCount)) return;
        submit(clear);
        if (!ok(hipGetLastError(), "kernel launch")) return;
        if (finish(clear, indexCount)) break;
        stats_.retries++;
        if (!clear) {
            setError("frame exceeded the binning buffers while drawing without clear");
            break;
        }
        if (attempt == kAttempts - 1)
            setError("frame still exceeded the binning buffers after growing them three times; "
                     "the outputs are incomplete");
    }
    copyOutToHost();
}

void TileRasterRenderer::drawAsync(int indexCount, bool clear) {
    if (gsBound_ && clipAt_ < 0) return drawGeometry(indexCount, clear, true);
    if (!prepare(indexCount)) return;
    submit(clear);
    ok(hipGetLastError(), "kernel launch");
}

void TileRasterRenderer::synchronize() {
    if (!stream_) return;
    // an asynchronous draw is not re-run: an overflow leaves an incomplete frame and is reported
    // (the buffers have grown, so the next draw fits)
    if (!finish(true, curTris_ * 3) && error_.empty())
        setError("frame exceeded the binning buffers (they have grown for the next draw); the outputs are incomplete");
}

void TileRasterRenderer::copyOutToHost() {
    if (!fb_ || !batch_.empty()) return;
    const size_t px = (size_t)width_ * height_;
    if (!fb_->color.empty() && fb_->color[0] && fsBound_) {
        ImageF32* img = fb_->color[0];
        if (img->channel == 4 && img->data.size() >= px * 4)
            ok(hipMemcpy(img->data.data(), rgba32f_.ptr, px * 16, hipMemcpyDeviceToHost), "colour readback");
    }
    if (fb_->depth && fb_->depth->data.size() >= px) {
        std::vector<float> d(px * samples_);
        if (!ok(hipMemcpy(d.data(), visDepth_.ptr, d.size() * 4, hipMemcpyDeviceToHost), "depth readback")) return;
        // multi-sample depth resolves to the sample mean, as the CUDA path's
        // resolve does (TileRasterCoreInvocationCuda.cu:2319-2361)
        for (size_t i = 0; i < px; i++) {
            if (samples_ == 1) {
                fb_->depth->data[i] = d[i];
            } else {
                double acc = 0;
                for (int s = 0; s < samples_; s++) acc += d[i * samples_ + s];
                fb_->depth->data[i] = (float)(acc / samples_);
            }
        }
    }
}

void TileRasterRenderer::clear() {
    // TileRasterRenderer::clear (TileRasterRenderer.cpp:430-434): colour 0, depth 255.
    if (fb_) {
        for (ImageF32* c : fb_->color)
            if (c) std::fill(c->data.begin(), c->data.end(), 0.0f);
        if (fb_->depth) std::fill(fb_->depth->data.begin(), fb_->depth->data.end(), 255.0f);
    }
}

void TileRasterRenderer::getViewVisibility(int view, uint32_t* ids, float* depth) {
    if (view < 0 || view >= (int)viewDesc_.size()) return setError("bad view index");
    const ViewDesc& D = viewDesc_[view];
    const size_t n = (size_t)D.width * D.height * samples_, off = (size_t)D.pixBase * samples_;
    if (ids) {
        launch_slot_to_prim(visSlot_.as<uint32_t>() + off, recPtrs(), n, visId_.as<uint32_t>() + off, stream_,
                            polygonMode_ != 0);
        if (!ok(hipStreamSynchronize(stream_), "id expansion")) return;
        ok(hipMemcpy(ids, visId_.as<uint32_t>() + off, n * 4, hipMemcpyDeviceToHost), "id readback");
    }
    if (depth) ok(hipMemcpy(depth, visDepth_.as<float>() + off, n * 4, hipMemcpyDeviceToHost), "depth readback");
}

void TileRasterRenderer::getViewColorRGBA8(int view, uint8_t* rgba) {
    if (view < 0 || view >= (int)viewDesc_.size()) return setError("bad view index");
    const ViewDesc& D = viewDesc_[view];
    ok(hipMemcpy(rgba, rgba8_.as<uint8_t>() + (size_t)D.pixBase * 4, (size_t)D.width * D.height * 4,
                 hipMemcpyDeviceToHost),
       "colour readback");
}

void TileRasterRenderer::getVisibility(uint32_t* ids, float* depth) { getViewVisibility(0, ids, depth); }

void TileRasterRenderer::getCoverageClass(uint8_t* cls) {
    if (!batch_.empty()) return setError("coverage classes are kept for single-view draws only");
    ok(hipMemcpy(cls, visClass_.ptr, (size_t)width_ * height_, hipMemcpyDeviceToHost), "class readback");
}

void TileRasterRenderer::getColorRGBA8(uint8_t* rgba) { getViewColorRGBA8(0, rgba); }

void TileRasterRenderer::getColorRGBA32F(float* rgba) {
    if (!batch_.empty() || (!floatOut_ && !blend_.blendEnable && !(fb_ && !fb_->color.empty() && fb_->color[0])))
        return setError("FP32 colour is written for a single target with iftrSetColorFloatOutput or blending");
    ok(hipMemcpy(rgba, rgba32f_.ptr, (size_t)width_ * height_ * 16, hipMemcpyDeviceToHost), "colour readback");
}

void TileRasterRenderer::setViews(const IfritViewDesc* views, int n) {
    if (n < 0 || n > kMaxViews || (n > 0 && !views)) return setError("bad view count");
    std::vector<HostView> b((size_t)n);
    for (int i = 0; i < n; i++) {
        const IfritViewDesc& v = views[i];
        if (v.width <= 0 || v.height <= 0 || v.width > 16384 || v.height > 16384)
            return setError("view sizes must be 1..16384");
        b[i].width = v.width;
        b[i].height = v.height;
        std::memcpy(b[i].mvp, v.mvp, sizeof(b[i].mvp));
        for (int k = 0; k < 3; k++) b[i].camera[k] = v.camera[k];
        b[i].shade = v.shade ? 1 : 0;
    }
    batch_ = std::move(b);
}

void TileRasterRenderer::setGraphMode(bool on) {
    if (!on) dropGraph();
    graphMode_ = on;
}

void TileRasterRenderer::setTuning(const IfritTuning& t) {
    const int part = t.rasterPartEntries;
    if ((part != 256 && part != 512 && part != 1024) || t.rasterGroupsPerCU < 1 || t.rasterGroupsPerCU > 8 ||
        (t.shadeTile != 8 && t.shadeTile != 16 && t.shadeTile != 32) || t.bigTriangleTiles < 1 ||
        (t.rasterTileThreads != 64 && t.rasterTileThreads != 256))
        return setError("tuning out of range: part 256/512/1024, 1..8 groups per CU, shading tile 8/16/32, "
                        "64 or 256 threads per tile");
    partShift_ = part == 256 ? 8 : (part == 512 ? 9 : 10);
    rasterGroupsPerCU_ = t.rasterGroupsPerCU;
    shadeTile_ = t.shadeTile;
    bigThreshold_ = t.bigTriangleTiles;
    tileThreads_ = t.rasterTileThreads;
    if (t.binCapacity && t.binCapacity != binCap_) {
        // an explicit capacity may also shrink the list (it grows again on overflow)
        binCap_ = t.binCapacity;
        binCapSet_ = true;
        binList_.release();
        DevBuf::generation++;
    }
}

void TileRasterRenderer::getTuning(IfritTuning* t) const {
    t->rasterPartEntries = 1 << partShift_;
    t->rasterGroupsPerCU = rasterGroupsPerCU_;
    t->shadeTile = shadeTile_;
    t->bigTriangleTiles = bigThreshold_;
    t->binCapacity = binCap_;
    t->rasterTileThreads = tileThreads_;
}

uint64_t TileRasterRenderer::digest() {
    if (viewDesc_.empty() || !digest_.ensure(3 * 8)) return 0;
    const size_t ps = totalPixels_ * samples_;
    // primitive IDs (not record slots, whose numbering depends on capacities)
    launch_slot_to_prim(visSlot_.as<uint32_t>(), recPtrs(), ps, visId_.as<uint32_t>(), stream_, polygonMode_ != 0);
    if (!ok(hipMemsetAsync(digest_.ptr, 0, 3 * 8, stream_), "digest")) return 0;
    // samples outside any view (64-pixel padding) hold stale values: digest views one by one
    for (const ViewDesc& D : viewDesc_) {
        const size_t off = (size_t)D.pixBase, n = (size_t)D.width * D.height;
        launch_digest(reinterpret_cast<const uint32_t*>(visDepth_.as<float>() + off * samples_), n * samples_,
                      digest_.as<unsigned long long>(), stream_);
        launch_digest(visId_.as<uint32_t>() + off * samples_, n * samples_, digest_.as<unsigned long long>() + 1,
                      stream_);
        if (D.shade && fsBound_ && fs_.desc.kind != IFHIP_FS_DEPTH_ONLY)
            launch_digest(reinterpret_cast<const uint32_t*>(rgba8_.as<uchar4>() + off), n,
                          digest_.as<unsigned long long>() + 2, stream_);
    }
    unsigned long long h[3] = {0, 0, 0};
    if (!ok(hipMemcpyAsync(h, digest_.ptr, sizeof(h), hipMemcpyDeviceToHost, stream_), "digest") ||
        !ok(hipStreamSynchronize(stream_), "digest"))
        return 0;
    return h[0] ^ (h[1] * 0x9E3779B97F4A7C15ull) ^ (h[2] * 0xC2B2AE3D27D4EB4Full);
}

// Coordinate search: each parameter in turn over its values with the others
// held at the best so far, two rounds; every configuration must reproduce the
// first one's digest.
void TileRasterRenderer::autotune(int indexCount, int frames, IfritAutotuneReport* out) {
    IfritAutotuneReport rep{};
    const bool graphSaved = graphMode_, timingSaved = stageTiming_;
    setGraphMode(true);
    stageTiming_ = false;
    frames = std::max(1, frames);
    IfritTuning best{};
    getTuning(&best);
    unsigned long long ref = 0;
    bool haveRef = false;
    rep.identical = 1;
    hipEvent_t e0 = ev_[0], e1 = ev_[1];
    auto measure = [&](IfritTuning t) -> float {
        t.binCapacity = 0;  // keep the current capacity (it grows on overflow)
        setTuning(t);
        drawElements(indexCount, true);  // captures the graph, grows buffers if needed
        if (!error_.empty()) return -1.0f;
        const unsigned long long d = digest();
        if (!haveRef) {
            ref = d;
            haveRef = true;
        } else if (d != ref) {
            rep.identical = 0;
        }
        drawAsync(indexCount, true);  // warm replay
        (void)hipEventRecord(e0, stream_);
        for (int f = 0; f < frames; f++) drawAsync(indexCount, true);
        (void)hipEventRecord(e1, stream_);
        if (!ok(hipEventSynchronize(e1), "autotune")) return -1.0f;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)frames;
        if (rep.candidates < 64) {
            rep.tried[rep.candidates] = t;
            rep.ms[rep.candidates] = ms;
        }
        rep.candidates++;
        return ms;
    };
    float bestMs = measure(best);
    rep.defaultMs = bestMs;
    // candidate values per parameter (threads per tile only matter at 1 sample per pixel)
    const int values[5][3] = {{256, 512, 1024}, {2, 4, 8}, {8, 16, 32}, {16, 64, 256}, {256, 64, 64}};
    const int knobs = samples_ == 1 ? 5 : 4;
    for (int round = 0; round < 2 && bestMs >= 0.0f; round++) {
        for (int knob = 0; knob < knobs; knob++) {
            for (int i = 0; i < 3; i++) {
                IfritTuning t = best;
                int* const fields[5] = {&t.rasterPartEntries, &t.rasterGroupsPerCU, &t.shadeTile, &t.bigTriangleTiles,
                                        &t.rasterTileThreads};
                int* field = fields[knob];
                const int val = values[knob][i];
                if (*field == val) continue;
                *field = val;
                const float ms = measure(t);
                if (ms < 0.0f) break;
                if (ms < bestMs) {
                    bestMs = ms;
                    best = t;
                }
            }
        }
    }
    // bin capacity: the measured demand plus 1/8, in 64 Ki-entry steps
    setTuning(best);
    drawElements(indexCount, true);
    rep.binDemand = (uint32_t)stats_.binEntries;
    const unsigned long long want = ((unsigned long long)rep.binDemand * 9 / 8 + 0xFFFF) & ~0xFFFFull;
    best.binCapacity = (uint32_t)std::min<unsigned long long>(std::max<unsigned long long>(want, 1u << 20), 0xFFFFFFF0ull);
    setTuning(best);
    drawElements(indexCount, true);
    if (digest() != ref) rep.identical = 0;
    getTuning(&rep.best);
    rep.bestMs = bestMs;
    setGraphMode(graphSaved);
    stageTiming_ = timingSaved;
    if (out) *out = rep;
}

}  // namespace ifrit_hip
