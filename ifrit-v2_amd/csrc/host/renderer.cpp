// TileRasterRenderer for MI355X: buffer management and frame orchestration.
// The stage order follows the reference's draw call
// (projects/softgraphics/src/engine/tileraster/TileRasterRenderer.cpp:436-457,
// TileRasterWorker.cpp:108-135): vertex processing, geometry processing and
// binning, then per-tile rasterization, depth test and shading, with each
// stage a kernel on the renderer's stream instead of a thread barrier.
#include "renderer.hpp"

#include "../common/x86_fp.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ifrit_hip {

namespace {
constexpr int kTile = 16;
constexpr int kBigThreshold = 64;

// Sample offsets from the pixel corner: 1x samples the corner itself, like
// the CPU path; 4x and 8x are the standard patterns of
// projects/softgraphics/src/engine/tilerastercuda/TileRasterCoreInvocationCuda.cu:210-216.
const float kS1[2] = {0.0f, 0.0f};
const float kS4[8] = {0.375f, 0.125f, 0.875f, 0.375f, 0.125f, 0.625f, 0.625f, 0.875f};
const float kS8[16] = {0.5625f, 0.3125f, 0.4375f, 0.6875f, 0.8125f, 0.5625f, 0.3125f, 0.1875f,
                       0.1875f, 0.8125f, 0.0625f, 0.4375f, 0.6875f, 0.9375f, 0.9375f, 0.0625f};
}  // namespace

bool DevBuf::ensure(size_t n) {
    if (n <= bytes && ptr) return true;
    release();
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
    for (DevBuf* b : {&vtx_, &idx_, &clip_, &recs_[0], &recs_[1], &recs_[2], &recs_[3], &recs_[4], &recs_[5], &bary_, &tileBox_, &slotOf_, &binCount_, &binStart_, &binCursor_, &binList_,
                      &bigList_, &ctr_, &clipList_, &statPart_, &items_, &tilePartial_, &multiList_, &partials_, &visId_, &visDepth_, &visSlot_, &visClass_, &rgba8_, &rgba32f_, &texDesc_, &nnParams_, &nnGrads_, &nnM_, &nnV_,
                      &nnPacked_, &nnQinfo_, &lat32_, &latGrad_, &latM_, &latV_, &lat16_, &nnData_, &nnFeat_, &nnCount_,
                      &nnLoss_, &nnEvalIn_, &nnEvalOut_, &nnGradPart_})
        b->release();
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

void TileRasterRenderer::setBlendFunc(const IfritColorAttachmentBlendState& s) {
    blend_ = s;
    if (s.blendEnable) setError("alpha blending is not supported by the GPU backend");
}

void TileRasterRenderer::SetDepthFunc(int f) {
    depthFuncSaved_ = f;
    if (depthTest_) depthFunc_ = f;
}

void TileRasterRenderer::optSetDepthTestEnable(bool opt) {
    depthTest_ = opt;
    depthFunc_ = opt ? depthFuncSaved_ : IF_COMPARE_OP_ALWAYS;
}

void TileRasterRenderer::setMsaaSamples(int s) {
    if (s != 1 && s != 4 && s != 8) {
        setError("MSAA sample count must be 1, 4 or 8");
        return;
    }
    samples_ = s;
}

int TileRasterRenderer::createTexture(int w, int h, int levels) {
    Init();
    if ((w & (w - 1)) || (h & (h - 1)) || w <= 0 || h <= 0 || levels <= 0 || levels > 16) {
        setError("textures must have power-of-two sizes and 1..16 levels");
        return -1;
    }
    TextureSlot t;
    t.width = w;
    t.height = h;
    t.levels = levels;
    size_t off = 0;
    for (int l = 0; l < levels; l++) {
        t.levelOffset.push_back(off);
        off += (size_t)std::max(1, w >> l) * std::max(1, h >> l);
    }
    t.bytes = off * 4;
    if (!ok(hipMalloc(&t.dev, t.bytes), "hipMalloc texture")) return -1;
    textures_.push_back(t);
    texDescDirty_ = true;
    return (int)textures_.size() - 1;
}

void TileRasterRenderer::copyHostBufferToImage(int tex, const void* rgba8) {
    if (tex < 0 || tex >= (int)textures_.size()) return setError("bad texture id");
    const TextureSlot& t = textures_[tex];
    ok(hipMemcpy(t.dev, rgba8, (size_t)t.width * t.height * 4, hipMemcpyHostToDevice), "texture upload");
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
        launch_mip_level(t.dev + t.levelOffset[l - 1] * 4, sw, sh, t.dev + t.levelOffset[l] * 4, dw, dh, stream_);
    }
    ok(hipGetLastError(), "mip launch");
    ok(hipStreamSynchronize(stream_), "mip sync");
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

bool TileRasterRenderer::prepare(int indexCount) {
    Init();
    if (!initialized_) return false;
    if (width_ <= 0 || height_ <= 0) {
        setError("no frame buffer bound");
        return false;
    }
    if (width_ > 16384 || height_ > 16384) {
        setError("frame buffer larger than 16384 x 16384");
        return false;
    }
    if (depthFunc_ != IF_COMPARE_OP_LESS) {
        setError("the GPU backend implements the LESS depth test only");
        return false;
    }
    if (blend_.blendEnable) {
        setError("alpha blending is not supported by the GPU backend");
        return false;
    }
    // inputs
    if (hostVs_ && hostVb_) {
        runHostVertexShader();
        vtxDirty_ = false;
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
    }
    if (idxDirty_ && hostIb_) {
        if (!idx_.ensure(hostIb_->data.size())) return setError("hipMalloc index buffer"), false;
        if (!ok(hipMemcpy(idx_.ptr, hostIb_->data.data(), hostIb_->data.size(), hipMemcpyHostToDevice),
                "index upload"))
            return false;
        nidx_ = (int)(hostIb_->data.size() / 4);
        idxDirty_ = false;
    }
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
    if (fsBound_ && fs_.desc.kind >= IFHIP_FS_GGX_LAYERED && (stride_ < 11 || fs_.desc.offTangent + 3 > stride_)) {
        setError("the layered-GGX and neural programs need a tangent in the vertex record");
        return false;
    }
    if (!hostVs_ && vs_.desc.kind == IFHIP_VS_CLIP_PASSTHROUGH && stride_ < 4) {
        setError("the clip pass-through vertex program needs 4 floats per vertex");
        return false;
    }
    const int ntri = indexCount / 3;
    curTris_ = ntri;
    const int tiles = ((width_ + kTile - 1) / kTile) * ((height_ + kTile - 1) / kTile);
    const size_t px = (size_t)width_ * height_;
    const size_t ps = px * samples_;
    clipCap_ = std::max<unsigned>(clipCap_, std::max(4096u, (unsigned)(ntri / 8)));
    binCap_ = std::max<unsigned>(binCap_, std::max(1u << 20, (unsigned)std::min<size_t>(0xFFFFFFF0ull, (size_t)ntri * 6)));
    bigCap_ = std::max<unsigned>(bigCap_, std::max(1024u, (unsigned)(ntri / 16)));
    itemCap_ = std::max<unsigned>(itemCap_, (unsigned)tiles + binCap_ / kRasterPart + 1u);
    partialCap_ = std::max<unsigned>(partialCap_, std::max(256u, binCap_ / (8u * kRasterPart)));
    const size_t slots = (size_t)ntri + clipCap_;
    bool okAll = clip_.ensure((size_t)nverts_ * 16) && recs_[0].ensure(slots * 16) && recs_[1].ensure(slots * 16) &&
                 recs_[2].ensure(slots * 16) && recs_[3].ensure(slots * 16) && recs_[4].ensure(slots * 16) &&
                 recs_[5].ensure(slots * 16) &&
                 bary_.ensure(slots * 32) && tileBox_.ensure(slots * 16) &&
                 slotOf_.ensure((size_t)std::max(1, ntri) * 4) && clipList_.ensure((size_t)clipCap_ * 4) &&
                 statPart_.ensure(((size_t)(ntri + 255) / 256 + 2 * (size_t)rasterGrid() + 2) * 8) &&
                 items_.ensure((size_t)itemCap_ * 8) && tilePartial_.ensure((size_t)tiles * 4) &&
                 multiList_.ensure((size_t)tiles * 8) && partials_.ensure((size_t)partialCap_ * 256 * samples_ * 8) && binCount_.ensure((size_t)tiles * 4) &&
                 binStart_.ensure((size_t)tiles * 4) && binCursor_.ensure((size_t)tiles * 4) &&
                 binList_.ensure((size_t)binCap_ * 4) && bigList_.ensure((size_t)bigCap_ * 4) &&
                 visId_.ensure(ps * 4) && visDepth_.ensure(ps * 4) && visSlot_.ensure(ps * 4) &&
                 visClass_.ensure(px) && rgba8_.ensure(px * 4) && rgba32f_.ensure(px * 16);
    if (!okAll) {
        setError("device allocation failed");
        return false;
    }
    if (texDescDirty_) {
        std::vector<TexDesc> d(std::max<size_t>(1, textures_.size()));
        for (size_t i = 0; i < textures_.size(); i++) {
            const TextureSlot& t = textures_[i];
            TexDesc td{};
            td.base = (unsigned long long)(uintptr_t)t.dev;
            td.width = t.width;
            td.height = t.height;
            td.levels = t.levels;
            for (int l = 0; l < t.levels; l++) td.levelOffset[l] = (unsigned)t.levelOffset[l];
            d[i] = td;
        }
        if (!texDesc_.ensure(d.size() * sizeof(TexDesc))) return setError("hipMalloc textures"), false;
        if (!ok(hipMemcpy(texDesc_.ptr, d.data(), d.size() * sizeof(TexDesc), hipMemcpyHostToDevice), "tex desc"))
            return false;
        texDescDirty_ = false;
    }
    return true;
}

TriRecs TileRasterRenderer::recPtrs() const {
    return TriRecs{recs_[0].as<float4>(), recs_[1].as<float4>(), recs_[2].as<float4>(),
                   recs_[3].as<float4>(), recs_[4].as<float4>(), recs_[5].as<float4>()};
}

StatPartials TileRasterRenderer::statPtrs() const {
    const int tilesX = (width_ + kTile - 1) / kTile, tilesY = (height_ + kTile - 1) / kTile;
    StatPartials st{};
    (void)tilesX;
    (void)tilesY;
    st.setupBlocks = (curTris_ + 255) / 256;
    st.rasterBlocks = rasterGrid();
    unsigned long long* base = statPart_.as<unsigned long long>();
    st.setupAssembled = base;
    st.rasterFrags = base + st.setupBlocks;
    st.rasterCovs = base + st.setupBlocks + st.rasterBlocks;
    return st;
}

int TileRasterRenderer::rasterGrid() const { return numCUs_ * 4; }

void TileRasterRenderer::enqueue(bool clear) {
    const int tilesX = (width_ + kTile - 1) / kTile, tilesY = (height_ + kTile - 1) / kTile;
    const int tiles = tilesX * tilesY;
    hipStream_t s = stream_;
    auto mark = [&](int i) {
        if (stageTiming_) (void)hipEventRecord(ev_[i], s);
    };
    mark(0);
    if (!hostVs_)
        launch_vertex(vtx_.as<float>(), stride_, nverts_, vs_.desc.mvp, vs_.desc.kind == IFHIP_VS_CLIP_PASSTHROUGH ? 1 : 0,
                      clip_.as<float4>(), binCount_.as<unsigned>(), binCursor_.as<unsigned>(), tiles,
                      ctr_.as<DeviceCounters>(), s);
    mark(1);
    if (hostVs_) launch_reset(binCount_.as<unsigned>(), binCursor_.as<unsigned>(), tiles, ctr_.as<DeviceCounters>(), s);
    mark(2);
    SetupParams sp{};
    sp.width = width_;
    sp.height = height_;
    sp.tilesX = tilesX;
    sp.tilesY = tilesY;
    sp.ntri = curTris_;
    sp.nverts = nverts_;
    sp.ccw = 0;
    sp.msaa = samples_ > 1 ? 1 : 0;
    sp.clipCap = clipCap_;
    sp.invW = 1.0f / (float)width_;  // TileRasterRenderer.cpp:59-60
    sp.invH = 1.0f / (float)height_;
    sp.invTileWidth = ifx::rcp_nr_ss(16.0f);
    launch_setup(sp, clip_.as<float4>(), idx_.as<int>(), recPtrs(), bary_.as<float4>(), tileBox_.as<int4>(),
                 slotOf_.as<unsigned>(), clipList_.as<unsigned>(), statPtrs(), ctr_.as<DeviceCounters>(), s);
    mark(3);
    RasterQueue q{items_.as<uint2>(), tilePartial_.as<int>(), multiList_.as<unsigned>(),
                  partials_.as<unsigned long long>(), itemCap_, partialCap_, rasterGrid()};
    BinParams bp{tilesX, tilesY, curTris_, clipCap_, binCap_, bigCap_, kBigThreshold};
    launch_bin(bp, recPtrs(), tileBox_.as<int4>(), binCount_.as<unsigned>(), binStart_.as<unsigned>(),
               binCursor_.as<unsigned>(), binList_.as<unsigned>(), bigList_.as<unsigned>(), q, ctr_.as<DeviceCounters>(),
               s);
    mark(4);
    RasterParams rp{};
    rp.width = width_;
    rp.height = height_;
    rp.tilesX = tilesX;
    rp.tilesY = tilesY;
    rp.depthFunc = depthFunc_;
    rp.clear = clear ? 1 : 0;
    rp.binCap = binCap_;
    const float* st = samples_ == 8 ? kS8 : (samples_ == 4 ? kS4 : kS1);
    for (int i = 0; i < samples_; i++) {
        rp.sx[i] = st[2 * i];
        rp.sy[i] = st[2 * i + 1];
    }
    launch_raster(rp, samples_, recPtrs(), slotOf_.as<unsigned>(), binStart_.as<unsigned>(),
                  binCount_.as<unsigned>(), binList_.as<unsigned>(), q, visDepth_.as<float>(), visSlot_.as<uint32_t>(),
                  visClass_.as<uint8_t>(), statPtrs(), ctr_.as<DeviceCounters>(), s);
    mark(5);
    launch_raster_merge(rp, samples_, recPtrs(), slotOf_.as<unsigned>(), binStart_.as<unsigned>(),
                        binCount_.as<unsigned>(), q, visDepth_.as<float>(), visSlot_.as<uint32_t>(),
                        visClass_.as<uint8_t>(), ctr_.as<DeviceCounters>(), s);
    mark(6);
    if (fsBound_ && fs_.desc.kind != IFHIP_FS_DEPTH_ONLY) {
        ShadeParams sh{};
        const IfritGpuFragmentProgramDesc& d = fs_.desc;
        sh.width = width_;
        sh.height = height_;
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
        sh.numTextures = (int)textures_.size();
        for (int k = 0; k < 3; k++) sh.camera[k] = d.camera[k];
        sh.ambient = d.ambient;
        sh.kd = d.kd;
        sh.ks = d.ks;
        sh.shininess = d.shininess;
        sh.numLights = std::min(4, std::max(0, d.numLights));
        std::memcpy(sh.lightPos, d.lightPos, sizeof(sh.lightPos));
        std::memcpy(sh.lightColor, d.lightColor, sizeof(sh.lightColor));
        const bool wantFloat = fb_ && !fb_->color.empty() && fb_->color[0];
        sh.writeFloat = wantFloat ? 1 : 0;
        sh.ntri = curTris_;
        sh.neural = neuralShadeParams();
        launch_shade(sh, recPtrs(), bary_.as<float4>(), visSlot_.as<uint32_t>(), vtx_.as<float>(),
                     idx_.as<int>(), texDesc_.as<TexDesc>(), nullptr, rgba8_.as<uchar4>(),
                     wantFloat ? rgba32f_.as<float4>() : nullptr, s);
    }
    mark(7);
}

bool TileRasterRenderer::finish(bool clear, int indexCount) {
    (void)clear;
    (void)indexCount;
    // frame statistics are reduced from the per-workgroup partials only when read
    launch_finalize(statPtrs(), ctr_.as<DeviceCounters>(), stream_);
    if (!ok(hipStreamSynchronize(stream_), "frame")) return false;
    DeviceCounters c{};
    if (!ok(hipMemcpy(&c, ctr_.ptr, sizeof(c), hipMemcpyDeviceToHost), "counters")) return false;
    stats_.inputTriangles = curTris_;
    stats_.assembledTriangles = (int64_t)c.assembled;
    stats_.clippedTriangles = (int64_t)c.clipCount;
    stats_.binEntries = c.binTotal;
    stats_.fragments = (int64_t)c.fragments;
    stats_.coveredSamples = (int64_t)c.coveredSamples;
    stats_.status = c.status;
    if (stageTiming_) {
        for (int i = 0; i < 7; i++) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev_[i], ev_[i + 1]);
            stats_.stageMs[i] = ms;
        }
    }
    if (c.status & kStatusClipOverflow) clipCap_ = std::max(clipCap_ * 2, c.clipCount + 1024u);
    if (c.status & kStatusBinOverflow) binCap_ = std::max<unsigned>(binCap_ * 2, c.binTotal + (1u << 20));
    if (c.status & kStatusBigOverflow) bigCap_ = std::max(bigCap_ * 2, c.bigCount + 1024u);
    if (c.status & kStatusRasterOverflow) {
        itemCap_ = std::max(itemCap_ * 2, c.numItems + 1024u);
        partialCap_ = std::max(partialCap_ * 2, c.numPartials + 256u);
    }
    if (c.status & kStatusBadIndex) setError("index buffer references a vertex beyond the vertex buffer");
    return (c.status & (kStatusClipOverflow | kStatusBinOverflow | kStatusBigOverflow | kStatusRasterOverflow)) == 0;
}

void TileRasterRenderer::drawElements(int indexCount, bool clear) {
    stats_.retries = 0;
    for (int attempt = 0; attempt < 4; attempt++) {
        if (!prepare(indexCount)) return;
        enqueue(clear);
        if (!ok(hipGetLastError(), "kernel launch")) return;
        if (finish(clear, indexCount)) break;
        stats_.retries++;
        if (!clear) {
            setError("frame exceeded the binning buffers while drawing without clear");
            break;
        }
    }
    copyOutToHost();
}

void TileRasterRenderer::drawAsync(int indexCount, bool clear) {
    if (!prepare(indexCount)) return;
    enqueue(clear);
    ok(hipGetLastError(), "kernel launch");
}

void TileRasterRenderer::synchronize() {
    if (!stream_) return;
    finish(true, curTris_ * 3);
}

void TileRasterRenderer::copyOutToHost() {
    if (!fb_) return;
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

void TileRasterRenderer::getVisibility(uint32_t* ids, float* depth) {
    const size_t n = (size_t)width_ * height_ * samples_;
    if (ids) {
        launch_slot_to_prim(visSlot_.as<uint32_t>(), recPtrs(), n, visId_.as<uint32_t>(), stream_);
        if (!ok(hipStreamSynchronize(stream_), "id expansion")) return;
        ok(hipMemcpy(ids, visId_.ptr, n * 4, hipMemcpyDeviceToHost), "id readback");
    }
    if (depth) ok(hipMemcpy(depth, visDepth_.ptr, n * 4, hipMemcpyDeviceToHost), "depth readback");
}

void TileRasterRenderer::getCoverageClass(uint8_t* cls) {
    ok(hipMemcpy(cls, visClass_.ptr, (size_t)width_ * height_, hipMemcpyDeviceToHost), "class readback");
}

void TileRasterRenderer::getColorRGBA8(uint8_t* rgba) {
    ok(hipMemcpy(rgba, rgba8_.ptr, (size_t)width_ * height_ * 4, hipMemcpyDeviceToHost), "colour readback");
}

}  // namespace ifrit_hip
