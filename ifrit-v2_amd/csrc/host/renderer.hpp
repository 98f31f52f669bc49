// Host side of the MI355X tile rasterizer: the objects the reference's API
// passes around (images, frame buffers, vertex buffers, varying descriptors,
// buffers) and TileRasterRenderer, which keeps the method names of
// include/ifrit/softgraphics/engine/tileraster/TileRasterRenderer.h:86-104 and
// the CUDA renderer's additions (TileRasterRendererCuda.h:81-115) and runs
// every draw as a sequence of HIP kernels on one stream.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/ifrit_hip.h"
#include "../kernels/pipeline.h"

namespace ifrit_hip {

// Core::Data::ImageF32 (include/ifrit/softgraphics/core/data/Image.h:101-115):
// row-major, channels interleaved.
struct ImageF32 {
    size_t width = 0, height = 0, channel = 0;
    std::vector<float> data;
};

// FrameBuffer (include/ifrit/softgraphics/engine/base/FrameBuffer.h:33-56).
struct FrameBuffer {
    std::vector<ImageF32*> color;
    ImageF32* depth = nullptr;
    int width = 0, height = 0;
};

// VertexBuffer (include/ifrit/softgraphics/engine/base/VertexBuffer.h:33-90):
// interleaved attributes with byte offsets.
struct VertexBuffer {
    std::vector<int> layout;   // TypeDescriptorEnum per attribute
    std::vector<int> offsets;  // byte offset per attribute
    int elementSize = 0;
    int vertexCount = 0;
    std::vector<uint8_t> buffer;
};

struct VaryingDescriptor {
    std::vector<int> types;
};

// TrivialBufferManager / IfritBuffer (include/ifrit/softgraphics/engine/bufferman/BufferManager.h).
struct BufferManager;
struct Buffer {
    BufferManager* manager = nullptr;
    std::vector<uint8_t> data;
};
struct BufferManager {
    std::vector<Buffer*> buffers;
};

constexpr uint32_t kVertexProgramTag = 0x49465653u;    // "IFVS"
constexpr uint32_t kFragmentProgramTag = 0x49464653u;  // "IFFS"
constexpr uint32_t kMeshProgramTag = 0x49464d53u;      // "IFMS"
constexpr uint32_t kTaskProgramTag = 0x49465453u;      // "IFTS"
constexpr uint32_t kGeometryProgramTag = 0x49464753u;  // "IFGS"

struct GpuVertexProgram {
    uint32_t tag = kVertexProgramTag;
    IfritGpuVertexProgramDesc desc;
};
struct GpuFragmentProgram {
    uint32_t tag = kFragmentProgramTag;
    IfritGpuFragmentProgramDesc desc;
};
struct GpuMeshProgram {
    uint32_t tag = kMeshProgramTag;
    IfritGpuMeshProgramDesc desc;
};
struct GpuTaskProgram {
    uint32_t tag = kTaskProgramTag;
    IfritGpuTaskProgramDesc desc;
};
struct GpuGeometryProgram {
    uint32_t tag = kGeometryProgramTag;
    IfritGpuGeometryProgramDesc desc;
};

struct TextureSlot {
    int width = 0, height = 0, levels = 0, layers = 1;
    uint8_t* dev = nullptr;
    std::vector<size_t> levelOffset;  // texel offset of each level's first layer
    size_t bytes = 0;
    size_t levelTexels(int l) const {  // texels of one layer of level l
        return (size_t)std::max(1, width >> l) * (size_t)std::max(1, height >> l);
    }
};

// CreateSampler's state (IfritSamplerT, Structures.h:52-61).
struct SamplerObject {
    SamplerMode mode;
    bool anisotropy = false;
    float maxAnisotropy = 1.0f;
};

// Grow-only device allocation.  Every (re)allocation bumps `generation`, so
// that captured graphs holding old addresses are recaptured.
struct DevBuf {
    static unsigned long long generation;
    void* ptr = nullptr;
    size_t bytes = 0;
    bool ensure(size_t n);
    void release();
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// Looks up the no-allocation scratch array on the current device (noalloc.hip).
bool noalloc_register_device(std::string* err);

class TileRasterRenderer {
public:
    TileRasterRenderer();
    ~TileRasterRenderer();

    // reference API (TileRasterRenderer.h:86-104)
    void Init();
    void bindFrameBuffer(FrameBuffer& fb);
    void bindVertexBuffer(const VertexBuffer& vb);
    void bindIndexBuffer(const Buffer& ib);
    void bindVertexShaderFunc(IfritVertexShaderFunctionalPtr f, VaryingDescriptor* vd);
    void bindFragmentShaderFunc(IfritFragmentShaderFunctionalPtr f);
    void bindVertexShader(const GpuVertexProgram& p);
    void bindFragmentShader(const GpuFragmentProgram& p);
    void setBlendFunc(const IfritColorAttachmentBlendState& s);
    void SetDepthFunc(int func);
    void optsetForceDeterministic(bool opt) { deterministic_ = opt; }
    bool fsWritesDepth() const { return fsBound_ && fs_.desc.depthWrite != 0; }
    void optSetDepthTestEnable(bool opt);
    void SetCullMode(int mode);  // TileRasterRendererCuda.h:109
    void setSetupReciprocal(int model);  // IFHIP_RCP_*: whose RCPPS the triangle setup reproduces
    void setRasterizerPolygonMode(int mode);  // TileRasterRendererCuda.h:106: 0 fill, 1 line, 2 point
    // device buffer slots and mesh work groups (TileRasterRendererCuda.h:88-89, 98-99, 115)
    int createBuffer(int slot, int size);
    void copyHostBufferToBuffer(const void* src, int slot, int size);
    void bindMeshShader(const GpuMeshProgram* p);
    void bindTaskShader(const GpuTaskProgram* p);
    void drawMeshTasks(int numGroups, int firstGroup, bool clear);
    void bindGeometryShader(const GpuGeometryProgram* p);  // TileRasterRendererCuda.h:87
    void setClearValues(const IfritVector4f* colors, size_t n, float depth);  // TileRasterRendererCuda.h:111
    void SetScissors(const IfritVector4f* rects, size_t n);                    // TileRasterRendererCuda.h:101
    void setScissorTestEnable(bool on) { scissorTest_ = on; }                  // TileRasterRendererCuda.h:102
    void drawElements(int indexCount, bool clear);
    void clear();

    // CUDA renderer additions and backend extensions
    void setMsaaSamples(int s);
    void bindVertexDataFloat(const float* data, int strideFloats, int numVertices);
    void bindIndexData(const int32_t* data, int numIndices);
    void setRenderTarget(int width, int height);
    int createTexture(int w, int h, int levels, int layers = 1);
    void copyHostBufferToImageLayer(int tex, int layer, const void* rgba8);
    void blitImage(int src, int dst, const IfritImageBlit& region, int filter);
    int CreateSampler(const IfritSamplerDesc& d);
    void copyHostBufferToImage(int tex, const void* rgba8);
    void uploadTextureLevels(int tex, const void* rgba8);
    void generateFbmTexture(int tex, uint32_t seed);
    void generateMipmap(int tex);
    void downloadTexture(int tex, void* rgba8);
    void getVisibility(uint32_t* ids, float* depth);
    void getCoverageClass(uint8_t* cls);
    void getColorRGBA8(uint8_t* rgba);
    void getColorRGBA32F(float* rgba);
    void setColorFloatOutput(bool on) { floatOut_ = on; }
    void getFrameStats(IfritFrameStats* out) const { *out = stats_; }
    void setStageTiming(bool on) { stageTiming_ = on; }
    void drawAsync(int indexCount, bool clear);
    void synchronize();
    const char* lastError();
    void setError(const std::string& e);
    bool ok(hipError_t e, const char* what);

    // multi-view batches, graph replay and tuning (extensions)
    void setViews(const IfritViewDesc* views, int n);
    void getViewVisibility(int view, uint32_t* ids, float* depth);
    void getViewColorRGBA8(int view, uint8_t* rgba);
    void setGraphMode(bool on);
    void setTuning(const IfritTuning& t);
    void getTuning(IfritTuning* t) const;
    void autotune(int indexCount, int frames, IfritAutotuneReport* out);
    uint64_t digest();

    // the draw with every intermediate buffer in the static scratch array (noalloc.hip)
    void drawNoAlloc(const IfritNoAllocDraw& d);

    // neural material (neural.cpp)
    void trainNeural(const IfritNeuralTrainDesc& d, IfritNeuralTrainReport* report);
    void setNeuralParams(const float* params);
    void getNeuralParams(float* params);
    void evalNeural(const float* in, float* out, int n, int precision);
    float validateNeural(int precision, float minPsnrDb, int indexCount);
    void validateNeuralEx(int precision, float minPsnrDb, int maxErrLsb, int indexCount, IfritNeuralValidation* out);

private:
    bool prepare(int indexCount);
    bool layoutViews(int ntri);
    void enqueue(bool clear);
    void enqueuePolygons(bool clear, const Views& V);
    void submit(bool clear);
    std::vector<uint8_t> graphKey(bool clear) const;
    void dropGraph();
    Views deviceViews() const;
    ShadeParams shadeParams(bool writeFloat) const;
    bool finish(bool clear, int indexCount);
    void copyOutToHost();
    void runHostVertexShader();

    hipStream_t stream_ = nullptr;
    bool initialized_ = false;
    bool deterministic_ = true;
    int depthFunc_ = IF_COMPARE_OP_LESS;
    int depthFuncSaved_ = IF_COMPARE_OP_LESS;
    bool depthTest_ = true;
    int cullMode_ = kCullBack;  // the CPU path's fixed back-face cull
    int rcpModel_ = 0;          // ifx::RcpModel of the setup's reciprocals
    int polygonMode_ = 0;       // IfritPolygonMode: 0 fill, 1 line, 2 point
    float clearColor_[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // TileRasterWorker.cpp:113-118
    float clearDepth_ = 3.3961514e38f;                  // 0x7f7f7f7f, TileRasterWorker.cpp:40
    std::vector<IfritVector4f> scissors_;  // x, y, width, height in pixels
    DevBuf scissorDev_;                    // the same on the device, read by the raster and shading kernels
    bool scissorTest_ = false;
    Scissor activeScissor() const;
    bool needSwapFlags() const { return cullMode_ == kCullNone || cullMode_ == kCullFront; }
    IfritColorAttachmentBlendState blend_{};
    BlendCoefs blendCoefs_{};
    bool floatOut_ = false;  // shading also writes RGBA32F
    int samples_ = 1;
    int width_ = 0, height_ = 0;
    FrameBuffer* fb_ = nullptr;

    // inputs
    const VertexBuffer* hostVb_ = nullptr;
    const Buffer* hostIb_ = nullptr;
    bool vtxDirty_ = true, idxDirty_ = true;
    bool boundsDirty_ = true;  // group bounds to recompute (geometry changed)
    int boundsTris_ = 0;       // triangles the group bounds cover
    int nverts_ = 0, stride_ = 0, nidx_ = 0;
    IfritVertexShaderFunctionalPtr hostVs_ = nullptr;
    VaryingDescriptor* hostVsVary_ = nullptr;
    bool hostFs_ = false;
    GpuVertexProgram vs_{};
    GpuFragmentProgram fs_{};
    bool vsBound_ = false, fsBound_ = false;

    // views: batch_ empty = one view from the bound target, vertex and fragment programs
    struct HostView {
        int width, height;
        float mvp[16];
        float camera[3];
        int shade;
    };
    std::vector<HostView> batch_;
    std::vector<ViewDesc> viewDesc_;        // device table as last uploaded
    std::vector<int> viewLayout_;           // (width, height, shade) per view + shading tile
    unsigned totalTiles_ = 0, shadeTiles_ = 0;
    size_t totalPixels_ = 0;
    DevBuf viewDev_, tileView_, shadeView_, viewCtr_, digest_, planSums_;

    // launch parameters (IfritTuning) and graph replay
    int partShift_ = kRasterPartShift;
    int rasterGroupsPerCU_ = 4;
    int shadeTile_ = 8;
    int bigThreshold_ = 64;
    int tileThreads_ = 256;   // raster threads per tile at 1 sample per pixel
    bool binCapSet_ = false;  // binCap_ given by setTuning: no default floor
    bool graphMode_ = false;
    bool capturing_ = false;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graphExec_ = nullptr;
    std::vector<uint8_t> graphKey_;

    // device buffers
    DevBuf vtx_, idx_, clip_, recs_[6], bary_, binQueue_, slotOf_, binCount_, binStart_, binCursor_, binList_, bigList_, ctr_;
    DevBuf groupBounds_, groupOut_;  // per triangle group: bounds; per view and group: outside flag
    DevBuf swapped_;                 // per record slot: set up with vertices 0 and 2 exchanged
    DevBuf polyKeys_;                // point and line modes: per pixel the nearest candidate's key
    static constexpr int kBufferSlots = 64;
    DevBuf bufSlots_[kBufferSlots];  // CreateBuffer slots
    int bufSize_[kBufferSlots] = {};
    DevBuf meshVtx_, meshIdx_;       // the emitted vertices of DrawMeshTasks and geometry programs
    GpuMeshProgram ms_{};
    GpuTaskProgram ts_{};
    GpuGeometryProgram gs_{};
    bool msBound_ = false, tsBound_ = false, gsBound_ = false;
    int clipAt_ = -1;  // >= 0 while an emitted list is drawn: the clip position's float offset in a record
    // Draws meshVtx_ / meshIdx_ (nv records of `stride` floats, clip position at clipAt, ni indices)
    // with the bound pipeline state, then restores the bound vertex and index inputs.
    void drawEmitted(int nv, int stride, int clipAt, int ni, bool clear, bool async);
    void drawGeometry(int indexCount, bool clear, bool async);
    bool uploadInputs();
    DevBuf clipList_, statPart_, items_, tilePartial_, multiList_, partials_, visId_, visDepth_, visSlot_, visClass_, rgba8_, rgba32f_, texDesc_;
    unsigned clipCap_ = 0, binCap_ = 0, bigCap_ = 0, itemCap_ = 0, partialCap_ = 0;
    int numCUs_ = 256;
    int curTris_ = 0;
    int curIdx_ = 0;  // indices of the current draw (the point mode draws every index)
    TriRecs recPtrs() const;
    StatPartials statPtrs() const;
    int rasterGrid() const;
    std::vector<TextureSlot> textures_;
    std::vector<SamplerObject> samplers_;
    bool texDescDirty_ = true;
    uint8_t* texPool_ = nullptr;  // lowest texture allocation: TexDesc bases are byte offsets from it

    // neural material state
    DevBuf nnParams_, nnGrads_, nnM_, nnV_, nnPacked_, nnQinfo_, lat32_, latGrad_, latM_, latV_, lat16_;
    DevBuf nnData_, nnFeat_, nnCount_, nnLoss_, nnEvalIn_, nnEvalOut_, nnGradPart_;
    NeuralLatent lat16Desc_{}, lat32Desc_{};
    int nnLodShift_ = 0;
    int nnSingleTap_ = 0;
    bool nnFallback_[2] = {false, false};  // per precision: below the PSNR bound, shade with the material
    bool nnReady_ = false;
    bool neuralLatentAlloc(int size);
    bool neuralPackAll(unsigned dataCount);
    NeuralShadeParams neuralShadeParams() const;

    IfritFrameStats stats_{};
    bool stageTiming_ = false;
    hipEvent_t ev_[8] = {};  // stage boundaries: vertex (with counter reset), setup, bin, raster, merge, shade
    std::string error_;
    std::string errorOut_;
};

}  // namespace ifrit_hip
