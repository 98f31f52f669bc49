// The no-allocation draw (iftrDrawNoAlloc): the draw of the bound single
// target, from the caller's device-resident vertices and indices into the
// caller's device outputs, with every intermediate buffer carved out of one
// static device array.  It calls no hipMalloc or hipFree, so callers that
// cannot allocate (serving loops, harnesses that forbid it) can draw.
//
// The array `scratch` below is a device symbol: in host code its name is not
// a device address.  Its address on each device is looked up once with
// hipGetSymbolAddress (at iftrInit, with that device current) and only that
// address is used: as kernel arguments, for copies and for clears.  Every
// renderer on a GPU shares the GPU's array, and calls keep the last call's
// contents; the kernels clear what must start at zero, and calls that overlap
// on one GPU race.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "../common/x86_fp.h"
#include "renderer.hpp"

namespace ifrit_hip {

static __device__ __align__(256) unsigned char scratch[IFHIP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void* g_scratchAddr[kMaxDevices] = {};

constexpr int kTile = 16;

// Bump allocation of 256-byte aligned ranges inside the scratch array.
struct Arena {
    uint8_t* base;
    size_t cap;
    size_t used = 0;
    void* take(size_t n) {
        used = (used + 255) & ~(size_t)255;
        void* p = base + used;
        used += std::max<size_t>(n, 1);
        return p;
    }
    bool fits() const { return used <= cap; }
};

}  // namespace

// Looks up the scratch array's address on the current device (once per device).
bool noalloc_register_device(std::string* err) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *err = "no-allocation scratch: no current HIP device below index 64";
        return false;
    }
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchAddr[dev]) return true;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        *err = std::string("no-allocation scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e);
        return false;
    }
    g_scratchAddr[dev] = p;
    return true;
}

static void* noalloc_scratch_address(std::string* err) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *err = "no-allocation scratch: no current HIP device below index 64";
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> lock(g_scratchMutex);
        if (g_scratchAddr[dev]) return g_scratchAddr[dev];
    }
    // a device other than the one current at iftrInit: look it up now
    if (!noalloc_register_device(err)) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    return g_scratchAddr[dev];
}

void TileRasterRenderer::drawNoAlloc(const IfritNoAllocDraw& d) {
    if (!initialized_) return setError("iftrDrawNoAlloc: call iftrInit first");
    if (!d.vertices || !d.indices || !d.depth) return setError("iftrDrawNoAlloc: vertices, indices and depth are required");
    if (!batch_.empty()) return setError("iftrDrawNoAlloc draws the single bound target: clear the view batch first");
    if (width_ <= 0 || height_ <= 0) return setError("iftrDrawNoAlloc: no render target set");
    if (width_ > 16384 || height_ > 16384) return setError("iftrDrawNoAlloc: render target larger than 16384 x 16384");
    if (hostVs_ || !vsBound_) return setError("iftrDrawNoAlloc needs a GPU vertex program");
    if (blend_.blendEnable) return setError("iftrDrawNoAlloc does not blend");
    if (polygonMode_ != 0) return setError("iftrDrawNoAlloc draws filled triangles only");
    if (gsBound_) return setError("iftrDrawNoAlloc does not run geometry programs");
    if (depthFunc_ < IF_COMPARE_OP_NEVER || depthFunc_ > IF_COMPARE_OP_ALWAYS)
        return setError("unknown depth function");
    if (d.numIndices < 0 || d.numVertices <= 0 || d.strideFloats <= 0)
        return setError("iftrDrawNoAlloc: bad vertex or index count");
    if (vs_.desc.kind == IFHIP_VS_CLIP_PASSTHROUGH ? d.strideFloats < 4 : d.strideFloats < 3)
        return setError("iftrDrawNoAlloc: vertex records too short for the vertex program");
    const bool shade = d.rgba8 && fsBound_ && fs_.desc.kind != IFHIP_FS_DEPTH_ONLY;
    if (shade) {
        const int k = fs_.desc.kind;
        if (k == IFHIP_FS_NEURAL && (!nnReady_ || textures_.size() < 3))
            return setError("the neural material has not been trained or its textures are not bound");
        if ((k == IFHIP_FS_GGX_LAYERED || k == IFHIP_FS_NEURAL) &&
            (textures_.size() < 3 || d.strideFloats < 11 || fs_.desc.offTangent + 3 > d.strideFloats))
            return setError("the layered-GGX and neural programs need three textures and a tangent");
    }
    const int ntri = d.numIndices / 3;
    if (samples_ == 1 && ntri >= (1 << 28)) return setError("too many triangles at 1 sample per pixel");
    std::string err;
    uint8_t* base = static_cast<uint8_t*>(noalloc_scratch_address(&err));
    if (!base) return setError(err);

    // everything the draw reads from the renderer's own buffers and sizes, restored on every exit
    DevBuf* const bufs[] = {&vtx_, &idx_, &clip_, &recs_[0], &recs_[1], &recs_[2], &recs_[3], &recs_[4], &recs_[5],
                            &bary_, &binQueue_, &slotOf_, &binCount_, &binStart_, &binCursor_, &binList_, &bigList_,
                            &ctr_, &groupBounds_, &groupOut_, &clipList_, &statPart_, &items_, &tilePartial_,
                            &multiList_, &partials_, &visDepth_, &visSlot_, &visClass_, &rgba8_, &rgba32f_,
                            &texDesc_, &viewDev_, &tileView_, &shadeView_, &viewCtr_, &planSums_, &swapped_};
    constexpr size_t kBufs = sizeof(bufs) / sizeof(bufs[0]);
    struct Restore {
        TileRasterRenderer& r;
        DevBuf* const* bufs;
        DevBuf saved[kBufs];
        unsigned caps[5];
        int ints[5];
        unsigned tiles, shadeTiles;
        size_t pixels;
        std::vector<ViewDesc> views;
        FrameBuffer* fb;
        bool floatOut;
        uint8_t* pool;
        Restore(TileRasterRenderer& rr, DevBuf* const* b) : r(rr), bufs(b) {
            for (size_t i = 0; i < kBufs; i++) saved[i] = *bufs[i];
            const unsigned c[5] = {r.clipCap_, r.binCap_, r.bigCap_, r.itemCap_, r.partialCap_};
            std::memcpy(caps, c, sizeof(caps));
            const int n[5] = {r.curTris_, r.nverts_, r.stride_, r.nidx_, r.boundsTris_};
            std::memcpy(ints, n, sizeof(ints));
            tiles = r.totalTiles_;
            shadeTiles = r.shadeTiles_;
            pixels = r.totalPixels_;
            views = r.viewDesc_;
            fb = r.fb_;
            floatOut = r.floatOut_;
            pool = r.texPool_;
        }
        ~Restore() {
            for (size_t i = 0; i < kBufs; i++) *bufs[i] = saved[i];
            r.clipCap_ = caps[0];
            r.binCap_ = caps[1];
            r.bigCap_ = caps[2];
            r.itemCap_ = caps[3];
            r.partialCap_ = caps[4];
            r.curTris_ = ints[0];
            r.nverts_ = ints[1];
            r.stride_ = ints[2];
            r.nidx_ = ints[3];
            r.boundsTris_ = ints[4];
            r.totalTiles_ = tiles;
            r.shadeTiles_ = shadeTiles;
            r.totalPixels_ = pixels;
            r.viewDesc_ = views;
            r.fb_ = fb;
            r.floatOut_ = floatOut;
            r.texPool_ = pool;
        }
    } restore(*this, bufs);

    // sizes, as prepare() and layoutViews() set them for one view
    curTris_ = ntri;
    nverts_ = d.numVertices;
    stride_ = d.strideFloats;
    nidx_ = d.numIndices;
    boundsTris_ = vs_.desc.kind == IFHIP_VS_TRANSFORM ? ntri : 0;
    fb_ = nullptr;
    floatOut_ = false;
    clipCap_ = std::max(4096u, (unsigned)(ntri / 8));
    const unsigned part = 1u << partShift_;
    binCap_ = std::max(1u << 20, (unsigned)std::min<size_t>(0xFFFFFFF0ull, (size_t)ntri * 6));
    bigCap_ = std::max(1024u, (unsigned)(ntri / 16));
    const int tilesX = (width_ + kTile - 1) / kTile, tilesY = (height_ + kTile - 1) / kTile;
    totalTiles_ = (unsigned)(tilesX * tilesY);
    shadeTiles_ = shade ? (unsigned)(((width_ + shadeTile_ - 1) / shadeTile_) * ((height_ + shadeTile_ - 1) / shadeTile_)) : 0u;
    totalPixels_ = ((size_t)width_ * height_ + 63) & ~(size_t)63;
    itemCap_ = totalTiles_ + binCap_ / part + 1u;
    partialCap_ = std::max(256u, binCap_ / (8u * part));
    ViewDesc vd;
    std::memset(&vd, 0, sizeof(vd));
    std::memcpy(vd.mvp, vs_.desc.mvp, sizeof(vd.mvp));
    for (int k = 0; k < 3; k++) vd.camera[k] = fs_.desc.camera[k];
    vd.shade = shade ? 1 : 0;
    vd.width = width_;
    vd.height = height_;
    vd.tilesX = tilesX;
    vd.tilesY = tilesY;
    vd.invW = 1.0f / (float)width_;  // TileRasterRenderer.cpp:59-60
    vd.invH = 1.0f / (float)height_;
    viewDesc_.assign(1, vd);

    // the scratch layout: every buffer of the draw except the caller's inputs and outputs
    const size_t slots = (size_t)ntri + clipCap_, tiles = totalTiles_, px = (size_t)width_ * height_;
    const size_t groups = (size_t)(boundsTris_ + kGroupTris - 1) / kGroupTris;
    Arena A{base, (size_t)IFHIP_NOALLOC_SCRATCH_BYTES};
    auto put = [&](DevBuf& b, size_t n) {
        b.ptr = A.take(n);
        b.bytes = n;
    };
    put(ctr_, sizeof(DeviceCounters));
    put(viewDev_, sizeof(ViewDesc));
    put(viewCtr_, sizeof(ViewCounters));
    put(tileView_, std::max<size_t>(1, tiles) * 2);
    put(shadeView_, std::max<size_t>(1, shadeTiles_) * 2);
    put(texDesc_, std::max<size_t>(1, textures_.size()) * sizeof(TexDesc));
    put(groupBounds_, std::max<size_t>(1, groups) * 32);
    put(groupOut_, std::max<size_t>(1, groups));
    put(clip_, (size_t)nverts_ * 16);
    for (int k = 0; k < 6; k++) put(recs_[k], slots * 16);
    put(bary_, slots * 32);
    put(binQueue_, slots * 16);
    put(slotOf_, (size_t)std::max(1, ntri) * 4);
    if (needSwapFlags()) put(swapped_, slots);
    put(clipList_, (size_t)clipCap_ * 4);
    put(statPart_, ((size_t)(ntri + 1023) / 1024 + 2 * (size_t)rasterGrid() + 2) * 8);
    put(items_, (size_t)itemCap_ * 16);
    put(tilePartial_, tiles * 4);
    put(planSums_, ((tiles + kPlanTiles - 1) / kPlanTiles + 1) * 6 * 4);
    put(multiList_, tiles * 8);
    put(partials_, (size_t)partialCap_ * 256 * samples_ * 8);
    put(binCount_, tiles * 4);
    put(binStart_, tiles * 4);
    put(binCursor_, tiles * 4);
    put(binList_, (size_t)binCap_ * 4);
    put(bigList_, (size_t)bigCap_ * 4);
    put(visSlot_, px * samples_ * 4);
    put(visClass_, px);
    if (!A.fits())
        return setError("iftrDrawNoAlloc: the draw needs " + std::to_string(A.used) + " bytes of scratch; the "
                        "no-allocation entry point holds " + std::to_string((size_t)IFHIP_NOALLOC_SCRATCH_BYTES));
    // the caller's inputs and outputs
    vtx_.ptr = const_cast<float*>(d.vertices);
    vtx_.bytes = (size_t)d.numVertices * d.strideFloats * 4;
    idx_.ptr = const_cast<int32_t*>(d.indices);
    idx_.bytes = (size_t)d.numIndices * 4;
    visDepth_.ptr = d.depth;
    visDepth_.bytes = px * samples_ * 4;
    rgba8_.ptr = d.rgba8;
    rgba8_.bytes = d.rgba8 ? px * 4 : 0;
    rgba32f_ = DevBuf{};

    // tables: the view, its tile maps (all view 0), the texture descriptors
    if (!ok(hipStreamSynchronize(stream_), "iftrDrawNoAlloc")) return;
    std::vector<TexDesc> td(std::max<size_t>(1, textures_.size()));
    texPool_ = nullptr;
    for (const TextureSlot& t : textures_)
        if (!texPool_ || t.dev < texPool_) texPool_ = t.dev;
    for (size_t i = 0; i < textures_.size(); i++) {
        const TextureSlot& t = textures_[i];
        td[i] = TexDesc{};
        td[i].base = (unsigned long long)(t.dev - texPool_);
        td[i].width = t.width;
        td[i].height = t.height;
        td[i].levels = t.levels;
        td[i].layers = t.layers;
    }
    if (!ok(hipMemcpy(viewDev_.ptr, &vd, sizeof(vd), hipMemcpyHostToDevice), "iftrDrawNoAlloc view table") ||
        !ok(hipMemcpy(texDesc_.ptr, td.data(), td.size() * sizeof(TexDesc), hipMemcpyHostToDevice),
            "iftrDrawNoAlloc texture table") ||
        !ok(hipMemsetAsync(tileView_.ptr, 0, tileView_.bytes, stream_), "iftrDrawNoAlloc tile map") ||
        !ok(hipMemsetAsync(shadeView_.ptr, 0, shadeView_.bytes, stream_), "iftrDrawNoAlloc shading map"))
        return;
    if (boundsTris_ > 0)
        launch_group_bounds(vtx_.as<float>(), stride_, nverts_, idx_.as<int>(), ntri, groupBounds_.as<float4>(), stream_);
    enqueue(true);
    if (d.primIds && px)
        launch_slot_to_prim(visSlot_.as<uint32_t>(), recPtrs(), px * samples_, d.primIds, stream_);
    if (!ok(hipGetLastError(), "kernel launch")) return;
    if (!finish(true, d.numIndices))
        setError("iftrDrawNoAlloc: the frame overflowed its scratch buffers (clip, bin or raster queue); the "
                 "outputs are incomplete");
}

}  // namespace ifrit_hip
