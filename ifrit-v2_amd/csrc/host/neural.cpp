// Neural material: training orchestration and state of TileRasterRenderer.
//
// Training collects fragments of the bound layered-GGX material from random
// views (the shading pass writes one record per covered pixel: latent
// footprint, view and light directions in the tangent frame, target colour),
// then runs Adam steps of the MFMA training kernel over random batches, and
// finally packs the network into BF16 and FP8 operands (FP8 scales from a
// calibration pass) and the latent into FP16.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>

#include "renderer.hpp"

namespace ifrit_hip {

void neural_pack(const float* params, float* qinfo, NeuralPacked* out, int withFp8, hipStream_t s);
void neural_train_step(const NeuralRecord* data, unsigned numData, const NeuralLatent& latent, float* latentGrad,
                       const NeuralPacked* packed, float* gradPartial, float* grads, float* loss, unsigned step,
                       unsigned batch, unsigned chunksPerWave, hipStream_t s);
void neural_adam(float* p, float* g, float* m, float* v, size_t n, float lr, float b1, float b2, int t, hipStream_t s);
void neural_init_latent(float* lat, size_t n, unsigned seed, float a, hipStream_t s);
void neural_to_half(const float* src, void* dst, size_t n, hipStream_t s);
void neural_calibrate(const NeuralRecord* data, unsigned numData, const NeuralLatent& latent,
                      const NeuralPacked* packed, float* qinfo, unsigned samples, hipStream_t s);
void neural_eval(const float* in, float* out, int n, const NeuralPacked* packed, int fp8, hipStream_t s);
void neural_append(const NeuralRecord* feat, int npix, NeuralRecord* data, unsigned* count, unsigned cap,
                   hipStream_t s);

namespace {

// Perspective and look-at with the reference's conventions
// (include/ifrit/core/math/LinalgOps.h:83-149): left-handed, z in [0, 1],
// w = view-space z; rows s, u, f, translated by -eye.
void perspective(double fovy, double aspect, double n, double f, double m[16]) {
    std::fill(m, m + 16, 0.0);
    const double top = n * std::tan(fovy / 2.0), right = top * aspect;
    m[0] = n / right;
    m[5] = n / top;
    m[10] = f / (f - n);
    m[11] = -(f * n) / (f - n);
    m[14] = 1.0;
}

void look_at(const double eye[3], const double at[3], double m[16]) {
    double f[3] = {at[0] - eye[0], at[1] - eye[1], at[2] - eye[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    const double up[3] = {0, 1, 0};
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    const double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& x : s) x /= sl;
    const double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    const double* rows[3] = {s, u, f};
    std::fill(m, m + 16, 0.0);
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) m[4 * r + c] = rows[r][c];
        m[4 * r + 3] = -(rows[r][0] * eye[0] + rows[r][1] * eye[1] + rows[r][2] * eye[2]);
    }
    m[15] = 1.0;
}

int ilog2(int v) {
    int l = 0;
    while ((1 << (l + 1)) <= v) l++;
    return l;
}

}  // namespace

bool TileRasterRenderer::neuralLatentAlloc(int size) {
    if (size <= 0 || (size & (size - 1))) {
        setError("the latent size must be a power of two");
        return false;
    }
    NeuralLatent d{};
    d.size = size;
    d.levels = ilog2(size) + 1;
    unsigned off = 0;
    for (int l = 0; l < d.levels; l++) {
        d.levelOffset[l] = off;
        off += (unsigned)std::max(1, size >> l) * (unsigned)std::max(1, size >> l);
    }
    const size_t n = (size_t)off * kLatentCh;
    if (!(lat32_.ensure(n * 4) && latGrad_.ensure(n * 4) && latM_.ensure(n * 4) && latV_.ensure(n * 4) &&
          lat16_.ensure(n * 2))) {
        setError("hipMalloc latent");
        return false;
    }
    lat32Desc_ = d;
    lat32Desc_.base = lat32_.ptr;
    lat16Desc_ = d;
    lat16Desc_.base = lat16_.ptr;
    return true;
}

NeuralShadeParams TileRasterRenderer::neuralShadeParams() const {
    NeuralShadeParams p{};
    p.latent = lat16Desc_;
    p.packed = nnPacked_.as<NeuralPacked>();
    p.precision = fs_.desc.precision == 1 ? 1 : 0;
    p.lodShift = nnLodShift_;
    p.singleTap = nnSingleTap_;
    p.featOut = nnFeat_.as<float4>();
    return p;
}

bool TileRasterRenderer::neuralPackAll(unsigned dataCount) {
    hipStream_t s = stream_;
    float* qinfo = nnQinfo_.as<float>();
    if (!ok(hipMemsetAsync(qinfo, 0, 8 * sizeof(float), s), "qinfo")) return false;
    neural_pack(nnParams_.as<float>(), qinfo, nnPacked_.as<NeuralPacked>(), 0, s);
    if (dataCount > 0)
        neural_calibrate(nnData_.as<NeuralRecord>(), dataCount, lat32Desc_, nnPacked_.as<NeuralPacked>(), qinfo,
                         std::min(dataCount, 1u << 18), s);
    neural_pack(nnParams_.as<float>(), qinfo, nnPacked_.as<NeuralPacked>(), 1, s);
    const size_t n = (size_t)(lat32_.bytes / 4);
    neural_to_half(lat32_.as<float>(), lat16_.ptr, n, s);
    return ok(hipStreamSynchronize(s), "neural pack");
}

void TileRasterRenderer::trainNeural(const IfritNeuralTrainDesc& d, IfritNeuralTrainReport* report) {
    Init();
    if (!initialized_) return;
    if (!fsBound_ || fs_.desc.kind != IFHIP_FS_GGX_LAYERED) {
        setError("neural training needs the layered-GGX fragment program bound");
        return;
    }
    if (textures_.size() < 3 || stride_ < 11 || nidx_ <= 0) {
        setError("neural training needs the three material textures and vertices with tangents");
        return;
    }
    if (d.views <= 0 || d.viewWidth <= 0 || d.viewHeight <= 0 || d.maxFragments <= 0 || d.steps <= 0 ||
        d.batch < 256 || d.batch % 256 != 0) {
        setError("invalid neural training description");
        return;
    }
    if (!neuralLatentAlloc(d.latentSize)) return;
    nnLodShift_ = ilog2(textures_[0].width) - ilog2(d.latentSize);
    nnSingleTap_ = d.singleTap ? 1 : 0;
    const size_t npix = (size_t)d.viewWidth * d.viewHeight;
    if (!(nnParams_.ensure(kNumParams * 4) && nnGrads_.ensure(kNumParams * 4) && nnM_.ensure(kNumParams * 4) &&
          nnV_.ensure(kNumParams * 4) && nnPacked_.ensure(sizeof(NeuralPacked)) && nnQinfo_.ensure(8 * 4) &&
          nnData_.ensure((size_t)d.maxFragments * sizeof(NeuralRecord)) && nnFeat_.ensure(npix * sizeof(NeuralRecord)) &&
          nnCount_.ensure(4) && nnLoss_.ensure((size_t)d.steps * 4))) {
        setError("hipMalloc neural training buffers");
        return;
    }
    hipStream_t s = stream_;
    auto t0 = std::chrono::steady_clock::now();

    // 1. training fragments from random views
    FrameBuffer* fbSaved = fb_;
    std::vector<HostView> batchSaved;
    batchSaved.swap(batch_);  // training views are single-view draws
    const int wSaved = width_, hSaved = height_, sSaved = samples_;
    const IfritGpuVertexProgramDesc vsSaved = vs_.desc;
    const IfritGpuFragmentProgramDesc fsSaved = fs_.desc;
    fb_ = nullptr;
    width_ = d.viewWidth;
    height_ = d.viewHeight;
    samples_ = 1;
    fs_.desc.kind = 5;  // layered GGX plus one training record per pixel
    std::mt19937 rng(d.seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    if (!ok(hipMemsetAsync(nnCount_.ptr, 0, 4, s), "count")) return;
    unsigned collected = 0;
    for (int v = 0; v < d.views && error_.empty(); v++) {
        // eye uniform on a sphere shell, look-at point jittered around the origin
        const double z = 2.0 * uni(rng) - 1.0, phi = 2.0 * M_PI * uni(rng);
        const double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
        const double dist = d.eyeDistMin + (d.eyeDistMax - d.eyeDistMin) * uni(rng);
        const double eye[3] = {dist * rr * std::cos(phi), dist * z, dist * rr * std::sin(phi)};
        const double at[3] = {d.targetJitter * (2 * uni(rng) - 1), d.targetJitter * (2 * uni(rng) - 1),
                              d.targetJitter * (2 * uni(rng) - 1)};
        // field of view log-uniform in [fovYMin, fovY]: pixel footprints from the framing view to close-ups
        const double fmin = d.fovYMin > 0 ? std::min((double)d.fovYMin, (double)d.fovY) : d.fovY;
        const double fov = std::exp(std::log(fmin) + (std::log((double)d.fovY) - std::log(fmin)) * uni(rng));
        double P[16], V[16];
        perspective(fov, (double)d.viewWidth / d.viewHeight, 0.1, 200.0, P);
        look_at(eye, at, V);
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) {
                double acc = 0;
                for (int k = 0; k < 4; k++) acc += P[4 * r + k] * V[4 * k + c];
                vs_.desc.mvp[4 * r + c] = (float)acc;
            }
        for (int k = 0; k < 3; k++) fs_.desc.camera[k] = (float)eye[k];
        drawElements(nidx_, true);
        neural_append(nnFeat_.as<NeuralRecord>(), (int)npix, nnData_.as<NeuralRecord>(), nnCount_.as<unsigned>(),
                      (unsigned)d.maxFragments, s);
        if ((v & 15) == 15 || v == d.views - 1) {
            if (!ok(hipMemcpyAsync(&collected, nnCount_.ptr, 4, hipMemcpyDeviceToHost, s), "count") ||
                !ok(hipStreamSynchronize(s), "count"))
                break;
            if (collected >= (unsigned)d.maxFragments) break;
        }
    }
    fb_ = fbSaved;
    batch_.swap(batchSaved);
    width_ = wSaved;
    height_ = hSaved;
    samples_ = sSaved;
    vs_.desc = vsSaved;
    fs_.desc = fsSaved;
    if (!error_.empty()) return;
    const unsigned numData = std::min(collected, (unsigned)d.maxFragments);
    if (numData < 1024) {
        setError("too few training fragments");
        return;
    }
    auto t1 = std::chrono::steady_clock::now();

    // 2. initial parameters: Xavier-uniform weights, zero biases, small uniform latent
    std::vector<float> params(kNumParams, 0.0f);
    auto xavier = [&](int off, int rows, int cols, int ld, int fanIn, int fanOut) {
        const double a = std::sqrt(6.0 / (fanIn + fanOut));
        for (int r = 0; r < rows; r++)
            for (int c = 0; c < cols; c++) params[off + r * ld + c] = (float)(a * (2 * uni(rng) - 1));
    };
    xavier(kW1, kMlpIn, 64, 64, kMlpIn, 64);
    xavier(kW2, 64, 64, 64, 64, 64);
    xavier(kW3, 64, 3, 4, 64, 3);
    if (!ok(hipMemcpyAsync(nnParams_.ptr, params.data(), kNumParams * 4, hipMemcpyHostToDevice, s), "params")) return;
    const size_t nlat = lat32_.bytes / 4;
    neural_init_latent(lat32_.as<float>(), nlat, d.seed ^ 0x5bd1e995u, 0.1f, s);
    if (!(ok(hipMemsetAsync(nnGrads_.ptr, 0, kNumParams * 4, s), "grads") &&
          ok(hipMemsetAsync(nnM_.ptr, 0, kNumParams * 4, s), "m") && ok(hipMemsetAsync(nnV_.ptr, 0, kNumParams * 4, s), "v") &&
          ok(hipMemsetAsync(latGrad_.ptr, 0, nlat * 4, s), "latent grads") &&
          ok(hipMemsetAsync(latM_.ptr, 0, nlat * 4, s), "latent m") &&
          ok(hipMemsetAsync(latV_.ptr, 0, nlat * 4, s), "latent v") &&
          ok(hipMemsetAsync(nnLoss_.ptr, 0, (size_t)d.steps * 4, s), "loss") &&
          ok(hipMemsetAsync(nnQinfo_.ptr, 0, 32, s), "qinfo")))
        return;

    // 3. Adam steps, cosine learning-rate decay to 5 %
    const unsigned chunksPerWave = 1;
    const unsigned batch = (unsigned)d.batch / (64u * chunksPerWave) * (64u * chunksPerWave);
    if (!nnGradPart_.ensure((size_t)(batch / (64u * chunksPerWave)) * kNumParams * 4)) {
        setError("hipMalloc gradient partials");
        return;
    }
    for (int step = 1; step <= d.steps; step++) {
        const float decay = 0.05f + 0.95f * 0.5f * (1.0f + std::cos((float)M_PI * (step - 1) / d.steps));
        neural_pack(nnParams_.as<float>(), nnQinfo_.as<float>(), nnPacked_.as<NeuralPacked>(), 0, s);
        neural_train_step(nnData_.as<NeuralRecord>(), numData, lat32Desc_, latGrad_.as<float>(),
                          nnPacked_.as<NeuralPacked>(), nnGradPart_.as<float>(), nnGrads_.as<float>(),
                          nnLoss_.as<float>() + (step - 1),
                          (unsigned)step, batch, chunksPerWave, s);
        neural_adam(nnParams_.as<float>(), nnGrads_.as<float>(), nnM_.as<float>(), nnV_.as<float>(), kNumParams,
                    d.lrMlp * decay, 0.9f, 0.99f, step, s);
        neural_adam(lat32_.as<float>(), latGrad_.as<float>(), latM_.as<float>(), latV_.as<float>(), nlat,
                    d.lrLatent * decay, 0.9f, 0.99f, step, s);
        if (!ok(hipGetLastError(), "training launch")) return;
    }
    std::vector<float> loss((size_t)d.steps);
    if (!ok(hipMemcpyAsync(loss.data(), nnLoss_.ptr, loss.size() * 4, hipMemcpyDeviceToHost, s), "loss") ||
        !ok(hipStreamSynchronize(s), "training"))
        return;
    auto t2 = std::chrono::steady_clock::now();

    // 4. inference operands
    if (!neuralPackAll(numData)) return;
    nnReady_ = true;
    nnFallback_[0] = nnFallback_[1] = false;
    if (report) {
        report->fragments = numData;
        report->lossFirst = loss.front() / (3.0f * batch);
        report->lossLast = loss.back() / (3.0f * batch);
        report->datasetMs = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();
        report->trainMs = (float)std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
}

void TileRasterRenderer::setNeuralParams(const float* params) {
    Init();
    if (!initialized_) return;
    if (!(nnParams_.ensure(kNumParams * 4) && nnPacked_.ensure(sizeof(NeuralPacked)) && nnQinfo_.ensure(32))) {
        setError("hipMalloc neural parameters");
        return;
    }
    if (!lat32_.ptr) {
        if (!neuralLatentAlloc(16) || !ok(hipMemset(lat32_.ptr, 0, lat32_.bytes), "latent")) return;
    }
    if (!ok(hipMemcpy(nnParams_.ptr, params, kNumParams * 4, hipMemcpyHostToDevice), "params")) return;
    neuralPackAll(0);
    // FP8 activation scales without a data set: inputs and activations taken as within +-4
    const float q[6] = {0, 0, 0, 4.0f, 4.0f, 4.0f};
    if (!ok(hipMemcpy(nnQinfo_.ptr, q, sizeof(q), hipMemcpyHostToDevice), "qinfo")) return;
    neural_pack(nnParams_.as<float>(), nnQinfo_.as<float>(), nnPacked_.as<NeuralPacked>(), 1, stream_);
    if (ok(hipStreamSynchronize(stream_), "neural pack")) nnReady_ = true;
    nnFallback_[0] = nnFallback_[1] = false;
}

void TileRasterRenderer::getNeuralParams(float* params) {
    if (!nnParams_.ptr) {
        setError("no neural parameters");
        return;
    }
    ok(hipMemcpy(params, nnParams_.ptr, kNumParams * 4, hipMemcpyDeviceToHost), "params");
}

void TileRasterRenderer::evalNeural(const float* in, float* out, int n, int precision) {
    if (!nnReady_) {
        setError("the neural material has not been trained or set");
        return;
    }
    if (n <= 0) return;
    if (!(nnEvalIn_.ensure((size_t)n * 64) && nnEvalOut_.ensure((size_t)n * 16))) {
        setError("hipMalloc neural evaluation");
        return;
    }
    if (!ok(hipMemcpy(nnEvalIn_.ptr, in, (size_t)n * 64, hipMemcpyHostToDevice), "eval in")) return;
    neural_eval(nnEvalIn_.as<float>(), nnEvalOut_.as<float>(), n, nnPacked_.as<NeuralPacked>(), precision == 1, stream_);
    if (!ok(hipStreamSynchronize(stream_), "eval")) return;
    ok(hipMemcpy(out, nnEvalOut_.ptr, (size_t)n * 16, hipMemcpyDeviceToHost), "eval out");
}

float TileRasterRenderer::validateNeural(int precision, float minPsnrDb, int indexCount) {
    IfritNeuralValidation v{};
    validateNeuralEx(precision, minPsnrDb, 255, indexCount, &v);
    return v.psnrDb;
}

void TileRasterRenderer::validateNeuralEx(int precision, float minPsnrDb, int maxErrLsb, int indexCount,
                                          IfritNeuralValidation* out) {
    IfritNeuralValidation res{};
    if (out) *out = res;
    if (!nnReady_) {
        setError("the neural material has not been trained or set");
        return;
    }
    if (!fsBound_ || (fs_.desc.kind != IFHIP_FS_NEURAL && fs_.desc.kind != IFHIP_FS_GGX_LAYERED)) {
        setError("validation needs the layered-GGX or neural program bound");
        return;
    }
    const int p = precision == 1 ? 1 : 0;
    if (!batch_.empty()) {
        setError("validation renders the single bound target: clear the view batch first");
        return;
    }
    const IfritGpuFragmentProgramDesc saved = fs_.desc;
    FrameBuffer* fbSaved = fb_;
    fb_ = nullptr;
    const size_t px = (size_t)width_ * height_;
    std::vector<uint8_t> plain(px * 4), neural(px * 4);
    std::vector<uint32_t> slot0(px);
    fs_.desc.kind = IFHIP_FS_GGX_LAYERED;
    drawElements(indexCount, true);
    getColorRGBA8(plain.data());
    {
        std::vector<uint32_t> slots(px * samples_);
        ok(hipMemcpy(slots.data(), visSlot_.ptr, slots.size() * 4, hipMemcpyDeviceToHost), "slots");
        for (size_t i = 0; i < px; i++) {
            uint32_t any = 0xffffffffu;
            for (int s = 0; s < samples_; s++) any &= slots[i * samples_ + s];
            slot0[i] = any;  // all samples empty only when every slot is empty
        }
    }
    fs_.desc.kind = IFHIP_FS_NEURAL;
    fs_.desc.precision = p;
    const bool fb = nnFallback_[p];
    nnFallback_[p] = false;
    drawElements(indexCount, true);
    getColorRGBA8(neural.data());
    nnFallback_[p] = fb;
    fs_.desc = saved;
    fb_ = fbSaved;
    if (!error_.empty()) return;
    // squared error over the RGB of covered pixels, and the distribution of each pixel's
    // largest channel error (in RGBA8 steps)
    double se = 0.0;
    size_t n = 0;
    unsigned long long hist[256] = {};
    for (size_t i = 0; i < px; i++) {
        if (slot0[i] == 0xffffffffu) continue;
        int worst = 0;
        for (int c = 0; c < 3; c++) {
            const int e = std::abs((int)neural[4 * i + c] - (int)plain[4 * i + c]);
            worst = std::max(worst, e);
            const double d = e / 255.0;
            se += d * d;
        }
        hist[worst]++;
        n += 3;
    }
    const double mse = n ? se / (double)n : 0.0;
    res.psnrDb = (float)(mse > 0 ? 10.0 * std::log10(1.0 / mse) : 120.0);
    res.pixels = (int64_t)(n / 3);
    for (int e = 255; e >= 0; e--)
        if (hist[e]) {
            res.maxErrLsb = e;
            break;
        }
    // the smallest error that 99.9 % of the covered pixels stay within
    unsigned long long acc = 0;
    const unsigned long long want = (unsigned long long)std::ceil(0.999 * (double)res.pixels);
    for (int e = 0; e < 256; e++) {
        acc += hist[e];
        if (acc >= want) {
            res.p999ErrLsb = e;
            break;
        }
    }
    res.fallback = (res.psnrDb < minPsnrDb || res.maxErrLsb > maxErrLsb) ? 1 : 0;
    nnFallback_[p] = res.fallback != 0;
    if (out) *out = res;
}

}  // namespace ifrit_hip
