// Neural material on gfx950: operand packing (BF16 and FP8), training
// (forward and data gradients on MFMA, weight gradients in FP32), Adam,
// latent conversion, FP8 calibration, feature collection and a stand-alone
// MLP evaluation.
//
// Training step, one wave per 64 fragments of a random batch:
//   1. latent footprint of each fragment sampled from the FP32 master latent;
//   2. forward pass on MFMA (BF16 operands, FP32 accumulation), ReLU outputs
//      kept in LDS;
//   3. loss gradient 2 (y - t) / batch;
//   4. data gradients dH2 = W3 dY, dH1 = W2 dZ2, dX = W1 dZ1 on MFMA (the
//      accumulators of each product are the next product's B operand);
//   5. weight gradients as FP32 outer products over the wave's fragments (a
//      lane owns one row of each weight matrix);
//   6. latent gradients scattered to the footprint texels with atomics.
// The weight gradients of a wave are added to the global gradient once per
// step; Adam then updates every parameter and clears its gradient.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "neural_device.h"

namespace ifrit_hip {

namespace {

using namespace nn;

// ---------------------------------------------------------------------------
// Operand packing
// ---------------------------------------------------------------------------
__device__ __forceinline__ float w1(const float* p, int k, int u) { return k < kMlpIn ? p[kW1 + k * 64 + u] : 0.0f; }
__device__ __forceinline__ float w2(const float* p, int u, int m) { return p[kW2 + u * 64 + m]; }
__device__ __forceinline__ float w3(const float* p, int m, int o) { return o < kMlpOut ? p[kW3 + m * 4 + o] : 0.0f; }

// Per-layer max |W| into absmax[0..2] (one workgroup).
__global__ void __launch_bounds__(256) k_absmax(const float* __restrict__ p, float* __restrict__ absmax) {
    __shared__ float red[3][256];
    float m[3] = {0, 0, 0};
    for (int i = threadIdx.x; i < 1024; i += 256) m[0] = fmaxf(m[0], fabsf(p[kW1 + i]));
    for (int i = threadIdx.x; i < 4096; i += 256) m[1] = fmaxf(m[1], fabsf(p[kW2 + i]));
    for (int i = threadIdx.x; i < 256; i += 256) m[2] = fmaxf(m[2], fabsf(p[kW3 + i]));
    for (int k = 0; k < 3; k++) red[k][threadIdx.x] = m[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int k = 0; k < 3; k++) red[k][threadIdx.x] = fmaxf(red[k][threadIdx.x], red[k][threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x < 3) absmax[threadIdx.x] = red[threadIdx.x][0];
}

// qinfo: [0..2] max |W| per layer, [3..5] max activation (inputs, H1, H2).
__global__ void __launch_bounds__(256) k_pack(const float* __restrict__ p, const float* __restrict__ qinfo,
                                              NeuralPacked* __restrict__ out, int withFp8) {
    const int tid = blockIdx.x * 256 + threadIdx.x;
    const int nthreads = gridDim.x * 256;
    float sc[6];
    for (int k = 0; k < 6; k++) sc[k] = qinfo[k] > 0.0f ? 240.0f / qinfo[k] : 1.0f;
    if (tid < 6) out->qscale[tid] = sc[tid];
    // one item = (array, tile, step, lane): 8 values
    // items: A1 128, A2 512, A3 256, AT3 128, AT2 512, AT1 256
    for (int it = tid; it < 1792; it += nthreads) {
        float v[8];
        uint16_t* dst16 = nullptr;
        uint8_t* dst8 = nullptr;
        float qs = 1.0f;
        int i = it;
        if (i < 128) {  // A1[mt][l]
            const int mt = i >> 6, l = i & 63, row = 32 * mt + (l & 31);
            for (int q = 0; q < 8; q++) v[q] = w1(p, 8 * (l >> 5) + q, row);
            dst16 = &out->A1[mt][l][0];
            dst8 = &out->Q1[mt][l][0];
            qs = sc[0];
        } else if ((i -= 128) < 512) {  // A2[mt][s][l]
            const int mt = i >> 8, s = (i >> 6) & 3, l = i & 63, row = 32 * mt + (l & 31);
            for (int q = 0; q < 8; q++) v[q] = w2(p, unit_of_k(16 * s + 8 * (l >> 5) + q), row);
            dst16 = &out->A2[mt][s][l][0];
            dst8 = &out->Q2[mt][s][l][0];
            qs = sc[1];
        } else if ((i -= 512) < 256) {  // A3[s][l]
            const int s = i >> 6, l = i & 63, row = l & 31;
            for (int q = 0; q < 8; q++) v[q] = w3(p, unit_of_k(16 * s + 8 * (l >> 5) + q), row);
            dst16 = &out->A3[s][l][0];
            dst8 = &out->Q3[s][l][0];
            qs = sc[2];
        } else if ((i -= 256) < 128) {  // AT3[mt][l]: rows m, k = output
            const int mt = i >> 6, l = i & 63, row = 32 * mt + (l & 31);
            for (int q = 0; q < 8; q++) v[q] = w3(p, row, 8 * (l >> 5) + q);
            dst16 = &out->AT3[mt][l][0];
        } else if ((i -= 128) < 512) {  // AT2[mt][s][l]: rows u, k -> layer-2 unit
            const int mt = i >> 8, s = (i >> 6) & 3, l = i & 63, row = 32 * mt + (l & 31);
            for (int q = 0; q < 8; q++) v[q] = w2(p, row, unit_of_k(16 * s + 8 * (l >> 5) + q));
            dst16 = &out->AT2[mt][s][l][0];
        } else {  // AT1[s][l]: rows = latent inputs 0..7, k -> layer-1 unit
            i -= 512;
            const int s = i >> 6, l = i & 63, row = l & 31;
            for (int q = 0; q < 8; q++) v[q] = row < kLatentCh ? w1(p, row, unit_of_k(16 * s + 8 * (l >> 5) + q)) : 0.0f;
            dst16 = &out->AT1[s][l][0];
        }
        for (int q = 0; q < 8; q++) dst16[q] = bf16_bits(v[q]);
        if (dst8 && withFp8) {
            const long b = to_fp8x8(v, qs);
            for (int q = 0; q < 8; q++) dst8[q] = (uint8_t)(b >> (8 * q));
        }
    }
    // biases in accumulator order
    for (int i = tid; i < 2 * 2 * 16; i += nthreads) {
        const int mt = i >> 5, g = (i >> 4) & 1, r = i & 15, row = 32 * mt + acc_row(g, r);
        out->C1[mt][g][r] = p[kB1 + row];
        out->C2[mt][g][r] = p[kB2 + row];
    }
    for (int i = tid; i < 2 * 16; i += nthreads) {
        const int g = i >> 4, r = i & 15, row = acc_row(g, r);
        out->C3[g][r] = row < kMlpOut ? p[kB3 + row] : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Training
// ---------------------------------------------------------------------------
struct TrainArgs {
    const NeuralRecord* data;
    unsigned numData;
    NeuralLatent latent;  // FP32 master (base = float*)
    float* latentGrad;
    const NeuralPacked* packed;
    float* gradPartial;   // kNumParams per wave, reduced by k_reduce_grads
    float* loss;          // one float for this step
    unsigned step, chunksPerWave;
    float invBatch;
};

__device__ __forceinline__ unsigned hash32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ void latent_sample_f(const NeuralLatent& L, const float4 a, const float4 b, float* z) {
    const float* tex = static_cast<const float*>(L.base);
#pragma unroll
    for (int k = 0; k < kLatentCh; k++) z[k] = 0.0f;
    latent_visit(L, a.x, a.y, a.z, a.w, b.x, (int)b.y, [&](unsigned t, float w) {
        const float4 c0 = reinterpret_cast<const float4*>(tex)[2 * t], c1 = reinterpret_cast<const float4*>(tex)[2 * t + 1];
        z[0] = fmaf(w, c0.x, z[0]);
        z[1] = fmaf(w, c0.y, z[1]);
        z[2] = fmaf(w, c0.z, z[2]);
        z[3] = fmaf(w, c0.w, z[3]);
        z[4] = fmaf(w, c1.x, z[4]);
        z[5] = fmaf(w, c1.y, z[5]);
        z[6] = fmaf(w, c1.z, z[6]);
        z[7] = fmaf(w, c1.w, z[7]);
    });
}

__global__ void __launch_bounds__(64) k_train(TrainArgs A) {
    __shared__ float xs[64 * 16];  // inputs; later the latent gradients
    __shared__ float hA[64 * 64];  // H1, then dZ1
    __shared__ float hB[64 * 64];  // H2, then dZ2
    __shared__ float ys[64 * 4];
    __shared__ float dys[64 * 4];
    const int l = threadIdx.x, g = l >> 5, c = l & 31;
    const NeuralPacked* W = A.packed;

    float gW1[16], gW2[64], gW3[3];
    float gB1 = 0, gB2 = 0, gB3 = 0, lossAcc = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) gW1[k] = 0;
#pragma unroll
    for (int k = 0; k < 64; k++) gW2[k] = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) gW3[k] = 0;

    for (unsigned ch = 0; ch < A.chunksPerWave; ch++) {
        const unsigned gi = (blockIdx.x * A.chunksPerWave + ch) * 64u + (unsigned)l;
        const unsigned idx = hash32(gi * 0x9E3779B1u + A.step * 0x85EBCA77u + 0x1234567u) % A.numData;
        const NeuralRecord rec = A.data[idx];
        float z[8];
        latent_sample_f(A.latent, rec.a, rec.b, z);
        __syncthreads();  // LDS of the previous chunk is free
#pragma unroll
        for (int k = 0; k < 8; k++) xs[l * 16 + k] = z[k];
        xs[l * 16 + 8] = rec.b.z;
        xs[l * 16 + 9] = rec.b.w;
        xs[l * 16 + 10] = rec.c.x;
        xs[l * 16 + 11] = rec.c.y;
        xs[l * 16 + 12] = rec.c.z;
        xs[l * 16 + 13] = rec.c.w;
        xs[l * 16 + 14] = 0.0f;
        xs[l * 16 + 15] = 0.0f;
        __syncthreads();
        mlp_forward<false>(W, xs, hA, hB, ys);
        __syncthreads();
        {
            const float t[3] = {rec.d.x, rec.d.y, rec.d.z};
#pragma unroll
            for (int o = 0; o < 3; o++) {
                const float d = ys[l * 4 + o] - t[o];
                lossAcc = fmaf(d, d, lossAcc);
                dys[l * 4 + o] = 2.0f * d * A.invBatch;
            }
            dys[l * 4 + 3] = 0.0f;
        }
        __syncthreads();
        // dW3 (lane owns layer-2 unit m = l) and db3 (lanes 0..2)
        for (int f = 0; f < 64; f++) {
            const float h = hB[f * 64 + l];
            const float4 d = *reinterpret_cast<const float4*>(&dys[f * 4]);
            gW3[0] = fmaf(h, d.x, gW3[0]);
            gW3[1] = fmaf(h, d.y, gW3[1]);
            gW3[2] = fmaf(h, d.z, gW3[2]);
            if (l < 3) gB3 += dys[f * 4 + l];
        }
        // dH2 = W3 dY, masked by H2 > 0
        f32x16 dz2[2][2];
        {
            bf16x8 bdy[2];
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                float v[8];
#pragma unroll
                for (int q = 0; q < 8; q++) v[q] = (g == 0 && q < 3) ? dys[(32 * nt + c) * 4 + q] : 0.0f;
                bdy[nt] = to_bf16x8(v);
            }
#pragma unroll
            for (int mt = 0; mt < 2; mt++) {
                const bf16x8 a = load_bf16x8(&W->AT3[mt][l][0]);
#pragma unroll
                for (int nt = 0; nt < 2; nt++) {
                    f32x16 acc = {};
                    acc = mfma_bf16(a, bdy[nt], acc);
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const int f = 32 * nt + c, m = 32 * mt + acc_row(g, r);
                        acc[r] = hB[f * 64 + m] > 0.0f ? acc[r] : 0.0f;
                    }
                    dz2[mt][nt] = acc;
                }
            }
        }
        __syncthreads();  // every read of H2 is done
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
#pragma unroll
                for (int r = 0; r < 16; r++) hB[(32 * nt + c) * 64 + 32 * mt + acc_row(g, r)] = dz2[mt][nt][r];
        __syncthreads();
        // dW2 (lane owns layer-1 unit u = l) and db2 (lane owns m = l)
        for (int f = 0; f < 64; f++) {
            const float h = hA[f * 64 + l];
            const float4* dz = reinterpret_cast<const float4*>(&hB[f * 64]);
#pragma unroll
            for (int m4 = 0; m4 < 16; m4++) {
                const float4 d = dz[m4];
                gW2[4 * m4] = fmaf(h, d.x, gW2[4 * m4]);
                gW2[4 * m4 + 1] = fmaf(h, d.y, gW2[4 * m4 + 1]);
                gW2[4 * m4 + 2] = fmaf(h, d.z, gW2[4 * m4 + 2]);
                gW2[4 * m4 + 3] = fmaf(h, d.w, gW2[4 * m4 + 3]);
            }
            gB2 += hB[f * 64 + l];
        }
        // dH1 = W2 dZ2, masked by H1 > 0
        f32x16 dz1[2][2];
        {
            bf16x8 bb[4][2];
#pragma unroll
            for (int s = 0; s < 4; s++)
#pragma unroll
                for (int nt = 0; nt < 2; nt++) {
                    float v[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) v[q] = dz2[s >> 1][nt][8 * (s & 1) + q];
                    bb[s][nt] = to_bf16x8(v);
                }
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
#pragma unroll
                for (int nt = 0; nt < 2; nt++) {
                    f32x16 acc = {};
#pragma unroll
                    for (int s = 0; s < 4; s++) acc = mfma_bf16(load_bf16x8(&W->AT2[mt][s][l][0]), bb[s][nt], acc);
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const int f = 32 * nt + c, u = 32 * mt + acc_row(g, r);
                        acc[r] = hA[f * 64 + u] > 0.0f ? acc[r] : 0.0f;
                    }
                    dz1[mt][nt] = acc;
                }
        }
        __syncthreads();  // every read of H1 is done
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
#pragma unroll
                for (int r = 0; r < 16; r++) hA[(32 * nt + c) * 64 + 32 * mt + acc_row(g, r)] = dz1[mt][nt][r];
        __syncthreads();
        // dW1 (lane owns layer-1 unit u = l) and db1
        for (int f = 0; f < 64; f++) {
            const float d = hA[f * 64 + l];
#pragma unroll
            for (int k = 0; k < 16; k++) gW1[k] = fmaf(d, xs[f * 16 + k], gW1[k]);
            gB1 += d;
        }
        // dX = W1 dZ1 for the latent inputs (rows 0..7: lane group g, registers 0..3)
        f32x16 dx[2];
        {
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                f32x16 acc = {};
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    float v[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) v[q] = dz1[s >> 1][nt][8 * (s & 1) + q];
                    acc = mfma_bf16(load_bf16x8(&W->AT1[s][l][0]), to_bf16x8(v), acc);
                }
                dx[nt] = acc;
            }
        }
        __syncthreads();  // xs reads of dW1 are done
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int r = 0; r < 4; r++) xs[(32 * nt + c) * 16 + 4 * g + r] = dx[nt][r];
        __syncthreads();
        {
            float d8[8];
#pragma unroll
            for (int k = 0; k < 8; k++) d8[k] = xs[l * 16 + k];
            latent_visit(A.latent, rec.a.x, rec.a.y, rec.a.z, rec.a.w, rec.b.x, (int)rec.b.y, [&](unsigned t, float w) {
#pragma unroll
                for (int k = 0; k < 8; k++) atomicAdd(&A.latentGrad[8 * (size_t)t + k], w * d8[k]);
            });
        }
    }
    // this wave's weight gradients (every parameter written by one lane)
    float* gp = A.gradPartial + (size_t)blockIdx.x * kNumParams;
    for (int k = 0; k < 16; k++) gp[kW1 + k * 64 + l] = k < kMlpIn ? gW1[k] : 0.0f;
    gp[kB1 + l] = gB1;
    for (int m = 0; m < 64; m++) gp[kW2 + l * 64 + m] = gW2[m];
    gp[kB2 + l] = gB2;
    for (int o = 0; o < 4; o++) gp[kW3 + l * 4 + o] = o < 3 ? gW3[o] : 0.0f;
    if (l < 4) gp[kB3 + l] = l < 3 ? gB3 : 0.0f;
    for (int off = 32; off > 0; off >>= 1) lossAcc += __shfl_down(lossAcc, off, 64);
    if (l == 0) atomicAdd(A.loss, lossAcc);
}

// Sum of the per-wave gradients.
__global__ void __launch_bounds__(256) k_reduce_grads(const float* __restrict__ part, int waves,
                                                      float* __restrict__ grads) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kNumParams) return;
    float a = 0.0f;
    for (int w = 0; w < waves; w++) a += part[(size_t)w * kNumParams + i];
    grads[i] = a;
}

// Adam on n parameters; the gradient is read and cleared.
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ p, float* __restrict__ gr, float* __restrict__ m,
                                              float* __restrict__ v, size_t n, float lr, float b1, float b2,
                                              float c1, float c2, float eps) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float g = gr[i];
        gr[i] = 0.0f;
        const float mi = fmaf(b1, m[i], (1.0f - b1) * g);
        const float vi = fmaf(b2, v[i], (1.0f - b2) * g * g);
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    }
}

// Uniform(-a, a) initial latent from a counter hash.
__global__ void __launch_bounds__(256) k_init_latent(float* __restrict__ lat, size_t n, unsigned seed, float a) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const unsigned h = hash32((unsigned)i * 0x9E3779B1u ^ hash32(seed + (unsigned)(i >> 32)));
        lat[i] = a * (2.0f * ((float)(h >> 8) * (1.0f / 16777216.0f)) - 1.0f);
    }
}

__global__ void __launch_bounds__(256) k_to_half(const float* __restrict__ src, __half* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = __float2half(src[i]);
}

// Largest |input| (latent part), H1 and H2 over records, for the FP8 scales
// (qinfo[3..5], float bits with atomic max on non-negative values).
__global__ void __launch_bounds__(64) k_calibrate(const NeuralRecord* __restrict__ data, unsigned numData,
                                                  NeuralLatent latent, const NeuralPacked* __restrict__ W,
                                                  float* __restrict__ qinfo, unsigned stride) {
    __shared__ float xs[64 * 16];
    __shared__ float hA[64 * 64];
    __shared__ float hB[64 * 64];
    __shared__ float ys[64 * 4];
    const int l = threadIdx.x;
    const unsigned idx = ((blockIdx.x * 64u + l) * stride) % numData;
    const NeuralRecord rec = data[idx];
    float z[8];
    latent_sample_f(latent, rec.a, rec.b, z);
    float mx = 0.0f;
    for (int k = 0; k < 8; k++) {
        xs[l * 16 + k] = z[k];
        mx = fmaxf(mx, fabsf(z[k]));
    }
    const float dirs[6] = {rec.b.z, rec.b.w, rec.c.x, rec.c.y, rec.c.z, rec.c.w};
    for (int k = 0; k < 6; k++) {
        xs[l * 16 + 8 + k] = dirs[k];
        mx = fmaxf(mx, fabsf(dirs[k]));
    }
    xs[l * 16 + 14] = xs[l * 16 + 15] = 0.0f;
    __syncthreads();
    mlp_forward<false>(W, xs, hA, hB, ys);
    __syncthreads();
    float m1 = 0.0f, m2 = 0.0f;
    for (int k = 0; k < 64; k++) {
        m1 = fmaxf(m1, hA[l * 64 + k]);
        m2 = fmaxf(m2, hB[l * 64 + k]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mx = fmaxf(mx, __shfl_down(mx, off, 64));
        m1 = fmaxf(m1, __shfl_down(m1, off, 64));
        m2 = fmaxf(m2, __shfl_down(m2, off, 64));
    }
    if (l == 0) {
        atomicMax(reinterpret_cast<int*>(&qinfo[3]), __float_as_int(mx));
        atomicMax(reinterpret_cast<int*>(&qinfo[4]), __float_as_int(m1));
        atomicMax(reinterpret_cast<int*>(&qinfo[5]), __float_as_int(m2));
    }
}

// MLP on given inputs ([n][16] floats) -> outputs ([n][4] floats).
template <bool FP8>
__global__ void __launch_bounds__(64) k_eval(const float* __restrict__ in, float* __restrict__ out, int n,
                                             const NeuralPacked* __restrict__ W) {
    __shared__ float xs[64 * 16];
    __shared__ float ys[64 * 4];
    const int l = threadIdx.x;
    const int i = blockIdx.x * 64 + l;
    for (int k = 0; k < 16; k++) xs[l * 16 + k] = i < n && k < kMlpIn ? in[(size_t)i * 16 + k] : 0.0f;
    __syncthreads();
    mlp_forward<FP8>(W, xs, nullptr, nullptr, ys);
    __syncthreads();
    if (i < n)
        for (int o = 0; o < 4; o++) out[(size_t)i * 4 + o] = o < 3 ? ys[l * 4 + o] : 0.0f;
}

// Appends the valid feature records of a frame (one per pixel; taps = 0
// marks an empty pixel) to the data set.
__global__ void __launch_bounds__(256) k_append(const NeuralRecord* __restrict__ feat, int npix,
                                                NeuralRecord* __restrict__ data, unsigned* __restrict__ count,
                                                unsigned cap) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    bool valid = false;
    NeuralRecord r;
    if (i < npix) {
        r = feat[i];
        valid = r.b.y > 0.0f;
    }
    const unsigned pos = dev::wave_inc(count, 0, valid);
    if (valid && pos < cap) data[pos] = r;
}

}  // namespace

void neural_pack(const float* params, float* qinfo, NeuralPacked* out, int withFp8, hipStream_t s) {
    if (withFp8) hipLaunchKernelGGL(k_absmax, dim3(1), dim3(256), 0, s, params, qinfo);
    hipLaunchKernelGGL(k_pack, dim3(8), dim3(256), 0, s, params, qinfo, out, withFp8);
}

void neural_train_step(const NeuralRecord* data, unsigned numData, const NeuralLatent& latent, float* latentGrad,
                       const NeuralPacked* packed, float* gradPartial, float* grads, float* loss, unsigned step,
                       unsigned batch, unsigned chunksPerWave, hipStream_t s) {
    TrainArgs a{data, numData, latent, latentGrad, packed, gradPartial, loss, step, chunksPerWave,
                1.0f / (float)batch};
    const unsigned waves = batch / (64u * chunksPerWave);
    hipLaunchKernelGGL(k_train, dim3(waves), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_reduce_grads, dim3((kNumParams + 255) / 256), dim3(256), 0, s, gradPartial, (int)waves, grads);
}

void neural_adam(float* p, float* g, float* m, float* v, size_t n, float lr, float b1, float b2, int t, hipStream_t s) {
    const float c1 = 1.0f / (1.0f - powf(b1, (float)t)), c2 = 1.0f / (1.0f - powf(b2, (float)t));
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_adam, dim3(blocks), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, c1, c2, 1e-8f);
}

void neural_init_latent(float* lat, size_t n, unsigned seed, float a, hipStream_t s) {
    hipLaunchKernelGGL(k_init_latent, dim3(2048), dim3(256), 0, s, lat, n, seed, a);
}

void neural_to_half(const float* src, void* dst, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_to_half, dim3(2048), dim3(256), 0, s, src, static_cast<__half*>(dst), n);
}

void neural_calibrate(const NeuralRecord* data, unsigned numData, const NeuralLatent& latent,
                      const NeuralPacked* packed, float* qinfo, unsigned samples, hipStream_t s) {
    const unsigned waves = std::max(1u, samples / 64u);
    const unsigned stride = std::max(1u, numData / std::max(1u, samples)) | 1u;
    hipLaunchKernelGGL(k_calibrate, dim3(waves), dim3(64), 0, s, data, numData, latent, packed, qinfo, stride);
}

void neural_eval(const float* in, float* out, int n, const NeuralPacked* packed, int fp8, hipStream_t s) {
    const dim3 grid((n + 63) / 64);
    if (fp8)
        hipLaunchKernelGGL(k_eval<true>, grid, dim3(64), 0, s, in, out, n, packed);
    else
        hipLaunchKernelGGL(k_eval<false>, grid, dim3(64), 0, s, in, out, n, packed);
}

void neural_append(const NeuralRecord* feat, int npix, NeuralRecord* data, unsigned* count, unsigned cap,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_append, dim3((npix + 255) / 256), dim3(256), 0, s, feat, npix, data, count, cap);
}

}  // namespace ifrit_hip
