// Neural material: an 8-channel latent texture with a mip chain and an MLP
// 14 -> 64 -> 64 -> 3 (ReLU) that replaces the layered-GGX shader.  Inputs:
// the filtered latent (8), the view and light directions in the tangent
// frame (3 + 3).  Output: linear RGB, clamped like every shaded colour.
//
// MFMA formulation (v_mfma_f32_32x32x16_bf16 / _fp8_fp8, wave64).  Every layer
// computes Y^T = W^T X^T for 64 fragments: A = weights (rows = output units,
// k = inputs), B = activations (k = inputs, columns = fragments), so a
// 32-fragment tile is one N block.  Lane l of a 32x32x16 operand holds row or
// column l % 32 and k = 8 (l / 32) + q, q = 0..7; the 16 accumulator values of
// lane l are column l % 32, rows i(g, r) = 8 (r / 4) + 4 g + r % 4 with
// g = l / 32.  A layer's accumulators become the next layer's B operand in
// place: register r of row tile t and lane group g is taken as
// k = 32 t + 16 (r / 8) + 8 g + r % 8, and the next layer's weight columns are
// permuted to match (unit_of_k below), so no data moves between layers.
#pragma once
#include <stdint.h>

namespace ifrit_hip {

constexpr int kLatentCh = 8;
constexpr int kMlpIn = 14;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 3;

// FP32 master parameters (training state, also the source of the packed
// operands), offsets in floats.
constexpr int kW1 = 0;      // [16][64]: input k (rows 14, 15 stay 0), unit u
constexpr int kB1 = 1024;   // [64]
constexpr int kW2 = 1088;   // [64][64]: layer-1 unit u, layer-2 unit m
constexpr int kB2 = 5184;   // [64]
constexpr int kW3 = 5248;   // [64][4]: layer-2 unit m, output o (o = 3 unused)
constexpr int kB3 = 5504;   // [4]
constexpr int kNumParams = 5508;

// Layer unit held by the k-th B-operand value taken from the accumulators of a
// 64-row layer (see the header comment).
__host__ __device__ inline int unit_of_k(int k) {
    const int t = k >> 5, s = (k >> 4) & 1, g = (k >> 3) & 1, q = k & 7;
    const int r = 8 * s + q;
    return 32 * t + 8 * (r >> 2) + 4 * g + (r & 3);
}
// Accumulator row of register r in lane group g.
__host__ __device__ inline int acc_row(int g, int r) { return 8 * (r >> 2) + 4 * g + (r & 3); }

// Operands in lane order for one wave (bf16 as raw 16-bit words).
struct NeuralPacked {
    uint16_t A1[2][64][8];     // forward: W1^T row tiles
    uint16_t A2[2][4][64][8];  // forward: W2^T, [row tile][k step]
    uint16_t A3[4][64][8];     // forward: W3^T (rows 0..2)
    float C1[2][2][16];        // biases in accumulator order [row tile][lane group][r]
    float C2[2][2][16];
    float C3[2][16];
    uint16_t AT3[2][64][8];    // backward: W3 (rows m, k = output)
    uint16_t AT2[2][4][64][8]; // backward: W2 (rows u, k = layer-2 units)
    uint16_t AT1[4][64][8];    // backward: W1 (rows = latent inputs 0..7, k = layer-1 units)
    uint8_t Q1[2][64][8];      // FP8 (E4M3) forward operands, scaled by qscale
    uint8_t Q2[2][4][64][8];
    uint8_t Q3[4][64][8];
    float qscale[8];           // weight scales sW1, sW2, sW3; activation scales sX, sH1, sH2
};

// Latent texture: square, power of two, levels packed, 8 channels per texel
// (FP32 while training, FP16 for shading).
struct NeuralLatent {
    const void* base;
    int size, levels;
    unsigned levelOffset[16];  // texel offsets
};

struct NeuralShadeParams {
    NeuralLatent latent;        // FP16
    const NeuralPacked* packed;
    int precision;              // 0: BF16, 1: FP8
    int lodShift;               // log2(material layer size / latent size)
    int singleTap;              // 1: one trilinear latent tap at the isotropic footprint
    float4* featOut;            // feature records (program 5), one per pixel
};

// One training fragment (64 bytes): latent footprint (u, v, au, av), (lod on
// the latent, taps, view x, view y), (view z, light xyz), (target rgb, 0).
struct NeuralRecord {
    float4 a, b, c, d;
};

}  // namespace ifrit_hip
