// Texture preparation on the device: procedural fBm level 0 and box-filter
// mip levels (the image operations of
// projects/softgraphics/src/engine/tilerastercuda/TileRasterImageOpInvocationsCuda.cu:36-156,
// with a 2x2 box filter in place of the point-like blit, so that every mip
// texel is an exact integer mean of its four parents).
#include <hip/hip_runtime.h>

#include "pipeline.h"

namespace ifrit_hip {

namespace {

__device__ __forceinline__ uint32_t hash2(uint32_t ix, uint32_t iy, uint32_t seed) {
    uint32_t h = (ix * 0x8DA6B343u) ^ (iy * 0xD8163841u) ^ (seed * 0x9E3779B1u);
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

// Tileable value-noise fBm in 16-bit fixed point: octave o has cells of
// size >> (o + 2) texels and contributes value >> (o + 1); channel c uses
// seed * 4 + c; alpha is 255.  Integer-exact, so any implementation of the
// same definition produces the same bytes.
__global__ void __launch_bounds__(256) k_fbm(uint8_t* __restrict__ out, int size, int log2size, unsigned seed) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= size || y >= size) return;
    uchar4 px;
    unsigned char ch[3];
    for (int c = 0; c < 3; c++) {
        long long acc = 0;
        for (int o = 0; o < 6; o++) {
            const int shift = max(0, log2size - (o + 2));
            const int cell = 1 << shift;
            const int ncell = size >> shift;
            const int cx = x >> shift, cy = y >> shift;
            const long long wx = (long long)(x & (cell - 1)) << (16 - shift);
            const long long wy = (long long)(y & (cell - 1)) << (16 - shift);
            const uint32_t s = (seed * 4u + (uint32_t)c) * 16u + (uint32_t)o;
            const int cx1 = (cx + 1) % ncell, cy1 = (cy + 1) % ncell;
            const long long v00 = hash2((uint32_t)(cx % ncell), (uint32_t)(cy % ncell), s) & 0xFFFF;
            const long long v10 = hash2((uint32_t)cx1, (uint32_t)(cy % ncell), s) & 0xFFFF;
            const long long v01 = hash2((uint32_t)(cx % ncell), (uint32_t)cy1, s) & 0xFFFF;
            const long long v11 = hash2((uint32_t)cx1, (uint32_t)cy1, s) & 0xFFFF;
            const long long top = v00 * (65536 - wx) + v10 * wx;
            const long long bot = v01 * (65536 - wx) + v11 * wx;
            const long long n = (top * (65536 - wy) + bot * wy) >> 32;
            acc += n >> (o + 1);
        }
        const long long b = acc >> 8;
        ch[c] = (unsigned char)(b > 255 ? 255 : b);
    }
    px = make_uchar4(ch[0], ch[1], ch[2], 255);
    reinterpret_cast<uchar4*>(out)[(size_t)y * size + x] = px;
}

__global__ void __launch_bounds__(256) k_mip(const uchar4* __restrict__ src, int sw, int sh, uchar4* __restrict__ dst,
                                             int dw, int dh) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= dw || y >= dh) return;
    const int x0 = min(2 * x, sw - 1), x1 = min(2 * x + 1, sw - 1);
    const int y0 = min(2 * y, sh - 1), y1 = min(2 * y + 1, sh - 1);
    const uchar4 a = src[(size_t)y0 * sw + x0], b = src[(size_t)y0 * sw + x1];
    const uchar4 c = src[(size_t)y1 * sw + x0], d = src[(size_t)y1 * sw + x1];
    dst[(size_t)y * dw + x] = make_uchar4((unsigned char)((a.x + b.x + c.x + d.x + 2) >> 2),
                                          (unsigned char)((a.y + b.y + c.y + d.y + 2) >> 2),
                                          (unsigned char)((a.z + b.z + c.z + d.z + 2) >> 2),
                                          (unsigned char)((a.w + b.w + c.w + d.w + 2) >> 2));
}

__global__ void __launch_bounds__(256) k_blit(const uchar4* __restrict__ src, uchar4* __restrict__ dst, BlitArgs a) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int rw = (int)(a.dx1 - a.dx0), rh = (int)(a.dy1 - a.dy0);
    if (x >= rw || y >= rh) return;
    const float px = (float)x / (float)rw, py = (float)y / (float)rh;
    const float cx = px * (float)(a.sx1 - a.sx0) + (float)a.sx0;
    const float cy = py * (float)(a.sy1 - a.sy0) + (float)a.sy0;
    uchar4 out;
    if (a.linear) {
        const int ix = (int)cx, iy = (int)cy;
        const float fx = cx - (float)ix, fy = cy - (float)iy;
        const int ix1 = min(ix + 1, a.sw - 1), iy1 = min(iy + 1, a.sh - 1);
        const uchar4 c00 = src[(size_t)iy * a.sw + ix], c01 = src[(size_t)iy * a.sw + ix1];
        const uchar4 c10 = src[(size_t)iy1 * a.sw + ix], c11 = src[(size_t)iy1 * a.sw + ix1];
        auto lerp = [](float p, float q, float t) { return p + (q - p) * t; };
        auto ch = [&](float a00, float a01, float a10, float a11) {
            const float v = lerp(lerp(a00, a01, fx), lerp(a10, a11, fx), fy);
            return (unsigned char)fminf(fmaxf(floorf(v + 0.5f), 0.0f), 255.0f);
        };
        out = make_uchar4(ch(c00.x, c01.x, c10.x, c11.x), ch(c00.y, c01.y, c10.y, c11.y),
                          ch(c00.z, c01.z, c10.z, c11.z), ch(c00.w, c01.w, c10.w, c11.w));
    } else {
        const int sx = min((int)roundf(cx), a.sw - 1), sy = min((int)roundf(cy), a.sh - 1);
        out = src[(size_t)sy * a.sw + sx];
    }
    dst[(size_t)(a.dy0 + y) * a.dw + a.dx0 + x] = out;
}

}  // namespace

void launch_blit(const uint8_t* src, uint8_t* dst, const BlitArgs& a, hipStream_t s) {
    const int rw = (int)(a.dx1 - a.dx0), rh = (int)(a.dy1 - a.dy0);
    if (rw <= 0 || rh <= 0) return;
    const dim3 grid((rw + 15) / 16, (rh + 15) / 16);
    hipLaunchKernelGGL(k_blit, grid, dim3(256), 0, s, reinterpret_cast<const uchar4*>(src), reinterpret_cast<uchar4*>(dst),
                       a);
}

void launch_fbm_texture(uint8_t* level0, int size, unsigned seed, hipStream_t s) {
    int log2size = 0;
    while ((1 << log2size) < size) log2size++;
    const dim3 grid((size + 15) / 16, (size + 15) / 16);
    hipLaunchKernelGGL(k_fbm, grid, dim3(256), 0, s, level0, size, log2size, seed);
}

void launch_mip_level(const uint8_t* src, int sw, int sh, uint8_t* dst, int dw, int dh, hipStream_t s) {
    const dim3 grid((dw + 15) / 16, (dh + 15) / 16);
    hipLaunchKernelGGL(k_mip, grid, dim3(256), 0, s, reinterpret_cast<const uchar4*>(src), sw, sh,
                       reinterpret_cast<uchar4*>(dst), dw, dh);
}

}  // namespace ifrit_hip
