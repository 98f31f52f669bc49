// DrawMeshTasks with the built-in meshlet program (the reference's meshlet
// demo, demo/shader/MeshletDemoShaders.cu, on the mesh pipeline of
// TileRasterCoreInvocationCuda.cu:355-500).  One workgroup per mesh work
// group: it emits its meshlet's vertices (each the input record followed by
// its clip position) and its indices rebased onto the emitted vertices.  Meshlets land in
// work-group order, so the list the triangle pipeline then draws is the
// concatenation of the meshlets.
#include "pipeline.h"

namespace ifrit_hip {
namespace {

__global__ void __launch_bounds__(256) k_mesh_expand(MeshParams P, const float* __restrict__ records,
                                                     const int* __restrict__ indices, const int* __restrict__ vo,
                                                     const int* __restrict__ io, float* __restrict__ outVtx,
                                                     int* __restrict__ outIdx, DeviceCounters* ctr) {
    const int m = P.first + (int)blockIdx.x;  // one mesh work group per task work group, payload = its id
    const int v0 = vo[m], v1 = vo[m + 1], i0 = io[m], i1 = io[m + 1];
    const int base = v0 - P.vertexBase;  // first emitted vertex of this meshlet
    const int stride = 4 + P.recordFloats;
    const float* M = P.mvp;
    for (int k = threadIdx.x; k < v1 - v0; k += blockDim.x) {
        const float* r = records + (size_t)(v0 + k) * P.recordFloats;
        float* out = outVtx + (size_t)(base + k) * stride;
        const float x = r[0], y = r[1], z = r[2];
        for (int f = 0; f < P.recordFloats; f++) out[f] = r[f];
        // the transform program's summation order (LinalgOps.h:40-48)
        float* c = out + P.recordFloats;
        c[0] = M[0] * x + M[1] * y + M[2] * z + M[3] * 1.0f;
        c[1] = M[4] * x + M[5] * y + M[6] * z + M[7] * 1.0f;
        c[2] = M[8] * x + M[9] * y + M[10] * z + M[11] * 1.0f;
        c[3] = M[12] * x + M[13] * y + M[14] * z + M[15] * 1.0f;
    }
    for (int k = threadIdx.x; k < i1 - i0; k += blockDim.x) {
        const int local = indices[i0 + k];
        // an index outside its meshlet is flagged (the draw reports it) and left out of range
        if (local < 0 || local >= v1 - v0) atomicOr(&ctr->status, kStatusBadIndex);
        outIdx[(i0 - P.indexBase) + k] = (local < 0 || local >= v1 - v0) ? -1 : base + local;
    }
}

// The geometry stage with the built-in clip-offset program: one lane per
// emitted vertex.
__global__ void __launch_bounds__(256) k_geometry_expand(GeometryParams P, const float* __restrict__ vtx,
                                                         const int* __restrict__ idx, float* __restrict__ outVtx,
                                                         int* __restrict__ outIdx, DeviceCounters* ctr) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= 3 * P.ntri) return;
    int vi = idx[j];
    if (vi < 0 || vi >= P.nverts) {
        atomicOr(&ctr->status, kStatusBadIndex);
        vi = 0;
    }
    const float* r = vtx + (size_t)vi * P.stride;
    float* out = outVtx + (size_t)j * (4 + P.stride);
    float c[4];
    if (P.passthrough) {
        for (int k = 0; k < 4; k++) c[k] = r[k];
    } else {
        const float* M = P.mvp;
        const float x = r[0], y = r[1], z = r[2];
        c[0] = M[0] * x + M[1] * y + M[2] * z + M[3] * 1.0f;
        c[1] = M[4] * x + M[5] * y + M[6] * z + M[7] * 1.0f;
        c[2] = M[8] * x + M[9] * y + M[10] * z + M[11] * 1.0f;
        c[3] = M[12] * x + M[13] * y + M[14] * z + M[15] * 1.0f;
    }
    for (int f = 0; f < P.stride; f++) out[f] = r[f];
    for (int k = 0; k < 4; k++) out[P.stride + k] = c[k] + P.offset[k];
    outIdx[j] = j;
}

}  // namespace

void launch_geometry_expand(const GeometryParams& p, const float* vtx, const int* idx, float* outVtx, int* outIdx,
                            DeviceCounters* ctr, hipStream_t s) {
    if (p.ntri <= 0) return;
    hipLaunchKernelGGL(k_geometry_expand, dim3((3 * p.ntri + 255) / 256), dim3(256), 0, s, p, vtx, idx, outVtx, outIdx,
                       ctr);
}

void launch_mesh_expand(const MeshParams& p, const float* records, const int* indices, const int* vo, const int* io,
                        float* outVtx, int* outIdx, DeviceCounters* ctr, hipStream_t s) {
    if (p.groups <= 0) return;
    hipLaunchKernelGGL(k_mesh_expand, dim3(p.groups), dim3(256), 0, s, p, records, indices, vo, io, outVtx, outIdx,
                       ctr);
}

}  // namespace ifrit_hip
