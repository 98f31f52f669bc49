// DrawMeshTasks with the built-in meshlet program (the reference's meshlet
// demo, demo/shader/MeshletDemoShaders.cu, on the mesh pipeline of
// TileRasterCoreInvocationCuda.cu:355-500).  One workgroup per mesh work
// group: it emits its meshlet's vertices, transformed, with their records,
// and its indices rebased onto the emitted vertices.  Meshlets land in
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
        // the transform program's summation order (LinalgOps.h:40-48)
        out[0] = M[0] * x + M[1] * y + M[2] * z + M[3] * 1.0f;
        out[1] = M[4] * x + M[5] * y + M[6] * z + M[7] * 1.0f;
        out[2] = M[8] * x + M[9] * y + M[10] * z + M[11] * 1.0f;
        out[3] = M[12] * x + M[13] * y + M[14] * z + M[15] * 1.0f;
        for (int f = 0; f < P.recordFloats; f++) out[4 + f] = r[f];
    }
    for (int k = threadIdx.x; k < i1 - i0; k += blockDim.x) {
        const int local = indices[i0 + k];
        // an index outside its meshlet is flagged (the draw reports it) and left out of range
        if (local < 0 || local >= v1 - v0) atomicOr(&ctr->status, kStatusBadIndex);
        outIdx[(i0 - P.indexBase) + k] = (local < 0 || local >= v1 - v0) ? -1 : base + local;
    }
}

}  // namespace

void launch_mesh_expand(const MeshParams& p, const float* records, const int* indices, const int* vo, const int* io,
                        float* outVtx, int* outIdx, DeviceCounters* ctr, hipStream_t s) {
    if (p.groups <= 0) return;
    hipLaunchKernelGGL(k_mesh_expand, dim3(p.groups), dim3(256), 0, s, p, records, indices, vo, io, outVtx, outIdx,
                       ctr);
}

}  // namespace ifrit_hip
