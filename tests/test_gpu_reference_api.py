"""The reference's own C ABI end to end on the GPU.

A frame buffer of FP32 images, a VertexBuffer object, an index buffer from
the buffer manager and a host vertex-shader function pointer (run on the
host, TileRasterRendererExport.h:44-47) are drawn with iftrDrawLegacy.  The
host shader computes the transform in FP32 in the same order as the GPU
vertex program, so primitive IDs, depth bits and colour must equal those of
the GPU vertex-program path bit for bit; the depth attachment must hold each
pixel's own depth and the colour attachment the unquantised colour.
"""
import ctypes as C

import numpy as np
import pytest

import ifrit_hip
from ifrit_hip import scenes

pytestmark = pytest.mark.gpu

IFTP_FLOAT4 = 4  # include/ifrit/softgraphics/engine/base/TypeDescriptor.h:25-39


class Vec4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


VSFUNC = C.CFUNCTYPE(None, C.POINTER(C.c_void_p), C.POINTER(Vec4), C.POINTER(C.POINTER(Vec4)))


def test_reference_objects_and_host_vertex_shader(renderer_factory):
    s = scenes.config1(width=160, height=120)
    W, H = s.width, s.height
    # GPU vertex program path
    r = renderer_factory()
    r.set_target(W, H, 1)
    r.bind_geometry(s.vertices, s.indices, s.mvp)
    r.bind_fragment_program(s, program=scenes.PROGRAM_FLAT)
    r.draw()
    ids_a, depth_a = r.visibility()
    col_a = r.color()

    L = ifrit_hip.lib()
    h = L.iftrCreateInstance()
    col = L.ifcrCreateImageFP32(W, H, 4)
    dep = L.ifcrCreateImageFP32(W, H, 1)
    fb = L.iftrCreateFrameBuffer()
    vb = L.iftrCreateVertexBuffer()
    mgr = L.ifbufCreateBufferManager()
    vd = L.iftrCreateVaryingDescriptor()
    try:
        L.iftrInit(h)
        L.iftrSetFrameBufferColorAttachmentFP32(fb, (C.c_void_p * 1)(col), 1)
        L.iftrSetFrameBufferDepthAttachmentFP32(fb, dep)
        L.iftrBindFrameBuffer(h, fb)
        pos = np.ascontiguousarray(s.vertices[:, :3], np.float32)
        L.iftrSetVertexBufferLayout(vb, (C.c_int * 1)(IFTP_FLOAT4), 1)
        L.iftrSetVertexBufferSize(vb, pos.shape[0])
        L.iftrAllocateVertexBuffer(vb)
        for i, p in enumerate(pos):
            L.iftrSetVertexBufferValueFloat4(vb, i, 0, (C.c_float * 4)(p[0], p[1], p[2], 1.0))
        L.iftrBindVertexBuffer(h, vb)
        idx = np.ascontiguousarray(s.indices, np.int32)
        ib = L.ifbufCreateBuffer(mgr, idx.nbytes)
        L.ifbufBufferData(ib, idx.ctypes.data, 0, idx.nbytes)
        L.iftrBindIndexBuffer(h, ib)
        L.iftrWriteVaryingDescriptor(vd, (C.c_int * 1)(IFTP_FLOAT4), 1)
        M = np.asarray(s.mvp, np.float32)
        one = np.float32(1.0)

        def vs(inp, out_pos, out_vary):
            p = C.cast(inp[0], C.POINTER(C.c_float))
            x, y, z = np.float32(p[0]), np.float32(p[1]), np.float32(p[2])
            # FP32, summed left to right like the GPU program (no fused multiply-add)
            c = [((M[k, 0] * x + M[k, 1] * y) + M[k, 2] * z) + M[k, 3] * one for k in range(4)]
            out_pos[0] = Vec4(float(c[0]), float(c[1]), float(c[2]), float(c[3]))
            out_vary[0][0] = Vec4(float(x), float(y), float(z), 1.0)

        cb = VSFUNC(vs)
        L.iftrBindVertexShaderFunc(h, C.cast(cb, C.c_void_p), vd)
        d = ifrit_hip.FragmentProgramDesc()
        d.kind = scenes.PROGRAM_FLAT
        prog = L.iftrCreateGpuFragmentProgram(C.byref(d))
        L.iftrBindFragmentShader(h, prog)
        L.iftrDrawLegacy(h, idx.size, 1)
        err = L.iftrGetLastError(h)
        assert err is None, err
        ids_b = np.empty(W * H, np.uint32)
        depth_b = np.empty(W * H, np.float32)
        L.iftrGetVisibility(h, ids_b.ctypes.data, depth_b.ctypes.data)
        col_b = np.empty((H, W, 4), np.uint8)
        L.iftrGetColorRGBA8(h, col_b.ctypes.data)
        depth_img = np.ctypeslib.as_array(L.ifcrGetImageRawDataFP32(dep), shape=(W * H,)).copy()
        col_img = np.ctypeslib.as_array(L.ifcrGetImageRawDataFP32(col), shape=(H, W, 4)).copy()
        L.iftrDestroyGpuProgram(prog)
    finally:
        L.iftrDestroyInstance(h)
        L.iftrDestroyFrameBuffer(fb)
        L.ifcrDestroyImageFP32(col)
        L.ifcrDestroyImageFP32(dep)
        L.iftrDestroyVertexBuffer(vb)
        L.iftrDestroyVaryingDescriptor(vd)
        L.ifbufDestroyBufferManager(mgr)
    covered = ids_a != 0xFFFFFFFF
    print(f"[reference ABI] covered {covered.sum()} id mismatches {(ids_a != ids_b).sum()} depth mismatches "
          f"{(depth_a.view(np.uint32) != depth_b.view(np.uint32)).sum()}", flush=True)
    assert covered.sum() > 1000
    assert np.array_equal(ids_a, ids_b)
    assert np.array_equal(depth_a.view(np.uint32), depth_b.view(np.uint32))
    assert np.array_equal(col_a, col_b)
    assert np.array_equal(depth_img.view(np.uint32), depth_b.view(np.uint32))
    assert np.array_equal(np.floor(col_img * 255.0 + 0.5).astype(np.uint8), col_b)
