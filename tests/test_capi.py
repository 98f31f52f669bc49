"""The C-ABI library: it loads, exports every function include/ifrit_hip.h
declares (the reference's iftr* / ifcr* / ifbuf* entry points,
projects/softgraphics/include/ifrit/softgraphics/engine/export/TileRasterRendererExport.h:33-76,
EngineCommonExport.h:33-70, CommonExport.h:26-31, BufferManagerExport.h:25-35,
plus the extensions), and its host-only objects behave; no kernel is
launched here."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ifrit_hip.h")


@pytest.fixture(scope="module")
def lib():
    import ifrit_hip

    if not os.path.exists(ifrit_hip.LIB_PATH):
        subprocess.run([sys.executable, os.path.join(ROOT, "ifrit-v2_amd", "build.py")], check=True)
    return ifrit_hip.lib()


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_\s\*]*?)\b((?:iftr|ifcr|ifbuf)[A-Za-z0-9_]*)\s*\(", text):
        if "typedef" in m.group(1):
            continue
        names.add(m.group(2))
    return names


def test_header_declares_the_reference_entry_points():
    names = _declared_functions()
    reference = {"iftrCreateInstance", "iftrDestroyInstance", "iftrBindFrameBuffer", "iftrBindVertexBuffer",
                 "iftrBindIndexBuffer", "iftrBindVertexShaderFunc", "iftrBindFragmentShaderFunc", "iftrSetBlendFunc",
                 "iftrSetDepthFunc", "iftrOptsetForceDeterministic", "iftrOptSetDepthTestEnable", "iftrDrawLegacy",
                 "iftrClear", "iftrInit", "iftrTest", "iftrBindVertexShader", "iftrBindFragmentShader",
                 "iftrCreateFrameBuffer", "iftrDestroyFrameBuffer", "iftrSetFrameBufferColorAttachmentFP32",
                 "iftrSetFrameBufferDepthAttachmentFP32", "iftrCreateVaryingDescriptor",
                 "iftrDestroyVaryingDescriptor", "iftrWriteVaryingDescriptor", "iftrCreateVertexBuffer",
                 "iftrDestroyVertexBuffer", "iftrSetVertexBufferLayout", "iftrSetVertexBufferSize",
                 "iftrAllocateVertexBuffer", "iftrSetVertexBufferValueFloat4", "ifcrCreateImageFP32",
                 "ifcrDestroyImageFP32", "ifcrGetImageRawDataFP32", "ifbufCreateBufferManager",
                 "ifbufDestroyBufferManager", "ifbufCreateBuffer", "ifbufDestroyBuffer", "ifbufBufferData"}
    assert reference <= names, sorted(reference - names)


def test_library_exports_every_declared_symbol(lib):
    import ifrit_hip

    names = _declared_functions()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names == set(ifrit_hip.EXPORTED_SYMBOLS)
    # plain C names: no C++ mangling on the exported entry points
    out = subprocess.run(["nm", "-D", "--defined-only", ifrit_hip.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert names <= exported


def test_test_callback_receives_constant(lib):
    got = []
    CB = C.CFUNCTYPE(None, C.c_int)
    cb = CB(lambda v: got.append(v))
    lib.iftrTest(cb)
    assert got == [114514]  # TileRasterRendererExport.cpp:151-155


def test_fp32_image_objects(lib):
    img = lib.ifcrCreateImageFP32(5, 3, 4)
    assert img
    p = lib.ifcrGetImageRawDataFP32(img)
    arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(3, 5, 4))
    assert np.all(arr == 0)
    arr[2, 4, 3] = 7.5
    again = np.ctypeslib.as_array(C.cast(lib.ifcrGetImageRawDataFP32(img), C.POINTER(C.c_float)), shape=(60,))
    assert again[-1] == 7.5
    lib.ifcrDestroyImageFP32(img)


def test_buffer_manager_objects(lib):
    m = lib.ifbufCreateBufferManager()
    b = lib.ifbufCreateBuffer(m, 64)
    assert m and b
    data = np.arange(16, dtype=np.int32)
    lib.ifbufBufferData(b, data.ctypes.data, 0, 64)
    lib.ifbufBufferData(b, data.ctypes.data, 32, 32)
    lib.ifbufDestroyBuffer(b)
    lib.ifbufDestroyBufferManager(m)
    # destroying the manager first leaves the buffer valid for destruction
    m = lib.ifbufCreateBufferManager()
    b = lib.ifbufCreateBuffer(m, 8)
    lib.ifbufDestroyBufferManager(m)
    lib.ifbufDestroyBuffer(b)


def test_vertex_buffer_and_framebuffer_objects(lib):
    vb = lib.iftrCreateVertexBuffer()
    IFTP_FLOAT4 = 4  # include/ifrit/softgraphics/engine/base/TypeDescriptor.h:27-41
    layout = (C.c_int * 2)(IFTP_FLOAT4, IFTP_FLOAT4)
    lib.iftrSetVertexBufferLayout(vb, layout, 2)
    lib.iftrSetVertexBufferSize(vb, 4)
    lib.iftrAllocateVertexBuffer(vb)
    v = (C.c_float * 4)(1, 2, 3, 4)
    lib.iftrSetVertexBufferValueFloat4(vb, 3, 1, v)
    lib.iftrDestroyVertexBuffer(vb)
    fb = lib.iftrCreateFrameBuffer()
    img = lib.ifcrCreateImageFP32(8, 8, 4)
    dep = lib.ifcrCreateImageFP32(8, 8, 1)
    arr = (C.c_void_p * 1)(img)
    lib.iftrSetFrameBufferColorAttachmentFP32(fb, arr, 1)
    lib.iftrSetFrameBufferDepthAttachmentFP32(fb, dep)
    vd = lib.iftrCreateVaryingDescriptor()
    lib.iftrWriteVaryingDescriptor(vd, (C.c_int * 1)(IFTP_FLOAT4), 1)
    lib.iftrDestroyVaryingDescriptor(vd)
    lib.iftrDestroyFrameBuffer(fb)
    lib.ifcrDestroyImageFP32(img)
    lib.ifcrDestroyImageFP32(dep)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_init_without_device_reports_error_instead_of_aborting(lib):
    h = lib.iftrCreateInstance()
    lib.iftrInit(h)
    msg = lib.iftrGetLastError(h)
    assert msg, "Init without a device must leave an error message"
    lib.iftrDestroyInstance(h)


def test_python_renderer_fails_loudly_without_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    from ifrit_hip import Renderer

    with pytest.raises(Exception):
        r = Renderer()
        r.set_target(16, 16)
        r.draw()


def test_batch_and_tuning_struct_layouts():
    """The ctypes mirrors match the C layouts of include/ifrit_hip.h (compiled here with gcc)."""
    import os
    import tempfile

    import ifrit_hip

    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "ifrit_hip.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(IfritViewDesc), offsetof(IfritViewDesc, camera),
           offsetof(IfritViewDesc, shade), sizeof(IfritTuning), sizeof(IfritAutotuneReport),
           offsetof(IfritAutotuneReport, tried), offsetof(IfritAutotuneReport, ms),
           sizeof(IfritGpuFragmentProgramDesc), offsetof(IfritGpuFragmentProgramDesc, samplerId),
           offsetof(IfritGpuFragmentProgramDesc, depthWrite), offsetof(IfritGpuFragmentProgramDesc, depthBias),
           sizeof(IfritSamplerDesc), sizeof(IfritImageBlit));
    return 0;
}
'''
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        with open(c, "w") as f:
            f.write(src)
        subprocess.run(["gcc", "-I", inc, c, "-o", exe], check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(ifrit_hip.ViewDesc), ifrit_hip.ViewDesc.camera.offset, ifrit_hip.ViewDesc.shade.offset,
            C.sizeof(ifrit_hip.Tuning), C.sizeof(ifrit_hip.AutotuneReport), ifrit_hip.AutotuneReport.tried.offset,
            ifrit_hip.AutotuneReport.ms.offset, C.sizeof(ifrit_hip.FragmentProgramDesc),
            ifrit_hip.FragmentProgramDesc.samplerId.offset, ifrit_hip.FragmentProgramDesc.depthWrite.offset,
            ifrit_hip.FragmentProgramDesc.depthBias.offset, C.sizeof(ifrit_hip.SamplerDesc),
            C.sizeof(ifrit_hip.ImageBlit)]
    assert got == want


def test_batch_calls_without_device_report_errors(lib):
    import ifrit_hip

    h = lib.iftrCreateInstance()
    views = (ifrit_hip.ViewDesc * 2)()
    views[0].width, views[0].height = 0, 10  # invalid size
    lib.iftrSetViews(h, C.cast(views, C.c_void_p), 2)
    assert lib.iftrGetLastError(h)
    t = ifrit_hip.Tuning(300, 4, 8, 64, 0, 256)  # 300 is not a permitted part size
    lib.iftrSetTuning(h, C.byref(t))
    assert lib.iftrGetLastError(h)
    lib.iftrGetTuning(h, C.byref(t))
    assert (t.rasterPartEntries, t.rasterGroupsPerCU, t.shadeTile) == (512, 4, 8)
    lib.iftrDestroyInstance(h)
