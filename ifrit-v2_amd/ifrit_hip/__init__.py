"""Python binding of the MI355X tile rasterizer (libifrit_hip.so, C ABI in
include/ifrit_hip.h).

This is a thin ctypes layer for tests and benchmarks; the renderer itself is
C++/HIP.  There is no fallback: if the library is missing or no GPU is
present, loading or drawing raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
# IFRIT_HIP_LIB names another build of the library (an A/B build of the same sources)
LIB_PATH = os.environ.get("IFRIT_HIP_LIB") or os.path.join(PKG_ROOT, "lib", "libifrit_hip.so")

_lib = None


class VertexProgramDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("mvp", C.c_float * 16)]


class FragmentProgramDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("offWorld", C.c_int32), ("offNormal", C.c_int32), ("offUv", C.c_int32),
                ("trisPerMaterial", C.c_int32), ("sampler", C.c_int32), ("maxAniso", C.c_int32),
                ("camera", C.c_float * 3), ("ambient", C.c_float), ("kd", C.c_float), ("ks", C.c_float),
                ("shininess", C.c_float), ("numLights", C.c_int32), ("lightPos", (C.c_float * 4) * 4),
                ("lightColor", (C.c_float * 4) * 4), ("offTangent", C.c_int32), ("precision", C.c_int32),
                ("samplerId", C.c_int32), ("lod", C.c_float), ("depthWrite", C.c_int32), ("depthScale", C.c_float),
                ("depthBias", C.c_float)]


class MeshProgramDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("mvp", C.c_float * 16), ("vertexSlot", C.c_int32), ("indexSlot", C.c_int32),
                ("vertexOffsetSlot", C.c_int32), ("indexOffsetSlot", C.c_int32), ("recordFloats", C.c_int32)]


class GeometryProgramDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("offset", C.c_float * 4)]


class SamplerDesc(C.Structure):
    _fields_ = [("filterMode", C.c_int32), ("addressModeU", C.c_int32), ("addressModeV", C.c_int32),
                ("addressModeW", C.c_int32), ("borderColor", C.c_int32), ("anisotropyEnable", C.c_int32),
                ("maxAnisotropy", C.c_float)]


class ImageBlit(C.Structure):
    _fields_ = [("srcMip", C.c_uint32), ("srcLayer", C.c_uint32), ("srcSt", C.c_uint32 * 3), ("srcEd", C.c_uint32 * 3),
                ("dstMip", C.c_uint32), ("dstLayer", C.c_uint32), ("dstSt", C.c_uint32 * 3),
                ("dstEd", C.c_uint32 * 3)]


class NeuralTrainDesc(C.Structure):
    _fields_ = [("views", C.c_int32), ("viewWidth", C.c_int32), ("viewHeight", C.c_int32),
                ("maxFragments", C.c_int32), ("steps", C.c_int32), ("batch", C.c_int32), ("lrLatent", C.c_float),
                ("lrMlp", C.c_float), ("seed", C.c_uint32), ("latentSize", C.c_int32), ("eyeDistMin", C.c_float),
                ("eyeDistMax", C.c_float), ("fovY", C.c_float), ("targetJitter", C.c_float), ("fovYMin", C.c_float),
                ("singleTap", C.c_int32)]


class NeuralTrainReport(C.Structure):
    _fields_ = [("fragments", C.c_int64), ("lossFirst", C.c_float), ("lossLast", C.c_float),
                ("datasetMs", C.c_float), ("trainMs", C.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


NEURAL_PARAMS = 5508  # W1[16][64], b1[64], W2[64][64], b2[64], W3[64][4], b3[4]


class FrameStats(C.Structure):
    _fields_ = [("inputTriangles", C.c_int64), ("assembledTriangles", C.c_int64), ("clippedTriangles", C.c_int64),
                ("binEntries", C.c_int64), ("fragments", C.c_int64), ("coveredSamples", C.c_int64),
                ("status", C.c_uint32), ("retries", C.c_uint32), ("stageMs", C.c_float * 8)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "stageMs"}
        d["stageMs"] = list(self.stageMs)
        return d


class BlendState(C.Structure):
    _fields_ = [("blendEnable", C.c_uint8), ("srcColorBlendFactor", C.c_int32), ("dstColorBlendFactor", C.c_int32),
                ("srcAlphaBlendFactor", C.c_int32), ("dstAlphaBlendFactor", C.c_int32)]


class ViewDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("mvp", C.c_float * 16), ("camera", C.c_float * 3),
                ("shade", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("rasterPartEntries", C.c_int32), ("rasterGroupsPerCU", C.c_int32), ("shadeTile", C.c_int32),
                ("bigTriangleTiles", C.c_int32), ("binCapacity", C.c_uint32), ("rasterTileThreads", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AutotuneReport(C.Structure):
    _fields_ = [("candidates", C.c_int32), ("identical", C.c_int32), ("defaultMs", C.c_float), ("bestMs", C.c_float),
                ("binDemand", C.c_uint32), ("best", Tuning), ("tried", Tuning * 64), ("ms", C.c_float * 64)]

    def as_dict(self):
        n = min(self.candidates, 64)
        return {"candidates": self.candidates, "identical": bool(self.identical), "default_ms": self.defaultMs,
                "best_ms": self.bestMs, "bin_demand": self.binDemand, "best": self.best.as_dict(),
                "tried": [dict(self.tried[i].as_dict(), ms=self.ms[i]) for i in range(n)]}


# Every symbol declared in include/ifrit_hip.h.
EXPORTED_SYMBOLS = [
    "iftrCreateInstance", "iftrDestroyInstance", "iftrBindFrameBuffer", "iftrBindVertexBuffer", "iftrBindIndexBuffer",
    "iftrBindVertexShaderFunc", "iftrBindFragmentShaderFunc", "iftrSetBlendFunc", "iftrSetDepthFunc",
    "iftrOptsetForceDeterministic", "iftrOptSetDepthTestEnable", "iftrOptsetDepthTestEnable", "iftrDrawLegacy", "iftrClear", "iftrInit", "iftrTest",
    "iftrBindVertexShader", "iftrBindFragmentShader", "iftrCreateFrameBuffer", "iftrDestroyFrameBuffer",
    "iftrSetFrameBufferColorAttachmentFP32", "iftrSetFrameBufferDepthAttachmentFP32", "iftrCreateVaryingDescriptor",
    "iftrDestroyVaryingDescriptor", "iftrWriteVaryingDescriptor", "iftrCreateVertexBuffer", "iftrDestroyVertexBuffer",
    "iftrSetVertexBufferLayout", "iftrSetVertexBufferSize", "iftrAllocateVertexBuffer",
    "iftrSetVertexBufferValueFloat4", "ifcrCreateImageFP32", "ifcrDestroyImageFP32", "ifcrGetImageRawDataFP32",
    "ifbufCreateBufferManager", "ifbufDestroyBufferManager", "ifbufCreateBuffer", "ifbufDestroyBuffer",
    "ifbufBufferData", "iftrCreateGpuVertexProgram", "iftrCreateGpuFragmentProgram", "iftrDestroyGpuProgram",
    "iftrSetMsaaSamples", "iftrBindVertexDataFloat", "iftrBindIndexData", "iftrCreateTexture",
    "iftrCopyHostBufferToImage", "iftrUploadTextureLevels", "iftrGenerateFbmTexture", "iftrGenerateMipmap",
    "iftrDownloadTexture", "iftrSetRenderTarget", "iftrGetVisibility", "iftrGetCoverageClass", "iftrGetColorRGBA8",
    "iftrGetFrameStats", "iftrSetStageTiming", "iftrDrawAsync", "iftrSynchronize", "iftrGetLastError",
    "iftrNeuralTrain", "iftrNeuralSetParams", "iftrNeuralGetParams", "iftrNeuralEval", "iftrNeuralValidate",
    "iftrSetViews", "iftrGetViewVisibility", "iftrGetViewColorRGBA8", "iftrSetGraphMode", "iftrSetTuning",
    "iftrGetTuning", "iftrAutotune", "iftrGetDigest", "iftrSetColorFloatOutput", "iftrGetColorRGBA32F",
    "iftrDrawNoAlloc", "iftrNeuralValidateEx", "iftrSetCullMode", "iftrSetRasterizerPolygonMode",
    "iftrCreateBuffer", "iftrCopyHostBufferToBuffer", "iftrCreateGpuMeshProgram", "iftrCreateGpuTaskProgram",
    "iftrBindMeshShader", "iftrBindTaskShader", "iftrDrawMeshTasks", "iftrCreateGpuGeometryProgram",
    "iftrBindGeometryShader", "iftrSetClearValues", "iftrSetScissors",
    "iftrSetScissorTestEnable", "iftrCreateTextureLayers", "iftrCopyHostBufferToImageLayer", "iftrCreateSampler",
    "iftrBlitImage", "iftrSetSetupReciprocal",
]


class NeuralValidation(C.Structure):
    _fields_ = [("psnrDb", C.c_float), ("maxErrLsb", C.c_int32), ("p999ErrLsb", C.c_int32), ("fallback", C.c_int32),
                ("pixels", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

NOALLOC_SCRATCH_BYTES = 48 << 20  # IFHIP_NOALLOC_SCRATCH_BYTES


class NoAllocDraw(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("strideFloats", C.c_int32), ("numVertices", C.c_int32),
                ("indices", C.c_void_p), ("numIndices", C.c_int32), ("primIds", C.c_void_p), ("depth", C.c_void_p),
                ("rgba8", C.c_void_p)]


def lib():
    """Loads libifrit_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing; build it with `python ifrit-v2_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "iftrCreateInstance": ([], vp), "iftrDestroyInstance": ([vp], None), "iftrInit": ([vp], None),
        "iftrBindFrameBuffer": ([vp, vp], None), "iftrBindVertexBuffer": ([vp, vp], None),
        "iftrBindIndexBuffer": ([vp, vp], None), "iftrBindVertexShaderFunc": ([vp, vp, vp], None),
        "iftrBindFragmentShaderFunc": ([vp, vp], None), "iftrSetBlendFunc": ([vp, vp], None),
        "iftrSetDepthFunc": ([vp, C.c_int32], None), "iftrOptsetForceDeterministic": ([vp, C.c_int], None),
        "iftrOptSetDepthTestEnable": ([vp, C.c_int], None), "iftrDrawLegacy": ([vp, C.c_int, C.c_int], None),
        "iftrClear": ([vp], None), "iftrBindVertexShader": ([vp, vp], None),
        "iftrBindFragmentShader": ([vp, vp], None), "iftrCreateFrameBuffer": ([], vp),
        "iftrDestroyFrameBuffer": ([vp], None), "iftrSetFrameBufferColorAttachmentFP32": ([vp, vp, C.c_size_t], None),
        "iftrSetFrameBufferDepthAttachmentFP32": ([vp, vp], None), "iftrCreateVaryingDescriptor": ([], vp),
        "iftrDestroyVaryingDescriptor": ([vp], None), "iftrWriteVaryingDescriptor": ([vp, vp, C.c_size_t], None),
        "iftrCreateVertexBuffer": ([], vp), "iftrDestroyVertexBuffer": ([vp], None),
        "iftrSetVertexBufferLayout": ([vp, vp, C.c_size_t], None), "iftrSetVertexBufferSize": ([vp, C.c_size_t], None),
        "iftrAllocateVertexBuffer": ([vp], None), "iftrSetVertexBufferValueFloat4": ([vp, C.c_int, C.c_int, vp], None),
        "ifcrCreateImageFP32": ([C.c_size_t, C.c_size_t, C.c_size_t], vp), "ifcrDestroyImageFP32": ([vp], None),
        "ifcrGetImageRawDataFP32": ([vp], C.POINTER(C.c_float)), "ifbufCreateBufferManager": ([], vp),
        "ifbufDestroyBufferManager": ([vp], None), "ifbufCreateBuffer": ([vp, C.c_size_t], vp),
        "ifbufDestroyBuffer": ([vp], None), "ifbufBufferData": ([vp, vp, C.c_size_t, C.c_size_t], None),
        "iftrCreateGpuVertexProgram": ([vp], vp), "iftrCreateGpuFragmentProgram": ([vp], vp),
        "iftrDestroyGpuProgram": ([vp], None), "iftrSetMsaaSamples": ([vp, C.c_int], None),
        "iftrBindVertexDataFloat": ([vp, vp, C.c_int, C.c_int], None), "iftrBindIndexData": ([vp, vp, C.c_int], None),
        "iftrCreateTexture": ([vp, C.c_int, C.c_int, C.c_int], C.c_int),
        "iftrCopyHostBufferToImage": ([vp, C.c_int, vp], None), "iftrUploadTextureLevels": ([vp, C.c_int, vp], None),
        "iftrGenerateFbmTexture": ([vp, C.c_int, C.c_uint32], None), "iftrGenerateMipmap": ([vp, C.c_int], None),
        "iftrDownloadTexture": ([vp, C.c_int, vp], None), "iftrSetRenderTarget": ([vp, C.c_int, C.c_int], None),
        "iftrGetVisibility": ([vp, vp, vp], None), "iftrGetCoverageClass": ([vp, vp], None),
        "iftrGetColorRGBA8": ([vp, vp], None), "iftrGetFrameStats": ([vp, C.POINTER(FrameStats)], None),
        "iftrSetStageTiming": ([vp, C.c_int], None), "iftrDrawAsync": ([vp, C.c_int, C.c_int], None),
        "iftrSynchronize": ([vp], None), "iftrGetLastError": ([vp], C.c_char_p),
        "iftrNeuralTrain": ([vp, C.POINTER(NeuralTrainDesc), C.POINTER(NeuralTrainReport)], None),
        "iftrNeuralSetParams": ([vp, vp], None), "iftrNeuralGetParams": ([vp, vp], None),
        "iftrNeuralEval": ([vp, vp, vp, C.c_int, C.c_int], None),
        "iftrNeuralValidate": ([vp, C.c_int, C.c_float, C.c_int], C.c_float),
        "iftrSetViews": ([vp, vp, C.c_int], None), "iftrGetViewVisibility": ([vp, C.c_int, vp, vp], None),
        "iftrGetViewColorRGBA8": ([vp, C.c_int, vp], None), "iftrSetGraphMode": ([vp, C.c_int], None),
        "iftrSetTuning": ([vp, C.POINTER(Tuning)], None), "iftrGetTuning": ([vp, C.POINTER(Tuning)], None),
        "iftrAutotune": ([vp, C.c_int, C.c_int, C.POINTER(AutotuneReport)], None),
        "iftrGetDigest": ([vp], C.c_uint64),
        "iftrSetColorFloatOutput": ([vp, C.c_int], None), "iftrGetColorRGBA32F": ([vp, vp], None),
        "iftrDrawNoAlloc": ([vp, C.POINTER(NoAllocDraw)], None),
        "iftrSetCullMode": ([vp, C.c_int], None), "iftrSetRasterizerPolygonMode": ([vp, C.c_int], None),
        "iftrSetSetupReciprocal": ([vp, C.c_int], None),
        "iftrCreateBuffer": ([vp, C.c_int, C.c_int], C.c_int),
        "iftrCopyHostBufferToBuffer": ([vp, vp, C.c_int, C.c_int], None),
        "iftrCreateGpuMeshProgram": ([vp], vp), "iftrCreateGpuTaskProgram": ([vp], vp),
        "iftrBindMeshShader": ([vp, vp], None), "iftrBindTaskShader": ([vp, vp], None),
        "iftrDrawMeshTasks": ([vp, C.c_int, C.c_int, C.c_int], None),
        "iftrCreateGpuGeometryProgram": ([vp], vp), "iftrBindGeometryShader": ([vp, vp], None),
        "iftrSetClearValues": ([vp, vp, C.c_size_t, C.c_float], None),
        "iftrSetScissors": ([vp, vp, C.c_size_t], None), "iftrSetScissorTestEnable": ([vp, C.c_int], None),
        "iftrCreateTextureLayers": ([vp, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "iftrCopyHostBufferToImageLayer": ([vp, C.c_int, C.c_int, vp], None),
        "iftrCreateSampler": ([vp, C.POINTER(SamplerDesc)], C.c_int),
        "iftrBlitImage": ([vp, C.c_int, C.c_int, C.POINTER(ImageBlit), C.c_int], None),
        "iftrNeuralValidateEx": ([vp, C.c_int, C.c_float, C.c_int, C.c_int, C.POINTER(NeuralValidation)], None),
        "iftrOptsetDepthTestEnable": ([vp, C.c_int], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class IfritError(RuntimeError):
    pass


class Renderer:
    """One renderer instance (one HIP stream) drawing a Scene."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.iftrCreateInstance()
        self.L.iftrInit(self.h)
        self._check()
        self._programs = []
        self.scene = None
        self.samples = 1

    def close(self):
        if self.h:
            self.L.iftrDestroyInstance(self.h)
            self.h = None
        for p in self._programs:
            self.L.iftrDestroyGpuProgram(p)
        self._programs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self):
        e = self.L.iftrGetLastError(self.h)
        if e:
            raise IfritError(e.decode())

    def set_depth_func(self, func: int):
        """IF_COMPARE_OP_* (include/ifrit_hip.h); LESS (1) is the default."""
        self.L.iftrSetDepthFunc(self.h, func)
        self._check()

    def set_deterministic(self, on: bool):
        '''optsetForceDeterministic: False selects the tagging mode's depth expressions at 1 sample per
        pixel (equal-depth ties still resolve to the lowest primitive).'''
        self.L.iftrOptsetForceDeterministic(self.h, 1 if on else 0)
        self._check()

    def create_buffer(self, slot: int, data: np.ndarray):
        '''CreateBuffer + copyHostBufferToBuffer of one array into a slot.'''
        a = np.ascontiguousarray(data)
        self.L.iftrCreateBuffer(self.h, slot, max(a.nbytes, 4))
        self._check()
        self.L.iftrCopyHostBufferToBuffer(self.h, a.ctypes.data, slot, a.nbytes)
        self._check()

    def bind_meshlets(self, mvp, record_floats: int, slots=(0, 1, 2, 3), task: bool = True):
        '''Binds the built-in meshlet mesh program (and the one-mesh-group task program).'''
        d = MeshProgramDesc(1, (C.c_float * 16)(*np.asarray(mvp, np.float32).reshape(-1)), *slots, record_floats)
        p = self.L.iftrCreateGpuMeshProgram(C.byref(d))
        self._programs.append(p)
        self.L.iftrBindMeshShader(self.h, p)
        if task:
            t = (C.c_int32 * 1)(1)
            tp = self.L.iftrCreateGpuTaskProgram(t)
            self._programs.append(tp)
            self.L.iftrBindTaskShader(self.h, tp)
        self._check()

    def draw_mesh_tasks(self, groups: int, first: int = 0, clear: bool = True):
        self.L.iftrDrawMeshTasks(self.h, groups, first, 1 if clear else 0)
        self._check()
        st = self.stats()
        if st["status"]:
            raise IfritError(f"frame incomplete, status bits {st['status']:#x}")
        return st

    def bind_clip_offset_geometry(self, offset=(0.03, 0.0, 0.0, 0.0)):
        '''Binds the built-in geometry program that moves every clip position by offset (None unbinds).'''
        if offset is None:
            self.L.iftrBindGeometryShader(self.h, None)
        else:
            d = GeometryProgramDesc(1, (C.c_float * 4)(*[float(x) for x in offset]))
            p = self.L.iftrCreateGpuGeometryProgram(C.byref(d))
            self._programs.append(p)
            self.L.iftrBindGeometryShader(self.h, p)
        self._check()

    def set_polygon_mode(self, mode: int):
        '''IF_POLYGON_MODE_*: 0 fill (the default), 1 line, 2 point.'''
        self.L.iftrSetRasterizerPolygonMode(self.h, mode)
        self._check()

    def set_cull_mode(self, mode: int):
        """IF_CULL_MODE_*: 0 none, 1 front, 2 back (the default), 3 front and back."""
        self.L.iftrSetCullMode(self.h, mode)
        self._check()

    RCP_MODELS = {"intel": 0, "amd": 1, "host": 2}

    def set_setup_reciprocal(self, model="intel"):
        """Whose RCPPS the triangle setup reproduces: "intel" (the default), "amd", or "host" (the
        CPU this process runs on).  The instruction is implementation-defined, and the reference's
        setup values follow the CPU it runs on."""
        self.L.iftrSetSetupReciprocal(self.h, self.RCP_MODELS[model] if isinstance(model, str) else int(model))
        self._check()

    def set_clear_values(self, color=(0.0, 0.0, 0.0, 0.0), depth=None):
        '''Colour and depth a draw with clear starts from (depth None: the CPU path's 0x7f7f7f7f).'''
        d = float(np.frombuffer(np.uint32(0x7F7F7F7F).tobytes(), np.float32)[0]) if depth is None else depth
        c = np.ascontiguousarray(color, np.float32).reshape(1, 4)
        self.L.iftrSetClearValues(self.h, _p(c), 1, d)
        self._check()

    def set_scissors(self, rects, enable=True):
        '''Scissor rectangles (x, y, width, height) in pixels and the test on/off.'''
        r = np.ascontiguousarray(rects, np.float32).reshape(-1, 4)
        self.L.iftrSetScissors(self.h, _p(r), r.shape[0])
        self.L.iftrSetScissorTestEnable(self.h, 1 if enable else 0)
        self._check()

    def set_blend(self, factors=None):
        """factors = (src colour, dst colour, src alpha, dst alpha) as IfritBlendFactor values, or None
        to draw without blending."""
        st = BlendState(0, 0, 0, 0, 0) if factors is None else BlendState(1, *[int(f) for f in factors])
        self.L.iftrSetBlendFunc(self.h, C.byref(st))
        self._check()

    def set_float_output(self, on: bool):
        self.L.iftrSetColorFloatOutput(self.h, 1 if on else 0)
        self._check()

    def color_f32(self):
        out = np.empty((self.height, self.width, 4), np.float32)
        self.L.iftrGetColorRGBA32F(self.h, _p(out))
        self._check()
        return out

    def set_target(self, width: int, height: int, samples: int = 1):
        self.L.iftrSetRenderTarget(self.h, width, height)
        self.L.iftrSetMsaaSamples(self.h, samples)
        self.width, self.height, self.samples = width, height, samples
        self._check()

    def bind_geometry(self, vertices: np.ndarray, indices: np.ndarray, mvp: np.ndarray, clip_passthrough=False):
        v = np.ascontiguousarray(vertices, np.float32)
        i = np.ascontiguousarray(indices, np.int32)
        self.L.iftrBindVertexDataFloat(self.h, _p(v), v.shape[1], v.shape[0])
        self.L.iftrBindIndexData(self.h, _p(i), i.size)
        d = VertexProgramDesc()
        d.kind = 1 if clip_passthrough else 0
        d.mvp[:] = [float(x) for x in np.asarray(mvp, np.float32).reshape(-1)]
        p = self.L.iftrCreateGpuVertexProgram(C.byref(d))
        self._programs.append(p)
        self.L.iftrBindVertexShader(self.h, p)
        self.num_indices = int(i.size)
        self._check()

    def create_sampler(self, filter=1, address_u=0, address_v=0, border=0, anisotropy=False, max_anisotropy=1.0):
        '''CreateSampler: IF_FILTER_* within a level, IF_SAMPLER_ADDRESS_MODE_* per axis, IF_BORDER_COLOR_*,
        anisotropic taps up to max_anisotropy.  Returns the sampler id.'''
        d = SamplerDesc(filter, address_u, address_v, 0, border, 1 if anisotropy else 0, max_anisotropy)
        sid = self.L.iftrCreateSampler(self.h, C.byref(d))
        self._check()
        return sid

    def blit(self, src, dst, src_region, dst_region, src_level=0, dst_level=0, src_layer=0, dst_layer=0, linear=True):
        '''blitImage: regions are (x0, y0, x1, y1) of the given levels and layers.'''
        b = ImageBlit(src_level, src_layer, (C.c_uint32 * 3)(src_region[0], src_region[1], 0),
                      (C.c_uint32 * 3)(src_region[2], src_region[3], 1), dst_level, dst_layer,
                      (C.c_uint32 * 3)(dst_region[0], dst_region[1], 0), (C.c_uint32 * 3)(dst_region[2], dst_region[3], 1))
        self.L.iftrBlitImage(self.h, src, dst, C.byref(b), 1 if linear else 0)
        self._check()

    def bind_fragment_program(self, scene, program=None, precision=0):
        d = FragmentProgramDesc()
        d.kind = scene.program if program is None else program
        d.precision = precision
        d.samplerId = getattr(scene, "sampler_id", 0)
        d.lod = getattr(scene, "lod", 0.0)
        dw = getattr(scene, "depth_write", None)  # (scale, bias): the program writes depth * scale + bias
        if dw is not None:
            d.depthWrite, d.depthScale, d.depthBias = 1, float(dw[0]), float(dw[1])
        d.offWorld, d.offNormal, d.offUv = 0, 3, 6
        d.offTangent = 8 if scene.stride >= 11 else 0
        d.trisPerMaterial = scene.tris_per_material
        d.sampler, d.maxAniso = scene.sampler, scene.max_aniso
        d.camera[:] = [float(x) for x in scene.camera]
        d.ambient, d.kd, d.ks, d.shininess = scene.ambient, scene.kd, scene.ks, scene.shininess
        d.numLights = len(scene.lights)
        for k, (lp, lc) in enumerate(scene.lights):
            d.lightPos[k][:] = list(lp)
            d.lightColor[k][:] = list(lc)
        p = self.L.iftrCreateGpuFragmentProgram(C.byref(d))
        self._programs.append(p)
        self.L.iftrBindFragmentShader(self.h, p)
        self._check()

    def train_neural(self, views=512, view_size=(256, 256), max_fragments=16 << 20, steps=2000, batch=65536,
                     lr_latent=1e-2, lr_mlp=2e-3, seed=1, latent_size=1024, eye_dist=(20.0, 45.0),
                     fov_y=1.0471975511965976, target_jitter=6.0, fov_y_min=0.14, single_tap=False) -> dict:
        """Trains the neural material on the bound layered-GGX program (see iftrNeuralTrain)."""
        d = NeuralTrainDesc(views, view_size[0], view_size[1], max_fragments, steps, batch, lr_latent, lr_mlp, seed,
                            latent_size, eye_dist[0], eye_dist[1], fov_y, target_jitter, fov_y_min,
                            1 if single_tap else 0)
        r = NeuralTrainReport()
        self.L.iftrNeuralTrain(self.h, C.byref(d), C.byref(r))
        self._check()
        return r.as_dict()

    def validate_neural(self, precision: int = 0, min_psnr_db: float = 40.0) -> float:
        """PSNR of the neural frame against the layered-GGX frame of the current view; below the bound
        the neural program at this precision falls back to the material."""
        v = self.L.iftrNeuralValidate(self.h, precision, min_psnr_db, self.num_indices)
        self._check()
        return float(v)

    def validate_neural_ex(self, precision: int = 0, min_psnr_db: float = 40.0, max_err_lsb: int = 255) -> dict:
        """PSNR, largest and 99.9th-percentile per-pixel error of the neural frame against the
        layered-GGX frame; the neural program falls back below the PSNR bound or above the
        per-pixel bound."""
        v = NeuralValidation()
        self.L.iftrNeuralValidateEx(self.h, precision, min_psnr_db, max_err_lsb, self.num_indices, C.byref(v))
        self._check()
        return v.as_dict()

    def set_neural_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32).reshape(-1)
        assert p.size == NEURAL_PARAMS
        self.L.iftrNeuralSetParams(self.h, _p(p))
        self._check()

    def get_neural_params(self) -> np.ndarray:
        p = np.empty(NEURAL_PARAMS, np.float32)
        self.L.iftrNeuralGetParams(self.h, _p(p))
        self._check()
        return p

    def eval_neural(self, inputs: np.ndarray, precision: int = 0) -> np.ndarray:
        x = np.zeros((inputs.shape[0], 16), np.float32)
        x[:, :inputs.shape[1]] = inputs
        out = np.empty((inputs.shape[0], 4), np.float32)
        self.L.iftrNeuralEval(self.h, _p(x), _p(out), x.shape[0], precision)
        self._check()
        return out[:, :3]

    def add_texture(self, tex) -> int:
        tid = self.L.iftrCreateTextureLayers(self.h, tex.width, tex.height, tex.levels, getattr(tex, "layers", 1))
        self._check()
        if tex.texels is not None:
            arr = np.ascontiguousarray(tex.texels, np.uint8)
            self.L.iftrUploadTextureLevels(self.h, tid, _p(arr))
        else:
            self.L.iftrGenerateFbmTexture(self.h, tid, tex.fbm_seed)
            self.L.iftrGenerateMipmap(self.h, tid)
        self._check()
        return tid

    def download_texture(self, tid: int, tex) -> np.ndarray:
        n = sum(max(1, tex.width >> l) * max(1, tex.height >> l) for l in range(tex.levels)) * 4 * getattr(tex, "layers", 1)
        out = np.empty((n,), np.uint8)
        self.L.iftrDownloadTexture(self.h, tid, _p(out))
        self._check()
        return out

    def load_scene(self, scene):
        self.scene = scene
        self.set_target(scene.width, scene.height, scene.samples)
        self.bind_geometry(scene.vertices, scene.indices, scene.mvp)
        for t in scene.textures:
            self.add_texture(t)
        if scene.sampler_state is not None:
            scene.sampler_id = self.create_sampler(**scene.sampler_state)
        self.bind_fragment_program(scene)

    def draw(self, clear: bool = True):
        self.L.iftrDrawLegacy(self.h, self.num_indices, 1 if clear else 0)
        self._check()
        st = self.stats()
        if st["status"]:
            raise IfritError(f"frame incomplete, status bits {st['status']:#x}")
        return st

    def draw_async(self, clear: bool = True):
        self.L.iftrDrawAsync(self.h, self.num_indices, 1 if clear else 0)

    def synchronize(self):
        self.L.iftrSynchronize(self.h)
        self._check()
        st = self.stats()
        if st["status"]:
            raise IfritError(f"frame incomplete, status bits {st['status']:#x}")

    def stage_timing(self, on: bool):
        self.L.iftrSetStageTiming(self.h, 1 if on else 0)

    def stats(self) -> dict:
        s = FrameStats()
        self.L.iftrGetFrameStats(self.h, C.byref(s))
        return s.as_dict()

    def visibility(self):
        n = self.width * self.height * self.samples
        ids = np.empty((n,), np.uint32)
        depth = np.empty((n,), np.float32)
        self.L.iftrGetVisibility(self.h, _p(ids), _p(depth))
        self._check()
        return ids, depth

    def coverage_class(self):
        out = np.empty((self.width * self.height,), np.uint8)
        self.L.iftrGetCoverageClass(self.h, _p(out))
        self._check()
        return out

    def color(self):
        out = np.empty((self.height, self.width, 4), np.uint8)
        self.L.iftrGetColorRGBA8(self.h, _p(out))
        self._check()
        return out

    # multi-view batches
    def set_views(self, views):
        """views: sequence of dicts with width, height, mvp (4x4), camera (3) and shade (bool); an
        empty sequence returns to the single target."""
        arr = (ViewDesc * max(1, len(views)))()
        for i, v in enumerate(views):
            arr[i].width = int(v["width"])
            arr[i].height = int(v["height"])
            arr[i].mvp[:] = [float(x) for x in np.asarray(v["mvp"], np.float32).reshape(-1)]
            arr[i].camera[:] = [float(x) for x in np.asarray(v.get("camera", (0, 0, 0)), np.float32)]
            arr[i].shade = 1 if v.get("shade", True) else 0
        self.L.iftrSetViews(self.h, C.cast(arr, C.c_void_p), len(views))
        self._check()
        self.views = [dict(v) for v in views]

    def view_visibility(self, view: int):
        v = self.views[view]
        n = int(v["width"]) * int(v["height"]) * self.samples
        ids = np.empty((n,), np.uint32)
        depth = np.empty((n,), np.float32)
        self.L.iftrGetViewVisibility(self.h, view, _p(ids), _p(depth))
        self._check()
        return ids, depth

    def view_color(self, view: int):
        v = self.views[view]
        out = np.empty((int(v["height"]), int(v["width"]), 4), np.uint8)
        self.L.iftrGetViewColorRGBA8(self.h, view, _p(out))
        self._check()
        return out

    def set_graph_mode(self, on: bool):
        self.L.iftrSetGraphMode(self.h, 1 if on else 0)
        self._check()

    def set_tuning(self, **kw):
        t = Tuning()
        self.L.iftrGetTuning(self.h, C.byref(t))
        for k, v in kw.items():
            setattr(t, k, v)
        self.L.iftrSetTuning(self.h, C.byref(t))
        self._check()

    def tuning(self) -> dict:
        t = Tuning()
        self.L.iftrGetTuning(self.h, C.byref(t))
        return t.as_dict()

    def autotune(self, frames: int = 3) -> dict:
        rep = AutotuneReport()
        self.L.iftrAutotune(self.h, self.num_indices, frames, C.byref(rep))
        self._check()
        return rep.as_dict()

    def digest(self) -> int:
        d = self.L.iftrGetDigest(self.h)
        self._check()
        return int(d)
