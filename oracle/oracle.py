"""ctypes bindings of the oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module; the product never does.  It wraps
  * oracle/_build/libifrit_oracle.so  -- the CPU restatement (oracle_core.cpp,
    oracle_shade.cpp), built by oracle/Makefile;
  * oracle/_ref/libifrit_ref.so       -- the reference CPU rasterizer compiled
    from its own sources by oracle/ref/Makefile (optional: absent where the
    reference sources are not available).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libifrit_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libifrit_ref.so")
REFERENCE_ROOT = "/root/reference"


CLEAR_DEPTH = float(np.frombuffer(np.uint32(0x7F7F7F7F).tobytes(), np.float32)[0])  # TileRasterWorker.cpp:40


def _rcp_mode(live_rcp, rcp_table):
    """RCPSS of the setup: this CPU's own instruction (live_rcp), else the captured Intel or AMD table."""
    if live_rcp:
        return 1
    return {"intel": 0, "amd": 2}[rcp_table]


class RasterConfig(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("samples", C.c_int), ("ccw", C.c_int),
                ("depth_func", C.c_int), ("clear", C.c_int), ("num_threads", C.c_int), ("live_rcp", C.c_int),
                ("cull_mode", C.c_int), ("clear_depth", C.c_float), ("tagging", C.c_int), ("depth_write", C.c_int),
                ("depth_scale", C.c_float), ("depth_bias", C.c_float)]


class Stats(C.Structure):
    _fields_ = [("input_triangles", C.c_int64), ("assembled_triangles", C.c_int64), ("clipped_inputs", C.c_int64),
                ("bin_entries", C.c_int64), ("fragments", C.c_int64), ("covered_samples", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Tri(C.Structure):
    _fields_ = [("f1", C.c_float * 4), ("f2", C.c_float * 4), ("f3", C.c_float * 4), ("bx", C.c_float * 3),
                ("by", C.c_float * 3), ("e1", C.c_float * 3), ("e2", C.c_float * 3), ("e3", C.c_float * 3),
                ("bbox", C.c_float * 4), ("prim", C.c_int32), ("fan", C.c_int32), ("swapped", C.c_int32)]


TRI_DTYPE = np.dtype([("f1", "<f4", 4), ("f2", "<f4", 4), ("f3", "<f4", 4), ("bx", "<f4", 3), ("by", "<f4", 3),
                      ("e1", "<f4", 3), ("e2", "<f4", 3), ("e3", "<f4", 3), ("bbox", "<f4", 4), ("prim", "<i4"),
                      ("fan", "<i4"), ("swapped", "<i4")])


class ShadeParams(C.Structure):
    _fields_ = [("program", C.c_int), ("stride_floats", C.c_int), ("off_world", C.c_int), ("off_normal", C.c_int),
                ("off_uv", C.c_int), ("tris_per_material", C.c_int), ("sampler", C.c_int), ("max_aniso", C.c_int),
                ("camera", C.c_float * 3), ("ambient", C.c_float), ("kd", C.c_float), ("ks", C.c_float),
                ("shininess", C.c_float), ("num_lights", C.c_int), ("light_pos", (C.c_float * 4) * 4),
                ("light_color", (C.c_float * 4) * 4), ("off_tangent", C.c_int), ("clear_color", C.c_float * 4),
                ("addr_u", C.c_int), ("addr_v", C.c_int), ("nearest", C.c_int), ("border", C.c_float * 4),
                ("lod", C.c_float)]


class TextureDesc(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("levels", C.c_int),
                ("layers", C.c_int)]


def sampler_params(state):
    """(sampler kind, max taps, addr_u, addr_v, nearest, border rgba) of a Renderer.create_sampler
    argument dict (the library's mapping: anisotropic taps when enabled, else trilinear)."""
    import math

    aniso = bool(state.get("anisotropy", False))
    border = (1.0, 1.0, 1.0, 1.0) if state.get("border", 0) == 1 else (0.0, 0.0, 0.0, 1.0)
    return (2 if aniso else 1, max(1, min(16, math.ceil(state.get("max_anisotropy", 1.0)))),
            state.get("address_u", 0), state.get("address_v", 0), 1 if state.get("filter", 1) == 0 else 0, border)


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference build when its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    if os.path.isdir(REFERENCE_ROOT):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(HERE, "ref")] + (["-B"] if force else []), check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.oracle_frame_create.restype = C.c_void_p
        _lib.oracle_frame_destroy.argtypes = [C.c_void_p]
        _lib.oracle_num_tris.argtypes = [C.c_void_p]
        _lib.oracle_get_tris.argtypes = [C.c_void_p, C.c_void_p]
        _lib.oracle_render.argtypes = [C.c_void_p, C.POINTER(RasterConfig), C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Stats)]
        _lib.oracle_render_blend.argtypes = [C.c_void_p, C.POINTER(RasterConfig), C.c_void_p, C.c_int, C.c_void_p,
                                             C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p]
        _lib.oracle_render_polygons.argtypes = [C.c_void_p, C.POINTER(RasterConfig), C.c_int, C.c_void_p, C.c_int,
                                                C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.POINTER(Stats)]
        _lib.oracle_vertex_transform.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        _lib.oracle_shade.argtypes = [C.c_void_p, C.POINTER(ShadeParams), C.c_int, C.c_int, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        _lib.oracle_flat_hash.argtypes = [C.c_uint32]
        _lib.oracle_flat_hash.restype = C.c_uint32
        _lib.oracle_build_mips.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        _lib.oracle_sample_positions.argtypes = [C.c_int, C.c_void_p]
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def host_cpu() -> str:
    """Vendor and model of this host's CPU (its RCPPS decides the reference's setup values)."""
    model = vendor = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("vendor_id") and vendor == "unknown":
                    vendor = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return f"{vendor} {model}"


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
        _ref.ref_render_visibility.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                               C.c_int, C.c_void_p, C.c_void_p]
        _ref.ref_render_timed.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                          C.c_void_p, C.c_int, C.c_void_p]
        _ref.ref_render_timed.restype = C.c_double
        _ref.ref_render_blend.argtypes = [C.c_int] * 7 + [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        if hasattr(_ref, "ref_render_visibility_depthwrite"):
            _ref.ref_render_visibility_depthwrite.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                                              C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        if hasattr(_ref, "ref_render_visibility_tagging"):
            _ref.ref_render_visibility_tagging.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                                           C.c_void_p, C.c_void_p]
    return _ref


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def vertex_transform(vertices: np.ndarray, mvp: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32)
    m = np.ascontiguousarray(mvp, np.float32)
    out = np.empty((v.shape[0], 4), np.float32)
    lib().oracle_vertex_transform(_p(v), v.shape[1], v.shape[0], _p(m), _p(out))
    return out


def sample_positions(samples: int) -> np.ndarray:
    out = np.empty((samples, 2), np.float32)
    lib().oracle_sample_positions(samples, _p(out))
    return out


class Frame:
    """One oracle frame: visibility outputs plus the assembled triangles."""

    def __init__(self):
        self._h = lib().oracle_frame_create()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_frame_destroy(self._h)
            self._h = None

    def render(self, width, height, samples, clip, indices, depth_func=1, ccw=0, threads=None, live_rcp=False,
               clear=True, depth_in=None, cull_mode=2, clear_depth=CLEAR_DEPTH, tagging=False, depth_write=None,
               rcp_table="intel"):
        clip = np.ascontiguousarray(clip, np.float32)
        indices = np.ascontiguousarray(indices, np.int32)
        n = width * height
        ids = np.empty((n * samples,), np.uint32)
        depth = np.empty((n * samples,), np.float32)
        tri = np.empty((n * samples,), np.uint32)
        cls = np.empty((n,), np.uint8)
        st = Stats()
        cfg = RasterConfig(width, height, samples, ccw, depth_func, 1 if clear else 0,
                           threads or os.cpu_count() or 1, _rcp_mode(live_rcp, rcp_table), cull_mode, clear_depth,
                           1 if tagging else 0, *_depth_write(depth_write))
        din = None if depth_in is None else np.ascontiguousarray(depth_in, np.float32)
        rc = lib().oracle_render(self._h, C.byref(cfg), _p(clip), clip.shape[0], _p(indices), indices.size,
                                 None if din is None else _p(din), _p(ids), _p(depth), _p(tri), _p(cls), C.byref(st))
        if rc != 0:
            raise RuntimeError(f"oracle_render failed ({rc})")
        self.width, self.height, self.samples = width, height, samples
        self.ids, self.depth, self.tri, self.cls, self.stats = ids, depth, tri, cls, st.as_dict()
        return self

    def render_blend(self, width, height, clip, indices, factors, depth_func=1, threads=None, clear=True,
                     depth_in=None, rgba_in=None, live_rcp=False, depth_write=None, rcp_table="intel"):
        """Blending of the translucent flat colour in primitive order (1 sample per pixel);
        factors = (src colour, dst colour, src alpha, dst alpha) as IfritBlendFactor values."""
        clip = np.ascontiguousarray(clip, np.float32)
        indices = np.ascontiguousarray(indices, np.int32)
        n = width * height
        ids = np.empty((n,), np.uint32)
        depth = np.empty((n,), np.float32)
        rgba = np.empty((n, 4), np.float32)
        cfg = RasterConfig(width, height, 1, 0, depth_func, 1 if clear else 0, threads or os.cpu_count() or 1,
                           _rcp_mode(live_rcp, rcp_table), 2, CLEAR_DEPTH, 0, *_depth_write(depth_write))
        bl = (C.c_int * 4)(*factors)
        din = None if depth_in is None else np.ascontiguousarray(depth_in, np.float32)
        cin = None if rgba_in is None else np.ascontiguousarray(rgba_in, np.float32)
        rc = lib().oracle_render_blend(self._h, C.byref(cfg), _p(clip), clip.shape[0], _p(indices), indices.size, bl,
                                       None if din is None else _p(din), None if cin is None else _p(cin), _p(ids),
                                       _p(depth), _p(rgba))
        if rc != 0:
            raise RuntimeError(f"oracle_render_blend failed ({rc})")
        self.ids, self.depth, self.rgba = ids, depth, rgba.reshape(height, width, 4)
        return self

    def render_polygons(self, width, height, clip, indices, mode, threads=None, clear=True, depth_in=None,
                        clear_depth=CLEAR_DEPTH):
        """Polygon mode 1 (lines) or 2 (points), 1 sample per pixel."""
        clip = np.ascontiguousarray(clip, np.float32)
        indices = np.ascontiguousarray(indices, np.int32)
        n = width * height
        ids = np.empty((n,), np.uint32)
        depth = np.empty((n,), np.float32)
        tri = np.empty((n,), np.uint32)
        st = Stats()
        cfg = RasterConfig(width, height, 1, 0, 1, 1 if clear else 0, threads or os.cpu_count() or 1, 0, 2,
                           clear_depth, 0, 0, 1.0, 0.0)
        din = None if depth_in is None else np.ascontiguousarray(depth_in, np.float32)
        rc = lib().oracle_render_polygons(self._h, C.byref(cfg), mode, _p(clip), clip.shape[0], _p(indices),
                                          indices.size, None if din is None else _p(din), _p(ids), _p(depth), _p(tri),
                                          C.byref(st))
        if rc != 0:
            raise RuntimeError(f"oracle_render_polygons failed ({rc})")
        self.width, self.height, self.samples = width, height, 1
        self.ids, self.depth, self.tri, self.stats = ids, depth, tri, st.as_dict()
        self.cls = np.zeros((n,), np.uint8)
        return self

    def tris(self) -> np.ndarray:
        n = lib().oracle_num_tris(self._h)
        out = np.zeros((n,), TRI_DTYPE)
        if n:
            lib().oracle_get_tris(self._h, _p(out))
        return out

    def shade(self, scene, threads=None, clear_color=(0.0, 0.0, 0.0, 0.0)):
        """RGBA8 colour and boundary flags from the FP64 shading oracle; samples no triangle covers
        count as clear_color."""
        p = ShadeParams()
        p.clear_color[:] = [float(x) for x in clear_color]
        p.program = scene.program
        p.stride_floats = scene.stride
        p.off_world, p.off_normal, p.off_uv = 0, 3, 6
        p.off_tangent = 8 if scene.stride >= 11 else 0
        p.tris_per_material = scene.tris_per_material
        p.sampler, p.max_aniso = scene.sampler, scene.max_aniso
        p.border[:] = [0.0, 0.0, 0.0, 1.0]
        if getattr(scene, "sampler_state", None) is not None:
            kind, taps, au, av, nearest, border = sampler_params(scene.sampler_state)
            p.sampler, p.max_aniso, p.addr_u, p.addr_v, p.nearest = kind, taps, au, av, nearest
            p.border[:] = list(border)
        p.lod = getattr(scene, "lod", 0.0)
        p.camera[:] = list(scene.camera)
        p.ambient, p.kd, p.ks, p.shininess = scene.ambient, scene.kd, scene.ks, scene.shininess
        p.num_lights = len(scene.lights)
        for i, (lp, lc) in enumerate(scene.lights):
            p.light_pos[i][:] = list(lp)
            p.light_color[i][:] = list(lc)
        descs = (TextureDesc * max(1, len(scene.textures)))()
        keep = []
        for i, t in enumerate(scene.textures):
            if t.texels is None:
                raise ValueError("texture texels must be materialised for the shading oracle")
            arr = np.ascontiguousarray(t.texels, np.uint8)
            keep.append(arr)
            descs[i] = TextureDesc(arr.ctypes.data, t.width, t.height, t.levels, getattr(t, "layers", 1))
        n = self.width * self.height
        rgba = np.empty((n * 4,), np.uint8)
        flags = np.empty((n,), np.uint8)
        verts = np.ascontiguousarray(scene.vertices, np.float32)
        idx = np.ascontiguousarray(scene.indices, np.int32)
        lib().oracle_shade(self._h, C.byref(p), self.width, self.height, self.samples, _p(self.tri), _p(idx),
                           _p(verts), descs, len(scene.textures), _p(rgba), _p(flags), threads or os.cpu_count() or 1)
        return rgba.reshape(self.height, self.width, 4), flags.reshape(self.height, self.width)


def _depth_write(dw):
    """(scale, bias) of a depth-writing fragment program, or None."""
    return (0, 1.0, 0.0) if dw is None else (1, float(dw[0]), float(dw[1]))


def render_scene(scene, threads=None, live_rcp=False, cull_mode=2, clear_depth=CLEAR_DEPTH, tagging=False,
                 rcp_table="intel"):
    clip = vertex_transform(scene.vertices, scene.mvp)
    f = Frame().render(scene.width, scene.height, scene.samples, clip, scene.indices, threads=threads,
                       live_rcp=live_rcp, cull_mode=cull_mode, clear_depth=clear_depth, tagging=tagging,
                       rcp_table=rcp_table,
                       depth_write=getattr(scene, "depth_write", None))
    return f, clip


def ref_visibility(width, height, clip, indices, depth_func=1):
    clip = np.ascontiguousarray(clip, np.float32)
    indices = np.ascontiguousarray(indices, np.int32)
    ids = np.empty((width * height,), np.uint32)
    depth = np.empty((width * height,), np.float32)
    bad = ref().ref_render_visibility(width, height, 0, depth_func, _p(clip), clip.shape[0], _p(indices),
                                      indices.size, _p(ids), _p(depth))
    return ids, depth, bad


def ref_visibility_tagging(width, height, clip, indices, depth_func=1):
    """The reference build drawing in its tagging mode (optsetForceDeterministic(false)): primitive
    ID per pixel and the depth attachment (the tagging mode writes depth back at the right pixels)."""
    clip = np.ascontiguousarray(clip, np.float32)
    indices = np.ascontiguousarray(indices, np.int32)
    ids = np.empty((width * height,), np.uint32)
    depth = np.empty((width * height,), np.float32)
    bad = ref().ref_render_visibility_tagging(width, height, depth_func, _p(clip), _p(indices), indices.size,
                                              _p(ids), _p(depth))
    return ids, depth, bad


def ref_visibility_depthwrite(width, height, clip, indices, scale, bias, depth_func=1):
    """The reference build with a fragment shader that writes depth * scale + bias through its depth
    out-parameter: primitive ID per pixel and the written depth (read back from the colour)."""
    clip = np.ascontiguousarray(clip, np.float32)
    indices = np.ascontiguousarray(indices, np.int32)
    ids = np.empty((width * height,), np.uint32)
    depth = np.empty((width * height,), np.float32)
    bad = ref().ref_render_visibility_depthwrite(width, height, depth_func, scale, bias, _p(clip), _p(indices),
                                                 indices.size, _p(ids), _p(depth))
    return ids, depth, bad


def ggx_light(n, v, l, albedo, alpha, kd, ks):
    """The layered-GGX material's one-light term (FP64), times n.l."""
    L = lib()
    f = L.oracle_ggx_light
    f.restype = None
    a = [np.ascontiguousarray(x, np.float64) for x in (n, v, l, albedo)]
    out = np.zeros(3, np.float64)
    f(*[_p(x) for x in a], C.c_double(alpha), C.c_double(kd), C.c_double(ks), _p(out))
    return out


def ref_blend(width, height, clip, indices, factors, depth_func=1):
    """FP32 colour attachment of the reference build drawing the translucent flat colour with blending."""
    clip = np.ascontiguousarray(clip, np.float32)
    indices = np.ascontiguousarray(indices, np.int32)
    out = np.empty((height, width, 4), np.float32)
    rc = ref().ref_render_blend(width, height, depth_func, *[int(f) for f in factors], _p(clip), _p(indices),
                                indices.size, _p(out))
    if rc != 0:
        raise RuntimeError(f"ref_render_blend failed ({rc})")
    return out


def sample_tex(texels, width, height, levels, sampler, max_aniso, u, v, dudx, dvdx, dudy, dvdy):
    """One FP64 texture lookup of the oracle's sampler (RGBA8 texels with the
    packed mip chain); returns (rgba in 0..1, too-close-to-call flag)."""
    L = lib()
    f = L.oracle_sample_tex
    f.restype = C.c_int
    t = np.ascontiguousarray(texels, np.uint8)
    out = np.zeros(4, np.float64)
    flag = f(_p(t), C.c_int(width), C.c_int(height), C.c_int(levels), C.c_int(sampler), C.c_int(max_aniso),
             *[C.c_double(x) for x in (u, v, dudx, dvdx, dudy, dvdy)], _p(out))
    return out, bool(flag)


def sample_tex_mode(texels, width, height, levels, sampler, max_aniso, addr_u, addr_v, nearest, border, u, v, dudx,
                    dvdx, dudy, dvdy):
    """One FP64 lookup under a sampler's addressing, filter and border colour."""
    L = lib()
    f = L.oracle_sample_tex_mode
    f.restype = C.c_int
    t = np.ascontiguousarray(texels, np.uint8)
    b = np.ascontiguousarray(border, np.float32)
    out = np.zeros(4, np.float64)
    flag = f(_p(t), C.c_int(width), C.c_int(height), C.c_int(levels), C.c_int(sampler), C.c_int(max_aniso),
             C.c_int(addr_u), C.c_int(addr_v), C.c_int(nearest), _p(b),
             *[C.c_double(x) for x in (u, v, dudx, dvdx, dudy, dvdy)], _p(out))
    return out, bool(flag)


def cube_lookup(texels, size, levels, d, lod, addr_u=2, addr_v=2, nearest=0, border=(0, 0, 0, 1)):
    """FP64 cube-map lookup of a six-layer texture along direction d; returns (rgba, too-close flag)."""
    L = lib()
    f = L.oracle_cube_lookup
    f.restype = C.c_int
    t = np.ascontiguousarray(texels, np.uint8)
    b = np.ascontiguousarray(border, np.float32)
    dd = np.ascontiguousarray(d, np.float64)
    out = np.zeros(4, np.float64)
    flag = f(_p(t), C.c_int(size), C.c_int(levels), C.c_int(addr_u), C.c_int(addr_v), C.c_int(nearest), _p(b), _p(dd),
             C.c_double(lod), _p(out))
    return out, bool(flag)


def flat_hash(prim: int) -> int:
    return lib().oracle_flat_hash(prim)


def blit(src, dst, src_region, dst_region, linear=True):
    """blitImage (TileRasterImageOpInvocationsCuda.cu:36-82) restated in float32 on the 0..255 values:
    src, dst are (h, w, 4) uint8 levels, regions (x0, y0, x1, y1).  Destination texel (dx0 + x, dy0 + y)
    reads the source at x / (dx1 - dx0) * (sx1 - sx0) + sx0 (likewise y); linear blends the four
    neighbours (right and lower clamped to the level) and rounds half up, nearest takes the texel at
    the rounded (half away from zero) coordinate, clamped.  Returns the new destination."""
    f32 = np.float32
    sx0, sy0, sx1, sy1 = src_region
    dx0, dy0, dx1, dy1 = dst_region
    sh, sw = src.shape[:2]
    out = dst.copy()
    rw, rh = dx1 - dx0, dy1 - dy0
    y, x = np.mgrid[0:rh, 0:rw]
    px = x.astype(f32) / f32(rw)
    py = y.astype(f32) / f32(rh)
    cx = px * f32(sx1 - sx0) + f32(sx0)
    cy = py * f32(sy1 - sy0) + f32(sy0)
    S = src.astype(f32)
    if linear:
        ix, iy = cx.astype(np.int64), cy.astype(np.int64)  # truncation: the coordinates are >= 0
        fx, fy = cx - ix.astype(f32), cy - iy.astype(f32)
        ix1, iy1 = np.minimum(ix + 1, sw - 1), np.minimum(iy + 1, sh - 1)
        fx, fy = fx[..., None], fy[..., None]

        def lerp(p, q, t):
            return p + (q - p) * t

        v = lerp(lerp(S[iy, ix], S[iy, ix1], fx), lerp(S[iy1, ix], S[iy1, ix1], fx), fy)
        res = np.clip(np.floor(v + f32(0.5)), 0, 255).astype(np.uint8)
    else:
        def rnd(c):
            fl = np.floor(c)
            return (fl + (c - fl >= f32(0.5))).astype(np.int64)

        res = src[np.minimum(rnd(cy), sh - 1), np.minimum(rnd(cx), sw - 1)]
    out[dy0:dy1, dx0:dx1] = res
    return out


def geometry_clip_offset(vertices, indices, mvp, offset, passthrough=False):
    """The geometry stage (TileRasterCoreInvocationCuda.cu:502-546) running the demo's geometry
    shader (demo/shader/DefaultDemoShaders.cu:82-99): every triangle is emitted once, its three
    vertices unshared, each clip position (mvp * (x, y, z, 1), or the record's first four floats
    for clip pass-through) plus `offset`, each varying record passed on unchanged.  Triangles land
    in input order, the order the deterministic sort keys (:539-542) give.  Returns (clip, records,
    indices) of the emitted list (indices 0 .. 3n - 1)."""
    v = np.ascontiguousarray(vertices, np.float32)
    idx = np.ascontiguousarray(indices, np.int32)[: len(indices) // 3 * 3]
    clip = v[:, :4].copy() if passthrough else vertex_transform(v, mvp)
    out = clip[idx] + np.asarray(offset, np.float32).reshape(1, 4)
    return out.astype(np.float32), v[idx].copy(), np.arange(idx.size, dtype=np.int32)
