"""Seeded synthetic scenes for the tile rasterizer workloads.

Every scene is a plain set of arrays -- an interleaved FP32 vertex buffer,
an int32 index buffer, a 4x4 view-projection matrix and the shading
parameters -- so the same bytes feed the MI355X pipeline and the CPU
checkers.  Camera matrices follow the reference's left-handed conventions
(include/ifrit/core/math/LinalgOps.h:83-149: LookAt rows s, u, f and a
perspective with z in [0, 1] and w = view-space z); they are computed in
float64 and rounded once to float32.

Vertex layout: position xyz, normal xyz, texture coordinate uv (32 bytes),
or position only (12 bytes) for the depth/flat-colour stress scene.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

STRIDE_FULL = 8
STRIDE_POS = 3

PROGRAM_FLAT = 0
PROGRAM_BLINN_PHONG = 1
PROGRAM_DEPTH_ONLY = 2
PROGRAM_GGX_LAYERED = 3
PROGRAM_FLAT_ALPHA = 6
PROGRAM_NORMAL_LAMBERT = 7
PROGRAM_CUBE_MAP = 8
SAMPLER_BILINEAR = 0
SAMPLER_TRILINEAR = 1
SAMPLER_ANISO = 2


def perspective(fovy: float, aspect: float, znear: float, zfar: float) -> np.ndarray:
    """LinalgOps.h:124-149 (left-handed, z in [0, 1], w = z_view)."""
    top = znear * math.tan(fovy / 2.0)
    right = top * aspect
    m = np.zeros((4, 4), dtype=np.float64)
    m[0, 0] = 2 * znear / (2 * right)
    m[1, 1] = 2 * znear / (2 * top)
    m[2, 2] = zfar / (zfar - znear)
    m[2, 3] = -(zfar * znear) / (zfar - znear)
    m[3, 2] = 1.0
    return m


def look_at(eye, center, up) -> np.ndarray:
    """LinalgOps.h:83-122: rows s, u, f then translation by -eye."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(center, np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    r = np.eye(4)
    r[0, :3], r[1, :3], r[2, :3] = s, u, f
    t = np.eye(4)
    t[:3, 3] = -eye
    return r @ t


def orthographic(half_w: float, half_h: float, znear: float, zfar: float) -> np.ndarray:
    """Orthographic projection in the conventions of `perspective`: left-handed,
    x, y in [-1, 1] over [-half_w, half_w] x [-half_h, half_h], z in [0, 1] over
    [znear, zfar], w = 1."""
    m = np.zeros((4, 4), dtype=np.float64)
    m[0, 0] = 1.0 / half_w
    m[1, 1] = 1.0 / half_h
    m[2, 2] = 1.0 / (zfar - znear)
    m[2, 3] = -znear / (zfar - znear)
    m[3, 3] = 1.0
    return m


@dataclasses.dataclass
class Texture:
    width: int
    height: int
    levels: int
    texels: np.ndarray | None = None      # uint8, all levels packed (layers one after another per level),
                                          # or None when generated on the device
    fbm_seed: int | None = None           # device-generated fBm texture seed
    layers: int = 1                       # 6 for a cube map (+x -x +y -y +z -z)


@dataclasses.dataclass
class Scene:
    name: str
    width: int
    height: int
    samples: int
    vertices: np.ndarray          # float32 [V, stride]
    indices: np.ndarray           # int32 [3T]
    mvp: np.ndarray               # float32 [4, 4] row-major, clip = mvp @ (p, 1)
    program: int = PROGRAM_FLAT
    sampler: int = SAMPLER_TRILINEAR
    max_aniso: int = 1
    tris_per_material: int = 0
    camera: tuple = (0.0, 0.0, 0.0)
    ambient: float = 0.1
    kd: float = 0.8
    ks: float = 0.3
    shininess: float = 32.0
    lights: list = dataclasses.field(default_factory=list)   # [(pos4, color4)]
    textures: list = dataclasses.field(default_factory=list)
    sampler_state: dict | None = None   # Renderer.create_sampler arguments (None: the program's own sampler)
    sampler_id: int = 0                 # set when the scene is loaded (0: none)
    lod: float = 0.0                    # cube-map program: level of detail
    depth_write: tuple | None = None    # (scale, bias): the program writes depth * scale + bias

    @property
    def stride(self) -> int:
        return int(self.vertices.shape[1])

    @property
    def num_triangles(self) -> int:
        return int(self.indices.size // 3)


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------

def uv_sphere(seg_u: int = 64, seg_v: int = 32, radius: float = 1.0):
    """(seg_u+1) x (seg_v+1) vertices with seam duplicates; 2 triangles per
    cell including the degenerate halves at the poles."""
    u = np.arange(seg_u + 1) / seg_u
    v = np.arange(seg_v + 1) / seg_v
    uu, vv = np.meshgrid(u, v)
    theta, phi = uu * 2 * np.pi, vv * np.pi
    n = np.stack([np.sin(phi) * np.cos(theta), np.cos(phi), np.sin(phi) * np.sin(theta)], -1).reshape(-1, 3)
    verts = np.concatenate([n * radius, n, np.stack([uu, vv], -1).reshape(-1, 2)], 1)
    idx = []
    w = seg_u + 1
    for i in range(seg_v):
        for j in range(seg_u):
            a, b, c, d = i * w + j, (i + 1) * w + j, (i + 1) * w + j + 1, i * w + j + 1
            idx += [a, b, c, a, c, d]
    return verts.astype(np.float32), np.asarray(idx, np.int32)


def icosphere(level: int):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.asarray(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    p = np.asarray(verts)
    uv = np.stack([np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi) + 0.5, np.arccos(np.clip(p[:, 1], -1, 1)) / np.pi], -1)
    return p, uv, np.asarray(faces, np.int32)


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ---------------------------------------------------------------------------
# textures (integer-exact so host and device generators agree bit for bit)
# ---------------------------------------------------------------------------

def _hash2(ix: np.ndarray, iy: np.ndarray, seed: int) -> np.ndarray:
    h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (iy.astype(np.uint32) * np.uint32(0xD8163841)) \
        ^ np.uint32(seed * 0x9E3779B1 & 0xFFFFFFFF)
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7FEB352D)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846CA68B)
    h ^= h >> np.uint32(16)
    return h


def fbm_texture(size: int, seed: int, octaves: int = 6, channels: int = 3) -> np.ndarray:
    """Tileable value-noise fBm in 16-bit fixed point; RGBA8 level 0.

    Octave o has cells of size >> (o + 2) texels; corner values are 16-bit
    hashes; bilinear weights are exact 16-bit fractions; octave o contributes
    value >> (o + 1).  Channel c uses seed * 4 + c; alpha is 255."""
    y, x = np.meshgrid(np.arange(size, dtype=np.int64), np.arange(size, dtype=np.int64), indexing="ij")
    out = np.zeros((size, size, 4), np.uint8)
    for ch in range(channels):
        acc = np.zeros((size, size), np.int64)
        for o in range(octaves):
            cell = max(1, size >> (o + 2))
            shift = cell.bit_length() - 1
            ncell = size // cell
            cx, cy = x >> shift, y >> shift
            wx = (x & (cell - 1)) << (16 - shift)
            wy = (y & (cell - 1)) << (16 - shift)
            s = (seed * 4 + ch) * 16 + o
            v00 = (_hash2(cx % ncell, cy % ncell, s) & 0xFFFF).astype(np.int64)
            v10 = (_hash2((cx + 1) % ncell, cy % ncell, s) & 0xFFFF).astype(np.int64)
            v01 = (_hash2(cx % ncell, (cy + 1) % ncell, s) & 0xFFFF).astype(np.int64)
            v11 = (_hash2((cx + 1) % ncell, (cy + 1) % ncell, s) & 0xFFFF).astype(np.int64)
            top = v00 * (65536 - wx) + v10 * wx
            bot = v01 * (65536 - wx) + v11 * wx
            n = (top * (65536 - wy) + bot * wy) >> 32
            acc += n >> (o + 1)
        out[..., ch] = np.minimum(acc >> 8, 255).astype(np.uint8)
    out[..., 3] = 255
    return out


def normal_map(height: np.ndarray, strength: float = 4.0) -> np.ndarray:
    """Tangent-space normal map (RGBA8, n encoded as 0.5 n + 0.5) from a
    uint8 height field with wrap-around central differences."""
    h = height.astype(np.float64) / 255.0
    dx = (np.roll(h, -1, axis=1) - np.roll(h, 1, axis=1)) * strength
    dy = (np.roll(h, -1, axis=0) - np.roll(h, 1, axis=0)) * strength
    n = np.stack([-dx, -dy, np.ones_like(h)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    out = np.empty(h.shape + (4,), np.uint8)
    out[..., :3] = np.floor((n * 0.5 + 0.5) * 255.0 + 0.5).astype(np.uint8)
    out[..., 3] = 255
    return out


def checker_texture(size: int = 512, checks: int = 8) -> np.ndarray:
    y, x = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
    c = ((x // (size // checks)) + (y // (size // checks))) & 1
    out = np.empty((size, size, 4), np.uint8)
    out[..., 0] = np.where(c, 230, 40)
    out[..., 1] = np.where(c, 200, 60)
    out[..., 2] = np.where(c, 170, 90)
    out[..., 3] = 255
    return out


def num_levels(w: int, h: int) -> int:
    return int(math.floor(math.log2(max(w, h)))) + 1


def build_mips(level0: np.ndarray) -> np.ndarray:
    """Box filter, (a + b + c + d + 2) >> 2 per channel; levels packed."""
    h, w = level0.shape[:2]
    out = [level0.reshape(-1)]
    cur = level0.astype(np.int32)
    for _ in range(1, num_levels(w, h)):
        ph, pw = cur.shape[:2]
        nh, nw = max(1, ph >> 1), max(1, pw >> 1)
        ys0 = np.minimum(np.arange(nh) * 2, ph - 1)
        ys1 = np.minimum(np.arange(nh) * 2 + 1, ph - 1)
        xs0 = np.minimum(np.arange(nw) * 2, pw - 1)
        xs1 = np.minimum(np.arange(nw) * 2 + 1, pw - 1)
        s = cur[ys0][:, xs0] + cur[ys0][:, xs1] + cur[ys1][:, xs0] + cur[ys1][:, xs1]
        cur = (s + 2) >> 2
        out.append(cur.astype(np.uint8).reshape(-1))
    return np.concatenate(out)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

def _mvp(eye, center, fov_deg, w, h, near=0.1, far=100.0):
    return (perspective(math.radians(fov_deg), w / h, near, far) @ look_at(eye, center, (0, 1, 0))).astype(np.float32)


def config1(width: int = 640, height: int = 480) -> Scene:
    """UV sphere 64x32 (4096 tris) + 2-tri ground quad; 512^2 checker with
    mips, trilinear; Blinn-Phong with one directional light; 1 spp."""
    sv, si = uv_sphere(64, 32, 1.0)
    g = np.array([[-4, -1, -4, 0, 1, 0, 0, 0], [4, -1, -4, 0, 1, 0, 4, 0], [4, -1, 4, 0, 1, 0, 4, 4],
                  [-4, -1, 4, 0, 1, 0, 0, 4]], np.float32)
    gi = np.array([0, 2, 1, 0, 3, 2], np.int32) + sv.shape[0]
    verts = np.concatenate([sv, g])
    idx = np.concatenate([si, gi])
    eye = (0.6, 1.2, -3.2)
    tex = checker_texture(512, 8)
    return Scene("config1", width, height, 1, verts, idx, _mvp(eye, (0, 0, 0), 60.0, width, height),
                 program=PROGRAM_BLINN_PHONG, sampler=SAMPLER_TRILINEAR, max_aniso=1, camera=eye,
                 lights=[((0.4, 0.8, -0.45, 0.0), (1.0, 1.0, 1.0, 0.0))],
                 textures=[Texture(512, 512, num_levels(512, 512), build_mips(tex))])


def config2(width: int = 1920, height: int = 1080, samples: int = 8, spheres: int = 256, level: int = 4,
            tex_size: int = 2048, seed: int = 20241202, host_textures: bool = False) -> Scene:
    """Seeded icospheres (level 4: 2562 verts / 5120 tris each) at U[-20,20]^3,
    log-uniform scale [0.2, 3], uniform rotation; per-sphere fBm texture with
    mips, 16x anisotropic filtering, Blinn-Phong with 4 point lights."""
    rng = np.random.default_rng(seed)
    p, uv, f = icosphere(level)
    nv, nt = p.shape[0], f.shape[0]
    verts = np.empty((spheres * nv, 8), np.float32)
    idx = np.empty((spheres * nt * 3,), np.int32)
    for s in range(spheres):
        pos = rng.uniform(-20, 20, 3)
        scale = math.exp(rng.uniform(math.log(0.2), math.log(3.0)))
        rot = random_rotation(rng)
        wp = (p * scale) @ rot.T + pos
        wn = p @ rot.T
        verts[s * nv:(s + 1) * nv, 0:3] = wp
        verts[s * nv:(s + 1) * nv, 3:6] = wn
        verts[s * nv:(s + 1) * nv, 6:8] = uv
        idx[s * nt * 3:(s + 1) * nt * 3] = (f + s * nv).reshape(-1)
    lights = []
    for _ in range(4):
        lp = rng.uniform(-25, 25, 3)
        lc = rng.uniform(0.4, 1.0, 3)
        lights.append(((float(lp[0]), float(lp[1]), float(lp[2]), 1.0), (float(lc[0]), float(lc[1]), float(lc[2]), 0.002)))
    eye = (0.0, 0.0, -38.0)
    levels = num_levels(tex_size, tex_size)
    textures = []
    for s in range(spheres):
        tseed = int(rng.integers(0, 2 ** 31))
        t = Texture(tex_size, tex_size, levels, None, tseed)
        if host_textures:
            t.texels = build_mips(fbm_texture(tex_size, tseed))
        textures.append(t)
    return Scene("config2", width, height, samples, verts, idx, _mvp(eye, (0, 0, 0), 60.0, width, height),
                 program=PROGRAM_BLINN_PHONG, sampler=SAMPLER_ANISO, max_aniso=16, tris_per_material=nt,
                 camera=eye, lights=lights, textures=textures, ambient=0.08, kd=0.9, ks=0.25)


def config4_views(seed: int = 20241204, cascades: int = 4, cascade_size: int = 4096, cube_faces: int = 6,
                  cube_size: int = 1024, cameras: int = 54, max_size=(1920, 1080), min_size: int = 256):
    """The views of config 4 over the config-2 scene (sphere centres in [-20, 20]^3):
    - shadow cascades, depth only: orthographic views along a seeded light direction
      towards the origin, extents 6, 12, 24, 48 (each twice the previous), square;
    - cube faces, depth only: 90-degree views from the origin along +x, -x, +y, -y,
      +z, -z (the origin may lie inside a sphere: triangles cross w = 0);
    - seeded shaded cameras: eye on a shell of radius U[25, 45], look-at point in
      [-6, 6]^3, vertical field of view U[40, 75] degrees, width U[min, max_w],
      height U[min, max_h].
    Returns dicts for Renderer.set_views."""
    rng = np.random.default_rng(seed)
    views = []
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.9 else (1.0, 0.0, 0.0)
    eye = -60.0 * d
    for c in range(cascades):
        ext = 6.0 * 2.0 ** c
        m = orthographic(ext, ext, 1.0, 120.0) @ look_at(eye, (0.0, 0.0, 0.0), up)
        views.append(dict(width=cascade_size, height=cascade_size, mvp=m.astype(np.float32),
                          camera=tuple(float(x) for x in eye), shade=False, kind="cascade"))
    dirs = [((1, 0, 0), (0, 1, 0)), ((-1, 0, 0), (0, 1, 0)), ((0, 1, 0), (0, 0, -1)), ((0, -1, 0), (0, 0, 1)),
            ((0, 0, 1), (0, 1, 0)), ((0, 0, -1), (0, 1, 0))]
    for f in range(cube_faces):
        fwd, upv = dirs[f % 6]
        m = perspective(math.pi / 2, 1.0, 0.1, 100.0) @ look_at((0.0, 0.0, 0.0), fwd, upv)
        views.append(dict(width=cube_size, height=cube_size, mvp=m.astype(np.float32), camera=(0.0, 0.0, 0.0),
                          shade=False, kind="cube"))
    for _ in range(cameras):
        w = int(rng.integers(min_size, max_size[0] + 1))
        h = int(rng.integers(min_size, max_size[1] + 1))
        z = rng.uniform(-1, 1)
        phi = rng.uniform(0, 2 * math.pi)
        r = rng.uniform(25, 45)
        rr = math.sqrt(max(0.0, 1 - z * z))
        e = (r * rr * math.cos(phi), r * z, r * rr * math.sin(phi))
        at = tuple(rng.uniform(-6, 6, 3))
        fov = math.radians(rng.uniform(40, 75))
        m = perspective(fov, w / h, 0.1, 100.0) @ look_at(e, at, (0, 1, 0))
        views.append(dict(width=w, height=h, mvp=m.astype(np.float32), camera=tuple(float(x) for x in e),
                          shade=True, kind="camera"))
    return views


def sphere_tangents(p: np.ndarray) -> np.ndarray:
    """Unit tangents d p / d u of the icosphere parameterisation
    (u = atan2(z, x) / 2 pi + 0.5): direction (-z, 0, x), (1, 0, 0) at the poles."""
    t = np.stack([-p[:, 2], np.zeros(len(p)), p[:, 0]], -1)
    n = np.linalg.norm(t, axis=1, keepdims=True)
    return np.where(n > 1e-9, t / np.maximum(n, 1e-300), np.array([1.0, 0.0, 0.0]))


def material_layers(size: int, seed: int, host: bool = True):
    """The three 2^k x 2^k layers of the layered-GGX material: fBm albedo,
    a normal map from an fBm height field, fBm roughness (red channel)."""
    rng = np.random.default_rng(seed)
    sa, sh, sr = (int(x) for x in rng.integers(0, 2 ** 31, 3))
    levels = num_levels(size, size)
    albedo = Texture(size, size, levels, None, sa)
    rough = Texture(size, size, levels, None, sr)
    if host:
        albedo.texels = build_mips(fbm_texture(size, sa))
        rough.texels = build_mips(fbm_texture(size, sr, channels=1))
    normal = Texture(size, size, levels, build_mips(normal_map(fbm_texture(size, sh, channels=1)[..., 0])))
    return [albedo, normal, rough]


def config5(width: int = 1920, height: int = 1080, samples: int = 1, spheres: int = 256, level: int = 4,
            tex_size: int = 2048, seed: int = 20241202, host_textures: bool = False) -> Scene:
    """Config-2 geometry with per-vertex tangents (stride 11: pos, normal, uv,
    tangent) and one layered-GGX material shared by every sphere (albedo,
    tangent-space normal, roughness layers at tex_size^2 with mips, 16x
    anisotropic filtering), one directional light plus ambient; 1 sample per
    pixel for the shade pass."""
    base = config2(width, height, samples, spheres, level, tex_size, seed)
    rng = np.random.default_rng(seed)
    p, uv, f = icosphere(level)
    tl = sphere_tangents(p)
    nv = p.shape[0]
    verts = np.empty((spheres * nv, 11), np.float32)
    verts[:, :8] = base.vertices
    for s in range(spheres):
        # the same draws as config2, in the same order
        rng.uniform(-20, 20, 3)
        rng.uniform(math.log(0.2), math.log(3.0))
        rot = random_rotation(rng)
        verts[s * nv:(s + 1) * nv, 8:11] = tl @ rot.T
    eye = base.camera
    return Scene("config5", width, height, samples, verts, base.indices, base.mvp, program=PROGRAM_GGX_LAYERED,
                 sampler=SAMPLER_ANISO, max_aniso=16, camera=eye, ambient=0.1, kd=1.0, ks=1.0,
                 lights=[((0.4, 0.8, -0.45, 0.0), (3.6, 3.45, 3.3, 0.0))],
                 textures=material_layers(tex_size, seed + 5, host=host_textures))


def config3(width: int = 1920, height: int = 1080, samples: int = 4, gx: int = 2048, gy: int = 1024,
            big: int = 16, seed: int = 7) -> Scene:
    """2048x1024-cell heightfield displaced by 6-octave fBm (amplitude 0.05 of
    the extent) plus 16 large triangles covering 10-60 % of the screen;
    position-only vertices, flat colour."""
    rng = np.random.default_rng(seed)
    extent = 20.0
    xs = np.linspace(-extent, extent, gx + 1)
    zs = np.linspace(-extent / 2, extent / 2, gy + 1)
    X, Z = np.meshgrid(xs, zs)
    noise = fbm_texture(1024, seed)[..., 0].astype(np.float64) / 255.0
    iy = (np.arange(gy + 1) * 1023 // gy)
    ix = (np.arange(gx + 1) * 1023 // gx)
    Y = (noise[iy][:, ix] - 0.5) * 0.05 * extent * 2
    verts = np.stack([X, Y, Z], -1).reshape(-1, 3)
    w = gx + 1
    jj, ii = np.meshgrid(np.arange(gx), np.arange(gy))
    a = (ii * w + jj).reshape(-1)
    quads = np.stack([a, a + w, a + w + 1, a, a + w + 1, a + 1], 1).reshape(-1)
    eye = np.array([0.0, 9.0, -17.0])
    mvp64 = perspective(math.radians(60.0), width / height, 0.1, 100.0) @ look_at(eye, (0, -1.0, 0.5), (0, 1, 0))
    inv = np.linalg.inv(mvp64)
    bverts, bidx = [], []
    base = verts.shape[0]
    for k in range(big):
        frac = rng.uniform(0.10, 0.60)
        # a random triangle in NDC with area frac * 4 (the NDC square has area 4)
        while True:
            c = rng.uniform(-0.6, 0.6, 2)
            ang = rng.uniform(0, 2 * np.pi)
            r = math.sqrt(frac * 4 / (1.299038105676658 * rng.uniform(0.8, 1.2)))
            pts = np.array([[c[0] + r * math.cos(ang + i * 2.0944), c[1] + r * math.sin(ang + i * 2.0944)] for i in range(3)])
            area = 0.5 * abs((pts[1, 0] - pts[0, 0]) * (pts[2, 1] - pts[0, 1]) - (pts[2, 0] - pts[0, 0]) * (pts[1, 1] - pts[0, 1]))
            if 0.10 * 4 <= area <= 0.60 * 4:
                break
        zn = rng.uniform(0.90, 0.995)
        for q in pts:
            ndc = np.array([q[0], q[1], zn, 1.0])
            wp = inv @ ndc
            bverts.append(wp[:3] / wp[3])
        bidx += [base + 3 * k, base + 3 * k + 1, base + 3 * k + 2]
    verts = np.concatenate([verts, np.asarray(bverts)]).astype(np.float32)
    idx = np.concatenate([quads, np.asarray(bidx)]).astype(np.int32)
    return Scene("config3", width, height, samples, verts, idx, mvp64.astype(np.float32), program=PROGRAM_FLAT,
                 camera=tuple(eye))


def normal_lambert(scene: Scene) -> Scene:
    """`scene` shaded like the reference CPU path's timing workload: the
    normal-Lambert program (no textures) with light direction (0.3, 0.8, -0.52)
    used as given, ambient 0.1, diffuse weights (0.8, 0.6, 0.4) and the
    highlight exponent 32."""
    return dataclasses.replace(scene, program=PROGRAM_NORMAL_LAMBERT, ambient=0.1, shininess=32.0, ks=1.0, kd=1.0,
                               lights=[((0.3, 0.8, -0.52, 0.0), (0.8, 0.6, 0.4, 0.0))], textures=[])


def reference_workload(width: int = 1920, height: int = 1080, spheres: int = 256, level: int = 4) -> Scene:
    """The config-2 geometry at one sample per pixel with the normal-Lambert
    shading: the frame the reference CPU rasterizer is timed on."""
    return normal_lambert(config2(width, height, 1, spheres=spheres, level=level))


def layered_mips(faces) -> np.ndarray:
    """The packed chain of a layered texture: each layer's box-filtered mips
    (build_mips), level by level with the layers of a level one after another
    (the library's layout; a cube map's layers are +x -x +y -y +z -z)."""
    chains = [build_mips(f) for f in faces]
    h, w = faces[0].shape[:2]
    out, off = [], 0
    for l in range(num_levels(w, h)):
        n = max(1, w >> l) * max(1, h >> l) * 4
        out.extend(c[off:off + n] for c in chains)
        off += n
    return np.concatenate(out)


def cube_texture(size: int = 64, seed: int = 11) -> Texture:
    """A seeded cube map: per face an fBm pattern tinted by a face-specific colour, with mips."""
    tints = [(255, 80, 80), (80, 255, 80), (80, 80, 255), (255, 255, 80), (255, 80, 255), (80, 255, 255)]
    faces = []
    for k, tint in enumerate(tints):
        base = fbm_texture(size, seed + 17 * k).astype(np.int32)
        f = (base * np.array(tint + (255,), np.int32)) // 255
        faces.append(f.astype(np.uint8))
    return Texture(size, size, num_levels(size, size), layered_mips(faces), None, layers=6)


def sphere_scene(width: int = 256, height: int = 192, samples: int = 1, seg: int = 48) -> Scene:
    """One UV sphere in front of the camera (normals cover every direction), for the cube-map program."""
    sv, si = uv_sphere(seg, seg // 2, 1.0)
    eye = (0.0, 0.3, -3.0)
    return Scene("sphere", width, height, samples, sv, si, _mvp(eye, (0, 0, 0), 50.0, width, height),
                 program=PROGRAM_CUBE_MAP, camera=eye)


def build_meshlets(vertices: np.ndarray, indices: np.ndarray, max_vertices: int = 64, max_triangles: int = 126):
    """Greedy meshlets of a triangle list in its order (a new meshlet starts when the next triangle
    would exceed either bound), as the buffers of the reference's meshlet demo: the meshlets'
    vertex records one after another (a vertex used by several meshlets is repeated), their
    meshlet-local indices, and the int32 vertex and index offsets (one more than meshlets).
    Returns (records, local_indices, vertex_offsets, index_offsets, global_indices) where
    global_indices indexes records and lists the same triangles in the same order."""
    tris = np.asarray(indices, np.int64).reshape(-1, 3)
    recs, loc, glob, vo, io = [], [], [], [0], [0]
    cur = {}
    ntri = 0

    def close():
        nonlocal cur, ntri
        vo.append(vo[-1] + len(cur))
        io.append(len(loc))
        cur = {}
        ntri = 0

    for t in tris:
        new = sum(1 for v in t if int(v) not in cur)
        if cur and (len(cur) + new > max_vertices or ntri + 1 > max_triangles):
            close()
        for v in t:
            v = int(v)
            if v not in cur:
                cur[v] = len(cur)
                recs.append(vertices[v])
            loc.append(cur[v])
            glob.append(vo[-1] + cur[v])
        ntri += 1
    if cur:
        close()
    return (np.asarray(recs, np.float32), np.asarray(loc, np.int32), np.asarray(vo, np.int32),
            np.asarray(io, np.int32), np.asarray(glob, np.int32))
