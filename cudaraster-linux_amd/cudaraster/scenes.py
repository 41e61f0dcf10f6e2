"""Seeded synthetic workloads for the rasterization path.

Every scene is a clip-space vertex buffer (records of ``stride`` bytes, all
float4: clipPos first, varyings after, as ShadedVertexBase / GouraudVertex
require -- cudaraster/cuda/PixelPipe.hpp:41-47, 130-133), an ``int3`` index
buffer (cudaraster/cuda/TriangleSetup.inl:225, 246) and the pipe state it is
drawn with.  The same buffers feed the GPU path and the CPU oracle.

Workloads:
  cube()          12-triangle cube of test/App.cpp:208-239 seen through the
                  demo camera (eye (0,2,4) -> origin, 60 deg, 0.1..100;
                  test/App.cpp:83-84, 97-102; test/engine/Camera.cpp:57-110)
  random_tris()   independent triangles: centres uniform over the viewport,
                  screen area log-uniform, uniform orientation and winding,
                  view depth uniform per triangle (+-10 % per vertex)
  skewed()        many sub-pixel triangles plus a set of near-full-screen
                  triangles at random depths, interleaved by seed
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# RenderModeFlag_* (cudaraster/cuda/PixelPipe.hpp:30-35)
FLAG_DEPTH = 1
FLAG_LERP = 2
FLAG_QUADS = 4

BLEND_REPLACE, BLEND_SRCOVER, BLEND_ADDITIVE, BLEND_DEPTHONLY = 0, 1, 2, 3
SHADER_GOURAUD, SHADER_CONSTANT, SHADER_TEXPHONG = 0, 1, 2

# TextureSpec and Material (test/shader/Shaders.hpp:33-51).
TEXSPEC_DTYPE = np.dtype([("size", "<f4", (2,)), ("pad", "<u4", (2,)), ("mipLevels", "<f4", (16, 4))])
MATERIAL_DTYPE = np.dtype([("diffuseColor", "<f4", (4,)), ("specularColor", "<f4", (3,)), ("glossiness", "<f4"),
                           ("diffuseTexture", TEXSPEC_DTYPE), ("alphaTexture", TEXSPEC_DTYPE)])
assert TEXSPEC_DTYPE.itemsize == 272 and MATERIAL_DTYPE.itemsize == 576

ZNEAR, ZFAR = 0.1, 100.0


@dataclasses.dataclass
class Scene:
    name: str
    width: int
    height: int
    samples_log2: int
    flags: int
    blend: int
    shader: int
    vertices: np.ndarray          # float32 [V, stride/4]
    indices: np.ndarray           # int32 [T, 3]
    clear_color: tuple = (0.2, 0.4, 0.8, 1.0)   # test/SceneCR.cpp:287
    clear_depth: float = 1.0
    deferred_clear: bool = True
    const_color: int = 0xFF0000FF
    # texPhong shader constants (None for the other shaders).
    materials: np.ndarray = None        # MATERIAL_DTYPE [M]
    tri_material: np.ndarray = None     # int32 [T]
    atlas: np.ndarray = None            # float32 [H, W, 4]

    @property
    def stride(self) -> int:
        return int(self.vertices.shape[1]) * 4

    @property
    def num_tris(self) -> int:
        return int(self.indices.shape[0])

    @property
    def num_samples(self) -> int:
        return 1 << self.samples_log2

    @property
    def surface_shape(self):
        """(rows, texels per row) of a surface: ceil8(H) x ceil8(W)*N."""
        wr = (self.width + 7) & ~7
        hr = (self.height + 7) & ~7
        return hr, wr * self.num_samples

    def input_bytes(self) -> int:
        return int(self.vertices.nbytes + self.indices.nbytes)


# ---------------------------------------------------------------------------
# Camera (glm 0.9.3 conventions: perspective() takes fovy in degrees,
# thirdparty/glm/gtc/matrix_transform.inl:223-244)
# ---------------------------------------------------------------------------

def perspective(fovy_deg, aspect, znear, zfar):
    rng = math.tan(math.radians(fovy_deg / 2.0)) * znear
    left, right = -rng * aspect, rng * aspect
    bottom, top = -rng, rng
    m = np.zeros((4, 4), dtype=np.float64)
    m[0, 0] = 2.0 * znear / (right - left)
    m[1, 1] = 2.0 * znear / (top - bottom)
    m[2, 2] = -(zfar + znear) / (zfar - znear)
    m[3, 2] = -1.0
    m[2, 3] = -(2.0 * zfar * znear) / (zfar - znear)
    return m


def look_at(eye, center, up):
    eye, center, up = (np.asarray(v, dtype=np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def demo_camera(width, height):
    """View-projection of the demo (Camera::setViewParams / setProjectionParams)."""
    eye, target = np.array([0.0, 2.0, 4.0]), np.array([0.0, 0.0, 0.0])
    d = target - eye
    d /= np.linalg.norm(d)
    up = np.array([0.0, 1.0, 0.0])
    left = np.cross(up, d)
    left /= np.linalg.norm(left)
    up = np.cross(d, left)
    up /= np.linalg.norm(up)
    return perspective(60.0, width / height, ZNEAR, ZFAR) @ look_at(eye, target, up)


CUBE_POSITIONS = np.array([
    [-1, -1, -1], [-1, -1, +1], [-1, +1, -1], [-1, +1, +1],
    [+1, -1, -1], [+1, -1, +1], [+1, +1, -1], [+1, +1, +1]], dtype=np.float64)

CUBE_INDICES = np.array([
    [7, 3, 1], [7, 1, 5], [7, 5, 6], [6, 5, 4], [6, 4, 2], [2, 4, 0],
    [2, 0, 3], [3, 0, 1], [3, 7, 6], [3, 6, 2], [5, 1, 0], [5, 0, 4]], dtype=np.int32)


def cube(width=1024, height=768, samples_log2=0, seed=0, mvp=None) -> Scene:
    """The demo cube: 8 vertices, 12 triangles, per-vertex float3 colours
    (seeded, alpha padded to 1), Gouraud, Replace, Depth|Lerp."""
    rng = np.random.default_rng(seed)
    m = demo_camera(width, height) if mvp is None else np.asarray(mvp, dtype=np.float64)
    pos = np.concatenate([CUBE_POSITIONS, np.ones((8, 1))], axis=1)
    clip = (pos @ m.T).astype(np.float32)
    col = np.concatenate([rng.random((8, 3)), np.ones((8, 1))], axis=1).astype(np.float32)
    verts = np.concatenate([clip, col], axis=1).astype(np.float32)
    return Scene("cube", width, height, samples_log2, FLAG_DEPTH | FLAG_LERP,
                 BLEND_REPLACE, SHADER_GOURAUD, np.ascontiguousarray(verts), CUBE_INDICES.copy())


def _proj_z(w):
    """Clip z for eye depth w (= -z_eye) under the demo projection."""
    a = (ZFAR + ZNEAR) / (ZFAR - ZNEAR)
    b = 2.0 * ZFAR * ZNEAR / (ZFAR - ZNEAR)
    return a * w - b


def _triangle_shapes(rng, n, area):
    """Random triangle shapes of the given screen areas (px^2), centred at 0,
    uniform orientation and random winding."""
    ang0 = rng.random(n) * 2.0 * math.pi
    jit = (rng.random((n, 3)) - 0.5) * (math.pi / 3.0)
    ang = ang0[:, None] + np.array([0.0, 2.0 * math.pi / 3.0, 4.0 * math.pi / 3.0])[None, :] + jit
    rad = 0.6 + 0.4 * rng.random((n, 3))
    x = rad * np.cos(ang)
    y = rad * np.sin(ang)
    flip = rng.random(n) < 0.5
    x[flip] = x[flip][:, ::-1]
    y[flip] = y[flip][:, ::-1]
    a0 = 0.5 * np.abs((x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0]))
    scale = np.sqrt(area / np.maximum(a0, 1e-12))
    x = x - x.mean(axis=1, keepdims=True)
    y = y - y.mean(axis=1, keepdims=True)
    return x * scale[:, None], y * scale[:, None]


def _to_clip(px, py, w, width, height):
    """Pixel positions (y up, origin bottom-left) at eye depth w -> clip."""
    nx = px / width * 2.0 - 1.0
    ny = py / height * 2.0 - 1.0
    return np.stack([nx * w, ny * w, _proj_z(w), w], axis=-1)


def random_tris(num_tris, width=2048, height=2048, samples_log2=2, seed=1,
                area_range=(0.5, 256.0), depth_range=(ZNEAR, ZFAR), alpha_range=None,
                flags=FLAG_DEPTH | FLAG_LERP, blend=BLEND_REPLACE, name="random") -> Scene:
    """Independent triangles (3 vertices each): centre uniform over the
    viewport, area log-uniform in area_range, uniform orientation (about half
    back-facing), eye depth uniform in depth_range per triangle with +-10 %
    per-vertex variation, float4 colours uniform in [0,1] (alpha uniform in
    alpha_range when given)."""
    rng = np.random.default_rng(seed)
    n = int(num_tris)
    cx = rng.random(n) * width
    cy = rng.random(n) * height
    la, lb = math.log(area_range[0]), math.log(area_range[1])
    area = np.exp(la + (lb - la) * rng.random(n))
    sx, sy = _triangle_shapes(rng, n, area)
    wt = depth_range[0] + (depth_range[1] - depth_range[0]) * rng.random(n)
    wv = np.clip(wt[:, None] * (0.9 + 0.2 * rng.random((n, 3))), depth_range[0], depth_range[1])
    clip = _to_clip(cx[:, None] + sx, cy[:, None] + sy, wv, width, height)   # [n,3,4]
    col = rng.random((n, 3, 4))
    if alpha_range is not None:
        col[..., 3] = alpha_range[0] + (alpha_range[1] - alpha_range[0]) * rng.random((n, 3))
    verts = np.concatenate([clip, col], axis=-1).reshape(n * 3, 8).astype(np.float32)
    idx = np.arange(n * 3, dtype=np.int32).reshape(n, 3)
    return Scene(name, width, height, samples_log2, flags, blend, SHADER_GOURAUD,
                 np.ascontiguousarray(verts), idx)


def skewed(num_small=4_000_000, num_big=1024, width=2048, height=2048, samples_log2=3,
           seed=2, area_range=(0.05, 1.0), alpha_range=(0.2, 0.8)) -> Scene:
    """Sub-pixel triangles uniform over the screen plus near-full-screen
    triangles at random depths, interleaved at seeded positions; SrcOver."""
    rng = np.random.default_rng(seed)
    small = random_tris(num_small, width, height, samples_log2, seed=seed + 1000,
                        area_range=area_range, alpha_range=alpha_range)
    nb = int(num_big)
    # Big triangles: one vertex near each of three screen regions, spanning
    # roughly the whole viewport (some exceed the guard band and are clipped).
    u = rng.random((nb, 3, 2))
    base = np.array([[-0.15, -0.15], [1.15, -0.10], [0.5, 1.2]])
    jitter = (u - 0.5) * 0.5
    pts = (base[None] + jitter) * np.array([width, height])
    # Force counter-clockwise winding so every big triangle is front-facing.
    px, py = pts[..., 0], pts[..., 1]
    area2 = (px[:, 1] - px[:, 0]) * (py[:, 2] - py[:, 0]) - (px[:, 2] - px[:, 0]) * (py[:, 1] - py[:, 0])
    cw = area2 < 0
    px[cw] = px[cw][:, ::-1]
    py[cw] = py[cw][:, ::-1]
    wt = ZNEAR + (ZFAR - ZNEAR) * rng.random(nb)
    wv = np.clip(wt[:, None] * (0.9 + 0.2 * rng.random((nb, 3))), ZNEAR, ZFAR)
    clip = _to_clip(px, py, wv, width, height)
    col = rng.random((nb, 3, 4))
    col[..., 3] = alpha_range[0] + (alpha_range[1] - alpha_range[0]) * rng.random((nb, 3))
    big = np.concatenate([clip, col], axis=-1).reshape(nb * 3, 8).astype(np.float32)

    total = small.num_tris + nb
    is_big = np.zeros(total, dtype=bool)
    is_big[rng.choice(total, size=nb, replace=False)] = True
    verts = np.empty((total * 3, 8), dtype=np.float32)
    vb = np.repeat(is_big, 3)
    verts[vb] = big
    verts[~vb] = small.vertices
    idx = np.arange(total * 3, dtype=np.int32).reshape(total, 3)
    return Scene("skewed", width, height, samples_log2, FLAG_DEPTH | FLAG_LERP,
                 BLEND_SRCOVER, SHADER_GOURAUD, verts, idx)


def clip_mix(num_tris, width=256, height=256, samples_log2=2, seed=3) -> Scene:
    """Clip-space triangles that straddle the frustum: vertices behind the
    near plane, beyond the far plane and far outside the guard band, so the
    setup's clip path (frustum clip + fan) and frustum cull are exercised."""
    rng = np.random.default_rng(seed)
    n = int(num_tris)
    w = 0.05 + 2.0 * rng.random((n, 3))
    xy = (rng.random((n, 3, 2)) * 2.0 - 1.0) * w[..., None] * (1.0 + 6.0 * rng.random((n, 1, 1)))
    zndc = rng.random((n, 3)) * 2.6 - 1.3
    pos = np.concatenate([xy, (zndc * w)[..., None], w[..., None]], axis=-1)
    col = rng.random((n, 3, 4))
    col[..., 3] = 0.2 + 0.6 * rng.random((n, 3))
    verts = np.concatenate([pos, col], axis=-1).reshape(n * 3, 8).astype(np.float32)
    idx = np.arange(n * 3, dtype=np.int32).reshape(n, 3)
    return Scene("clip_mix", width, height, samples_log2, FLAG_DEPTH | FLAG_LERP, BLEND_SRCOVER,
                 SHADER_GOURAUD, verts, idx)


def snap_ties(num_tris=300, width=64, height=64, samples_log2=0, seed=12):
    """Triangles whose every vertex lies exactly half-way between two
    subpixel positions: p = k + 1/2 subpixels with integer k of both signs
    and parities, and w in {0.5, 1, 2} so that x * (1/w) * (W * 8) is exact.
    The snap's rounding rule (cvt.rni: nearest, ties to even,
    TriangleSetup.inl:19-31) alone decides every snapped coordinate.
    Returns (scene, k) with k int64 [T, 3, 2] the integer parts."""
    rng = np.random.default_rng(seed)
    sx, sy = width * 8, height * 8                  # subpixels from the centre to an edge
    ks = np.empty((num_tris, 3, 2), np.int64)
    clip = np.empty((num_tris, 3, 4), np.float64)
    for t in range(num_tris):
        while True:
            a = rng.integers(-sx + 8, sx - 8, size=3)
            b = rng.integers(-sy + 8, sy - 8, size=3)
            area = (a[1] - a[0]) * (b[2] - b[0]) - (b[1] - b[0]) * (a[2] - a[0])
            if abs(area) > 256:
                break
        if area < 0:                                 # counter-clockwise = front-facing
            a[[1, 2]] = a[[2, 1]]
            b[[1, 2]] = b[[2, 1]]
        w = float(rng.choice([0.5, 1.0, 2.0]))
        ks[t, :, 0], ks[t, :, 1] = a, b
        clip[t, :, 0] = (a + 0.5) / sx * w
        clip[t, :, 1] = (b + 0.5) / sy * w
        clip[t, :, 2] = (rng.random(3) - 0.5) * w
        clip[t, :, 3] = w
    col = rng.random((num_tris, 3, 4))
    verts = np.concatenate([clip, col], axis=-1).reshape(num_tris * 3, 8).astype(np.float32)
    assert np.array_equal(verts.reshape(num_tris, 3, 8)[..., :2].astype(np.float64), clip[..., :2])
    idx = np.arange(num_tris * 3, dtype=np.int32).reshape(num_tris, 3)
    return (Scene("snap_ties", width, height, samples_log2, FLAG_DEPTH | FLAG_LERP, BLEND_REPLACE,
                  SHADER_GOURAUD, np.ascontiguousarray(verts), idx), ks)


# Snapped boxes (x0, x1, y0, y1) in subpixels relative to the viewport centre
# at 2048^2 (one NDC unit = 16384 subpixels), each drawn as the triangle
# (x0,y0) (x1,y0) (x0,y1), and whether the fast path takes it: the snapped
# box must lie in [-32768, 32767] with max(x1, y1) - min(x0, y0) <= 2^15 - 1
# (TriangleSetup.inl:285-296).  The extent mixes the axes.
GUARD_BAND_CASES = [
    ((-16384, 16383, -100, 100), True),     # extent 32767
    ((-16384, 16384, -100, 100), False),    # extent 32768
    ((-16384, 0, 0, 16383), True),          # mixed axes: 16383 - (-16384)
    ((-16384, 0, 0, 16384), False),         # mixed axes: 32768, each axis 16384
    ((-32768, -1, -200, -1), True),         # corner at the lower limit
    ((-32769, -2, -200, -2), False),        # one subpixel beyond it
    ((1, 32767, 1, 200), True),             # corner at the upper limit
    ((1, 32768, 1, 200), False),            # one subpixel beyond it
]


def guard_band(width=2048, height=2048, samples_log2=0) -> Scene:
    """One triangle per GUARD_BAND_CASES entry, w = 1 (snapped coordinates
    are the given integers exactly), front-facing, inside the depth range."""
    n = len(GUARD_BAND_CASES)
    clip = np.zeros((n, 3, 4), np.float64)
    for t, ((x0, x1, y0, y1), _) in enumerate(GUARD_BAND_CASES):
        pts = [(x0, y0), (x1, y0), (x0, y1)]
        for k, (x, y) in enumerate(pts):
            clip[t, k] = (x / (width * 8), y / (height * 8), 0.25 + 0.05 * t, 1.0)
    col = np.random.default_rng(5).random((n, 3, 4))
    verts = np.concatenate([clip, col], axis=-1).reshape(n * 3, 8).astype(np.float32)
    idx = np.arange(n * 3, dtype=np.int32).reshape(n, 3)
    return Scene("guard_band", width, height, samples_log2, FLAG_DEPTH | FLAG_LERP, BLEND_REPLACE,
                 SHADER_GOURAUD, np.ascontiguousarray(verts), idx)


def config(i: int, scale: float = 1.0) -> Scene:
    """The plain-method workloads, optionally scaled down (triangle counts and
    resolution) for oracle-sized parity runs."""
    if i == 0:
        return cube()
    if i == 1:
        n = max(1, int(1_000_000 * scale))
        res = max(64, int(round(2048 * math.sqrt(scale) / 8)) * 8) if scale < 1.0 else 2048
        return random_tris(n, res, res, samples_log2=2, seed=1, name="random_1m")
    if i == 2:
        ns = max(1, int(4_000_000 * scale))
        nb = max(1, int(1024 * scale)) if scale < 1.0 else 1024
        res = max(64, int(round(2048 * math.sqrt(scale) / 8)) * 8) if scale < 1.0 else 2048
        return skewed(ns, nb, res, res)
    raise ValueError(f"no plain-method workload {i}")


def multi_view(num_views=64, seed=3, resolutions=(256, 512, 1024, 2048), tri_range=(10_000, 500_000),
               samples_choices=(0, 2), scale=1.0) -> list:
    """Batched multi-view workload: per view a resolution drawn from
    `resolutions` (square), a triangle count log-uniform in `tri_range`, and a
    sample count 2^s with s drawn from `samples_choices`, all from one seed;
    each view's triangles follow the 1M-triangle workload's distribution
    (random_tris defaults).  `scale` shrinks triangle counts and resolutions
    for oracle-sized runs."""
    rng = np.random.default_rng(seed)
    res = rng.choice(np.asarray(resolutions), size=num_views)
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    tris = np.exp(lo + (hi - lo) * rng.random(num_views))
    sl2 = rng.choice(np.asarray(samples_choices), size=num_views)
    seeds = rng.integers(0, 2 ** 31, size=num_views)
    views = []
    for k in range(num_views):
        r = int(res[k])
        n = int(tris[k])
        if scale < 1.0:
            r = max(64, int(round(r * math.sqrt(scale) / 8)) * 8)
            n = max(1, int(n * scale))
        views.append(random_tris(n, r, r, samples_log2=int(sl2[k]), seed=int(seeds[k]), name=f"view{k}"))
    return views


@dataclasses.dataclass
class DiffView:
    """One view of the differentiable workload: clip positions [V, 4],
    attributes [V, C] uniform in [-1, 1], triangles [T, 3], and an L2 target
    image [H, W, C] (row 0 at the bottom)."""
    name: str
    width: int
    height: int
    pos: np.ndarray
    attr: np.ndarray
    tris: np.ndarray
    target: np.ndarray

    def raster_scene(self) -> Scene:
        """The visibility draw: positions only, depth test, 1x."""
        return Scene(self.name, self.width, self.height, 0, FLAG_DEPTH, BLEND_REPLACE, SHADER_CONSTANT,
                     np.ascontiguousarray(self.pos, dtype=np.float32), np.ascontiguousarray(self.tris))


def diff_view(num_tris=200_000, width=512, height=512, channels=8, seed=4, area_range=(0.5, 256.0)) -> DiffView:
    """Random mesh with the 1M-triangle workload's distribution at the given
    resolution, `channels` FP32 attributes per vertex uniform in [-1, 1] and a
    seeded random target image."""
    base = random_tris(num_tris, width, height, samples_log2=0, seed=seed, area_range=area_range)
    rng = np.random.default_rng(seed + 7919)
    V = base.vertices.shape[0]
    attr = rng.uniform(-1.0, 1.0, size=(V, channels)).astype(np.float32)
    target = rng.uniform(-1.0, 1.0, size=(height, width, channels)).astype(np.float32)
    pos = np.ascontiguousarray(base.vertices[:, :4])
    return DiffView(f"diff{seed}", width, height, pos, attr, np.ascontiguousarray(base.indices), target)


def diff_views(num_views=32, num_tris=200_000, width=512, height=512, channels=8, seed=4, scale=1.0) -> list:
    """The differentiable workload: `num_views` independent views."""
    n = max(1, int(num_tris * scale))
    return [diff_view(n, width, height, channels, seed=seed + 101 * k) for k in range(num_views)]


def _mip_chain(rng, size, kind):
    """Level 0 .. 1x1 of a size x size float4 texture (box-filtered)."""
    y, x = np.mgrid[0:size, 0:size].astype(np.float32) / size
    if kind == "checker":
        c = ((np.floor(x * 8) + np.floor(y * 8)) % 2)[..., None]
        base = rng.random((1, 1, 4), dtype=np.float32) * 0.5
        tex = base + c * (1.0 - base) * rng.random((1, 1, 4), dtype=np.float32)
        tex = tex + 0.1 * rng.random((size, size, 4), dtype=np.float32)
    else:   # alpha: radial bands in the second channel
        r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2)
        tex = np.zeros((size, size, 4), np.float32)
        tex[..., 1] = 0.5 + 0.5 * np.cos(r * 40.0)
        tex[..., 0] = rng.random((size, size), dtype=np.float32)
    levels = [tex.astype(np.float32)]
    while levels[-1].shape[0] > 1:
        t = levels[-1]
        levels.append((0.25 * (t[0::2, 0::2] + t[1::2, 0::2] + t[0::2, 1::2] + t[1::2, 1::2])).astype(np.float32))
    return levels


def texture_atlas(rng, size=32, kinds=("checker", "alpha")):
    """An atlas holding one mip chain per kind, each in a 2*size x size
    region (level l at x offset 2*size - 2*size/2^l), and the TextureSpec of
    each (mip levels past the 1x1 level repeat it)."""
    aw, ah = 2 * size * len(kinds), size
    atlas = np.zeros((ah, aw, 4), np.float32)
    specs = []
    for k, kind in enumerate(kinds):
        spec = np.zeros((), TEXSPEC_DTYPE)
        spec["size"] = (size, size)
        x0 = 2 * size * k
        for l, lev in enumerate(_mip_chain(rng, size, kind)):
            n = lev.shape[0]
            atlas[0:n, x0:x0 + n] = lev
            spec["mipLevels"][l] = (n / aw, n / ah, x0 / aw, 0.0)
            x0 += n
            last = l
        for l in range(last + 1, 16):
            spec["mipLevels"][l] = spec["mipLevels"][last]
        specs.append(spec)
    return atlas, specs


def tex_phong(num_tris, width=256, height=192, samples_log2=0, seed=5, quads=False, area_range=(4.0, 3000.0),
              blend=BLEND_REPLACE, flags=FLAG_DEPTH | FLAG_LERP, tex_scale=(0.05, 20.0), name="tex_phong") -> Scene:
    """Textured Phong triangles (PixelPipe_texPhong, test/shader/Shaders.cu:
    122-199) over the random_tris() layout: ShadedVertex_texPhong records
    (clip position, camera position, camera normal, texture coordinate),
    texture coordinates at log-uniform densities (mip levels 0..5 and
    beyond), four materials -- diffuse texture, diffuse + alpha texture
    (alpha-tested), untextured, untextured with alpha 0.3 (always
    discarded) -- and a few triangles with material indices outside the
    table (discarded)."""
    rng = np.random.default_rng(seed)
    base = random_tris(num_tris, width, height, samples_log2, seed=seed + 77, area_range=area_range)
    n = base.num_tris
    clip = base.vertices[:, :4].reshape(n, 3, 4)
    w = clip[..., 3:4]
    cam = np.concatenate([clip[..., :2], -w, np.ones_like(w)], axis=-1)
    nrm = rng.normal(size=(n, 3, 3)).astype(np.float32)
    nrm[..., 2] = -np.abs(nrm[..., 2]) - 0.3
    nrm = np.concatenate([nrm, np.zeros((n, 3, 1), np.float32)], axis=-1)
    scale = np.exp(np.log(tex_scale[0]) + (np.log(tex_scale[1]) - np.log(tex_scale[0])) * rng.random(n))
    uv0 = rng.random((n, 1, 2)) * 4.0 - 2.0
    uv = uv0 + rng.random((n, 3, 2)) * scale[:, None, None]
    tc = np.concatenate([uv, np.zeros((n, 3, 1)), np.ones((n, 3, 1))], axis=-1)
    verts = np.concatenate([clip, cam, nrm, tc], axis=-1).reshape(n * 3, 16).astype(np.float32)
    atlas, (tex_a, tex_b) = texture_atlas(rng)
    mats = np.zeros(4, MATERIAL_DTYPE)
    mats["diffuseColor"] = [(0.8, 0.7, 0.6, 1.0), (0.5, 0.6, 0.9, 1.0), (0.3, 0.8, 0.4, 0.9), (0.9, 0.2, 0.2, 0.3)]
    mats["specularColor"] = [(0.3, 0.3, 0.3), (0.6, 0.5, 0.4), (0.2, 0.2, 0.2), (0.5, 0.5, 0.5)]
    mats["glossiness"] = [8.0, 32.0, 2.5, 16.0]
    mats["diffuseTexture"][0] = tex_a
    mats["diffuseTexture"][1] = tex_a
    mats["alphaTexture"][1] = tex_b
    tri_mat = rng.integers(0, 4, size=n).astype(np.int32)
    bad = rng.random(n) < 0.01
    tri_mat[bad] = rng.choice(np.array([-1, 4, 1000], np.int32), size=int(bad.sum()))
    return Scene(name, width, height, samples_log2, flags | (FLAG_QUADS if quads else 0), blend, SHADER_TEXPHONG,
                 np.ascontiguousarray(verts), base.indices, materials=mats, tri_material=tri_mat,
                 atlas=np.ascontiguousarray(atlas))
