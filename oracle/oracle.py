"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, the smoke entry point and bench.py's CPU-baseline leg import
this module; it checks the HIP path, it is never the product.  The C library
(liboracle.so, built from cr_oracle.c by the Makefile beside it) restates the
reference's rasterization rules; see cr_oracle.h for the file:line map.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class Pipe(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("samplesLog2", ctypes.c_int32), ("renderModeFlags", ctypes.c_uint32),
                ("blend", ctypes.c_int32), ("shader", ctypes.c_int32),
                ("constColor", ctypes.c_uint32), ("vertexStride", ctypes.c_int32),
                ("deferredClear", ctypes.c_int32), ("clearColor", ctypes.c_uint32),
                ("clearDepth", ctypes.c_uint32), ("numMaterials", ctypes.c_int32),
                ("atlasWidth", ctypes.c_int32), ("atlasHeight", ctypes.c_int32),
                ("materials", ctypes.c_void_p), ("triangleMaterialIdx", ctypes.c_void_p),
                ("atlas", ctypes.c_void_p)]


class Header(ctypes.Structure):
    _fields_ = [("v0x", ctypes.c_int16), ("v0y", ctypes.c_int16),
                ("v1x", ctypes.c_int16), ("v1y", ctypes.c_int16),
                ("v2x", ctypes.c_int16), ("v2y", ctypes.c_int16),
                ("misc", ctypes.c_uint32)]


class Data(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("zx", "zy", "zb", "zslope")] + \
               [(n, ctypes.c_int32) for n in ("wx", "wy", "wb", "ux", "uy", "ub", "vx", "vy", "vb")] + \
               [(n, ctypes.c_uint32) for n in ("vi0", "vi1", "vi2")]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("numTris", "numSubtris", "culledFrustum",
                                              "culledBackface", "culledBetween", "clipped",
                                              "fragments", "coveredSamples", "zKills", "trisVisible")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


HEADER_DTYPE = np.dtype([("v0x", "<i2"), ("v0y", "<i2"), ("v1x", "<i2"), ("v1y", "<i2"),
                         ("v2x", "<i2"), ("v2y", "<i2"), ("misc", "<u4")])
DATA_DTYPE = np.dtype([(n, "<u4") for n in ("zx", "zy", "zb", "zslope")] +
                      [(n, "<i4") for n in ("wx", "wy", "wb", "ux", "uy", "ub", "vx", "vy", "vb")] +
                      [(n, "<u4") for n in ("vi0", "vi1", "vi2")])


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        f4 = ctypes.POINTER(ctypes.c_float)
        L.cro_to_abgr.restype = ctypes.c_uint32
        L.cro_to_abgr.argtypes = [f4]
        L.cro_to_abgr_host.restype = ctypes.c_uint32
        L.cro_to_abgr_host.argtypes = [f4]
        L.cro_blend.restype = ctypes.c_uint32
        L.cro_blend.argtypes = [ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32]
        for fn in ("cro_encode_depth", "cro_decode_depth"):
            getattr(L, fn).restype = ctypes.c_uint32
            getattr(L, fn).argtypes = [ctypes.c_uint32]
        L.cro_clear_depth.restype = ctypes.c_uint32
        L.cro_clear_depth.argtypes = [ctypes.c_float]
        L.cro_msaa_pattern.restype = ctypes.c_int32
        L.cro_msaa_pattern.argtypes = [ctypes.c_int32, ctypes.c_int32]
        L.cro_select_centroid.restype = ctypes.c_int32
        L.cro_select_centroid.argtypes = [ctypes.c_int32, ctypes.c_uint32]
        L.cro_setup_pleq.restype = None
        L.cro_setup_pleq.argtypes = [f4] + [ctypes.c_int32] * 6 + [ctypes.c_float, ctypes.c_int32,
                                                                     ctypes.POINTER(ctypes.c_uint32)]
        L.cro_clip_triangle.restype = ctypes.c_int32
        L.cro_clip_triangle.argtypes = [f4, f4, f4, f4]
        L.cro_cover_point.restype = ctypes.c_int32
        L.cro_cover_point.argtypes = [ctypes.POINTER(Header), ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int64, ctypes.c_int64]
        L.cro_cover_pixel.restype = ctypes.c_uint32
        L.cro_cover_pixel.argtypes = [ctypes.POINTER(Header), ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.cro_setup_triangle.restype = ctypes.c_int32
        L.cro_setup_triangle.argtypes = [ctypes.POINTER(Pipe), f4, f4, f4, ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(Header), ctypes.POINTER(Data),
                                         ctypes.POINTER(ctypes.c_int32)]
        L.cro_setup.restype = ctypes.c_int64
        L.cro_setup.argtypes = [ctypes.POINTER(Pipe), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32]
        L.cro_render.restype = ctypes.c_int32
        L.cro_render.argtypes = [ctypes.POINTER(Pipe), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                 ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_int32, ctypes.POINTER(Stats)]
        for fn in ("cro_log2f", "cro_exp2f"):
            getattr(L, fn).restype = ctypes.c_float
            getattr(L, fn).argtypes = [ctypes.c_float]
        L.cro_powf.restype = ctypes.c_float
        L.cro_powf.argtypes = [ctypes.c_float, ctypes.c_float]
        L.cro_tex_sample.restype = None
        L.cro_tex_sample.argtypes = [f4, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_float, f4]
        L.cro_texture2d.restype = None
        L.cro_texture2d.argtypes = [ctypes.c_void_p, f4, ctypes.c_int32, ctypes.c_int32, f4, f4, f4, f4,
                                    ctypes.POINTER(ctypes.c_float)]
        L.cro_lighting.restype = None
        L.cro_lighting.argtypes = [f4, f4, ctypes.c_void_p, f4, f4]
        L.cro_vertex_passthrough.restype = None
        L.cro_vertex_passthrough.argtypes = [f4, ctypes.c_int64, f4, f4, ctypes.c_int64]
        _lib = L
    return _lib


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def to_abgr(c) -> int:
    c = np.ascontiguousarray(c, dtype=np.float32)
    return int(lib().cro_to_abgr(_fptr(c)))


def to_abgr_host(c) -> int:
    c = np.ascontiguousarray(c, dtype=np.float32)
    return int(lib().cro_to_abgr_host(_fptr(c)))


def blend(mode, src, dst) -> int:
    return int(lib().cro_blend(int(mode), int(src) & 0xFFFFFFFF, int(dst) & 0xFFFFFFFF))


def clear_depth(d: float) -> int:
    return int(lib().cro_clear_depth(float(d)))


def encode_depth(d: int) -> int:
    return int(lib().cro_encode_depth(int(d)))


def decode_depth(d: int) -> int:
    return int(lib().cro_decode_depth(int(d)))


def msaa_pattern(s, i) -> int:
    return int(lib().cro_msaa_pattern(s, i))


def select_centroid(s, mask) -> int:
    return int(lib().cro_select_centroid(s, mask))


def setup_pleq(values, v0, d1, d2, area_rcp, samples_log2):
    vals = np.ascontiguousarray(values, dtype=np.float32)
    out = (ctypes.c_uint32 * 3)()
    lib().cro_setup_pleq(_fptr(vals), int(v0[0]), int(v0[1]), int(d1[0]), int(d1[1]),
                         int(d2[0]), int(d2[1]), float(np.float32(area_rcp)), int(samples_log2), out)
    return [int(out[0]), int(out[1]), int(out[2])]


def clip_triangle(v0, v1, v2):
    bary = np.zeros(18, dtype=np.float32)
    vs = [np.ascontiguousarray(v, dtype=np.float32) for v in (v0, v1, v2)]
    n = lib().cro_clip_triangle(_fptr(bary), *(_fptr(v) for v in vs))
    return bary[: 2 * n].reshape(n, 2)


def make_pipe(scene, deferred_clear=None) -> Pipe:
    p = Pipe()
    p.width, p.height = scene.width, scene.height
    p.samplesLog2 = scene.samples_log2
    p.renderModeFlags = scene.flags
    p.blend = scene.blend
    p.shader = scene.shader
    p.constColor = scene.const_color
    p.vertexStride = scene.stride
    dc = scene.deferred_clear if deferred_clear is None else deferred_clear
    p.deferredClear = 1 if dc else 0
    p.clearColor = to_abgr_host(scene.clear_color)
    p.clearDepth = clear_depth(scene.clear_depth)
    if getattr(scene, "materials", None) is not None:
        # texPhong constants: the arrays stay referenced by the Pipe object.
        mats = np.ascontiguousarray(scene.materials)
        tri_mat = np.ascontiguousarray(scene.tri_material, dtype=np.int32)
        atlas = np.ascontiguousarray(scene.atlas, dtype=np.float32)
        p._keep = (mats, tri_mat, atlas)
        p.numMaterials = mats.shape[0]
        p.materials = mats.ctypes.data
        p.triangleMaterialIdx = tri_mat.ctypes.data
        p.atlas = atlas.ctypes.data
        p.atlasHeight, p.atlasWidth = atlas.shape[0], atlas.shape[1]
    return p


def setup_triangle(scene, t):
    """Set up triangle t of a scene; returns (headers, datas, cull code)."""
    p = make_pipe(scene)
    idx = np.ascontiguousarray(scene.indices[t], dtype=np.int32)
    v = [np.ascontiguousarray(scene.vertices[i, :4], dtype=np.float32) for i in idx]
    hdr = (Header * 7)()
    dat = (Data * 7)()
    cull = ctypes.c_int32(0)
    n = lib().cro_setup_triangle(ctypes.byref(p), *(_fptr(x) for x in v),
                                 idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), hdr, dat,
                                 ctypes.byref(cull))
    return list(hdr[:n]), list(dat[:n]), int(cull.value)


def setup(scene, nthreads=None):
    """Whole-buffer setup.  Returns (triSubtris u8[T], headers, datas, subtriTri)."""
    T = scene.num_tris
    max_slots = T * 7 + 1
    tri_subtris = np.zeros(max(T, 1), dtype=np.uint8)
    hdr = np.zeros(max_slots, dtype=HEADER_DTYPE)
    dat = np.zeros(max_slots, dtype=DATA_DTYPE)
    sub = np.zeros(max_slots, dtype=np.uint32)
    p = make_pipe(scene)
    verts = np.ascontiguousarray(scene.vertices, dtype=np.float32)
    idx = np.ascontiguousarray(scene.indices, dtype=np.int32)
    n = lib().cro_setup(ctypes.byref(p), verts.ctypes.data, verts.shape[0], idx.ctypes.data, T,
                        tri_subtris.ctypes.data, hdr.ctypes.data, dat.ctypes.data, sub.ctypes.data,
                        max_slots, nthreads or os.cpu_count() or 1)
    if n < 0:
        raise RuntimeError("oracle setup overflow")
    return tri_subtris[:T], hdr[:n], dat[:n], sub[:n]


def render(scene, color=None, depth=None, triid=True, nthreads=None, deferred_clear=None):
    """Draw a scene.  Returns dict(color, depth, triid, stats) with surfaces
    shaped (ceil8(H), ceil8(W)*N) as uint32."""
    shape = scene.surface_shape
    color = np.zeros(shape, dtype=np.uint32) if color is None else np.array(color, dtype=np.uint32)
    depth = np.zeros(shape, dtype=np.uint32) if depth is None else np.array(depth, dtype=np.uint32)
    tid = None
    if triid is not None and triid is not False:
        tid = np.full(shape, 0xFFFFFFFF, dtype=np.uint32) if triid is True else np.array(triid, dtype=np.uint32)
    p = make_pipe(scene, deferred_clear)
    verts = np.ascontiguousarray(scene.vertices, dtype=np.float32)
    idx = np.ascontiguousarray(scene.indices, dtype=np.int32)
    st = Stats()
    rc = lib().cro_render(ctypes.byref(p), verts.ctypes.data, verts.shape[0], idx.ctypes.data,
                          scene.num_tris, color.ctypes.data, depth.ctypes.data,
                          tid.ctypes.data if tid is not None else None,
                          nthreads or os.cpu_count() or 1, ctypes.byref(st))
    if rc != 0:
        raise ValueError("oracle rejected the draw arguments")
    return {"color": color, "depth": depth, "triid": tid, "stats": st.as_dict()}


def vertex_passthrough(model_pos, pos_to_clip, out=None, stride_floats=4):
    """Clip positions of model positions [V, 3] under the 4x4 matrix
    pos_to_clip (row i, column j as a numpy array): records of stride_floats
    floats, clip position first (out's other floats are kept when given)."""
    p = np.ascontiguousarray(model_pos, dtype=np.float32).reshape(-1, 3)
    m = np.ascontiguousarray(np.asarray(pos_to_clip, dtype=np.float32).T).reshape(16)   # Mat4f order
    n = p.shape[0]
    o = np.zeros((n, stride_floats), np.float32) if out is None else np.array(out, dtype=np.float32).reshape(n, stride_floats)
    lib().cro_vertex_passthrough(_fptr(p), n, _fptr(m), _fptr(o), stride_floats)
    return o


def resolve(color, width, height, samples_log2):
    """Box MSAA resolve of an RGBA8 surface (texture layout of
    CudaSurface.cpp:86-91: per 8-pixel tile column, sample 0 of its 8
    pixels, then sample 1, ...).  The reference's own resolve is commented
    out (CudaSurface.hpp:73); per channel this is the mean of the samples
    rounded half up, (sum + N/2) >> log2(N).  Returns uint32 [height, width]."""
    n = 1 << samples_log2
    c = np.asarray(color, dtype=np.uint32)
    rw = c.shape[1] // n                                   # rounded width
    t = c[:height].reshape(height, rw // 8, n, 8)          # [y, tile column, sample, x in tile]
    t = t.transpose(0, 1, 3, 2).reshape(height, rw, n)[:, :width]   # [y, x, sample]
    out = np.zeros((height, width), np.uint32)
    for ch in range(4):
        v = ((t >> np.uint32(8 * ch)) & np.uint32(0xFF)).astype(np.uint32).sum(axis=2)
        out |= ((v + np.uint32(n >> 1)) >> np.uint32(samples_log2)) << np.uint32(8 * ch)
    return out


def log2f(x) -> float:
    return float(lib().cro_log2f(float(x)))


def exp2f(y) -> float:
    return float(lib().cro_exp2f(float(y)))


def powf(x, y) -> float:
    return float(lib().cro_powf(float(x), float(y)))


def tex_sample(atlas, u, v):
    a = np.ascontiguousarray(atlas, dtype=np.float32)
    out = np.zeros(4, np.float32)
    lib().cro_tex_sample(_fptr(a), a.shape[1], a.shape[0], float(u), float(v), _fptr(out))
    return out


def texture2d(spec, atlas, tex, tex_dx, tex_dy):
    """texture2D of a TEXSPEC_DTYPE record; returns (rgba, lod)."""
    a = np.ascontiguousarray(atlas, dtype=np.float32)
    sp = np.ascontiguousarray(np.asarray(spec).reshape(1))
    vs = [np.ascontiguousarray(v, dtype=np.float32) for v in (tex, tex_dx, tex_dy)]
    out = np.zeros(4, np.float32)
    lod = ctypes.c_float()
    lib().cro_texture2d(sp.ctypes.data, _fptr(a), a.shape[1], a.shape[0], *(_fptr(v) for v in vs), _fptr(out),
                        ctypes.byref(lod))
    return out, float(lod.value)


def lighting(camera_pos, camera_normal, material, diffuse):
    m = np.ascontiguousarray(np.asarray(material).reshape(1))
    vs = [np.ascontiguousarray(v, dtype=np.float32) for v in (camera_pos, camera_normal, diffuse)]
    out = np.zeros(3, np.float32)
    lib().cro_lighting(_fptr(vs[0]), _fptr(vs[1]), m.ctypes.data, _fptr(vs[2]), _fptr(out))
    return out
