"""Python front end of libcr.so (the C ABI in include/cr/cr.h).

Mirrors the reference's host classes FW::CudaRaster and FW::CudaSurface
(src/cudaraster/CudaRaster.hpp:143-168, CudaSurface.hpp:60-71) with the same
method names, so tests read like uses of the reference API.  Errors the
reference turns into fail() raise CrError here.

Device memory for vertex and index buffers comes from torch (plumbing
only); torch is imported before libcr.so is loaded so that one HIP runtime
serves both.  Everything the rasterizer computes runs in libcr.so's HIP
kernels; there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import scenes  # noqa: F401  (re-exported workloads)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libcr.so")

PROFILING_DEFAULT, PROFILING_COUNTERS = 0, 1
# cr_profile_counter order (include/cr/cr.h).
PROFILE_COUNTERS = [
    "setup_viewport_cull", "setup_backface_cull", "setup_between_pixels_cull", "setup_clipped",
    "setup_samples_per_tri", "bin_queue_overflows", "bin_tris_per_chunk", "bin_tri_bb_area",
    "bin_tri_single_bin", "coarse_bins", "coarse_rounds_per_bin", "coarse_tris_per_round",
    "coarse_tiles_per_round", "coarse_emits_per_round", "coarse_emits_per_tri", "coarse_case_a",
    "coarse_case_b", "coarse_case_c", "fine_tri_per_tile", "fine_early_z_cull", "fine_empty_cull",
    "fine_z_kill", "fine_warp_util", "fine_blend_rounds", "fine_frag_per_tri", "fine_frag_per_tile",
]
CR_OK, CR_ERR_INVALID, CR_ERR_STATE, CR_ERR_DEVICE, CR_ERR_OVERFLOW, CR_ERR_UNSUPPORTED, CR_ERR_NOMEM = range(7)
FORMAT_RGBA8, FORMAT_DEPTH32, FORMAT_TRIANGLE_ID = 0, 1, 2
RENDER_DEPTH, RENDER_LERP, RENDER_QUADS = 1, 2, 4
SHADER_GOURAUD, SHADER_CONSTANT, SHADER_TEX_PHONG = 0, 1, 2
BLEND_REPLACE, BLEND_SRC_OVER, BLEND_ADDITIVE, BLEND_DEPTH_ONLY = 0, 1, 2, 3

# The exported C symbols (include/cr/cr.h).
EXPORTS = [
    "cr_context_create", "cr_context_destroy", "cr_last_error",
    "cr_surface_create", "cr_surface_destroy", "cr_surface_get_info", "cr_surface_get_device_ptr",
    "cr_surface_download", "cr_surface_upload",
    "cr_set_surfaces", "cr_set_triangle_id_surface", "cr_deferred_clear", "cr_set_pixel_pipe",
    "cr_pipe_lookup", "cr_set_vertex_buffer", "cr_set_index_buffer",
    "cr_draw_triangles", "cr_draw_triangles_async", "cr_draw_check", "cr_reserve",
    "cr_get_stats", "cr_get_counters", "cr_get_profiling_info", "cr_set_debug_params", "cr_abi_version",
    "cr_batch_create", "cr_batch_destroy", "cr_batch_draw", "cr_batch_autotune", "cr_batch_reserve",
    "cr_batch_use_graph", "cr_batch_get_info", "cr_batch_get_view_counters",
    "cr_interpolate", "cr_interpolate_backward", "cr_batch_configure", "cr_vertex_shader_passthrough",
    "cr_surface_resolve", "cr_draw_triangles_noalloc", "cr_get_profile_counters", "cr_set_shader_constants",
    "cr_interpolate_ids", "cr_interpolate_backward_deterministic",
]


class CrError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"[{status}] {message}")
        self.status = status


class PipeDesc(ctypes.Structure):
    _fields_ = [("vertex_struct_size", ctypes.c_int32), ("fragment_shader", ctypes.c_int32),
                ("blend_shader", ctypes.c_int32), ("samples_log2", ctypes.c_int32),
                ("render_mode_flags", ctypes.c_uint32), ("const_color", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [("setupTime", ctypes.c_float), ("binTime", ctypes.c_float),
                ("coarseTime", ctypes.c_float), ("fineTime", ctypes.c_float)]


class Counters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("num_tris", "num_subtris", "bin_entries", "tile_entries",
                                              "fragments", "retries", "visible_subtris", "tile_list_entries")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class DebugParams(ctypes.Structure):
    _fields_ = [("disable_hiz", ctypes.c_int32), ("count_fragments", ctypes.c_int32),
                ("coarse_threads", ctypes.c_int32), ("profiling_mode", ctypes.c_int32)]


class ShaderConstants(ctypes.Structure):
    """cr_shader_constants: the texPhong shader's materials, per-triangle
    material indices and float4 texture atlas (device pointers)."""
    _fields_ = [("materials", ctypes.c_void_p), ("num_materials", ctypes.c_int32),
                ("atlas_width", ctypes.c_int32), ("atlas_height", ctypes.c_int32),
                ("triangle_material_idx", ctypes.c_void_p), ("num_triangle_materials", ctypes.c_int64),
                ("atlas", ctypes.c_void_p)]


class ProfileCounters(ctypes.Structure):
    """cr_profile_counters: numerator and denominator per counter."""
    _fields_ = [("num", ctypes.c_int64 * len(PROFILE_COUNTERS)),
                ("denom", ctypes.c_int64 * len(PROFILE_COUNTERS))]

    def as_dict(self):
        return {n: (int(self.num[i]), int(self.denom[i])) for i, n in enumerate(PROFILE_COUNTERS)}


class View(ctypes.Structure):
    """cr_view: one view of a batched draw."""
    _fields_ = [("color", ctypes.c_void_p), ("depth", ctypes.c_void_p), ("tri_id", ctypes.c_void_p),
                ("vertices", ctypes.c_void_p), ("num_vertices", ctypes.c_int64),
                ("indices", ctypes.c_void_p), ("num_tris", ctypes.c_int32), ("deferred_clear", ctypes.c_int32),
                ("pipe", PipeDesc), ("clear_color", ctypes.c_float * 4), ("clear_depth", ctypes.c_float)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("num_views", ctypes.c_int32), ("fine_tiles_per_group", ctypes.c_int32),
                ("uses_graph", ctypes.c_int32), ("tuned", ctypes.c_int32), ("retries", ctypes.c_int64),
                ("reserved_bytes", ctypes.c_int64), ("last_gpu_ms", ctypes.c_float),
                ("tuned_gpu_ms", ctypes.c_float), ("fine_banded", ctypes.c_int32),
                ("coarse_threads", ctypes.c_int32), ("view_groups", ctypes.c_int32)]

    def as_dict(self):
        return {n: (float(getattr(self, n)) if t is ctypes.c_float else int(getattr(self, n)))
                for n, t in self._fields_}


_lib = None


def lib_path() -> str:
    return LIB_PATH


def load(path: str | None = None):
    """Load libcr.so.  Raises if the library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libcr.so not built ({p}); run __graft_entry__.build() or "
                           "python cudaraster-linux_amd/build.py")
    try:  # one HIP runtime per process: let torch's runtime be the one libcr binds to
        import torch  # noqa: F401
    except Exception:  # pragma: no cover - torch is part of the environment
        pass
    L = ctypes.CDLL(p)
    vp, i32, u32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64
    sig = {
        "cr_context_create": [ctypes.POINTER(vp)],
        "cr_context_destroy": [vp],
        "cr_surface_create": [i32, i32, i32, i32, ctypes.POINTER(vp)],
        "cr_surface_destroy": [vp],
        "cr_surface_get_info": [vp] + [ctypes.POINTER(i32)] * 7,
        "cr_surface_get_device_ptr": [vp, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t)],
        "cr_surface_download": [vp, vp, ctypes.c_size_t],
        "cr_surface_upload": [vp, vp, ctypes.c_size_t],
        "cr_set_surfaces": [vp, vp, vp],
        "cr_set_triangle_id_surface": [vp, vp],
        "cr_deferred_clear": [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_float],
        "cr_set_pixel_pipe": [vp, ctypes.POINTER(PipeDesc)],
        "cr_pipe_lookup": [ctypes.c_char_p, i32, i32, u32, ctypes.POINTER(PipeDesc)],
        "cr_set_vertex_buffer": [vp, vp, i64, i64],
        "cr_set_index_buffer": [vp, vp, i64, i32],
        "cr_draw_triangles": [vp, vp],
        "cr_draw_triangles_async": [vp, vp],
        "cr_draw_check": [vp],
        "cr_reserve": [vp, i64, i64, i64],
        "cr_get_stats": [vp, ctypes.POINTER(Stats)],
        "cr_get_counters": [vp, ctypes.POINTER(Counters)],
        "cr_get_profiling_info": [vp, ctypes.c_char_p, ctypes.c_size_t],
        "cr_set_debug_params": [vp, ctypes.POINTER(DebugParams)],
        "cr_batch_create": [vp, ctypes.POINTER(View), i32, ctypes.POINTER(vp)],
        "cr_batch_destroy": [vp],
        "cr_batch_draw": [vp, vp],
        "cr_batch_autotune": [vp, vp],
        "cr_batch_reserve": [vp, i32, i64, i64, i64],
        "cr_batch_use_graph": [vp, i32],
        "cr_batch_get_info": [vp, ctypes.POINTER(BatchInfo)],
        "cr_batch_get_view_counters": [vp, i32, ctypes.POINTER(Counters)],
        "cr_interpolate": [vp, vp, vp, i64, vp, i32, vp, i32, vp, vp, vp],
        "cr_batch_configure": [vp, vp],
        "cr_interpolate_backward": [vp, vp, i32, i32, vp, i64, vp, i32, vp, i32, vp, vp, vp, vp],
        "cr_vertex_shader_passthrough": [vp, vp, i64, ctypes.POINTER(ctypes.c_float), vp, i32, vp],
        "cr_surface_resolve": [vp, vp, vp],
        "cr_draw_triangles_noalloc": [vp, vp],
        "cr_get_profile_counters": [vp, ctypes.POINTER(ProfileCounters)],
        "cr_set_shader_constants": [vp, ctypes.POINTER(ShaderConstants)],
        "cr_interpolate_ids": [vp, vp, i32, i32, vp, i64, vp, i32, vp, i32, vp, vp],
        "cr_interpolate_backward_deterministic": [vp, vp, i32, i32, vp, i64, vp, i32, vp, i32, vp, vp, vp, vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.cr_last_error.argtypes = [vp]
    L.cr_last_error.restype = ctypes.c_char_p
    L.cr_abi_version.argtypes = []
    L.cr_abi_version.restype = ctypes.c_int32
    _lib = L
    return L


def _check(status, ctx=None):
    if status != CR_OK:
        msg = load().cr_last_error(ctx)
        raise CrError(status, msg.decode() if msg else "")


def pipe_lookup(name, blend=BLEND_REPLACE, samples_log2=0, flags=RENDER_DEPTH | RENDER_LERP) -> PipeDesc:
    d = PipeDesc()
    _check(load().cr_pipe_lookup(name.encode(), blend, samples_log2, flags, ctypes.byref(d)))
    return d


def pipe_desc(vertex_struct_size=32, fragment_shader=SHADER_GOURAUD, blend_shader=BLEND_REPLACE,
              samples_log2=0, render_mode_flags=RENDER_DEPTH | RENDER_LERP, const_color=0xFF0000FF) -> PipeDesc:
    return PipeDesc(vertex_struct_size, fragment_shader, blend_shader, samples_log2, render_mode_flags,
                    const_color & 0xFFFFFFFF)


class CudaSurface:
    """FW::CudaSurface (CudaSurface.hpp:36-78): RGBA8 / DEPTH32 / triangle-id
    render target with the tile-replicated MSAA layout."""

    def __init__(self, size, fmt=FORMAT_RGBA8, num_samples=1):
        L = load()
        h = ctypes.c_void_p()
        _check(L.cr_surface_create(int(size[0]), int(size[1]), int(fmt), int(num_samples), ctypes.byref(h)))
        self._h = h
        vals = [ctypes.c_int32() for _ in range(7)]
        _check(L.cr_surface_get_info(h, *(ctypes.byref(v) for v in vals)))
        (self.width, self.height, self.rounded_width, self.rounded_height, self.texture_width,
         self.num_samples, self.format) = (v.value for v in vals)

    @property
    def handle(self):
        return self._h

    def getSize(self):
        return (self.width, self.height)

    def getRoundedSize(self):
        return (self.rounded_width, self.rounded_height)

    def getTextureSize(self):
        return (self.texture_width, self.rounded_height)

    def getNumSamples(self):
        return self.num_samples

    def getSamplesLog2(self):
        return int(self.num_samples - 1).bit_count()

    @property
    def shape(self):
        return (self.rounded_height, self.texture_width)

    def device_ptr(self):
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        _check(load().cr_surface_get_device_ptr(self._h, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.uint32)
        _check(load().cr_surface_download(self._h, out.ctypes.data, out.nbytes))
        return out

    def resolve(self, out=None, stream=None):
        """Box MSAA resolve on the GPU (cr_surface_resolve): returns a device
        int32 tensor [height, width] of RGBA8 words (row 0 at the bottom)."""
        import torch
        w, h = self.width, self.height
        if out is None:
            out = torch.empty((h, w), dtype=torch.int32, device="cuda")
        elif not (hasattr(out, "is_cuda") and out.is_cuda and out.is_contiguous() and
                  out.dtype in (torch.int32, torch.uint32) and out.numel() >= h * w):
            raise ValueError(f"out must be a contiguous 32-bit device tensor of at least {h}x{w} elements")
        s = None if stream is None else (stream if isinstance(stream, int) else stream.cuda_stream)
        _check(load().cr_surface_resolve(self._h, _ptr(out), s))
        return out

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint32)
        if a.shape != self.shape:
            raise ValueError(f"surface expects {self.shape}, got {a.shape}")
        _check(load().cr_surface_upload(self._h, a.ctypes.data, a.nbytes))

    def close(self):
        if self._h:
            load().cr_surface_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(buf):
    if buf is None:
        return None
    if isinstance(buf, int):
        return buf
    if hasattr(buf, "data_ptr"):
        if not buf.is_cuda:
            raise ValueError("vertex/index buffers must be device tensors")
        return buf.data_ptr()
    raise TypeError("expected a device tensor or a raw device pointer")


class CudaRaster:
    """FW::CudaRaster (CudaRaster.hpp:42-190) over the C ABI."""

    def __init__(self):
        self._ctx = None
        self._keep = {}

    def init(self):
        h = ctypes.c_void_p()
        _check(load().cr_context_create(ctypes.byref(h)))
        self._ctx = h
        return self

    @property
    def ctx(self):
        if self._ctx is None:
            self.init()
        return self._ctx

    def setSurfaces(self, color, depth):
        self._keep["color"], self._keep["depth"] = color, depth
        _check(load().cr_set_surfaces(self.ctx, color.handle if color else None,
                                      depth.handle if depth else None), self.ctx)

    def setTriangleIdSurface(self, surf):
        self._keep["triid"] = surf
        _check(load().cr_set_triangle_id_surface(self.ctx, surf.handle if surf else None), self.ctx)

    def deferredClear(self, color=(0.0, 0.0, 0.0, 0.0), depth=1.0):
        c = (ctypes.c_float * 4)(*[float(x) for x in color])
        _check(load().cr_deferred_clear(self.ctx, c, float(depth)), self.ctx)

    def setPixelPipe(self, pipe):
        if isinstance(pipe, str):
            pipe = pipe_lookup(pipe)
        _check(load().cr_set_pixel_pipe(self.ctx, ctypes.byref(pipe) if pipe is not None else None), self.ctx)
        self._stride = pipe.vertex_struct_size if pipe is not None else None

    def setShaderConstants(self, materials, tri_material, atlas):
        """The texPhong shader's constants (cr_set_shader_constants):
        `materials` a device uint8 tensor of 576-byte Material records
        (scenes.MATERIAL_DTYPE), `tri_material` device int32 [T], `atlas`
        device float32 [H, W, 4]; None clears them."""
        L = load()
        if materials is None:
            self._keep.pop("shader_constants", None)
            _check(L.cr_set_shader_constants(self.ctx, None), self.ctx)
            return
        for name, t, dt in (("materials", materials, "uint8"), ("tri_material", tri_material, "int32"),
                            ("atlas", atlas, "float32")):
            if not (hasattr(t, "is_cuda") and t.is_cuda and t.is_contiguous() and str(t.dtype) == f"torch.{dt}"):
                raise ValueError(f"{name} must be a contiguous device tensor of {dt}")
        if materials.numel() % 576:
            raise ValueError("materials must hold whole 576-byte Material records")
        if atlas.dim() != 3 or atlas.shape[2] != 4:
            raise ValueError("atlas must be [height, width, 4] float32")
        c = ShaderConstants(materials.data_ptr(), materials.numel() // 576, int(atlas.shape[1]), int(atlas.shape[0]),
                            tri_material.data_ptr(), tri_material.numel(), atlas.data_ptr())
        self._keep["shader_constants"] = (materials, tri_material, atlas)
        _check(L.cr_set_shader_constants(self.ctx, ctypes.byref(c)), self.ctx)

    def setVertexBuffer(self, buf, ofs=0, num_vertices=None):
        """num_vertices defaults to the whole tensor divided by the pipe's
        vertex struct size (set the pipe first)."""
        self._keep["vb"] = buf
        if num_vertices is None:
            if buf is None:
                num_vertices = 0
            else:
                stride = getattr(self, "_stride", None)
                if not stride:
                    raise CrError(CR_ERR_STATE, "set the pixel pipe before sizing the vertex buffer")
                num_vertices = (buf.numel() * buf.element_size() - int(ofs)) // stride
        _check(load().cr_set_vertex_buffer(self.ctx, _ptr(buf), int(ofs), int(num_vertices)), self.ctx)

    def setIndexBuffer(self, buf, ofs, num_tris):
        self._keep["ib"] = buf
        _check(load().cr_set_index_buffer(self.ctx, _ptr(buf), int(ofs), int(num_tris)), self.ctx)

    def drawTriangles(self, stream=None):
        s = None if stream is None else (stream if isinstance(stream, int) else stream.cuda_stream)
        _check(load().cr_draw_triangles(self.ctx, s), self.ctx)

    def drawTrianglesNoAlloc(self, stream=None):
        """drawTriangles with the intermediates in the library's fixed
        scratch array (cr_draw_triangles_noalloc): no device allocation."""
        s = None if stream is None else (stream if isinstance(stream, int) else stream.cuda_stream)
        _check(load().cr_draw_triangles_noalloc(self.ctx, s), self.ctx)

    def drawTrianglesAsync(self, stream=None):
        s = None if stream is None else (stream if isinstance(stream, int) else stream.cuda_stream)
        _check(load().cr_draw_triangles_async(self.ctx, s), self.ctx)

    def drawCheck(self):
        return load().cr_draw_check(self.ctx)

    def vertexShaderPassthrough(self, model_pos, pos_to_clip, out, out_stride=16, num_vertices=None, stream=None):
        """The demo's passthrough vertex shader (test/shader/PassThrough.cu:16-35)
        on the GPU: clip position pos_to_clip * (p, 1) of each model position
        (device buffer of Vec3f) into the first 16 bytes of each out_stride-byte
        record of `out` (a vertex buffer for setVertexBuffer).  pos_to_clip
        is a 4x4 matrix indexed [row][column] (numpy or nested lists)."""
        import torch
        m = np.ascontiguousarray(np.asarray(pos_to_clip, dtype=np.float32).reshape(4, 4).T).reshape(16)
        if not (hasattr(model_pos, "is_cuda") and model_pos.is_cuda and model_pos.dtype == torch.float32 and
                model_pos.is_contiguous()):
            raise ValueError("model_pos must be a contiguous float32 device tensor of Vec3f positions")
        n = (model_pos.numel() // 3) if num_vertices is None else int(num_vertices)
        if n < 0 or model_pos.numel() < 3 * n:
            raise ValueError(f"model_pos holds {model_pos.numel()} floats, {n} vertices need {3 * n}")
        if not (hasattr(out, "is_cuda") and out.is_cuda and out.is_contiguous()):
            raise ValueError("out must be a contiguous device tensor")
        if out.numel() * out.element_size() < n * int(out_stride):
            raise ValueError(f"out holds {out.numel() * out.element_size()} bytes, {n} records of "
                             f"{int(out_stride)} bytes need {n * int(out_stride)}")
        s = None if stream is None else (stream if isinstance(stream, int) else stream.cuda_stream)
        _check(load().cr_vertex_shader_passthrough(self.ctx, _ptr(model_pos), n,
                                                   m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _ptr(out),
                                                   int(out_stride), s), self.ctx)

    def reserve(self, max_subtris, max_bin_entries, max_tile_entries):
        _check(load().cr_reserve(self.ctx, int(max_subtris), int(max_bin_entries), int(max_tile_entries)), self.ctx)

    def getStats(self) -> dict:
        s = Stats()
        _check(load().cr_get_stats(self.ctx, ctypes.byref(s)), self.ctx)
        return {"setupTime": s.setupTime, "binTime": s.binTime, "coarseTime": s.coarseTime,
                "fineTime": s.fineTime}

    def getCounters(self) -> dict:
        c = Counters()
        _check(load().cr_get_counters(self.ctx, ctypes.byref(c)), self.ctx)
        return c.as_dict()

    def getProfilingInfo(self) -> str:
        buf = ctypes.create_string_buffer(8192)
        _check(load().cr_get_profiling_info(self.ctx, buf, 8192), self.ctx)
        return buf.value.decode()

    def getProfileCounters(self) -> dict:
        """Counters of the last PROFILING_COUNTERS draw: name -> (num, denom)."""
        c = ProfileCounters()
        _check(load().cr_get_profile_counters(self.ctx, ctypes.byref(c)), self.ctx)
        return c.as_dict()

    def setDebugParams(self, disable_hiz=False, count_fragments=False, coarse_threads=0,
                       profiling_mode=PROFILING_DEFAULT):
        p = DebugParams(1 if disable_hiz else 0, 1 if count_fragments else 0, int(coarse_threads),
                        int(profiling_mode))
        _check(load().cr_set_debug_params(self.ctx, ctypes.byref(p)), self.ctx)

    def close(self):
        if self._ctx is not None:
            load().cr_context_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_buffers(scene, device="cuda"):
    """Upload a scene's vertex and index buffers as device tensors."""
    import torch
    v = torch.from_numpy(np.ascontiguousarray(scene.vertices, dtype=np.float32)).to(device)
    i = torch.from_numpy(np.ascontiguousarray(scene.indices, dtype=np.int32)).to(device)
    return v, i


def scene_pipe(scene) -> PipeDesc:
    return pipe_desc(scene.stride, scene.shader, scene.blend, scene.samples_log2, scene.flags,
                     scene.const_color)


class SceneRenderer:
    """Sets up a CudaRaster for a scene the way test/SceneCR.cpp:149-195 and
    :243-317 do (surfaces, pipe, deferred clear, buffers, draw)."""

    def __init__(self, scene, triangle_ids=True, device="cuda", debug=None):
        self.scene = scene
        n = scene.num_samples
        self.color = CudaSurface((scene.width, scene.height), FORMAT_RGBA8, n)
        self.depth = CudaSurface((scene.width, scene.height), FORMAT_DEPTH32, n)
        self.triid = CudaSurface((scene.width, scene.height), FORMAT_TRIANGLE_ID, n) if triangle_ids else None
        self.cr = CudaRaster().init()
        self.cr.setSurfaces(self.color, self.depth)
        if self.triid is not None:
            self.cr.setTriangleIdSurface(self.triid)
        self.cr.setPixelPipe(scene_pipe(scene))
        if debug:
            self.cr.setDebugParams(**debug)
        self.vbuf, self.ibuf = device_buffers(scene, device)
        if getattr(scene, "materials", None) is not None:
            import torch
            self.shader_constants = (
                torch.from_numpy(np.ascontiguousarray(scene.materials).view(np.uint8).reshape(-1).copy()).to(device),
                torch.from_numpy(np.ascontiguousarray(scene.tri_material, dtype=np.int32)).to(device),
                torch.from_numpy(np.ascontiguousarray(scene.atlas, dtype=np.float32)).to(device))
            self.cr.setShaderConstants(*self.shader_constants)
        self.cr.setVertexBuffer(self.vbuf, 0, scene.vertices.shape[0])
        self.cr.setIndexBuffer(self.ibuf, 0, scene.num_tris)

    def draw(self, clear=None, stream=None):
        sc = self.scene
        if clear if clear is not None else sc.deferred_clear:
            self.cr.deferredClear(sc.clear_color, sc.clear_depth)
        self.cr.drawTriangles(stream)

    def read(self):
        out = {"color": self.color.download(), "depth": self.depth.download()}
        out["triid"] = self.triid.download() if self.triid is not None else None
        return out

    def close(self):
        self.cr.close()
        for s in (self.color, self.depth, self.triid):
            if s is not None:
                s.close()


def _stream_handle(stream):
    return None if stream is None else (stream if isinstance(stream, int) else stream.cuda_stream)


class Batch:
    """A batched multi-view draw (cr_batch_*): every view's surfaces, buffers
    and pipe are fixed at creation; draw() renders all of them, with one
    launch per stage over all views, replayed from a captured hipGraph."""

    def __init__(self, raster, views, keep=()):
        self._raster = raster
        self._keep = list(keep)
        arr = (View * len(views))(*views)
        h = ctypes.c_void_p()
        _check(load().cr_batch_create(raster.ctx, arr, len(views), ctypes.byref(h)), raster.ctx)
        self._h = h

    def draw(self, stream=None):
        _check(load().cr_batch_draw(self._h, _stream_handle(stream)), self._raster.ctx)

    def autotune(self, stream=None):
        _check(load().cr_batch_autotune(self._h, _stream_handle(stream)), self._raster.ctx)

    def reserve(self, view, max_subtris, max_bin_entries, max_tile_entries):
        _check(load().cr_batch_reserve(self._h, int(view), int(max_subtris), int(max_bin_entries),
                                       int(max_tile_entries)), self._raster.ctx)

    def configure(self, fine_tiles_per_group=4, fine_banded=0, coarse_threads=256, view_groups=1):
        c = (ctypes.c_int32 * 4)(fine_tiles_per_group, fine_banded, coarse_threads, view_groups)
        _check(load().cr_batch_configure(self._h, c), self._raster.ctx)

    def use_graph(self, enable=True):
        _check(load().cr_batch_use_graph(self._h, 1 if enable else 0), self._raster.ctx)

    def info(self) -> dict:
        i = BatchInfo()
        _check(load().cr_batch_get_info(self._h, ctypes.byref(i)), self._raster.ctx)
        return i.as_dict()

    def view_counters(self, view) -> dict:
        c = Counters()
        _check(load().cr_batch_get_view_counters(self._h, int(view), ctypes.byref(c)), self._raster.ctx)
        return c.as_dict()

    def close(self):
        if self._h:
            load().cr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchRenderer:
    """All scenes of a list as one batch: per scene its own surfaces, device
    buffers and pipe (what a SceneRenderer per scene would hold)."""

    def __init__(self, scene_list, triangle_ids=True, device="cuda"):
        self.scenes = list(scene_list)
        self.cr = CudaRaster().init()
        self.surfaces, self.buffers, views = [], [], []
        for sc in self.scenes:
            n = sc.num_samples
            col = CudaSurface((sc.width, sc.height), FORMAT_RGBA8, n)
            dep = CudaSurface((sc.width, sc.height), FORMAT_DEPTH32, n)
            tid = CudaSurface((sc.width, sc.height), FORMAT_TRIANGLE_ID, n) if triangle_ids else None
            vb, ib = device_buffers(sc, device)
            self.surfaces.append((col, dep, tid))
            self.buffers.append((vb, ib))
            v = View()
            v.color, v.depth, v.tri_id = col.handle.value, dep.handle.value, (tid.handle.value if tid else None)
            v.vertices, v.num_vertices = vb.data_ptr(), sc.vertices.shape[0]
            v.indices, v.num_tris = ib.data_ptr(), sc.num_tris
            v.deferred_clear = 1 if sc.deferred_clear else 0
            v.pipe = scene_pipe(sc)
            v.clear_color[:] = [float(x) for x in sc.clear_color]
            v.clear_depth = float(sc.clear_depth)
            views.append(v)
        self.batch = Batch(self.cr, views)

    def draw(self, stream=None):
        self.batch.draw(stream)

    def read(self, i):
        col, dep, tid = self.surfaces[i]
        return {"color": col.download(), "depth": dep.download(),
                "triid": tid.download() if tid is not None else None}

    def close(self):
        self.batch.close()
        self.cr.close()
        for trio in self.surfaces:
            for s in trio:
                if s is not None:
                    s.close()


def rgba8_to_image(words) -> np.ndarray:
    """RGBA8 words [H, W] (the surfaces' ABGR packing, row 0 at the bottom)
    -> uint8 [H, W, 4] RGBA with row 0 at the top (image order)."""
    a = np.ascontiguousarray(np.asarray(words).astype(np.uint32))
    return a.view(np.uint8).reshape(a.shape[0], a.shape[1], 4)[::-1].copy()


def save_png(path, words):
    """Writes RGBA8 words [H, W] as an RGBA PNG (PIL)."""
    from PIL import Image
    Image.fromarray(rgba8_to_image(words), "RGBA").save(path)


def load_png(path) -> np.ndarray:
    """Reads an RGBA PNG back as RGBA8 words [H, W], row 0 at the bottom."""
    from PIL import Image
    img = np.ascontiguousarray(np.asarray(Image.open(path).convert("RGBA"))[::-1])
    return img.view(np.uint32).reshape(img.shape[0], img.shape[1])
