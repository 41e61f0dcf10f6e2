"""Differentiable rendering on top of the rasterizer (libcr.so's
cr_interpolate / cr_interpolate_backward behind a torch.autograd.Function).

A frame is rasterized exactly (coverage, depth test, per-pixel triangle ids:
the same kernels as any draw); C float attributes per vertex are then
interpolated with perspective-correct barycentrics at pixel centres.  The
backward pass propagates dL/dimage to the attributes and to the clip-space
positions (x, y, w) through the barycentrics, with visibility fixed by the
forward triangle ids.  torch provides device memory and streams only; both
passes are HIP kernels in libcr.so.
"""
from __future__ import annotations

import ctypes

from . import (CR_ERR_OVERFLOW, CR_OK, Batch, CrError, CudaRaster, CudaSurface, FORMAT_DEPTH32, FORMAT_RGBA8,
               FORMAT_TRIANGLE_ID, View, _check, load, pipe_desc, SHADER_CONSTANT, BLEND_REPLACE, RENDER_DEPTH)


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


class Interpolate:
    """Namespace for the autograd function (built lazily so that importing
    this module does not need torch)."""
    _fn = None

    @classmethod
    def function(cls):
        if cls._fn is not None:
            return cls._fn
        import torch

        class _Interpolate(torch.autograd.Function):
            @staticmethod
            def forward(ctx, pos, attr, tris, tri_image, raster, ids_given):
                H, W = tri_image.shape
                C = attr.shape[1]
                out = torch.empty((H, W, C), dtype=torch.float32, device=pos.device)
                if ids_given:
                    # Visibility from the caller's id image, not from whatever
                    # the rasterizer drew last.
                    raster._interpolate_ids(pos, attr, tris, out, tri_image)
                else:
                    raster._interpolate(pos, attr, tris, out, tri_image)
                ctx.save_for_backward(pos, attr, tris, tri_image)
                ctx.raster = raster
                return out

            @staticmethod
            def backward(ctx, grad_out):
                pos, attr, tris, tri_image = ctx.saved_tensors
                gp, ga = ctx.raster._interpolate_backward(pos, attr, tris, tri_image, grad_out.contiguous())
                return gp, ga, None, None, None, None

        cls._fn = _Interpolate
        return cls._fn


class InterpolateViews:
    """Autograd function over several views of one resolution: one forward
    and one backward call for all of them (the per-view host work is a few
    library calls, not a pass through the autograd engine)."""
    _fn = None

    @classmethod
    def function(cls):
        if cls._fn is not None:
            return cls._fn
        import torch

        class _InterpolateViews(torch.autograd.Function):
            @staticmethod
            def forward(ctx, raster, n, *tensors):
                pos, attr, tris = tensors[:n], tensors[n:2 * n], tensors[2 * n:3 * n]
                H, W, C = raster.height, raster.width, attr[0].shape[1]
                out = torch.empty((n, H, W, C), dtype=torch.float32, device=pos[0].device)
                ids = torch.empty((n, H, W), dtype=torch.int32, device=pos[0].device)
                # Visibility of every view in one batched draw, then each
                # view's interpolation from its own triangle-id surface.
                surfaces = raster._draw_views(pos, tris)
                for i in range(n):
                    raster._interpolate(pos[i].detach(), attr[i].detach(), tris[i], out[i], ids[i],
                                        surface=surfaces[i])
                ctx.save_for_backward(ids, *tensors)
                ctx.raster, ctx.n = raster, n
                ctx.mark_non_differentiable(ids)
                return out, ids

            @staticmethod
            def backward(ctx, grad_out, _grad_ids):
                n = ctx.n
                ids, *tensors = ctx.saved_tensors
                pos, attr, tris = tensors[:n], tensors[n:2 * n], tensors[2 * n:3 * n]
                grad_out = grad_out.contiguous()
                gp, ga = [], []
                for i in range(n):
                    a, b = ctx.raster._interpolate_backward(pos[i], attr[i], tris[i], ids[i], grad_out[i])
                    gp.append(a)
                    ga.append(b)
                return (None, None, *gp, *ga, *([None] * n))

        cls._fn = _InterpolateViews
        return cls._fn


class DiffRasterizer:
    """Single-sample rasterizer + differentiable interpolation for one
    resolution.  render(pos, attr, tris) -> image [H, W, C] (row 0 at the
    bottom); gradients flow to pos ([V, 4] clip space) and attr ([V, C]).
    deterministic=True makes the gradients bit-reproducible from run to run
    (fixed-point accumulation, cr_interpolate_backward_deterministic)."""

    def __init__(self, width, height, deterministic=False):
        self.width, self.height = int(width), int(height)
        self.deterministic = bool(deterministic)
        self._sized = False       # capacities known from a synchronous draw
        self.cr = CudaRaster().init()
        self.color = CudaSurface((width, height), FORMAT_RGBA8, 1)
        self.depth = CudaSurface((width, height), FORMAT_DEPTH32, 1)
        self.ids = CudaSurface((width, height), FORMAT_TRIANGLE_ID, 1)
        self.cr.setSurfaces(self.color, self.depth)
        self.cr.setTriangleIdSurface(self.ids)
        # Positions only (16-byte vertices), depth test on, constant colour.
        self.cr.setPixelPipe(pipe_desc(16, SHADER_CONSTANT, BLEND_REPLACE, 0, RENDER_DEPTH))
        # render_views: the batched visibility draw of the last view sizes
        # (signature, cr_batch, per-view surfaces, staging buffers).
        self._views = None

    def _draw_views(self, pos, tris):
        """Exact visibility of several views in one batched draw (cr_batch:
        one launch per stage over all views, the same images as drawing them
        in turn); returns each view's triangle-id surface.  The batch, its
        per-view surfaces and staging buffers for the positions and indices
        are kept for the next call with the same view sizes."""
        import torch
        for p, t in zip(pos, tris):
            if not (p.is_cuda and p.dtype == torch.float32 and p.dim() == 2 and p.shape[1] == 4):
                raise CrError(1, "pos must be a float32 [V, 4] device tensor")
            if not (t.is_cuda and t.dtype == torch.int32 and t.dim() == 2 and t.shape[1] == 3):
                raise CrError(1, "tris must be an int32 [T, 3] device tensor")
        key = tuple((int(p.shape[0]), int(t.shape[0]), p.device.index) for p, t in zip(pos, tris))
        if self._views is None or self._views[0] != key:
            self._close_views()
            vbuf = [torch.empty((k[0], 4), dtype=torch.float32, device=p.device) for k, p in zip(key, pos)]
            ibuf = [torch.empty((k[1], 3), dtype=torch.int32, device=p.device) for k, p in zip(key, pos)]
            surfaces, views = [], []
            pipe = pipe_desc(16, SHADER_CONSTANT, BLEND_REPLACE, 0, RENDER_DEPTH)
            try:
                for vb, ib in zip(vbuf, ibuf):
                    trio = (CudaSurface((self.width, self.height), FORMAT_RGBA8, 1),
                            CudaSurface((self.width, self.height), FORMAT_DEPTH32, 1),
                            CudaSurface((self.width, self.height), FORMAT_TRIANGLE_ID, 1))
                    surfaces.append(trio)
                    v = View()
                    v.color, v.depth, v.tri_id = (s.handle.value for s in trio)
                    v.vertices, v.num_vertices = (vb.data_ptr() if vb.numel() else None), vb.shape[0]
                    v.indices, v.num_tris = (ib.data_ptr() if ib.numel() else None), ib.shape[0]
                    v.deferred_clear = 1
                    v.pipe = pipe
                    v.clear_color[:] = [0.0, 0.0, 0.0, 0.0]
                    v.clear_depth = 1.0
                    views.append(v)
                batch = Batch(self.cr, views)
            except Exception:
                for trio in surfaces:
                    for sf in trio:
                        sf.close()
                raise
            self._views = (key, batch, surfaces, vbuf, ibuf)
        _, batch, surfaces, vbuf, ibuf = self._views
        for vb, ib, p, t in zip(vbuf, ibuf, pos, tris):
            vb.copy_(p.detach())
            ib.copy_(t)
        batch.draw(torch.cuda.current_stream())
        return [trio[2] for trio in surfaces]

    def _close_views(self):
        if self._views is not None:
            _, batch, surfaces, _, _ = self._views
            self._views = None
            batch.close()
            for trio in surfaces:
                for sf in trio:
                    sf.close()

    def _draw(self, pos, tris):
        """Exact visibility of pos/tris into the triangle-id surface."""
        import torch
        if pos.dtype != torch.float32 or pos.dim() != 2 or pos.shape[1] != 4 or not pos.is_contiguous():
            raise CrError(1, "pos must be a contiguous float32 [V, 4] device tensor")
        if tris.dtype != torch.int32 or tris.dim() != 2 or tris.shape[1] != 3 or not tris.is_contiguous():
            raise CrError(1, "tris must be a contiguous int32 [T, 3] device tensor")
        stream = torch.cuda.current_stream()
        self.sync()                       # the previous asynchronous draw fitted its capacities
        self.cr.deferredClear((0.0, 0.0, 0.0, 0.0), 1.0)
        self.cr.setVertexBuffer(pos.detach(), 0, pos.shape[0])
        self.cr.setIndexBuffer(tris, 0, tris.shape[0])
        if self._sized:
            # Queued behind the previous view's work on torch's stream; the
            # host does not wait for it (checked at the next call or sync()).
            self.cr.drawTrianglesAsync(stream)
        else:
            # First draw: synchronous, sizes the queues (grown and re-run on
            # overflow), then twice its demand is reserved for later views.
            self.cr.drawTriangles(stream)
            c = self.cr.getCounters()
            self.cr.reserve(2 * c["num_subtris"] + 4096, 2 * c["bin_entries"] + 4096, 2 * c["tile_entries"] + 4096)
            self._sized = True

    def rasterize(self, pos, tris):
        """Exact visibility: returns the triangle id image [H, W] (int32, -1
        where empty) by drawing pos/tris and reading the id surface."""
        import torch
        self._draw(pos, tris)
        tri_image = torch.empty((self.height, self.width), dtype=torch.int32, device=pos.device)
        dummy = torch.empty(1, dtype=torch.float32, device=pos.device)   # no channels: ids only
        self._interpolate(pos.detach(), None, tris, dummy, tri_image, channels=0)
        return tri_image

    def render(self, pos, attr, tris, tri_image=None):
        """Interpolated attributes [H, W, C].  Without tri_image, draws pos/tris
        first; the interpolation then also writes the triangle ids.  With
        tri_image (int32 [H, W], e.g. from rasterize()), interpolates over
        those ids, whatever has been drawn since."""
        import torch
        given = tri_image is not None
        if not given:
            self._draw(pos, tris)
            tri_image = torch.empty((self.height, self.width), dtype=torch.int32, device=pos.device)
        return Interpolate.function().apply(pos, attr, tris, tri_image, self, given)

    def render_views(self, pos, attr, tris):
        """Several views at once: lists of pos [V_i, 4], attr [V_i, C] and
        tris [T_i, 3] (same C for all) -> (images [n, H, W, C], triangle ids
        [n, H, W]); gradients flow to every pos and attr."""
        if not (len(pos) == len(attr) == len(tris)) or len(pos) == 0:
            raise CrError(1, "render_views needs equally long, non-empty lists")
        C = attr[0].shape[1] if attr[0].dim() == 2 else -1
        for a in attr:
            if a.dim() != 2 or a.shape[1] != C or a.dtype.itemsize != 4 or not a.is_contiguous():
                raise CrError(1, "render_views: every attr must be a contiguous float32 [V_i, C] tensor, same C")
        return InterpolateViews.function().apply(self, len(pos), *pos, *attr, *tris)

    # -- raw calls ------------------------------------------------------------
    def _check_inputs(self, pos, attr, tris, tri_image, channels):
        """Shapes, dtypes, devices and layout of what the kernels read, checked
        before any pointer reaches them."""
        import torch
        if not (pos.is_cuda and pos.dtype == torch.float32 and pos.dim() == 2 and pos.shape[1] == 4 and
                pos.is_contiguous()):
            raise CrError(1, "pos must be a contiguous float32 [V, 4] device tensor")
        if not (tris.is_cuda and tris.dtype == torch.int32 and tris.dim() == 2 and tris.shape[1] == 3 and
                tris.is_contiguous()):
            raise CrError(1, "tris must be a contiguous int32 [T, 3] device tensor")
        if channels > 0 and not (attr is not None and attr.is_cuda and attr.dtype == torch.float32 and
                                 attr.dim() == 2 and attr.shape == (pos.shape[0], channels) and attr.is_contiguous()):
            raise CrError(1, "attr must be a contiguous float32 [V, C] device tensor with one row per vertex")
        if not (tri_image.is_cuda and tri_image.dtype == torch.int32 and tri_image.is_contiguous() and
                tuple(tri_image.shape) == (self.height, self.width)):
            raise CrError(1, "tri_image must be a contiguous int32 [H, W] device tensor at the rasterizer's resolution")

    def _interpolate(self, pos, attr, tris, out, tri_image, channels=None, surface=None):
        """Interpolates from the triangle ids of `surface` (default: this
        rasterizer's own, from the last _draw) and copies them to tri_image."""
        C = attr.shape[1] if channels is None else channels
        self._check_inputs(pos, attr, tris, tri_image, C)
        hw = self.height * self.width
        if C > 0 and (out.numel() != hw * C or not out.is_contiguous() or not out.is_cuda):
            raise CrError(1, "interpolation output buffers do not match the rasterizer's resolution")
        ids = self.ids if surface is None else surface
        _check(load().cr_interpolate(self.cr.ctx, ids.handle, pos.data_ptr(), pos.shape[0],
                                     attr.data_ptr() if attr is not None else None, C, tris.data_ptr(),
                                     tris.shape[0], out.data_ptr(), tri_image.data_ptr(), _stream()), self.cr.ctx)

    def _interpolate_ids(self, pos, attr, tris, out, tri_image):
        C = attr.shape[1]
        self._check_inputs(pos, attr, tris, tri_image, C)
        if out.numel() != self.height * self.width * C or not out.is_contiguous():
            raise CrError(1, "interpolation output buffer does not match the rasterizer's resolution")
        _check(load().cr_interpolate_ids(self.cr.ctx, tri_image.data_ptr(), self.width, self.height, pos.data_ptr(),
                                         pos.shape[0], attr.data_ptr(), C, tris.data_ptr(), tris.shape[0],
                                         out.data_ptr(), _stream()), self.cr.ctx)

    def _interpolate_backward(self, pos, attr, tris, tri_image, grad_out):
        import torch
        self._check_inputs(pos, attr, tris, tri_image, attr.shape[1])
        if (grad_out.dtype != torch.float32 or grad_out.numel() != tri_image.numel() * attr.shape[1] or
                not grad_out.is_contiguous()):
            raise CrError(1, "gradient of the image does not match its shape")
        gp = torch.empty_like(pos)
        ga = torch.empty_like(attr)
        H, W = tri_image.shape
        fn = load().cr_interpolate_backward_deterministic if self.deterministic else load().cr_interpolate_backward
        _check(fn(self.cr.ctx, tri_image.data_ptr(), W, H, pos.data_ptr(), pos.shape[0], attr.data_ptr(),
                  attr.shape[1], tris.data_ptr(), tris.shape[0], grad_out.data_ptr(), gp.data_ptr(), ga.data_ptr(),
                  _stream()), self.cr.ctx)
        return gp, ga

    def sync(self):
        """Waits for the last asynchronous visibility draw; raises CrError if it
        exceeded the reserved queue capacities (its triangle ids are then
        incomplete).  The next draw is then synchronous again -- it grows the
        queues and re-runs until they fit, and reserves for later views -- so
        re-running the view succeeds."""
        st = self.cr.drawCheck()
        if st == CR_ERR_OVERFLOW:
            self._sized = False
            raise CrError(st, "visibility draw exceeded its queue capacities; re-run the view")
        if st != CR_OK:
            _check(st, self.cr.ctx)

    def close(self):
        try:
            self.sync()
        finally:
            self._close_views()
            self.cr.close()
            for s in (self.color, self.depth, self.ids):
                s.close()
