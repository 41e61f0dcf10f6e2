"""The passthrough vertex shader on the GPU (cr_vertex_shader_passthrough),
against the oracle: clip positions bit-identical, the rest of each vertex
record untouched, and the demo cube drawn from model positions through it."""
import numpy as np
import pytest

from test_gpu_parity import assert_parity

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_clip_positions_bit_identical(oracle_lib):
    torch = _torch()
    import cudaraster
    rng = np.random.default_rng(11)
    p = (rng.standard_normal((100_003, 3)) * 20.0).astype(np.float32)
    m = (rng.standard_normal((4, 4)) * np.array([1.0, 0.5, 3.0, 1e2])).astype(np.float32)
    cr = cudaraster.CudaRaster().init()
    try:
        dp = torch.from_numpy(p).cuda()
        out = torch.full((p.shape[0], 8), 7.5, dtype=torch.float32, device="cuda")   # 32-byte records
        cr.vertexShaderPassthrough(dp, m, out, out_stride=32)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
    finally:
        cr.close()
    ref = oracle_lib.vertex_passthrough(p, m)
    assert np.array_equal(got[:, :4].view(np.uint32), ref.view(np.uint32))
    assert np.all(got[:, 4:] == 7.5)


def test_argument_checks():
    torch = _torch()
    import cudaraster
    cr = cudaraster.CudaRaster().init()
    try:
        p = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
        out = torch.zeros((4, 8), dtype=torch.float32, device="cuda")
        for stride in (0, 8, 24):
            with pytest.raises(cudaraster.CrError):
                cr.vertexShaderPassthrough(p, np.eye(4), out, out_stride=stride)
        with pytest.raises(cudaraster.CrError):
            cr.vertexShaderPassthrough(p, np.eye(4), out.view(-1)[1:], out_stride=16)   # misaligned
        cr.vertexShaderPassthrough(p, np.eye(4), out, out_stride=16, num_vertices=0)      # nothing to do
    finally:
        cr.close()


def test_cube_from_model_positions(oracle_lib):
    """The demo's frame: model positions -> vertex shader -> drawTriangles,
    against the oracle drawing the oracle-transformed vertices."""
    torch = _torch()
    import dataclasses
    from cudaraster import scenes
    import cudaraster
    width, height = 1024, 768
    m = scenes.demo_camera(width, height).astype(np.float32)
    base = scenes.cube(width, height)
    verts = base.vertices.copy()
    verts[:, :4] = oracle_lib.vertex_passthrough(scenes.CUBE_POSITIONS.astype(np.float32), m)
    scene = dataclasses.replace(base, vertices=np.ascontiguousarray(verts))
    ref = oracle_lib.render(scene)

    r = cudaraster.SceneRenderer(base)        # vertex buffer: colours, positions to be overwritten
    try:
        dp = torch.from_numpy(scenes.CUBE_POSITIONS.astype(np.float32)).cuda()
        r.cr.vertexShaderPassthrough(dp, m, r.vbuf, out_stride=32)
        r.draw()
        got = r.read()
        assert np.array_equal(r.vbuf.cpu().numpy().view(np.uint32).reshape(verts.shape), verts.view(np.uint32))
    finally:
        r.close()
    assert_parity(got, ref, "cube through the vertex shader")


def test_wrapper_rejects_undersized_buffers():
    """vertexShaderPassthrough checks its buffers before the library call:
    an output smaller than num_vertices records, fewer model positions than
    vertices, or a wrong dtype raise ValueError and launch nothing."""
    torch = _torch()
    import cudaraster
    cr = cudaraster.CudaRaster().init()
    try:
        m = np.eye(4, dtype=np.float32)
        pos = torch.zeros(300, dtype=torch.float32, device="cuda")      # 100 vertices
        out = torch.zeros(100 * 8, dtype=torch.float32, device="cuda")  # 100 records of 32 bytes
        cr.vertexShaderPassthrough(pos, m, out, out_stride=32)
        for kw in ({"out": out[:-1]}, {"num_vertices": 101}, {"model_pos": pos.double()},
                   {"model_pos": pos.cpu()}, {"out_stride": 48}):
            args = {"model_pos": pos, "out": out, "out_stride": 32, **kw}
            with pytest.raises(ValueError):
                cr.vertexShaderPassthrough(args["model_pos"], m, args["out"], out_stride=args["out_stride"],
                                           num_vertices=args.get("num_vertices"))
    finally:
        cr.close()
