"""The texPhong fragment shader and RenderModeFlag_EnableQuads on the GPU
against the oracle (test/shader/Shaders.cu:122-199; PixelPipe.hpp:34,
59-69; FineRaster.inl:21-44): depth and triangle ids bit-exact (the alpha
test's discards included), colour within 1 LSB per channel, across sample
counts, analytic and numerical derivatives, blend modes and flags."""
import numpy as np
import pytest

from test_gpu_parity import assert_parity, render_gpu

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("quads", [False, True])
@pytest.mark.parametrize("samples_log2", [0, 1, 2, 3])
def test_tex_phong_against_oracle(oracle_lib, samples_log2, quads):
    _torch()
    from cudaraster import scenes
    sc = scenes.tex_phong(6000, 256, 192, samples_log2=samples_log2, seed=70 + samples_log2, quads=quads)
    print(f"texPhong s={samples_log2} quads={quads}", flush=True)
    got = render_gpu(sc)
    ref = oracle_lib.render(sc)
    assert_parity(got, ref, f"texPhong s={samples_log2} quads={quads}")
    ids = got["triid"][got["triid"] != 0xFFFFFFFF]
    assert ids.size > 1000


@pytest.mark.parametrize("blend", [1, 2, 3])
def test_tex_phong_blend_modes(oracle_lib, blend):
    _torch()
    from cudaraster import scenes
    sc = scenes.tex_phong(4000, 200, 136, samples_log2=2, seed=80 + blend, quads=blend == 2, blend=blend)
    print(f"texPhong blend={blend}", flush=True)
    assert_parity(render_gpu(sc), oracle_lib.render(sc), f"texPhong blend {blend}")


@pytest.mark.parametrize("flags", [0, 1, 2, 6])
def test_tex_phong_render_flags(oracle_lib, flags):
    """Without depth test, without interpolation (varyings of vertex 2), and
    quads without interpolation."""
    _torch()
    from cudaraster import scenes
    sc = scenes.tex_phong(3000, 128, 128, samples_log2=1, seed=90 + flags, flags=flags)
    print(f"texPhong flags={flags}", flush=True)
    assert_parity(render_gpu(sc), oracle_lib.render(sc), f"texPhong flags {flags}")


def test_tex_phong_counting_draw(oracle_lib):
    """The counting kernel variant of the texPhong pass: same image, exact
    fragment and visible-subtriangle totals."""
    _torch()
    from cudaraster import scenes
    sc = scenes.tex_phong(5000, 192, 160, samples_log2=2, seed=95, quads=True)
    got = render_gpu(sc, debug={"count_fragments": True})
    ref = oracle_lib.render(sc)
    assert_parity(got, ref, "texPhong counting")
    assert got["counters"]["fragments"] == ref["stats"]["fragments"]
    assert got["counters"]["visible_subtris"] == ref["stats"]["numSubtris"]


def test_tex_phong_needs_constants():
    _torch()
    import cudaraster
    from cudaraster import scenes
    sc = scenes.tex_phong(100, 64, 64)
    r = cudaraster.SceneRenderer(sc)
    try:
        r.cr.setShaderConstants(None, None, None)
        with pytest.raises(cudaraster.CrError) as e:
            r.cr.drawTriangles()
        assert e.value.status == cudaraster.CR_ERR_STATE
        r.cr.setShaderConstants(*r.shader_constants)
        r.cr.drawTriangles()
    finally:
        r.close()
    # Batches draw the stock Gouraud / constant pipes only.
    import ctypes
    raster = cudaraster.CudaRaster().init()
    col = cudaraster.CudaSurface((64, 64), cudaraster.FORMAT_RGBA8, 1)
    dep = cudaraster.CudaSurface((64, 64), cudaraster.FORMAT_DEPTH32, 1)
    vb, ib = cudaraster.device_buffers(sc)
    v = cudaraster.View()
    v.color, v.depth, v.tri_id = col.handle.value, dep.handle.value, None
    v.vertices, v.num_vertices, v.indices, v.num_tris = vb.data_ptr(), sc.vertices.shape[0], ib.data_ptr(), sc.num_tris
    v.pipe = cudaraster.scene_pipe(sc)
    with pytest.raises(cudaraster.CrError) as e:
        cudaraster.Batch(raster, [v])
    assert e.value.status == cudaraster.CR_ERR_UNSUPPORTED
    raster.close()
    col.close()
    dep.close()
