"""MSAA resolve on the GPU (cr_surface_resolve) against the oracle, and the
whole demo frame -- model positions, vertex shader, 4x draw, resolve --
against the golden image."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("w,h,sl2", [(64, 64, 0), (203, 77, 1), (256, 192, 2), (2048, 2048, 2), (129, 33, 3)])
def test_resolve_matches_oracle(oracle_lib, w, h, sl2):
    _torch()
    import cudaraster
    surf = cudaraster.CudaSurface((w, h), cudaraster.FORMAT_RGBA8, 1 << sl2)
    try:
        c = np.random.default_rng(w + sl2).integers(0, 2 ** 32, size=surf.shape, dtype=np.uint64).astype(np.uint32)
        surf.upload(c)
        got = surf.resolve().cpu().numpy().view(np.uint32)
    finally:
        surf.close()
    assert np.array_equal(got, oracle_lib.resolve(c, w, h, sl2))


def test_resolve_rejects_non_colour_surfaces():
    _torch()
    import cudaraster
    surf = cudaraster.CudaSurface((16, 16), cudaraster.FORMAT_DEPTH32, 4)
    try:
        with pytest.raises(cudaraster.CrError):
            surf.resolve()
    finally:
        surf.close()


def test_cube_c0_frame_matches_stored_surfaces():
    """The C0 frame (the demo cube at 1024x768, 1x, clip positions from the
    GPU vertex shader) drawn on the GPU equals the stored colour, depth and
    triangle-id surfaces."""
    torch = _torch()
    import sys
    import cudaraster
    from cudaraster import scenes
    sys.path.insert(0, GOLDEN)
    import make_golden
    scene, m = make_golden.cube_scene(1024, 768, 0)
    gold = make_golden.load_cube_frame()
    base = scenes.cube(scene.width, scene.height, samples_log2=0)
    r = cudaraster.SceneRenderer(base)
    try:
        dp = torch.from_numpy(scenes.CUBE_POSITIONS.astype(np.float32)).cuda()
        r.cr.vertexShaderPassthrough(dp, m, r.vbuf, out_stride=32)
        r.draw()
        got = r.read()
    finally:
        r.close()
    for k in ("color", "depth", "triid"):
        assert np.array_equal(got[k], gold[k]), k


def test_cube_frame_matches_golden():
    torch = _torch()
    import sys
    import cudaraster
    from cudaraster import scenes
    sys.path.insert(0, GOLDEN)
    import make_golden
    scene, m = make_golden.cube_scene()
    base = scenes.cube(scene.width, scene.height, samples_log2=scene.samples_log2)
    r = cudaraster.SceneRenderer(base)              # colours in place; positions from the vertex shader
    try:
        dp = torch.from_numpy(scenes.CUBE_POSITIONS.astype(np.float32)).cuda()
        r.cr.vertexShaderPassthrough(dp, m, r.vbuf, out_stride=32)
        r.draw()
        got = r.color.resolve().cpu().numpy().view(np.uint32)
    finally:
        r.close()
    assert np.array_equal(got, cudaraster.load_png(os.path.join(GOLDEN, "cube_256x192_4x_resolved.png")))


def test_resolve_into_an_unaligned_buffer(oracle_lib):
    """Output rows that are not 16-byte aligned take the per-pixel stores."""
    torch = _torch()
    import cudaraster
    surf = cudaraster.CudaSurface((64, 8), cudaraster.FORMAT_RGBA8, 4)
    try:
        c = np.random.default_rng(5).integers(0, 2 ** 32, size=surf.shape, dtype=np.uint64).astype(np.uint32)
        surf.upload(c)
        buf = torch.zeros(64 * 8 + 1, dtype=torch.int32, device="cuda")
        surf.resolve(buf[1:].view(8, 64))
        got = buf.cpu().numpy().view(np.uint32)
    finally:
        surf.close()
    assert got[0] == 0
    assert np.array_equal(got[1:].reshape(8, 64), oracle_lib.resolve(c, 64, 8, 2))


@pytest.mark.parametrize("name", ["random_96x64_1x_resolved.png", "random_96x64_2x_resolved.png",
                                  "random_96x64_4x_resolved.png", "random_96x64_8x_resolved.png"])
def test_random_frames_match_goldens(name):
    _torch()
    import sys
    import cudaraster
    sys.path.insert(0, GOLDEN)
    import make_golden
    r = cudaraster.SceneRenderer(make_golden.random_scene(name))
    try:
        r.draw()
        got = r.color.resolve().cpu().numpy().view(np.uint32)
    finally:
        r.close()
    assert np.array_equal(got, cudaraster.load_png(os.path.join(GOLDEN, name)))


def test_resolve_rejects_wrong_outputs():
    """resolve(out=...) checks size, dtype and layout before the library
    call; the library refuses a surface that is not RGBA8."""
    torch = _torch()
    import cudaraster
    col = cudaraster.CudaSurface((40, 24), cudaraster.FORMAT_RGBA8, 4)
    dep = cudaraster.CudaSurface((40, 24), cudaraster.FORMAT_DEPTH32, 4)
    try:
        col.resolve(torch.empty((24, 40), dtype=torch.int32, device="cuda"))
        for out in (torch.empty((24, 39), dtype=torch.int32, device="cuda"),
                    torch.empty((24, 40), dtype=torch.float64, device="cuda"),
                    torch.empty((40, 24), dtype=torch.int32, device="cuda").t(),
                    torch.empty((24, 40), dtype=torch.int32)):
            with pytest.raises(ValueError):
                col.resolve(out)
        with pytest.raises(cudaraster.CrError):
            dep.resolve()
    finally:
        col.close()
        dep.close()
