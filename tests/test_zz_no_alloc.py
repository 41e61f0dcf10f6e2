"""cr_draw_triangles_noalloc against cr_draw_triangles, through the C ABI.

The no-allocation draw keeps its intermediates in the library's fixed
per-device scratch array; its frames must be bit-identical to the allocating
draw's on the same state, on consecutive calls that share the array (the
second with a smaller input, so stale contents of the first stay behind it),
including a draw whose queues overflow the first division of the array and
are re-divided, and draws too large for the array are refused.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _frames(r):
    out = r.read()
    return out["color"], out["depth"], out["triid"]


def _draw_both(scene):
    """Draws `scene` with the allocating entry point, then with the
    no-allocation one on the same renderer (deferred clear both times, so
    both start from the same surfaces); returns both frames and the second
    draw's counters."""
    import cudaraster
    r = cudaraster.SceneRenderer(scene)
    try:
        r.cr.deferredClear(scene.clear_color, scene.clear_depth)
        r.cr.drawTriangles()
        ref = _frames(r)
        r.color.upload(np.full(r.color.shape, 0x12345678, np.uint32))
        r.depth.upload(np.zeros(r.depth.shape, np.uint32))
        r.cr.deferredClear(scene.clear_color, scene.clear_depth)
        r.cr.drawTrianglesNoAlloc()
        got = _frames(r)
        counters = r.cr.getCounters()
    finally:
        r.close()
    return ref, got, counters


def _assert_same(ref, got, name):
    for k, a, b in zip(("color", "depth", "triid"), ref, got):
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"{name}: {k} differs at {len(bad)} texels, first {tuple(bad[0])}")


def test_noalloc_matches_draw_twice_smaller_second():
    _torch()
    from cudaraster import scenes
    first = scenes.random_tris(200_000, 1024, 768, samples_log2=2, seed=41)
    second = scenes.random_tris(3_000, 200, 120, samples_log2=3, seed=42, alpha_range=(0.2, 0.8),
                                blend=scenes.BLEND_SRCOVER, area_range=(2.0, 2000.0))
    for name, sc in (("first (200k, 1024x768, 4x)", first), ("second (3k, 200x120, 8x src-over)", second)):
        print(f"no-alloc: {name}", flush=True)
        ref, got, _ = _draw_both(sc)
        _assert_same(ref, got, name)


def test_noalloc_regrows_queues_inside_the_array():
    """900 near-full-screen triangles at 1024^2 need about 9.5M tile-list
    entries (38 MB), more than the array's first division gives the tile
    lists (about 7.5M) and less than the whole array: the draw must re-divide
    the array by the exact demand and still match."""
    _torch()
    from cudaraster import scenes
    sc = scenes.skewed(num_small=2_000, num_big=900, width=1024, height=1024, samples_log2=0, seed=43)
    print("no-alloc: skewed 2k + 900 big, 1024^2", flush=True)
    ref, got, counters = _draw_both(sc)
    _assert_same(ref, got, "skewed")
    assert counters["retries"] >= 1


def test_noalloc_refuses_draws_beyond_the_array():
    """Refused on the host before any launch: 4M triangles need 4M slots of
    80 bytes, more than CR_NOALLOC_SCRATCH_BYTES."""
    torch = _torch()
    import cudaraster
    from cudaraster import scenes
    L = cudaraster.load()
    sc = scenes.random_tris(16, 64, 64, samples_log2=0, seed=44)
    r = cudaraster.SceneRenderer(sc)
    try:
        r.cr.setIndexBuffer(r.ibuf, 0, 4_000_000)
        st = L.cr_draw_triangles_noalloc(r.cr.ctx, None)
        assert st == cudaraster.CR_ERR_INVALID
        assert b"scratch array" in L.cr_last_error(r.cr.ctx)
        # The context still draws normally afterwards.
        r.cr.setIndexBuffer(r.ibuf, 0, sc.num_tris)
        r.cr.deferredClear(sc.clear_color, sc.clear_depth)
        r.cr.drawTrianglesNoAlloc()
        torch.cuda.synchronize()
    finally:
        r.close()


def test_noalloc_counters_match():
    """The no-allocation draw reports the same per-draw counters (slots, bin
    and tile entries) as the allocating draw of the same state."""
    _torch()
    import cudaraster
    from cudaraster import scenes
    sc = scenes.random_tris(20_000, 320, 256, samples_log2=1, seed=45)
    r = cudaraster.SceneRenderer(sc)
    try:
        r.cr.deferredClear(sc.clear_color, sc.clear_depth)
        r.cr.drawTriangles()
        a = r.cr.getCounters()
        r.cr.deferredClear(sc.clear_color, sc.clear_depth)
        r.cr.drawTrianglesNoAlloc()
        b = r.cr.getCounters()
    finally:
        r.close()
    for k in ("num_tris", "num_subtris", "bin_entries"):
        assert a[k] == b[k], k
    # Tile lists are sized from bounding boxes: the reserved room equals too.
    assert a["tile_entries"] == b["tile_entries"]
