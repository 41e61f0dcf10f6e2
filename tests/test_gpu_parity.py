"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Coverage, depth and triangle ids must be bit-exact; colour within 1 LSB per
channel (the contract; in practice the two are bit-identical because both
evaluate the same operation sequence).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def render_gpu(scene, **kw):
    import cudaraster
    r = cudaraster.SceneRenderer(scene, debug=kw.pop("debug", None))
    r.draw(**kw)
    out = r.read()
    out["counters"] = r.cr.getCounters()
    out["stats"] = r.cr.getStats()
    r.close()
    return out


def assert_parity(got, ref, name):
    for k in ("depth", "triid"):
        if not np.array_equal(got[k], ref[k]):
            bad = np.argwhere(got[k] != ref[k])
            y, x = bad[0]
            raise AssertionError(f"{name}: {k} differs at {len(bad)} texels, first (row {y}, col {x}): "
                                 f"gpu {got[k][y, x]:#x} oracle {ref[k][y, x]:#x}")
    d = np.abs(got["color"].view(np.uint8).astype(np.int16) - ref["color"].view(np.uint8).astype(np.int16))
    assert d.max() <= 1, f"{name}: colour differs by {d.max()} LSB"


def test_cube_exact(oracle_lib):
    _torch()
    from cudaraster import scenes
    s = scenes.cube()
    got = render_gpu(s)
    ref = oracle_lib.render(s)
    assert_parity(got, ref, "cube")
    assert np.array_equal(got["color"], ref["color"])


@pytest.mark.parametrize("samples_log2", [0, 1, 2, 3])
def test_random_tris_all_sample_counts(oracle_lib, samples_log2):
    _torch()
    from cudaraster import scenes
    s = scenes.random_tris(20000, 320, 256, samples_log2=samples_log2, seed=11 + samples_log2)
    got = render_gpu(s, debug={"count_fragments": True})
    ref = oracle_lib.render(s)
    assert_parity(got, ref, f"random s={samples_log2}")
    assert got["counters"]["fragments"] == ref["stats"]["fragments"]
    assert got["counters"]["visible_subtris"] == ref["stats"]["numSubtris"]


@pytest.mark.parametrize("blend", [0, 1, 2, 3])
def test_blend_modes(oracle_lib, blend):
    _torch()
    from cudaraster import scenes
    s = scenes.random_tris(8000, 256, 256, samples_log2=2, seed=21 + blend, alpha_range=(0.2, 0.8),
                           area_range=(4.0, 4000.0), blend=blend)
    assert_parity(render_gpu(s), oracle_lib.render(s), f"blend {blend}")


@pytest.mark.parametrize("blend,lerp", [(0, True), (0, False), (1, True), (1, False)])
def test_extreme_colours(oracle_lib, blend, lerp):
    """Vertex colours that exercise every branch of toABGR's saturation:
    NaN, +-inf, negative, above 1, huge, and values a few ulp either side of
    the k/255 rounding boundaries -- through deferred shading (replace) and
    in-loop shading (src-over), interpolated or flat (without LERP the
    fragment takes the third vertex's colour).  Colours must match the oracle
    exactly."""
    _torch()
    from cudaraster import scenes
    flags = scenes.FLAG_DEPTH | (scenes.FLAG_LERP if lerp else 0)
    s = scenes.random_tris(5000, 192, 128, samples_log2=2, seed=91 + blend, area_range=(2.0, 1500.0),
                           blend=blend, flags=flags)
    rng = np.random.default_rng(5 + 2 * blend + lerp)
    special = np.array([np.nan, np.inf, -np.inf, -0.5, -0.0, 1.5, 1e30, -1e30, 0.0, 1.0], np.float32)
    k = rng.integers(0, 256, size=4096)
    boundary = ((k + 0.5) / 255.0).astype(np.float32)
    near = np.nextafter(boundary, np.where(rng.random(4096) < 0.5, np.float32(0), np.float32(2)))
    cols = s.vertices[:, 4:8].copy()
    u = rng.random(cols.shape)
    cols = np.where(u < 0.1, rng.choice(special, size=cols.shape), cols)
    cols = np.where((u >= 0.1) & (u < 0.3), rng.choice(boundary, size=cols.shape), cols)
    cols = np.where((u >= 0.3) & (u < 0.5), rng.choice(near, size=cols.shape), cols)
    s.vertices[:, 4:8] = cols.astype(np.float32)
    got, ref = render_gpu(s), oracle_lib.render(s)
    c = ref["color"].view(np.uint8)
    assert (c == 0).sum() > 20 and (c == 255).sum() > 20, "workload must reach both saturation ends"
    assert_parity(got, ref, f"extreme colours, blend {blend}, lerp {lerp}")
    assert np.array_equal(got["color"], ref["color"])


def pathological_scene(seed=77):
    """Random triangles with hostile clip positions: NaN, +-inf and huge
    components, w = 0, vertices behind the eye, vertices on the far plane,
    and degenerate triangles (a repeated vertex, three collinear vertices)."""
    from cudaraster import scenes
    s = scenes.random_tris(4000, 160, 120, samples_log2=2, seed=seed, area_range=(1.0, 4000.0))
    v = s.vertices.copy()
    rng = np.random.default_rng(seed)
    pick = rng.random(v.shape[0])
    specials = np.array([np.nan, np.inf, -np.inf, 1e30, -1e30, 0.0], np.float32)
    idx = np.where(pick < 0.03)[0]
    v[idx, rng.integers(0, 4, len(idx))] = rng.choice(specials, len(idx))
    idx = np.where((pick >= 0.03) & (pick < 0.05))[0]
    v[idx, 3] = 0.0
    idx = np.where((pick >= 0.05) & (pick < 0.07))[0]
    v[idx, 3] *= -1
    idx = np.where((pick >= 0.07) & (pick < 0.09))[0]
    v[idx, 2] = v[idx, 3]
    tri = v.reshape(-1, 3, 8)
    k = rng.choice(len(tri), 200, replace=False)
    tri[k, 2, :4] = tri[k, 1, :4]
    k = rng.choice(len(tri), 200, replace=False)
    tri[k, 2, :4] = 0.5 * (tri[k, 0, :4] + tri[k, 1, :4])
    s.vertices[:] = tri.reshape(-1, 8)
    return s


def test_pathological_vertices(oracle_lib):
    _torch()
    s = pathological_scene()
    ref = oracle_lib.render(s)
    assert ref["stats"]["clipped"] > 100 and ref["stats"]["numSubtris"] > 1000
    got = render_gpu(s, debug={"count_fragments": True})
    assert_parity(got, ref, "pathological vertices")
    assert got["counters"]["fragments"] == ref["stats"]["fragments"]


def quadrant_scene(samples_log2, flags, blend=0, seed=0):
    """Deferred-shading draws whose triangles sit on the tiles' 4x4-pixel
    quadrant boundaries: tiny triangles centred on pixel coordinates that are
    multiples of 4 (jittered by 0, +-1/16 or +-1/2 pixel), right triangles
    with both legs exactly on such lines, and larger triangles spanning
    several quadrants and tiles, in random order.  Without the depth test the
    last triangle of a sample wins, so every pixel's submission order is
    visible in the image."""
    from cudaraster import scenes
    rng = np.random.default_rng(seed)
    width, height, n = 96, 64, 3000
    kind = rng.integers(0, 3, n)
    cx = rng.integers(0, width // 4 + 1, n) * 4.0 + rng.choice([0.0, 1 / 16, -1 / 16, 0.5, -0.5], n)
    cy = rng.integers(0, height // 4 + 1, n) * 4.0 + rng.choice([0.0, 1 / 16, -1 / 16, 0.5, -0.5], n)
    area = np.where(kind == 0, 0.3 + 5.7 * rng.random(n), 20.0 + 180.0 * rng.random(n))
    sx, sy = scenes._triangle_shapes(rng, n, area)
    px, py = cx[:, None] + sx, cy[:, None] + sy
    # Right triangles: the corner on a boundary crossing, legs along x = 4i and y = 4j.
    legs = rng.choice([-1.0, 1.0], (n, 2)) * rng.integers(1, 9, (n, 2))
    r = kind == 1
    px[r] = np.stack([cx[r], cx[r] + legs[r, 0], cx[r]], axis=1).round()
    py[r] = np.stack([cy[r], cy[r], cy[r] + legs[r, 1]], axis=1).round()
    w = np.repeat(1.0 + 9.0 * rng.random((n, 1)), 3, axis=1)
    clip = scenes._to_clip(px, py, w, width, height)
    col = rng.random((n, 3, 4))
    verts = np.concatenate([clip, col], axis=-1).reshape(n * 3, 8).astype(np.float32)
    idx = np.arange(n * 3, dtype=np.int32).reshape(n, 3)       # about half face away (random winding)
    return scenes.Scene(f"quadrants s={samples_log2} f={flags} b={blend}", width, height, samples_log2, flags,
                        blend, scenes.SHADER_GOURAUD, np.ascontiguousarray(verts), idx)


@pytest.mark.parametrize("samples_log2", [0, 1, 2, 3])
@pytest.mark.parametrize("depth", [True, False])
def test_quadrant_boundaries(oracle_lib, samples_log2, depth):
    """The deferred-shading walk gives each 4x4 quadrant its own record queue;
    a record reaching a quadrant only across its boundary, or an edge on the
    boundary, must still be walked there, in submission order."""
    _torch()
    from cudaraster import scenes
    flags = scenes.FLAG_LERP | (scenes.FLAG_DEPTH if depth else 0)
    s = quadrant_scene(samples_log2, flags, seed=40 + samples_log2 + (4 if depth else 0))
    ref = oracle_lib.render(s)
    assert ref["stats"]["numSubtris"] > 1000
    assert_parity(render_gpu(s), ref, s.name)


def test_quadrant_boundaries_depth_only(oracle_lib):
    _torch()
    from cudaraster import scenes
    s = quadrant_scene(2, scenes.FLAG_DEPTH | scenes.FLAG_LERP, blend=scenes.BLEND_DEPTHONLY, seed=9)
    assert_parity(render_gpu(s), oracle_lib.render(s), s.name)


def deep_list_scene(samples_log2, flags, blend, seed=0):
    """Tiny triangles packed into a 32-tile region (4 x 8 tiles) of a 256 x 256
    view: every tile's list grows by about one 32-entry segment per coarse
    round (1024 bin entries), so its lists are long chains whose links to the
    segment after next are filled in a round or two later."""
    from cudaraster import scenes
    rng = np.random.default_rng(seed)
    width = height = 256
    n = 20000
    cx = 96.0 + 32.0 * rng.random(n)
    cy = 64.0 + 64.0 * rng.random(n)
    sx, sy = scenes._triangle_shapes(rng, n, 0.4 + 2.6 * rng.random(n))
    w = np.repeat(1.0 + 9.0 * rng.random((n, 1)), 3, axis=1)
    clip = scenes._to_clip(cx[:, None] + sx, cy[:, None] + sy, w, width, height)
    col = rng.random((n, 3, 4))
    col[..., 3] = 0.2 + 0.6 * rng.random((n, 3))
    verts = np.concatenate([clip, col], axis=-1).reshape(n * 3, 8).astype(np.float32)
    idx = np.arange(n * 3, dtype=np.int32).reshape(n, 3)
    return scenes.Scene(f"deep lists s={samples_log2} f={flags} b={blend}", width, height, samples_log2, flags,
                        blend, scenes.SHADER_GOURAUD, np.ascontiguousarray(verts), idx)


@pytest.mark.parametrize("samples_log2,depth,blend", [(3, False, 1), (2, True, 1), (0, False, 2), (2, False, 0)])
def test_deep_tile_lists(oracle_lib, samples_log2, depth, blend):
    """Long tile lists built one segment per round: the in-order copies
    (src-over, additive) read them two segments per step through the links
    to the segment after next; the deferred-shading copy (replace) one."""
    _torch()
    from cudaraster import scenes
    flags = scenes.FLAG_LERP | (scenes.FLAG_DEPTH if depth else 0)
    s = deep_list_scene(samples_log2, flags, blend, seed=60 + samples_log2)
    ref = oracle_lib.render(s)
    got = render_gpu(s)
    assert got["counters"]["tile_list_entries"] > 8 * 32 * 32      # > 8 segments per tile on average
    assert_parity(got, ref, s.name)


def test_zero_triangles(oracle_lib):
    """A draw without triangles still applies the deferred clear."""
    _torch()
    import dataclasses
    from cudaraster import scenes
    s = scenes.random_tris(10, 72, 40, samples_log2=1, seed=2)
    s = dataclasses.replace(s, indices=s.indices[:0])
    got = render_gpu(s)
    assert_parity(got, oracle_lib.render(s), "zero triangles")
    assert got["counters"]["num_tris"] == 0 and got["counters"]["tile_entries"] == 0


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_render_mode_flags(oracle_lib, flags):
    _torch()
    from cudaraster import scenes
    s = scenes.random_tris(6000, 200, 136, samples_log2=1, seed=31 + flags, flags=flags,
                           area_range=(1.0, 2000.0), alpha_range=(0.2, 0.8), blend=1)
    assert_parity(render_gpu(s), oracle_lib.render(s), f"flags {flags}")


def test_skewed_small(oracle_lib):
    _torch()
    from cudaraster import scenes
    s = scenes.skewed(60000, 48, 256, 256, samples_log2=3, seed=5)
    got = render_gpu(s)
    ref = oracle_lib.render(s)
    assert_parity(got, ref, "skewed")


@pytest.mark.parametrize("samples_log2", [0, 2, 3])
def test_clip_path(oracle_lib, samples_log2):
    _torch()
    from cudaraster import scenes
    s = scenes.clip_mix(4000, 192, 160, samples_log2=samples_log2, seed=71 + samples_log2)
    got = render_gpu(s, debug={"count_fragments": True})
    ref = oracle_lib.render(s)
    assert ref["stats"]["clipped"] > 100 and ref["stats"]["culledFrustum"] > 100
    assert_parity(got, ref, "clip")
    assert got["counters"]["fragments"] == ref["stats"]["fragments"]
    assert got["counters"]["visible_subtris"] == ref["stats"]["numSubtris"]
    assert got["counters"]["num_subtris"] > s.num_tris      # clipped fans took extra slots


def test_odd_viewport_and_no_clear(oracle_lib):
    _torch()
    import cudaraster
    from cudaraster import scenes
    s = scenes.random_tris(3000, 203, 77, samples_log2=2, seed=41, area_range=(1.0, 3000.0),
                           alpha_range=(0.2, 0.8), blend=1)
    r = cudaraster.SceneRenderer(s)
    r.draw()                          # deferred clear + draw
    s2 = scenes.random_tris(3000, 203, 77, samples_log2=2, seed=42, area_range=(1.0, 3000.0),
                            alpha_range=(0.2, 0.8), blend=1)
    import torch
    r.vbuf2 = torch.from_numpy(s2.vertices).cuda()
    r.cr.setVertexBuffer(r.vbuf2, 0, s2.vertices.shape[0])
    r.cr.drawTriangles()              # second draw without clear: accumulates
    got = r.read()
    r.close()
    ref1 = oracle_lib.render(s)
    ref = oracle_lib.render(s2, color=ref1["color"], depth=ref1["depth"], triid=ref1["triid"],
                            deferred_clear=False)
    assert_parity(got, ref, "two draws")


def test_overflow_retry(oracle_lib):
    """Many large triangles overflow the initial tile-queue capacity; the
    draw grows its buffers and re-runs (CudaRaster.cpp:277-339)."""
    _torch()
    from cudaraster import scenes
    s = scenes.random_tris(3000, 1024, 1024, samples_log2=0, seed=51, area_range=(2e4, 4e5))
    got = render_gpu(s)
    assert got["counters"]["retries"] >= 1
    assert_parity(got, oracle_lib.render(s), "overflow")


def test_subtriangle_capacity_limit():
    """Subtriangle slots are capped at 2^28 (the fine stage packs slot << 4
    with a centroid code into 32 bits): a larger reservation is refused
    with an error instead of wrapping."""
    _torch()
    import cudaraster
    cr = cudaraster.CudaRaster().init()
    try:
        cr.reserve(1 << 28, 1, 1)
        with pytest.raises(cudaraster.CrError, match="2\\^28"):
            cr.reserve((1 << 28) + 1, 1, 1)
    finally:
        cr.close()


def test_hiz_does_not_change_image(oracle_lib):
    _torch()
    from cudaraster import scenes
    s = scenes.random_tris(30000, 256, 256, samples_log2=2, seed=61, area_range=(1.0, 600.0))
    a = render_gpu(s)
    b = render_gpu(s, debug={"disable_hiz": True})
    for k in ("color", "depth", "triid"):
        assert np.array_equal(a[k], b[k]), k


def test_full_c1_against_oracle(oracle_lib):
    """1M triangles, 2048^2, 4x MSAA: full-size parity with the oracle."""
    _torch()
    from cudaraster import scenes
    s = scenes.config(1)
    got = render_gpu(s)
    ref = oracle_lib.render(s)
    assert_parity(got, ref, "c1 full")


def test_full_c2_against_oracle(oracle_lib):
    """4M sub-pixel triangles + 1024 near-full-screen ones, 2048^2, 8x MSAA,
    SrcOver: full-size parity with the oracle (queue growth, ordered merges
    of the large triangles' long tile lists, blending in submission order)."""
    _torch()
    from cudaraster import scenes
    s = scenes.config(2)
    got = render_gpu(s)
    ref = oracle_lib.render(s)
    assert_parity(got, ref, "c2 full")
