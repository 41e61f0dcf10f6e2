"""Pins of the oracle's texPhong shader (test/shader/Shaders.cu:37-48,
122-199) and of RenderModeFlag_EnableQuads (PixelPipe.hpp:34, 59-69)
against what the mathematics fixes: float64 references for log2, exp2 and
pow; texel-centre, constant, linear and wrap identities of the bilinear
sampler; mip selection at exact powers of two; closed forms of the lighting;
and whole draws whose colours are known in advance (the mip level a texture
density selects, analytic and numerical derivatives alike; the alpha test's
discard)."""
import numpy as np
import pytest

from cudaraster import scenes


def ulps(got, ref):
    ref = np.float64(ref)
    return abs(np.float64(got) - ref) / np.float64(np.spacing(np.float32(abs(ref))))


def test_log2_against_float64(oracle_lib):
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.exp(rng.uniform(-85, 85, 4000)), rng.uniform(0.5, 2.0, 4000),
                         1.0 + rng.uniform(-1e-3, 1e-3, 2000), [1.4142135, 1.4142137, 0.70710677],
                         np.float32(2.0) ** -np.arange(127, 149)]).astype(np.float32)
    worst = max(ulps(oracle_lib.log2f(float(x)), np.log2(np.float64(x))) for x in xs)
    assert worst <= 4.0, worst
    for k in range(-149, 128):
        assert oracle_lib.log2f(2.0 ** k) == k
    assert oracle_lib.log2f(0.0) == -np.inf and oracle_lib.log2f(np.inf) == np.inf
    assert np.isnan(oracle_lib.log2f(-1.0)) and np.isnan(oracle_lib.log2f(np.nan))


def test_exp2_and_pow_against_float64(oracle_lib):
    rng = np.random.default_rng(1)
    ys = rng.uniform(-126, 127.9, 6000).astype(np.float32)
    worst = max(ulps(oracle_lib.exp2f(float(y)), 2.0 ** np.float64(y)) for y in ys)
    assert worst <= 1.0, worst
    for k in range(-149, 128):
        assert oracle_lib.exp2f(float(k)) == 2.0 ** k
    assert oracle_lib.exp2f(128.0) == np.inf and oracle_lib.exp2f(-151.0) == 0.0
    # pow(x, g) = exp2(g log2 x): relative error grows with |g log2 x|.
    for x, g in zip(rng.uniform(1e-3, 1.0, 3000), rng.uniform(0.1, 64.0, 3000)):
        x, g = np.float32(x), np.float32(g)
        ref = np.float64(x) ** np.float64(g)
        if ref < 1e-37:
            continue
        got = oracle_lib.powf(float(x), float(g))
        bound = 2e-7 * (2.0 + abs(np.float64(g) * np.log2(np.float64(x)))) * 1.5
        assert abs(got - ref) <= bound * ref, (x, g, got, ref)
    assert oracle_lib.powf(0.0, 5.0) == 0.0 and oracle_lib.powf(0.0, 0.0) == 1.0
    assert oracle_lib.powf(0.3, 0.0) == 1.0 and oracle_lib.powf(1.0, 37.0) == 1.0


def test_bilinear_sampler_identities(oracle_lib):
    rng = np.random.default_rng(2)
    W, H = 16, 8
    atlas = rng.random((H, W, 4), dtype=np.float32)
    # Texel centres return the texel (weights exactly 0 and 1).
    for j in range(H):
        for i in range(W):
            got = oracle_lib.tex_sample(atlas, (i + 0.5) / W, (j + 0.5) / H)
            assert np.array_equal(got, atlas[j, i])
    # Wrap addressing: whole periods give the same lookup.
    for u, v in rng.random((50, 2)):
        a = oracle_lib.tex_sample(atlas, u, v)
        b = oracle_lib.tex_sample(atlas, u + 3.0, v - 2.0)
        assert np.allclose(a, b, atol=2e-6)
    # Between centres: the float64 bilinear interpolation.
    for u, v in rng.random((200, 2)):
        x, y = u * W - 0.5, v * H - 0.5
        i0, j0 = int(np.floor(x)), int(np.floor(y))
        a, b = x - i0, y - j0
        t = lambda i, j: atlas[j % H, i % W].astype(np.float64)
        ref = (1 - b) * ((1 - a) * t(i0, j0) + a * t(i0 + 1, j0)) + b * ((1 - a) * t(i0, j0 + 1) + a * t(i0 + 1, j0 + 1))
        assert np.allclose(oracle_lib.tex_sample(atlas, u, v), ref, atol=1e-6)
    # A constant texture is constant everywhere, NaN coordinates included.
    c = np.full((H, W, 4), 0.375, np.float32)
    assert np.array_equal(oracle_lib.tex_sample(c, 0.123, 0.77), np.full(4, 0.375, np.float32))
    assert np.array_equal(oracle_lib.tex_sample(c, np.inf, 0.5)[:1], np.array([np.nan], np.float32)) or \
        np.isnan(oracle_lib.tex_sample(c, np.inf, 0.5)).all()


def level_atlas(size=64):
    """A 64x64 texture whose mip level l is the constant colour
    (l / 10, 0.5, 1 - l / 10, 1): level l is an 8-row band of a 16-wide
    atlas, and its TextureSpec maps the texture onto the band's middle four
    rows, so no bilinear footprint (wrapping included) leaves the band."""
    levels = int(np.log2(size)) + 1
    W, H = 16, 8 * 16
    atlas = np.zeros((H, W, 4), np.float32)
    spec = np.zeros((), scenes.TEXSPEC_DTYPE)
    spec["size"] = (size, size)
    for l in range(16):
        lv = min(l, levels - 1)
        atlas[8 * l:8 * l + 8] = (lv / 10.0, 0.5, 1.0 - lv / 10.0, 1.0)
        spec["mipLevels"][l] = (1.0, 4.0 / H, 0.0, (8 * l + 2) / H)
    return atlas, spec, levels


@pytest.mark.parametrize("k", [0, 1, 2, 3, 4, 5])
def test_lod_selects_the_mip_level(oracle_lib, k):
    atlas, spec, levels = level_atlas()
    size = float(spec["size"][0])
    d = 2.0 ** k / size
    rgba, lod = oracle_lib.texture2d(spec, atlas, (0.3, 0.6), (d, 0.0), (0.0, d * 0.5))
    assert lod == k
    assert np.allclose(rgba, (k / 10.0, 0.5, 1.0 - k / 10.0, 1.0), atol=1e-6)
    # Half way between two levels: the lerp of the two level colours.
    if k + 1 < levels:
        rgba, lod = oracle_lib.texture2d(spec, atlas, (0.3, 0.6), (d * 2 ** 0.5, 0.0), (0.0, 0.0))
        assert lod == pytest.approx(k + 0.5, abs=1e-6)
        assert np.allclose(rgba[0], (k + 0.5) / 10.0, atol=1e-5)
    # Minification below level 0 clamps to 0; the texture coordinate wraps.
    rgba, lod = oracle_lib.texture2d(spec, atlas, (5.3, -2.4), (d / 64, 0.0), (0.0, 0.0))
    assert lod == 0.0 and np.allclose(rgba[0], 0.0, atol=1e-6)


def test_lighting_closed_forms(oracle_lib):
    m = np.zeros((), scenes.MATERIAL_DTYPE)
    m["specularColor"] = (0.25, 0.5, 0.125)
    m["glossiness"] = 13.0
    dc = np.array([0.4, 0.6, 0.8], np.float32)
    # Normal facing the viewer: diffuse 1, reflection straight back, specular 1.
    out = oracle_lib.lighting((0.0, 0.0, -3.0), (0.0, 0.0, 2.0), m, dc)
    assert np.array_equal(out, dc + np.array([0.25, 0.5, 0.125], np.float32))
    # Grazing normal: diffuse 1/4, reflection along the view ray, no specular.
    out = oracle_lib.lighting((0.0, 0.0, -3.0), (1.0, 0.0, 0.0), m, dc)
    assert np.array_equal(out, dc * np.float32(0.25))
    # General case against float64 Phong.
    rng = np.random.default_rng(4)
    for _ in range(200):
        p, n = rng.normal(size=3), rng.normal(size=3)
        I, N = p / np.linalg.norm(p), n / np.linalg.norm(n)
        R = I - 2 * np.dot(I, N) * N
        ref = (max(-np.dot(I, N), 0) * 0.75 + 0.25) * dc + max(-np.dot(I, R), 0) ** 13.0 * m["specularColor"]
        assert np.allclose(oracle_lib.lighting(p, n, m, dc), ref, atol=2e-5)


def flat_scene(density_x, density_y, quads, samples_log2=0, alpha=1.0, lerp=True):
    """A 64x64 viewport covered by two triangles at w = 1 with lighting that
    returns the diffuse colour: camera position (0,0,-1), normal (0,0,1), no
    specular.  Texture coordinates change by density_x / size per pixel in
    x along u and density_y / size per pixel in y along v."""
    W = H = 64
    atlas, spec, levels = level_atlas()
    size = float(spec["size"][0])
    quad = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float64)
    px = (quad + 1) * 0.5 * np.array([W, H])
    uv = np.stack([px[:, 0] * density_x / size, px[:, 1] * density_y / size], axis=1) + 0.25
    verts = np.zeros((4, 16), np.float32)
    verts[:, 0:2] = quad
    verts[:, 2] = 0.5
    verts[:, 3] = 1.0
    verts[:, 4:8] = (0.0, 0.0, -1.0, 1.0)
    verts[:, 8:12] = (0.0, 0.0, 1.0, 0.0)
    verts[:, 12:14] = uv
    verts[:, 15] = 1.0
    mats = np.zeros(1, scenes.MATERIAL_DTYPE)
    mats["diffuseColor"] = (1.0, 1.0, 1.0, alpha)
    mats["glossiness"] = 4.0
    mats["diffuseTexture"][0] = spec
    flags = (scenes.FLAG_DEPTH | (scenes.FLAG_LERP if lerp else 0) | (scenes.FLAG_QUADS if quads else 0))
    return scenes.Scene("flat", W, H, samples_log2, flags, scenes.BLEND_REPLACE, scenes.SHADER_TEXPHONG,
                        verts, np.array([[0, 1, 2], [0, 2, 3]], np.int32), materials=mats,
                        tri_material=np.zeros(2, np.int32), atlas=atlas)


@pytest.mark.parametrize("quads", [False, True])
@pytest.mark.parametrize("kx,ky,s", [(1, 0, 0), (3, 1, 0), (0, 4, 2), (2, 2, 3)])
def test_whole_draw_colour_is_the_selected_level(oracle_lib, quads, kx, ky, s):
    """Texture density 2^k texels per pixel selects mip level k = max(kx,
    ky), with analytic derivatives and with numerical quad derivatives."""
    sc = flat_scene(2.0 ** kx, 2.0 ** ky, quads, samples_log2=s)
    r = oracle_lib.render(sc)
    k = max(kx, ky)
    want = oracle_lib.to_abgr(np.array([k / 10.0, 0.5, 1.0 - k / 10.0, 1.0], np.float32))
    col = r["color"].view(np.uint8).reshape(r["color"].shape + (4,)).astype(int)
    ref = np.array([want & 0xFF, (want >> 8) & 0xFF, (want >> 16) & 0xFF, want >> 24])
    assert np.abs(col - ref).max() <= 1
    assert (r["triid"] != 0xFFFFFFFF).all()


def test_alpha_test_discards(oracle_lib):
    """Alpha below 1/2 discards every fragment: no depth, colour or id is
    written; exactly 1/2 passes; without LERP every fragment reads vertex 2."""
    clear = oracle_lib.render(flat_scene(1.0, 1.0, False, alpha=0.49))
    assert (clear["triid"] == 0xFFFFFFFF).all()
    assert (clear["depth"] == oracle_lib.clear_depth(1.0)).all()
    assert (clear["color"] == oracle_lib.to_abgr_host((0.2, 0.4, 0.8, 1.0))).all()
    drawn = oracle_lib.render(flat_scene(1.0, 1.0, True, alpha=0.5))
    assert (drawn["triid"] != 0xFFFFFFFF).all()
    flat = oracle_lib.render(flat_scene(4.0, 1.0, True, lerp=False))
    assert len(np.unique(flat["color"])) == 1


def test_material_index_outside_the_table_discards(oracle_lib):
    sc = flat_scene(1.0, 1.0, False)
    sc.tri_material = np.array([0, 5], np.int32)
    r = oracle_lib.render(sc)
    ids = np.unique(r["triid"])
    assert 0 in ids and 1 not in ids
