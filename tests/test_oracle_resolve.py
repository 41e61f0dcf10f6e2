"""The oracle's box MSAA resolve, the PNG store format and the golden image.

The resolve is pinned by closed forms (a single-sample surface is a crop,
equal samples resolve to themselves, hand-computed roundings) and by a
per-texel brute force that addresses samples with the surface layout's own
formula (texel = y * textureW + (x >> 3) * 8N + i * 8 + (x & 7)); the golden
cube image pins the whole oracle frame (vertex shader, draw, resolve)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _surface(rng, width, height, n):
    rw, rh = (width + 7) // 8 * 8, (height + 7) // 8 * 8
    return rng.integers(0, 2 ** 32, size=(rh, rw * n), dtype=np.uint64).astype(np.uint32)


def _brute(color, width, height, n):
    tw = color.shape[1]
    flat = color.reshape(-1)
    out = np.zeros((height, width), np.uint32)
    for y in range(height):
        for x in range(width):
            word = 0
            for ch in range(4):
                s = sum((int(flat[y * tw + (x >> 3) * 8 * n + i * 8 + (x & 7)]) >> (8 * ch)) & 0xFF for i in range(n))
                word |= ((s + n // 2) // n) << (8 * ch)
            out[y, x] = word
    return out


def test_single_sample_is_a_crop(oracle_lib):
    c = _surface(np.random.default_rng(1), 21, 13, 1)
    assert np.array_equal(oracle_lib.resolve(c, 21, 13, 0), c[:13, :21])


@pytest.mark.parametrize("sl2", [1, 2, 3])
def test_equal_samples_resolve_to_themselves(oracle_lib, sl2):
    n = 1 << sl2
    rng = np.random.default_rng(2)
    px = rng.integers(0, 2 ** 32, size=(16, 16), dtype=np.uint64).astype(np.uint32)
    c = np.repeat(px.reshape(16, 2, 1, 8), n, axis=2).reshape(16, 16 * n)   # every sample = its pixel
    assert np.array_equal(oracle_lib.resolve(c, 16, 16, sl2), px)


def test_hand_computed_rounding(oracle_lib):
    # 4x, pixel (0, 0): samples of the red channel 0, 255, 255, 255 -> (765 + 2) / 4 = 191;
    # green 1, 2, 0, 0 -> (3 + 2) / 4 = 1; blue 2, 2, 2, 0 -> (6 + 2) / 4 = 2 (half up).
    c = np.zeros((8, 32), np.uint32)
    for i, (r, g, b) in enumerate([(0, 1, 2), (255, 2, 2), (255, 0, 2), (255, 0, 0)]):
        c[0, i * 8] = r | (g << 8) | (b << 16) | (255 << 24)
    assert oracle_lib.resolve(c, 8, 8, 2)[0, 0] == 191 | (1 << 8) | (2 << 16) | (255 << 24)


@pytest.mark.parametrize("w,h,sl2", [(8, 8, 2), (13, 5, 1), (21, 11, 3), (9, 17, 2)])
def test_against_per_texel_brute_force(oracle_lib, w, h, sl2):
    c = _surface(np.random.default_rng(3 + sl2), w, h, 1 << sl2)
    assert np.array_equal(oracle_lib.resolve(c, w, h, sl2), _brute(c, w, h, 1 << sl2))


def test_png_round_trip(tmp_path):
    import cudaraster
    w = np.random.default_rng(4).integers(0, 2 ** 32, size=(7, 5), dtype=np.uint64).astype(np.uint32)
    cudaraster.save_png(tmp_path / "x.png", w)
    assert np.array_equal(cudaraster.load_png(tmp_path / "x.png"), w)


def test_oracle_reproduces_golden_cube():
    import sys
    import cudaraster
    sys.path.insert(0, GOLDEN)
    import make_golden
    got = make_golden.cube_golden()
    assert np.array_equal(got, cudaraster.load_png(os.path.join(GOLDEN, "cube_256x192_4x_resolved.png")))


@pytest.mark.parametrize("name", ["random_96x64_1x_resolved.png", "random_96x64_2x_resolved.png",
                                  "random_96x64_4x_resolved.png", "random_96x64_8x_resolved.png"])
def test_oracle_reproduces_random_goldens(name):
    import sys
    import cudaraster
    sys.path.insert(0, GOLDEN)
    import make_golden
    assert np.array_equal(make_golden.random_golden(name), cudaraster.load_png(os.path.join(GOLDEN, name)))


def test_cube_frame_surfaces_golden(oracle_lib):
    """The C0 frame (the demo cube at 1024x768, 1x) as stored: the oracle
    draws exactly the stored colour, depth and triangle-id surfaces; from
    the demo camera two faces (four of the cube's 12 triangles) are
    visible."""
    import sys
    sys.path.insert(0, GOLDEN)
    import make_golden
    _, surf = make_golden.cube_frame()
    gold = make_golden.load_cube_frame()
    for k in ("color", "depth", "triid"):
        assert gold[k].shape == (768, 1024) and np.array_equal(surf[k], gold[k]), k
    ids = set(np.unique(gold["triid"]).tolist())
    assert ids <= set(range(12)) | {0xFFFFFFFF} and len(ids - {0xFFFFFFFF}) == 4
