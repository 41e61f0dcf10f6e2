"""Pins the CPU oracle's rasterization rules against independent derivations.

The reference holds no tests, golden images or known-answer vectors for this
path, and its host emulation cannot be compiled here (it needs cuda.h and
GL).  Every rule the oracle restates is therefore checked against something
that does not share its code: exact rational arithmetic, published standard
tables, brute force, textbook definitions written independently, and
invariants that a dropped term or a flipped sign would break.
"""
import ctypes
import math
import random
from fractions import Fraction

import numpy as np
import pytest

F32_MAX = float(np.finfo(np.float32).max)


def f32(x):
    return float(np.float32(x))


# ---------------------------------------------------------------------------
# Quantisation: toABGR (Util.inl:30-38) and Vec4f::toABGR (Math.cpp:41-48)
# ---------------------------------------------------------------------------

def round_down_to_f32(q: Fraction) -> float:
    """Largest float32 <= q (exact), as fma.rm rounds."""
    x = np.float32(float(q))           # nearest double, then nearest float32
    while Fraction(float(x)) > q:
        x = np.nextafter(x, np.float32(-np.inf))
    while Fraction(float(np.nextafter(x, np.float32(np.inf)))) <= q and np.isfinite(np.nextafter(x, np.float32(np.inf))):
        x = np.nextafter(x, np.float32(np.inf))
    return float(x)


def ptx_channel(c: float) -> int:
    """The device instruction sequence of toABGR, restated bit by bit:
    fma.rm.f32(c, 255*2^24, 2^23) -> cvt.rmi.sat.u32.f32 -> byte 3."""
    if math.isnan(c):
        return 0
    if math.isinf(c):
        return 255 if c > 0 else 0
    exact = Fraction(c) * (255 * 2 ** 24) + 2 ** 23
    r = round_down_to_f32(exact) if abs(exact) < Fraction(F32_MAX) else (F32_MAX if exact > 0 else -F32_MAX)
    u = math.floor(r)
    u = 0 if u < 0 else min(u, 2 ** 32 - 1)
    return u >> 24


def probe_values(rng, n):
    vals = [0.0, -0.0, 1.0, -1.0, 0.5, 1.5, 2.0, 1e-30, -1e-30, 1e30, -1e30, float("inf"), -float("inf"),
            float("nan"), 255.5 / 255, 1.0 / 255, 0.5 / 255]
    for k in range(256):                      # around every rounding boundary (k+1/2)/255
        b = np.float32((k + 0.5) / 255.0)
        for _ in range(3):
            vals.append(float(b))
            b = np.nextafter(b, np.float32(np.inf))
        b = np.float32((k + 0.5) / 255.0)
        for _ in range(3):
            b = np.nextafter(b, np.float32(-np.inf))
            vals.append(float(b))
    vals += [f32(rng.uniform(-0.2, 1.2)) for _ in range(n)]
    vals += [float(np.frombuffer(np.uint32(rng.getrandbits(32)).tobytes(), np.float32)[0]) for _ in range(n)]
    return [f32(v) for v in vals]


def test_to_abgr_matches_ptx_sequence(oracle_lib):
    rng = random.Random(1)
    for c in probe_values(rng, 1500):
        v = oracle_lib.to_abgr([c, 0.0, 0.0, 0.0]) & 0xFF
        assert v == ptx_channel(c), c


def test_to_abgr_closed_form_and_packing(oracle_lib):
    # Channel = clamp(floor(255c + 1/2)) of the exact product; R in the low byte.
    rng = random.Random(2)
    for _ in range(2000):
        c = [f32(rng.uniform(-0.5, 1.5)) for _ in range(4)]
        want = 0
        for i, x in enumerate(c):
            q = math.floor(Fraction(x) * 255 + Fraction(1, 2))
            want |= max(0, min(255, q)) << (8 * i)
        assert oracle_lib.to_abgr(c) == want


def test_to_abgr_host_rounds_half_up(oracle_lib):
    rng = random.Random(3)
    for c in [0.0, 1.0, 0.2, 0.4, 0.8, 0.5 / 255, 1.5 / 255] + [rng.random() for _ in range(3000)]:
        c = f32(c)
        q = math.floor(Fraction(c) * 255 + Fraction(1, 2))
        assert oracle_lib.to_abgr_host([c, c, c, c]) == q * 0x01010101
    # The demo's clear colour (test/SceneCR.cpp:287).
    assert oracle_lib.to_abgr_host([0.2, 0.4, 0.8, 1.0]) == 0xFFCC6633


# ---------------------------------------------------------------------------
# Blending (Util.inl:42-60, PixelPipe.inl:73-85)
# ---------------------------------------------------------------------------

def test_blend_divide_by_255_identity_exhaustive():
    # ((x * 0x010101 + 0x800000) >> 24) is round-to-nearest of x / 255 for
    # every x the blend can form (0 .. 255*255); x/255 is never a tie.
    x = np.arange(255 * 255 + 1, dtype=np.uint64)
    lhs = (x * 0x010101 + 0x800000) >> 24
    rhs = (2 * x + 255) // 510        # floor(x/255 + 1/2)
    assert np.array_equal(lhs, rhs)


def test_blend_modes_against_exact_rounding(oracle_lib):
    rng = random.Random(4)
    cases = [(rng.getrandbits(32), rng.getrandbits(32)) for _ in range(20000)]
    cases += [((a << 24) | (s * 0x010101), d * 0x01010101) for a in range(0, 256, 5) for s in (0, 1, 128, 255)
              for d in (0, 77, 255)]
    for src, dst in cases:
        a = src >> 24
        so = ad = 0
        for ch in range(4):
            s, d = (src >> 8 * ch) & 255, (dst >> 8 * ch) & 255
            so |= round(Fraction(s * a + d * (255 - a), 255)) << 8 * ch
            ad |= round(Fraction(min(s * 255 + d * 255, 255 * 255), 255)) << 8 * ch
        assert oracle_lib.blend(1, src, dst) == so
        assert oracle_lib.blend(2, src, dst) == ad
        assert oracle_lib.blend(0, src, dst) == src
        assert oracle_lib.blend(3, src, dst) == dst


# ---------------------------------------------------------------------------
# MSAA sample positions and centroid (Util.hpp:27-36, 182-203)
# ---------------------------------------------------------------------------

# Direct3D 10.1 standard sample positions, in 1/16 pixel from the pixel centre.
D3D_STANDARD = {
    1: {(4, 4), (-4, -4)},
    2: {(-2, -6), (6, -2), (-6, 2), (2, 6)},
    3: {(1, -3), (-1, 3), (5, 1), (-3, -5), (-5, 5), (-7, -1), (3, 7), (7, -7)},
}


@pytest.mark.parametrize("s", [0, 1, 2, 3])
def test_msaa_patterns_are_the_standard_positions(oracle_lib, s):
    n = 1 << s
    cols = [oracle_lib.msaa_pattern(s, i) for i in range(n)]
    assert sorted(cols) == list(range(n))                  # one sample per row and column
    offs = {((2 * cols[i] + 1 - n) * 8 // n, (2 * i + 1 - n) * 8 // n) for i in range(n)}
    if s == 0:
        assert offs == {(0, 0)}
    else:
        assert offs == D3D_STANDARD[s]


@pytest.mark.parametrize("s", [1, 2, 3])
def test_centroid_brute_force(oracle_lib, s):
    """Centroid sample = the covered sample nearest the pixel centre, lowest
    index on ties, none for empty or full masks; positions taken in pixel
    units from the standard offsets matched to each sample's row."""
    n = 1 << s
    row_y = {i: Fraction(2 * i + 1 - n, 2 * n) for i in range(n)}
    std = {Fraction(y, 16): Fraction(x, 16) for x, y in D3D_STANDARD[s]}
    pos = [(std[row_y[i]], row_y[i]) for i in range(n)]
    for mask in range(1 << n):
        got = oracle_lib.select_centroid(s, mask)
        covered = [i for i in range(n) if mask >> i & 1]
        if not covered or len(covered) == n:
            assert got == -1
        else:
            want = min(covered, key=lambda i: (pos[i][0] ** 2 + pos[i][1] ** 2, i))
            assert got == want


# ---------------------------------------------------------------------------
# Depth encoding (Util.hpp:290-310, CudaRaster.cpp:178)
# ---------------------------------------------------------------------------

def test_depth_encoding_closed_forms(oracle_lib):
    dmin, dmax = 2200 * 8, 2 ** 32 - 1 - 2200 * 8
    assert oracle_lib.encode_depth(0) == dmin
    assert oracle_lib.encode_depth(2 ** 32 - 1) == dmax
    assert oracle_lib.clear_depth(1.0) == dmax
    assert oracle_lib.clear_depth(0.0) == dmin
    assert oracle_lib.clear_depth(2.0) == dmax and oracle_lib.clear_depth(-1.0) == dmin
    # deferredClear quantises depth d to min(floor(d * 2^32), 2^32 - 1) first.
    for d, q in [(0.5, 2 ** 31), (0.25, 2 ** 30), (0.75, 3 * 2 ** 30), (2.0 ** -20, 2 ** 12),
                 (f32(0.999999), math.floor(Fraction(f32(0.999999)) * 2 ** 32))]:
        assert oracle_lib.clear_depth(d) == oracle_lib.encode_depth(q), d
    prev = 0
    for x in range(0, 2 ** 32, 2 ** 32 // 4099):
        e = oracle_lib.encode_depth(x)
        want = math.floor(Fraction(x, 2 ** 32 - 1) * (dmax - dmin) + dmin)
        assert abs(e - want) <= 1 and e >= prev
        prev = e
        assert abs(oracle_lib.decode_depth(e) - x) <= 2


# ---------------------------------------------------------------------------
# Coverage: top-left rule (Util.inl:95-102, 340-357)
# ---------------------------------------------------------------------------

def make_header(oracle_lib, pts):
    h = oracle_lib.Header()
    (h.v0x, h.v0y), (h.v1x, h.v1y), (h.v2x, h.v2y) = pts
    return h


def d3d_covered(tri_img, p):
    """Textbook Direct3D rule in image coordinates (row index grows
    downward): the triangle is clockwise on screen; a point strictly inside
    is covered; a point on an edge is covered only if that edge is a top
    edge (horizontal, the third vertex below it) or a left edge (not
    horizontal, the triangle to its right)."""
    (ax, ay), (bx, by), (cx, cy) = tri_img
    px, py = p
    verts = [(ax, ay), (bx, by), (cx, cy)]
    for k in range(3):
        (x0, y0), (x1, y1) = verts[k], verts[(k + 1) % 3]
        ox, oy = verts[(k + 2) % 3]
        side = lambda x, y: (x1 - x0) * (y - y0) - (y1 - y0) * (x - x0)
        sp, so = side(px, py), side(ox, oy)
        if sp == 0:
            if y0 == y1:
                top = oy > y0                               # third vertex below (larger row)
                if not top:
                    return False
            else:
                # the interior lies to the right of a left edge: the third
                # vertex is on the +x side of the edge's supporting line
                xline = x0 + Fraction(x1 - x0, y1 - y0) * (oy - y0)
                if not ox > xline:
                    return False
        elif (sp > 0) != (so > 0):
            return False
    return True


def random_small_triangle(rng, span=40):
    while True:
        pts = [(rng.randint(-span, span), rng.randint(-span, span)) for _ in range(3)]
        (x0, y0), (x1, y1), (x2, y2) = pts
        area = (x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0)
        if area > 0:                   # counter-clockwise in y-up = front-facing
            return pts


def test_coverage_matches_textbook_top_left_rule(oracle_lib):
    """Vertices and samples on a coarse lattice so that many samples fall
    exactly on edges and vertices."""
    rng = random.Random(5)
    W = H = 16          # header coordinates are relative to (W*8, H*8)
    for _ in range(300):
        pts = [(8 * x, 8 * y) for x, y in random_small_triangle(rng, span=6)]
        h = make_header(oracle_lib, pts)
        tri_abs = [(x + W * 8, y + H * 8) for x, y in pts]
        for sx in range(0, 256, 8):
            for sy in range(0, 256, 8):
                want = d3d_covered(tri_abs, (sx, sy))
                got = oracle_lib.lib().cro_cover_point(ctypes.byref(h), W, H, sx, sy)
                assert bool(got) == want, (pts, sx, sy)


def test_shared_edges_are_watertight(oracle_lib):
    """A jittered grid of quads split into triangles (both diagonals used):
    every sample inside the grid is covered by exactly one triangle."""
    rng = random.Random(6)
    W = H = 64
    n = 6
    step = 96
    grid = [[(i * step + rng.randint(-12, 12) if 0 < i < n else i * step,
              j * step + rng.randint(-12, 12) if 0 < j < n else j * step) for j in range(n + 1)] for i in range(n + 1)]
    tris = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = grid[i][j], grid[i + 1][j], grid[i + 1][j + 1], grid[i][j + 1]
            if (i + j) % 2:
                tris += [(a, b, c), (a, c, d)]
            else:
                tris += [(a, b, d), (b, c, d)]
    L = oracle_lib.lib()
    hdrs = [make_header(oracle_lib, [(x - W * 8, y - H * 8) for x, y in t]) for t in tris]
    for s in range(0, 4):
        nS = 1 << s
        for px in range(0, n * step // 16):
            for py in range(0, n * step // 16):
                total = 0
                for h in hdrs:
                    total += bin(L.cro_cover_pixel(ctypes.byref(h), W, H, s, px, py)).count("1")
                assert total == nS, (s, px, py, total)


def test_axis_aligned_rectangle_counts(oracle_lib):
    """Rectangle [x0,x1) x [y0,y1) in pixels, split in two triangles, with
    corners on pixel centres' grid lines: covers exactly its pixels."""
    W = H = 32
    L = oracle_lib.lib()
    x0, y0, x1, y1 = 16 * 3, 16 * 5, 16 * 11, 16 * 9        # pixel borders
    a, b, c, d = (x0, y0), (x1, y0), (x1, y1), (x0, y1)
    tris = [(a, b, c), (a, c, d)]
    hdrs = [make_header(oracle_lib, [(x - W * 8, y - H * 8) for x, y in t]) for t in tris]
    covered = set()
    for px in range(W):
        for py in range(H):
            if any(L.cro_cover_pixel(ctypes.byref(h), W, H, 0, px, py) for h in hdrs):
                covered.add((px, py))
    assert covered == {(x, y) for x in range(3, 11) for y in range(5, 9)}


# ---------------------------------------------------------------------------
# Plane equations (Util.inl:65-91) and barycentrics (FineRaster.inl:21-44)
# ---------------------------------------------------------------------------

def plane_at(pleq, sx, sy):
    """Evaluate the U32 plane at fractional sample coordinates, modular."""
    zx, zy, zb = (Fraction(v if v < 2 ** 31 else v - 2 ** 32) for v in pleq)
    return zx * sx + zy * sy + zb


def test_setup_pleq_reproduces_vertex_values(oracle_lib):
    rng = random.Random(7)
    for _ in range(3000):
        s = rng.randint(0, 3)
        pts = random_small_triangle(rng, span=rng.choice([20, 200, 2000]))
        (x0, y0), (x1, y1), (x2, y2) = pts
        d1, d2 = (x1 - x0, y1 - y0), (x2 - x0, y2 - y0)
        area = d1[0] * d2[1] - d1[1] * d2[0]
        # Values of a linear function (depth-like: within [2^14, 2^32 - 2^14],
        # gradient small enough that no vertex difference exceeds 2^30).
        span = max(max(abs(x), abs(y)) for x, y in pts)
        g = 2.0 ** rng.randint(0, int(math.log2(2 ** 29 / (2 * span))))
        gx, gy = rng.uniform(-g, g), rng.uniform(-g, g)
        base = rng.uniform(2 ** 30, 3 * 2 ** 30)
        vals = [f32(base + gx * x + gy * y) for x, y in pts]
        rcp = f32(1.0 / np.float32(area))
        ox, oy = rng.randint(0, 30000), rng.randint(0, 30000)
        v0 = (x0 + ox, y0 + oy)
        pleq = oracle_lib.setup_pleq(vals, v0, d1, d2, rcp, s)
        unit = Fraction(16, 1 << s)        # subpixels per sample unit
        mx = max(vals)
        sh = min(max(int(math.floor(math.log2(mx))) - 22, 0), 8)
        # Error budget: values truncated to multiples of 2^sh (< 2^sh); each
        # gradient floored to an integer per sample unit (< 1 per unit of
        # distance from the bounding-box centre the plane is anchored at,
        # i.e. half the extent in sample units per axis); the reciprocal's
        # relative error 2^-24 over the value range (< 2^7).  Far below the
        # reference's own bound LERP_ERROR = 2200 << s.
        ext = (max(x for x, _ in pts) - min(x for x, _ in pts)) + (max(y for _, y in pts) - min(y for _, y in pts))
        tol = (1 << sh) + ext * (1 << s) / 32 + 2 ** 7
        assert tol < 2200 << s or ext > 1000
        for (px, py), v in zip([(x0, y0), (x1, y1), (x2, y2)], vals):
            sx, sy = Fraction(px + ox, 1) / unit, Fraction(py + oy, 1) / unit
            got = plane_at(pleq, sx, sy)
            err = (got - Fraction(v) + 2 ** 31) % 2 ** 32 - 2 ** 31
            assert abs(err) <= tol, (s, pts, vals, float(err))


def test_setup_pleq_constant_values(oracle_lib):
    for v in [f32(17600.0), f32(1e6), f32(2147483648.0), f32(4294949696.0)]:
        pleq = oracle_lib.setup_pleq([v, v, v], (100, 200), (50, 3), (-7, 40), f32(1.0 / (50 * 40 + 3 * 7)), 2)
        sh = min(max(int(math.floor(math.log2(v))) - 22, 0), 8)
        assert pleq[0] == 0 and pleq[1] == 0 and pleq[2] == (int(v) >> sh) << sh


def test_barycentrics_are_perspective_correct(oracle_lib):
    """Render one large slanted triangle with per-vertex colours (1,0,0),
    (0,1,0), (0,0,1): each pixel's colour is the perspective-correct
    barycentric triple at the pixel centre, to within the quantisation."""
    from cudaraster import scenes
    W = H = 64
    clip = np.array([[-0.8 * 1.0, -0.8 * 1.0, 0.5, 1.0], [0.9 * 4.0, -0.7 * 4.0, 3.0, 4.0],
                     [-0.6 * 2.0, 0.9 * 2.0, 1.0, 2.0]], dtype=np.float32)
    col = np.array([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1]], dtype=np.float32)
    s = scenes.Scene("bary", W, H, 0, 3, 0, 0, np.concatenate([clip, col], 1), np.array([[0, 1, 2]], np.int32))
    r = oracle_lib.render(s)
    img = r["color"].view(np.uint8).reshape(H, W, 4).astype(float) / 255.0
    ndc = clip[:, :2] / clip[:, 3:4]
    scr = (ndc * 0.5 + 0.5) * [W, H]
    checked = 0
    for py in range(H):
        for px in range(W):
            if r["triid"][py, px] != 0:
                continue
            p = np.array([px + 0.5, py + 0.5])
            T = np.array([scr[1] - scr[0], scr[2] - scr[0]]).T
            l1, l2 = np.linalg.solve(T, p - scr[0])
            l = np.array([1 - l1 - l2, l1, l2])          # screen-space barycentrics
            b = l / clip[:, 3]
            b /= b.sum()                                   # perspective correction
            assert np.abs(img[py, px, :3] - b).max() < 1.5 / 255 + 1e-3, (px, py, img[py, px], b)
            checked += 1
    assert checked > 500


# ---------------------------------------------------------------------------
# Clipping (Util.hpp:221-286) against a float64 Sutherland-Hodgman
# ---------------------------------------------------------------------------

def sh_clip_f64(v0, v1, v2):
    poly = [np.array(v, dtype=np.float64) for v in (v0, v1, v2)]
    for axis in range(3):
        if any(v[3] < abs(v[axis]) for v in (v0, v1, v2)):
            for sign in (1.0, -1.0):
                out = []
                n = len(poly)
                for i in range(n):
                    a, b = poly[i - 1], poly[i]
                    da, db = a[3] + sign * a[axis], b[3] + sign * b[axis]
                    if da * db < 0:
                        t = da / (da - db)
                        out.append(a + (b - a) * t)
                    if db >= 0:
                        out.append(b)
                poly = out
    return poly


def test_clip_triangle_against_float64(oracle_lib):
    rng = random.Random(8)
    checked = 0
    for _ in range(2000):
        vs = []
        for _ in range(3):
            w = rng.uniform(-1.0, 3.0)
            vs.append([rng.uniform(-3, 3) * abs(w) + rng.uniform(-1, 1), rng.uniform(-3, 3) * abs(w),
                       rng.uniform(-1.5, 1.5) * abs(w), w])
        vs = [[f32(x) for x in v] for v in vs]
        bary = oracle_lib.clip_triangle(*vs)
        ref = sh_clip_f64(*vs)
        V0, V1, V2 = (np.array(v, dtype=np.float64) for v in vs)
        got = [V0 + (V1 - V0) * b[0] + (V2 - V0) * b[1] for b in bary]
        if len(ref) != len(got):
            # a vertex within float rounding of a plane may be kept by one and
            # dropped by the other; those polygons differ by a sliver only
            continue
        for g, rr in zip(got, ref):
            assert np.allclose(g, rr, rtol=1e-4, atol=1e-4 * max(1.0, np.abs(rr).max()))
        for g in got:
            assert (np.abs(g[:3]) <= g[3] * (1 + 1e-4) + 1e-4).all()
        checked += len(got) >= 3
    assert checked > 500


def test_clip_vertices_on_planes(oracle_lib):
    """Vertices exactly on a clip plane are kept once (a vertex with
    w + x == 0 lies inside the closed half-space)."""
    cases = [
        ([1, 0, 0, 1], [3, 0.5, 0, 1], [0, 1, 0, 1]),      # v0 on x = w
        ([1, 0, 0, 1], [0, 1, 0, 1], [2, -1, 0, 1]),
        ([-1, -1, 0, 1], [3, -1, 0, 1], [-1, 3, 0, 1]),    # two vertices on planes
        ([0, 0, -1, 1], [0.5, 0, -3, 1], [0, 0.5, 0.5, 1]),  # on the near plane
    ]
    for vs in cases:
        vs = [[f32(x) for x in v] for v in vs]
        bary = oracle_lib.clip_triangle(*vs)
        ref = sh_clip_f64(*vs)
        assert len(bary) == len(ref), vs
        V0, V1, V2 = (np.array(v, dtype=np.float64) for v in vs)
        got = [V0 + (V1 - V0) * b[0] + (V2 - V0) * b[1] for b in bary]
        for g, r in zip(got, ref):
            assert np.allclose(g, r, atol=1e-6)


# ---------------------------------------------------------------------------
# Setup culls (TriangleSetup.inl:38-100, 262-281)
# ---------------------------------------------------------------------------

def one_tri_scene(clip, W=64, H=64, s=0):
    from cudaraster import scenes
    col = np.ones((3, 4), np.float32)
    return scenes.Scene("one", W, H, s, 3, 0, 0, np.concatenate([np.asarray(clip, np.float32), col], 1),
                        np.array([[0, 1, 2]], np.int32))


def test_setup_culls(oracle_lib):
    ccw = [[-0.5, -0.5, 0, 1], [0.5, -0.5, 0, 1], [0, 0.5, 0, 1]]
    assert oracle_lib.setup_triangle(one_tri_scene(ccw), 0)[2] == 0
    cw = [ccw[0], ccw[2], ccw[1]]
    assert oracle_lib.setup_triangle(one_tri_scene(cw), 0)[2] == 2          # backface
    degenerate = [[-0.5, -0.5, 0, 1], [0.5, 0.5, 0, 1], [0.0, 0.0, 0, 1]]
    assert oracle_lib.setup_triangle(one_tri_scene(degenerate), 0)[2] == 2
    outside = [[2, 0, 0, 1], [3, 0, 0, 1], [2.5, 1, 0, 1]]
    assert oracle_lib.setup_triangle(one_tri_scene(outside), 0)[2] == 1     # frustum
    # A sliver between pixel centres (64x64: pixel = 1/32 NDC): culled.
    c = 0.5 / 32
    between = [[c + 0.001, c + 0.001, 0, 1], [c + 0.02, c + 0.001, 0, 1], [c + 0.001, c + 0.02, 0, 1]]
    assert oracle_lib.setup_triangle(one_tri_scene(between), 0)[2] == 3
    # The same sliver moved over a pixel centre is kept.
    over = [[c - 0.003 + 1 / 32, c - 0.003 + 1 / 32, 0, 1], [c + 0.05 + 1 / 32, c - 0.003 + 1 / 32, 0, 1],
            [c - 0.003 + 1 / 32, c + 0.05 + 1 / 32, 0, 1]]
    assert oracle_lib.setup_triangle(one_tri_scene(over), 0)[2] == 0
    # Wide and flat: spans many pixel centres in x but none in y (screen
    # rows 32.6 .. 33.4 hold no centre): culled although its box is wide.
    def ndc(px, py):
        return [px / 32 - 1, py / 32 - 1, 0, 1]
    flat = [ndc(10.2, 32.6), ndc(50.3, 32.7), ndc(30.0, 33.4)]
    assert oracle_lib.setup_triangle(one_tri_scene(flat), 0)[2] == 3
    # Crossing the near plane takes the clip path and yields a fan.
    near = [[-0.5, -0.5, -2.0, 1], [0.5, -0.5, 0.5, 1], [0, 0.5, 0.5, 1]]
    h, d, code = oracle_lib.setup_triangle(one_tri_scene(near), 0)
    assert code == 5 and len(h) >= 1


def test_snap_rounds_ties_to_even(oracle_lib):
    """R4: cvt.rni.sat.s32 (TriangleSetup.inl:19-31, Util.hpp:88) rounds a
    subpixel position that lies exactly half-way to the even neighbour.
    Every vertex here is such a tie (k + 1/2, both signs and parities); the
    expected header coordinates come from Python's round(), which is
    round-half-even on exact binary fractions.  The host emulation's
    floor(x + 0.5) (CudaRaster.cpp:701-703) and round-half-away both
    disagree on half of these vertices."""
    from cudaraster import scenes
    for s in (0, 2):
        sc, ks = scenes.snap_ties(samples_log2=s, seed=12 + s)
        seen = {"even+": 0, "odd+": 0, "even-": 0, "odd-": 0}
        visible = 0
        for t in range(sc.num_tris):
            h, _, code = oracle_lib.setup_triangle(sc, t)
            if code != 0:
                continue
            visible += 1
            got = [(h[0].v0x, h[0].v0y), (h[0].v1x, h[0].v1y), (h[0].v2x, h[0].v2y)]
            for k in range(3):
                for axis in range(2):
                    kk = int(ks[t, k, axis])
                    want = round(Fraction(2 * kk + 1, 2))      # exact tie, half to even
                    assert want == (kk if kk % 2 == 0 else kk + 1)
                    assert got[k][axis] == want, (t, k, axis, kk)
                    seen[("even" if kk % 2 == 0 else "odd") + ("+" if kk >= 0 else "-")] += 1
        assert visible > 200 and min(seen.values()) > 100, seen


def test_guard_band_extent_decides_fast_path(oracle_lib):
    """R3: the fast path needs the snapped box inside [-32768, 32767] and
    max(hi.x, hi.y) - min(lo.x, lo.y) <= 2^15 - 1 (TriangleSetup.inl:285-296);
    anything else goes through the clip path (cull code 5) even when no
    vertex is outside the frustum.  Cases sit exactly on and one subpixel
    past each limit, including an extent that exceeds the limit only when
    the axes are mixed."""
    from cudaraster import scenes
    sc = scenes.guard_band()
    for t, (box, fast) in enumerate(scenes.GUARD_BAND_CASES):
        h, _, code = oracle_lib.setup_triangle(sc, t)
        assert code == (0 if fast else 5), (box, code)
        assert len(h) >= 1, box
        if fast:   # the snapped vertices are the box corners themselves
            x0, x1, y0, y1 = box
            assert (h[0].v0x, h[0].v0y, h[0].v1x, h[0].v1y, h[0].v2x, h[0].v2y) == (x0, y0, x1, y0, x0, y1)
    st = oracle_lib.render(sc)["stats"]
    assert st["clipped"] == sum(1 for _, fast in scenes.GUARD_BAND_CASES if not fast)


# ---------------------------------------------------------------------------
# Whole draw: an independent pure-Python sequential rasterizer
# ---------------------------------------------------------------------------

def python_render(oracle_lib, scene):
    """Sequential reference in plain Python loops: per subtriangle in
    submission order, per sample of its box, textbook top-left coverage in
    image coordinates, LESS depth from the oracle's plane equations, and the
    blend in exact rational arithmetic."""
    W, H, s = scene.width, scene.height, scene.samples_log2
    n = 1 << s
    wr, hr = (W + 7) & ~7, (H + 7) & ~7
    tri_sub, hdr, dat, sub = oracle_lib.setup(scene)
    color = np.full((hr, wr * n), oracle_lib.to_abgr_host(scene.clear_color), np.uint64)
    depth = np.full((hr, wr * n), oracle_lib.clear_depth(scene.clear_depth), np.uint64)
    tid = np.full((hr, wr * n), 0xFFFFFFFF, np.uint64)
    order = []
    for t, k in enumerate(tri_sub):
        order += [t] if k == 1 else [int(hdr[t]["misc"]) + j for j in range(k)]
    frags = samples = zkills = 0
    pat = [oracle_lib.msaa_pattern(s, i) for i in range(n)]
    for slot in order:
        h, d = hdr[slot], dat[slot]
        tri = [(int(h["v0x"]) + W * 8, int(h["v0y"]) + H * 8), (int(h["v1x"]) + W * 8, int(h["v1y"]) + H * 8),
               (int(h["v2x"]) + W * 8, int(h["v2y"]) + H * 8)]
        xs, ys = [p[0] for p in tri], [p[1] for p in tri]
        for py in range(max(0, min(ys) // 16), min(hr - 1, max(ys) // 16) + 1):
            for px in range(max(0, min(xs) // 16), min(wr - 1, max(xs) // 16) + 1):
                cov = [d3d_covered(tri, (((px * n + pat[i]) * 2 + 1) * 8 // n, ((py * n + i) * 2 + 1) * 8 // n))
                       for i in range(n)]
                if not any(cov):
                    continue
                frags += 1
                samples += sum(cov)
                col_base = (px >> 3) * 8 * n + (px & 7)
                written = 0
                for i in range(n):
                    if not cov[i]:
                        continue
                    z = (int(d["zx"]) * (px * n + pat[i]) + int(d["zy"]) * (py * n + i) + int(d["zb"])) % 2 ** 32
                    if z >= depth[py, col_base + 8 * i]:
                        continue
                    depth[py, col_base + 8 * i] = z
                    tid[py, col_base + 8 * i] = int(sub[slot])
                    written += 1
                zkills += written == 0
    python_render.counts = {"fragments": frags, "coveredSamples": samples, "zKills": zkills,
                            "trisVisible": sum(1 for k in tri_sub if k)}
    return depth.astype(np.uint32), tid.astype(np.uint32), frags


@pytest.mark.parametrize("s", [0, 2])
def test_oracle_equals_python_sequential_renderer(oracle_lib, s):
    from cudaraster import scenes
    sc = scenes.random_tris(150, 40, 32, samples_log2=s, seed=9 + s, area_range=(2.0, 120.0))
    r = oracle_lib.render(sc)
    depth, tid, frags = python_render(oracle_lib, sc)
    assert np.array_equal(r["depth"], depth)
    assert np.array_equal(r["triid"], tid)
    assert r["stats"]["fragments"] == frags
    for k, v in python_render.counts.items():
        assert r["stats"][k] == v, k


def test_profile_counts_closed_forms(oracle_lib):
    """The profiling counts of the oracle (covered samples, depth-killed
    fragments, triangles with a visible subtriangle) on scenes whose values
    are fixed by construction: every triangle drawn twice in a row kills
    every fragment of the second copy (LESS), and doubles the samples and
    visible triangles; a full-viewport quad covers every sample once."""
    from cudaraster import scenes
    for s in (0, 1, 3):
        sc = scenes.random_tris(80, 48, 40, samples_log2=s, seed=30 + s, area_range=(4.0, 300.0))
        one = oracle_lib.render(sc)["stats"]
        idx = np.repeat(sc.indices, 2, axis=0)
        dup = scenes.Scene("dup", sc.width, sc.height, s, sc.flags, sc.blend, sc.shader, sc.vertices,
                           np.ascontiguousarray(idx))
        two = oracle_lib.render(dup)["stats"]
        assert two["fragments"] == 2 * one["fragments"]
        assert two["coveredSamples"] == 2 * one["coveredSamples"]
        assert two["trisVisible"] == 2 * one["trisVisible"] == 2 * one["numSubtris"]
        assert two["zKills"] == one["zKills"] + one["fragments"]
        # Without the depth test nothing is killed.
        nod = scenes.Scene("nodepth", sc.width, sc.height, s, scenes.FLAG_LERP, sc.blend, sc.shader,
                           sc.vertices, np.ascontiguousarray(idx))
        assert oracle_lib.render(nod)["stats"]["zKills"] == 0
        # A quad over the whole (tile-rounded) surface: two triangles, each
        # sample covered exactly once.
        W, H = 48, 40
        quad = np.array([[-1, -1, 0.5, 1], [1, -1, 0.5, 1], [1, 1, 0.5, 1], [-1, 1, 0.5, 1]], np.float32)
        verts = np.concatenate([quad, np.ones((4, 4), np.float32)], axis=1)
        q = scenes.Scene("quad", W, H, s, sc.flags, sc.blend, sc.shader, np.ascontiguousarray(verts),
                         np.array([[0, 1, 2], [0, 2, 3]], np.int32))
        st = oracle_lib.render(q)["stats"]
        assert st["coveredSamples"] == W * H * (1 << s)
        assert st["zKills"] == 0 and st["trisVisible"] == 2


def test_depth_test_is_strictly_less(oracle_lib):
    """The same triangle drawn twice: the second copy fails LESS at every
    sample, so triangle ids keep the first copy."""
    from cudaraster import scenes
    sc = scenes.random_tris(60, 40, 32, samples_log2=2, seed=11, area_range=(4.0, 200.0))
    idx = np.repeat(sc.indices, 2, axis=0)
    dup = scenes.Scene("dup", sc.width, sc.height, sc.samples_log2, sc.flags, sc.blend, sc.shader,
                       sc.vertices, np.ascontiguousarray(idx))
    r = oracle_lib.render(dup)
    ids = r["triid"][r["triid"] != 0xFFFFFFFF]
    assert ids.size > 1000 and (ids % 2 == 0).all()


def test_hierarchical_z_bound_holds(oracle_lib):
    """The header's zmin (TriangleSetup.inl:147) never exceeds the depth the
    plane equation gives at a covered sample -- the premise of the fine
    stage's tile-level early rejection."""
    from cudaraster import scenes
    sc = scenes.random_tris(400, 48, 48, samples_log2=2, seed=10, area_range=(1.0, 500.0))
    tri_sub, hdr, dat, sub = oracle_lib.setup(sc)
    L = oracle_lib.lib()
    W = H = 48
    n = 4
    checked = 0
    for slot in range(len(hdr)):
        if slot < sc.num_tris and tri_sub[slot] != 1:
            continue
        h = oracle_lib.Header(*[int(hdr[slot][f]) for f in ("v0x", "v0y", "v1x", "v1y", "v2x", "v2y", "misc")])
        zmin = int(hdr[slot]["misc"]) & 0xFFFFF000
        d = dat[slot]
        xs = [int(hdr[slot][f]) + W * 8 for f in ("v0x", "v1x", "v2x")]
        ys = [int(hdr[slot][f]) + H * 8 for f in ("v0y", "v1y", "v2y")]
        for px in range(max(0, min(xs) // 16), min(W - 1, max(xs) // 16) + 1):
            for py in range(max(0, min(ys) // 16), min(H - 1, max(ys) // 16) + 1):
                m = L.cro_cover_pixel(ctypes.byref(h), W, H, 2, px, py)
                for i in range(n):
                    if m >> i & 1:
                        z = (int(d["zx"]) * (px * n + oracle_lib.msaa_pattern(2, i)) + int(d["zy"]) * (py * n + i)
                             + int(d["zb"])) % 2 ** 32
                        assert z >= zmin
                        checked += 1
    assert checked > 1000
