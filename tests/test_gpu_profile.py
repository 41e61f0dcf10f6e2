"""Profiling modes on the GPU: the counters of a CR_PROFILING_COUNTERS draw
against the oracle's counts of the same scene, their consistency across
stages, the getStats convention and the getProfilingInfo report
(CudaRaster.cpp:346-497, PrivateDefs.hpp:158-204).

Setup counts (viewport, backface and between-pixels culls, clipped
triangles), covered samples, fragments and depth-killed fragments are exact
functions of the scene and must equal the oracle's; the bin-box size is
checked against the bins the oracle's snapped headers overlap.
"""
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _profiled(scene, **debug):
    import cudaraster
    r = cudaraster.SceneRenderer(scene, debug=dict(profiling_mode=cudaraster.PROFILING_COUNTERS, **debug))
    try:
        r.draw()
        out = r.read()
        out["prof"] = r.cr.getProfileCounters()
        out["counters"] = r.cr.getCounters()
        out["report"] = r.cr.getProfilingInfo()
    finally:
        r.close()
    return out


def _bins_per_subtri(oracle_lib, scene):
    """Bins each visible subtriangle's snapped box overlaps: bins are
    2^b x 2^b tiles with the smallest b that keeps at most 256 bins."""
    tri_sub, hdr, _, _ = oracle_lib.setup(scene)
    W, H = scene.width, scene.height
    wt, ht = (W + 7) // 8, (H + 7) // 8
    b = 0
    while ((wt + (1 << b) - 1) >> b) * ((ht + (1 << b) - 1) >> b) > 256 and b < 4:
        b += 1
    wb, hb = (wt + (1 << b) - 1) >> b, (ht + (1 << b) - 1) >> b
    slots = []
    for t, k in enumerate(tri_sub):
        slots += [t] if k == 1 else [int(hdr[t]["misc"]) + j for j in range(k)]
    h = hdr[np.array(slots, dtype=np.int64)] if slots else hdr[:0]
    xs = np.stack([h["v0x"], h["v1x"], h["v2x"]]).astype(np.int64) + W * 8
    ys = np.stack([h["v0y"], h["v1y"], h["v2y"]]).astype(np.int64) + H * 8
    sh = 4 + 3 + b
    minx, maxx, miny, maxy = xs.min(0), xs.max(0), ys.min(0), ys.max(0)
    on = (maxx >= 0) & (maxy >= 0) & (minx < wt * 8 * 16) & (miny < ht * 8 * 16)
    nx = np.minimum(maxx >> sh, wb - 1) - np.maximum(minx >> sh, 0) + 1
    ny = np.minimum(maxy >> sh, hb - 1) - np.maximum(miny >> sh, 0) + 1
    n = np.where(on, nx * ny, 0)
    return n, len(slots)


SCENES = [
    ("random 1x", lambda s: s.random_tris(30_000, 320, 256, samples_log2=0, seed=61)),
    ("random 4x", lambda s: s.random_tris(30_000, 320, 256, samples_log2=2, seed=62)),
    ("clip mix 8x", lambda s: s.clip_mix(6_000, 256, 256, samples_log2=3, seed=63)),
    ("skewed 2x", lambda s: s.skewed(num_small=20_000, num_big=24, width=512, height=384, samples_log2=1, seed=64)),
]


@pytest.mark.parametrize("name,make", SCENES, ids=[n for n, _ in SCENES])
def test_counters_equal_oracle(oracle_lib, name, make):
    _torch()
    from cudaraster import scenes
    sc = make(scenes)
    print(f"profile: {name}", flush=True)
    got = _profiled(sc, disable_hiz=True)
    ref = oracle_lib.render(sc)
    st, p = ref["stats"], got["prof"]
    # The profiled draw draws the same image.
    assert np.array_equal(got["depth"], ref["depth"]) and np.array_equal(got["triid"], ref["triid"])
    T = sc.num_tris
    assert p["setup_viewport_cull"] == (100 * st["culledFrustum"], T)
    assert p["setup_backface_cull"] == (100 * st["culledBackface"], T)
    assert p["setup_between_pixels_cull"] == (100 * st["culledBetween"], T)
    assert p["setup_clipped"] == (100 * st["clipped"], T)
    assert p["setup_samples_per_tri"] == (st["coveredSamples"], st["trisVisible"])
    assert p["fine_z_kill"] == (100 * st["zKills"], st["fragments"])
    assert p["fine_frag_per_tile"][0] == st["fragments"]
    assert p["fine_frag_per_tri"][0] == st["fragments"]
    assert p["fine_early_z_cull"][0] == 0          # hierarchical Z off
    bins, nsub = _bins_per_subtri(oracle_lib, sc)
    assert nsub == st["numSubtris"]
    assert p["bin_tri_bb_area"] == (int(bins.sum()), nsub)
    assert p["bin_tri_single_bin"] == (100 * int((bins == 1).sum()), nsub)
    assert p["bin_tris_per_chunk"] == (nsub, (T + 1023) // 1024)
    assert got["counters"]["bin_entries"] == int(bins.sum())
    # Stage consistency: coarse walks every bin entry once and emits exactly
    # the entries fine walks; every walked entry is classified once.
    entries = p["coarse_tris_per_round"][0]
    assert entries == int(bins.sum())
    assert p["coarse_emits_per_tri"] == (p["coarse_emits_per_round"][0], entries)
    assert p["fine_tri_per_tile"][0] == p["coarse_emits_per_round"][0]
    assert p["coarse_case_a"][0] + p["coarse_case_b"][0] + p["coarse_case_c"][0] == 100 * entries
    assert p["coarse_rounds_per_bin"][0] == p["coarse_tris_per_round"][1]
    assert p["fine_empty_cull"][1] == p["fine_tri_per_tile"][0]
    assert p["fine_blend_rounds"][0] == p["fine_blend_rounds"][1] == st["fragments"] - st["zKills"]
    assert p["bin_queue_overflows"] == (got["counters"]["retries"], 1)


def test_hiz_counter_and_image(oracle_lib):
    """With hierarchical Z on, the profiled draw rejects some tile entries
    (a deep scene) and still draws the oracle's image."""
    _torch()
    from cudaraster import scenes
    sc = scenes.random_tris(60_000, 256, 256, samples_log2=2, seed=65, area_range=(20.0, 2000.0))
    got = _profiled(sc)
    ref = oracle_lib.render(sc)
    assert np.array_equal(got["depth"], ref["depth"]) and np.array_equal(got["triid"], ref["triid"])
    num, den = got["prof"]["fine_early_z_cull"]
    assert den == got["prof"]["fine_tri_per_tile"][0] and 0 < num < 100 * den


def _report_values(text):
    vals = {}
    for line in text.splitlines():
        m = re.match(r"^\s*- (.+?)\s+(-?[0-9.]+)%?$", line)
        if m:
            vals[m.group(1).strip()] = float(m.group(2))
    return vals


def test_counters_report_text():
    """getProfilingInfo in counters mode prints each counter as
    num / max(denom, 1) with the reference's labels and precision."""
    _torch()
    from cudaraster import scenes
    sc = scenes.clip_mix(4_000, 192, 160, samples_log2=2, seed=66)
    got = _profiled(sc)
    text, p = got["report"], got["prof"]
    assert "ProfilingMode_Counters" in text
    for head in ("TriangleSetup:", "BinRaster:", "CoarseRaster:", "FineRaster:"):
        assert head in text
    vals = _report_values(text)
    ratio = lambda k: p[k][0] / max(p[k][1], 1)
    checks = {"Viewport cull": ("setup_viewport_cull", 1), "Backface cull": ("setup_backface_cull", 1),
              "Between pixels cull": ("setup_between_pixels_cull", 1), "Clipped": ("setup_clipped", 1),
              "Avg. samples / tri": ("setup_samples_per_tri", 2), "Avg. tri bb size": ("bin_tri_bb_area", 1),
              "Emits / Triangle": ("coarse_emits_per_tri", 2), "Z kills": ("fine_z_kill", 1),
              "Avg. frag/tri": ("fine_frag_per_tri", 1), "Blend attempt rounds": ("fine_blend_rounds", 2)}
    for label, (key, digits) in checks.items():
        assert label in vals, label
        assert vals[label] == pytest.approx(round(ratio(key), digits), abs=10 ** -digits), label
    assert vals["Viewport cull"] > 0 and vals["Clipped"] > 0


def test_default_report_and_stats():
    """Default mode: getStats gives per-stage seconds with bin_time 0 (setup
    and binning are one kernel), the report prints them in the reference's
    '%-16s%.3f ms (%.0f%%)' layout with the queue sizes, and the profile
    counters of a default draw are zero."""
    _torch()
    import cudaraster
    from cudaraster import scenes
    sc = scenes.random_tris(100_000, 1024, 768, samples_log2=2, seed=67)
    r = cudaraster.SceneRenderer(sc)
    try:
        r.draw()
        stats = r.cr.getStats()
        counters = r.cr.getCounters()
        text = r.cr.getProfilingInfo()
        prof = r.cr.getProfileCounters()
    finally:
        r.close()
    assert stats["binTime"] == 0.0
    for k in ("setupTime", "coarseTime", "fineTime"):
        assert 0.0 < stats[k] < 0.5, k
    assert "ProfilingMode_Default" in text
    rows = {m.group(1): float(m.group(2)) for m in re.finditer(r"^(\w+)\s+([0-9.]+) ms \(", text, re.M)}
    assert rows["binRaster"] == 0.0
    assert rows["fineRaster"] == pytest.approx(stats["fineTime"] * 1e3, abs=1.5e-3)
    m = re.search(r"^numSubtris\s+(\d+)", text, re.M)
    assert m and int(m.group(1)) == counters["num_subtris"]
    m = re.search(r"^numTileEntries\s+(\d+)", text, re.M)
    assert m and int(m.group(1)) == counters["tile_entries"]
    assert all(v == (0, 0) for v in prof.values())
