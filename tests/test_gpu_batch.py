"""Batched multi-view draws (cr_batch_*): every view's surfaces must be
bit-identical to a plain draw of the same view, on every graph replay, with
the autotuner's capacities and workgroup shape, and after overflow retries."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def mixed_views(n=12, seed=5):
    """Small views covering every fine-stage variant: 1x/2x/4x/8x, replace
    (deferred shading) and src-over / additive (in-order shading), odd sizes,
    an empty view, a view drawn over existing contents."""
    from cudaraster import scenes
    vs = scenes.multi_view(n, seed=seed, resolutions=(64, 96, 136, 256), tri_range=(200, 6000),
                           samples_choices=(0, 1, 2, 3))
    out = []
    for k, v in enumerate(vs):
        if k % 4 == 1:
            v = scenes.random_tris(v.num_tris, v.width, v.height, samples_log2=v.samples_log2, seed=100 + k,
                                   alpha_range=(0.2, 0.8), blend=1, name=v.name)
        if k % 4 == 2:
            v = dataclasses.replace(v, blend=2)
        if k == 3:
            v = dataclasses.replace(v, width=203, height=77)
        if k == 5:
            v = dataclasses.replace(v, indices=v.indices[:0])
        if k == 7:
            v = dataclasses.replace(v, deferred_clear=False)
        out.append(v)
    return out


def plain(scene, preload=None):
    import cudaraster
    r = cudaraster.SceneRenderer(scene)
    if preload is not None:
        r.color.upload(preload["color"])
        r.depth.upload(preload["depth"])
        r.triid.upload(preload["triid"])
    r.draw()
    out = r.read()
    r.close()
    return out


def preload_for(scene):
    rng = np.random.default_rng(9)
    shape = scene.surface_shape
    return {"color": rng.integers(0, 2 ** 32, size=shape, dtype=np.uint64).astype(np.uint32),
            "depth": rng.integers(2 ** 31, 2 ** 32 - 1, size=shape, dtype=np.uint64).astype(np.uint32),
            "triid": np.full(shape, 0xFFFFFFFF, np.uint32)}


def load_preloads(br, views):
    pre = {}
    for i, v in enumerate(views):
        if not v.deferred_clear:
            pre[i] = preload_for(v)
            col, dep, tid = br.surfaces[i]
            col.upload(pre[i]["color"])
            dep.upload(pre[i]["depth"])
            tid.upload(pre[i]["triid"])
    return pre


def check_all(br, views, refs, what):
    for i in range(len(views)):
        got = br.read(i)
        for k in ("color", "depth", "triid"):
            assert np.array_equal(got[k], refs[i][k]), f"{what}: view {i} ({views[i].name}) {k} differs"


def test_batch_equals_plain_draws_on_every_replay():
    torch = _torch()
    import cudaraster
    views = mixed_views()
    pre = {i: preload_for(v) for i, v in enumerate(views) if not v.deferred_clear}
    refs = [plain(v, pre.get(i)) for i, v in enumerate(views)]
    br = cudaraster.BatchRenderer(views)
    try:
        for rep in range(3):                 # graph replays, with eager torch work between them
            load_preloads(br, views)
            br.draw()
            x = torch.randn(1 << 20, device="cuda")
            torch.cuda.synchronize()
            assert float((x * 0).sum()) == 0.0
            check_all(br, views, refs, f"replay {rep}")
        info = br.batch.info()
        assert info["uses_graph"] == 1 and info["num_views"] == len(views)
    finally:
        br.close()


@pytest.mark.parametrize("cfg", [(2, 1, 1024, 1), (8, 0, 256, 2), (4, 1, 256, 4), (4, 0, 1024, 3)])
def test_batch_configurations_give_identical_images(cfg):
    """Every launch configuration (fine shape, placement, coarse size,
    concurrent view groups) draws the same images, graph or direct."""
    _torch()
    import cudaraster
    views = mixed_views(9, seed=12)
    refs = [plain(v) if v.deferred_clear else None for v in views]
    br = cudaraster.BatchRenderer(views)
    try:
        br.batch.configure(*cfg)
        for graph in (True, False):
            br.batch.use_graph(graph)
            br.draw()
            info = br.batch.info()
            assert (info["fine_tiles_per_group"], info["fine_banded"], info["coarse_threads"],
                    info["view_groups"]) == cfg
            for i, v in enumerate(views):
                if refs[i] is None:
                    continue
                got = br.read(i)
                for k in ("color", "depth", "triid"):
                    assert np.array_equal(got[k], refs[i][k]), f"{cfg} graph={graph}: view {i} {k}"
    finally:
        br.close()


def test_batch_against_oracle(oracle_lib):
    _torch()
    import cudaraster
    views = mixed_views(6, seed=8)
    br = cudaraster.BatchRenderer(views)
    try:
        load_preloads(br, views)
        br.draw()
        for i, v in enumerate(views):
            got = br.read(i)
            pre = preload_for(v) if not v.deferred_clear else {}
            ref = oracle_lib.render(v, color=pre.get("color"), depth=pre.get("depth"), triid=pre.get("triid", True))
            assert np.array_equal(got["depth"], ref["depth"]), f"view {i} depth"
            assert np.array_equal(got["triid"], ref["triid"]), f"view {i} triid"
            assert np.array_equal(got["color"], ref["color"]), f"view {i} color"
    finally:
        br.close()


def test_batch_overflow_retry_redraws_only_overflowed_views():
    _torch()
    import cudaraster
    views = mixed_views(8, seed=6)
    refs = [plain(v) if v.deferred_clear else None for v in views]
    br = cudaraster.BatchRenderer(views)
    try:
        for i in (0, 2, 4):                  # tiny queues: subtriangles, bin and tile entries all overflow
            br.batch.reserve(i, 1, 16, 16)
        br.batch.use_graph(False)
        br.draw()
        assert br.batch.info()["retries"] >= 1
        for i, v in enumerate(views):
            if refs[i] is None:
                continue
            got = br.read(i)
            for k in ("color", "depth", "triid"):
                assert np.array_equal(got[k], refs[i][k]), f"view {i} {k}"
        br.batch.use_graph(True)
        br.draw()                            # capacities now fit: no retry, same images
        assert br.batch.info()["retries"] == 0
    finally:
        br.close()


def test_batch_autotune_keeps_images_and_restores_surfaces():
    _torch()
    import cudaraster
    views = mixed_views(10, seed=7)
    pre = {i: preload_for(v) for i, v in enumerate(views) if not v.deferred_clear}
    assert pre, "workload must hold a view drawn without clear"
    refs = [plain(v, pre.get(i)) for i, v in enumerate(views)]
    br = cudaraster.BatchRenderer(views)
    try:
        load_preloads(br, views)
        br.batch.autotune()
        for i in pre:                        # autotuning left non-clearing views untouched
            col, dep, tid = br.surfaces[i]
            assert np.array_equal(col.download(), pre[i]["color"])
            assert np.array_equal(dep.download(), pre[i]["depth"])
        info = br.batch.info()
        assert info["tuned"] == 1 and info["fine_tiles_per_group"] in (2, 4, 8) and info["tuned_gpu_ms"] > 0
        br.draw()
        check_all(br, views, refs, "tuned")
        for i, v in enumerate(views):
            c = br.batch.view_counters(i)
            assert c["num_tris"] == v.num_tris and c["num_subtris"] >= v.num_tris
    finally:
        br.close()


def test_batch_rejects_shared_or_mismatched_surfaces():
    _torch()
    import cudaraster
    from cudaraster import scenes
    sc = scenes.random_tris(100, 64, 64, samples_log2=0, seed=1)
    br = cudaraster.BatchRenderer([sc])
    try:
        v = cudaraster.View()
        col, dep, _ = br.surfaces[0]
        vb, ib = br.buffers[0]
        v.color, v.depth, v.tri_id = col.handle.value, dep.handle.value, None
        v.vertices, v.num_vertices, v.indices, v.num_tris = vb.data_ptr(), sc.vertices.shape[0], ib.data_ptr(), 100
        v.pipe = cudaraster.scene_pipe(sc)
        with pytest.raises(cudaraster.CrError, match="more than one view"):
            cudaraster.Batch(br.cr, [v, v])
        v2 = cudaraster.View.from_buffer_copy(v)
        v2.pipe = cudaraster.pipe_desc(samples_log2=2)
        with pytest.raises(cudaraster.CrError, match="multisampling"):
            cudaraster.Batch(br.cr, [v2])
        empty = cudaraster.Batch(br.cr, [])
        empty.draw()
        empty.close()
    finally:
        br.close()


def test_full_64_view_workload_equals_plain_draws():
    """The 64-view workload at full size (256^2-2048^2, 10k-500k triangles,
    1x/4x, 8.6M triangles): after the autotuner, two graph replays give every
    view exactly the images of a plain draw of it."""
    torch = _torch()
    import cudaraster
    from cudaraster import scenes
    views = scenes.multi_view(64, seed=3)
    br = cudaraster.BatchRenderer(views)
    try:
        br.batch.autotune()
        for rep in range(2):
            br.draw()
            torch.cuda.synchronize()
            for i, v in enumerate(views):
                ref = plain(v)
                got = br.read(i)
                for k in ("color", "depth", "triid"):
                    assert np.array_equal(got[k], ref[k]), f"replay {rep}: view {i} ({v.name}) {k} differs"
    finally:
        br.close()
