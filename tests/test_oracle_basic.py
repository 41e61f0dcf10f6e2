"""Oracle sanity on the demo cube and thread-count independence."""
import numpy as np


def test_cube_faces(oracle_lib):
    from cudaraster import scenes
    s = scenes.cube()
    r = oracle_lib.render(s)
    st = r["stats"]
    # Camera at (0,2,4) sees the +z and +y faces: 4 front-facing triangles,
    # 8 back-facing ones culled (R5).
    assert st["numSubtris"] == 4 and st["culledBackface"] == 8
    ids = np.unique(r["triid"])
    assert set(ids.tolist()) == {0, 1, 8, 9, 0xFFFFFFFF}
    # Clear colour (0.2, 0.4, 0.8, 1) -> ABGR 0xffcc6633 (test/SceneCR.cpp:287).
    assert r["color"][0, 0] == 0xFFCC6633
    assert (r["depth"][r["triid"] == 0xFFFFFFFF] == oracle_lib.clear_depth(1.0)).all()


def test_thread_count_independent(oracle_lib):
    from cudaraster import scenes
    s = scenes.random_tris(20000, 300, 200, samples_log2=2, seed=3, alpha_range=(0.2, 0.8), blend=1,
                           area_range=(1.0, 3000.0))
    a = oracle_lib.render(s, nthreads=1)
    b = oracle_lib.render(s, nthreads=8)
    for k in ("color", "depth", "triid"):
        assert np.array_equal(a[k], b[k]), k
    assert a["stats"] == b["stats"]
