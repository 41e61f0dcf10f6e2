"""The FP64 CPU reference of the differentiable interpolation, pinned
against closed forms and FP64 central finite differences (CPU only).  The
GPU kernels are checked against this reference in test_gpu_diff.py."""
import numpy as np
import pytest


def tri_image_from_oracle(oracle_lib, view):
    sc = view.raster_scene()
    r = oracle_lib.render(sc)
    ids = r["triid"][: view.height, : view.width].astype(np.int64)
    ids[ids == 0xFFFFFFFF] = -1
    return ids


@pytest.fixture(scope="module")
def small_view():
    from cudaraster import scenes
    return scenes.diff_view(3000, 64, 48, channels=5, seed=11, area_range=(2.0, 120.0))


def test_projection_invariant(oracle_lib, small_view):
    """Interpolating the clip coordinates themselves: (sum b x)/(sum b w) is
    the pixel centre's NDC x (same for y) -- the defining property of
    perspective-correct barycentrics."""
    import diff_ref
    v = small_view
    img = tri_image_from_oracle(oracle_lib, v)
    assert (img >= 0).sum() > 1000
    attr = v.pos[:, [0, 1, 3]].astype(np.float64)
    out = diff_ref.interpolate(v.pos, attr, v.tris, img)
    ys, xs = np.nonzero(img >= 0)
    qx = (2.0 * xs + 1.0) / v.width - 1.0
    qy = (2.0 * ys + 1.0) / v.height - 1.0
    o = out[ys, xs]
    assert np.allclose(o[:, 0] / o[:, 2], qx, rtol=0, atol=1e-9)
    assert np.allclose(o[:, 1] / o[:, 2], qy, rtol=0, atol=1e-9)


def test_affine_case_matches_screen_space_solve(oracle_lib):
    """All w = 1: barycentrics are the screen-space ones, from a 2x2 solve."""
    import diff_ref
    from cudaraster import scenes
    v = scenes.diff_view(400, 40, 40, channels=3, seed=12, area_range=(4.0, 200.0))
    pos = v.pos.astype(np.float64).copy()
    pos[:, :2] /= pos[:, 3:4]
    pos[:, 2] /= pos[:, 3]
    pos[:, 3] = 1.0
    view = scenes.DiffView(v.name, v.width, v.height, pos.astype(np.float32), v.attr, v.tris, v.target)
    img = tri_image_from_oracle(oracle_lib, view)
    out = diff_ref.interpolate(view.pos, v.attr, v.tris, img)
    ys, xs = np.nonzero(img >= 0)
    for y, x in list(zip(ys, xs))[::7]:
        i = v.tris[img[y, x]]
        P = view.pos[i, :2].astype(np.float64)
        q = np.array([(2.0 * x + 1.0) / v.width - 1.0, (2.0 * y + 1.0) / v.height - 1.0])
        l12 = np.linalg.solve(np.stack([P[1] - P[0], P[2] - P[0]], axis=1), q - P[0])
        b = np.array([1.0 - l12.sum(), l12[0], l12[1]])
        assert np.allclose(out[y, x], b @ v.attr[i].astype(np.float64), atol=1e-9)


def sample_params(rng, view, img, n):
    used = np.unique(view.tris[img[img >= 0]].ravel())
    out = []
    for _ in range(n):
        vtx = int(rng.choice(used))
        if rng.random() < 0.5:
            out.append(("attr", vtx, int(rng.integers(view.attr.shape[1]))))
        else:
            out.append(("pos", vtx, int(rng.choice([0, 1, 3]))))
    return out


def test_backward_matches_finite_differences(oracle_lib, small_view):
    import diff_ref
    v = small_view
    img = tri_image_from_oracle(oracle_lib, v)
    out = diff_ref.interpolate(v.pos, v.attr, v.tris, img)
    gp, ga = diff_ref.backward(v.pos, v.attr, v.tris, img, 2.0 * (out - v.target))
    rng = np.random.default_rng(13)
    params = sample_params(rng, v, img, 200)
    an = np.array([gp[vtx, c] if kind == "pos" else ga[vtx, c] for kind, vtx, c in params])
    # The bar of the workload: FP64 central differences with h = 1e-4, 1e-4 relative.
    fd = diff_ref.finite_difference(v.pos, v.attr, v.tris, img, v.target, params, h=1e-4)
    scale = np.abs(fd).max()
    err = np.abs(an - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    assert err.max() < 1e-4, (err.max(), params[int(err.argmax())])
    # Stricter: Richardson extrapolation removes the O(h^2) truncation term.
    fd2 = diff_ref.finite_difference(v.pos, v.attr, v.tris, img, v.target, params, h=5e-5)
    rich = (4.0 * fd2 - fd) / 3.0
    err = np.abs(an - rich) / np.maximum(np.abs(rich), 1e-3 * scale)
    assert err.max() < 1e-7, (err.max(), params[int(err.argmax())])


def test_depth_does_not_enter_barycentrics(oracle_lib, small_view):
    import diff_ref
    v = small_view
    img = tri_image_from_oracle(oracle_lib, v)
    out = diff_ref.interpolate(v.pos, v.attr, v.tris, img)
    gp, _ = diff_ref.backward(v.pos, v.attr, v.tris, img, 2.0 * (out - v.target))
    assert np.all(gp[:, 2] == 0.0)
    pos2 = v.pos.copy()
    pos2[:, 2] *= 0.5
    assert np.array_equal(diff_ref.interpolate(pos2, v.attr, v.tris, img), out)
