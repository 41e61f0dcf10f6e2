"""Differentiable interpolation on the GPU (cr_interpolate /
cr_interpolate_backward through torch.autograd): triangle ids exact against
the oracle, forward against the FP64 reference, gradients against the FP64
reference gradient and FP64 central finite differences (h = 1e-4, 1e-4
relative, on sampled parameters)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def run_view(view):
    torch = _torch()
    from cudaraster.diff import DiffRasterizer
    dr = DiffRasterizer(view.width, view.height)
    try:
        pos = torch.from_numpy(view.pos).cuda().requires_grad_(True)
        attr = torch.from_numpy(view.attr).cuda().requires_grad_(True)
        tris = torch.from_numpy(view.tris).cuda()
        target = torch.from_numpy(view.target).cuda()
        ids = dr.rasterize(pos, tris)
        out = dr.render(pos, attr, tris, ids)
        loss = ((out - target) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
        return {"ids": ids.cpu().numpy(), "out": out.detach().cpu().numpy(), "gp": pos.grad.cpu().numpy(),
                "ga": attr.grad.cpu().numpy(), "loss": float(loss.detach())}
    finally:
        dr.close()


def oracle_ids(oracle_lib, view):
    r = oracle_lib.render(view.raster_scene())
    ids = r["triid"][: view.height, : view.width].astype(np.int64)
    ids[ids == 0xFFFFFFFF] = -1
    return ids


def check_view(oracle_lib, view, n_fd, seed):
    import diff_ref
    got = run_view(view)
    ids = oracle_ids(oracle_lib, view)
    assert np.array_equal(got["ids"], ids), "triangle ids differ from the oracle"
    ref_out = diff_ref.interpolate(view.pos, view.attr, view.tris, ids)
    assert np.abs(got["out"] - ref_out).max() < 1e-6
    gp, ga = diff_ref.backward(view.pos, view.attr, view.tris, ids, 2.0 * (ref_out - view.target))
    for name, g, r in (("pos", got["gp"], gp), ("attr", got["ga"], ga)):
        scale = np.abs(r).max()
        assert np.abs(g - r).max() <= 1e-6 * scale + 1e-30, f"{name} gradient vs FP64 reference"
    assert np.all(got["gp"][:, 2] == 0.0)
    rng = np.random.default_rng(seed)
    used = np.unique(view.tris[ids[ids >= 0]].ravel())
    params = []
    for _ in range(n_fd):
        v = int(rng.choice(used))
        params.append(("attr", v, int(rng.integers(view.attr.shape[1]))) if rng.random() < 0.5
                      else ("pos", v, int(rng.choice([0, 1, 3]))))
    # FP64 central differences with h = 1e-4.  Near the near plane (w < 1) a
    # sub-pixel triangle is not much larger than h in clip units, and there
    # the O(h^2) truncation term of the difference itself exceeds 1e-4; such
    # parameters are recognised by D(h) and D(h/2) disagreeing by more than
    # 1e-5 and are checked against h = 1e-6 instead (all parameters are).
    fd = diff_ref.finite_difference(view.pos, view.attr, view.tris, ids, view.target, params, h=1e-4)
    fd_half = diff_ref.finite_difference(view.pos, view.attr, view.tris, ids, view.target, params, h=5e-5)
    fd_fine = diff_ref.finite_difference(view.pos, view.attr, view.tris, ids, view.target, params, h=1e-6)
    an = np.array([got["gp"][v, c] if k == "pos" else got["ga"][v, c] for k, v, c in params])
    floor = 1e-3 * np.abs(fd_fine).max()

    def rel(a, b):
        return np.abs(a - b) / np.maximum(np.abs(b), floor)

    converged = rel(fd_half, fd) <= 1e-5
    assert converged.mean() > 0.85, converged.mean()
    err = rel(an, fd)[converged]
    assert err.max() < 1e-4, err.max()
    err = rel(an, fd_fine)
    assert err.max() < 1e-4, (err.max(), params[int(err.argmax())])
    got["fd_converged"] = int(converged.sum())
    return got


def test_small_view(oracle_lib):
    from cudaraster import scenes
    check_view(oracle_lib, scenes.diff_view(20000, 128, 96, channels=8, seed=21), n_fd=300, seed=1)


def test_full_size_view(oracle_lib):
    """One view of the differentiable workload at full size: 200k triangles,
    512x512, 8 channels; 1000 sampled parameters against finite differences."""
    from cudaraster import scenes
    check_view(oracle_lib, scenes.diff_view(200_000, 512, 512, channels=8, seed=4), n_fd=1000, seed=2)


def test_gradients_accumulate_over_views_independently():
    """Two rasterizers (resolutions) in one autograd graph: gradients add."""
    torch = _torch()
    from cudaraster import scenes
    from cudaraster.diff import DiffRasterizer
    v = scenes.diff_view(5000, 96, 96, channels=4, seed=31)
    a, b = DiffRasterizer(96, 96), DiffRasterizer(64, 64)
    try:
        pos = torch.from_numpy(v.pos).cuda().requires_grad_(True)
        attr = torch.from_numpy(v.attr).cuda().requires_grad_(True)
        tris = torch.from_numpy(v.tris).cuda()
        oa = a.render(pos, attr, tris)
        ob = b.render(pos, attr, tris)
        (oa.square().sum() + ob.square().sum()).backward()
        g_both = attr.grad.clone()
        attr.grad = None
        pos.grad = None
        a.render(pos, attr, tris).square().sum().backward()
        g_a = attr.grad.clone()
        attr.grad = None
        b.render(pos, attr, tris).square().sum().backward()
        assert torch.allclose(g_both, g_a + attr.grad, rtol=1e-6, atol=1e-6)
    finally:
        a.close()
        b.close()


def test_render_views_matches_single_views():
    """render_views (one autograd call for several views, one DiffRasterizer)
    gives the per-view images, ids and gradients of render() on each view,
    including views of different triangle counts."""
    torch = _torch()
    from cudaraster import scenes
    from cudaraster.diff import DiffRasterizer
    views = [scenes.diff_view(seed=40 + k, num_tris=3000 + 2000 * k, width=96, height=80) for k in range(4)]
    singles = [run_view(v) for v in views]
    dr = DiffRasterizer(96, 80)
    try:
        pos = [torch.from_numpy(v.pos).cuda().requires_grad_(True) for v in views]
        attr = [torch.from_numpy(v.attr).cuda().requires_grad_(True) for v in views]
        tris = [torch.from_numpy(v.tris).cuda() for v in views]
        target = torch.stack([torch.from_numpy(v.target).cuda() for v in views])
        out, ids = dr.render_views(pos, attr, tris)
        ((out - target) ** 2).sum().backward()
        torch.cuda.synchronize()
        for k, s in enumerate(singles):
            assert np.array_equal(ids[k].cpu().numpy(), s["ids"]), f"view {k} ids"
            assert np.array_equal(out[k].detach().cpu().numpy(), s["out"]), f"view {k} image"
            for name, g, r in (("pos", pos[k].grad, s["gp"]), ("attr", attr[k].grad, s["ga"])):
                g = g.cpu().numpy()
                assert np.abs(g - r).max() <= 1e-6 * np.abs(r).max() + 1e-30, f"view {k} {name} gradient"
    finally:
        dr.close()


def test_render_views_full_size_32_views(oracle_lib):
    """The differentiable workload at full size (32 views, 200k triangles,
    512x512, 8 channels) through one render_views call: every view's ids
    equal the oracle's, its image the FP64 interpolation and its gradients
    the FP64 reference gradients of the summed L2 loss (1e-6 of the largest,
    as for single views)."""
    torch = _torch()
    from concurrent.futures import ThreadPoolExecutor
    import diff_ref
    from cudaraster import scenes
    from cudaraster.diff import DiffRasterizer
    views = scenes.diff_views(32)
    dr = DiffRasterizer(512, 512)
    try:
        pos = [torch.from_numpy(v.pos).cuda().requires_grad_(True) for v in views]
        attr = [torch.from_numpy(v.attr).cuda().requires_grad_(True) for v in views]
        tris = [torch.from_numpy(v.tris).cuda() for v in views]
        target = torch.stack([torch.from_numpy(v.target).cuda() for v in views])
        out, ids = dr.render_views(pos, attr, tris)
        ((out - target) ** 2).sum().backward()
        torch.cuda.synchronize()
        got = [(ids[k].cpu().numpy(), out[k].detach().cpu().numpy(), pos[k].grad.cpu().numpy(),
                attr[k].grad.cpu().numpy()) for k in range(len(views))]
    finally:
        dr.close()
    ref_ids = [oracle_ids(oracle_lib, v) for v in views]

    def reference(k):
        v, i = views[k], ref_ids[k]
        ref_out = diff_ref.interpolate(v.pos, v.attr, v.tris, i)
        return (ref_out,) + tuple(diff_ref.backward(v.pos, v.attr, v.tris, i, 2.0 * (ref_out - v.target)))

    with ThreadPoolExecutor(8) as ex:
        refs = list(ex.map(reference, range(len(views))))
    for k, ((g_ids, g_out, g_pos, g_attr), (r_out, r_pos, r_attr)) in enumerate(zip(got, refs)):
        assert np.array_equal(g_ids, ref_ids[k]), f"view {k} ids"
        assert np.abs(g_out - r_out).max() < 1e-6, f"view {k} image"
        for name, g, r in (("pos", g_pos, r_pos), ("attr", g_attr, r_attr)):
            assert np.abs(g - r).max() <= 1e-6 * np.abs(r).max() + 1e-30, f"view {k} {name} gradient"


def test_render_views_mixed_with_single_views(oracle_lib):
    """render_views draws its views' visibility as one batch kept for the
    next call with the same view sizes: single-view draws between such calls
    keep their own state, a call with other view sizes rebuilds the batch,
    and every view's ids equal the oracle's throughout."""
    torch = _torch()
    from cudaraster import scenes
    from cudaraster.diff import DiffRasterizer
    views_a = [scenes.diff_view(num_tris=3000, width=96, height=80, channels=3, seed=70 + k) for k in range(3)]
    views_b = [scenes.diff_view(num_tris=1500 + 700 * k, width=96, height=80, channels=3, seed=80 + k)
               for k in range(2)]
    ref = {id(v): oracle_ids(oracle_lib, v) for v in views_a + views_b}
    dr = DiffRasterizer(96, 80)
    try:
        def tensors(vs):
            return ([torch.from_numpy(v.pos).cuda() for v in vs], [torch.from_numpy(v.attr).cuda() for v in vs],
                    [torch.from_numpy(v.tris).cuda() for v in vs])
        for vs in (views_a, views_b, views_a, views_a):
            pos, attr, tris = tensors(vs)
            _, ids = dr.render_views(pos, attr, tris)
            for k, v in enumerate(vs):
                assert np.array_equal(ids[k].cpu().numpy(), ref[id(v)]), "render_views ids differ from the oracle"
            # A single view drawn between batched calls.
            single = dr.rasterize(pos[-1], tris[-1])
            assert np.array_equal(single.cpu().numpy(), ref[id(vs[-1])]), "single view after render_views"
    finally:
        dr.close()


def test_render_views_with_an_empty_view(oracle_lib):
    """A view without triangles in a render_views call: empty ids, a zero
    image and zero gradients for it; the other views as drawn alone."""
    torch = _torch()
    from cudaraster import scenes
    from cudaraster.diff import DiffRasterizer
    v = scenes.diff_view(num_tris=2000, width=64, height=48, channels=2, seed=90)
    dr = DiffRasterizer(64, 48)
    try:
        pos = [torch.from_numpy(v.pos).cuda().requires_grad_(True), torch.zeros((0, 4), device="cuda",
                                                                                 requires_grad=True)]
        attr = [torch.from_numpy(v.attr).cuda().requires_grad_(True), torch.zeros((0, 2), device="cuda",
                                                                                   requires_grad=True)]
        tris = [torch.from_numpy(v.tris).cuda(), torch.zeros((0, 3), dtype=torch.int32, device="cuda")]
        out, ids = dr.render_views(pos, attr, tris)
        out.sum().backward()
        torch.cuda.synchronize()
        assert np.array_equal(ids[0].cpu().numpy(), oracle_ids(oracle_lib, v))
        assert bool((ids[1] == -1).all()) and float(out[1].detach().abs().max()) == 0.0
        assert pos[1].grad.shape == (0, 4) and attr[1].grad.shape == (0, 2)
        assert float(pos[0].grad.abs().sum()) > 0.0
    finally:
        dr.close()


def test_async_visibility_overflow_is_reported_then_recovers():
    """After a small sizing view, a much larger view overflows the reserved
    queues in its asynchronous draw: sync() raises instead of returning
    incomplete ids, and re-rendering the view (queues now grown) matches a
    fresh rasterizer."""
    torch = _torch()
    from cudaraster import CrError, scenes
    from cudaraster.diff import DiffRasterizer
    small = scenes.diff_view(seed=61, num_tris=200, width=128, height=96)
    big = scenes.diff_view(seed=62, num_tris=60000, width=128, height=96, area_range=(4.0, 2000.0))
    expect = run_view(big)
    dr = DiffRasterizer(128, 96)
    try:
        def render(v):
            pos = torch.from_numpy(v.pos).cuda().requires_grad_(True)
            attr = torch.from_numpy(v.attr).cuda().requires_grad_(True)
            tris = torch.from_numpy(v.tris).cuda()
            return dr.render(pos, attr, tris)
        render(small)                      # synchronous sizing draw
        render(big)                        # asynchronous, exceeds the reserve
        with pytest.raises(CrError, match="capacities"):
            dr.sync()
        out = render(big)
        dr.sync()
        torch.cuda.synchronize()
        assert np.array_equal(out.detach().cpu().numpy(), expect["out"])
    finally:
        dr.close()


def test_render_uses_the_given_ids_not_the_last_draw():
    """rasterize(A), rasterize(B), then render(A, ids_A): the image and the
    gradients follow A's visibility although B was drawn last."""
    torch = _torch()
    import diff_ref
    from cudaraster import scenes
    from cudaraster.diff import DiffRasterizer
    a, b = (scenes.diff_view(4000, 96, 80, channels=4, seed=s) for s in (71, 72))
    dr = DiffRasterizer(96, 80)
    try:
        ta = {k: torch.from_numpy(getattr(a, k)).cuda() for k in ("pos", "attr", "tris")}
        tb = {k: torch.from_numpy(getattr(b, k)).cuda() for k in ("pos", "tris")}
        ids_a = dr.rasterize(ta["pos"], ta["tris"])
        dr.rasterize(tb["pos"], tb["tris"])
        pos = ta["pos"].clone().requires_grad_(True)
        attr = ta["attr"].clone().requires_grad_(True)
        out = dr.render(pos, attr, ta["tris"], ids_a)
        out.square().sum().backward()
        torch.cuda.synchronize()
        ids = ids_a.cpu().numpy().astype(np.int64)
        ref = diff_ref.interpolate(a.pos, a.attr, a.tris, ids)
        assert np.abs(out.detach().cpu().numpy() - ref).max() < 1e-6
        gp, ga = diff_ref.backward(a.pos, a.attr, a.tris, ids, 2.0 * ref)
        assert np.abs(attr.grad.cpu().numpy() - ga).max() <= 1e-6 * np.abs(ga).max()
        assert np.abs(pos.grad.cpu().numpy() - gp).max() <= 1e-6 * np.abs(gp).max()
    finally:
        dr.close()


@pytest.mark.parametrize("channels,offset", [(3, 0), (5, 0), (8, 1), (4, 3)])
def test_scalar_channel_path(channels, offset):
    """Channel counts that are not a multiple of 4, and attribute / output
    storage that is not 16-byte aligned, take the forward's per-channel path:
    forward and gradients against the FP64 reference."""
    torch = _torch()
    import diff_ref
    from cudaraster import scenes
    from cudaraster.diff import DiffRasterizer
    v = scenes.diff_view(3000, 80, 64, channels=channels, seed=80 + channels)
    dr = DiffRasterizer(80, 64)
    try:
        pos = torch.from_numpy(v.pos).cuda().requires_grad_(True)
        flat = torch.zeros(offset + v.attr.size, dtype=torch.float32, device="cuda")
        flat[offset:] = torch.from_numpy(v.attr.reshape(-1)).cuda()
        attr = flat[offset:].view(v.attr.shape).detach().requires_grad_(True)
        assert attr.data_ptr() % 16 == (4 * offset) % 16
        tris = torch.from_numpy(v.tris).cuda()
        target = torch.from_numpy(v.target).cuda()
        out = dr.render(pos, attr, tris)
        ids = dr.rasterize(pos, tris).cpu().numpy().astype(np.int64)
        ((out - target) ** 2).sum().backward()
        torch.cuda.synchronize()
        ref = diff_ref.interpolate(v.pos, v.attr, v.tris, ids)
        assert np.abs(out.detach().cpu().numpy() - ref).max() < 1e-6
        gp, ga = diff_ref.backward(v.pos, v.attr, v.tris, ids, 2.0 * (ref - v.target))
        assert np.abs(attr.grad.cpu().numpy() - ga).max() <= 1e-6 * np.abs(ga).max()
        assert np.abs(pos.grad.cpu().numpy() - gp).max() <= 1e-6 * np.abs(gp).max()
    finally:
        dr.close()


def test_interpolation_rejects_wrong_buffers():
    """Wrong dtypes, shapes or layouts are refused before a kernel sees them."""
    torch = _torch()
    from cudaraster import CrError, scenes
    from cudaraster.diff import DiffRasterizer
    v = scenes.diff_view(500, 32, 32, channels=4, seed=90)
    dr = DiffRasterizer(32, 32)
    try:
        pos = torch.from_numpy(v.pos).cuda()
        attr = torch.from_numpy(v.attr).cuda()
        tris = torch.from_numpy(v.tris).cuda()
        ids = dr.rasterize(pos, tris)
        bad = [(pos, attr.to(torch.int32), tris, ids), (pos, attr, tris.to(torch.int64), ids),
               (pos.double(), attr, tris, ids), (pos, attr, tris, ids.to(torch.float32)),
               (pos, attr[:-1], tris, ids), (pos, attr, tris, ids[:, :16].contiguous())]
        for p, a, t, i in bad:
            with pytest.raises(CrError):
                dr.render(p, a, t, i)
    finally:
        dr.close()


def test_deterministic_backward_is_bit_reproducible():
    """deterministic=True: two backward passes of a full-size view give
    bit-identical gradients (fixed-point accumulation), equal to the FP64
    reference within the same bound as the FP32-atomic backward."""
    torch = _torch()
    import diff_ref
    from cudaraster import scenes
    from cudaraster.diff import DiffRasterizer
    v = scenes.diff_view(200_000, 512, 512, channels=8, seed=4)
    dr = DiffRasterizer(512, 512, deterministic=True)
    try:
        pos = torch.from_numpy(v.pos).cuda().requires_grad_(True)
        attr = torch.from_numpy(v.attr).cuda().requires_grad_(True)
        tris = torch.from_numpy(v.tris).cuda()
        target = torch.from_numpy(v.target).cuda()
        ids = dr.rasterize(pos, tris)
        grads = []
        for _ in range(2):
            pos.grad = None
            attr.grad = None
            ((dr.render(pos, attr, tris, ids) - target) ** 2).sum().backward()
            torch.cuda.synchronize()
            grads.append((pos.grad.cpu().numpy().copy(), attr.grad.cpu().numpy().copy()))
        assert np.array_equal(grads[0][0].view(np.uint32), grads[1][0].view(np.uint32))
        assert np.array_equal(grads[0][1].view(np.uint32), grads[1][1].view(np.uint32))
        idn = ids.cpu().numpy().astype(np.int64)
        ref = diff_ref.interpolate(v.pos, v.attr, v.tris, idn)
        gp, ga = diff_ref.backward(v.pos, v.attr, v.tris, idn, 2.0 * (ref - v.target))
        assert np.abs(grads[0][0] - gp).max() <= 1e-6 * np.abs(gp).max()
        assert np.abs(grads[0][1] - ga).max() <= 1e-6 * np.abs(ga).max()
    finally:
        dr.close()
