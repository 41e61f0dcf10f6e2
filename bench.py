"""Throughput of the four-stage rasterization path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c2|c0]

A step is one CudaRaster::drawTriangles() of the workload -- deferred clear,
setup, bin, coarse, fine -- with the reference's synchronous semantics (the
call returns when the frame is complete and the overflow counters have been
read back).  Inputs are resident in HBM before timing starts; the draws
are timed by a native loop over the C ABI.  Prints one
JSON line: Mtris/s as `value`, Gfragments/s beside it, the roofline of the
dominant kernel (timed with HIP events on the draw stream) and the CPU
oracle's throughput on the same workload as `cpu_baseline`.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cudaraster-linux_amd"))

METRIC = "Mtris/s + Gfragments/s at 2048² 4xMSAA; batched/plain views/s ratio"
HBM_PEAK_GBS = 8000.0


def workload(name):
    from cudaraster import scenes
    if name == "c1":
        return scenes.config(1), "1M random triangles, 2048x2048, 4x MSAA, depth+lerp, Gouraud, replace"
    if name == "c2":
        return scenes.config(2), "4M sub-pixel + 1024 near-full-screen triangles, 2048x2048, 8x MSAA, src-over"
    if name == "c0":
        return scenes.cube(), "12-triangle cube, 1024x768, 1x"
    raise SystemExit(f"unknown workload {name}")


def fine_algorithmic_bytes(scene, visible_subtris, triid):
    """Compulsory HBM bytes of one fine-raster launch: every texel of the
    colour, depth (and triangle-id) surfaces written once (deferred clear:
    nothing read), and per visible subtriangle its 16 B header, 64 B
    plane-equation record and the three 16 B colour varyings read once."""
    hr, tw = scene.surface_shape
    surfaces = hr * tw * (12 if triid else 8)
    return surfaces + visible_subtris * (16 + 64 + 3 * 16)


def setup_algorithmic_bytes(scene, c, visible_subtris):
    """Setup + binning: indices (12 B) and three clip positions (16 B each)
    read per triangle; per visible subtriangle a 16 B header, a 64 B data
    record and its 4 B source index written; 16 B per bin-queue entry (slot
    and snapped vertices) and 8 B per (bin, chunk) count/offset pair written."""
    chunks = (scene.num_tris + 1023) // 1024
    bins = ((scene.width + 127) // 128) * ((scene.height + 127) // 128)
    return (scene.num_tris * (12 + 3 * 16) + visible_subtris * (16 + 64 + 4)
            + c["bin_entries"] * 16 + bins * chunks * 8)


def coarse_algorithmic_bytes(scene, c, visible_subtris):
    """Coarse: every 16 B bin-queue entry and the per-(bin, chunk)
    counts/offsets read, every tile-list entry (4 B) and per-tile start/count
    (8 B) written."""
    chunks = (scene.num_tris + 1023) // 1024
    bins = ((scene.width + 127) // 128) * ((scene.height + 127) // 128)
    tiles = ((scene.width + 7) // 8) * ((scene.height + 7) // 8)
    return c["bin_entries"] * 16 + bins * chunks * 8 + c["tile_list_entries"] * 4 + tiles * 8


def measured_traffic(kernel, workload_key, triid):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE counters and scaled by a copy-kernel calibration, as recorded
    in a profile summary under profiles/ (a line
    `TRAFFIC <kernel> <workload> triid=<0|1> bytes=<n>`), or None."""
    pdir = os.path.join(REPO, "profiles")
    if not os.path.isdir(pdir):
        return None
    want = f"TRAFFIC {kernel} {workload_key} triid={int(triid)} "
    for name in sorted(os.listdir(pdir)):
        if not name.endswith((".txt", ".md")):
            continue
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                if line.startswith(want):
                    return int(line.split("bytes=")[1].split()[0])
    return None


def measured_issue(kernel, workload_key, triid):
    """VALU instructions per launch of `kernel` (rocprofv3 SQ_INSTS_VALU) and
    the VALU issue fraction they imply (x 4 cycles per wave64 instruction over
    1024 SIMDs at 2.4 GHz and the kernel's traced duration), as recorded in a
    profile summary under profiles/ (a line `VALU <kernel> <workload>
    triid=<0|1> insts=<n> us=<t> issue_frac=<f>`), or None."""
    pdir = os.path.join(REPO, "profiles")
    if not os.path.isdir(pdir):
        return None
    want = f"VALU {kernel} {workload_key} triid={int(triid)} "
    for name in sorted(os.listdir(pdir)):
        if not name.endswith((".txt", ".md")):
            continue
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                if line.startswith(want):
                    kv = dict(t.split("=", 1) for t in line.split()[4:])
                    return {"insts_per_launch": int(kv["insts"]), "frac": float(kv["issue_frac"]),
                            "profiled_us": float(kv["us"]), "source": f"profiles/{name}"}
    return None


def copy_calibration():
    """A plain device-to-device copy of known size (256 MiB, three times):
    the reference point for scaling FETCH_SIZE / WRITE_SIZE readings."""
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "copies": 3}))


def frame_algorithmic_bytes(scene, triid):
    """B = V*stride + 12*T + S_r*N*(4+4[+4]): inputs read once, surfaces
    written once (deferred clear needs no read)."""
    hr, tw = scene.surface_shape
    return scene.vertices.nbytes + 12 * scene.num_tris + hr * tw * (12 if triid else 8)


def multi_view_ratio(reps=5, triid=True):
    """Batched multi-view workload: 64 seeded views (256^2..2048^2, 10k-500k
    triangles, 1x or 4x MSAA).  Times `reps` passes of 64 sequential plain
    draws (each synchronous, the reference's usage: set surfaces, pipe,
    buffers, deferredClear, drawTriangles) against `reps` batched draws
    (autotuned, graph replay, synchronous), after checking that every view's
    colour, depth and triangle-id surfaces are identical between the two."""
    import numpy as np
    import torch
    import cudaraster
    from cudaraster import scenes
    views = scenes.multi_view(64, seed=3)
    br = cudaraster.BatchRenderer(views, triangle_ids=triid)
    cr = cudaraster.CudaRaster().init()
    plain_surf = []
    for sc in views:
        n = sc.num_samples
        plain_surf.append((cudaraster.CudaSurface((sc.width, sc.height), cudaraster.FORMAT_RGBA8, n),
                           cudaraster.CudaSurface((sc.width, sc.height), cudaraster.FORMAT_DEPTH32, n),
                           cudaraster.CudaSurface((sc.width, sc.height), cudaraster.FORMAT_TRIANGLE_ID, n)
                           if triid else None))
    pipes = [cudaraster.scene_pipe(sc) for sc in views]

    def plain_pass():
        for k, sc in enumerate(views):
            col, dep, tid = plain_surf[k]
            vb, ib = br.buffers[k]
            cr.setSurfaces(col, dep)
            cr.setTriangleIdSurface(tid)
            cr.setPixelPipe(pipes[k])
            cr.deferredClear(sc.clear_color, sc.clear_depth)
            cr.setVertexBuffer(vb, 0, sc.vertices.shape[0])
            cr.setIndexBuffer(ib, 0, sc.num_tris)
            cr.drawTriangles()

    plain_pass()
    # GPU time of the plain draws: per view, the stage events of its draw
    # (reset, setup+bin, coarse, fine) summed over the 64 views.
    plain_gpu_ms = 0.0
    for k, sc in enumerate(views):
        col, dep, tid = plain_surf[k]
        vb, ib = br.buffers[k]
        cr.setSurfaces(col, dep)
        cr.setTriangleIdSurface(tid)
        cr.setPixelPipe(pipes[k])
        cr.deferredClear(sc.clear_color, sc.clear_depth)
        cr.setVertexBuffer(vb, 0, sc.vertices.shape[0])
        cr.setIndexBuffer(ib, 0, sc.num_tris)
        cr.drawTriangles()
        st = cr.getStats()
        plain_gpu_ms += (st["setupTime"] + st["binTime"] + st["coarseTime"] + st["fineTime"]) * 1e3
    br.batch.autotune()
    br.draw()
    identical = True
    for k in range(len(views)):
        got = br.read(k)
        col, dep, tid = plain_surf[k]
        identical &= bool(np.array_equal(got["color"], col.download()) and np.array_equal(got["depth"], dep.download()))
        if triid:
            identical &= bool(np.array_equal(got["triid"], tid.download()))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain_pass()
    torch.cuda.synchronize()
    t_plain_py = (time.perf_counter() - t0) / reps
    # The same sequential draws from a native loop over the C ABI (no
    # interpreter between the calls): the plain leg of the ratio.
    import ctypes
    bench_lib = _bench_lib()
    cviews = (cudaraster.View * len(views))()
    for k, sc in enumerate(views):
        col, dep, tid = plain_surf[k]
        vb, ib = br.buffers[k]
        v = cviews[k]
        v.color, v.depth, v.tri_id = col.handle.value, dep.handle.value, (tid.handle.value if tid else None)
        v.vertices, v.num_vertices = vb.data_ptr(), sc.vertices.shape[0]
        v.indices, v.num_tris = ib.data_ptr(), sc.num_tris
        v.deferred_clear = 1 if sc.deferred_clear else 0
        v.pipe = pipes[k]
        v.clear_color[:] = [float(x) for x in sc.clear_color]
        v.clear_depth = float(sc.clear_depth)
    st, bad = ctypes.c_int32(), ctypes.c_int32()
    t_plain = bench_lib.crbench_plain_draws(cr.ctx, cviews, len(views), reps, ctypes.byref(st), ctypes.byref(bad))
    if t_plain < 0:
        raise RuntimeError(f"native plain draws failed at view {bad.value}: status {st.value}")
    gpu_ms = []
    t0 = time.perf_counter()
    for _ in range(reps):
        br.draw()
        gpu_ms.append(br.batch.info()["last_gpu_ms"])
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    info = br.batch.info()
    out = {"views": len(views), "triangles": int(sum(s.num_tris for s in views)),
           "plain_views_per_s": len(views) / t_plain, "batched_views_per_s": len(views) / t_batch,
           "ratio": t_plain / t_batch, "plain_ms_per_pass": t_plain * 1e3, "batched_ms_per_pass": t_batch * 1e3,
           "plain_loop": "native C loop over the C ABI (cudaraster-linux_amd/bench_native/plain_draws.cpp)",
           "plain_ms_per_pass_python": t_plain_py * 1e3, "ratio_python_plain_loop": t_plain_py / t_batch,
           "plain_gpu_ms": plain_gpu_ms, "batched_gpu_ms": statistics.median(gpu_ms),
           "gpu_only_ratio": plain_gpu_ms / statistics.median(gpu_ms), "identical_images": identical,
           "fine_tiles_per_group": info["fine_tiles_per_group"], "uses_graph": bool(info["uses_graph"]),
           "reserved_bytes": info["reserved_bytes"]}
    br.close()
    cr.close()
    for trio in plain_surf:
        for s in trio:
            if s is not None:
                s.close()
    return out


def differentiable_views(reps=3):
    """Differentiable workload: 32 views at 512x512, 200k triangles each, 8
    FP32 channels; the exact visibility of all views in one batched draw,
    per view the interpolation forward, an L2 loss against the views'
    targets and the backward pass to every view's positions and attributes
    (one torch autograd call over the HIP kernels of all views:
    DiffRasterizer.render_views)."""
    import torch
    from cudaraster import scenes
    from cudaraster.diff import DiffRasterizer
    views = scenes.diff_views(32)
    dr = DiffRasterizer(512, 512)
    pos = [torch.from_numpy(v.pos).cuda().requires_grad_(True) for v in views]
    attr = [torch.from_numpy(v.attr).cuda().requires_grad_(True) for v in views]
    tris = [torch.from_numpy(v.tris).cuda() for v in views]
    target = torch.stack([torch.from_numpy(v.target).cuda() for v in views])

    def one_pass():
        for t in pos + attr:
            t.grad = None
        out, _ = dr.render_views(pos, attr, tris)
        ((out - target) ** 2).sum().backward()

    one_pass()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        one_pass()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    dr.close()
    return {"views": len(views), "triangles_per_view": int(views[0].tris.shape[0]), "resolution": 512,
            "channels": int(views[0].attr.shape[1]), "ms_per_pass": t * 1e3, "views_per_s": len(views) / t,
            "steps": "batched visibility draw of the views; per view interpolate forward; L2 loss; "
                     "backward to positions and attributes",
            "api": "DiffRasterizer.render_views (one autograd call for the 32 views)"}


def vertex_shader(num_vertices=3_000_000, reps=20):
    """The passthrough vertex shader over the 1M-triangle workload's vertex
    count (3 vertices per triangle): model positions (12 B) in, the clip
    position part of 32-byte vertex records (16 B) out, timed with HIP events
    on the stream it is launched on.  HBM-bound: 28 algorithmic bytes per
    vertex."""
    import numpy as np
    import torch
    import cudaraster
    rng = np.random.default_rng(21)
    p = torch.from_numpy((rng.standard_normal((num_vertices, 3)) * 20.0).astype(np.float32)).cuda()
    out = torch.zeros((num_vertices, 8), dtype=torch.float32, device="cuda")
    m = rng.standard_normal((4, 4)).astype(np.float32)
    cr = cudaraster.CudaRaster().init()
    stream = torch.cuda.current_stream()
    for _ in range(3):
        cr.vertexShaderPassthrough(p, m, out, out_stride=32, stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        cr.vertexShaderPassthrough(p, m, out, out_stride=32, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    cr.close()
    us = e0.elapsed_time(e1) / reps * 1e3
    bytes_ = 28 * num_vertices
    return {"vertices": num_vertices, "us": us, "algorithmic_bytes": bytes_,
            "GBps": bytes_ / (us * 1e-6) / 1e9, "peak_GBps": 8000.0}


def resolve_leg(reps=20):
    """Box MSAA resolve of a 2048x2048 4x RGBA8 surface (the 1M-triangle
    workload's colour surface) into 2048x2048 pixels, HIP events on the
    stream it runs on.  HBM-bound: 16 bytes read and 4 written per pixel."""
    import numpy as np
    import torch
    import cudaraster
    surf = cudaraster.CudaSurface((2048, 2048), cudaraster.FORMAT_RGBA8, 4)
    surf.upload(np.random.default_rng(22).integers(0, 2 ** 32, size=surf.shape, dtype=np.uint64).astype(np.uint32))
    out = torch.empty((2048, 2048), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(3):
        surf.resolve(out, stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        surf.resolve(out, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    surf.close()
    us = e0.elapsed_time(e1) / reps * 1e3
    bytes_ = 20 * 2048 * 2048
    return {"pixels": 2048 * 2048, "samples": 4, "us": us, "algorithmic_bytes": bytes_,
            "GBps": bytes_ / (us * 1e-6) / 1e9, "peak_GBps": 8000.0}


def host_cpu():
    """The host this process runs on: hardware threads, the threads it may
    use, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": os.cpu_count(), "affinity": allowed, "model": model}


def cpu_baseline(scene, runs=5, single_runs=3, max_threads=16):
    """The CPU oracle (plain C + pthreads, band-parallel) drawing the same
    frame on the host cores: all the threads this process may use, capped at
    the one-GPU host share of 16, median of `runs` after a warm-up; and one
    thread, median of `single_runs`."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    host = host_cpu()
    threads = max(1, min(host["affinity"], max_threads))
    oracle.render(scene, nthreads=threads)   # warm-up
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle.render(scene, nthreads=threads)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    t1 = []
    for _ in range(single_runs):
        t0 = time.perf_counter()
        oracle.render(scene, nthreads=1)
        t1.append(time.perf_counter() - t0)
    out = {"value": scene.num_tris / t / 1e6, "unit": "Mtris/s", "cores": threads, "kind": "port",
           "sample": f"whole {scene.num_tris}-triangle frame, median of {runs} runs after 1 warm-up "
                     f"({threads} threads)" + (f" and of {single_runs} runs on 1 thread" if single_runs else ""),
           "seconds_per_frame": t, "host": host}
    if t1:
        t_single = statistics.median(t1)
        out["single_thread"] = {"value": scene.num_tris / t_single / 1e6, "seconds_per_frame": t_single}
    return out


def _bench_lib():
    import ctypes
    lib = ctypes.CDLL(os.path.join(REPO, "cudaraster-linux_amd", "lib", "libcrbench.so"))
    lib.crbench_plain_draws.restype = ctypes.c_double
    lib.crbench_plain_draws.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    return lib


def native_draw_seconds(r, steps):
    """Seconds per synchronous drawTriangles() of a SceneRenderer's frame,
    timed by a native loop over the C ABI (bench_native/plain_draws.cpp:
    the calls a C++ caller makes per frame -- bind surfaces, pipe and
    buffers, deferredClear, drawTriangles -- with no interpreter between
    them)."""
    import ctypes
    import cudaraster
    sc = r.scene
    views = (cudaraster.View * 1)()
    v = views[0]
    v.color, v.depth = r.color.handle.value, r.depth.handle.value
    v.tri_id = r.triid.handle.value if r.triid is not None else None
    v.vertices, v.num_vertices = r.vbuf.data_ptr(), sc.vertices.shape[0]
    v.indices, v.num_tris = r.ibuf.data_ptr(), sc.num_tris
    v.deferred_clear = 1 if sc.deferred_clear else 0
    v.pipe = cudaraster.scene_pipe(sc)
    v.clear_color[:] = [float(x) for x in sc.clear_color]
    v.clear_depth = float(sc.clear_depth)
    st, bad = ctypes.c_int32(), ctypes.c_int32()
    t = _bench_lib().crbench_plain_draws(r.cr.ctx, views, 1, steps, ctypes.byref(st), ctypes.byref(bad))
    if t < 0:
        raise RuntimeError(f"native draw loop failed: status {st.value}")
    return t


def plain_workload_leg(name, steps=10, warmup=2, triid=True, cpu=True):
    """One plain-draw workload beside the headline one (the skewed 4M +
    1024-triangle 8x src-over frame): Mtris/s and Gfragments/s over `steps`
    synchronous draws, per-stage times and per-kernel algorithmic GB/s, and
    the CPU oracle on one frame."""
    import torch
    import cudaraster
    scene, desc = workload(name)
    r = cudaraster.SceneRenderer(scene, triangle_ids=triid)
    r.cr.setDebugParams(count_fragments=True)
    r.draw()
    cnt = r.cr.getCounters()
    fragments, visible = cnt["fragments"], cnt["visible_subtris"]
    r.cr.setDebugParams()
    for _ in range(warmup):
        r.draw()
    torch.cuda.synchronize()
    stage = {"setupTime": 0.0, "binTime": 0.0, "coarseTime": 0.0, "fineTime": 0.0}
    gpu_draw = []
    t0 = time.perf_counter()
    for _ in range(steps):
        r.draw()
        st = r.cr.getStats()
        for k in stage:
            stage[k] += st[k]
        gpu_draw.append(sum(st[k] for k in stage))
    torch.cuda.synchronize()
    elapsed_py = time.perf_counter() - t0
    counters = r.cr.getCounters()
    elapsed = native_draw_seconds(r, steps) * steps
    r.close()
    avg = {k: v / steps for k, v in stage.items()}
    kernels = {}
    for key, kname, b in (("setupTime", "setupBinKernel", setup_algorithmic_bytes(scene, counters, visible)),
                          ("coarseTime", "coarseKernel", coarse_algorithmic_bytes(scene, counters, visible)),
                          ("fineTime", "fineKernel", fine_algorithmic_bytes(scene, visible, triid))):
        kernels[kname] = {"ms": avg[key] * 1e3, "algorithmic_bytes": b,
                          "GBps": b / avg[key] / 1e9 if avg[key] > 0 else None}
        kernels[kname]["frac"] = kernels[kname]["GBps"] / HBM_PEAK_GBS if kernels[kname]["GBps"] else None
    ms = elapsed / steps * 1e3
    out = {"workload": desc, "triangles": scene.num_tris, "steps": steps, "ms_per_step": ms,
           "timing": "native C loop over the C ABI; stage_ms from the library's stage events",
           "ms_per_step_python": elapsed_py / steps * 1e3,
           "gpu_ms_per_draw": statistics.median(gpu_draw) * 1e3,
           "mtris_per_s_gpu_time": scene.num_tris / statistics.median(gpu_draw) / 1e6,
           "mtris_per_s": scene.num_tris * steps / elapsed / 1e6,
           "gfragments_per_s": fragments * steps / elapsed / 1e9, "fragments_per_frame": fragments,
           "stage_ms": {k: v * 1e3 for k, v in avg.items()}, "kernels": kernels,
           "frame_algorithmic_GBps": frame_algorithmic_bytes(scene, triid) / (ms * 1e-3) / 1e9,
           "counters": counters}
    out["frame_roofline_frac"] = out["frame_algorithmic_GBps"] / HBM_PEAK_GBS
    if cpu:
        # One multi-threaded frame (a 1-thread frame of this workload would
        # take minutes of host time).
        out["cpu_baseline"] = cpu_baseline(scene, runs=1, single_runs=0)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--no-triid", action="store_true", help="draw without the triangle-id surface")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-count", action="store_true",
                    help="skip the counting draw (fragment and visible-subtriangle totals)")
    ap.add_argument("--no-diff", action="store_true",
                    help="skip the differentiable workload (32 views, forward + backward)")
    ap.add_argument("--no-multi-view", action="store_true",
                    help="skip the batched/plain multi-view ratio (64 views)")
    ap.add_argument("--no-c2", action="store_true",
                    help="skip the 4M + 1024-triangle 8x src-over leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 256 MiB copy kernel (reference for HBM counter scaling)")
    ap.add_argument("--only", choices=("multi-view", "diff", "vertex", "resolve"),
                    help="run only that leg and print its JSON record (for kernel profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by view or screen region, see the design notes)", file=sys.stderr)
        raise SystemExit(2)

    import torch
    if args.calibrate_copy:
        copy_calibration()
        return
    if args.only == "multi-view":
        print(json.dumps({"multi_view": multi_view_ratio(triid=not args.no_triid)}), flush=True)
        return
    if args.only == "diff":
        print(json.dumps({"differentiable": differentiable_views()}), flush=True)
        return
    if args.only == "vertex":
        print(json.dumps({"vertex_shader": vertex_shader()}), flush=True)
        return
    if args.only == "resolve":
        print(json.dumps({"resolve": resolve_leg()}), flush=True)
        return
    import cudaraster
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    scene, desc = workload(args.workload)
    triid = not args.no_triid
    r = cudaraster.SceneRenderer(scene, triangle_ids=triid)

    # One counting draw (hierarchical Z off; a separate kernel variant) for
    # the exact fragment and visible-subtriangle totals.
    fragments = visible = None
    if not args.no_count:
        r.cr.setDebugParams(count_fragments=True)
        r.draw()
        cnt = r.cr.getCounters()
        fragments, visible = cnt["fragments"], cnt["visible_subtris"]
        r.cr.setDebugParams()
    for _ in range(args.warmup):
        r.draw()
    torch.cuda.synchronize()

    stage = {"setupTime": 0.0, "binTime": 0.0, "coarseTime": 0.0, "fineTime": 0.0}
    gpu_draw = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.draw()
        st = r.cr.getStats()
        for k in stage:
            stage[k] += st[k]
        gpu_draw.append(sum(st[k] for k in stage))
    torch.cuda.synchronize()
    elapsed_py = time.perf_counter() - t0
    counters = r.cr.getCounters()
    # `value`: the same draws timed by a native loop over the C ABI (the
    # Python loop above, which also reads the stage times, is reported as
    # ms_per_step_python).
    elapsed = native_draw_seconds(r, args.steps) * args.steps
    ms = elapsed / args.steps * 1e3

    avg = {k: v / args.steps for k, v in stage.items()}
    # Per-kernel roofline: algorithmic bytes of one launch over its average
    # duration (HIP events recorded around each kernel on the draw stream).
    kernels = {}
    for key, kname, fn in (("setupTime", "setupBinKernel", lambda: setup_algorithmic_bytes(scene, counters, visible)),
                           ("coarseTime", "coarseKernel", lambda: coarse_algorithmic_bytes(scene, counters, visible)),
                           ("fineTime", "fineKernel", lambda: fine_algorithmic_bytes(scene, visible, triid))):
        b = fn() if visible is not None else None
        kernels[kname] = {"ms": avg[key] * 1e3, "algorithmic_bytes": b,
                          "GBps": b / avg[key] / 1e9 if b is not None and avg[key] > 0 else None}
    kname = max(kernels, key=lambda k: kernels[k]["ms"])
    alg_bytes = kernels[kname]["algorithmic_bytes"]
    kernel_s = kernels[kname]["ms"] * 1e-3
    achieved = kernels[kname]["GBps"]
    traffic = measured_traffic(kname, args.workload, triid)
    line = {
        "metric": METRIC,
        "value": scene.num_tris * args.steps / elapsed / 1e6,
        "unit": "Mtris/s",
        "gfragments_per_s": fragments * args.steps / elapsed / 1e9 if fragments is not None else None,
        "fragments_per_frame": fragments,
        "batched_over_plain_views_ratio": None,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "timing": "native C loop over the C ABI (cudaraster-linux_amd/bench_native/plain_draws.cpp); "
                  "stage_ms from the library's stage events",
        "ms_per_step_python": elapsed_py / args.steps * 1e3,
        # The draw's GPU time from the first stage to the last (HIP events,
        # median over the steps) and the triangle rate over it.
        "gpu_ms_per_draw": statistics.median(gpu_draw) * 1e3,
        "mtris_per_s_gpu_time": scene.num_tris / statistics.median(gpu_draw) / 1e6,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "i32+f32",
        "data": "synthetic (seeded)",
        "config": {"workload": desc, "triangles": scene.num_tris, "width": scene.width,
                   "height": scene.height, "samples": scene.num_samples, "triangle_id_surface": triid},
        "stage_ms": {k: v * 1e3 for k, v in avg.items()},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved is not None else None,
                     "traffic": traffic, "algorithmic_bytes_per_launch": alg_bytes,
                     "kernel_ms": kernel_s * 1e3, "visible_subtris": visible,
                     # The fine stage is integer work limited by VALU issue, not by HBM: the
                     # fraction of the VALU issue capacity its instruction count fills.
                     "valu_issue": measured_issue(kname, args.workload, triid)},
        "kernels": kernels,
        "frame_algorithmic_GBps": frame_algorithmic_bytes(scene, triid) / (ms * 1e-3) / 1e9,
        "counters": counters,
    }
    # The whole frame against the HBM peak (compulsory bytes of the draw:
    # inputs read once, surfaces written once), beside the per-kernel figure.
    line["frame_roofline_frac"] = line["frame_algorithmic_GBps"] / HBM_PEAK_GBS
    r.close()
    if args.workload == "c1" and not args.no_c2:
        line["c2"] = plain_workload_leg("c2", triid=triid, cpu=not args.no_cpu_baseline)
    if not args.no_multi_view:
        mv = multi_view_ratio(triid=triid)
        line["batched_over_plain_views_ratio"] = mv["ratio"]
        line["multi_view"] = mv
    if not args.no_diff:
        line["differentiable"] = differentiable_views()
    line["vertex_shader"] = vertex_shader()
    line["resolve"] = resolve_leg()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
