"""Writes the golden images under tests/golden/ from the CPU oracle (test
fixtures: expected outputs of the drop-in path).  Store format: RGBA PNG,
upright (the surface's row 0, the bottom row, is the file's last row),
written and read by cudaraster.save_png / load_png.

    python tests/golden/make_golden.py

cube_256x192_4x_resolved.png: the demo cube (seed 0, demo camera) at
256x192, 4x MSAA, clip positions from the passthrough vertex shader
(float32 Mat4f * Vec4f), drawn by the oracle and box-resolved.
random_96x64_{1,2,4,8}x_resolved.png: 1500 seeded random triangles (areas
0.5-64 px^2, depth test, Gouraud with perspective-correct interpolation) at
each sample count, drawn by the oracle and box-resolved.
cube_1024x768_1x_surfaces.npz: the C0 frame -- the demo cube at 1024x768,
1x, clip positions from the passthrough vertex shader -- as drawn by the
oracle: the colour, depth and triangle-id surfaces (uint32 texels in the
reference's CudaSurface layout, row 0 the bottom row; numpy's compressed
archive, arrays `color`, `depth`, `triid`)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for _p in (os.path.join(REPO, "oracle"), os.path.join(REPO, "cudaraster-linux_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)


def cube_scene(width=256, height=192, samples_log2=2):
    import dataclasses
    from cudaraster import scenes
    import oracle
    m = scenes.demo_camera(width, height).astype(np.float32)
    base = scenes.cube(width, height, samples_log2=samples_log2)
    verts = base.vertices.copy()
    verts[:, :4] = oracle.vertex_passthrough(scenes.CUBE_POSITIONS.astype(np.float32), m)
    return dataclasses.replace(base, vertices=np.ascontiguousarray(verts)), m


def cube_golden():
    import oracle
    scene, _ = cube_scene()
    r = oracle.render(scene)
    return oracle.resolve(r["color"], scene.width, scene.height, scene.samples_log2)


# Small random scenes (the 1M-triangle workload's distribution) at every
# sample count: name -> (triangles, width, height, samples_log2, seed).
RANDOM_GOLDENS = {
    "random_96x64_1x_resolved.png": (1500, 96, 64, 0, 31),
    "random_96x64_2x_resolved.png": (1500, 96, 64, 1, 32),
    "random_96x64_4x_resolved.png": (1500, 96, 64, 2, 33),
    "random_96x64_8x_resolved.png": (1500, 96, 64, 3, 34),
}


def random_scene(name):
    from cudaraster import scenes
    n, w, h, sl2, seed = RANDOM_GOLDENS[name]
    return scenes.random_tris(n, w, h, samples_log2=sl2, seed=seed, area_range=(0.5, 64.0))


def random_golden(name):
    import oracle
    scene = random_scene(name)
    r = oracle.render(scene)
    return oracle.resolve(r["color"], scene.width, scene.height, scene.samples_log2)


CUBE_FRAME = "cube_1024x768_1x_surfaces.npz"


def cube_frame():
    """The C0 frame's scene and the oracle's surfaces for it."""
    import oracle
    scene, _ = cube_scene(1024, 768, 0)
    r = oracle.render(scene)
    return scene, {k: r[k] for k in ("color", "depth", "triid")}


def load_cube_frame():
    with np.load(os.path.join(HERE, CUBE_FRAME)) as z:
        return {k: z[k] for k in ("color", "depth", "triid")}


def main():
    import cudaraster
    cudaraster.save_png(os.path.join(HERE, "cube_256x192_4x_resolved.png"), cube_golden())
    _, surf = cube_frame()
    np.savez_compressed(os.path.join(HERE, CUBE_FRAME), **surf)
    for name in RANDOM_GOLDENS:
        cudaraster.save_png(os.path.join(HERE, name), random_golden(name))


if __name__ == "__main__":
    main()
