"""C ABI: libcr.so loads on a machine without a GPU and exports every entry
point include/cr/cr.h declares; argument checks that precede any device
call behave as the reference's fail() conditions (CudaSurface.cpp:44-62,
CudaRaster.cpp:183-216)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(REPO, "include", "cr", "cr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cr_[a-z_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import cudaraster
    return cudaraster.load()


def test_library_exports_header(lib):
    import cudaraster
    fns = header_functions()
    assert len(fns) >= 20
    for f in fns:
        assert hasattr(lib, f), f"libcr.so does not export {f}"
    assert sorted(cudaraster.EXPORTS) == fns


def test_abi_version(lib):
    assert lib.cr_abi_version() >> 16 == 1


@pytest.mark.parametrize("w,h,fmt,ns", [(0, 10, 0, 1), (10, -1, 0, 1), (2049, 10, 0, 1), (10, 10, 7, 1),
                                        (10, 10, 0, 16), (10, 10, 0, 3), (10, 10, 0, 0)])
def test_surface_argument_checks(lib, w, h, fmt, ns):
    import cudaraster
    out = ctypes.c_void_p()
    st = lib.cr_surface_create(w, h, fmt, ns, ctypes.byref(out))
    assert st == cudaraster.CR_ERR_INVALID
    assert out.value is None
    assert lib.cr_last_error(None)


def test_pipe_lookup(lib):
    import cudaraster
    d = cudaraster.pipe_lookup("PixelPipe_passthrough", cudaraster.BLEND_SRC_OVER, 2, 3)
    assert (d.vertex_struct_size, d.fragment_shader, d.const_color) == (16, cudaraster.SHADER_CONSTANT, 0xFF0000FF)
    assert (d.blend_shader, d.samples_log2, d.render_mode_flags) == (1, 2, 3)
    d = cudaraster.pipe_lookup("PixelPipe_gouraud")
    assert (d.vertex_struct_size, d.fragment_shader) == (32, cudaraster.SHADER_GOURAUD)
    with pytest.raises(cudaraster.CrError):
        cudaraster.pipe_lookup("PixelPipe_nonexistent")


def test_null_arguments(lib):
    import cudaraster
    assert lib.cr_draw_triangles(None, None) == cudaraster.CR_ERR_INVALID
    assert lib.cr_set_surfaces(None, None, None) == cudaraster.CR_ERR_INVALID
    assert lib.cr_context_create(None) == cudaraster.CR_ERR_INVALID
    assert lib.cr_context_destroy(None) == cudaraster.CR_OK


def test_cpp_facade_compiles_and_links(tmp_path):
    """The header-only C++ facade (reference class and method names) builds
    against cr.h and links against libcr.so with the host compiler."""
    import shutil
    import subprocess
    import cudaraster
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "client.cpp"
    src.write_text(r'''
#include "cr/CudaRaster.hpp"
int run(const void* vb, const void* ib, int nv, int nt) {
  cr::CudaSurface color(1024, 768, cr::CudaSurface::FormatRGBA8, 1);
  cr::CudaSurface depth(1024, 768, cr::CudaSurface::FormatDepth32, 1);
  cr::CudaRaster raster;
  raster.setSurfaces(&color, &depth);
  raster.setPixelPipe("PixelPipe_gouraud");
  const float clear[4] = {0.2f, 0.4f, 0.8f, 1.0f};
  raster.deferredClear(clear, 1.0f);
  raster.setVertexBuffer(vb, 0, nv);
  raster.setIndexBuffer(ib, 0, nt);
  const float mvp[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  raster.vertexShaderPassthrough(nullptr, 0, mvp, nullptr, 32);
  raster.drawTriangles();
  color.resolveToDevice(nullptr);
  return (int)raster.getStats().fineTime + (int)color.download().size();
}
int main(int argc, char**) { return argc > 5 ? run(nullptr, nullptr, 0, 0) : 0; }
''')
    libdir = os.path.dirname(cudaraster.lib_path())
    exe = tmp_path / "client"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"), str(src),
                        "-L", libdir, "-lcr", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match_ctypes(tmp_path):
    """The ctypes mirrors of cr_view, cr_batch_info, cr_counters and
    cr_pipe_desc have the C layout (size and every field offset)."""
    import shutil
    import subprocess
    import cudaraster
    if shutil.which("gcc") is None:
        pytest.skip("no host C compiler")
    mirrors = {"cr_view": cudaraster.View, "cr_batch_info": cudaraster.BatchInfo,
               "cr_counters": cudaraster.Counters, "cr_pipe_desc": cudaraster.PipeDesc,
               "cr_debug_params": cudaraster.DebugParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cr/cr.h"', 'int main(void) {']
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
