// Native timing loop for the multi-view benchmark's plain leg: the views
// drawn one after another through the C ABI, each the way a SceneCR-style
// caller draws a frame (test/SceneCR.cpp:267-292: bind surfaces, pipe and
// buffers, deferredClear, a synchronous drawTriangles).  Measuring the loop
// here keeps interpreter overhead out of the per-view cost.  Built into
// lib/libcrbench.so against libcr.so; used by bench.py only.
#include <chrono>
#include <cstdint>

#include "cr/cr.h"

extern "C" {

// Draws the n views `passes` times in order; returns the mean seconds per
// pass, or a negative value after the first failing call (*status: its
// cr_status, *failed_view: the view).
double crbench_plain_draws(cr_context* ctx, const cr_view* views, int32_t n, int32_t passes, int32_t* status,
                           int32_t* failed_view) {
  *status = CR_OK;
  *failed_view = -1;
  const auto t0 = std::chrono::steady_clock::now();
  for (int32_t p = 0; p < passes; p++) {
    for (int32_t k = 0; k < n; k++) {
      const cr_view& v = views[k];
      cr_status st = cr_set_surfaces(ctx, v.color, v.depth);
      if (st == CR_OK) st = cr_set_triangle_id_surface(ctx, v.tri_id);
      if (st == CR_OK) st = cr_set_pixel_pipe(ctx, &v.pipe);
      if (st == CR_OK && v.deferred_clear) st = cr_deferred_clear(ctx, v.clear_color, v.clear_depth);
      if (st == CR_OK) st = cr_set_vertex_buffer(ctx, v.vertices, 0, v.num_vertices);
      if (st == CR_OK) st = cr_set_index_buffer(ctx, v.indices, 0, v.num_tris);
      if (st == CR_OK) st = cr_draw_triangles(ctx, nullptr);
      if (st != CR_OK) {
        *status = st;
        *failed_view = k;
        return -1.0;
      }
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count() / (passes > 0 ? passes : 1);
}

}  // extern "C"
