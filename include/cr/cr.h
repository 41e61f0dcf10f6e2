/*
 * cr.h -- C ABI of the MI355X (gfx950) rasterizer, libcr.so.
 *
 * Drop-in boundary for the reference's host API, FW::CudaRaster
 * (src/cudaraster/CudaRaster.hpp:143-168) and FW::CudaSurface
 * (src/cudaraster/CudaSurface.hpp:60-71).  Plain pointers and sizes only.
 * Each entry point names the reference member it replaces.
 *
 * Errors: the reference calls fail() and exits (base/Defs.hpp:107-117).
 * Here every function returns a cr_status; cr_last_error() gives the text of
 * the last failure on that context (or of the last context-less failure when
 * ctx is NULL).  Misuse the reference rejects with fail() is rejected with
 * the same condition and CR_ERR_INVALID / CR_ERR_STATE.
 *
 * Threading: a context is single-threaded, as the reference (one host
 * thread, shared module globals).  Different contexts are independent.
 *
 * Memory: vertex and index buffers are caller-owned device pointers
 * (hipMalloc'ed or torch tensors).  Surfaces are device buffers owned by the
 * library, laid out exactly as the reference's CUDA arrays
 * (CudaSurface.cpp:86-91): (ceil8(W) * samples) x ceil8(H) 32-bit texels,
 * row 0 at the bottom, an 8x8 tile replicated horizontally once per sample:
 * texel (tileX*8*N + sample*8 + (x & 7), y).
 */
#ifndef CR_CR_H
#define CR_CR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cr_status {
    CR_OK = 0,
    CR_ERR_INVALID = 1,      /* bad argument */
    CR_ERR_STATE = 2,        /* call order / missing state */
    CR_ERR_DEVICE = 3,       /* HIP runtime error */
    CR_ERR_OVERFLOW = 4,     /* internal capacity exceeded beyond the limits */
    CR_ERR_UNSUPPORTED = 5,
    CR_ERR_NOMEM = 6
} cr_status;

typedef struct cr_context cr_context;
typedef struct cr_surface cr_surface;

/* CudaSurface::Format (CudaSurface.hpp:39-45) plus a 32-bit triangle-id
 * target (per-sample index of the last triangle written, 0xFFFFFFFF where
 * none was). */
typedef enum cr_format {
    CR_FORMAT_RGBA8 = 0,
    CR_FORMAT_DEPTH32 = 1,
    CR_FORMAT_TRIANGLE_ID = 2
} cr_format;

/* RenderModeFlag_* (cudaraster/cuda/PixelPipe.hpp:30-35). */
enum {
    CR_RENDER_DEPTH = 1 << 0,
    CR_RENDER_LERP = 1 << 1,
    CR_RENDER_QUADS = 1 << 2     /* numerical derivatives over 2x2 pixel quads (texPhong's texture LOD) */
};

/* Stock fragment shaders.  GOURAUD = GouraudShader (PixelPipe.inl:65-69):
 * toABGR of varying 0 interpolated at the centroid.  CONSTANT writes
 * pipe.const_color (FragmentShader_passthrough, test/shader/PassThrough.cu:44-54,
 * is CONSTANT with 0xFF0000FF). */
typedef enum cr_fragment_shader {
    CR_SHADER_GOURAUD = 0,
    CR_SHADER_CONSTANT = 1,
    CR_SHADER_TEX_PHONG = 2      /* FragmentShader_texPhong (test/shader/Shaders.cu:122-199), below */
} cr_fragment_shader;

/* ---- the texPhong shader (test/shader/Shaders.cu:37-48, 122-199) ----
 * Vertex records (ShadedVertex_texPhong, Shaders.hpp:84-90): clip position,
 * then varying 0 camera position, varying 1 camera normal, varying 2
 * texture coordinate (x, y used): vertex_struct_size >= 64.  Per fragment:
 * camera position and normal at the coverage centroid; the texture
 * coordinate at the centroid with analytic derivatives, or, with
 * CR_RENDER_QUADS, at the pixel centre with numerical derivatives over the
 * 2x2 pixel quad; the triangle's material (an index outside the table
 * discards the fragment); the diffuse colour from the diffuse texture and
 * alpha from the alpha texture's second channel when their size[0] != 0,
 * each a lerp of two bilinear lookups at the mip level the derivatives
 * select; alpha < 0.5 discards (no depth or colour write); Phong lighting
 * with the camera at the origin.  Arithmetic is IEEE FP32 with this
 * library's own log2/exp2 (the reference's -use_fast_math approximations
 * are not reproduced) and FP32 bilinear filtering with wrap addressing. */

/* TextureSpec (Shaders.hpp:33-38): a texture's place in the atlas. */
typedef struct cr_texture_spec {
    float size[2];               /* level-0 size in texels; size[0] == 0: no texture */
    uint32_t pad[2];
    float mip_levels[16][4];     /* per level: scaleX, scaleY, biasX, biasY (normalised atlas coordinates) */
} cr_texture_spec;

/* Material (Shaders.hpp:44-51). */
typedef struct cr_material {
    float diffuse_color[4];
    float specular_color[3];
    float glossiness;
    cr_texture_spec diffuse_texture;
    cr_texture_spec alpha_texture;
} cr_material;

/* The shader's constants (Shaders.hpp:57-65 Constants::materials and
 * triangleMaterialIdx, and t_textureAtlas); all device pointers that must
 * stay valid while draws use them. */
typedef struct cr_shader_constants {
    const cr_material* materials;            /* num_materials records, 16-byte aligned */
    int32_t num_materials;
    int32_t atlas_width, atlas_height;       /* texels, 1..16384 */
    const int32_t* triangle_material_idx;    /* material of input triangle t */
    int64_t num_triangle_materials;          /* entries (triangles beyond them are discarded) */
    const float* atlas;                      /* atlas_height x atlas_width float4 texels, 16-byte aligned */
} cr_shader_constants;

/* Stock blend shaders (PixelPipe.hpp:145-175, PixelPipe.inl:73-85). */
typedef enum cr_blend_shader {
    CR_BLEND_REPLACE = 0,
    CR_BLEND_SRC_OVER = 1,
    CR_BLEND_ADDITIVE = 2,
    CR_BLEND_DEPTH_ONLY = 3
} cr_blend_shader;

/* What CR_DEFINE_PIXEL_PIPE fixes at compile time in the reference
 * (PixelPipe.inl:241-275): vertex struct size, shaders, samples, flags.
 * All variants are compiled ahead of time for gfx950. */
typedef struct cr_pipe_desc {
    int32_t vertex_struct_size;  /* bytes, multiple of 16, >= 16 (>= 32 for GOURAUD) */
    int32_t fragment_shader;     /* cr_fragment_shader */
    int32_t blend_shader;        /* cr_blend_shader */
    int32_t samples_log2;        /* 0..3 */
    uint32_t render_mode_flags;  /* CR_RENDER_* */
    uint32_t const_color;        /* ABGR8888 for CR_SHADER_CONSTANT */
} cr_pipe_desc;

/* CudaRaster::Stats (CudaRaster.hpp:45-51): seconds per stage of the last
 * draw, from HIP events.  Triangle setup and binning run as one kernel:
 * setup_time covers both and bin_time is 0. */
typedef struct cr_stats {
    float setup_time;
    float bin_time;
    float coarse_time;
    float fine_time;
} cr_stats;

/* Counters of the last draw (the reference's ProfilingMode_Counters subset
 * that the pipeline keeps without slowing down). */
typedef struct cr_counters {
    int64_t num_tris;
    int64_t num_subtris;         /* slots used (numTris + clipped extra) */
    int64_t bin_entries;
    int64_t tile_entries;
    int64_t fragments;           /* (subtriangle, pixel) pairs with >= 1 covered sample;
                                    counted only by draws with count_fragments set, else 0 */
    int64_t retries;             /* capacity grow-and-retry rounds */
    int64_t visible_subtris;     /* subtriangles that survive setup's culls;
                                    counted only by draws with count_fragments set, else 0 */
    int64_t tile_list_entries;   /* (subtriangle, tile) entries of the tile lists; tile_entries
                                    is the room reserved for them (whole 32-entry segments) */
} cr_counters;

/* ProfilingMode (cudaraster/cuda/PrivateDefs.hpp:158-204, CudaRaster.cpp:296-323):
 * DEFAULT draws at full speed and reports stage times and queue sizes;
 * COUNTERS draws with the counting kernel variants and records the profile
 * counters below (cr_get_profile_counters, cr_get_profiling_info). */
enum {
    CR_PROFILING_DEFAULT = 0,
    CR_PROFILING_COUNTERS = 1
};

/* CudaRaster::DebugParams (CudaRaster.hpp:53-67) has per-stage host
 * emulation; this library has one GPU implementation per stage and keeps
 * the knobs that change execution but never the image. */
typedef struct cr_debug_params {
    int32_t disable_hiz;         /* 1: no tile-level early depth rejection */
    int32_t count_fragments;     /* 1: count fragments exactly (implies disable_hiz; slower) */
    int32_t coarse_threads;      /* coarse-stage workgroup size: 256, 1024, or 0 to choose from the draw */
    int32_t profiling_mode;      /* CR_PROFILING_* */
} cr_debug_params;

/* Profile counters of the last CR_PROFILING_COUNTERS draw, the reference's
 * CR_PROFILING_COUNTERS table (PrivateDefs.hpp:158-204) mapped onto this
 * pipeline: a numerator and a denominator each; the report prints
 * num / max(denom, 1).  Percentages count 100 per event.  Setup counters are
 * the reference's own; bin, coarse and fine counters describe this
 * library's stages (per-chunk binning, per-bin coarse rounds, one wave per
 * tile) where the reference's describe its own. */
typedef enum cr_profile_counter {
    CR_PROF_SETUP_VIEWPORT_CULL = 0,      /* % of triangles frustum-culled */
    CR_PROF_SETUP_BACKFACE_CULL,          /* % backfacing (area <= 0) */
    CR_PROF_SETUP_BETWEEN_PIXELS_CULL,    /* % covering no sample by the box test */
    CR_PROF_SETUP_CLIPPED,                /* % taking the clip path */
    CR_PROF_SETUP_SAMPLES_PER_TRI,        /* covered samples / triangles with a subtriangle */
    CR_PROF_BIN_QUEUE_OVERFLOWS,          /* capacity grow-and-retry rounds of the draw */
    CR_PROF_BIN_TRIS_PER_CHUNK,           /* subtriangles / 1024-triangle setup chunk */
    CR_PROF_BIN_TRI_BB_AREA,              /* bins / subtriangle (bounding box) */
    CR_PROF_BIN_TRI_SINGLE_BIN,           /* % of subtriangles inside one bin */
    CR_PROF_COARSE_BINS,                  /* bins with entries */
    CR_PROF_COARSE_ROUNDS_PER_BIN,        /* rounds (one entry per lane) / bin */
    CR_PROF_COARSE_TRIS_PER_ROUND,        /* bin entries / round */
    CR_PROF_COARSE_TILES_PER_ROUND,       /* tiles receiving entries / round */
    CR_PROF_COARSE_EMITS_PER_ROUND,       /* tile entries / round */
    CR_PROF_COARSE_EMITS_PER_TRI,         /* tile entries / bin entry */
    CR_PROF_COARSE_CASE_A,                /* % of entries walked one lane per box (boxes <= 2x2 tiles) */
    CR_PROF_COARSE_CASE_B,                /* % walked over the wave's union box */
    CR_PROF_COARSE_CASE_C,                /* % walked per lane with large boxes by the wave */
    CR_PROF_FINE_TRI_PER_TILE,            /* tile entries / tile with entries */
    CR_PROF_FINE_EARLY_Z_CULL,            /* % of tile entries rejected by hierarchical Z */
    CR_PROF_FINE_EMPTY_CULL,              /* % of walked entries covering no sample of the tile */
    CR_PROF_FINE_Z_KILL,                  /* % of fragments whose samples all fail the depth test */
    CR_PROF_FINE_WARP_UTIL,               /* % of lanes with a fragment per walked entry */
    CR_PROF_FINE_BLEND_ROUNDS,            /* ROP rounds / written fragment (always 1: in order per lane) */
    CR_PROF_FINE_FRAG_PER_TRI,            /* fragments / walked entry */
    CR_PROF_FINE_FRAG_PER_TILE,           /* fragments / tile with entries */
    CR_PROF_COUNT
} cr_profile_counter;

typedef struct cr_profile_counters {
    int64_t num[CR_PROF_COUNT];
    int64_t denom[CR_PROF_COUNT];
} cr_profile_counters;

/* ---- context (CudaRaster ctor + init, CudaRaster.cpp:98-128) ---- */
cr_status cr_context_create(cr_context** out);
cr_status cr_context_destroy(cr_context* ctx);
const char* cr_last_error(const cr_context* ctx);

/* ---- surfaces (CudaSurface ctor, CudaSurface.cpp:39-110) ----
 * Size 1..2048 per axis, samples 1, 2, 4 or 8. */
cr_status cr_surface_create(int32_t width, int32_t height, int32_t format, int32_t num_samples,
                            cr_surface** out);
cr_status cr_surface_destroy(cr_surface* surf);
/* getSize / getRoundedSize / getTextureSize / getNumSamples. */
cr_status cr_surface_get_info(const cr_surface* surf, int32_t* width, int32_t* height,
                              int32_t* rounded_width, int32_t* rounded_height,
                              int32_t* texture_width, int32_t* num_samples, int32_t* format);
/* Device pointer of the texels (replaces getCudaArray). */
cr_status cr_surface_get_device_ptr(const cr_surface* surf, void** dev_ptr, size_t* bytes);
/* Synchronous copies of the whole texture (texture_width x rounded_height u32). */
cr_status cr_surface_download(const cr_surface* surf, void* host, size_t bytes);
cr_status cr_surface_upload(cr_surface* surf, const void* host, size_t bytes);

/* ---- pipeline state ---- */
/* CudaRaster::setSurfaces (CudaRaster.cpp:132-170).  Both NULL unsets. */
cr_status cr_set_surfaces(cr_context* ctx, cr_surface* color, cr_surface* depth);
/* Optional triangle-id target (CR_FORMAT_TRIANGLE_ID, same size/samples);
 * NULL disables it.  No reference counterpart (per-sample triangle index
 * of the last written fragment). */
cr_status cr_set_triangle_id_surface(cr_context* ctx, cr_surface* tri_id);
/* CudaRaster::deferredClear (CudaRaster.cpp:174-179): rgba as Vec4f, depth
 * in [0,1]. Applies to the next draw. */
cr_status cr_deferred_clear(cr_context* ctx, const float rgba[4], float depth);
/* CudaRaster::setPixelPipe (CudaRaster.cpp:183-216). */
cr_status cr_set_pixel_pipe(cr_context* ctx, const cr_pipe_desc* pipe);
/* Stock pipe by name, as the demo's CR_DEFINE_PIXEL_PIPE names:
 * "PixelPipe_passthrough" (constant red, 16-byte vertices) and
 * "PixelPipe_gouraud" (GouraudVertex, GouraudShader) and "PixelPipe_texPhong"
 * (ShadedVertex_texPhong, FragmentShader_texPhong); blend, samples and
 * flags as the demo's -D defines (test/SceneCR.cpp:165-177). */
cr_status cr_pipe_lookup(const char* name, int32_t blend_shader, int32_t samples_log2,
                         uint32_t render_mode_flags, cr_pipe_desc* out);
/* Sets (or, with NULL, clears) the texPhong shader's constants: what the
 * demo writes into c_constants and binds to t_textureAtlas before drawing
 * with PixelPipe_texPhong. */
cr_status cr_set_shader_constants(cr_context* ctx, const cr_shader_constants* constants);
/* CudaRaster::setVertexBuffer (CudaRaster.cpp:220-224): device pointer +
 * byte offset; num_vertices bounds the indices (out-of-range triangles are
 * dropped). */
cr_status cr_set_vertex_buffer(cr_context* ctx, const void* dev_ptr, int64_t ofs, int64_t num_vertices);
/* CudaRaster::setIndexBuffer (CudaRaster.cpp:228-233): int3 per triangle. */
cr_status cr_set_index_buffer(cr_context* ctx, const void* dev_ptr, int64_t ofs, int32_t num_tris);

/* CudaRaster::drawTriangles (CudaRaster.cpp:237-342).  Launches the four
 * stages on `stream` (a hipStream_t; NULL = default stream), reads back the
 * overflow counters, grows the internal buffers and re-runs when needed, and
 * returns once the frame is complete -- the reference's synchronous
 * semantics. */
cr_status cr_draw_triangles(cr_context* ctx, void* stream);
/* Asynchronous variant (no counterpart in the reference): launches with the
 * current capacities and returns without waiting.  cr_draw_check() waits
 * for that draw and reports CR_ERR_OVERFLOW if it needs a re-run (the
 * surfaces are then unchanged from before the draw when deferred clear was
 * off, or cleared when it was on).  cr_reserve() pre-sizes the capacities. */
cr_status cr_draw_triangles_async(cr_context* ctx, void* stream);
cr_status cr_draw_check(cr_context* ctx);
cr_status cr_reserve(cr_context* ctx, int64_t max_subtris, int64_t max_bin_entries,
                     int64_t max_tile_entries);

/* drawTriangles that allocates nothing.  No counterpart in the reference:
 * for callers that cannot call hipMalloc (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  Same state, same
 * synchronous semantics and the same images as cr_draw_triangles, but the
 * draw's intermediate buffers live in one fixed device array of
 * CR_NOALLOC_SCRATCH_BYTES bytes that every GPU the library runs on holds
 * until the process ends (looked up per device by cr_context_create).
 * A draw whose intermediates do not fit is rejected with CR_ERR_INVALID
 * (its per-draw tables alone are too large) or CR_ERR_OVERFLOW (its queues
 * are), leaving the surfaces as cr_draw_triangles leaves them after an
 * overflowed attempt.  All calls on one GPU share the array: calls that
 * overlap on one GPU (from several threads or on several streams) race. */
#define CR_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
cr_status cr_draw_triangles_noalloc(cr_context* ctx, void* stream);

/* CudaRaster::getStats (CudaRaster.cpp:346-363). */
cr_status cr_get_stats(cr_context* ctx, cr_stats* out);
cr_status cr_get_counters(cr_context* ctx, cr_counters* out);
/* CudaRaster::getProfilingInfo (CudaRaster.cpp:367-497): text report in
 * the reference's layout -- stage times and queue sizes for
 * CR_PROFILING_DEFAULT, the counter table for CR_PROFILING_COUNTERS. */
cr_status cr_get_profiling_info(cr_context* ctx, char* buf, size_t len);
/* The profile counters of the last draw (all zero unless it was drawn with
 * CR_PROFILING_COUNTERS); the numbers behind getProfilingInfo's counter
 * table (CudaRaster.cpp:424-447). */
cr_status cr_get_profile_counters(cr_context* ctx, cr_profile_counters* out);
/* CudaRaster::setDebugParams (CudaRaster.cpp:501-504). */
cr_status cr_set_debug_params(cr_context* ctx, const cr_debug_params* p);

/* ---- batched multi-view draws (no reference counterpart) ----
 * The reference draws one view per drawTriangles() call, each synchronous.
 * A batch draws many independent views -- each with its own surfaces,
 * buffers, pipe, resolution and sample count -- with one launch per stage
 * over all views, replayed from a captured hipGraph.  Every view's output is
 * bit-identical to a plain draw of the same state (cr_set_surfaces,
 * cr_set_pixel_pipe, cr_set_*_buffer, cr_deferred_clear, cr_draw_triangles).
 *
 * Views may not share a colour, depth or triangle-id surface (they run
 * concurrently).  Vertex and index buffers are device pointers that must
 * stay valid, unchanged in size, for the life of the batch; their contents
 * may change between draws. */
typedef struct cr_view {
    cr_surface* color;           /* CR_FORMAT_RGBA8 */
    cr_surface* depth;           /* CR_FORMAT_DEPTH32, same size and samples */
    cr_surface* tri_id;          /* optional CR_FORMAT_TRIANGLE_ID, or NULL */
    const void* vertices;        /* device pointer (offset applied), 16-byte aligned */
    int64_t num_vertices;
    const void* indices;         /* device pointer (offset applied), int3 per triangle */
    int32_t num_tris;
    int32_t deferred_clear;      /* 1: clear to clear_color / clear_depth first (cr_deferred_clear) */
    cr_pipe_desc pipe;
    float clear_color[4];
    float clear_depth;
} cr_view;

typedef struct cr_batch cr_batch;

typedef struct cr_batch_info {
    int32_t num_views;
    int32_t fine_tiles_per_group;  /* fine-stage workgroup shape in use (tiles = waves per workgroup) */
    int32_t uses_graph;            /* 1: draws replay a captured hipGraph */
    int32_t tuned;                 /* 1: cr_batch_autotune has run */
    int64_t retries;               /* capacity grow rounds of the last draw */
    int64_t reserved_bytes;        /* device bytes of the batch's intermediate buffers */
    float last_gpu_ms;             /* GPU time of the last draw's launch (events around it) */
    float tuned_gpu_ms;            /* best GPU time seen by the autotuner */
    int32_t fine_banded;           /* 1: fine tile groups in per-XCD bands; 0: dealt round-robin */
    int32_t coarse_threads;        /* coarse-stage workgroup size in use */
    int32_t view_groups;           /* views split into this many concurrently launched groups */
} cr_batch_info;

/* Copies the view list; validates each view as cr_set_surfaces /
 * cr_set_pixel_pipe / cr_draw_triangles would. */
cr_status cr_batch_create(cr_context* ctx, const cr_view* views, int32_t num_views, cr_batch** out);
cr_status cr_batch_destroy(cr_batch* batch);
/* Draws every view; returns once all are complete (grows the capacities and
 * re-draws the views that overflowed, as cr_draw_triangles). */
cr_status cr_batch_draw(cr_batch* batch, void* stream);
/* Project-written autotuner: draws the batch to measure each view's exact
 * queue demand, sets per-view queue depths to it plus a margin, orders every
 * stage's views longest-first by the measured entries per bin and per tile,
 * times every fine-stage workgroup shape with banded and round-robin tile
 * groups and both coarse workgroup sizes, then 1, 2 or 4 cost-balanced view
 * groups whose launch chains run concurrently (forked streams, parallel
 * graph branches), keeps the fastest, then captures the graph.  Bin size is chosen
 * per view from its resolution (the smallest bins with at most 256 bins).
 * Surfaces of views without deferred clear are restored afterwards. */
cr_status cr_batch_autotune(cr_batch* batch, void* stream);
/* Sets one view's capacities exactly (subtriangle slots, bin-queue and
 * tile-list entries; each >= 1). Smaller than the demand is allowed: the
 * draw then grows them. */
cr_status cr_batch_reserve(cr_batch* batch, int32_t view, int64_t max_subtris, int64_t max_bin_entries,
                           int64_t max_tile_entries);
/* 0: launch the stages directly on each draw instead of replaying a graph. */
cr_status cr_batch_use_graph(cr_batch* batch, int32_t enable);
/* Sets the launch configuration the autotuner otherwise chooses (images do
 * not depend on it): fine tiles per workgroup (2, 4 or 8), fine tile groups
 * in XCD bands (1) or round-robin (0), coarse workgroup size (256 or 1024),
 * concurrent view groups (1..4). */
typedef struct cr_batch_config {
    int32_t fine_tiles_per_group;
    int32_t fine_banded;
    int32_t coarse_threads;
    int32_t view_groups;
} cr_batch_config;
cr_status cr_batch_configure(cr_batch* batch, const cr_batch_config* config);
cr_status cr_batch_get_info(const cr_batch* batch, cr_batch_info* out);
/* Counters of one view from the last draw (fragments and visible_subtris are 0). */
cr_status cr_batch_get_view_counters(const cr_batch* batch, int32_t view, cr_counters* out);

/* ---- differentiable interpolation (no reference counterpart) ----
 * The reference's fragment shaders interpolate varyings with the fixed-point
 * barycentric planes (FineRaster.inl:21-119) and have no backward pass.
 * These two calls make a rasterized frame differentiable with respect to
 * per-vertex attributes and clip-space positions, visibility fixed by the
 * rasterizer's triangle ids (interior barycentric gradients; no edge or
 * occlusion terms).  Images are row-major, row 0 at the bottom (the surface
 * convention), pixel (x, y) at NDC ((2x+1)/W - 1, (2y+1)/H - 1).  Both calls
 * are asynchronous on `stream`; barycentrics and per-pixel gradient terms
 * are evaluated in FP64.
 *
 * pos:  float4 clip position per vertex (16-byte aligned); attr: C floats per
 * vertex; tris: int3 per triangle (the index buffer the frame was drawn with).
 * With num_vertices == 0 (an empty view) pos, attr, grad_pos and grad_attr
 * may be NULL: nothing is read or written through them.
 *
 * Forward: out[y][x][c] = sum_k b_k attr[tris[t][k]][c] for the triangle t
 * in the single-sample id surface, 0 where no triangle is; tri_image[y][x]
 * = t or -1 (may be NULL). */
cr_status cr_interpolate(cr_context* ctx, const cr_surface* tri_id, const float* pos, int64_t num_vertices,
                         const float* attr, int32_t num_channels, const int32_t* tris, int32_t num_tris,
                         float* out, int32_t* tri_image, void* stream);
/* Forward from a triangle-id image the caller holds (int32 [height][width],
 * row 0 at the bottom, -1 or any id >= num_tris where empty -- e.g. the
 * tri_image an earlier cr_interpolate wrote): the same out as cr_interpolate
 * over the surface those ids were read from, whatever the context has
 * drawn since. */
cr_status cr_interpolate_ids(cr_context* ctx, const int32_t* tri_image, int32_t width, int32_t height,
                             const float* pos, int64_t num_vertices, const float* attr, int32_t num_channels,
                             const int32_t* tris, int32_t num_tris, float* out, void* stream);
/* Backward: overwrites grad_pos[V][4] (z component 0) and grad_attr[V][C]
 * with dL/dpos and dL/dattr given grad_out[y][x][c] = dL/dout.  The terms
 * of each triangle's pixels within a 16x16 tile are summed in FP64 and added
 * to the outputs with FP32 atomics (summation order varies between runs in
 * the last bits).  Uses no context scratch. */
cr_status cr_interpolate_backward(cr_context* ctx, const int32_t* tri_image, int32_t width, int32_t height,
                                  const float* pos, int64_t num_vertices, const float* attr,
                                  int32_t num_channels, const int32_t* tris, int32_t num_tris,
                                  const float* grad_out, float* grad_pos, float* grad_attr, void* stream);
/* The same gradients, bit-reproducible from run to run: each (tile,
 * triangle) FP64 sum is converted to a 128-bit fixed-point integer (2^-64
 * resolution; sums must stay below 2^63 in magnitude) and accumulated with
 * integer atomics, whose total does not depend on their order, then rounded
 * to FP32 once.  Uses a context scratch buffer of 16 bytes per output value
 * (grown with hipMalloc when needed). */
cr_status cr_interpolate_backward_deterministic(cr_context* ctx, const int32_t* tri_image, int32_t width,
                                                int32_t height, const float* pos, int64_t num_vertices,
                                                const float* attr, int32_t num_channels, const int32_t* tris,
                                                int32_t num_tris, const float* grad_out, float* grad_pos,
                                                float* grad_attr, void* stream);

/* Passthrough vertex shader, the step before drawTriangles: replaces
 * vertexShader_passthrough (test/shader/PassThrough.cu:16-35) as the demo
 * launches it (test/SceneCR.cpp:267-282).  For each of num_vertices model
 * positions (Vec3f, 12 bytes each) writes the clip position
 * pos_to_clip * (p, 1) -- Mat4f * Vec4f of base/Math.hpp:990-1001, bit for
 * bit -- as 4 floats at the start of the vertex's out_stride-byte output
 * record, leaving the rest of the record (varyings) untouched; the output is
 * a vertex buffer for cr_set_vertex_buffer.  pos_to_clip is 16 floats in
 * Mat4f memory order (column-major, the reference's Mat4f::getPtr()), in
 * host memory.
 * out_vertices must be 16-byte aligned and out_stride a positive multiple of
 * 16.  Asynchronous on `stream`. */
cr_status cr_vertex_shader_passthrough(cr_context* ctx, const float* model_pos, int64_t num_vertices,
                                       const float* pos_to_clip, void* out_vertices, int32_t out_stride,
                                       void* stream);

/* MSAA resolve, the step after drawTriangles (the reference declares one
 * but leaves it commented out, CudaSurface.hpp:73): writes width x height
 * RGBA8 words (row-major, row 0 at the bottom like the surface) to the
 * device buffer out_rgba8, each channel the mean of the pixel's samples
 * rounded half up, (sum + N/2) / N; a single-sample surface is copied.
 * `color` must be an RGBA8 surface.  Asynchronous on `stream`. */
cr_status cr_surface_resolve(const cr_surface* color, void* out_rgba8, void* stream);

/* Version of the ABI (major << 16 | minor). */
int32_t cr_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
