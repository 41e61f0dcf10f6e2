/*
 * CPU oracle for the four-stage rasterization path (test infrastructure).
 *
 * This library is the checker, never the product: only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py load it.  It restates the
 * reference's rasterization rules in plain C (C11 + pthreads):
 *
 *   triangle setup   cudaraster/cuda/TriangleSetup.inl:19-417
 *   plane equations  cudaraster/cuda/Util.inl:65-91 (GPU variant, not the
 *                    host emulation's F64 variant at CudaRaster.cpp:669-688)
 *   clipping         cudaraster/cuda/Util.hpp:221-286
 *   coverage         cudaraster/cuda/Util.inl:95-134, 340-357 (closed forms)
 *   MSAA / centroid  cudaraster/cuda/Util.hpp:27-36, 182-203
 *   depth encoding   cudaraster/cuda/Util.hpp:290-310, CudaRaster.cpp:174-179
 *   shading, blend   cudaraster/cuda/FineRaster.inl:21-119, Util.inl:30-60,
 *                    PixelPipe.inl:52-85
 *   ordering / ROP   sequential per-sample semantics of the host emulation,
 *                    CudaRaster.cpp:1184-1403 (triangle order, LESS, blend)
 *   surface layout   CudaSurface.cpp:86-91, FineRaster.inl:909, 1088-1100
 *
 * Rendering is "sort-nothing": every subtriangle is rasterized in submission
 * order over the samples of its bounding box.  The multi-threaded mode splits
 * the surface into horizontal bands; each band is owned by one thread and
 * walks the same ordered subtriangle list, so its output is bit-identical to
 * the single-threaded run.
 */
#ifndef CR_ORACLE_H
#define CR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CRO_FLAG_DEPTH = 1,   /* RenderModeFlag_EnableDepth (PixelPipe.hpp:32) */
    CRO_FLAG_LERP  = 2,   /* RenderModeFlag_EnableLerp  (PixelPipe.hpp:33) */
    CRO_FLAG_QUADS = 4,   /* RenderModeFlag_EnableQuads (PixelPipe.hpp:34): numerical
                             derivatives over 2x2 pixel quads (PixelPipe.hpp:59-69) */
};

enum {
    CRO_BLEND_REPLACE   = 0,  /* PixelPipe.hpp:145-150 */
    CRO_BLEND_SRCOVER   = 1,  /* PixelPipe.hpp:154-158, PixelPipe.inl:73-77 */
    CRO_BLEND_ADDITIVE  = 2,  /* PixelPipe.hpp:162-166, PixelPipe.inl:81-85 */
    CRO_BLEND_DEPTHONLY = 3,  /* PixelPipe.hpp:170-175 */
};

enum {
    CRO_SHADER_GOURAUD  = 0,  /* PixelPipe.inl:65-69: toABGR(lerp(varying 0)) */
    CRO_SHADER_CONSTANT = 1,  /* test/shader/PassThrough.cu:44-54 style */
    CRO_SHADER_TEXPHONG = 2,  /* FragmentShader_texPhong, test/shader/Shaders.cu:122-199 */
};

/* TextureSpec and Material (test/shader/Shaders.hpp:33-51), same layout. */
typedef struct {
    float    size[2];           /* level-0 texels; size[0] == 0: no texture */
    uint32_t pad[2];
    float    mipLevels[16][4];  /* scaleX, scaleY, biasX, biasY into the atlas */
} cro_texture_spec;

typedef struct {
    float            diffuseColor[4];
    float            specularColor[3];
    float            glossiness;
    cro_texture_spec diffuseTexture;
    cro_texture_spec alphaTexture;
} cro_material;

typedef struct {
    int32_t  width;           /* viewport width in pixels (1..2048)  */
    int32_t  height;          /* viewport height in pixels (1..2048) */
    int32_t  samplesLog2;     /* 0..3 */
    uint32_t renderModeFlags; /* CRO_FLAG_* */
    int32_t  blend;           /* CRO_BLEND_* */
    int32_t  shader;          /* CRO_SHADER_* */
    uint32_t constColor;      /* ABGR8888 for CRO_SHADER_CONSTANT */
    int32_t  vertexStride;    /* bytes per vertex record, multiple of 16 */
    int32_t  deferredClear;   /* 1: clear before drawing */
    uint32_t clearColor;      /* ABGR8888 */
    uint32_t clearDepth;      /* encoded depth */
    /* texPhong resources (Shaders.hpp:57-65 Constants, t_textureAtlas). */
    int32_t  numMaterials;
    int32_t  atlasWidth, atlasHeight;
    const cro_material* materials;
    const int32_t* triangleMaterialIdx;   /* per input triangle */
    const float* atlas;                    /* atlasHeight x atlasWidth float4 texels */
} cro_pipe;

/* CRTriangleHeader (PrivateDefs.hpp:26-36). */
typedef struct {
    int16_t  v0x, v0y, v1x, v1y, v2x, v2y;
    uint32_t misc;
} cro_header;

/* CRTriangleData (PrivateDefs.hpp:40-62). */
typedef struct {
    uint32_t zx, zy, zb, zslope;
    int32_t  wx, wy, wb;
    int32_t  ux, uy, ub;
    int32_t  vx, vy, vb;
    uint32_t vi0, vi1, vi2;
} cro_data;

typedef struct {
    int64_t numTris;
    int64_t numSubtris;        /* subtriangles emitted (visible) */
    int64_t culledFrustum;     /* TriangleSetup.inl:262-281 */
    int64_t culledBackface;    /* prepareTriangle == 1, fast path */
    int64_t culledBetween;     /* prepareTriangle == 2, fast path */
    int64_t clipped;           /* took the clip path */
    int64_t fragments;         /* (subtriangle, pixel) pairs with >= 1 covered sample */
    int64_t coveredSamples;    /* covered samples summed over those pairs (before the
                                  depth test; SetupSamplesPerTri's numerator,
                                  FineRaster.inl:684, 1032) */
    int64_t zKills;            /* pairs whose covered samples all fail LESS
                                  (FineZKill, FineRaster.inl:702, 1069) */
    int64_t trisVisible;       /* input triangles with >= 1 subtriangle
                                  (SetupSamplesPerTri's denominator, BinRaster.inl:96) */
} cro_stats;

/* ---- single rules (exported for the pinning tests) ---- */
/* The texPhong shader's transcendental functions, written from published
 * methods with IEEE-exact operations only (the reference compiles its
 * shaders with nvcc -use_fast_math, whose approximations are not
 * restated): log2 by the atanh series ln m = 2 atanh((m-1)/(m+1)) on the
 * mantissa, exp2 by range reduction and the Taylor series of e^z, and
 * pow(x, y) = exp2(y log2 x). */
float    cro_log2f(float x);
float    cro_exp2f(float y);
float    cro_powf(float x, float y);
/* Bilinear lookup of the float4 atlas at normalised (u, v), wrap addressing
 * (tex2D on t_textureAtlas, Shaders.cu:140-141). */
void     cro_tex_sample(const float* atlas, int32_t w, int32_t h, float u, float v, float out[4]);
/* texture2D (Shaders.cu:125-148): LOD from the texture-coordinate
 * derivatives, two bilinear lookups, lerp.  *lod receives the LOD. */
void     cro_texture2d(const cro_texture_spec* spec, const float* atlas, int32_t w, int32_t h,
                       const float tex[2], const float texDX[2], const float texDY[2],
                       float out[4], float* lod);
/* evaluateLighting (Shaders.cu:37-48). */
void     cro_lighting(const float cameraPos[3], const float cameraNormal[3], const cro_material* m,
                      const float diffuse[3], float out[3]);
uint32_t cro_to_abgr(const float c[4]);            /* Util.inl:30-38 (device)   */
uint32_t cro_to_abgr_host(const float c[4]);       /* Math.cpp:41-48 (host)     */
uint32_t cro_blend(int32_t mode, uint32_t src, uint32_t dst);
uint32_t cro_encode_depth(uint32_t depth);         /* Util.hpp:290-298 */
uint32_t cro_decode_depth(uint32_t depth);         /* Util.hpp:302-310 */
uint32_t cro_clear_depth(float depth);             /* CudaRaster.cpp:178 */
int32_t  cro_msaa_pattern(int32_t samplesLog2, int32_t i);
int32_t  cro_select_centroid(int32_t samplesLog2, uint32_t mask);
void     cro_setup_pleq(const float values[3], int32_t v0x, int32_t v0y,
                        int32_t d1x, int32_t d1y, int32_t d2x, int32_t d2y,
                        float areaRcp, int32_t samplesLog2, uint32_t out[3]);
int32_t  cro_clip_triangle(float bary[18], const float v0[4], const float v1[4],
                           const float v2[4]);
/* Per-sample coverage of a header triangle; sample given in absolute
 * subpixel coordinates (pixel i spans [16i, 16i+16)). */
int32_t  cro_cover_point(const cro_header* h, int32_t width, int32_t height,
                         int64_t sx, int64_t sy);
/* Sample coverage mask of pixel (px,py) for a header triangle. */
uint32_t cro_cover_pixel(const cro_header* h, int32_t width, int32_t height,
                         int32_t samplesLog2, int32_t px, int32_t py);

/* Set up one triangle.  Writes up to 7 subtriangles (fan order) and returns
 * their number; *cull receives 0 visible, 1 frustum, 2 backface, 3 between
 * samples, 4 clipped-to-nothing, 5 clip path. */
int32_t  cro_setup_triangle(const cro_pipe* pipe, const float v0[4], const float v1[4],
                            const float v2[4], const int32_t vidx[3],
                            cro_header hdr[7], cro_data data[7], int32_t* cull);

/* Whole-buffer setup: slot t holds triangle t when it has one subtriangle;
 * triangles with >= 2 subtriangles get consecutive slots starting at
 * numTris + (prefix over earlier clipped triangles) and hdr[t].misc = base.
 * subtriTri[slot] = source triangle index.  Returns total slots used
 * (numTris + clipped extra) or -1 if maxSlots is too small. */
int64_t  cro_setup(const cro_pipe* pipe, const void* vertices, int64_t numVertices,
                   const int32_t* indices, int64_t numTris,
                   uint8_t* triSubtris, cro_header* hdr, cro_data* data,
                   uint32_t* subtriTri, int64_t maxSlots, int32_t nthreads);

/* Full draw: setup + in-order rasterization into surfaces laid out as
 * (ceil8(W) * N) x ceil8(H) texels (CudaSurface.cpp:86-91).  color/depth
 * must hold the previous contents when deferredClear == 0.  triid may be
 * NULL; when given it receives the source triangle index of the last
 * fragment that passed for each sample, 0xFFFFFFFF where none did (and is
 * cleared to 0xFFFFFFFF by a deferred clear).  Returns 0 or -1 on bad
 * arguments. */
int32_t  cro_render(const cro_pipe* pipe, const void* vertices, int64_t numVertices,
                    const int32_t* indices, int64_t numTris,
                    uint32_t* color, uint32_t* depth, uint32_t* triid,
                    int32_t nthreads, cro_stats* stats);

/* Passthrough vertex shader (test/shader/PassThrough.cu:16-35): the clip
 * position of model position p (Vec3f) is posToClip * (p, 1) with
 * Mat4f * Vec4f of base/Math.hpp:990-1001 -- per row i, rr = 0, then
 * rr += m(i, j) * v[j] for j = 0..3, each step one fused multiply-add (nvcc's
 * default contraction).  posToClip is in Mat4f memory order (column-major:
 * m(r, c) = m[r + 4c]).  Clip positions are written as 4 floats at the
 * start of each outStrideFloats-float output record; the rest of the record
 * is left as it is. */
void     cro_vertex_passthrough(const float* modelPos, int64_t numVertices, const float posToClip[16],
                                float* out, int64_t outStrideFloats);

#ifdef __cplusplus
}
#endif
#endif
