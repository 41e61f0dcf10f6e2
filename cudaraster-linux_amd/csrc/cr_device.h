// Device-side data layouts and rasterization rules for the gfx950 pipeline.
//
// The rules are the reference's (cudaraster/cuda/TriangleSetup.inl,
// Util.hpp, Util.inl, FineRaster.inl); the layouts of the per-triangle
// records keep CRTriangleHeader / CRTriangleData (PrivateDefs.hpp:26-62) so
// that setup output can be compared field by field.  Everything else --
// queues, lists, counters -- is this library's own design.
//
// Float arithmetic: the library is compiled with -ffp-contract=off, IEEE
// division and denormals preserved.  Fused multiply-adds appear only as
// explicit fmaf() at the places where nvcc's default contraction fuses the
// reference's `a*b + c` expressions (left product fused in `a*b + c*d`).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace crhip {

// Constants.hpp:21-84
constexpr int kSubpixelLog2 = 4;
constexpr int kTileLog2 = 3;                  // 8x8 pixels per tile
constexpr int kBinLog2 = 4;                   // largest bin: 16x16 tiles (per draw: Params::binLog2)
constexpr int kMaxViewport = 2048;
constexpr int kMaxBinsLog2 = 4;               // <= 16x16 bins
constexpr int kMaxBins = 1 << (2 * kMaxBinsLog2);
constexpr int kBinTiles = 1 << (2 * kBinLog2);
// Tile lists are chains of 32-entry segments (the fine stage's chunk).
constexpr int kTileSegLog2 = 5;
constexpr int kTileSeg = 1 << kTileSegLog2;
constexpr uint32_t kNoSeg = 0xFFFFFFFFu;
constexpr uint32_t kDepthMin = 2200u << 3;
constexpr uint32_t kDepthMax = 0xFFFFFFFFu - (2200u << 3);
constexpr int kBaryMax = (1 << (30 - kSubpixelLog2)) - 1;
__host__ __device__ constexpr uint32_t lerpError(int s) { return 2200u << s; }

enum : uint32_t { kFlagDepth = 1, kFlagLerp = 2, kFlagQuads = 4 };
enum : int32_t { kBlendReplace = 0, kBlendSrcOver = 1, kBlendAdditive = 2, kBlendDepthOnly = 3 };
enum : int32_t { kShaderGouraud = 0, kShaderConstant = 1, kShaderTexPhong = 2 };

// TextureSpec and Material (test/shader/Shaders.hpp:33-51), the layout of
// cr_texture_spec / cr_material.
struct TexSpec {
  float size[2];         // level-0 texels; size[0] == 0: no texture
  uint32_t pad[2];
  float4 mip[16];        // per level: scaleX, scaleY, biasX, biasY in atlas coordinates
};
struct Material {
  float4 diffuse;
  float specular[3];
  float glossiness;
  TexSpec diffuseTexture;
  TexSpec alphaTexture;
};
static_assert(sizeof(TexSpec) == 272 && sizeof(Material) == 576, "material layout");

struct Header {          // CRTriangleHeader
  int16_t v0x, v0y, v1x, v1y, v2x, v2y;
  uint32_t misc;         // 1 subtri: zmin[31:12] | flips; >= 2: first slot
};
struct TriData {         // CRTriangleData
  uint32_t zx, zy, zb, zslope;
  int32_t wx, wy, wb;
  int32_t ux, uy, ub;
  int32_t vx, vy, vb;
  uint32_t vi0, vi1, vi2;
};
static_assert(sizeof(Header) == 16, "header layout");
static_assert(sizeof(TriData) == 64, "data layout");

// Device counters; one 256-byte block per context.
struct Counters {
  uint32_t numSubtris;     // slot allocator for clipped subtriangles (starts at numTris)
  uint32_t tileEntries;    // tile-list allocator: entries reserved, in whole 32-entry segments
  uint32_t overflow;       // bit 0 subtris, bit 1 bin entries, bit 2 tile entries
  uint32_t binEntries;     // bin-queue entry allocator
  uint32_t tileListEntries;  // tile-list entries written (the lists' total length)
  uint32_t pad0[3];
  uint64_t fragments;      // (subtriangle, pixel) pairs with >= 1 covered sample
  uint64_t visibleSubtris; // subtriangles that survive setup (counting draws only)
  uint64_t pad[26];
};
static_assert(sizeof(Counters) == 256, "counter block");

// Profiling counters (ProfilingMode_Counters): each one a numerator and a
// denominator, accumulated by the stages of a profiling draw, in the order of
// the report (cr_profile_counter in cr.h has the same order).  Percentages
// add 100 to the numerator per event and 1 to the denominator per candidate,
// as the reference's CR_COUNT (PixelPipe.inl:91-101, PrivateDefs.hpp:158-204).
enum ProfCounter : int {
  kProfSetupViewportCull = 0,
  kProfSetupBackfaceCull,
  kProfSetupBetweenPixelsCull,
  kProfSetupClipped,
  kProfSetupSamplesPerTri,
  kProfBinQueueOverflows,
  kProfBinTrisPerChunk,
  kProfBinTriBBArea,
  kProfBinTriSingleBin,
  kProfCoarseBins,
  kProfCoarseRoundsPerBin,
  kProfCoarseTrisPerRound,
  kProfCoarseTilesPerRound,
  kProfCoarseEmitsPerRound,
  kProfCoarseEmitsPerTri,
  kProfCoarseCaseA,
  kProfCoarseCaseB,
  kProfCoarseCaseC,
  kProfFineTriPerTile,
  kProfFineEarlyZCull,
  kProfFineEmptyCull,
  kProfFineZKill,
  kProfFineWarpUtil,
  kProfFineBlendRounds,
  kProfFineFragPerTri,
  kProfFineFragPerTile,
  kProfCount
};

// Adds (num, denom) to counter c of the profile block (64-bit atomics).
__device__ __forceinline__ void profAdd(unsigned long long* prof, int c, uint64_t num, uint64_t denom) {
  if (num) atomicAdd(&prof[2 * c], (unsigned long long)num);
  if (denom) atomicAdd(&prof[2 * c + 1], (unsigned long long)denom);
}

struct Params {
  // Input.
  int32_t numTris;
  int32_t numVertices;
  int32_t vertexStride;            // bytes
  const uint8_t* vertices;         // already offset
  const int32_t* indices;          // int3 per triangle, already offset
  // Geometry of the surfaces.
  int32_t viewportWidth, viewportHeight;
  int32_t widthPixels, heightPixels;   // rounded to tiles
  int32_t widthTiles, heightTiles, numTiles;
  int32_t binLog2;                 // bin = 2^binLog2 x 2^binLog2 tiles (<= kBinLog2; <= 256 bins)
  int32_t widthBins, heightBins, numBins;
  // Pipe.
  int32_t samplesLog2;
  uint32_t flags;
  int32_t blend;
  int32_t shader;
  uint32_t constColor;
  int32_t deferredClear;
  uint32_t clearColor, clearDepth;
  int32_t hiz;                     // 1: tile-level early depth rejection
  int32_t coarseThreads;           // coarse workgroup size (256 or 1024), 0: chosen from the draw
  uint32_t* tileFrags;             // per-tile fragment counts (counting draws only), else null
  uint32_t* chunkSubtris;          // per-setup-chunk visible subtriangles (counting draws only), else null
  unsigned long long* prof;        // profile counters, 2 * kProfCount (profiling draws only), else null
  // texPhong shader constants (Shaders.hpp:57-65): materials, per-triangle
  // material index, float4 texture atlas.
  const Material* materials;
  int32_t numMaterials;
  const int32_t* triMaterial;
  int64_t numTriMaterials;
  const float4* atlas;
  int32_t atlasWidth, atlasHeight;
  // Setup output.
  int32_t maxSubtris;
  uint8_t* triSubtris;
  Header* triHeader;
  TriData* triData;
  uint32_t* subtriTri;             // source triangle of slots >= numTris
  // Bin queue: chunk c (kSetupChunkTris triangles) owns a region starting at
  // chunkBase[c]; bin b's entries of chunk c are binCnt[b][c] slots at
  // region offset binOff[b][c], in submission order.
  int32_t numChunks;
  uint32_t* chunkBase;             // [numChunks]
  uint32_t* binCnt;                // [numBins][numChunks]
  uint32_t* binOff;                // [numBins][numChunks]
  int32_t maxBinEntries;
  uint4* binList;                  // per entry: slot, then the header's three snapped vertices
  // Tile lists in submission order, each a chain of kTileSeg-entry
  // segments: segment s holds tileList[s * kTileSeg ...]; tileSegNext[2s] is
  // the next segment of its tile (kNoSeg after the last) and
  // tileSegNext[2s + 1] the one after that (written once it exists; never
  // read past the list's end).
  uint32_t* tileStart;             // [numTiles] first segment (kNoSeg: empty list)
  uint32_t* tileCount;             // [numTiles] entries in the list
  int32_t maxTileEntries;          // tileList capacity; maxTileEntries / kTileSeg segments
  uint32_t* tileList;
  uint32_t* tileSegNext;           // [2 * maxTileEntries / kTileSeg]
  // Surfaces: (widthPixels * N) x heightPixels texels, row 0 at the bottom.
  uint32_t* color;
  uint32_t* depth;
  uint32_t* triid;                 // optional
  Counters* counters;
};

// ---------------------------------------------------------------------------
// Exact conversions (PTX cvt semantics restated).
// ---------------------------------------------------------------------------

// cvt.rni.sat.s32.f32: nearest-even, saturate, NaN -> 0.
__device__ __forceinline__ int32_t cvtRniSatS32(float a) {
  if (a != a) return 0;
  float r = __builtin_rintf(a);
  if (r >= 2147483648.0f) return INT32_MAX;
  if (r < -2147483648.0f) return INT32_MIN;
  return (int32_t)r;
}

// cvt.rni.sat.u32.f32
__device__ __forceinline__ uint32_t cvtRniSatU32(float a) {
  if (a != a) return 0u;
  float r = __builtin_rintf(a);
  if (r >= 4294967296.0f) return 0xFFFFFFFFu;
  if (r <= 0.0f) return 0u;
  return (uint32_t)r;
}

// (U32)float in the reference's device code: cvt.rzi.u32.f32 (saturating).
__device__ __forceinline__ uint32_t cvtRziSatU32(float a) {
  if (a != a || a <= 0.0f) return 0u;
  if (a >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)a;
}

__device__ __forceinline__ int32_t asr32(int32_t a, int s) { return a >> s; }   // arithmetic on AMDGPU
__device__ __forceinline__ int64_t asr64(int64_t a, int s) { return a >> s; }
__device__ __forceinline__ int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
__device__ __forceinline__ int64_t wmul64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
__device__ __forceinline__ uint32_t wabs32(int32_t a) { return a < 0 ? 0u - (uint32_t)a : (uint32_t)a; }

// toABGR (Util.inl:30-38): per channel floor(255*c + 1/2) of the exact
// product, saturated to [0,255], NaN -> 0.  Evaluated exactly in F64
// (24x8-bit product + 1/2), then v_cvt_u32_f64, which truncates toward zero
// (floor for x >= 0), saturates out-of-range values to [0, 2^32-1] and maps
// NaN to 0, and a min with 255: no branches.
__device__ __forceinline__ uint32_t abgrChannel(float c) {
  const double x = fma((double)c, 255.0, 0.5);
  uint32_t u;
  asm("v_cvt_u32_f64 %0, %1" : "=v"(u) : "v"(x));
  return min(u, 255u);
}
__device__ __forceinline__ uint32_t toABGR(float r, float g, float b, float a) {
  return abgrChannel(r) | (abgrChannel(g) << 8) | (abgrChannel(b) << 16) | (abgrChannel(a) << 24);
}

// blendABGR / blendABGRClamp with the stock factors (Util.inl:42-60,
// PixelPipe.inl:73-85): per channel x = s*a + d*(255-a) (src-over) or
// x = min(255*(s+d), 255*255) (additive), then (x * 0x010101 + 0x800000) >> 24.
// Both x fit 16 bits, so two channels share one packed 16-bit operation:
//  - round(x / 255) = (t + (t >> 8)) >> 8 with t = x + 128 (no intermediate
//    above 65407) equals (x * 0x010101 + 0x800000) >> 24 for every x in
//    [0, 255*255] (checked exhaustively);
//  - for additive x = 255*k with k = min(s+d, 255) and that rounding gives k
//    back, so the channel is the saturating sum min(s+d, 255).
typedef unsigned short crU16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ crU16x2 bytes02(uint32_t v) {   // bytes 0 and 2, one per half
  return __builtin_bit_cast(crU16x2, v & 0x00FF00FFu);
}
__device__ __forceinline__ crU16x2 bytes13(uint32_t v) {   // bytes 1 and 3 (selector 0x0C: zero)
  return __builtin_bit_cast(crU16x2, __builtin_amdgcn_perm(0u, v, 0x0C030C01u));
}
__device__ __forceinline__ uint32_t packBytes(crU16x2 lo, crU16x2 hi) {
  return __builtin_bit_cast(uint32_t, lo) | (__builtin_bit_cast(uint32_t, hi) << 8);
}
__device__ __forceinline__ uint32_t blendPixel(int32_t mode, uint32_t src, uint32_t dst) {
  if (mode == kBlendReplace) return src;
  if (mode == kBlendDepthOnly) return dst;
  const crU16x2 s02 = bytes02(src), s13 = bytes13(src), d02 = bytes02(dst), d13 = bytes13(dst);
  if (mode == kBlendSrcOver) {
    const unsigned short a = (unsigned short)(src >> 24), ia = (unsigned short)(255u - a);
    const crU16x2 va = {a, a}, via = {ia, ia};
    // The bias is kept opaque so that it folds into the first multiply-add
    // instead of costing an add of its own.
    uint32_t bias = 0x00800080u;
    __asm__("" : "+s"(bias));
    const crU16x2 half = __builtin_bit_cast(crU16x2, bias);
    crU16x2 x02 = s02 * va + (d02 * via + half);
    crU16x2 x13 = s13 * va + (d13 * via + half);
    x02 = (x02 + (x02 >> (unsigned short)8)) >> (unsigned short)8;
    x13 = (x13 + (x13 >> (unsigned short)8)) >> (unsigned short)8;
    return packBytes(x02, x13);
  }
  const crU16x2 cap = {(unsigned short)255, (unsigned short)255};
  return packBytes(__builtin_elementwise_min(s02 + d02, cap), __builtin_elementwise_min(s13 + d13, cap));
}

// c_msaaPatterns (Util.hpp:27-36), packed 4 bits per sample.
__host__ __device__ constexpr uint32_t msaaPatternPacked(int s) {
  return s == 0 ? 0x0u : s == 1 ? 0x10u : s == 2 ? 0x2031u : 0x51360427u;
}
__host__ __device__ constexpr int msaaPattern(int s, int i) {
  return (int)((msaaPatternPacked(s) >> (4 * i)) & 0xFu);
}

// selectMSAACentroid (Util.hpp:182-203)
__host__ __device__ inline int selectCentroid(int s, uint32_t mask) {
  int n = 1 << s;
  if (mask == 0 || mask == (1u << n) - 1u) return -1;
  int best = -1, bestDist = 0x7FFFFFFF;
  for (int i = 0; i < n; i++) {
    if (mask & (1u << i)) {
      int dx = msaaPattern(s, i) * 2 + 1 - n;
      int dy = i * 2 + 1 - n;
      int dist = dx * dx + dy * dy;
      if (dist < bestDist) { best = i; bestDist = dist; }
    }
  }
  return best;
}

// Host-side launchers, one per stage file.
void launchReset(const Params& P, hipStream_t stream);
void launchSetup(const Params& P, hipStream_t stream);     // setup + bin
void launchCoarse(const Params& P, hipStream_t stream);
void launchFine(const Params& P, hipStream_t stream);
void launchFragmentSum(const Params& P, hipStream_t stream);
constexpr int kSetupChunkTris = 1024;

// Batched draws: `views` is a device array of per-view Params, each stage
// one launch over the workgroups of all views.  A stage lists its views in
// its own order: the k-th listed view is views[order[k]] and owns the
// workgroups [blockStart[k], blockStart[k+1]).
void launchResetBatch(const Params* views, int numViews, hipStream_t stream);
void launchSetupBatch(const Params* views, const int32_t* blockStart, const int32_t* order, int numViews,
                      int totalBlocks, hipStream_t stream);
void launchCoarseBatch(int threads, const Params* views, const int32_t* blockStart, const int32_t* order,
                       int numViews, int totalBlocks, hipStream_t stream);
constexpr int kCoarseShapes[2] = {256, 1024};    // coarse workgroup sizes (lanes) compiled
int coarseShapeFor(const Params& P);             // the plain draw's choice
// Fine-stage variant of a draw: (samplesLog2 * 2 + deferred shading) * 2 + depth test.
constexpr int kFineVariants = 16;
// Tile groups per workgroup of a batched fine pass: the workgroup walks them
// in turn, so its view lookup and parameter loads serve several groups and
// fewer workgroups are dispatched.
constexpr int kFineRounds = 4;
int fineVariant(const Params& P);
void launchFineBatch(int variant, int tilesPerGroup, int banded, const Params* views, const int32_t* groupStart,
                     const int32_t* order, int numViews, int totalGroups, hipStream_t stream);
constexpr int kFineBatchShapes[3] = {2, 4, 8};   // tiles (waves) per workgroup the batch path can use

// Differentiable interpolation over a single-sample triangle-id surface.
struct InterpGeom {
  int32_t width, height;        // viewport = surface size
  int32_t texW;                 // texel pitch of the id surface (rounded width, one sample)
  int32_t numChannels;
  int32_t numTris;
  int64_t numVertices;
};
// Box MSAA resolve of an RGBA8 surface into width x height pixels.
void launchResolve(const uint32_t* tex, int32_t textureW, int32_t width, int32_t height, int32_t samplesLog2,
                   uint32_t* out, hipStream_t stream);
// Passthrough vertex shader: clip position posToClip * (p, 1) of each
// model position into the first 16 bytes of each outStride-byte record.
void launchVertexPassthrough(const float* modelPos, int64_t numVertices, const float posToClip[16], void* out,
                             int64_t outStride, hipStream_t stream);
void launchInterpForward(const InterpGeom& G, const uint32_t* ids, const float* pos, const float* attr,
                         const int32_t* tris, float* out, int32_t* triImage, hipStream_t stream);
// gradPos (4 per vertex) and gradAttr (C per vertex) are overwritten.  The
// per-(tile, triangle) FP64 sums are added with FP32 atomics (acc null), or,
// bit-reproducibly, as 128-bit fixed-point integers in acc (2 * (4 + C) * V
// 64-bit words) rounded to FP32 at the end.
void launchInterpBackward(const InterpGeom& G, const int32_t* triImage, const float* pos, const float* attr,
                          const int32_t* tris, const float* gradOut, float* gradPos, float* gradAttr,
                          unsigned long long* acc, hipStream_t stream);

// Largest v with start[v] <= b: the view owning workgroup (or group) b.
// start has numViews + 1 nondecreasing entries; views without work have
// start[v] == start[v + 1] and are never returned for b < start[numViews].
__device__ __forceinline__ int findView(const int32_t* __restrict__ start, int numViews, int b) {
  int lo = 0, hi = numViews - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace crhip
