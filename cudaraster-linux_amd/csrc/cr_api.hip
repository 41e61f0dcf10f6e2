// Host runtime behind the C ABI (include/cr/cr.h): state, validation,
// capacity management with grow-and-retry, stage launches and statistics.
//
// Mirrors FW::CudaRaster (src/cudaraster/CudaRaster.cpp): the same state
// machine, the same validity checks, the same overflow-retry contract
// (:277-339), per-stage events for getStats (:346-363).  Buffers are plain
// device allocations owned by the context (the reference's lazily migrating
// FW::Buffer is not needed: inputs are device pointers, surfaces are device
// buffers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "cr/cr.h"
#include "cr_device.h"
#include "cr_host.h"

using namespace crhip;

namespace crhost {

thread_local std::string g_lastError;

cr_status fail(cr_context* ctx, cr_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  g_lastError = buf;
  return st;
}

namespace {

int popc8(int v) {
  int c = 0;
  for (int i = 0; i < 8; i++) c += (v >> i) & 1;
  return c;
}

// Vec4f::toABGR (base/Math.cpp:41-48), the host-side quantiser used by
// deferredClear.
uint32_t hostChannel(float c) {
  float cl = c;
  if (!(cl >= 0.0f)) cl = 0.0f;
  if (cl > 1.0f) cl = 1.0f;
  uint64_t q = (uint64_t)((double)cl * 72057594037927936.0);
  return (uint32_t)((((q * 255u) >> 55) + 1u) >> 1);
}

// encodeDepth (cudaraster/cuda/Util.hpp:290-298)
uint32_t encodeDepth(uint32_t depth) {
  double v = (double)depth;
  v /= 65536.0 * 65536.0 - 1.0;
  v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
  v *= (double)(kDepthMax - kDepthMin);
  v += (double)kDepthMin;
  return (uint32_t)v;
}

}  // namespace

void clearValues(const float rgba[4], float depth, uint32_t* color, uint32_t* encodedDepth) {
  *color = hostChannel(rgba[0]) | (hostChannel(rgba[1]) << 8) | (hostChannel(rgba[2]) << 16) |
           (hostChannel(rgba[3]) << 24);
  double d = (double)depth * 4294967296.0;
  uint64_t q;
  if (!(d > 0.0)) q = 0;
  else if (d >= 18446744073709551616.0) q = UINT64_MAX;
  else q = (uint64_t)d;
  *encodedDepth = encodeDepth((uint32_t)std::min<uint64_t>(q, 0xFFFFFFFFull));
}

cr_status validatePipe(cr_context* ctx, const cr_pipe_desc* p) {
  if (p->samples_log2 < 0 || p->samples_log2 > 3) return fail(ctx, CR_ERR_INVALID, "pipe: samples_log2 must be 0..3");
  if (p->vertex_struct_size < 16 || (p->vertex_struct_size & 15))
    return fail(ctx, CR_ERR_INVALID, "pipe: vertex_struct_size must be a positive multiple of 16 (Vec4f records)");
  if (p->fragment_shader != CR_SHADER_GOURAUD && p->fragment_shader != CR_SHADER_CONSTANT &&
      p->fragment_shader != CR_SHADER_TEX_PHONG)
    return fail(ctx, CR_ERR_INVALID, "pipe: unknown fragment shader %d", p->fragment_shader);
  if (p->fragment_shader == CR_SHADER_GOURAUD && p->vertex_struct_size < 32)
    return fail(ctx, CR_ERR_INVALID, "pipe: GouraudShader needs a colour varying (vertex_struct_size >= 32)");
  if (p->fragment_shader == CR_SHADER_TEX_PHONG && p->vertex_struct_size < 64)
    return fail(ctx, CR_ERR_INVALID,
                "pipe: texPhong needs camera position, normal and texture coordinate varyings (vertex_struct_size >= 64)");
  if (p->blend_shader < CR_BLEND_REPLACE || p->blend_shader > CR_BLEND_DEPTH_ONLY)
    return fail(ctx, CR_ERR_INVALID, "pipe: unknown blend shader %d", p->blend_shader);
  if (p->render_mode_flags & ~(uint32_t)(CR_RENDER_DEPTH | CR_RENDER_LERP | CR_RENDER_QUADS))
    return fail(ctx, CR_ERR_INVALID, "pipe: unknown render mode flags 0x%x", p->render_mode_flags);
  return CR_OK;
}

}  // namespace crhost

using namespace crhost;

namespace crhost {

bool countingDraw(const cr_context* ctx) {
  return ctx->dbg.count_fragments != 0 || ctx->dbg.profiling_mode == CR_PROFILING_COUNTERS;
}

bool profilingDraw(const cr_context* ctx) { return ctx->dbg.profiling_mode == CR_PROFILING_COUNTERS; }

cr_status validateDraw(cr_context* ctx) {
  if (!ctx->color) return fail(ctx, CR_ERR_STATE, "CudaRaster: Surfaces not set!");
  if (!ctx->pipeSet) return fail(ctx, CR_ERR_STATE, "CudaRaster: Pixel pipe not set!");
  if (!ctx->vbuf && ctx->numTris > 0) return fail(ctx, CR_ERR_STATE, "CudaRaster: Vertex buffer not set!");
  if (!ctx->ibuf && ctx->numTris > 0) return fail(ctx, CR_ERR_STATE, "CudaRaster: Index buffer not set!");
  if (ctx->pipe.samples_log2 != ctx->color->samplesLog2)
    return fail(ctx, CR_ERR_STATE, "CudaRaster: Mismatch in multisampling between pixel pipe and surface!");
  if (ctx->triid && (ctx->triid->width != ctx->color->width || ctx->triid->height != ctx->color->height ||
                     ctx->triid->samples != ctx->color->samples))
    return fail(ctx, CR_ERR_STATE, "triangle-id surface does not match the colour surface");
  if (ctx->pipe.fragment_shader == CR_SHADER_TEX_PHONG && !ctx->shaderConstantsSet)
    return fail(ctx, CR_ERR_STATE, "texPhong: shader constants (materials, atlas) not set");
  return CR_OK;
}

}  // namespace crhost

namespace {

cr_status ensureBuffers(cr_context* ctx, const Geometry& g, int32_t numChunks) {
  int64_t nt = ctx->numTris;
  CR_HIP(ctx, ctx->triSubtris.ensure((size_t)std::max<int64_t>(nt, 1)));
  CR_HIP(ctx, ctx->triHeader.ensure((size_t)ctx->maxSubtris * sizeof(Header)));
  CR_HIP(ctx, ctx->triData.ensure((size_t)ctx->maxSubtris * sizeof(TriData)));
  CR_HIP(ctx, ctx->subtriTri.ensure((size_t)std::max<int64_t>(ctx->maxSubtris - nt, 1) * 4));
  CR_HIP(ctx, ctx->chunkBase.ensure((size_t)std::max(numChunks, 1) * 4));
  CR_HIP(ctx, ctx->binCnt.ensure((size_t)g.nb * std::max(numChunks, 1) * 4));
  CR_HIP(ctx, ctx->binOff.ensure((size_t)g.nb * std::max(numChunks, 1) * 4));
  CR_HIP(ctx, ctx->binList.ensure((size_t)ctx->maxBinEntries * 16));
  CR_HIP(ctx, ctx->tileStart.ensure((size_t)g.nt * 4));
  CR_HIP(ctx, ctx->tileCount.ensure((size_t)g.nt * 4));
  CR_HIP(ctx, ctx->tileList.ensure((size_t)ctx->maxTileEntries * 4));
  CR_HIP(ctx, ctx->tileSegNext.ensure((size_t)std::max<int64_t>(ctx->maxTileEntries >> kTileSegLog2, 1) * 8));
  if (countingDraw(ctx)) {
    CR_HIP(ctx, ctx->tileFrags.ensure((size_t)g.nt * 4));
    CR_HIP(ctx, ctx->chunkSubtris.ensure((size_t)std::max(numChunks, 1) * 4));
  }
  if (profilingDraw(ctx)) CR_HIP(ctx, ctx->profBuf.ensure(2 * kProfCount * sizeof(unsigned long long)));
  return CR_OK;
}

}  // namespace

namespace crhost {

DrawBuffers contextBuffers(cr_context* ctx) {
  DrawBuffers B;
  B.maxSubtris = ctx->maxSubtris;
  B.maxBinEntries = ctx->maxBinEntries;
  B.maxTileEntries = ctx->maxTileEntries;
  B.triSubtris = ctx->triSubtris.as<uint8_t>();
  B.triHeader = ctx->triHeader.as<Header>();
  B.triData = ctx->triData.as<TriData>();
  B.subtriTri = ctx->subtriTri.as<uint32_t>();
  B.chunkBase = ctx->chunkBase.as<uint32_t>();
  B.binCnt = ctx->binCnt.as<uint32_t>();
  B.binOff = ctx->binOff.as<uint32_t>();
  B.binList = ctx->binList.as<uint4>();
  B.tileStart = ctx->tileStart.as<uint32_t>();
  B.tileCount = ctx->tileCount.as<uint32_t>();
  B.tileList = ctx->tileList.as<uint32_t>();
  B.tileSegNext = ctx->tileSegNext.as<uint32_t>();
  B.tileFrags = countingDraw(ctx) ? ctx->tileFrags.as<uint32_t>() : nullptr;
  B.chunkSubtris = countingDraw(ctx) ? ctx->chunkSubtris.as<uint32_t>() : nullptr;
  B.prof = profilingDraw(ctx) ? ctx->profBuf.as<unsigned long long>() : nullptr;
  B.counters = ctx->counters.as<Counters>();
  return B;
}

Params makeParams(cr_context* ctx, const Geometry& g, int32_t numChunks, const DrawBuffers& B) {
  Params P;
  memset(&P, 0, sizeof(P));
  P.numTris = ctx->numTris;
  P.numVertices = (int32_t)std::min<int64_t>(ctx->numVerts, INT32_MAX);
  P.vertexStride = ctx->pipe.vertex_struct_size;
  P.vertices = ctx->vbuf ? ctx->vbuf + ctx->vofs : nullptr;
  P.indices = ctx->ibuf ? reinterpret_cast<const int32_t*>(ctx->ibuf + ctx->iofs) : nullptr;
  P.viewportWidth = g.vw;
  P.viewportHeight = g.vh;
  P.widthPixels = g.wp;
  P.heightPixels = g.hp;
  P.widthTiles = g.wt;
  P.heightTiles = g.ht;
  P.numTiles = g.nt;
  P.binLog2 = g.bl;
  P.widthBins = g.wb;
  P.heightBins = g.hb;
  P.numBins = g.nb;
  P.samplesLog2 = ctx->pipe.samples_log2;
  P.flags = ctx->pipe.render_mode_flags & (kFlagDepth | kFlagLerp | kFlagQuads);
  P.blend = ctx->pipe.blend_shader;
  P.shader = ctx->pipe.fragment_shader;
  P.constColor = ctx->pipe.const_color;
  if (ctx->pipe.fragment_shader == CR_SHADER_TEX_PHONG) {
    const cr_shader_constants& k = ctx->shaderConstants;
    P.materials = reinterpret_cast<const Material*>(k.materials);
    P.numMaterials = k.num_materials;
    P.triMaterial = k.triangle_material_idx;
    P.numTriMaterials = k.num_triangle_materials;
    P.atlas = reinterpret_cast<const float4*>(k.atlas);
    P.atlasWidth = k.atlas_width;
    P.atlasHeight = k.atlas_height;
  }
  P.deferredClear = ctx->deferredClear ? 1 : 0;
  P.clearColor = ctx->clearColor;
  P.clearDepth = ctx->clearDepth;
  P.hiz = (ctx->dbg.disable_hiz || ctx->dbg.count_fragments) ? 0 : 1;
  P.coarseThreads = ctx->dbg.coarse_threads;
  P.tileFrags = countingDraw(ctx) ? B.tileFrags : nullptr;
  P.chunkSubtris = countingDraw(ctx) ? B.chunkSubtris : nullptr;
  P.prof = profilingDraw(ctx) ? B.prof : nullptr;
  P.maxSubtris = (int32_t)B.maxSubtris;
  P.triSubtris = B.triSubtris;
  P.triHeader = B.triHeader;
  P.triData = B.triData;
  P.subtriTri = B.subtriTri;
  P.numChunks = numChunks;
  P.chunkBase = B.chunkBase;
  P.binCnt = B.binCnt;
  P.binOff = B.binOff;
  P.maxBinEntries = (int32_t)B.maxBinEntries;
  P.binList = B.binList;
  P.tileStart = B.tileStart;
  P.tileCount = B.tileCount;
  P.maxTileEntries = (int32_t)B.maxTileEntries;
  P.tileList = B.tileList;
  P.tileSegNext = B.tileSegNext;
  P.color = ctx->color->dev;
  P.depth = ctx->depth->dev;
  P.triid = ctx->triid ? ctx->triid->dev : nullptr;
  P.counters = B.counters;
  return P;
}

cr_status launchAll(cr_context* ctx, const Params& P, hipStream_t stream) {
  CR_HIP(ctx, hipEventRecord(ctx->ev[0], stream));
  launchReset(P, stream);
  launchSetup(P, stream);                       // setup and binning are one kernel
  CR_HIP(ctx, hipEventRecord(ctx->ev[1], stream));   // (no separate bin stage to bracket)
  launchCoarse(P, stream);
  CR_HIP(ctx, hipEventRecord(ctx->ev[3], stream));
  launchFine(P, stream);
  CR_HIP(ctx, hipEventRecord(ctx->ev[4], stream));
  launchFragmentSum(P, stream);
  CR_HIP(ctx, hipGetLastError());
  CR_HIP(ctx, hipMemcpyAsync(ctx->hostCounters, P.counters, sizeof(Counters), hipMemcpyDeviceToHost, stream));
  if (P.prof)
    CR_HIP(ctx, hipMemcpyAsync(ctx->hostProf, P.prof, 2 * kProfCount * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, stream));
  ctx->evRecorded = true;
  ctx->lastProfiled = P.prof != nullptr;
  return CR_OK;
}


uint32_t recordCounters(cr_context* ctx) {
  const Counters& c = *ctx->hostCounters;
  ctx->last.num_tris = ctx->numTris;
  ctx->last.num_subtris = c.numSubtris;
  ctx->last.bin_entries = c.binEntries;
  ctx->last.tile_entries = c.tileEntries;
  ctx->last.tile_list_entries = c.tileListEntries;
  ctx->last.fragments = (int64_t)c.fragments;
  ctx->last.visible_subtris = (int64_t)c.visibleSubtris;
  memset(&ctx->lastProf, 0, sizeof(ctx->lastProf));
  if (ctx->lastProfiled) {
    for (int i = 0; i < kProfCount; i++) {
      ctx->lastProf.num[i] = (int64_t)ctx->hostProf[2 * i];
      ctx->lastProf.denom[i] = (int64_t)ctx->hostProf[2 * i + 1];
    }
  }
  return c.overflow;
}

}  // namespace crhost

namespace {

cr_status finishDraw(cr_context* ctx, bool* overflow) {
  const Counters& c = *ctx->hostCounters;
  *overflow = recordCounters(ctx) != 0;
  if (!*overflow) return CR_OK;
  if (c.overflow & 1u) ctx->maxSubtris = std::max<int64_t>(ctx->maxSubtris, (int64_t)c.numSubtris + kSubtrisSlack);
  if (c.overflow & 2u) ctx->maxBinEntries = std::max<int64_t>(ctx->maxBinEntries, (int64_t)c.binEntries + kBinSlack);
  if (c.overflow & 4u) ctx->maxTileEntries = std::max<int64_t>(ctx->maxTileEntries, (int64_t)c.tileEntries + kTileSlack);
  if (ctx->maxSubtris > kMaxSubtris)
    return fail(ctx, CR_ERR_OVERFLOW, "CudaRaster: more than 2^28 subtriangles in one draw");
  if (ctx->maxBinEntries > kMaxEntries || ctx->maxTileEntries > kMaxEntries)
    return fail(ctx, CR_ERR_OVERFLOW, "CudaRaster: internal queue capacity exceeded (2^31 entries)");
  return CR_OK;
}

cr_status drawImpl(cr_context* ctx, hipStream_t stream, bool sync) {
  cr_status st = validateDraw(ctx);
  if (st != CR_OK) return st;
  CR_HIP(ctx, hipSetDevice(ctx->device));
  if (ctx->pending) {
    cr_status s2 = cr_draw_check(ctx);
    if (s2 != CR_OK && s2 != CR_ERR_OVERFLOW) return s2;
  }
  Geometry g = geometryOf(ctx->color);
  int32_t numChunks = (ctx->numTris + kSetupChunkTris - 1) / kSetupChunkTris;
  initialCapacities(ctx->numTris, g, ctx->maxSubtris, ctx->maxBinEntries, ctx->maxTileEntries);
  ctx->last.retries = 0;
  for (;;) {
    st = ensureBuffers(ctx, g, numChunks);
    if (st != CR_OK) return st;
    Params P = makeParams(ctx, g, numChunks, contextBuffers(ctx));
    st = launchAll(ctx, P, stream);
    if (st != CR_OK) return st;
    if (!sync) {
      // Waiting on this event (not on the stream) lets cr_draw_check return
      // as soon as the draw is done, whatever was queued after it.
      CR_HIP(ctx, hipEventRecord(ctx->evPending, stream));
      ctx->pending = true;
      ctx->pendingStream = stream;
      ctx->deferredClear = false;
      return CR_OK;
    }
    CR_HIP(ctx, hipStreamSynchronize(stream));
    bool overflow = false;
    st = finishDraw(ctx, &overflow);
    if (st != CR_OK) return st;
    if (!overflow) break;
    ctx->last.retries++;
  }
  ctx->deferredClear = false;
  return CR_OK;
}

}  // namespace

extern "C" {

int32_t cr_abi_version(void) { return (1 << 16) | 6; }

const char* cr_last_error(const cr_context* ctx) { return ctx ? ctx->err.c_str() : g_lastError.c_str(); }

cr_status cr_context_create(cr_context** out) {
  if (!out) return fail(nullptr, CR_ERR_INVALID, "cr_context_create: out is NULL");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(nullptr, CR_ERR_DEVICE, "CudaRaster: No HIP-capable devices found!");
  cr_context* ctx = new (std::nothrow) cr_context();
  if (!ctx) return fail(nullptr, CR_ERR_NOMEM, "out of host memory");
  hipError_t e = hipGetDevice(&ctx->device);
  // The no-allocation draw's scratch array on this device (looked up once;
  // a failed lookup is reported by that entry point, not here).
  if (e == hipSuccess) scratchInit(ctx->device);
  if (e == hipSuccess) e = ctx->counters.ensure(sizeof(Counters));
  if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->hostCounters, sizeof(Counters), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&ctx->hostProf, 2 * kProfCount * sizeof(unsigned long long), hipHostMallocDefault);
  for (int i = 0; i < 5 && e == hipSuccess; i++) e = hipEventCreate(&ctx->ev[i]);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->evPending, hipEventDisableTiming);
  if (e != hipSuccess) {
    cr_context_destroy(ctx);
    return fail(nullptr, CR_ERR_DEVICE, "cr_context_create: %s", hipGetErrorString(e));
  }
  memset(ctx->hostCounters, 0, sizeof(Counters));
  memset(ctx->hostProf, 0, 2 * kProfCount * sizeof(unsigned long long));
  *out = ctx;
  return CR_OK;
}

cr_status cr_context_destroy(cr_context* ctx) {
  if (!ctx) return CR_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->pending) (void)hipStreamSynchronize(ctx->pendingStream);
  for (DevBuf* b : {&ctx->triSubtris, &ctx->triHeader, &ctx->triData, &ctx->subtriTri, &ctx->chunkBase,
                    &ctx->binCnt, &ctx->binOff, &ctx->binList, &ctx->tileStart, &ctx->tileCount,
                    &ctx->tileList, &ctx->tileSegNext, &ctx->tileFrags, &ctx->chunkSubtris, &ctx->counters,
                    &ctx->profBuf, &ctx->interpAcc})
    b->release();
  if (ctx->hostCounters) (void)hipHostFree(ctx->hostCounters);
  if (ctx->hostProf) (void)hipHostFree(ctx->hostProf);
  for (int i = 0; i < 5; i++)
    if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
  if (ctx->evPending) (void)hipEventDestroy(ctx->evPending);
  delete ctx;
  return CR_OK;
}

cr_status cr_surface_create(int32_t width, int32_t height, int32_t format, int32_t numSamples, cr_surface** out) {
  if (!out) return fail(nullptr, CR_ERR_INVALID, "cr_surface_create: out is NULL");
  *out = nullptr;
  // CudaSurface.cpp:44-62
  if (std::min(width, height) <= 0) return fail(nullptr, CR_ERR_INVALID, "CudaSurface: Size must be positive!");
  if (std::max(width, height) > kMaxViewport)
    return fail(nullptr, CR_ERR_INVALID, "CudaSurface: CR_MAXVIEWPORT_SIZE exceeded!");
  if (format < 0 || format > CR_FORMAT_TRIANGLE_ID) return fail(nullptr, CR_ERR_INVALID, "CudaSurface: Invalid format!");
  if (numSamples > 8) return fail(nullptr, CR_ERR_INVALID, "CudaSurface: numSamples cannot exceed 8!");
  if (numSamples < 1 || popc8(numSamples) != 1)
    return fail(nullptr, CR_ERR_INVALID, "CudaSurface: numSamples must be a power of two!");
  cr_surface* s = new (std::nothrow) cr_surface();
  if (!s) return fail(nullptr, CR_ERR_NOMEM, "out of host memory");
  s->width = width;
  s->height = height;
  s->format = format;
  s->samples = numSamples;
  s->samplesLog2 = popc8(numSamples - 1);
  s->roundedW = (width + 7) & ~7;
  s->roundedH = (height + 7) & ~7;
  s->textureW = s->roundedW * numSamples;
  s->bytes = (size_t)s->textureW * s->roundedH * 4;
  hipError_t e = hipGetDevice(&s->device);
  if (e == hipSuccess) e = hipMalloc((void**)&s->dev, s->bytes);
  if (e == hipSuccess) e = hipMemset(s->dev, 0, s->bytes);
  if (e != hipSuccess) {
    if (s->dev) (void)hipFree(s->dev);
    delete s;
    return fail(nullptr, CR_ERR_DEVICE, "cr_surface_create: %s", hipGetErrorString(e));
  }
  *out = s;
  return CR_OK;
}

cr_status cr_surface_destroy(cr_surface* s) {
  if (!s) return CR_OK;
  if (s->dev) {
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(s->dev);
  }
  delete s;
  return CR_OK;
}

cr_status cr_surface_get_info(const cr_surface* s, int32_t* width, int32_t* height, int32_t* rw, int32_t* rh,
                              int32_t* tw, int32_t* ns, int32_t* fmt) {
  if (!s) return fail(nullptr, CR_ERR_INVALID, "cr_surface_get_info: surface is NULL");
  if (width) *width = s->width;
  if (height) *height = s->height;
  if (rw) *rw = s->roundedW;
  if (rh) *rh = s->roundedH;
  if (tw) *tw = s->textureW;
  if (ns) *ns = s->samples;
  if (fmt) *fmt = s->format;
  return CR_OK;
}

cr_status cr_surface_get_device_ptr(const cr_surface* s, void** p, size_t* bytes) {
  if (!s) return fail(nullptr, CR_ERR_INVALID, "cr_surface_get_device_ptr: surface is NULL");
  if (p) *p = s->dev;
  if (bytes) *bytes = s->bytes;
  return CR_OK;
}

cr_status cr_surface_download(const cr_surface* s, void* host, size_t bytes) {
  if (!s || !host) return fail(nullptr, CR_ERR_INVALID, "cr_surface_download: NULL argument");
  if (bytes != s->bytes) return fail(nullptr, CR_ERR_INVALID, "cr_surface_download: %zu bytes, surface holds %zu", bytes, s->bytes);
  if (hipSetDevice(s->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(host, s->dev, bytes, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(nullptr, CR_ERR_DEVICE, "cr_surface_download: copy failed");
  return CR_OK;
}

cr_status cr_surface_upload(cr_surface* s, const void* host, size_t bytes) {
  if (!s || !host) return fail(nullptr, CR_ERR_INVALID, "cr_surface_upload: NULL argument");
  if (bytes != s->bytes) return fail(nullptr, CR_ERR_INVALID, "cr_surface_upload: %zu bytes, surface holds %zu", bytes, s->bytes);
  if (hipSetDevice(s->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(s->dev, host, bytes, hipMemcpyHostToDevice) != hipSuccess)
    return fail(nullptr, CR_ERR_DEVICE, "cr_surface_upload: copy failed");
  return CR_OK;
}

cr_status cr_set_surfaces(cr_context* ctx, cr_surface* color, cr_surface* depth) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_set_surfaces: context is NULL");
  ctx->color = nullptr;
  ctx->depth = nullptr;
  if (!color && !depth) return CR_OK;
  // CudaRaster.cpp:142-159
  if (!color) return fail(ctx, CR_ERR_INVALID, "CudaRaster: No color buffer specified!");
  if (!depth) return fail(ctx, CR_ERR_INVALID, "CudaRaster: No depth buffer specified!");
  if (color->format != CR_FORMAT_RGBA8) return fail(ctx, CR_ERR_INVALID, "CudaRaster: Unsupported color buffer format!");
  if (depth->format != CR_FORMAT_DEPTH32) return fail(ctx, CR_ERR_INVALID, "CudaRaster: Unsupported depth buffer format!");
  if (color->width != depth->width || color->height != depth->height)
    return fail(ctx, CR_ERR_INVALID, "CudaRaster: Mismatch in size between surfaces!");
  if (color->samples != depth->samples)
    return fail(ctx, CR_ERR_INVALID, "CudaRaster: Mismatch in multisampling between surfaces!");
  ctx->color = color;
  ctx->depth = depth;
  return CR_OK;
}

cr_status cr_set_triangle_id_surface(cr_context* ctx, cr_surface* tri) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_set_triangle_id_surface: context is NULL");
  if (tri && tri->format != CR_FORMAT_TRIANGLE_ID)
    return fail(ctx, CR_ERR_INVALID, "triangle-id surface must have CR_FORMAT_TRIANGLE_ID");
  ctx->triid = tri;
  return CR_OK;
}

cr_status cr_deferred_clear(cr_context* ctx, const float rgba[4], float depth) {
  if (!ctx || !rgba) return fail(ctx, CR_ERR_INVALID, "cr_deferred_clear: NULL argument");
  // CudaRaster.cpp:174-179
  ctx->deferredClear = true;
  clearValues(rgba, depth, &ctx->clearColor, &ctx->clearDepth);
  return CR_OK;
}

cr_status cr_set_pixel_pipe(cr_context* ctx, const cr_pipe_desc* p) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_set_pixel_pipe: context is NULL");
  if (!p) {
    ctx->pipeSet = false;
    return CR_OK;
  }
  cr_status st = validatePipe(ctx, p);
  if (st != CR_OK) return st;
  ctx->pipe = *p;
  ctx->pipeSet = true;
  return CR_OK;
}

cr_status cr_pipe_lookup(const char* name, int32_t blend, int32_t samplesLog2, uint32_t flags, cr_pipe_desc* out) {
  if (!name || !out) return fail(nullptr, CR_ERR_INVALID, "cr_pipe_lookup: NULL argument");
  cr_pipe_desc d{};
  d.blend_shader = blend;
  d.samples_log2 = samplesLog2;
  d.render_mode_flags = flags;
  if (strcmp(name, "PixelPipe_passthrough") == 0) {
    d.vertex_struct_size = 16;
    d.fragment_shader = CR_SHADER_CONSTANT;
    d.const_color = 0xFF0000FFu;   // toABGR(Vec4f(1,0,0,1))
  } else if (strcmp(name, "PixelPipe_gouraud") == 0) {
    d.vertex_struct_size = 32;
    d.fragment_shader = CR_SHADER_GOURAUD;
  } else if (strcmp(name, "PixelPipe_texPhong") == 0) {
    d.vertex_struct_size = 64;                 // ShadedVertex_texPhong (Shaders.hpp:84-90)
    d.fragment_shader = CR_SHADER_TEX_PHONG;
  } else {
    return fail(nullptr, CR_ERR_INVALID, "CudaRaster: Invalid pixel pipe!");
  }
  *out = d;
  return CR_OK;
}

cr_status cr_set_vertex_buffer(cr_context* ctx, const void* p, int64_t ofs, int64_t numVertices) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_set_vertex_buffer: context is NULL");
  if (ofs < 0 || (ofs & 15) || numVertices < 0)
    return fail(ctx, CR_ERR_INVALID, "cr_set_vertex_buffer: offset must be a non-negative multiple of 16");
  ctx->vbuf = static_cast<const uint8_t*>(p);
  ctx->vofs = ofs;
  ctx->numVerts = numVertices;
  return CR_OK;
}

cr_status cr_set_index_buffer(cr_context* ctx, const void* p, int64_t ofs, int32_t numTris) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_set_index_buffer: context is NULL");
  if (ofs < 0 || (ofs & 3) || numTris < 0)
    return fail(ctx, CR_ERR_INVALID, "cr_set_index_buffer: bad offset or triangle count");
  if (numTris > kMaxTris) return fail(ctx, CR_ERR_INVALID, "cr_set_index_buffer: more than 2^28 triangles");
  ctx->ibuf = static_cast<const uint8_t*>(p);
  ctx->iofs = ofs;
  ctx->numTris = numTris;
  return CR_OK;
}

cr_status cr_draw_triangles(cr_context* ctx, void* stream) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_draw_triangles: context is NULL");
  return drawImpl(ctx, static_cast<hipStream_t>(stream), true);
}

cr_status cr_draw_triangles_async(cr_context* ctx, void* stream) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_draw_triangles_async: context is NULL");
  return drawImpl(ctx, static_cast<hipStream_t>(stream), false);
}

cr_status cr_draw_check(cr_context* ctx) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_draw_check: context is NULL");
  if (!ctx->pending) return CR_OK;
  CR_HIP(ctx, hipEventSynchronize(ctx->evPending));
  ctx->pending = false;
  bool overflow = false;
  cr_status st = finishDraw(ctx, &overflow);
  if (st != CR_OK) return st;
  if (overflow) return fail(ctx, CR_ERR_OVERFLOW, "draw exceeded the reserved capacities; capacities grown, re-run it");
  return CR_OK;
}

cr_status cr_reserve(cr_context* ctx, int64_t maxSubtris, int64_t maxBin, int64_t maxTile) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_reserve: context is NULL");
  if (maxSubtris > kMaxSubtris || maxBin > kMaxEntries || maxTile > kMaxEntries)
    return fail(ctx, CR_ERR_INVALID, "cr_reserve: capacity above 2^28 subtriangles or 2^31 queue entries");
  ctx->maxSubtris = std::max(ctx->maxSubtris, maxSubtris);
  ctx->maxBinEntries = std::max(ctx->maxBinEntries, maxBin);
  ctx->maxTileEntries = std::max(ctx->maxTileEntries, maxTile);
  return CR_OK;
}

cr_status cr_get_stats(cr_context* ctx, cr_stats* out) {
  if (!ctx || !out) return fail(ctx, CR_ERR_INVALID, "cr_get_stats: NULL argument");
  memset(out, 0, sizeof(*out));
  if (!ctx->evRecorded) return CR_OK;
  CR_HIP(ctx, hipEventSynchronize(ctx->ev[4]));
  // Binning runs inside the setup kernel: its time is part of setup_time and
  // bin_time is 0 (an extra event pair between the two would cost the GPU
  // a few microseconds per draw).
  float setupMs = 0.0f, coarseMs = 0.0f, fineMs = 0.0f;
  CR_HIP(ctx, hipEventElapsedTime(&setupMs, ctx->ev[0], ctx->ev[1]));
  CR_HIP(ctx, hipEventElapsedTime(&coarseMs, ctx->ev[1], ctx->ev[3]));
  CR_HIP(ctx, hipEventElapsedTime(&fineMs, ctx->ev[3], ctx->ev[4]));
  out->setup_time = setupMs * 1e-3f;
  out->bin_time = 0.0f;
  out->coarse_time = coarseMs * 1e-3f;
  out->fine_time = fineMs * 1e-3f;
  return CR_OK;
}

cr_status cr_get_counters(cr_context* ctx, cr_counters* out) {
  if (!ctx || !out) return fail(ctx, CR_ERR_INVALID, "cr_get_counters: NULL argument");
  *out = ctx->last;
  return CR_OK;
}

cr_status cr_get_profiling_info(cr_context* ctx, char* buf, size_t len) {
  if (!ctx || !buf || len == 0) return fail(ctx, CR_ERR_INVALID, "cr_get_profiling_info: NULL argument");
  std::string out("\n");
  char line[256];
  auto add = [&](const char* fmt, auto... v) {
    snprintf(line, sizeof(line), fmt, v...);
    out += line;
  };
  if (!ctx->pipeSet) out += "Pixel pipe not set!\n";
  if (!ctx->lastProfiled) {
    // ProfilingMode_Default (CudaRaster.cpp:386-422): stage times with their
    // share of the draw, then the queue sizes.  Setup and binning are one
    // kernel here, so binRaster reads 0.
    cr_stats s;
    cr_status st = cr_get_stats(ctx, &s);
    if (st != CR_OK) return st;
    const cr_counters& c = ctx->last;
    const double total = (double)s.setup_time + s.bin_time + s.coarse_time + s.fine_time;
    const double pct = total > 0.0 ? 100.0 / total : 0.0;
    out += "ProfilingMode_Default\n---------------------\n\n";
    add("%-16s%.3f ms (%.0f%%)\n", "triangleSetup", s.setup_time * 1e3, s.setup_time * pct);
    add("%-16s%.3f ms (%.0f%%)\n", "binRaster", s.bin_time * 1e3, s.bin_time * pct);
    add("%-16s%.3f ms (%.0f%%)\n", "coarseRaster", s.coarse_time * 1e3, s.coarse_time * pct);
    add("%-16s%.3f ms (%.0f%%)\n", "fineRaster", s.fine_time * 1e3, s.fine_time * pct);
    out += "\n";
    const double mb = 1.0 / (1024.0 * 1024.0);
    add("%-16s%-10lld(%.1f MB)\n", "numSubtris", (long long)c.num_subtris,
        (double)c.num_subtris * (1 + sizeof(Header) + sizeof(TriData)) * mb);
    add("%-16s%-10lld(%.1f MB)\n", "numBinEntries", (long long)c.bin_entries, (double)c.bin_entries * 16 * mb);
    add("%-16s%-10lld(%.1f MB)\n", "numTileEntries", (long long)c.tile_entries, (double)c.tile_entries * 4 * mb);
    if (countingDraw(ctx)) add("%-16s%lld\n", "fragments", (long long)c.fragments);
    add("%-16s%lld\n", "retries", (long long)c.retries);
  } else {
    // ProfilingMode_Counters (CudaRaster.cpp:424-447): num / max(denom, 1)
    // per counter, in the reference's layout.
    cr_profile_counters pc;
    cr_status st = cr_get_profile_counters(ctx, &pc);
    if (st != CR_OK) return st;
    auto v = [&](int i) { return (double)pc.num[i] / std::max((double)pc.denom[i], 1.0); };
    out += "ProfilingMode_Counters\n----------------------\n\n";
    out += "TriangleSetup:\n";
    add("- Viewport cull        %.1f%%\n", v(CR_PROF_SETUP_VIEWPORT_CULL));
    add("- Backface cull        %.1f%%\n", v(CR_PROF_SETUP_BACKFACE_CULL));
    add("- Between pixels cull  %.1f%%\n", v(CR_PROF_SETUP_BETWEEN_PIXELS_CULL));
    add("- Clipped              %.1f%%\n", v(CR_PROF_SETUP_CLIPPED));
    add("- Avg. samples / tri   %.2f\n\n", v(CR_PROF_SETUP_SAMPLES_PER_TRI));
    out += "BinRaster:\n";
    add("- Queue overflows      %.0f\n", v(CR_PROF_BIN_QUEUE_OVERFLOWS));
    add("- Avg. triangles/chunk %.1f\n", v(CR_PROF_BIN_TRIS_PER_CHUNK));
    add("- Avg. tri bb size     %.1f\n", v(CR_PROF_BIN_TRI_BB_AREA));
    add("- Single-bin boxes     %.1f%%\n\n", v(CR_PROF_BIN_TRI_SINGLE_BIN));
    out += "CoarseRaster:\n";
    add("- Bins                 %.0f\n", v(CR_PROF_COARSE_BINS));
    add("- Rounds / Bin         %.1f\n", v(CR_PROF_COARSE_ROUNDS_PER_BIN));
    add("- Triangles / Round    %.1f\n", v(CR_PROF_COARSE_TRIS_PER_ROUND));
    add("- Tiles / Round        %.1f\n", v(CR_PROF_COARSE_TILES_PER_ROUND));
    add("- Emits / Round        %.1f\n", v(CR_PROF_COARSE_EMITS_PER_ROUND));
    add("- Emits / Triangle     %.2f\n", v(CR_PROF_COARSE_EMITS_PER_TRI));
    add("- Case A               %.0f%%\n", v(CR_PROF_COARSE_CASE_A));
    add("- Case B               %.0f%%\n", v(CR_PROF_COARSE_CASE_B));
    add("- Case C               %.0f%%\n\n", v(CR_PROF_COARSE_CASE_C));
    out += "FineRaster:\n";
    out += "- Triangles culled\n";
    add("  - Early Z kill       %.1f%%\n", v(CR_PROF_FINE_EARLY_Z_CULL));
    add("  - Empty coverage     %.1f%%\n", v(CR_PROF_FINE_EMPTY_CULL));
    add("- Z kills              %.1f%%\n", v(CR_PROF_FINE_Z_KILL));
    add("- Post-kill warp util. %.1f%%\n", v(CR_PROF_FINE_WARP_UTIL));
    add("- Blend attempt rounds %.2f\n", v(CR_PROF_FINE_BLEND_ROUNDS));
    add("- Avg. tri/tile        %.0f\n", v(CR_PROF_FINE_TRI_PER_TILE));
    add("- Avg. frag/tri        %.1f\n", v(CR_PROF_FINE_FRAG_PER_TRI));
    add("- Avg. frag/tile       %.0f\n", v(CR_PROF_FINE_FRAG_PER_TILE));
  }
  out += "\n";
  if (out.size() + 1 > len) {
    buf[0] = 0;
    return fail(ctx, CR_ERR_INVALID, "cr_get_profiling_info: the report needs %zu bytes", out.size() + 1);
  }
  memcpy(buf, out.c_str(), out.size() + 1);
  return CR_OK;
}

cr_status cr_get_profile_counters(cr_context* ctx, cr_profile_counters* out) {
  if (!ctx || !out) return fail(ctx, CR_ERR_INVALID, "cr_get_profile_counters: NULL argument");
  static_assert((int)CR_PROF_COUNT == (int)kProfCount, "profile counter tables");
  *out = ctx->lastProf;
  if (ctx->lastProfiled) {
    // Host-side counter: grow-and-retry rounds of the draw.
    out->num[CR_PROF_BIN_QUEUE_OVERFLOWS] = ctx->last.retries;
    out->denom[CR_PROF_BIN_QUEUE_OVERFLOWS] = 1;
  }
  return CR_OK;
}

cr_status cr_interpolate(cr_context* ctx, const cr_surface* triId, const float* pos, int64_t numVertices,
                         const float* attr, int32_t numChannels, const int32_t* tris, int32_t numTris, float* out,
                         int32_t* triImage, void* stream) {
  if (!ctx || !triId || !out) return fail(ctx, CR_ERR_INVALID, "cr_interpolate: NULL argument");
  if (triId->format != CR_FORMAT_TRIANGLE_ID || triId->samples != 1)
    return fail(ctx, CR_ERR_INVALID, "cr_interpolate: needs a single-sample CR_FORMAT_TRIANGLE_ID surface");
  // (Without vertices the kernels read no position or attribute: an empty
  // view may pass NULL for them.)
  if (numChannels < 0 || numVertices < 0 || numTris < 0 || (numChannels > 0 && numVertices > 0 && !attr) ||
      (numTris > 0 && (!pos || !tris)) || ((uintptr_t)pos & 15) || ((uintptr_t)tris & 3))
    return fail(ctx, CR_ERR_INVALID, "cr_interpolate: bad buffers or counts");
  InterpGeom G{triId->width, triId->height, triId->textureW, numChannels, numTris, numVertices};
  CR_HIP(ctx, hipSetDevice(ctx->device));
  launchInterpForward(G, triId->dev, pos, attr, tris, out, triImage, static_cast<hipStream_t>(stream));
  CR_HIP(ctx, hipGetLastError());
  return CR_OK;
}

cr_status cr_interpolate_ids(cr_context* ctx, const int32_t* triImage, int32_t width, int32_t height,
                             const float* pos, int64_t numVertices, const float* attr, int32_t numChannels,
                             const int32_t* tris, int32_t numTris, float* out, void* stream) {
  if (!ctx || !triImage || !out) return fail(ctx, CR_ERR_INVALID, "cr_interpolate_ids: NULL argument");
  if (width <= 0 || height <= 0 || width > kMaxViewport || height > kMaxViewport || numChannels < 0 ||
      numVertices < 0 || numTris < 0 || (numChannels > 0 && numVertices > 0 && !attr) ||
      (numTris > 0 && (!pos || !tris)) || ((uintptr_t)pos & 15) || ((uintptr_t)tris & 3) ||
      ((uintptr_t)triImage & 3))
    return fail(ctx, CR_ERR_INVALID, "cr_interpolate_ids: bad sizes or buffers");
  // An int32 id image is a single-sample id surface with a pitch of `width`
  // texels (-1 reads as 0xFFFFFFFF: no triangle).
  InterpGeom G{width, height, width, numChannels, numTris, numVertices};
  CR_HIP(ctx, hipSetDevice(ctx->device));
  launchInterpForward(G, reinterpret_cast<const uint32_t*>(triImage), pos, attr, tris, out, nullptr,
                      static_cast<hipStream_t>(stream));
  CR_HIP(ctx, hipGetLastError());
  return CR_OK;
}

namespace {

cr_status interpolateBackward(cr_context* ctx, const char* name, bool deterministic, const int32_t* triImage,
                              int32_t width, int32_t height, const float* pos, int64_t numVertices,
                              const float* attr, int32_t numChannels, const int32_t* tris, int32_t numTris,
                              const float* gradOut, float* gradPos, float* gradAttr, void* stream) {
  // (Without vertices there is no gradient to write and no attribute to read.)
  if (!ctx || !triImage || !gradOut || (numVertices > 0 && (!gradPos || (numChannels > 0 && (!attr || !gradAttr)))))
    return fail(ctx, CR_ERR_INVALID, "%s: NULL argument", name);
  if (width <= 0 || height <= 0 || width > kMaxViewport || height > kMaxViewport || numChannels < 0 ||
      numVertices < 0 || numTris < 0 || (numTris > 0 && (!pos || !tris)) || ((uintptr_t)pos & 15) ||
      ((uintptr_t)tris & 3))
    return fail(ctx, CR_ERR_INVALID, "%s: bad sizes or buffers", name);
  CR_HIP(ctx, hipSetDevice(ctx->device));
  InterpGeom G{width, height, width, numChannels, numTris, numVertices};
  unsigned long long* acc = nullptr;
  if (deterministic) {
    const int64_t words = 2 * numVertices * (4 + (int64_t)numChannels);
    CR_HIP(ctx, ctx->interpAcc.ensure((size_t)std::max<int64_t>(words, 1) * 8));
    acc = ctx->interpAcc.as<unsigned long long>();
  }
  launchInterpBackward(G, triImage, pos, attr, tris, gradOut, gradPos, gradAttr, acc,
                       static_cast<hipStream_t>(stream));
  CR_HIP(ctx, hipGetLastError());
  return CR_OK;
}

}  // namespace

cr_status cr_interpolate_backward(cr_context* ctx, const int32_t* triImage, int32_t width, int32_t height,
                                  const float* pos, int64_t numVertices, const float* attr, int32_t numChannels,
                                  const int32_t* tris, int32_t numTris, const float* gradOut, float* gradPos,
                                  float* gradAttr, void* stream) {
  return interpolateBackward(ctx, "cr_interpolate_backward", false, triImage, width, height, pos, numVertices,
                             attr, numChannels, tris, numTris, gradOut, gradPos, gradAttr, stream);
}

cr_status cr_interpolate_backward_deterministic(cr_context* ctx, const int32_t* triImage, int32_t width,
                                                int32_t height, const float* pos, int64_t numVertices,
                                                const float* attr, int32_t numChannels, const int32_t* tris,
                                                int32_t numTris, const float* gradOut, float* gradPos,
                                                float* gradAttr, void* stream) {
  return interpolateBackward(ctx, "cr_interpolate_backward_deterministic", true, triImage, width, height, pos,
                             numVertices, attr, numChannels, tris, numTris, gradOut, gradPos, gradAttr, stream);
}

cr_status cr_surface_resolve(const cr_surface* color, void* out, void* stream) {
  if (!color || !out) return fail(nullptr, CR_ERR_INVALID, "cr_surface_resolve: NULL argument");
  if (color->format != CR_FORMAT_RGBA8 || ((uintptr_t)out & 3))
    return fail(nullptr, CR_ERR_INVALID, "cr_surface_resolve: not an RGBA8 surface, or a misaligned output");
  CR_HIP(nullptr, hipSetDevice(color->device));
  launchResolve(color->dev, color->textureW, color->width, color->height, color->samplesLog2,
                static_cast<uint32_t*>(out), static_cast<hipStream_t>(stream));
  CR_HIP(nullptr, hipGetLastError());
  return CR_OK;
}

cr_status cr_vertex_shader_passthrough(cr_context* ctx, const float* modelPos, int64_t numVertices,
                                       const float* posToClip, void* outVertices, int32_t outStride,
                                       void* stream) {
  if (!ctx || !posToClip || (numVertices > 0 && (!modelPos || !outVertices)))
    return fail(ctx, CR_ERR_INVALID, "cr_vertex_shader_passthrough: NULL argument");
  if (numVertices < 0 || outStride < 16 || (outStride & 15) || ((uintptr_t)outVertices & 15) ||
      ((uintptr_t)modelPos & 3))
    return fail(ctx, CR_ERR_INVALID, "cr_vertex_shader_passthrough: bad size, stride or alignment");
  CR_HIP(ctx, hipSetDevice(ctx->device));
  launchVertexPassthrough(modelPos, numVertices, posToClip, outVertices, outStride, static_cast<hipStream_t>(stream));
  CR_HIP(ctx, hipGetLastError());
  return CR_OK;
}

cr_status cr_set_shader_constants(cr_context* ctx, const cr_shader_constants* c) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_set_shader_constants: context is NULL");
  if (!c) {
    ctx->shaderConstantsSet = false;
    return CR_OK;
  }
  static_assert(sizeof(cr_material) == sizeof(Material) && sizeof(cr_texture_spec) == sizeof(TexSpec),
                "material layout");
  if (c->num_materials < 0 || (c->num_materials > 0 && !c->materials) || c->num_triangle_materials < 0 ||
      (c->num_triangle_materials > 0 && !c->triangle_material_idx) || ((uintptr_t)c->materials & 15) ||
      ((uintptr_t)c->triangle_material_idx & 3))
    return fail(ctx, CR_ERR_INVALID, "cr_set_shader_constants: bad material table or index buffer");
  if (!c->atlas || ((uintptr_t)c->atlas & 15) || c->atlas_width < 1 || c->atlas_height < 1 ||
      c->atlas_width > 16384 || c->atlas_height > 16384)
    return fail(ctx, CR_ERR_INVALID, "cr_set_shader_constants: the atlas must be a 16-byte aligned device "
                                     "buffer of 1..16384 x 1..16384 float4 texels");
  ctx->shaderConstants = *c;
  ctx->shaderConstantsSet = true;
  return CR_OK;
}

cr_status cr_set_debug_params(cr_context* ctx, const cr_debug_params* p) {
  if (!ctx || !p) return fail(ctx, CR_ERR_INVALID, "cr_set_debug_params: NULL argument");
  if (p->coarse_threads != 0 && p->coarse_threads != 256 && p->coarse_threads != 1024)
    return fail(ctx, CR_ERR_INVALID, "cr_set_debug_params: coarse_threads must be 0, 256 or 1024");
  if (p->profiling_mode != CR_PROFILING_DEFAULT && p->profiling_mode != CR_PROFILING_COUNTERS)
    return fail(ctx, CR_ERR_INVALID, "cr_set_debug_params: unknown profiling mode %d", p->profiling_mode);
  ctx->dbg = *p;
  return CR_OK;
}

}  // extern "C"
