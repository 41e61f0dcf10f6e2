// drawTriangles without allocating: cr_draw_triangles_noalloc.
//
// For callers that cannot call hipMalloc (code under HIP graph capture,
// serving loops, harnesses that forbid allocation).  The draw's intermediate
// buffers -- subtriangle records, bin queue, tile lists, counters -- are
// carved out of one fixed device array, `scratch`, of CR_NOALLOC_SCRATCH_BYTES
// bytes.  Every GPU the library runs on holds its own copy of the array until
// the process ends.  The host looks up each device's copy once
// (hipGetSymbolAddress at context creation, with that device current) and
// passes that address to the kernels, never the array's host-side name.
//
// All calls on one GPU share that GPU's array and it keeps the last call's
// contents: the counter block at its start is cleared by a kernel at the
// start of every draw, and everything else is written before it is read.
// Calls that overlap on one GPU (several host threads, or draws queued on
// different streams) race on the array.
//
// Capacities: the space left after the per-draw tables is shared between
// subtriangle slots, bin-queue entries and tile-list entries in proportion to
// the first-draw estimates the allocating entry point uses.  A draw that
// overflows one of them is re-run with the space re-divided by the exact
// demand the counters report, as drawTriangles grows and retries
// (CudaRaster.cpp:277-339); a demand beyond the array is an error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "cr/cr.h"
#include "cr_device.h"
#include "cr_host.h"

using namespace crhip;

static constexpr size_t kScratchBytes = CR_NOALLOC_SCRATCH_BYTES;
// The array itself.  In host code `scratch` only names the host-side
// stand-in the runtime registers; its device address comes from
// hipGetSymbolAddress (scratchInit) and is what the kernels receive.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
bool g_scratchLooked[kMaxDevices] = {};
void* g_scratchPtr[kMaxDevices] = {};
hipError_t g_scratchErr[kMaxDevices] = {};

constexpr size_t kAlign = 256;
inline size_t alignUp(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// Bytes of the per-draw tables (everything but the three queues whose
// capacity is chosen).
size_t fixedBytes(int64_t numTris, const crhost::Geometry& g, int32_t numChunks, bool counting, bool profiling) {
  const size_t nc = (size_t)std::max(numChunks, 1);
  size_t b = alignUp(sizeof(Counters));
  b += alignUp((size_t)std::max<int64_t>(numTris, 1));      // triSubtris
  b += alignUp(nc * 4);                                       // chunkBase
  b += 2 * alignUp((size_t)g.nb * nc * 4);                    // binCnt, binOff
  b += 2 * alignUp((size_t)g.nt * 4);                         // tileStart, tileCount
  if (counting) b += alignUp((size_t)g.nt * 4) + alignUp(nc * 4);   // tileFrags, chunkSubtris
  if (profiling) b += alignUp(2 * kProfCount * sizeof(unsigned long long));
  return b;
}

// Bytes of the queues at the given capacities.
size_t queueBytes(int64_t numTris, int64_t subtris, int64_t bin, int64_t tile) {
  return alignUp((size_t)subtris * sizeof(Header)) + alignUp((size_t)subtris * sizeof(TriData)) +
         alignUp((size_t)std::max<int64_t>(subtris - numTris, 1) * 4) + alignUp((size_t)bin * 16) +
         alignUp((size_t)tile * 4) + alignUp((size_t)std::max<int64_t>(tile >> kTileSegLog2, 1) * 8);
}

// Lays the draw out in the array at `base`.
crhost::DrawBuffers layout(uint8_t* base, int64_t numTris, const crhost::Geometry& g, int32_t numChunks,
                           bool counting, bool profiling, int64_t subtris, int64_t bin, int64_t tile) {
  crhost::DrawBuffers B;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    uint8_t* p = base + off;
    off += alignUp(bytes);
    return p;
  };
  const size_t nc = (size_t)std::max(numChunks, 1);
  B.maxSubtris = subtris;
  B.maxBinEntries = bin;
  B.maxTileEntries = tile;
  B.counters = reinterpret_cast<Counters*>(take(sizeof(Counters)));
  B.triSubtris = take((size_t)std::max<int64_t>(numTris, 1));
  B.chunkBase = reinterpret_cast<uint32_t*>(take(nc * 4));
  B.binCnt = reinterpret_cast<uint32_t*>(take((size_t)g.nb * nc * 4));
  B.binOff = reinterpret_cast<uint32_t*>(take((size_t)g.nb * nc * 4));
  B.tileStart = reinterpret_cast<uint32_t*>(take((size_t)g.nt * 4));
  B.tileCount = reinterpret_cast<uint32_t*>(take((size_t)g.nt * 4));
  if (counting) {
    B.tileFrags = reinterpret_cast<uint32_t*>(take((size_t)g.nt * 4));
    B.chunkSubtris = reinterpret_cast<uint32_t*>(take(nc * 4));
  }
  if (profiling) B.prof = reinterpret_cast<unsigned long long*>(take(2 * kProfCount * sizeof(unsigned long long)));
  B.triHeader = reinterpret_cast<Header*>(take((size_t)subtris * sizeof(Header)));
  B.triData = reinterpret_cast<TriData*>(take((size_t)subtris * sizeof(TriData)));
  B.subtriTri = reinterpret_cast<uint32_t*>(take((size_t)std::max<int64_t>(subtris - numTris, 1) * 4));
  B.binList = reinterpret_cast<uint4*>(take((size_t)bin * 16));
  B.tileList = reinterpret_cast<uint32_t*>(take((size_t)tile * 4));
  B.tileSegNext = reinterpret_cast<uint32_t*>(take((size_t)std::max<int64_t>(tile >> kTileSegLog2, 1) * 8));
  return B;
}

// Scales the capacities (each at least its demand) up so that the queues
// fill `room` bytes; false when even the demands do not fit.
bool fitCapacities(int64_t numTris, size_t room, int64_t& subtris, int64_t& bin, int64_t& tile) {
  if (queueBytes(numTris, subtris, bin, tile) > room) return false;
  const double f = (double)room / (double)queueBytes(numTris, subtris, bin, tile);
  int64_t s = std::min<int64_t>((int64_t)((double)subtris * f), crhost::kMaxSubtris);
  int64_t b = std::min<int64_t>((int64_t)((double)bin * f), crhost::kMaxEntries);
  int64_t t = std::min<int64_t>((int64_t)((double)tile * f), crhost::kMaxEntries);
  // Alignment padding can push the scaled layout a little past the room:
  // step back until it fits (never below the demands, which fit).
  for (int i = 0; i < 64 && queueBytes(numTris, s, b, t) > room; i++) {
    s = std::max(subtris, s - (s >> 6) - 1);
    b = std::max(bin, b - (b >> 6) - 1);
    t = std::max(tile, t - (t >> 6) - 1);
  }
  if (queueBytes(numTris, s, b, t) > room) return false;
  subtris = s;
  bin = b;
  tile = t;
  return true;
}

}  // namespace

namespace crhost {

void scratchInit(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (g_scratchLooked[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratchErr[device] = e;
  g_scratchPtr[device] = (e == hipSuccess) ? p : nullptr;
  g_scratchLooked[device] = true;
}

}  // namespace crhost

using namespace crhost;

extern "C" cr_status cr_draw_triangles_noalloc(cr_context* ctx, void* stream) {
  if (!ctx) return fail(nullptr, CR_ERR_INVALID, "cr_draw_triangles_noalloc: context is NULL");
  cr_status st = validateDraw(ctx);
  if (st != CR_OK) return st;
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return fail(ctx, CR_ERR_UNSUPPORTED, "cr_draw_triangles_noalloc: device %d beyond the %d supported", ctx->device,
                kMaxDevices);
  void* base = nullptr;
  hipError_t lookErr = hipSuccess;
  {
    std::lock_guard<std::mutex> lock(g_scratchMu);
    base = g_scratchPtr[ctx->device];
    lookErr = g_scratchLooked[ctx->device] ? g_scratchErr[ctx->device] : hipErrorNotInitialized;
  }
  if (!base)
    return fail(ctx, CR_ERR_DEVICE, "cr_draw_triangles_noalloc: scratch array of device %d unavailable (%s)",
                ctx->device, hipGetErrorString(lookErr));
  hipStream_t s = static_cast<hipStream_t>(stream);
  CR_HIP(ctx, hipSetDevice(ctx->device));
  if (ctx->pending) {
    // An earlier asynchronous draw of this context still owns the context's
    // pinned counter block: wait for it (its overflow, if any, is its own).
    CR_HIP(ctx, hipEventSynchronize(ctx->evPending));
    ctx->pending = false;
    (void)recordCounters(ctx);
  }

  const Geometry g = geometryOf(ctx->color);
  const int64_t nt = ctx->numTris;
  const int32_t numChunks = (int32_t)((nt + kSetupChunkTris - 1) / kSetupChunkTris);
  const bool counting = countingDraw(ctx), profiling = profilingDraw(ctx);
  const size_t fixed = fixedBytes(nt, g, numChunks, counting, profiling);
  // Demands: the least each queue must hold (first the allocating entry
  // point's first-draw estimates, later the exact counts of an overflowed
  // draw); capacities: the demands scaled up to fill the array.
  int64_t dSub = 0, dBin = 0, dTile = 0;
  initialCapacities(nt, g, dSub, dBin, dTile);
  // Every input triangle owns slot t, so numTris slots (plus one queue entry
  // each) are the least a draw can run with.
  const int64_t needSub = std::max<int64_t>(nt, 1);
  if (fixed + queueBytes(nt, needSub, 1, 1) > kScratchBytes)
    return fail(ctx, CR_ERR_INVALID,
                "cr_draw_triangles_noalloc: %lld triangles at %dx%d need more than the %zu-byte scratch array",
                (long long)nt, g.vw, g.vh, kScratchBytes);
  const size_t room = kScratchBytes - fixed;
  if (queueBytes(nt, dSub, dBin, dTile) > room) {
    // Shrink the estimates to the room; the retry below raises a queue to
    // its exact demand if the draw needs more of it.
    const double f = (double)(room - queueBytes(nt, needSub, 1, 1)) / (double)queueBytes(nt, dSub, dBin, dTile);
    dSub = std::max<int64_t>(needSub, (int64_t)((double)dSub * f));
    dBin = std::max<int64_t>(1, (int64_t)((double)dBin * f));
    dTile = std::max<int64_t>(1, (int64_t)((double)dTile * f));
  }

  ctx->last.retries = 0;
  for (int attempt = 0;; attempt++) {
    int64_t subtris = dSub, bin = dBin, tile = dTile;
    if (dSub > kMaxSubtris || attempt > 8 || !fitCapacities(nt, room, subtris, bin, tile))
      return fail(ctx, CR_ERR_OVERFLOW,
                  "cr_draw_triangles_noalloc: the draw needs %lld subtriangle slots, %lld bin and %lld tile "
                  "entries, more than the %zu-byte scratch array holds",
                  (long long)dSub, (long long)dBin, (long long)dTile, kScratchBytes);
    const DrawBuffers B =
        layout(static_cast<uint8_t*>(base), nt, g, numChunks, counting, profiling, subtris, bin, tile);
    const Params P = makeParams(ctx, g, numChunks, B);
    st = launchAll(ctx, P, s);
    if (st != CR_OK) return st;
    CR_HIP(ctx, hipStreamSynchronize(s));
    const uint32_t ov = recordCounters(ctx);
    if (ov == 0) break;
    const Counters& c = *ctx->hostCounters;
    // Exact demands of the queues that overflowed; a queue that a later stage
    // never reached (coarse skips after a setup overflow) keeps its share.
    if (ov & 1u) dSub = std::max<int64_t>(dSub, (int64_t)c.numSubtris);
    if (ov & 2u) dBin = std::max<int64_t>(dBin, (int64_t)c.binEntries);
    if (ov & 4u) dTile = std::max<int64_t>(dTile, (int64_t)c.tileEntries);
    ctx->last.retries++;
  }
  ctx->deferredClear = false;
  return CR_OK;
}
