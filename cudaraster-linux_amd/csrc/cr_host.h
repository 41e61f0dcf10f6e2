// Host-side pieces shared by the C ABI implementation (cr_api.hip) and the
// batched-draw runtime (cr_batch.hip): surface and context objects, device
// buffers, error reporting, the reference's host quantisers and the draw
// geometry.  Internal to libcr.so.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "cr/cr.h"
#include "cr_device.h"

struct cr_surface {
  int32_t width = 0, height = 0, format = 0, samples = 1, samplesLog2 = 0;
  int32_t roundedW = 0, roundedH = 0, textureW = 0;
  uint32_t* dev = nullptr;
  size_t bytes = 0;
  int device = 0;
};

namespace crhost {

// A device allocation that only grows.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t need) {
    if (need <= bytes && ptr) return hipSuccess;
    if (ptr) {
      hipError_t e = hipFree(ptr);
      if (e != hipSuccess) return e;
      ptr = nullptr;
      bytes = 0;
    }
    size_t alloc = std::max<size_t>(need, 256);
    hipError_t e = hipMalloc(&ptr, alloc);
    if (e == hipSuccess) bytes = alloc;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

struct Geometry {
  int32_t vw, vh, wp, hp, wt, ht, nt, bl, wb, hb, nb;
};

inline Geometry geometryOf(const cr_surface* s) {
  Geometry g;
  g.vw = s->width;
  g.vh = s->height;
  g.wp = s->roundedW;
  g.hp = s->roundedH;
  g.wt = g.wp >> crhip::kTileLog2;
  g.ht = g.hp >> crhip::kTileLog2;
  g.nt = g.wt * g.ht;
  // The smallest bins that keep the count within kMaxBins: small surfaces
  // get small bins, so that coarse work spreads over as many workgroups as a
  // large surface's (the image does not depend on the bin size).
  for (g.bl = 0;; g.bl++) {
    g.wb = (g.wt + (1 << g.bl) - 1) >> g.bl;
    g.hb = (g.ht + (1 << g.bl) - 1) >> g.bl;
    g.nb = g.wb * g.hb;
    if (g.nb <= crhip::kMaxBins || g.bl == crhip::kBinLog2) break;
  }
  return g;
}

// Slack added when growing, as the reference (CudaRaster.cpp:239-241).
constexpr int64_t kSubtrisSlack = 4096;
constexpr int64_t kBinSlack = 256 * 512;
constexpr int64_t kTileSlack = 4096 * 32;
constexpr int64_t kMaxEntries = (int64_t)INT32_MAX - 1;
constexpr int32_t kMaxTris = 1 << 28;
// Subtriangle slots: the fine stage packs slot << 4 | centroid code into one
// 32-bit word per sample.
constexpr int64_t kMaxSubtris = (int64_t)1 << 28;

// First-draw capacities of a draw of nt triangles over geometry g.
inline void initialCapacities(int64_t nt, const Geometry& g, int64_t& subtris, int64_t& bin, int64_t& tile) {
  subtris = std::max(subtris, std::min(nt + kSubtrisSlack, kMaxSubtris));
  bin = std::max(bin, nt + (int64_t)g.nb * 64 + kBinSlack);
  // Tile lists: two entries per triangle, plus the unused room of each
  // list's last 32-entry segment.
  tile = std::max(tile, 2 * nt + (int64_t)g.nt * crhip::kTileSeg + kTileSlack);
}

// Records `fmt` as the context's (and the thread's) last error; returns st.
cr_status fail(cr_context* ctx, cr_status st, const char* fmt, ...);

// Vec4f::toABGR of the clear colour and the encoded clear depth
// (CudaRaster.cpp:174-179, base/Math.cpp:41-48, cuda/Util.hpp:290-298).
void clearValues(const float rgba[4], float depth, uint32_t* color, uint32_t* encodedDepth);

// setPixelPipe's checks of a pipe descriptor (CudaRaster.cpp:183-216).
cr_status validatePipe(cr_context* ctx, const cr_pipe_desc* p);

// Where one draw's intermediates live and how many entries they hold.
struct DrawBuffers {
  int64_t maxSubtris = 0, maxBinEntries = 0, maxTileEntries = 0;
  uint8_t* triSubtris = nullptr;
  crhip::Header* triHeader = nullptr;
  crhip::TriData* triData = nullptr;
  uint32_t* subtriTri = nullptr;
  uint32_t* chunkBase = nullptr;
  uint32_t* binCnt = nullptr;
  uint32_t* binOff = nullptr;
  uint4* binList = nullptr;
  uint32_t* tileStart = nullptr;
  uint32_t* tileCount = nullptr;
  uint32_t* tileList = nullptr;
  uint32_t* tileSegNext = nullptr;    // maxTileEntries / 32 segments: next and next-but-one links
  uint32_t* tileFrags = nullptr;      // counting draws only
  uint32_t* chunkSubtris = nullptr;   // counting draws only
  unsigned long long* prof = nullptr; // profiling draws only: 2 * kProfCount counters
  crhip::Counters* counters = nullptr;
};

// Whether the context's next draw runs the counting kernel variants, and
// whether it records the profile counters.
bool countingDraw(const cr_context* ctx);
bool profilingDraw(const cr_context* ctx);

// drawTriangles' state checks (CudaRaster.cpp:243-260).
cr_status validateDraw(cr_context* ctx);
// The context's own (growable) intermediate buffers.
DrawBuffers contextBuffers(cr_context* ctx);
// Kernel parameters of the context's current state over buffers B.
crhip::Params makeParams(cr_context* ctx, const Geometry& g, int32_t numChunks, const DrawBuffers& B);
// Reset, setup+bin, coarse and fine launches with the stage events, then the
// counter block's copy into ctx->hostCounters, all on `stream`.
cr_status launchAll(cr_context* ctx, const crhip::Params& P, hipStream_t stream);
// Copies the read-back counters into ctx->last; returns their overflow bits.
uint32_t recordCounters(cr_context* ctx);
// Looks up this device's copy of the no-allocation draw's fixed scratch
// array (once per device; the device must be current).
void scratchInit(int device);

}  // namespace crhost

#define CR_HIP(ctx, call)                                                                      \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return crhost::fail((ctx), CR_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(e_));       \
  } while (0)

struct cr_context {
  std::string err;
  int device = 0;
  cr_surface* color = nullptr;
  cr_surface* depth = nullptr;
  cr_surface* triid = nullptr;
  bool deferredClear = false;
  uint32_t clearColor = 0, clearDepth = 0;
  bool pipeSet = false;
  cr_pipe_desc pipe{};
  const uint8_t* vbuf = nullptr;
  int64_t vofs = 0, numVerts = 0;
  const uint8_t* ibuf = nullptr;
  int64_t iofs = 0;
  int32_t numTris = 0;
  cr_debug_params dbg{};
  cr_shader_constants shaderConstants{};
  bool shaderConstantsSet = false;

  int64_t maxSubtris = 1, maxBinEntries = 1, maxTileEntries = 1;
  crhost::DevBuf triSubtris, triHeader, triData, subtriTri;
  crhost::DevBuf chunkBase, binCnt, binOff, binList;
  crhost::DevBuf tileStart, tileCount, tileList, tileSegNext, tileFrags, chunkSubtris;
  crhost::DevBuf counters, profBuf;
  crhost::DevBuf interpAcc;                  // fixed-point gradient accumulators (deterministic backward)
  crhip::Counters* hostCounters = nullptr;   // pinned
  unsigned long long* hostProf = nullptr;    // pinned, 2 * kProfCount
  cr_profile_counters lastProf{};
  bool lastProfiled = false;                 // the last launch recorded profile counters
  hipEvent_t ev[5] = {};
  bool evRecorded = false;
  hipStream_t pendingStream = nullptr;
  hipEvent_t evPending = nullptr;     // after an asynchronous draw's counter readback
  bool pending = false;
  cr_counters last{};
};
