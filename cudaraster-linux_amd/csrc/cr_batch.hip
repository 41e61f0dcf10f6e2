// Batched multi-view draws: many independent views, one launch per stage.
//
// The reference draws one view per CudaRaster::drawTriangles() call, each
// with its own four launches and a blocking read of the overflow counters
// (CudaRaster.cpp:237-342).  A batch lays the intermediate buffers of all its
// views out in shared allocations (per-view regions, per-view queue depths),
// keeps a device table of per-view Params, and runs each stage as one grid
// over the work of every view: setup+bin over all triangle chunks, coarse
// over all bins, fine over all tile groups of each pipe variant (the
// workgroup -> view lookup is a binary search over a prefix table).  The
// launch sequence is captured once in a hipGraph and replayed.
//
// Per view the kernels are the plain draw's, with the same Params, so every
// image is bit-identical to a plain draw of that view.  Overflow is per view:
// views whose counters report it are grown and re-drawn alone; the others
// are complete (the fine stage skips a view whose counters overflowed, so
// its surfaces were not touched).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <set>
#include <vector>

#include "cr/cr.h"
#include "cr_device.h"
#include "cr_host.h"

using namespace crhip;
using namespace crhost;

namespace {

struct ViewState {
  cr_view desc;
  Geometry g;
  int32_t numChunks = 0;
  uint32_t clearColor = 0, clearDepth = 0;
  int64_t maxSubtris = 1, maxBin = 1, maxTile = 1;
  // Element offsets of the view's regions in the shared buffers.
  size_t oTri = 0, oSlot = 0, oClip = 0, oChunk = 0, oBinCC = 0, oBin = 0, oTile = 0, oTileList = 0;
};

// One launch set: the device Params table of a list of views and, per
// stage, the order in which the stage lists them with its workgroup prefix.
struct Tables {
  struct Stage {
    int variant = -1;          // fine stages: pipe variant
    int n = 0;                 // listed views
    int blocks = 0;            // workgroups (fine: tile groups)
    int ofs = 0, orderOfs = 0; // prefix (n + 1 entries) and order (n entries) in `starts`
  };
  int n = 0;
  Stage setup, coarse;
  std::vector<Stage> fine;
  DevBuf params, starts;
};

size_t align64(size_t x) { return (x + 63) & ~(size_t)63; }

constexpr int kMaxGroups = 4;   // concurrent view groups (streams / graph branches)

}  // namespace

struct cr_batch {
  cr_context* ctx = nullptr;
  int device = 0;
  std::vector<ViewState> views;
  DevBuf triSubtris, triHeader, triData, subtriTri, chunkBase, binCnt, binOff, binList;
  DevBuf tileStart, tileCount, tileList, tileSegNext, counters, saved;
  Counters* hostCounters = nullptr;   // pinned, one block per view
  Tables groups[kMaxGroups];          // views split into cost-balanced groups, one launch chain each
  int numGroups = 1;
  Tables subset;                      // views re-drawn after an overflow
  hipStream_t side[kMaxGroups] = {};  // groups 1.. run on these (group 0 on the draw's stream)
  hipEvent_t evFork = nullptr, evJoin[kMaxGroups] = {};
  int tilesPerGroup = 4;
  int banded = 0;                     // fine tile groups in XCD bands (1) or round-robin (0)
  int coarseThreads = 256;            // coarse workgroup size
  bool haveCounts = false;            // hostCounters hold every view's demand from a complete draw
  bool laidOut = false, useGraph = true, graphValid = false, tuned = false;
  hipStream_t capStream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float lastMs = 0.0f, tunedMs = 0.0f;
  int64_t retries = 0;
  size_t reserved = 0;
};

namespace {

cr_status layout(cr_batch* b) {
  cr_context* ctx = b->ctx;
  size_t nTri = 0, nSlot = 0, nClip = 0, nChunk = 0, nBinCC = 0, nBin = 0, nTile = 0, nTileList = 0;
  for (ViewState& v : b->views) {
    const int64_t nt = v.desc.num_tris;
    v.oTri = nTri;
    nTri += align64((size_t)std::max<int64_t>(nt, 1));
    v.oSlot = nSlot;
    nSlot += align64((size_t)v.maxSubtris);
    v.oClip = nClip;
    nClip += align64((size_t)std::max<int64_t>(v.maxSubtris - nt, 1));
    v.oChunk = nChunk;
    nChunk += align64((size_t)std::max(v.numChunks, 1));
    v.oBinCC = nBinCC;
    nBinCC += align64((size_t)v.g.nb * std::max(v.numChunks, 1));
    v.oBin = nBin;
    nBin += align64((size_t)v.maxBin);
    v.oTile = nTile;
    nTile += align64((size_t)v.g.nt);
    v.oTileList = nTileList;
    nTileList += align64((size_t)v.maxTile);
  }
  CR_HIP(ctx, b->triSubtris.ensure(nTri));
  CR_HIP(ctx, b->triHeader.ensure(nSlot * sizeof(Header)));
  CR_HIP(ctx, b->triData.ensure(nSlot * sizeof(TriData)));
  CR_HIP(ctx, b->subtriTri.ensure(nClip * 4));
  CR_HIP(ctx, b->chunkBase.ensure(nChunk * 4));
  CR_HIP(ctx, b->binCnt.ensure(nBinCC * 4));
  CR_HIP(ctx, b->binOff.ensure(nBinCC * 4));
  CR_HIP(ctx, b->binList.ensure(nBin * 16));
  CR_HIP(ctx, b->tileStart.ensure(nTile * 4));
  CR_HIP(ctx, b->tileCount.ensure(nTile * 4));
  CR_HIP(ctx, b->tileList.ensure(nTileList * 4));
  CR_HIP(ctx, b->tileSegNext.ensure(std::max<size_t>(nTileList >> kTileSegLog2, 1) * 8));
  CR_HIP(ctx, b->counters.ensure(b->views.size() * sizeof(Counters)));
  b->reserved = b->triSubtris.bytes + b->triHeader.bytes + b->triData.bytes + b->subtriTri.bytes +
                b->chunkBase.bytes + b->binCnt.bytes + b->binOff.bytes + b->binList.bytes + b->tileStart.bytes +
                b->tileCount.bytes + b->tileList.bytes + b->tileSegNext.bytes + b->counters.bytes;
  b->graphValid = false;
  b->laidOut = true;
  return CR_OK;
}

Params viewParams(const cr_batch* b, int i) {
  const ViewState& v = b->views[i];
  const cr_view& d = v.desc;
  Params P;
  memset(&P, 0, sizeof(P));
  P.numTris = d.num_tris;
  P.numVertices = (int32_t)std::min<int64_t>(d.num_vertices, INT32_MAX);
  P.vertexStride = d.pipe.vertex_struct_size;
  P.vertices = static_cast<const uint8_t*>(d.vertices);
  P.indices = static_cast<const int32_t*>(d.indices);
  P.viewportWidth = v.g.vw;
  P.viewportHeight = v.g.vh;
  P.widthPixels = v.g.wp;
  P.heightPixels = v.g.hp;
  P.widthTiles = v.g.wt;
  P.heightTiles = v.g.ht;
  P.numTiles = v.g.nt;
  P.binLog2 = v.g.bl;
  P.widthBins = v.g.wb;
  P.heightBins = v.g.hb;
  P.numBins = v.g.nb;
  P.samplesLog2 = d.pipe.samples_log2;
  P.flags = d.pipe.render_mode_flags & (kFlagDepth | kFlagLerp);
  P.blend = d.pipe.blend_shader;
  P.shader = d.pipe.fragment_shader;
  P.constColor = d.pipe.const_color;
  P.deferredClear = d.deferred_clear ? 1 : 0;
  P.clearColor = v.clearColor;
  P.clearDepth = v.clearDepth;
  P.hiz = 1;
  P.maxSubtris = (int32_t)v.maxSubtris;
  P.triSubtris = b->triSubtris.as<uint8_t>() + v.oTri;
  P.triHeader = b->triHeader.as<Header>() + v.oSlot;
  P.triData = b->triData.as<TriData>() + v.oSlot;
  P.subtriTri = b->subtriTri.as<uint32_t>() + v.oClip;
  P.numChunks = v.numChunks;
  P.chunkBase = b->chunkBase.as<uint32_t>() + v.oChunk;
  P.binCnt = b->binCnt.as<uint32_t>() + v.oBinCC;
  P.binOff = b->binOff.as<uint32_t>() + v.oBinCC;
  P.maxBinEntries = (int32_t)v.maxBin;
  P.binList = b->binList.as<uint4>() + v.oBin;
  P.tileStart = b->tileStart.as<uint32_t>() + v.oTile;
  P.tileCount = b->tileCount.as<uint32_t>() + v.oTile;
  P.maxTileEntries = (int32_t)v.maxTile;
  P.tileList = b->tileList.as<uint32_t>() + v.oTileList;
  P.tileSegNext = b->tileSegNext.as<uint32_t>() + 2 * (v.oTileList >> kTileSegLog2);   // oTileList: a multiple of 64
  P.color = d.color->dev;
  P.depth = d.depth->dev;
  P.triid = d.tri_id ? d.tri_id->dev : nullptr;
  P.counters = b->counters.as<Counters>() + i;
  return P;
}

// Appends one stage: the listed Params indices `list` (in launch order) with
// `blocks[k]` workgroups each.
void appendStage(std::vector<int32_t>& st, Tables::Stage& S, const std::vector<int>& list,
                 const std::vector<int64_t>& blocks, int64_t& most) {
  S.n = (int)list.size();
  S.ofs = (int)st.size();
  int64_t acc = 0;
  for (int k : list) {
    st.push_back((int32_t)acc);
    acc += blocks[k];
  }
  st.push_back((int32_t)acc);
  S.orderOfs = (int)st.size();
  for (int k : list) st.push_back(k);
  S.blocks = (int)std::min<int64_t>(acc, INT32_MAX);
  most = std::max(most, acc);
}

// Params indices sorted by decreasing cost per workgroup, so that the
// longest workgroups start first (ties keep submission order).
std::vector<int> longestFirst(const std::vector<double>& cost, const std::vector<int>& list) {
  std::vector<int> out = list;
  std::stable_sort(out.begin(), out.end(), [&](int a, int c) { return cost[a] > cost[c]; });
  return out;
}

cr_status buildTables(cr_batch* b, Tables& T, const std::vector<int>& idx) {
  cr_context* ctx = b->ctx;
  const int n = (int)idx.size();
  const int W = b->tilesPerGroup;
  std::vector<Params> ps(n);
  for (int k = 0; k < n; k++) ps[k] = viewParams(b, idx[k]);
  // Cost per workgroup: measured entries per bin / per tile once a draw has
  // counted them, else triangles per bin / per tile.
  std::vector<double> binCost(n), tileCost(n);
  std::vector<int64_t> chunks(n), bins(n), groups(n);
  for (int k = 0; k < n; k++) {
    const Params& P = ps[k];
    const Counters& c = b->hostCounters[idx[k]];
    const bool measured = b->haveCounts && !c.overflow;
    binCost[k] = (measured ? (double)c.binEntries : (double)P.numTris) / std::max(P.numBins, 1);
    tileCost[k] = (measured ? (double)c.tileEntries : 2.0 * P.numTris) * (1 << P.samplesLog2) /
                  std::max(P.numTiles, 1);
    chunks[k] = P.numTris > 0 ? P.numChunks : 0;
    bins[k] = P.numBins;
    groups[k] = ((P.numTiles + W - 1) / W + kFineRounds - 1) / kFineRounds;
  }
  std::vector<int> all(n);
  for (int k = 0; k < n; k++) all[k] = k;
  std::vector<int32_t> st;
  int64_t most = 0;
  appendStage(st, T.setup, all, chunks, most);
  appendStage(st, T.coarse, longestFirst(binCost, all), bins, most);
  T.fine.clear();
  for (int var = 0; var < kFineVariants; var++) {
    std::vector<int> list;
    for (int k = 0; k < n; k++)
      if (fineVariant(ps[k]) == var) list.push_back(k);
    if (list.empty()) continue;
    Tables::Stage S;
    S.variant = var;
    appendStage(st, S, longestFirst(tileCost, list), groups, most);
    if (S.blocks > 0) T.fine.push_back(S);
  }
  if (most >= (int64_t)1 << 30) return fail(ctx, CR_ERR_OVERFLOW, "batch: more than 2^30 workgroups in one stage");
  CR_HIP(ctx, T.params.ensure((size_t)std::max(n, 1) * sizeof(Params)));
  CR_HIP(ctx, T.starts.ensure(st.size() * sizeof(int32_t)));
  if (n > 0) CR_HIP(ctx, hipMemcpy(T.params.ptr, ps.data(), (size_t)n * sizeof(Params), hipMemcpyHostToDevice));
  CR_HIP(ctx, hipMemcpy(T.starts.ptr, st.data(), st.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  T.n = n;
  return CR_OK;
}

std::vector<int> allViews(const cr_batch* b) {
  std::vector<int> idx(b->views.size());
  for (size_t i = 0; i < idx.size(); i++) idx[i] = (int)i;
  return idx;
}

// Splits the views into numGroups groups of similar cost (longest first to
// the least loaded group) and builds each group's tables.
cr_status buildAll(cr_batch* b) {
  const int n = (int)b->views.size();
  const int K = std::max(1, std::min(b->numGroups, n));
  std::vector<double> cost(n);
  for (int i = 0; i < n; i++) {
    const Counters& c = b->hostCounters[i];
    const int ns = 1 << b->views[i].desc.pipe.samples_log2;
    cost[i] = (b->haveCounts && !c.overflow) ? (double)c.binEntries + (double)c.tileEntries * ns
                                             : (double)b->views[i].desc.num_tris * (1 + ns);
  }
  std::vector<int> order = allViews(b);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
  std::vector<std::vector<int>> members(K);
  std::vector<double> load(K, 0.0);
  for (int i : order) {
    const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    members[g].push_back(i);
    load[g] += cost[i];
  }
  for (int g = 0; g < K; g++) {
    std::sort(members[g].begin(), members[g].end());
    cr_status st = buildTables(b, b->groups[g], members[g]);
    if (st != CR_OK) return st;
  }
  b->numGroups = K;
  b->graphValid = false;
  return CR_OK;
}

cr_status launchTables(cr_batch* b, const Tables& T, hipStream_t stream) {
  const Params* pv = T.params.as<Params>();
  const int32_t* s = T.starts.as<int32_t>();
  launchResetBatch(pv, T.n, stream);
  launchSetupBatch(pv, s + T.setup.ofs, s + T.setup.orderOfs, T.setup.n, T.setup.blocks, stream);
  launchCoarseBatch(b->coarseThreads, pv, s + T.coarse.ofs, s + T.coarse.orderOfs, T.coarse.n, T.coarse.blocks,
                    stream);
  for (const Tables::Stage& f : T.fine)
    launchFineBatch(f.variant, b->tilesPerGroup, b->banded, pv, s + f.ofs, s + f.orderOfs, f.n, f.blocks, stream);
  CR_HIP(b->ctx, hipGetLastError());
  return CR_OK;
}

// All groups: group 0 on `stream`, the others on side streams forked from
// and joined back into it (parallel branches when captured in a graph).
cr_status launchAllGroups(cr_batch* b, hipStream_t stream) {
  if (b->numGroups <= 1) return launchTables(b, b->groups[0], stream);
  CR_HIP(b->ctx, hipEventRecord(b->evFork, stream));
  for (int g = 1; g < b->numGroups; g++) CR_HIP(b->ctx, hipStreamWaitEvent(b->side[g], b->evFork, 0));
  cr_status st = launchTables(b, b->groups[0], stream);
  for (int g = 1; g < b->numGroups && st == CR_OK; g++) st = launchTables(b, b->groups[g], b->side[g]);
  for (int g = 1; g < b->numGroups; g++) {
    CR_HIP(b->ctx, hipEventRecord(b->evJoin[g], b->side[g]));
    CR_HIP(b->ctx, hipStreamWaitEvent(stream, b->evJoin[g], 0));
  }
  return st;
}

cr_status capture(cr_batch* b) {
  cr_context* ctx = b->ctx;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  CR_HIP(ctx, hipStreamBeginCapture(b->capStream, hipStreamCaptureModeThreadLocal));
  cr_status st = launchAllGroups(b, b->capStream);
  hipError_t e = hipStreamEndCapture(b->capStream, &b->graph);
  if (st != CR_OK) return st;
  CR_HIP(ctx, e);
  CR_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graphValid = true;
  return CR_OK;
}

cr_status readCounters(cr_batch* b, hipStream_t stream) {
  CR_HIP(b->ctx, hipMemcpyAsync(b->hostCounters, b->counters.ptr, b->views.size() * sizeof(Counters),
                                hipMemcpyDeviceToHost, stream));
  CR_HIP(b->ctx, hipStreamSynchronize(stream));
  return CR_OK;
}

cr_status prepare(cr_batch* b) {
  if (b->laidOut) return CR_OK;
  cr_status st = layout(b);
  if (st != CR_OK) return st;
  return buildAll(b);
}

// Grows the capacities of every view whose counters overflowed and re-draws
// those views (only) until none does.
cr_status settleOverflow(cr_batch* b, hipStream_t stream) {
  for (;;) {
    std::vector<int> over;
    for (size_t i = 0; i < b->views.size(); i++) {
      const Counters& c = b->hostCounters[i];
      if (!c.overflow) continue;
      ViewState& v = b->views[i];
      if (c.overflow & 1u) v.maxSubtris = std::max<int64_t>(v.maxSubtris, (int64_t)c.numSubtris + kSubtrisSlack);
      if (c.overflow & 2u) v.maxBin = std::max<int64_t>(v.maxBin, (int64_t)c.binEntries + kBinSlack);
      if (c.overflow & 4u) v.maxTile = std::max<int64_t>(v.maxTile, (int64_t)c.tileEntries + kTileSlack);
      if (v.maxSubtris > kMaxSubtris)
        return fail(b->ctx, CR_ERR_OVERFLOW, "batch view %zu: more than 2^28 subtriangles", i);
      if (v.maxBin > kMaxEntries || v.maxTile > kMaxEntries)
        return fail(b->ctx, CR_ERR_OVERFLOW, "batch view %zu: internal queue capacity exceeded (2^31 entries)", i);
      over.push_back((int)i);
    }
    if (over.empty()) {
      b->haveCounts = true;
      return CR_OK;
    }
    b->retries++;
    cr_status st = layout(b);
    if (st == CR_OK) st = buildAll(b);
    if (st == CR_OK) st = buildTables(b, b->subset, over);
    if (st == CR_OK) st = launchTables(b, b->subset, stream);
    if (st == CR_OK) st = readCounters(b, stream);
    if (st != CR_OK) return st;
  }
}

cr_status drawBatch(cr_batch* b, hipStream_t stream, bool graph) {
  cr_context* ctx = b->ctx;
  CR_HIP(ctx, hipSetDevice(ctx->device));
  cr_status st = prepare(b);
  if (st != CR_OK) return st;
  b->retries = 0;
  if (graph && !b->graphValid) {
    st = capture(b);
    if (st != CR_OK) return st;
  }
  CR_HIP(ctx, hipEventRecord(b->ev0, stream));
  if (graph) {
    CR_HIP(ctx, hipGraphLaunch(b->exec, stream));
  } else {
    st = launchAllGroups(b, stream);
    if (st != CR_OK) return st;
  }
  CR_HIP(ctx, hipEventRecord(b->ev1, stream));
  st = readCounters(b, stream);
  if (st != CR_OK) return st;
  CR_HIP(ctx, hipEventElapsedTime(&b->lastMs, b->ev0, b->ev1));
  return settleOverflow(b, stream);
}

// Device-to-device copies of the surfaces of views drawn without deferred
// clear (the autotuner draws the batch several times).
cr_status saveSurfaces(cr_batch* b, hipStream_t stream, bool restore) {
  size_t need = 0;
  for (const ViewState& v : b->views)
    if (!v.desc.deferred_clear)
      need += v.desc.color->bytes + v.desc.depth->bytes + (v.desc.tri_id ? v.desc.tri_id->bytes : 0);
  if (need == 0) return CR_OK;
  if (!restore) CR_HIP(b->ctx, b->saved.ensure(need));
  uint8_t* p = b->saved.as<uint8_t>();
  for (const ViewState& v : b->views) {
    if (v.desc.deferred_clear) continue;
    for (const cr_surface* s : {v.desc.color, v.desc.depth, v.desc.tri_id}) {
      if (!s) continue;
      if (restore) CR_HIP(b->ctx, hipMemcpyAsync(s->dev, p, s->bytes, hipMemcpyDeviceToDevice, stream));
      else CR_HIP(b->ctx, hipMemcpyAsync(p, s->dev, s->bytes, hipMemcpyDeviceToDevice, stream));
      p += s->bytes;
    }
  }
  return CR_OK;
}

cr_status validateView(cr_context* ctx, const cr_view& v, int i) {
  if (!v.color || !v.depth) return fail(ctx, CR_ERR_INVALID, "batch view %d: colour and depth surfaces are required", i);
  if (v.color->format != CR_FORMAT_RGBA8) return fail(ctx, CR_ERR_INVALID, "batch view %d: unsupported color buffer format", i);
  if (v.depth->format != CR_FORMAT_DEPTH32) return fail(ctx, CR_ERR_INVALID, "batch view %d: unsupported depth buffer format", i);
  if (v.color->width != v.depth->width || v.color->height != v.depth->height)
    return fail(ctx, CR_ERR_INVALID, "batch view %d: mismatch in size between surfaces", i);
  if (v.color->samples != v.depth->samples)
    return fail(ctx, CR_ERR_INVALID, "batch view %d: mismatch in multisampling between surfaces", i);
  if (v.tri_id && (v.tri_id->format != CR_FORMAT_TRIANGLE_ID || v.tri_id->width != v.color->width ||
                   v.tri_id->height != v.color->height || v.tri_id->samples != v.color->samples))
    return fail(ctx, CR_ERR_INVALID, "batch view %d: triangle-id surface does not match the colour surface", i);
  cr_status st = validatePipe(ctx, &v.pipe);
  if (st != CR_OK) return st;
  if (v.pipe.fragment_shader == CR_SHADER_TEX_PHONG)
    return fail(ctx, CR_ERR_UNSUPPORTED, "batch view %d: the texPhong shader draws through cr_draw_triangles", i);
  if (v.pipe.samples_log2 != v.color->samplesLog2)
    return fail(ctx, CR_ERR_INVALID, "batch view %d: mismatch in multisampling between pixel pipe and surface", i);
  if (v.num_tris < 0 || v.num_tris > kMaxTris || v.num_vertices < 0)
    return fail(ctx, CR_ERR_INVALID, "batch view %d: bad triangle or vertex count", i);
  if (v.num_tris > 0 && (!v.vertices || !v.indices))
    return fail(ctx, CR_ERR_INVALID, "batch view %d: vertex and index buffers are required", i);
  if (((uintptr_t)v.vertices & 15) || ((uintptr_t)v.indices & 3))
    return fail(ctx, CR_ERR_INVALID, "batch view %d: misaligned vertex or index pointer", i);
  return CR_OK;
}

}  // namespace

extern "C" {

cr_status cr_batch_create(cr_context* ctx, const cr_view* views, int32_t numViews, cr_batch** out) {
  if (!ctx || !out || (!views && numViews > 0) || numViews < 0)
    return fail(ctx, CR_ERR_INVALID, "cr_batch_create: bad arguments");
  *out = nullptr;
  std::set<const cr_surface*> seen;
  for (int i = 0; i < numViews; i++) {
    cr_status st = validateView(ctx, views[i], i);
    if (st != CR_OK) return st;
    for (const cr_surface* s : {views[i].color, views[i].depth, views[i].tri_id})
      if (s && !seen.insert(s).second)
        return fail(ctx, CR_ERR_INVALID, "batch view %d: a surface appears in more than one view", i);
  }
  cr_batch* b = new (std::nothrow) cr_batch();
  if (!b) return fail(ctx, CR_ERR_NOMEM, "out of host memory");
  b->ctx = ctx;
  b->device = ctx->device;
  b->views.resize(numViews);
  for (int i = 0; i < numViews; i++) {
    ViewState& v = b->views[i];
    v.desc = views[i];
    v.g = geometryOf(v.desc.color);
    v.numChunks = (v.desc.num_tris + kSetupChunkTris - 1) / kSetupChunkTris;
    if (v.desc.deferred_clear) clearValues(v.desc.clear_color, v.desc.clear_depth, &v.clearColor, &v.clearDepth);
    initialCapacities(v.desc.num_tris, v.g, v.maxSubtris, v.maxBin, v.maxTile);
  }
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->hostCounters, std::max<size_t>(numViews, 1) * sizeof(Counters),
                                         hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->capStream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->evFork, hipEventDisableTiming);
  for (int g = 1; g < kMaxGroups && e == hipSuccess; g++) {
    e = hipStreamCreateWithFlags(&b->side[g], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->evJoin[g], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    cr_batch_destroy(b);
    return fail(ctx, CR_ERR_DEVICE, "cr_batch_create: %s", hipGetErrorString(e));
  }
  memset(b->hostCounters, 0, std::max<size_t>(numViews, 1) * sizeof(Counters));
  *out = b;
  return CR_OK;
}

cr_status cr_batch_destroy(cr_batch* b) {
  if (!b) return CR_OK;
  (void)hipSetDevice(b->device);
  (void)hipDeviceSynchronize();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (DevBuf* d : {&b->triSubtris, &b->triHeader, &b->triData, &b->subtriTri, &b->chunkBase, &b->binCnt, &b->binOff,
                    &b->binList, &b->tileStart, &b->tileCount, &b->tileList, &b->tileSegNext, &b->counters,
                    &b->saved,
                    &b->subset.params, &b->subset.starts})
    d->release();
  for (Tables& t : b->groups) {
    t.params.release();
    t.starts.release();
  }
  if (b->hostCounters) (void)hipHostFree(b->hostCounters);
  if (b->capStream) (void)hipStreamDestroy(b->capStream);
  for (int g = 1; g < kMaxGroups; g++) {
    if (b->side[g]) (void)hipStreamDestroy(b->side[g]);
    if (b->evJoin[g]) (void)hipEventDestroy(b->evJoin[g]);
  }
  if (b->evFork) (void)hipEventDestroy(b->evFork);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  delete b;
  return CR_OK;
}

cr_status cr_batch_draw(cr_batch* b, void* stream) {
  if (!b) return fail(nullptr, CR_ERR_INVALID, "cr_batch_draw: batch is NULL");
  if (b->views.empty()) return CR_OK;
  return drawBatch(b, static_cast<hipStream_t>(stream), b->useGraph);
}

cr_status cr_batch_autotune(cr_batch* b, void* stream_) {
  if (!b) return fail(nullptr, CR_ERR_INVALID, "cr_batch_autotune: batch is NULL");
  if (b->views.empty()) return CR_OK;
  cr_context* ctx = b->ctx;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  CR_HIP(ctx, hipSetDevice(ctx->device));
  cr_status st = prepare(b);
  if (st == CR_OK) st = saveSurfaces(b, stream, false);
  // 1. Exact per-view queue demand (the draw grows whatever overflows).
  if (st == CR_OK) st = drawBatch(b, stream, false);
  if (st != CR_OK) return st;
  // 2. Queue depths = demand + 1/16 + a small constant margin.
  for (size_t i = 0; i < b->views.size(); i++) {
    const Counters& c = b->hostCounters[i];
    ViewState& v = b->views[i];
    const int64_t sub = std::max<int64_t>(c.numSubtris, v.desc.num_tris);
    v.maxSubtris = std::min<int64_t>(sub + sub / 16 + 1024, kMaxSubtris);
    v.maxBin = std::min<int64_t>((int64_t)c.binEntries + c.binEntries / 16 + 4096, kMaxEntries);
    v.maxTile = std::min<int64_t>((int64_t)c.tileEntries + c.tileEntries / 16 + 4096, kMaxEntries);
  }
  st = layout(b);
  if (st != CR_OK) return st;
  // 3. Fine-stage workgroup shape and placement: the fastest combination of
  //    the compiled shapes and banded / round-robin tile groups (the stages
  //    list views longest-first by the measured entries per bin and tile).
  int bestW = b->tilesPerGroup, bestBanded = b->banded, bestCoarse = b->coarseThreads;
  float bestMs = 1e30f;
  for (int banded = 0; banded < 2; banded++) {
    for (int trial = 0; trial < 3 + 1; trial++) {
      // three fine shapes with the current coarse size, then the other
      // coarse size with the best fine shape so far
      const int W = trial < 3 ? kFineBatchShapes[trial] : bestW;
      b->tilesPerGroup = W;
      b->banded = banded;
      b->coarseThreads = trial < 3 ? bestCoarse : (bestCoarse == 256 ? 1024 : 256);
      st = buildAll(b);
      if (st != CR_OK) return st;
      float ms[3];
      for (int rep = 0; rep < 4; rep++) {
        st = saveSurfaces(b, stream, true);
        if (st == CR_OK) st = drawBatch(b, stream, false);
        if (st != CR_OK) return st;
        if (rep > 0) ms[rep - 1] = b->lastMs;
      }
      std::sort(ms, ms + 3);
      if (ms[1] < bestMs) {
        bestMs = ms[1];
        bestW = W;
        bestBanded = banded;
        bestCoarse = b->coarseThreads;
      }
      b->coarseThreads = bestCoarse;
    }
  }
  b->tilesPerGroup = bestW;
  b->banded = bestBanded;
  b->coarseThreads = bestCoarse;
  // 4. Concurrent view groups: the views split into K cost-balanced groups
  //    whose launch chains run side by side (K = 1, 2 or 4).
  int bestK = 1;
  for (int K : {2, 4}) {
    if (K > (int)b->views.size()) break;
    b->numGroups = K;
    st = buildAll(b);
    if (st != CR_OK) return st;
    float ms[3];
    for (int rep = 0; rep < 4; rep++) {
      st = saveSurfaces(b, stream, true);
      if (st == CR_OK) st = drawBatch(b, stream, false);
      if (st != CR_OK) return st;
      if (rep > 0) ms[rep - 1] = b->lastMs;
    }
    std::sort(ms, ms + 3);
    if (ms[1] < bestMs) {
      bestMs = ms[1];
      bestK = K;
    }
  }
  b->numGroups = bestK;
  st = buildAll(b);
  if (st == CR_OK) st = saveSurfaces(b, stream, true);
  if (st != CR_OK) return st;
  CR_HIP(ctx, hipStreamSynchronize(stream));
  b->saved.release();
  b->graphValid = false;
  b->tuned = true;
  b->tunedMs = bestMs;
  return CR_OK;
}

cr_status cr_batch_reserve(cr_batch* b, int32_t view, int64_t maxSubtris, int64_t maxBin, int64_t maxTile) {
  if (!b) return fail(nullptr, CR_ERR_INVALID, "cr_batch_reserve: batch is NULL");
  if (view < 0 || view >= (int32_t)b->views.size()) return fail(b->ctx, CR_ERR_INVALID, "cr_batch_reserve: bad view");
  if (maxSubtris < 1 || maxBin < 1 || maxTile < 1 || maxSubtris > kMaxEntries || maxBin > kMaxEntries ||
      maxTile > kMaxEntries)
    return fail(b->ctx, CR_ERR_INVALID, "cr_batch_reserve: capacities must be in 1 .. 2^31-2");
  ViewState& v = b->views[view];
  // Slot t holds triangle t: at least one slot per input triangle.
  v.maxSubtris = std::max<int64_t>(maxSubtris, v.desc.num_tris);
  v.maxBin = maxBin;
  v.maxTile = maxTile;
  b->laidOut = false;
  b->graphValid = false;
  return CR_OK;
}

cr_status cr_batch_configure(cr_batch* b, const cr_batch_config* c) {
  if (!b || !c) return fail(nullptr, CR_ERR_INVALID, "cr_batch_configure: NULL argument");
  if ((c->fine_tiles_per_group != 2 && c->fine_tiles_per_group != 4 && c->fine_tiles_per_group != 8) ||
      (c->fine_banded != 0 && c->fine_banded != 1) || (c->coarse_threads != 256 && c->coarse_threads != 1024) ||
      c->view_groups < 1 || c->view_groups > kMaxGroups)
    return fail(b->ctx, CR_ERR_INVALID, "cr_batch_configure: fine tiles 2/4/8, banded 0/1, coarse 256/1024, groups 1..4");
  b->tilesPerGroup = c->fine_tiles_per_group;
  b->banded = c->fine_banded;
  b->coarseThreads = c->coarse_threads;
  b->numGroups = c->view_groups;
  b->laidOut = false;                 // tables are rebuilt on the next draw
  b->graphValid = false;
  return CR_OK;
}

cr_status cr_batch_use_graph(cr_batch* b, int32_t enable) {
  if (!b) return fail(nullptr, CR_ERR_INVALID, "cr_batch_use_graph: batch is NULL");
  b->useGraph = enable != 0;
  return CR_OK;
}

cr_status cr_batch_get_info(const cr_batch* b, cr_batch_info* out) {
  if (!b || !out) return fail(nullptr, CR_ERR_INVALID, "cr_batch_get_info: NULL argument");
  memset(out, 0, sizeof(*out));
  out->num_views = (int32_t)b->views.size();
  out->fine_tiles_per_group = b->tilesPerGroup;
  out->fine_banded = b->banded;
  out->coarse_threads = b->coarseThreads;
  out->view_groups = b->numGroups;
  out->uses_graph = b->useGraph ? 1 : 0;
  out->tuned = b->tuned ? 1 : 0;
  out->retries = b->retries;
  out->reserved_bytes = (int64_t)b->reserved;
  out->last_gpu_ms = b->lastMs;
  out->tuned_gpu_ms = b->tunedMs;
  return CR_OK;
}

cr_status cr_batch_get_view_counters(const cr_batch* b, int32_t view, cr_counters* out) {
  if (!b || !out) return fail(nullptr, CR_ERR_INVALID, "cr_batch_get_view_counters: NULL argument");
  if (view < 0 || view >= (int32_t)b->views.size())
    return fail(b->ctx, CR_ERR_INVALID, "cr_batch_get_view_counters: bad view");
  const Counters& c = b->hostCounters[view];
  memset(out, 0, sizeof(*out));
  out->num_tris = b->views[view].desc.num_tris;
  out->num_subtris = c.numSubtris;
  out->bin_entries = c.binEntries;
  out->tile_entries = c.tileEntries;
  out->tile_list_entries = c.tileListEntries;
  out->retries = b->retries;
  return CR_OK;
}

}  // extern "C"
