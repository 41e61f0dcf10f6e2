// Stage 3: coarse raster -- per-bin, order-preserving tile lists.
//
// The reference merges 16 bin streams by lowest triangle index and emits
// 32-entry linked tile segments (cudaraster/cuda/CoarseRaster.inl:28-809).
// Here one block of waves owns a bin and walks the bin's entries once, in
// submission order, in rounds of one entry per lane:
//
//   test    each entry's tiles: exact-sign edge tests of the tiles' closed
//           squares inside the entry's box (bin-local);
//   rank    per-wave 64-bit emit masks per tile, a prefix over the waves:
//           every entry's position in its tile's list;
//   extend  per tile, the 32-entry segments the round needs beyond the
//           tile's last, partly filled segment: one block scan over the
//           tiles and one atomic reserve them for the whole bin, and they
//           are chained to the tile's list;
//   write   every entry to its segment and place in it.
//
// Lists are chains of segments of the fine stage's chunk size, filled in
// order, so no count pass (and no second read of the bin queue) is needed
// to size them; a list wastes at most its last segment's unused room.
//
// The bin's entries are the concatenation, in chunk order, of the segments
// the fused setup/bin stage wrote per chunk of triangles; a segment table is
// scanned into LDS and every thread finds its entry by a binary search of
// the segment starts, one round ahead.  Large boxes are shared out over the
// workgroup's waves through an LDS queue.  Binning never changes the image;
// the fine stage is exact.
#include "cr_device.h"

namespace crhip {

// Workgroup sizes the coarse stage is compiled for (kCoarseShapes in
// cr_device.h): 1024 lanes suit bins with thousands of entries; 256 lanes
// need a quarter of the LDS, so several workgroups share a CU and a block
// barrier waits for 4 waves instead of 16.

// Set bits of `mask` below this lane (v_mbcnt).
__device__ __forceinline__ uint32_t rankBelow(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <int NT>
__device__ __forceinline__ uint32_t blockScanC(uint32_t v, uint32_t* sums, uint32_t& total) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sums[wave] = x;
  __syncthreads();
  if (threadIdx.x < 64) {
    const uint32_t w = (threadIdx.x < NW) ? sums[threadIdx.x] : 0u;
    uint32_t sc = w;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(sc, o, 64);
      if ((int)threadIdx.x >= o) sc += y;
    }
    if (threadIdx.x < NW) sums[threadIdx.x] = sc - w;
    if (threadIdx.x == NW - 1) sums[NW] = sc;
  }
  __syncthreads();
  total = sums[NW];
  const uint32_t r = sums[wave] + x - v;
  __syncthreads();
  return r;
}

// Clears the counter block (and a profiling draw's profile counters) with
// plain stores: graph-safe, no memset node.
__global__ void __launch_bounds__(64) resetCountersKernel(Counters* c, uint32_t numTris, unsigned long long* prof) {
  uint32_t* w = reinterpret_cast<uint32_t*>(c);
  for (int i = threadIdx.x; i < (int)(sizeof(Counters) / 4); i += 64) w[i] = 0;
  if (prof != nullptr)
    for (int i = threadIdx.x; i < 2 * kProfCount; i += 64) prof[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) c->numSubtris = numTris;
}

// Per-lane tile test data of one triangle inside bin (bx, by).
struct TileTest {
  int32_t x0, y0, x1, y1;   // bin-local tile box (inclusive); empty when x0 > x1
  int32_t G[3];             // clamped edge thresholds at the box origin
  int32_t sx[3], sy[3];     // threshold change per tile step in x / y
  __device__ __forceinline__ bool hit(int x, int y) const {
    const int32_t rx = x - x0, ry = y - y0;
    return (rx * sx[0] + ry * sy[0] <= G[0]) & (rx * sx[1] + ry * sy[1] <= G[1]) &
           (rx * sx[2] + ry * sy[2] <= G[2]);
  }
};

// Tile (x, y) of the bin holds a covered sample only if its closed square
// reaches the inner side of all three edges: min over the square of
// p x d <= a x d.  With the corner that minimises p x d, p x d at tile
// (x, y) = p x d at the box origin + rel, |rel| < 2^29 (<= 15 tile steps of
// 128 subpixels times |d| < 2^17), so the threshold c - p0 x d clamped to
// +-2^30 decides every tile exactly in 32 bits.
__device__ __forceinline__ TileTest tileTestOf(const Params& P, uint4 h, int bx, int by) {
  TileTest r;
  const int32_t ox = P.viewportWidth << (kSubpixelLog2 - 1), oy = P.viewportHeight << (kSubpixelLog2 - 1);
  int32_t ax[3], ay[3];
  ax[0] = (int32_t)(int16_t)(h.x & 0xFFFFu) + ox; ay[0] = (int32_t)(int16_t)(h.x >> 16) + oy;
  ax[1] = (int32_t)(int16_t)(h.y & 0xFFFFu) + ox; ay[1] = (int32_t)(int16_t)(h.y >> 16) + oy;
  ax[2] = (int32_t)(int16_t)(h.z & 0xFFFFu) + ox; ay[2] = (int32_t)(int16_t)(h.z >> 16) + oy;
  const int32_t minx = min(min(ax[0], ax[1]), ax[2]), maxx = max(max(ax[0], ax[1]), ax[2]);
  const int32_t miny = min(min(ay[0], ay[1]), ay[2]), maxy = max(max(ay[0], ay[1]), ay[2]);
  const int ts = kSubpixelLog2 + kTileLog2;
  const int bl = P.binLog2;
  const int32_t tx0 = bx << bl, ty0 = by << bl;
  const int32_t lastX = min(P.widthTiles - tx0, 1 << bl) - 1;
  const int32_t lastY = min(P.heightTiles - ty0, 1 << bl) - 1;
  r.x0 = max((minx >> ts) - tx0, 0); r.x1 = min((maxx >> ts) - tx0, lastX);
  r.y0 = max((miny >> ts) - ty0, 0); r.y1 = min((maxy >> ts) - ty0, lastY);
  const int64_t BX = (int64_t)(tx0 + r.x0) << ts, BY = (int64_t)(ty0 + r.y0) << ts;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int j = (k + 1) % 3;
    const int32_t dx = ax[j] - ax[k], dy = ay[j] - ay[k];
    const int64_t c = (int64_t)ax[k] * dy - (int64_t)ay[k] * dx;
    const int64_t px = BX + ((dy > 0) ? 0 : (1 << ts));
    const int64_t py = BY + ((dx > 0) ? (1 << ts) : 0);
    int64_t g = c - (px * dy - py * dx);
    g = g > (1ll << 30) ? (1ll << 30) : (g < -(1ll << 30) ? -(1ll << 30) : g);
    r.G[k] = (int32_t)g;
    r.sx[k] = dy * (1 << ts);
    r.sy[k] = -dx * (1 << ts);
  }
  return r;
}

// Boxes larger than this many tiles are handled by the whole wave, one
// triangle at a time, instead of by their own lane.
constexpr int kCoopArea = 16;

// Lane j's tile test, broadcast to every lane of the wave (j uniform).
__device__ __forceinline__ TileTest readTileTest(const TileTest& r, int j) {
  TileTest q;
  q.x0 = __builtin_amdgcn_readlane(r.x0, j);
  q.y0 = __builtin_amdgcn_readlane(r.y0, j);
  q.x1 = __builtin_amdgcn_readlane(r.x1, j);
  q.y1 = __builtin_amdgcn_readlane(r.y1, j);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    q.G[k] = __builtin_amdgcn_readlane(r.G[k], j);
    q.sx[k] = __builtin_amdgcn_readlane(r.sx[k], j);
    q.sy[k] = __builtin_amdgcn_readlane(r.sy[k], j);
  }
  return q;
}

// Calls f(x, y) for every tile of q's box that q hits, the wave's lanes
// covering the box 16 x 4 tiles at a time (bins are at most 16 x 16 tiles).
template <class F>
__device__ __forceinline__ void forEachHitCoop(const TileTest& q, F&& f) {
  const int lane = threadIdx.x & 63;
  const int x = q.x0 + (lane & 15);
  for (int y0 = q.y0; y0 <= q.y1; y0 += 4) {
    const int y = y0 + (lane >> 4);
    if (x <= q.x1 && y <= q.y1 && q.hit(x, y)) f(x, y);
  }
}

// How a wave walked its entries' tiles (the profiling counters' coarse
// cases): A every box at most 2x2 tiles, one lane per box; B the union box
// of the wave, one ballot per tile; C small boxes per lane and large boxes
// by the whole wave.
enum CoarseCase : int { kCaseA = 0, kCaseB = 1, kCaseC = 2 };

// A large box queued for the workgroup's waves to share: its tile test,
// bin entry slot and (wave << 6 | lane) of the entry it came from.
struct BigEntry {
  uint4 a, b, c, d;
};

__device__ __forceinline__ void putBig(BigEntry* e, const TileTest& r, uint32_t slot, uint32_t id) {
  e->a = make_uint4(r.x0, r.y0, r.x1, r.y1);
  e->b = make_uint4(r.G[0], r.G[1], r.G[2], r.sx[0]);
  e->c = make_uint4(r.sx[1], r.sx[2], r.sy[0], r.sy[1]);
  e->d = make_uint4(r.sy[2], slot, id, 0u);
}

__device__ __forceinline__ TileTest getBig(const BigEntry* e, uint32_t& slot, uint32_t& id) {
  const uint4 a = e->a, b = e->b, c = e->c, d = e->d;
  TileTest q;
  q.x0 = (int32_t)a.x; q.y0 = (int32_t)a.y; q.x1 = (int32_t)a.z; q.y1 = (int32_t)a.w;
  q.G[0] = (int32_t)b.x; q.G[1] = (int32_t)b.y; q.G[2] = (int32_t)b.z;
  q.sx[0] = (int32_t)b.w; q.sx[1] = (int32_t)c.x; q.sx[2] = (int32_t)c.y;
  q.sy[0] = (int32_t)c.z; q.sy[1] = (int32_t)c.w; q.sy[2] = (int32_t)d.x;
  slot = d.y;
  id = d.z;
  return q;
}

// One round of up to one bin entry per lane: set bit `lane` of
// emitMask[wave][tile] for every tile the entry's edges reach.  Large boxes
// of a scattered wave go to the workgroup's queue `big` (while it has room;
// *queued tells the lane) and are walked after the next barrier by
// whichever waves are free, spreading them over the workgroup.  A large box
// that reaches every tile of a complete bin (*whole) sets no mask bits: its
// wave's word fullBits[wave] holds it for every tile, and its slot goes to
// wholeSlot[wave * 64 + lane] for the writes.  Returns the CoarseCase the
// wave took.
__device__ __forceinline__ int coarseRecord(const TileTest& r, bool valid, const int bl,
                                            uint64_t (*emitMask)[kBinTiles], BigEntry* big,
                                            uint32_t* bigCount, uint32_t bigCap, uint32_t slot,
                                            bool& queued, bool binComplete, uint64_t* fullBits,
                                            uint32_t* wholeSlot, uint32_t* wholeCount, bool& whole) {
  queued = false;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  int w = valid ? r.x1 - r.x0 + 1 : 0, h = valid ? r.y1 - r.y0 + 1 : 0;
  // Every tile of the bin hit: the box is the bin and the tile test's left
  // side, linear in the tile offsets, stays within each threshold at the
  // box's far corners (|15 * s| < 2^28).
  const int side = 1 << bl;
  whole = binComplete && valid && w * h > kCoopArea && r.x0 == 0 && r.y0 == 0 && w == side && h == side;
  uint64_t wholeM = __ballot(whole);
  if (wholeM != 0) {                     // (the edge tests only where some box is the bin)
#pragma unroll
    for (int k = 0; k < 3; k++)
      whole = whole && max(0, (side - 1) * r.sx[k]) + max(0, (side - 1) * r.sy[k]) <= r.G[k];
    wholeM = __ballot(whole);
  }
  if (wholeM != 0) {
    if (lane == 0) {
      fullBits[wave] = wholeM;
      atomicAdd(wholeCount, (uint32_t)__popcll(wholeM));
    }
    if (whole) {
      wholeSlot[(wave << 6) + lane] = slot;
      w = h = 0;
    }
  }
  const bool has = valid && w > 0 && h > 0;
  const uint64_t bit = 1ull << lane;
  const bool allSmall = __ballot(has && (w > 2 || h > 2)) == 0;
  bool coop = false;
  int ux0 = 0, ux1 = -1, uy0 = 0, uy1 = -1;
  if (!allSmall) {
    // Walk the union box of the wave's triangles tile by tile (one ballot
    // per tile) when that costs less than the scattered path below: its own
    // box per lane for small boxes (as many steps as the largest, each an
    // LDS atomic that lanes hitting the same tile serialise: weight 2), plus
    // one wave step per 64 tiles for each large box.
    const bool bigOne = has && w * h > kCoopArea;
    int a0 = has ? r.x0 : 1 << 20, a1 = has ? r.x1 : -1, b0 = has ? r.y0 : 1 << 20, b1 = has ? r.y1 : -1;
    int smallArea = (has && !bigOne) ? w * h : 0;
    int bigSteps = bigOne ? (h + 3) >> 2 : 0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      a0 = min(a0, __shfl_xor(a0, o, 64)); a1 = max(a1, __shfl_xor(a1, o, 64));
      b0 = min(b0, __shfl_xor(b0, o, 64)); b1 = max(b1, __shfl_xor(b1, o, 64));
      smallArea = max(smallArea, __shfl_xor(smallArea, o, 64));
      bigSteps += __shfl_xor(bigSteps, o, 64);
    }
    ux0 = a0; ux1 = a1; uy0 = b0; uy1 = b1;
    coop = (ux1 - ux0 + 1) * (uy1 - uy0 + 1) <= 2 * smallArea + bigSteps;
  }
  if (coop) {
    for (int y = uy0; y <= uy1; y++)
      for (int x = ux0; x <= ux1; x++) {
        const bool hit = has && x >= r.x0 && x <= r.x1 && y >= r.y0 && y <= r.y1 && r.hit(x, y);
        const uint64_t m = __ballot(hit);
        if (m != 0 && lane == 0) emitMask[wave][x + (y << bl)] = m;
      }
    return kCaseB;
  }
  // Scattered boxes: small ones by their own lane, large ones queued (or,
  // past the queue's room, by the wave).
  const bool isBig = has && w * h > kCoopArea;
  const uint64_t bigM = __ballot(isBig);
  if (bigM != 0) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(bigCount, (uint32_t)__popcll(bigM));
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
    const uint32_t idx = base + rankBelow(bigM);
    if (isBig && idx < bigCap) {
      putBig(&big[idx], r, slot, ((uint32_t)wave << 6) | (uint32_t)lane);
      queued = true;
    }
  }
  if (has && !isBig) {
    // One loop over the box (row by row): as many wave steps as the
    // largest small box has tiles.
    int x = r.x0, y = r.y0;
    for (int i = 0; i < w * h; i++) {
      if (r.hit(x, y)) atomicOr((unsigned long long*)&emitMask[wave][x + (y << bl)], (unsigned long long)bit);
      if (++x > r.x1) { x = r.x0; y++; }
    }
  }
  for (uint64_t b = __ballot(isBig && !queued); b != 0; b &= b - 1) {
    const int j = (int)__builtin_ctzll(b);
    const TileTest q = readTileTest(r, j);
    const uint64_t bj = 1ull << j;
    forEachHitCoop(q, [&](int x, int y) {
      atomicOr((unsigned long long*)&emitMask[wave][x + (y << bl)], (unsigned long long)bj);
    });
  }
  return allSmall ? kCaseA : kCaseC;
}

// Coarse raster of bin `bin` of one draw (one CT-lane workgroup).
template <int CT>
__device__ __forceinline__ void coarseBody(const Params& P, const int bin) {
  constexpr int kCoarseThreads = CT;
  constexpr int kCoarseWaves = CT / 64;
  constexpr int kSegBatch = 2 * CT;     // chunks per segment-table batch (two per lane)
  __shared__ uint32_t segStart[kSegBatch + 1];
  __shared__ uint32_t segAddr[kSegBatch];
  __shared__ uint64_t emitMask[kCoarseWaves][kBinTiles];
  __shared__ uint32_t emitPre[kCoarseWaves][kBinTiles];
  // Per tile of the bin: list length so far, first, last and second-to-last
  // segment (the last two wait for their link to the segment after next).
  __shared__ uint32_t tTotal[kBinTiles];
  __shared__ uint32_t tFirst[kBinTiles];
  __shared__ uint32_t tLast[kBinTiles];
  __shared__ uint32_t tPrev[kBinTiles];
  // The round's segment map of each tile (one LDS read per lookup):
  // segments the list held before the round, its last one then, the
  // round's first new one.
  __shared__ uint4 tRound[kBinTiles];
  __shared__ uint32_t sums[kCoarseWaves + 1];
  __shared__ uint32_t sSegBase;
  __shared__ uint32_t sWritten;
  // Large boxes of the round, shared out over the waves (a quarter of the
  // round's entries fit; more are walked by their own wave).
  constexpr uint32_t kBigCap = CT / 4;
  __shared__ BigEntry bigQ[kBigCap];
  __shared__ uint32_t sBigCount;
  // Entries of the round that reach every tile of the bin: per wave, their
  // lanes (or-ed into every tile's mask once per round) and, per lane, their
  // slots for the writes.
  __shared__ uint64_t sFull[kCoarseWaves];
  __shared__ uint32_t wholeSlot[CT];
  __shared__ uint32_t sWholeCount;
  // Profiling draws: entries, rounds, emits, tiles with emits, entries per case.
  __shared__ uint32_t sProf[8];

  if (P.counters->overflow & 3u) return;
  const bool prof = P.prof != nullptr;
  if (prof && threadIdx.x < 8) sProf[threadIdx.x] = 0;
  const int bl = P.binLog2, tpb = 1 << (2 * bl);   // tiles per bin side log2, tiles per bin
  const int bx = bin % P.widthBins, by = bin / P.widthBins;
  // Every tile of the bin's square lies in the viewport.
  const bool binComplete = ((bx + 1) << bl) <= P.widthTiles && ((by + 1) << bl) <= P.heightTiles;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint32_t* bCnt = P.binCnt + (size_t)bin * P.numChunks;
  const uint32_t* bOff = P.binOff + (size_t)bin * P.numChunks;
  const uint32_t segCap = (uint32_t)P.maxTileEntries >> kTileSegLog2;   // segments the list buffer holds

  for (int i = threadIdx.x; i < kBinTiles; i += kCoarseThreads) {
    tTotal[i] = 0;
    tFirst[i] = kNoSeg;
    tLast[i] = kNoSeg;
    tPrev[i] = kNoSeg;
  }
  if (threadIdx.x == 0) sWritten = 0;
  __syncthreads();

  for (int cb = 0; cb < P.numChunks; cb += kSegBatch) {
    // Segment table of chunks [cb, cb + nb): start of each segment in the
    // bin's entry order and its address in the bin queue.
    const int nb = min(kSegBatch, P.numChunks - cb);
    const int i0 = 2 * threadIdx.x, i1 = i0 + 1;
    const uint32_t n0 = (i0 < nb) ? bCnt[cb + i0] : 0u, n1 = (i1 < nb) ? bCnt[cb + i1] : 0u;
    if (i0 < nb) segAddr[i0] = P.chunkBase[cb + i0] + bOff[cb + i0];
    if (i1 < nb) segAddr[i1] = P.chunkBase[cb + i1] + bOff[cb + i1];
    uint32_t segTotal;
    const uint32_t s0 = blockScanC<kCoarseThreads>(n0 + n1, sums, segTotal);
    if (i0 < nb) segStart[i0] = s0;
    if (i1 < nb) segStart[i1] = s0 + n0;
    if (threadIdx.x == 0) segStart[nb] = segTotal;
    __syncthreads();

    // Entry g of the batch lies in the last segment that starts at or before
    // it (an empty segment starts where its successor does, so that one is
    // never the last): a binary search of the segment table, which stays in
    // LDS for the whole batch.  The entry carries the slot and the snapped
    // vertices: one contiguous 16-byte read, no dependent header fetch.
    auto entryAt = [&](uint32_t g) -> uint4 {
      int lo = 0, hi = nb - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segStart[mid] <= g) lo = mid;
        else hi = mid - 1;
      }
      return P.binList[segAddr[lo] + (g - segStart[lo])];
    };
    // Each round's entry is loaded one round ahead, so that its latency
    // overlaps the previous round's ranking and writes.
    uint4 eNext = make_uint4(0u, 0u, 0u, 0u);
    if (threadIdx.x < segTotal) eNext = entryAt(threadIdx.x);

    for (uint32_t r0 = 0; r0 < segTotal; r0 += kCoarseThreads) {
      for (int i = threadIdx.x; i < kCoarseWaves * kBinTiles; i += kCoarseThreads)
        emitMask[i / kBinTiles][i % kBinTiles] = 0ull;
      if (threadIdx.x == 0) {
        sBigCount = 0;
        sWholeCount = 0;
      }
      if (threadIdx.x < (unsigned)kCoarseWaves) sFull[threadIdx.x] = 0ull;
      __syncthreads();
      const uint32_t g = r0 + threadIdx.x;
      bool valid = g < segTotal;
      const uint4 e = eNext;
      uint32_t slot = 0;
      TileTest r;
      r.x0 = 1; r.x1 = 0; r.y0 = 1; r.y1 = 0;
      if (valid) {
        slot = e.x;
        valid = slot < (uint32_t)P.maxSubtris;
        if (valid) r = tileTestOf(P, make_uint4(e.y, e.z, e.w, 0u), bx, by);
      }
      bool queued, whole;
      const int wcase = coarseRecord(r, valid, bl, emitMask, bigQ, &sBigCount, kBigCap, slot, queued, binComplete,
                                     sFull, wholeSlot, &sWholeCount, whole);
      if (g + kCoarseThreads < segTotal) eNext = entryAt(g + kCoarseThreads);
      if (prof) {
        const uint32_t nv = (uint32_t)__popcll(__ballot(valid));
        if (lane == 0 && nv) {
          atomicAdd(&sProf[0], nv);
          atomicAdd(&sProf[5 + wcase], nv);
        }
        if (threadIdx.x == 0) sProf[1]++;
      }
      __syncthreads();
      // The queued large boxes, one per wave at a time: a bit per tile hit
      // in the mask of the wave the entry came from.
      const uint32_t nBig = min(sBigCount, kBigCap);
      const uint32_t nWhole = sWholeCount;
      // Entries that reach every tile: their wave's bits into every tile's
      // mask, one word per thread at a time (atomically: the large boxes
      // below may set bits in the same words).
      if (nWhole > 0) {
        for (int i = threadIdx.x; i < kCoarseWaves * tpb; i += kCoarseThreads) {
          const int q = i >> (2 * bl), t = i & (tpb - 1);
          const uint64_t f = sFull[q];
          if (f) atomicOr((unsigned long long*)&emitMask[q][t], (unsigned long long)f);
        }
      }
      if (nBig > 0) {
        for (uint32_t i = (uint32_t)wave; i < nBig; i += kCoarseWaves) {
          uint32_t sj, id;
          const TileTest q = getBig(&bigQ[i], sj, id);
          const uint64_t bj = 1ull << (id & 63);
          uint64_t* mrow = emitMask[id >> 6];
          forEachHitCoop(q, [&](int x, int y) {
            atomicOr((unsigned long long*)&mrow[x + (y << bl)], (unsigned long long)bj);
          });
        }
      }
      if (nBig > 0 || nWhole > 0) __syncthreads();
      // Rank: each wave's entries of tile t follow the earlier waves' and
      // the tile's list so far; the segments the round needs past the
      // tile's last, partly filled one.
      uint32_t need = 0, kOld = 0, oldLast = kNoSeg;
      if (threadIdx.x < (unsigned)tpb) {
        const uint32_t old = tTotal[threadIdx.x];
        uint32_t acc = old;
#pragma unroll
        for (int q = 0; q < kCoarseWaves; q++) {
          emitPre[q][threadIdx.x] = acc;
          acc += (uint32_t)__popcll(emitMask[q][threadIdx.x]);
        }
        kOld = (old + kTileSeg - 1) >> kTileSegLog2;
        need = ((acc + kTileSeg - 1) >> kTileSegLog2) - kOld;
        oldLast = tLast[threadIdx.x];
        tTotal[threadIdx.x] = acc;
        if (prof && acc != old) {
          atomicAdd(&sProf[2], acc - old);
          atomicAdd(&sProf[3], 1u);
        }
      }
      // Extend: one reservation for the whole round.  The tiles (at most
      // 256: the first four waves) scan their segment needs -- a wave scan,
      // then the four wave totals -- so the reservation costs two barriers.
      static_assert(kBinTiles <= 256 && kCoarseWaves >= 4, "tiles in the first four waves");
      uint32_t x = need;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
      }
      if (wave < 4 && lane == 63) sums[wave] = x;
      __syncthreads();
      const uint32_t w0 = sums[0], w1 = sums[1], w2 = sums[2], w3 = sums[3];
      const uint32_t segsRound = w0 + w1 + w2 + w3;
      const uint32_t segEx = x - need + (wave >= 1 ? w0 : 0u) + (wave >= 2 ? w1 : 0u) + (wave >= 3 ? w2 : 0u);
      if (threadIdx.x == 0) {
        uint32_t base = 0;
        if (segsRound) {
          base = atomicAdd(&P.counters->tileEntries, segsRound << kTileSegLog2);
          if ((uint64_t)base + (segsRound << kTileSegLog2) > (uint64_t)P.maxTileEntries)
            atomicOr(&P.counters->overflow, 4u);
        }
        sSegBase = base >> kTileSegLog2;
      }
      __syncthreads();
      if (threadIdx.x < (unsigned)tpb) {
        const uint32_t first = sSegBase + segEx;
        tRound[threadIdx.x] = make_uint4(kOld, oldLast, first, 0u);
        if (need > 0) {
          // Chain the round's segments (consecutive ids) to the tile's
          // list: each segment's next one and the one after that.  The
          // list's last two segments learn the latter from a later round
          // (or never need it).  Segments past the buffer (an overflowing
          // draw, re-run larger) are never written.
          uint32_t* links = P.tileSegNext;
          const uint32_t last = first + need - 1, prev = tPrev[threadIdx.x];
          if (oldLast == kNoSeg) {
            tFirst[threadIdx.x] = first;
          } else if (oldLast < segCap) {
            links[2 * oldLast] = first;
            if (need > 1) links[2 * oldLast + 1] = first + 1;
          }
          if (prev != kNoSeg && prev < segCap) links[2 * prev + 1] = first;
          for (uint32_t j = first; j < last; j++) {
            if (j < segCap) {
              links[2 * j] = j + 1;
              if (j + 1 < last) links[2 * j + 1] = j + 2;
            }
          }
          if (last < segCap) links[2 * last] = kNoSeg;
          tPrev[threadIdx.x] = need > 1 ? last - 1 : oldLast;
          tLast[threadIdx.x] = last;
        }
      }
      __syncthreads();
      // Write the entries: list position e of tile ti is entry e & 31 of
      // the tile's (e >> 5)-th segment -- the old last segment or one of the
      // round's.  A tile's mask, rank base and segment map are read
      // together, before the test that uses them, so that each tile costs
      // one LDS round trip.  Large boxes are walked by the whole wave, one
      // entry at a time.
      auto put = [&](const uint4& q, uint32_t e, uint32_t sl) {
        const uint32_t k = e >> kTileSegLog2;
        const uint32_t sg = k < q.x ? q.y : q.z + (k - q.x);
        if (sg < segCap) P.tileList[((size_t)sg << kTileSegLog2) + (e & (kTileSeg - 1))] = sl;
      };
      const bool hasBox = valid && !whole && r.x0 <= r.x1 && r.y0 <= r.y1;
      const int boxArea = hasBox ? (r.x1 - r.x0 + 1) * (r.y1 - r.y0 + 1) : 0;
      const bool bigBox = boxArea > kCoopArea;
      if (hasBox && !bigBox) {
        const uint64_t bit = 1ull << lane;
        // One loop over the box (row by row): the wave runs as many steps
        // as its largest box has tiles.
        int x = r.x0, y = r.y0;
        for (int i = 0; i < boxArea; i++) {
          const int ti = x + (y << bl);
          const uint64_t m = emitMask[wave][ti];
          const uint32_t pre = emitPre[wave][ti];
          const uint4 q = tRound[ti];
          if (m & bit) put(q, pre + rankBelow(m), slot);
          if (++x > r.x1) { x = r.x0; y++; }
        }
      }
      for (uint64_t b = __ballot(bigBox && !queued); b != 0; b &= b - 1) {
        const int j = (int)__builtin_ctzll(b);
        const uint64_t bj = 1ull << j, ltj = bj - 1;
        const uint32_t sj = (uint32_t)__builtin_amdgcn_readlane((int)slot, j);
        const int qx0 = __builtin_amdgcn_readlane(r.x0, j), qx1 = __builtin_amdgcn_readlane(r.x1, j);
        const int qy0 = __builtin_amdgcn_readlane(r.y0, j), qy1 = __builtin_amdgcn_readlane(r.y1, j);
        const int x = qx0 + (lane & 15);
        for (int y0 = qy0; y0 <= qy1; y0 += 4) {
          const int y = y0 + (lane >> 4);
          if (x <= qx1 && y <= qy1) {
            const int ti = x + (y << bl);
            const uint64_t m = emitMask[wave][ti];
            const uint32_t pre = emitPre[wave][ti];
            const uint4 q = tRound[ti];
            if (m & bj) put(q, pre + (uint32_t)__popcll(m & ltj), sj);
          }
        }
      }
      for (uint32_t i = (uint32_t)wave; i < nBig; i += kCoarseWaves) {
        uint32_t sj, id;
        const TileTest qt = getBig(&bigQ[i], sj, id);
        const uint64_t bj = 1ull << (id & 63), ltj = bj - 1;
        const uint64_t* mrow = emitMask[id >> 6];
        const uint32_t* prow = emitPre[id >> 6];
        const int x = qt.x0 + (lane & 15);
        for (int y0 = qt.y0; y0 <= qt.y1; y0 += 4) {
          const int y = y0 + (lane >> 4);
          if (x <= qt.x1 && y <= qt.y1) {
            const int ti = x + (y << bl);
            const uint64_t m = mrow[ti];
            const uint32_t pre = prow[ti];
            const uint4 q = tRound[ti];
            if (m & bj) put(q, pre + (uint32_t)__popcll(m & ltj), sj);
          }
        }
      }
      // Entries that reach every tile: per (source wave, 64 tiles of a
      // complete bin of at least 8 x 8 tiles), one lane per tile reads the
      // tile's mask, rank base and segment map once and writes every such
      // entry of that wave; the pairs are shared out over the waves.
      if (nWhole > 0) {
        const int chunksLog2 = 2 * bl - 6;
        for (int i = wave; i < (kCoarseWaves << chunksLog2); i += kCoarseWaves) {
          const int q = i >> chunksLog2;
          const uint64_t f = sFull[q];
          if (f == 0) continue;
          const int ti = ((i & ((1 << chunksLog2) - 1)) << 6) + lane;
          const uint64_t m = emitMask[q][ti];
          const uint32_t pre = emitPre[q][ti];
          const uint4 tq = tRound[ti];
          for (uint64_t b = f; b != 0; b &= b - 1) {
            const int j = (int)__builtin_ctzll(b);
            put(tq, pre + (uint32_t)__popcll(m & ((1ull << j) - 1)), wholeSlot[(q << 6) + j]);
          }
        }
      }
      __syncthreads();
    }
    __syncthreads();
  }
  // The bin's lists: first segment and length per tile, and the entries
  // written (the lists' total length) into the counter block.
  if (threadIdx.x < (unsigned)tpb) {
    const int x = threadIdx.x & ((1 << bl) - 1), y = threadIdx.x >> bl;
    const int gx = (bx << bl) + x, gy = (by << bl) + y;
    if (gx < P.widthTiles && gy < P.heightTiles) {
      P.tileStart[gx + gy * P.widthTiles] = tFirst[threadIdx.x];
      P.tileCount[gx + gy * P.widthTiles] = tTotal[threadIdx.x];
    }
    if (tTotal[threadIdx.x]) atomicAdd(&sWritten, tTotal[threadIdx.x]);
  }
  __syncthreads();
  if (threadIdx.x == 0 && sWritten) atomicAdd(&P.counters->tileListEntries, sWritten);
  if (prof && threadIdx.x == 0) {
    // (The rounds ended with a barrier: every wave's additions are in.)
    const uint32_t entries = sProf[0], rounds = sProf[1], emits = sProf[2];
    profAdd(P.prof, kProfCoarseBins, entries ? 1 : 0, 0);
    profAdd(P.prof, kProfCoarseRoundsPerBin, rounds, entries ? 1 : 0);
    profAdd(P.prof, kProfCoarseTrisPerRound, entries, rounds);
    profAdd(P.prof, kProfCoarseTilesPerRound, sProf[3], rounds);
    profAdd(P.prof, kProfCoarseEmitsPerRound, emits, rounds);
    profAdd(P.prof, kProfCoarseEmitsPerTri, emits, entries);
    profAdd(P.prof, kProfCoarseCaseA, 100ull * sProf[5], entries);
    profAdd(P.prof, kProfCoarseCaseB, 100ull * sProf[6], entries);
    profAdd(P.prof, kProfCoarseCaseC, 100ull * sProf[7], entries);
  }
}

template <int CT>
__global__ void __launch_bounds__(CT) coarseKernel(const Params P) {
  // XCD-aware order: workgroup b runs on XCD b % 8, so each XCD gets one
  // contiguous band of bins.  Neighbouring bins' runs of a chunk share the
  // bin queue's cache lines, which are then read into one L2 instead of
  // several.
  const int per = (P.numBins + 7) >> 3;
  const int bin = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (bin >= P.numBins) return;
  coarseBody<CT>(P, bin);
}

template <int CT>
__global__ void __launch_bounds__(CT) coarseBatchKernel(const Params* __restrict__ views,
                                                        const int32_t* __restrict__ blockStart,
                                                        const int32_t* __restrict__ order, int numViews) {
  const int k = findView(blockStart, numViews, (int)blockIdx.x);
  coarseBody<CT>(views[order[k]], (int)blockIdx.x - blockStart[k]);
}

// One workgroup per view: clears the view's counter block.
__global__ void __launch_bounds__(64) resetBatchKernel(const Params* __restrict__ views) {
  const Params& P = views[blockIdx.x];
  uint32_t* w = reinterpret_cast<uint32_t*>(P.counters);
  for (int i = threadIdx.x; i < (int)(sizeof(Counters) / 4); i += 64) w[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) P.counters->numSubtris = (uint32_t)P.numTris;
}

void launchReset(const Params& P, hipStream_t stream) {
  hipLaunchKernelGGL(resetCountersKernel, dim3(1), dim3(64), 0, stream, P.counters, (uint32_t)P.numTris, P.prof);
}

int coarseShapeFor(const Params& P) {
  if (P.coarseThreads == 256 || P.coarseThreads == 1024) return P.coarseThreads;
  // A single draw leaves the GPU room for the large workgroups (measured
  // faster from 32k triangles at 256^2 to 1M at 2048^2); the small ones pay
  // off when many draws share the GPU (batches tune it).
  return 1024;
}

void launchCoarse(const Params& P, hipStream_t stream) {
  if (coarseShapeFor(P) == 1024)
    hipLaunchKernelGGL(coarseKernel<1024>, dim3(((P.numBins + 7) >> 3) * 8), dim3(1024), 0, stream, P);
  else
    hipLaunchKernelGGL(coarseKernel<256>, dim3(((P.numBins + 7) >> 3) * 8), dim3(256), 0, stream, P);
}

void launchResetBatch(const Params* views, int numViews, hipStream_t stream) {
  if (numViews <= 0) return;
  hipLaunchKernelGGL(resetBatchKernel, dim3(numViews), dim3(64), 0, stream, views);
}

void launchCoarseBatch(int threads, const Params* views, const int32_t* blockStart, const int32_t* order,
                       int numViews, int totalBlocks, hipStream_t stream) {
  if (totalBlocks <= 0) return;
  if (threads == 1024)
    hipLaunchKernelGGL(coarseBatchKernel<1024>, dim3(totalBlocks), dim3(1024), 0, stream, views, blockStart, order,
                       numViews);
  else
    hipLaunchKernelGGL(coarseBatchKernel<256>, dim3(totalBlocks), dim3(256), 0, stream, views, blockStart, order,
                       numViews);
}

}  // namespace crhip
