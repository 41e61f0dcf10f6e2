// Stages 2 and 3: order-preserving bin and tile queues.
//
// The reference streams triangles through 16 persistent bin blocks into
// linked 512-entry segments per (bin, stream) and merges the 16 streams per
// bin in the coarse stage (cudaraster/cuda/BinRaster.inl:19-469,
// CoarseRaster.inl:28-809).  Here both levels are counting sorts that keep
// submission order by construction and produce one contiguous list per bin
// and per tile:
//
//   binCount    chunk of kBinChunk triangles per block: per-bin emit counts
//   binScan     per bin, exclusive scan of the chunk counts; the last block
//               to finish also computes the bin starts (prefix over bins)
//   binScatter  chunk per block: subtriangles of each 256-triangle round are
//               compacted in LDS (wave prefix sums), then every bin's entries
//               are ranked with per-wave 64-bit emit masks + popcounts and
//               written in (triangle, subtriangle) order
//   coarse      one block (16 waves) per bin: count per tile, allocate the
//               bin's tile lists with one atomic, then emit in rounds of 1024
//               entries with the same emit-mask ranking
//
// Binning is conservative (bounding boxes, plus a tile-level edge test in
// the coarse stage); the fine stage computes exact coverage, so neither
// level can change the image.
#include "cr_device.h"

namespace crhip {

constexpr int kBinChunk = kBinChunkTris;   // triangles per bin chunk
constexpr int kBinThreads = 256;
constexpr int kCoarseThreads = 1024;
constexpr int kCoarseWaves = kCoarseThreads / 64;

// Slot of subtriangle s of triangle t (BinRaster.inl:148-157 encodes the
// same choice as t*8+s).
__device__ __forceinline__ uint32_t subtriSlot(const Params& P, uint32_t t, int k, int s) {
  return (k == 1) ? t : P.triHeader[t].misc + (uint32_t)s;
}

struct Box { int32_t x0, y0, x1, y1; };   // inclusive, empty when x0 > x1

// Absolute subpixel bounding box of a header triangle.
__device__ __forceinline__ void headerBounds(const Params& P, const Header& h, int32_t& minx, int32_t& miny,
                                             int32_t& maxx, int32_t& maxy) {
  int32_t ox = P.viewportWidth << (kSubpixelLog2 - 1), oy = P.viewportHeight << (kSubpixelLog2 - 1);
  minx = min(min((int32_t)h.v0x, (int32_t)h.v1x), (int32_t)h.v2x) + ox;
  maxx = max(max((int32_t)h.v0x, (int32_t)h.v1x), (int32_t)h.v2x) + ox;
  miny = min(min((int32_t)h.v0y, (int32_t)h.v1y), (int32_t)h.v2y) + oy;
  maxy = max(max((int32_t)h.v0y, (int32_t)h.v1y), (int32_t)h.v2y) + oy;
}

// Bins (128x128 px) touched by the triangle's bounding box.
__device__ __forceinline__ Box binBox(const Params& P, const Header& h) {
  int32_t minx, miny, maxx, maxy;
  headerBounds(P, h, minx, miny, maxx, maxy);
  Box b;
  const int sh = kSubpixelLog2 + kTileLog2 + kBinLog2;
  if (maxx < 0 || maxy < 0 || minx >= (P.widthPixels << kSubpixelLog2) || miny >= (P.heightPixels << kSubpixelLog2)) {
    b.x0 = 1; b.x1 = 0; b.y0 = 1; b.y1 = 0;
    return b;
  }
  b.x0 = max(minx >> sh, 0); b.x1 = min(maxx >> sh, P.widthBins - 1);
  b.y0 = max(miny >> sh, 0); b.y1 = min(maxy >> sh, P.heightBins - 1);
  return b;
}

__device__ __forceinline__ uint64_t laneMaskLt() {
  uint32_t lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(64) resetCountersKernel(Counters* c, uint32_t numTris) {
  uint32_t* w = reinterpret_cast<uint32_t*>(c);
  for (int i = threadIdx.x; i < (int)(sizeof(Counters) / 4); i += 64) w[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) c->numSubtris = numTris;
}

__global__ void __launch_bounds__(kBinThreads) binCountKernel(const Params P) {
  __shared__ uint32_t cnt[kMaxBins];
  if (P.counters->overflow & 1u) return;   // clipped slots may lie past the buffer
  for (int i = threadIdx.x; i < P.numBins; i += kBinThreads) cnt[i] = 0;
  __syncthreads();
  int chunk = blockIdx.x;
  for (int r = 0; r < kBinChunk; r += kBinThreads) {
    int t = chunk * kBinChunk + r + threadIdx.x;
    if (t >= P.numTris) break;
    int k = P.triSubtris[t];
    for (int s = 0; s < k; s++) {
      Header h = P.triHeader[subtriSlot(P, t, k, s)];
      Box b = binBox(P, h);
      for (int y = b.y0; y <= b.y1; y++)
        for (int x = b.x0; x <= b.x1; x++) atomicAdd(&cnt[x + y * P.widthBins], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P.numBins; i += kBinThreads)
    P.binChunkCount[(size_t)i * P.numChunks + chunk] = cnt[i];
}

// Block-wide exclusive scan of one value per thread (1024 threads).
template <int NT>
__device__ __forceinline__ uint32_t blockExclusiveScan(uint32_t v, uint32_t* warpSums, uint32_t& total) {
  constexpr int NW = NT / 64;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) warpSums[wave] = x;
  __syncthreads();
  if (threadIdx.x < 64) {
    uint32_t w = (threadIdx.x < NW) ? warpSums[threadIdx.x] : 0u;
    uint32_t s = w;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t y = __shfl_up(s, o, 64);
      if ((int)threadIdx.x >= o) s += y;
    }
    if (threadIdx.x < NW) warpSums[threadIdx.x] = s - w;
    if (threadIdx.x == NW - 1) warpSums[NW] = s;
  }
  __syncthreads();
  total = warpSums[NW];
  uint32_t r = warpSums[wave] + x - v;
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(1024) binScanKernel(const Params P, uint32_t* binStart) {
  __shared__ uint32_t sums[17];
  __shared__ uint32_t isLast;
  int b = blockIdx.x;
  uint32_t* row = P.binChunkCount + (size_t)b * P.numChunks;
  uint32_t carry = 0;
  for (int base = 0; base < P.numChunks; base += 1024) {
    int i = base + threadIdx.x;
    uint32_t v = (i < P.numChunks) ? row[i] : 0u;
    uint32_t total;
    uint32_t ex = blockExclusiveScan<1024>(v, sums, total);
    if (i < P.numChunks) row[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) {
    P.binTotal[b] = carry;
    __threadfence();
    isLast = (atomicAdd(&P.counters->binScanDone, 1u) == (uint32_t)P.numBins - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (!isLast) return;
  __threadfence();
  // Last block: bin starts (prefix over bins) and the grand total.
  uint32_t v = (threadIdx.x < (unsigned)P.numBins) ? __atomic_load_n(&P.binTotal[threadIdx.x], __ATOMIC_RELAXED) : 0u;
  uint32_t total;
  uint32_t ex = blockExclusiveScan<1024>(v, sums, total);
  if (threadIdx.x < (unsigned)P.numBins) binStart[threadIdx.x] = ex;
  if (threadIdx.x == 0) {
    P.counters->binEntries = total;
    if (total > (uint32_t)P.maxBinEntries) atomicOr(&P.counters->overflow, 2u);
  }
}

__global__ void __launch_bounds__(kBinThreads) binScatterKernel(const Params P, const uint32_t* binStart) {
  constexpr int NW = kBinThreads / 64;
  __shared__ uint32_t outBase[kMaxBins];
  __shared__ uint32_t run[kMaxBins];
  __shared__ uint64_t emitMask[NW][kMaxBins];
  __shared__ uint32_t emitPre[NW][kMaxBins];
  __shared__ uint32_t list[kBinThreads * 7];
  __shared__ uint32_t sums[NW + 1];

  if (P.counters->overflow & 3u) return;
  int chunk = blockIdx.x;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < P.numBins; i += kBinThreads) {
    outBase[i] = binStart[i] + P.binChunkCount[(size_t)i * P.numChunks + chunk];
    run[i] = 0;
  }
  __syncthreads();

  for (int r = 0; r < kBinChunk; r += kBinThreads) {
    int t0 = chunk * kBinChunk + r;
    if (t0 >= P.numTris) break;
    int t = t0 + threadIdx.x;
    int k = (t < P.numTris) ? P.triSubtris[t] : 0;
    uint32_t total;
    uint32_t pos = blockExclusiveScan<kBinThreads>((uint32_t)k, sums, total);
    for (int s = 0; s < k; s++) list[pos + s] = subtriSlot(P, t, k, s);
    __syncthreads();

    for (uint32_t sub = 0; sub < total; sub += kBinThreads) {
      for (int i = threadIdx.x; i < NW * P.numBins; i += kBinThreads)
        emitMask[i / P.numBins][i % P.numBins] = 0ull;
      __syncthreads();
      uint32_t e = sub + threadIdx.x;
      bool valid = e < total;
      uint32_t slot = valid ? list[e] : 0u;
      Box b = {1, 1, 0, 0};
      if (valid) {
        b = binBox(P, P.triHeader[slot]);
        uint64_t bit = 1ull << lane;
        for (int y = b.y0; y <= b.y1; y++)
          for (int x = b.x0; x <= b.x1; x++) atomicOr((unsigned long long*)&emitMask[wave][x + y * P.widthBins], bit);
      }
      __syncthreads();
      for (int i = threadIdx.x; i < P.numBins; i += kBinThreads) {
        uint32_t acc = run[i];
#pragma unroll
        for (int w = 0; w < NW; w++) {
          emitPre[w][i] = acc;
          acc += (uint32_t)__popcll(emitMask[w][i]);
        }
        run[i] = acc;
      }
      __syncthreads();
      if (valid) {
        uint64_t lt = laneMaskLt();
        for (int y = b.y0; y <= b.y1; y++)
          for (int x = b.x0; x <= b.x1; x++) {
            int bin = x + y * P.widthBins;
            uint32_t p = outBase[bin] + emitPre[wave][bin] + (uint32_t)__popcll(emitMask[wave][bin] & lt);
            if (p < (uint32_t)P.maxBinEntries) P.binList[p] = slot;
          }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Coarse stage
// ---------------------------------------------------------------------------

// Tiles of bin (bx, by) touched by a triangle: bounding box clipped to the
// bin and to the surface, then a conservative edge test per tile against the
// closed tile square (no sample lies on a tile border, so a tile whose
// square is strictly outside one edge holds no covered sample).
struct TriTiles {
  int32_t x0, y0, x1, y1;         // bin-local tile box (inclusive)
  int32_t ax[3], ay[3], dx[3], dy[3];
  int64_t c[3];                   // a_k x d_k (absolute subpixels)
};

__device__ __forceinline__ TriTiles triTiles(const Params& P, const Header& h, int bx, int by) {
  TriTiles r;
  int32_t ox = P.viewportWidth << (kSubpixelLog2 - 1), oy = P.viewportHeight << (kSubpixelLog2 - 1);
  r.ax[0] = h.v0x + ox; r.ay[0] = h.v0y + oy;
  r.ax[1] = h.v1x + ox; r.ay[1] = h.v1y + oy;
  r.ax[2] = h.v2x + ox; r.ay[2] = h.v2y + oy;
  int32_t minx = min(min(r.ax[0], r.ax[1]), r.ax[2]), maxx = max(max(r.ax[0], r.ax[1]), r.ax[2]);
  int32_t miny = min(min(r.ay[0], r.ay[1]), r.ay[2]), maxy = max(max(r.ay[0], r.ay[1]), r.ay[2]);
  const int ts = kSubpixelLog2 + kTileLog2;
  int32_t tx0 = bx << kBinLog2, ty0 = by << kBinLog2;
  int32_t lastX = min(P.widthTiles - tx0, 1 << kBinLog2) - 1;
  int32_t lastY = min(P.heightTiles - ty0, 1 << kBinLog2) - 1;
  r.x0 = max((minx >> ts) - tx0, 0); r.x1 = min((maxx >> ts) - tx0, lastX);
  r.y0 = max((miny >> ts) - ty0, 0); r.y1 = min((maxy >> ts) - ty0, lastY);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int j = (k + 1) % 3;
    r.dx[k] = r.ax[j] - r.ax[k];
    r.dy[k] = r.ay[j] - r.ay[k];
    r.c[k] = (int64_t)r.ax[k] * r.dy[k] - (int64_t)r.ay[k] * r.dx[k];
  }
  return r;
}

// Does the closed square of bin-local tile (x, y) reach the inner side of
// every edge?  Inner side: (a - p) x d >= 0, i.e. p x d <= a x d.
__device__ __forceinline__ bool tileHit(const TriTiles& r, int bx, int by, int x, int y) {
  const int ts = kSubpixelLog2 + kTileLog2;
  int64_t X0 = (int64_t)(((bx << kBinLog2) + x) << ts), Y0 = (int64_t)(((by << kBinLog2) + y) << ts);
  int64_t X1 = X0 + (1 << ts), Y1 = Y0 + (1 << ts);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int64_t px = (r.dy[k] > 0) ? X0 : X1;      // minimises px*dy
    int64_t py = (r.dx[k] > 0) ? Y1 : Y0;      // maximises py*dx
    if (px * r.dy[k] - py * r.dx[k] > r.c[k]) return false;
  }
  return true;
}

__global__ void __launch_bounds__(kCoarseThreads) coarseKernel(const Params P, const uint32_t* binStart) {
  __shared__ uint32_t cnt[kBinTiles];
  __shared__ uint32_t run[kBinTiles];
  __shared__ uint64_t emitMask[kCoarseWaves][kBinTiles];
  __shared__ uint32_t emitPre[kCoarseWaves][kBinTiles];
  __shared__ uint32_t sums[kCoarseWaves + 1];
  __shared__ uint32_t sBase;

  if (P.counters->overflow & 3u) return;
  int bin = blockIdx.x;
  int bx = bin % P.widthBins, by = bin / P.widthBins;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t first = binStart[bin];
  uint32_t num = P.binTotal[bin];

  for (int i = threadIdx.x; i < kBinTiles; i += kCoarseThreads) cnt[i] = 0;
  __syncthreads();

  // Pass 1: count entries per tile.
  for (uint32_t e = threadIdx.x; e < num; e += kCoarseThreads) {
    uint32_t slot = P.binList[first + e];
    if (slot >= (uint32_t)P.maxSubtris) continue;
    TriTiles r = triTiles(P, P.triHeader[slot], bx, by);
    for (int y = r.y0; y <= r.y1; y++)
      for (int x = r.x0; x <= r.x1; x++)
        if (tileHit(r, bx, by, x, y)) atomicAdd(&cnt[x + (y << kBinLog2)], 1u);
  }
  __syncthreads();

  // Allocate the bin's tile lists: exclusive scan over the 256 tiles.
  uint32_t v = (threadIdx.x < kBinTiles) ? cnt[threadIdx.x] : 0u;
  uint32_t total;
  uint32_t ex = blockExclusiveScan<kCoarseThreads>(v, sums, total);
  if (threadIdx.x == 0) {
    uint32_t base = total ? atomicAdd(&P.counters->tileEntries, total) : 0u;
    if (total && (uint64_t)base + total > (uint64_t)P.maxTileEntries) {
      atomicOr(&P.counters->overflow, 4u);
      base = 0xFFFFFFFFu;
    }
    sBase = base;
  }
  __syncthreads();
  uint32_t base = sBase;
  bool ok = (base != 0xFFFFFFFFu);
  if (threadIdx.x < kBinTiles) {
    int x = threadIdx.x & ((1 << kBinLog2) - 1), y = threadIdx.x >> kBinLog2;
    int gx = (bx << kBinLog2) + x, gy = (by << kBinLog2) + y;
    if (gx < P.widthTiles && gy < P.heightTiles) {
      int tile = gx + gy * P.widthTiles;
      P.tileStart[tile] = ok ? base + ex : 0u;
      P.tileCount[tile] = ok ? v : 0u;
    }
    run[threadIdx.x] = base + ex;
  }
  if (!ok || total == 0) return;
  __syncthreads();

  // Pass 2: emit in submission order, 1024 entries per round.
  for (uint32_t r0 = 0; r0 < num; r0 += kCoarseThreads) {
    for (int i = threadIdx.x; i < kCoarseWaves * kBinTiles; i += kCoarseThreads)
      emitMask[i / kBinTiles][i % kBinTiles] = 0ull;
    __syncthreads();
    uint32_t e = r0 + threadIdx.x;
    bool valid = e < num;
    uint32_t slot = 0;
    TriTiles r;
    r.x0 = 1; r.x1 = 0; r.y0 = 1; r.y1 = 0;
    if (valid) {
      slot = P.binList[first + e];
      valid = slot < (uint32_t)P.maxSubtris;
    }
    if (valid) {
      r = triTiles(P, P.triHeader[slot], bx, by);
      uint64_t bit = 1ull << lane;
      for (int y = r.y0; y <= r.y1; y++)
        for (int x = r.x0; x <= r.x1; x++)
          if (tileHit(r, bx, by, x, y)) atomicOr((unsigned long long*)&emitMask[wave][x + (y << kBinLog2)], bit);
    }
    __syncthreads();
    if (threadIdx.x < kBinTiles) {
      uint32_t acc = run[threadIdx.x];
#pragma unroll
      for (int w = 0; w < kCoarseWaves; w++) {
        emitPre[w][threadIdx.x] = acc;
        acc += (uint32_t)__popcll(emitMask[w][threadIdx.x]);
      }
      run[threadIdx.x] = acc;
    }
    __syncthreads();
    if (valid) {
      uint64_t lt = laneMaskLt();
      for (int y = r.y0; y <= r.y1; y++)
        for (int x = r.x0; x <= r.x1; x++) {
          int ti = x + (y << kBinLog2);
          uint64_t m = emitMask[wave][ti];
          uint32_t p = emitPre[wave][ti] + (uint32_t)__popcll(m & lt);
          if ((m & (1ull << lane)) && p < (uint32_t)P.maxTileEntries) P.tileList[p] = slot;
        }
    }
    __syncthreads();
  }
}

void launchReset(const Params& P, hipStream_t stream) {
  hipLaunchKernelGGL(resetCountersKernel, dim3(1), dim3(64), 0, stream, P.counters, (uint32_t)P.numTris);
}

void launchBin(const Params& P, uint32_t* binStart, hipStream_t stream) {
  if (P.numChunks > 0) hipLaunchKernelGGL(binCountKernel, dim3(P.numChunks), dim3(kBinThreads), 0, stream, P);
  hipLaunchKernelGGL(binScanKernel, dim3(P.numBins), dim3(1024), 0, stream, P, binStart);
  if (P.numChunks > 0) hipLaunchKernelGGL(binScatterKernel, dim3(P.numChunks), dim3(kBinThreads), 0, stream, P, (const uint32_t*)binStart);
}

void launchCoarse(const Params& P, const uint32_t* binStart, hipStream_t stream) {
  hipLaunchKernelGGL(coarseKernel, dim3(P.numBins), dim3(kCoarseThreads), 0, stream, P, binStart);
}

}  // namespace crhip
