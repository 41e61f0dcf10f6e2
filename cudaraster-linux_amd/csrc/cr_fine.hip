// Stage 4: fine raster, one wave64 per 8x8 tile, one lane per pixel.
//
// Same contract as fineRasterImpl_SingleSample / _MultiSample
// (cudaraster/cuda/FineRaster.inl:499-752, 855-1126): for every triangle of
// the tile, in submission order, exact sample coverage (top-left rule), LESS
// depth test, the fragment shader at the pixel centre / coverage centroid,
// and the blend ROP.
//
// Wave64 design (not the reference's fragment-distribution warp):
//  * each lane owns one pixel and keeps all its samples (depth, colour or
//    visibility, triangle id) in registers for the whole tile, so fragments
//    of one pixel are applied in order by construction -- no intra-wave
//    conflict resolution and exactly one coalesced read/write of the tile;
//  * the tile's triangle list is consumed in chunks of 32: lane j tests
//    entry j against the tile's hierarchical Z; the survivors, gathered over
//    chunks into batches of up to 32, are set up one per lane (tile-relative
//    edge thresholds per sample, depth plane per sample) into an LDS record
//    array, and the wave then walks the records with broadcast LDS reads --
//    per (triangle, tile) the lanes only do integer compares, one depth add
//    per sample and register selects;
//  * for Replace and DepthOnly blending the colour written to a sample is
//    the shaded colour of the last fragment that passed there, so shading is
//    deferred: the loop records (subtriangle slot, centroid) per sample and
//    the tile is shaded once per (pixel, distinct triangle) at the end.
//    SrcOver and Additive blend in submission order inside the loop.
#include <utility>

#include "cr_device.h"
#include "cr_texphong.h"

namespace crhip {

// Per-(triangle, tile) record in LDS, written by the lane that sets the
// triangle up and read by every lane of the wave.  Everything that does not
// depend on the pixel is precomputed there: the per-record loop only does
// what each pixel needs.
template <int S>
struct Rec {
  static constexpr int N = 1 << S;
  static constexpr int R = 0;                  // 3*N coverage thresholds, R[k*N+i]
  static constexpr int D = 3 * N;              // per edge: dy, -dx (so that L = lx16*dy + ly16*(-dx))
  static constexpr int Z = D + 6;              // depth plane: (zx lo16, zy lo16), (zx hi16, zy hi16)
  static constexpr int ZS = Z + 2;             // depth at the tile's first pixel, per sample [N]
  static constexpr int SLOT = ZS + N;          // subtriangle slot (deferred shading: slot << 4)
  static constexpr int TRI = SLOT + 1;         // source triangle index
  static constexpr int SH = TRI + 1;           // in-order shading: W/U/V planes (9) + colours (12)
  static constexpr int SIZE_VIS = (SH + 3) & ~3;
  static constexpr int SIZE_FULL = (SH + 21 + 3) & ~3;
};

// x * K (mod 2^32) for a constant K < 16 as shifts and adds (no 32-bit
// multiply).
template <int K>
__device__ __forceinline__ uint32_t mulSmall(uint32_t x) {
  static_assert(K >= 0 && K < 16, "small constant");
  uint32_t r = 0;
  if (K & 1) r += x;
  if (K & 2) r += x << 1;
  if (K & 4) r += x << 2;
  if (K & 8) r += x << 3;
  return r;
}

// out[i] = base + zx * pattern(i) + zy * i for every sample i.
template <int S, int... I>
__device__ __forceinline__ void sampleDepths(uint32_t* out, uint32_t base, uint32_t zx, uint32_t zy,
                                             std::integer_sequence<int, I...>) {
  ((out[I] = base + mulSmall<msaaPattern(S, I)>(zx) + mulSmall<I>(zy)), ...);
}

// (hi << 16) + lo as one full-rate instruction.  Written out so that the
// compiler does not fold the shift back into a quarter-rate 32-bit multiply
// of the 24-bit products above it.
__device__ __forceinline__ uint32_t shl16Add(uint32_t hi, uint32_t lo) {
  uint32_t r;
  asm("v_lshl_add_u32 %0, %1, 16, %2" : "=v"(r) : "v"(hi), "v"(lo));
  return r;
}

// a.lo*b.lo + a.hi*b.hi + c over unsigned 16-bit halves, modulo 2^32
// (v_dot2_u32_u16 without clamping).
typedef unsigned short crUShort2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t dot2u(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_udot2(__builtin_bit_cast(crUShort2, a), __builtin_bit_cast(crUShort2, b), c, false);
}

__device__ __forceinline__ uint32_t waveMax(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

// Set bits of `mask` below this lane (v_mbcnt: no 64-bit lane mask is built).
__device__ __forceinline__ uint32_t rankBelow(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Sample order by distance to the pixel centre, lowest index first on ties:
// the centroid of a coverage mask (selectMSAACentroid, Util.hpp:182-203) is
// its first covered sample in this order, or the pixel centre (code 15) when
// none or all samples are covered.  The fine stage evaluates it as a
// compile-time priority chain instead of the reference's search loop.
template <int S>
struct CentroidOrder {
  int order[8];
  constexpr CentroidOrder() : order() {
    const int n = 1 << S;
    bool used[8] = {false, false, false, false, false, false, false, false};
    for (int k = 0; k < n; k++) {
      int best = -1, bestDist = 1 << 30;
      for (int i = 0; i < n; i++) {
        if (used[i]) continue;
        const int dx = msaaPattern(S, i) * 2 + 1 - n, dy = i * 2 + 1 - n;
        const int d = dx * dx + dy * dy;
        if (d < bestDist) { best = i; bestDist = d; }
      }
      used[best] = true;
      order[k] = best;
    }
  }
};

// Perspective-correct barycentrics at the centroid (computeBarys,
// FineRaster.inl:21-44, 93-114) and GouraudShader + toABGR
// (PixelPipe.inl:52-69).
template <int S>
__device__ __forceinline__ uint32_t shadeGouraud(bool lerpOn, int code, int px, int py, const uint4& t1,
                                                 const uint4& t2, const uint4& t3, const float4& c0,
                                                 const float4& c1, const float4& c2) {
  constexpr int N = 1 << S;
  float b0, b1, b2;
  if (lerpOn) {
    const int cx = (code == 15) ? N : 2 * msaaPattern(S, code) + 1;
    const int cy = (code == 15) ? N : 2 * code + 1;
    const int32_t sx = px * (2 * N) + cx, sy = py * (2 * N) + cy;
    const int32_t W = wadd(wadd(wmul((int32_t)t1.x, sx), wmul((int32_t)t1.y, sy)), (int32_t)t1.z);
    const int32_t U = wadd(wadd(wmul((int32_t)t1.w, sx), wmul((int32_t)t2.x, sy)), (int32_t)t2.y);
    const int32_t V = wadd(wadd(wmul((int32_t)t2.z, sx), wmul((int32_t)t2.w, sy)), (int32_t)t3.x);
    const float w = 1.0f / (float)W;
    const float u = w * (float)U;
    const float v = w * (float)V;
    b0 = 1.0f - u - v; b1 = u; b2 = v;
  } else {
    b0 = 0.0f; b1 = 0.0f; b2 = 1.0f;                  // FineRaster.inl:72-81
  }
  return toABGR(fmaf(c2.x, b2, fmaf(c0.x, b0, c1.x * b1)), fmaf(c2.y, b2, fmaf(c0.y, b0, c1.y * b1)),
                fmaf(c2.z, b2, fmaf(c0.z, b0, c1.z * b1)), fmaf(c2.w, b2, fmaf(c0.w, b0, c1.w * b1)));
}

__device__ __forceinline__ float4 varying0(const uint8_t* verts, uint32_t vi, int stride) {
  // Unsigned 32x32 -> 64-bit product: one v_mad_u64_u32 with the base.
  return *reinterpret_cast<const float4*>(verts + (uint64_t)vi * (uint32_t)stride + 16);
}

constexpr int kFineWaves = 4;     // tiles (waves) per workgroup
constexpr int kChunk = 32;        // triangles set up per chunk
constexpr int kMaxDefer = 8;      // chunks hierarchical-Z survivors may wait before their walk

// LDS ordering inside one wave: every lane's record stores must land before
// any lane reads them; a wavefront-scope fence plus the wave barrier orders
// the LDS operations of the wave (no cross-wave synchronisation is needed:
// each wave owns its tile and its LDS slice).
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so
// workgroup b runs on XCD b % 8; group (b % 8) * per + b / 8 gives each XCD
// one contiguous band of tile groups, and triangles shared by neighbouring
// tiles stay in one L2.
__device__ __forceinline__ int xcdGroup(int groups) {
  const int per = (groups + 7) >> 3;
  return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// Waves per SIMD the 1x-4x deferred-shading copies are held to (at most
// 64 registers each): the batched copies' extra round loop would otherwise
// cost them an eighth of their occupancy.
template <int S, bool VIS>
constexpr int fineMinWaves() { return (VIS && S <= 2) ? 8 : 1; }

// LDS words of one wave's record array.
template <int S, bool VIS>
constexpr int recWords() { return kChunk * (VIS ? Rec<S>::SIZE_VIS : Rec<S>::SIZE_FULL); }

// Fine raster of tile group `group` (W tiles, one wave each) of one draw.
// COUNT: counting draw (fragment totals per tile into P.tileFrags).
// BL: blend mode known at compile time (in-order shading), or -1 for P.blend.
// TEX: the texPhong fragment shader (in-order shading; it may discard).
// recAll: the workgroup's LDS record arrays, W * recWords<S, VIS>() words.
template <int S, bool VIS, bool COUNT, int W, bool DEPTH, int BL, bool TEX = false>
__device__ __forceinline__ void fineBody(const Params& P, const int group, uint32_t* __restrict__ recAll) {
  static_assert(!(TEX && VIS), "texPhong shades in order");
  static_assert(!(COUNT && VIS), "counting draws shade in order (their statistics are per record)");
  constexpr int N = 1 << S;
  using L = Rec<S>;
  constexpr int RS = VIS ? L::SIZE_VIS : L::SIZE_FULL;
  const uint4* __restrict__ hdr = reinterpret_cast<const uint4*>(P.triHeader);
  const uint4* __restrict__ tdata = reinterpret_cast<const uint4*>(P.triData);
  const uint32_t* __restrict__ tileList = P.tileList;
  const uint8_t* __restrict__ verts = P.vertices;
  const uint32_t* __restrict__ subtriTri = P.subtriTri;
  uint32_t* __restrict__ color = P.color;
  uint32_t* __restrict__ depth = P.depth;
  uint32_t* __restrict__ triid = P.triid;
  const int32_t blend = BL >= 0 ? BL : P.blend;

  const int lane = threadIdx.x & 63;
  // Wave-uniform, so the tile's coordinates and addresses are scalars.
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  uint32_t* rec = recAll + wave * (kChunk * RS);
  const int tile = group * W + wave;
  if (tile >= P.numTiles) return;
  // The overflow flag, the tile's entry count and its first segment do not
  // depend on each other: issued together, the tile's start waits for one
  // load latency instead of a chain of three.
  const uint32_t overflowed = P.counters->overflow;
  const uint32_t cntLoaded = P.tileCount[tile];
  const uint32_t startLoaded = P.tileStart[tile];
  // (An empty asm that takes all three keeps the compiler from sinking the
  // later loads below the branches that follow.)
  __asm__ volatile("" ::"s"(overflowed), "s"(cntLoaded), "s"(startLoaded));
  if (overflowed) return;
  // Wave-uniform (one tile per wave): held in scalar registers.
  const uint32_t cnt = (uint32_t)__builtin_amdgcn_readfirstlane((int)cntLoaded);
  if (cnt == 0 && !P.deferredClear) {
    if (COUNT && lane == 0) P.tileFrags[tile] = 0;
    return;
  }

  const int lx = lane & 7, ly = lane >> 3;
  const int tx = tile % P.widthTiles, ty = tile / P.widthTiles;
  const int px = (tx << kTileLog2) + lx, py = (ty << kTileLog2) + ly;
  const size_t pitch = (size_t)P.widthPixels * N;
  const size_t texel = (size_t)py * pitch + (size_t)tx * (8 * N) + lx;
  const bool lerpOn = (P.flags & kFlagLerp) != 0;
  const bool idOn = triid != nullptr;
  const bool gouraud = P.shader == kShaderGouraud;

  uint32_t col[N], dep[N], aux[N];   // aux: VIS -> (slot << 4 | centroid) or ~0; else triangle id
  if (P.deferredClear) {
#pragma unroll
    for (int i = 0; i < N; i++) { col[i] = P.clearColor; dep[i] = P.clearDepth; aux[i] = 0xFFFFFFFFu; }
  } else {
#pragma unroll
    for (int i = 0; i < N; i++) {
      col[i] = color[texel + i * 8];
      dep[i] = depth[texel + i * 8];
      aux[i] = VIS ? 0xFFFFFFFFu : (idOn ? triid[texel + i * 8] : 0xFFFFFFFFu);
    }
  }

  // Tile origin relative to the viewport centre (header coordinates).
  const int32_t Tx = (tx << (kTileLog2 + kSubpixelLog2)) - (P.viewportWidth << (kSubpixelLog2 - 1));
  const int32_t Ty = (ty << (kTileLog2 + kSubpixelLog2)) - (P.viewportHeight << (kSubpixelLog2 - 1));
  const int32_t lx16 = lx << kSubpixelLog2, ly16 = ly << kSubpixelLog2;
  const uint32_t lxN = (uint32_t)(lx * N), lyN = (uint32_t)(ly * N);
  const uint32_t pixN = lxN | (lyN << 16);
  const uint32_t quadShift = (uint32_t)((((ly >> 2) << 1) | (lx >> 2)) * 8);   // byte of this lane's quadrant
  const uint32_t tileSx = (uint32_t)(tx * 8 * N), tileSy = (uint32_t)(ty * 8 * N);
  constexpr CentroidOrder<S> co;

  uint64_t frags = 0;
  // Profiling draws (COUNT with a profile block): records culled by
  // hierarchical Z, records walked, walked records with no covered sample,
  // covered samples, fragments whose covered samples all fail the depth
  // test, fragments written.
  uint64_t pHiz = 0, pWalked = 0, pEmpty = 0, pSamples = 0, pZKill = 0, pWritten = 0;
  // The tile's list: a chain of kTileSeg-entry segments, read one step
  // ahead -- the next step's entries and the link after them are in flight
  // while this step's are tested, so a step waits for one load latency, not
  // a chain.  The in-order copies (deep, mostly culled lists) take two
  // segments per step, 64 entries, one per lane: each half of the wave
  // follows every other segment -- lanes 0-31 the even ones, lanes 32-63
  // the odd ones -- through the links to the segment after next.  The
  // deferred-shading copies take one segment per step (their lists are
  // short, and the second half's state would not fit their registers).
  constexpr uint32_t kHalves = VIS ? 1u : 2u;
  static_assert(kChunk == kTileSeg, "one list segment per half step");
  // (The links stay in vector registers: a scalar copy would wait for each
  // load as soon as it is issued.)  The first segments are loaded without a
  // branch: segment 0 (always allocated) stands in for a missing one (an
  // empty list's start word was not written by this draw); entries past the
  // count and links past the last segment are never used.
  const uint32_t half = kHalves == 2 ? (uint32_t)lane >> 5 : 0u, hl = (uint32_t)lane & 31u;
  const uint32_t seg0 = cnt > 0 ? (uint32_t)__builtin_amdgcn_readfirstlane((int)startLoaded) : 0u;
  const uint2 link0 = *reinterpret_cast<const uint2*>(P.tileSegNext + 2 * (size_t)seg0);
  const uint32_t mySeg = half ? (cnt > (uint32_t)kChunk ? link0.x : 0u) : seg0;
  uint32_t cur = tileList[((size_t)mySeg << kTileSegLog2) + hl];
  uint32_t myNext = kHalves == 1 ? link0.x : (half ? P.tileSegNext[2 * (size_t)mySeg + 1] : link0.y);
  const bool hiz = DEPTH && P.hiz;

  // Hierarchical Z: a triangle whose conservative zmin is not nearer than
  // the tile's farthest stored depth fails LESS at every sample.  The stored
  // depths change only when records are walked, so the tile's maximum (a
  // scalar) is recomputed only after a walk.
  uint32_t tileZMax = 0xFFFFFFFFu;
  bool zStale = true;
  // Survivors of hierarchical Z are set up in batches: their slots queue up
  // (in list order) until a batch of kChunk has gathered, the list ends or
  // kMaxDefer chunks have passed, so that a tile whose entries are mostly
  // culled pays for one set-up pass per batch rather than one per chunk.
  // The depth bound a queued entry was tested against is at least the
  // current one (depths only decrease), so the test stays conservative.
  // Queue entry p lives in record p % kChunk, in a word that set-up never
  // writes (Q0 for p < kChunk, Q1 above): the queue outlives the first
  // batch's records, and up to 2 * kChunk - 1 entries fit.
  constexpr int Q0 = VIS ? L::TRI : L::SLOT, Q1 = RS - 1;
  static_assert(Q0 != Q1 && Q1 >= L::SH + (VIS ? 0 : 21), "two record words free for the queue");
  uint32_t queued = 0, sinceWalk = 0;

  for (uint32_t c = 0; c < cnt; c += kHalves * kChunk) {
    // ---- lane j loads entry c + j of the list ------------------------------
    const uint32_t n = min(kHalves * kChunk, cnt - c);
    const bool lastStep = c + kHalves * kChunk >= cnt;
    const uint32_t entry = cur;
    const bool valid = lane < (int)n && entry < (uint32_t)P.maxSubtris;
    // The header's last word: the triangle's conservative zmin (high 20
    // bits), issued before the next step's loads so that its wait does not
    // cover them, and by every lane (slot 0 stands in for an empty one): a
    // load under a branch would be waited for at the branch's end.
    const uint32_t zmin = reinterpret_cast<const uint32_t*>(hdr)[(size_t)(valid ? entry : 0u) * 4 + 3];
    // This half's next segment (entries c + kHalves * 32 + 32 * half on) and
    // the link after it, loaded by every lane without a branch (so that
    // waiting for zmin does not wait for them); segment 0 stands in past the
    // list.
    const uint32_t segSafe = c + kHalves * kChunk + kChunk * half < cnt ? myNext : 0u;
    const uint32_t nxt = tileList[((size_t)segSafe << kTileSegLog2) + hl];
    const uint32_t after = P.tileSegNext[2 * (size_t)segSafe + (kHalves - 1)];
    cur = nxt;
    myNext = after;
    // The two halves are tested and queued in list order, each followed by a
    // walk when it is due (the queue holds at most 2 * kChunk - 1 entries),
    // so the second half is tested against the depths the first one left.
    const int halves = kHalves == 2 && n > (uint32_t)kChunk ? 2 : 1;
    for (int h = 0; h < halves; h++) {
      if (hiz && zStale) {
        if (c == 0 && h == 0 && P.deferredClear) {
          tileZMax = P.clearDepth;           // every sample still holds the clear depth
        } else {
          uint32_t mx = dep[0];
#pragma unroll
          for (int i = 1; i < N; i++) mx = max(mx, dep[i]);
          tileZMax = (uint32_t)__builtin_amdgcn_readfirstlane((int)waveMax(mx));
        }
        zStale = false;
      }
      // (Bitwise, not short-circuit: the load's value is used on every path.)
      const bool keep = valid & ((int)half == h) & (!hiz | ((zmin & 0xFFFFF000u) < tileZMax));
      const uint64_t km = __ballot(keep);
      const uint32_t m = (uint32_t)__popcll(km);
      if (COUNT) {
        pHiz += min((uint32_t)kChunk, n - (uint32_t)h * kChunk) - m;
        pWalked += m;
      }
      if (keep) {
        const uint32_t p = queued + rankBelow(km);
        rec[(p & (kChunk - 1)) * RS + (p < (uint32_t)kChunk ? Q0 : Q1)] = entry;
      }
      queued += m;
      const bool lastChunk = lastStep && h == halves - 1;
      const bool walkNow = !hiz || lastChunk || queued >= (uint32_t)kChunk || ++sinceWalk >= (uint32_t)kMaxDefer;
      if (!walkNow || queued == 0) continue;
      const uint32_t total = queued;
      queued = 0;
      sinceWalk = 0;
      zStale = true;

      for (uint32_t b = 0; b < total; b += kChunk) {
        // ---- set up the batch: lane j, record j, queue entry b + j -------------
        const uint32_t nb = min((uint32_t)kChunk, total - b);
        uint32_t* r = rec + lane * RS;
        const bool su = lane < (int)nb;
        // Deferred shading: the 4x4-pixel quadrants of the tile this record
        // may cover (bit (ly >= 4) * 2 + (lx >= 4)).
        uint32_t quads = 0;
        waveSync();                            // the queue stores (or the last walk's reads) are done
        if (su) {
          const uint32_t slot = r[b == 0 ? Q0 : Q1];
          // Header and depth plane, read again here (cache hits): only the
          // set-up lanes of a batch hold them.
          const uint4 h = hdr[slot];
          const uint4 z = DEPTH ? tdata[(size_t)slot * 4] : make_uint4(0, 0, 0, 0);
          int32_t ax[3], ay[3];
          ax[0] = (int32_t)(int16_t)(h.x & 0xFFFFu) - Tx; ay[0] = (int32_t)(int16_t)(h.x >> 16) - Ty;
          ax[1] = (int32_t)(int16_t)(h.y & 0xFFFFu) - Tx; ay[1] = (int32_t)(int16_t)(h.y >> 16) - Ty;
          ax[2] = (int32_t)(int16_t)(h.z & 0xFFFFu) - Tx; ay[2] = (int32_t)(int16_t)(h.z >> 16) - Ty;
          // Vertices within 2^14 subpixels of the tile origin (all but large
          // triangles): a x d fits 31 bits, so it is two 24-bit products and
          // needs no clamp.  Uniform over the wave's set-up lanes.
          const uint32_t span = (uint32_t)(abs(ax[0]) | abs(ay[0]) | abs(ax[1]) | abs(ay[1]) | abs(ax[2]) | abs(ay[2]));
          const bool near = __ballot(span >= (1u << 14)) == 0;
          // Whole tile covered: at every sample of every pixel each edge's
          // L = lx16*dy + ly16*(-dx) (lx16, ly16 in 0..112) is at most the
          // sample's threshold, i.e. its largest value over the tile is at most
          // the smallest threshold.  Exactly the per-sample test's outcome.
          bool full = true;
          if (VIS) quads = 15u;
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const int kn = (k + 1) % 3;
            const int32_t dx = ax[kn] - ax[k], dy = ay[kn] - ay[k];
            // Edge function at the tile origin, a x d, minus 1 on exclusive
            // edges (Util.inl:100).  |a|,|d| <= 2^16.  Clamping to +-2^30
            // keeps the sign of a x d - s x d exact for every in-tile sample
            // (|s x d| < 2^25), and every threshold below fits in 32 bits.
            // With |a| < 2^14 (so |d| < 2^15), |a x d| < 2^30: the clamp never
            // applies and 32-bit arithmetic gives the same value.
            const int32_t excl = (dy > 0 || (dy == 0 && dx <= 0)) ? 1 : 0;
            int32_t ck;
            if (near) {
              ck = __mul24(ax[k], dy) - __mul24(ay[k], dx) - excl;
            } else {
              int64_t cc = (int64_t)ax[k] * dy - (int64_t)ay[k] * dx - excl;
              cc = cc > (1ll << 30) ? (1ll << 30) : (cc < -(1ll << 30) ? -(1ll << 30) : cc);
              ck = (int32_t)cc;
            }
            int32_t minR = INT32_MAX, maxR = INT32_MIN;
#pragma unroll
            for (int i = 0; i < N; i++) {
              const int32_t ox = (2 * msaaPattern(S, i) + 1) << (3 - S);
              const int32_t oy = (2 * i + 1) << (3 - S);
              // |ox|, |oy| < 16 and |dx|, |dy| < 2^17: exact 24-bit products.
              const int32_t rk = ck - (__mul24(ox, dy) - __mul24(oy, dx));
              r[L::R + k * N + i] = (uint32_t)rk;
              minR = min(minR, rk);
              maxR = max(maxR, rk);
            }
            full = full && __mul24(112, max(dy, 0)) + __mul24(112, max(-dx, 0)) <= minR;
            if (VIS) {
              // A quadrant none of whose samples can pass this edge: the
              // smallest L over its pixels (lx16 in 64*qx + {0..48}, ly16 in
              // 64*qy + {0..48}) exceeds the largest threshold.  |d| < 2^17:
              // every term below fits 25 bits.
              const int32_t ndx = -dx;
              const int32_t m00 = (dy < 0 ? __mul24(48, dy) : 0) + (ndx < 0 ? __mul24(48, ndx) : 0);
              const int32_t ex = dy * 64, ey = ndx * 64;
              quads &= (m00 <= maxR ? 1u : 0u) | (m00 + ex <= maxR ? 2u : 0u) | (m00 + ey <= maxR ? 4u : 0u) |
                       (m00 + ex + ey <= maxR ? 8u : 0u);
            }
            r[L::D + 2 * k] = (uint32_t)dy;
            r[L::D + 2 * k + 1] = (uint32_t)(-dx);
          }
          if (DEPTH) {
            r[L::Z] = (z.x & 0xFFFFu) | (z.y << 16);
            r[L::Z + 1] = (z.x >> 16) | (z.y & 0xFFFF0000u);
            // Depth at each sample of the tile's first pixel: one full 32-bit
            // product per axis, then the sample offsets (< 8) as shifts and adds.
            const uint32_t base = z.x * tileSx + z.y * tileSy + z.z;
            sampleDepths<S>(r + L::ZS, base, z.x, z.y, std::make_integer_sequence<int, N>{});
          }
          // Deferred shading records the (slot << 4 | centroid code) words the
          // walk stores per sample; at 1x the code is always the pixel centre.
          if (VIS) r[L::SLOT] = (slot << 4) | (N == 1 ? 15u : 0u);
          // In-order shading: the source triangle, bit 31 set when the triangle
          // covers the whole tile (source indices stay below 2^28).
          if (!VIS)
            r[L::TRI] = ((slot < (uint32_t)P.numTris) ? slot : subtriTri[slot - (uint32_t)P.numTris]) |
                        (full ? 0x80000000u : 0u);
          if (TEX) {
            // Barycentric planes and vertex indices; the shader reads the
            // varyings, material and textures itself.
            const uint4 t1 = tdata[(size_t)slot * 4 + 1], t2 = tdata[(size_t)slot * 4 + 2],
                        t3 = tdata[(size_t)slot * 4 + 3];
            uint32_t* q = r + L::SH;
            q[0] = t1.x; q[1] = t1.y; q[2] = t1.z; q[3] = t1.w; q[4] = t2.x; q[5] = t2.y; q[6] = t2.z;
            q[7] = t2.w; q[8] = t3.x; q[9] = t3.y; q[10] = t3.z; q[11] = t3.w;
          } else if (!VIS && gouraud) {
            const uint4 t1 = tdata[(size_t)slot * 4 + 1], t2 = tdata[(size_t)slot * 4 + 2],
                        t3 = tdata[(size_t)slot * 4 + 3];
            const float4 c0 = varying0(verts, t3.y, P.vertexStride), c1 = varying0(verts, t3.z, P.vertexStride),
                         c2 = varying0(verts, t3.w, P.vertexStride);
            uint32_t* q = r + L::SH;
            q[0] = t1.x; q[1] = t1.y; q[2] = t1.z; q[3] = t1.w; q[4] = t2.x; q[5] = t2.y; q[6] = t2.z;
            q[7] = t2.w; q[8] = t3.x;
            q[9] = __float_as_uint(c0.x); q[10] = __float_as_uint(c0.y); q[11] = __float_as_uint(c0.z);
            q[12] = __float_as_uint(c0.w); q[13] = __float_as_uint(c1.x); q[14] = __float_as_uint(c1.y);
            q[15] = __float_as_uint(c1.z); q[16] = __float_as_uint(c1.w); q[17] = __float_as_uint(c2.x);
            q[18] = __float_as_uint(c2.y); q[19] = __float_as_uint(c2.z); q[20] = __float_as_uint(c2.w);
          }
          // Quadrant queue slots start empty (this record's queue entry was
          // read above).
          if (VIS) r[Q0] = 0xFFFFFFFFu;
        }
        waveSync();

        // Deferred shading: per 4x4 quadrant, the batch's records that may
        // cover it, in submission order -- byte q of record i's word Q0 is the
        // index of quadrant q's i-th record (0xFF past its end).  Step i of the
        // walk takes the i-th record of every quadrant, so the quadrants walk
        // their own records side by side; their pixels are disjoint, so every
        // pixel still sees its triangles in submission order, and a step
        // serves up to four small triangles.
        uint32_t steps = nb;
        if (VIS) {
          steps = 0;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const uint64_t m = __ballot((quads >> q) & 1u);
            steps = max(steps, (uint32_t)__popcll(m));
            if ((quads >> q) & 1u)
              reinterpret_cast<uint8_t*>(rec + rankBelow(m) * RS + Q0)[q] = (uint8_t)lane;
          }
          waveSync();
        }

        // ---- walk the batch's triangles in submission order --------------------
        for (uint32_t j = 0; j < steps; j++) {
          // In-order shading: record j for the whole wave (uniform address).
          // Deferred shading: this lane's quadrant's j-th record; lanes whose
          // quadrant has no j-th record read record 31 and take no part.
          uint64_t validM = ~0ull;
          const uint32_t* r = rec + j * RS;
          if (VIS) {
            const uint32_t t = (r[Q0] >> quadShift) & 0xFFu;
            validM = __ballot(t < (uint32_t)kChunk);
            r = rec + min(t, (uint32_t)kChunk - 1) * RS;
          }
          const uint32_t* cur = r;
          // |lx16|,|ly16| < 2^7 and |d| < 2^17: 24-bit multiplies are exact.
          // In-order shading: a record covering the whole tile skips the
          // per-sample edge tests (a scalar branch: the flag is uniform).
          const uint32_t triWord = VIS ? 0u : (uint32_t)__builtin_amdgcn_readfirstlane((int)r[L::TRI]);
          const bool fullTile = !VIS && (triWord >> 31) != 0;
          // Per-sample coverage and write masks are kept as explicit 64-bit lane
          // masks (scalar registers): a per-lane bool carried from one block to
          // the next would be rebuilt from a vector register before each ballot.
          uint64_t covM[N];
          uint64_t anyM = 0;
          if (fullTile) {
#pragma unroll
            for (int i = 0; i < N; i++) covM[i] = ~0ull;
            anyM = ~0ull;
          } else {
            int32_t Lk[3];
#pragma unroll
            for (int k = 0; k < 3; k++)
              Lk[k] = __mul24(lx16, (int32_t)cur[L::D + 2 * k]) + __mul24(ly16, (int32_t)cur[L::D + 2 * k + 1]);
#pragma unroll
            for (int i = 0; i < N; i++) {
              covM[i] = __ballot(Lk[0] <= (int32_t)cur[L::R + i]) & __ballot(Lk[1] <= (int32_t)cur[L::R + N + i]) &
                        __ballot(Lk[2] <= (int32_t)cur[L::R + 2 * N + i]) & validM;
              anyM |= covM[i];
            }
          }
          if (COUNT) {
            pEmpty += anyM == 0;
#pragma unroll
            for (int i = 0; i < N; i++) pSamples += (uint64_t)__popcll(covM[i]);
          }
          if (anyM == 0) continue;
          if (COUNT) frags += (uint64_t)__popcll(anyM);

          uint64_t wrM[N];
          uint32_t dz[N];
          if (DEPTH) {
            // zx*lxN + zy*lyN mod 2^32 from the 16-bit halves of zx and zy, one
            // dot product per half (lxN, lyN < 64: no product or sum overflows).
            const uint32_t pl = shl16Add(dot2u(r[L::Z + 1], pixN, 0u), dot2u(r[L::Z], pixN, 0u));
            uint64_t anyWrM = 0;
#pragma unroll
            for (int i = 0; i < N; i++) {
              dz[i] = r[L::ZS + i] + pl;
              wrM[i] = covM[i] & __ballot(dz[i] < dep[i]);     // LESS (FineRaster.inl:697)
              anyWrM |= wrM[i];
            }
            if (COUNT) {
              pZKill += (uint64_t)__popcll(anyM & ~anyWrM);
              pWritten += (uint64_t)__popcll(anyWrM);
            }
            if (anyWrM == 0) continue;
          } else {
#pragma unroll
            for (int i = 0; i < N; i++) { wrM[i] = covM[i]; dz[i] = 0; }
            if (COUNT) pWritten += (uint64_t)__popcll(anyM);
          }
          bool cov[N], wr[N];
          bool anyWr = false;
#pragma unroll
          for (int i = 0; i < N; i++) {
            cov[i] = __builtin_amdgcn_inverse_ballot_w64(covM[i]);
            wr[i] = __builtin_amdgcn_inverse_ballot_w64(wrM[i]);
            anyWr = anyWr | wr[i];
          }
          // Centroid (selectMSAACentroid): first covered sample in distance
          // order, the pixel centre (15) when none or all are covered.
          int code = 15;
          if (!fullTile) {
            bool all = true;
#pragma unroll
            for (int k = N - 1; k >= 0; k--) code = cov[co.order[k]] ? co.order[k] : code;
#pragma unroll
            for (int i = 0; i < N; i++) all = all & cov[i];
            if (all) code = 15;
          }

          if (VIS) {
            const uint32_t v = N == 1 ? r[L::SLOT] : (r[L::SLOT] | (uint32_t)code);
#pragma unroll
            for (int i = 0; i < N; i++) {
              if (DEPTH) dep[i] = wr[i] ? dz[i] : dep[i];
              aux[i] = wr[i] ? v : aux[i];
            }
          } else {
            uint32_t src = P.constColor;
            const uint32_t tri = triWord & 0x7FFFFFFFu;
            bool discard = false;
            if (TEX) {
              if (anyWr) {
                const uint32_t* q = r + L::SH;
                const Planes pl = {(int32_t)q[0], (int32_t)q[1], (int32_t)q[2], (int32_t)q[3], (int32_t)q[4],
                                   (int32_t)q[5], (int32_t)q[6], (int32_t)q[7], (int32_t)q[8]};
                discard = shadeTexPhong<S>(P, pl, q[9], q[10], q[11], tri, code, px, py, &src);
              }
            } else if (anyWr && gouraud) {
              const uint32_t* q = r + L::SH;
              const uint4 t1 = make_uint4(q[0], q[1], q[2], q[3]);
              const uint4 t2 = make_uint4(q[4], q[5], q[6], q[7]);
              const uint4 t3 = make_uint4(q[8], 0, 0, 0);
              const float4 c0 = make_float4(__uint_as_float(q[9]), __uint_as_float(q[10]), __uint_as_float(q[11]),
                                            __uint_as_float(q[12]));
              const float4 c1 = make_float4(__uint_as_float(q[13]), __uint_as_float(q[14]), __uint_as_float(q[15]),
                                            __uint_as_float(q[16]));
              const float4 c2 = make_float4(__uint_as_float(q[17]), __uint_as_float(q[18]), __uint_as_float(q[19]),
                                            __uint_as_float(q[20]));
              src = shadeGouraud<S>(lerpOn, code, px, py, t1, t2, t3, c0, c1, c2);
            }
            // A discarded fragment writes neither depth nor colour.
#pragma unroll
            for (int i = 0; i < N; i++) {
              if (wr[i] && !discard) {
                if (DEPTH) dep[i] = dz[i];
                col[i] = blendPixel(blend, src, col[i]);
                aux[i] = tri;
              }
            }
          }
        }
        waveSync();
      }
    }
  }

  // The pixel's coordinates and surface offset, derived again from the lane
  // index (v_mbcnt) instead of being held in registers through the walk.
  const int laneE = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  const int pxE = (tx << kTileLog2) + (laneE & 7), pyE = (ty << kTileLog2) + (laneE >> 3);
  const size_t texelE = (size_t)pyE * pitch + (size_t)tx * (8 * N) + (laneE & 7);

  // ---- deferred shading: once per (pixel, distinct winning triangle) -------
  uint32_t tid[N];
  if (VIS) {
    // Samples that no triangle reached keep the surface's id.
#pragma unroll
    for (int i = 0; i < N; i++) {
      tid[i] = 0xFFFFFFFFu;
      if (!P.deferredClear && idOn && aux[i] == 0xFFFFFFFFu) tid[i] = triid[texelE + i * 8];
    }
    // Each pass shades the winner of this lane's first unresolved sample and
    // resolves every sample it won (aux -> ~0); a lane leaves when none is
    // left.
    for (;;) {
      uint32_t v = aux[N - 1];
#pragma unroll
      for (int i = N - 2; i >= 0; i--) v = (aux[i] != 0xFFFFFFFFu) ? aux[i] : v;
      if (v == 0xFFFFFFFFu) break;
      const uint32_t slot = v >> 4;
      const uint32_t tri = (slot < (uint32_t)P.numTris) ? slot : subtriTri[slot - (uint32_t)P.numTris];
      uint32_t src = P.constColor;
      if (gouraud) {
        const uint4* td = tdata + (size_t)slot * 4;
        const uint4 t1 = td[1], t2 = td[2], t3 = td[3];
        src = shadeGouraud<S>(lerpOn, (int)(v & 15u), pxE, pyE, t1, t2, t3, varying0(verts, t3.y, P.vertexStride),
                              varying0(verts, t3.z, P.vertexStride), varying0(verts, t3.w, P.vertexStride));
      }
#pragma unroll
      for (int i = 0; i < N; i++) {
        const bool won = aux[i] == v;
        if (blend != kBlendDepthOnly) col[i] = won ? src : col[i];
        tid[i] = won ? tri : tid[i];
        aux[i] = won ? 0xFFFFFFFFu : aux[i];
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < N; i++) tid[i] = aux[i];
  }

  const bool storeColor = blend != kBlendDepthOnly || P.deferredClear;
#pragma unroll
  for (int i = 0; i < N; i++) {
    if (storeColor) color[texelE + i * 8] = col[i];
    depth[texelE + i * 8] = dep[i];
    if (idOn) triid[texelE + i * 8] = tid[i];
  }
  // Fragment statistics only on counting draws: one plain store per tile
  // (a device-wide atomic per tile serialises across the XCDs).
  if (COUNT && lane == 0) P.tileFrags[tile] = (uint32_t)frags;
  if (COUNT && P.prof != nullptr && lane == 0 && cnt > 0) {
    profAdd(P.prof, kProfSetupSamplesPerTri, pSamples, 0);
    profAdd(P.prof, kProfFineTriPerTile, cnt, 1);
    profAdd(P.prof, kProfFineEarlyZCull, 100 * pHiz, cnt);
    profAdd(P.prof, kProfFineEmptyCull, 100 * pEmpty, pWalked);
    profAdd(P.prof, kProfFineZKill, 100 * pZKill, frags);
    profAdd(P.prof, kProfFineWarpUtil, 100 * frags, 64 * pWalked);
    profAdd(P.prof, kProfFineBlendRounds, pWritten, pWritten);
    profAdd(P.prof, kProfFineFragPerTri, frags, pWalked);
    profAdd(P.prof, kProfFineFragPerTile, frags, 1);
  }
}

// In-order shading blends in the per-triangle loop: SrcOver and Additive
// get their own copy of the body with the blend mode folded in (the choice
// is uniform per draw); counting draws and the rare in-order Replace /
// DepthOnly case use the generic copy.
template <int S, bool VIS, bool COUNT, int W, bool DEPTH>
__device__ __forceinline__ void fineDispatch(const Params& P, const int group, uint32_t* recAll) {
  if (!VIS && !COUNT) {
    if (P.blend == kBlendSrcOver) return fineBody<S, VIS, COUNT, W, DEPTH, kBlendSrcOver>(P, group, recAll);
    if (P.blend == kBlendAdditive) return fineBody<S, VIS, COUNT, W, DEPTH, kBlendAdditive>(P, group, recAll);
  }
  fineBody<S, VIS, COUNT, W, DEPTH, -1>(P, group, recAll);
}

template <int S, bool VIS, bool COUNT, bool DEPTH>
__global__ void __launch_bounds__(64 * kFineWaves, (fineMinWaves<S, VIS>())) fineKernel(const Params P) {
  __shared__ __attribute__((aligned(16))) uint32_t recAll[kFineWaves * recWords<S, VIS>()];
  const int groups = (P.numTiles + kFineWaves - 1) / kFineWaves;
  const int group = xcdGroup(groups);
  if (group >= groups) return;
  fineDispatch<S, VIS, COUNT, kFineWaves, DEPTH>(P, group, recAll);
}

// The texPhong shader: in-order shading with the shader compiled in, blend
// mode from the draw.
template <int S, bool COUNT, bool DEPTH>
__global__ void __launch_bounds__(64 * kFineWaves) fineTexKernel(const Params P) {
  __shared__ __attribute__((aligned(16))) uint32_t recAll[kFineWaves * recWords<S, false>()];
  const int groups = (P.numTiles + kFineWaves - 1) / kFineWaves;
  const int group = xcdGroup(groups);
  if (group >= groups) return;
  fineBody<S, false, COUNT, kFineWaves, DEPTH, -1, true>(P, group, recAll);
}

template <int S, bool COUNT, bool DEPTH>
static void launchFineTex(const Params& P, hipStream_t stream) {
  const int groups = (P.numTiles + kFineWaves - 1) / kFineWaves;
  const int per = (groups + 7) >> 3;
  hipLaunchKernelGGL((fineTexKernel<S, COUNT, DEPTH>), dim3(per * 8), dim3(64 * kFineWaves), 0, stream, P);
}

// Batched draws: the tile groups of the views of one pipe variant form one
// grid; the k-th listed view is views[order[k]] and owns groups
// [groupStart[k], groupStart[k+1]).  banded: XCD-aware bands (as a plain
// draw); otherwise groups are dealt round-robin over the XCDs in list order,
// which spreads the first-listed (heaviest) groups over all of them.
template <int S, bool VIS, int W, bool DEPTH>
__global__ void __launch_bounds__(64 * W, (fineMinWaves<S, VIS>()))
    fineBatchKernel(const Params* __restrict__ views, const int32_t* __restrict__ groupStart,
                    const int32_t* __restrict__ order, int numViews, int totalGroups, int banded) {
  __shared__ __attribute__((aligned(16))) uint32_t recAll[W * recWords<S, VIS>()];
  const int g = banded ? xcdGroup(totalGroups) : (int)blockIdx.x;
  if (g >= totalGroups) return;
  const int k = findView(groupStart, numViews, g);
  const Params& P = views[order[k]];
  const int first = (g - groupStart[k]) * kFineRounds;
  for (int t = 0; t < kFineRounds; t++) fineDispatch<S, VIS, false, W, DEPTH>(P, first + t, recAll);
}

// Variant v = (samplesLog2 * 2 + deferred shading) * 2 + depth test.
template <int V>
struct FineVariant {
  static constexpr int S = V >> 2;
  static constexpr bool VIS = ((V >> 1) & 1) != 0;
  static constexpr bool DEPTH = (V & 1) != 0;
};

template <int V, bool COUNT>
static void launchFineV(const Params& P, hipStream_t stream) {
  using FV = FineVariant<V>;
  const int groups = (P.numTiles + kFineWaves - 1) / kFineWaves;
  const int per = (groups + 7) >> 3;
  hipLaunchKernelGGL((fineKernel<FV::S, FV::VIS, COUNT, FV::DEPTH>), dim3(per * 8), dim3(64 * kFineWaves), 0,
                     stream, P);
}

template <int V, int W>
static void launchFineBatchVW(const Params* views, const int32_t* groupStart, const int32_t* order, int numViews,
                              int totalGroups, int banded, hipStream_t stream) {
  using FV = FineVariant<V>;
  const int per = (totalGroups + 7) >> 3;
  hipLaunchKernelGGL((fineBatchKernel<FV::S, FV::VIS, W, FV::DEPTH>), dim3(banded ? per * 8 : totalGroups),
                     dim3(64 * W), 0, stream, views, groupStart, order, numViews, totalGroups, banded);
}

using FineLaunch = void (*)(const Params&, hipStream_t);
using FineBatchLaunch = void (*)(const Params*, const int32_t*, const int32_t*, int, int, int, hipStream_t);

template <int... V>
struct FineTables {
  static constexpr FineLaunch plain[] = {&launchFineV<V, false>...};
  static constexpr FineBatchLaunch batch2[] = {&launchFineBatchVW<V, 2>...};
  static constexpr FineBatchLaunch batch4[] = {&launchFineBatchVW<V, 4>...};
  static constexpr FineBatchLaunch batch8[] = {&launchFineBatchVW<V, 8>...};
};
using AllFine = FineTables<0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15>;
static_assert(kFineVariants == 16, "variant table");

int fineVariant(const Params& P) {
  const bool vis = (P.blend == kBlendReplace || P.blend == kBlendDepthOnly) && P.maxSubtris < (1 << 28) &&
                   P.shader != kShaderTexPhong;
  const bool depth = (P.flags & kFlagDepth) != 0;
  return (P.samplesLog2 * 2 + (vis ? 1 : 0)) * 2 + (depth ? 1 : 0);
}

void launchFineBatch(int variant, int tilesPerGroup, int banded, const Params* views, const int32_t* groupStart,
                     const int32_t* order, int numViews, int totalGroups, hipStream_t stream) {
  if (totalGroups <= 0 || numViews <= 0 || variant < 0 || variant >= kFineVariants) return;
  const FineBatchLaunch* t = tilesPerGroup == 2 ? AllFine::batch2 : (tilesPerGroup == 8 ? AllFine::batch8
                                                                                         : AllFine::batch4);
  t[variant](views, groupStart, order, numViews, totalGroups, banded, stream);
}

// Host-side launcher.  Deferred shading needs the visibility word
// (slot << 4 | centroid) to hold every slot: slots < 2^28.
void launchFine(const Params& P, hipStream_t stream) {
  if (P.shader == kShaderTexPhong) {
    using L = void (*)(const Params&, hipStream_t);
    static constexpr L table[16] = {
        &launchFineTex<0, false, false>, &launchFineTex<0, false, true>, &launchFineTex<0, true, false>,
        &launchFineTex<0, true, true>,   &launchFineTex<1, false, false>, &launchFineTex<1, false, true>,
        &launchFineTex<1, true, false>,  &launchFineTex<1, true, true>,   &launchFineTex<2, false, false>,
        &launchFineTex<2, false, true>,  &launchFineTex<2, true, false>,  &launchFineTex<2, true, true>,
        &launchFineTex<3, false, false>, &launchFineTex<3, false, true>,  &launchFineTex<3, true, false>,
        &launchFineTex<3, true, true>};
    const int count = P.tileFrags != nullptr ? 1 : 0, depth = (P.flags & kFlagDepth) ? 1 : 0;
    table[P.samplesLog2 * 4 + count * 2 + depth](P, stream);
    return;
  }
  if (P.tileFrags != nullptr) {
    // Counting draws run the in-order variant with per-tile totals.
    const bool depth = (P.flags & kFlagDepth) != 0;
    switch (P.samplesLog2 * 2 + (depth ? 1 : 0)) {
      case 0: launchFineV<0, true>(P, stream); break;
      case 1: launchFineV<1, true>(P, stream); break;
      case 2: launchFineV<4, true>(P, stream); break;
      case 3: launchFineV<5, true>(P, stream); break;
      case 4: launchFineV<8, true>(P, stream); break;
      case 5: launchFineV<9, true>(P, stream); break;
      case 6: launchFineV<12, true>(P, stream); break;
      default: launchFineV<13, true>(P, stream); break;
    }
    return;
  }
  AllFine::plain[fineVariant(P)](P, stream);
}

// Sums a u32 array into a u64 counter (one workgroup; counting draws only).
__global__ void __launch_bounds__(1024) sumCountsKernel(const uint32_t* v, int n, uint64_t* out) {
  __shared__ unsigned long long part[16];
  unsigned long long acc = 0;
  for (int i = threadIdx.x; i < n; i += 1024) acc += v[i];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < 16; w++) t += part[w];
    *out = t;
  }
}

void launchFragmentSum(const Params& P, hipStream_t stream) {
  if (P.tileFrags != nullptr)
    hipLaunchKernelGGL(sumCountsKernel, dim3(1), dim3(1024), 0, stream, (const uint32_t*)P.tileFrags, P.numTiles,
                       &P.counters->fragments);
  if (P.chunkSubtris != nullptr && P.numChunks > 0)
    hipLaunchKernelGGL(sumCountsKernel, dim3(1), dim3(1024), 0, stream, (const uint32_t*)P.chunkSubtris,
                       P.numChunks, &P.counters->visibleSubtris);
}

}  // namespace crhip
