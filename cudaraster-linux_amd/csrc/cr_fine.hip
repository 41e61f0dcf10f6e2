// Stage 4: fine raster, one wave64 per 8x8 tile, one lane per pixel.
//
// Same contract as fineRasterImpl_SingleSample / _MultiSample
// (cudaraster/cuda/FineRaster.inl:499-752, 855-1126): for every triangle of
// the tile, in submission order, exact sample coverage (top-left rule), LESS
// depth test, the fragment shader at the pixel centre / coverage centroid,
// and the blend ROP.  The wave64 design differs from the reference's
// fragment-distribution warp: each lane owns one pixel and keeps all of its
// samples (colour, depth, triangle id) in registers for the whole tile, so
// fragments of the same pixel are applied in order by construction -- no
// intra-warp conflict resolution, no read-modify-write of the surface per
// sample, one coalesced read and one coalesced write of the tile per draw.
// The triangle is wave-uniform (scalar loads); coverage is three integer edge
// tests per sample in 32-bit arithmetic (see coverage bounds below); the
// 64-bit ballot of per-pixel coverage gives empty-triangle rejection.
#include "cr_device.h"

namespace crhip {

__device__ __forceinline__ uint32_t waveMax(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

template <int S>
__device__ __forceinline__ int centroidSample(uint32_t mask) {
  return selectCentroid(S, mask);
}

template <int S>
__global__ void __launch_bounds__(64) fineKernel(const Params P, const uint4* __restrict__ hdr,
                                                 const uint4* __restrict__ tdata,
                                                 const uint32_t* __restrict__ tileList,
                                                 const uint8_t* __restrict__ verts,
                                                 const uint32_t* __restrict__ subtriTri,
                                                 uint32_t* __restrict__ color, uint32_t* __restrict__ depth,
                                                 uint32_t* __restrict__ triid) {
  constexpr int N = 1 << S;
  if (P.counters->overflow) return;
  // XCD-aware order: consecutive workgroups go round-robin to the 8 XCDs;
  // give each XCD one contiguous band of tiles so that triangles shared by
  // neighbouring tiles stay in one L2.
  const int per = (P.numTiles + 7) >> 3;
  const int tile = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (tile >= P.numTiles) return;
  const uint32_t cnt = P.tileCount[tile];
  if (cnt == 0 && !P.deferredClear) return;

  const int lane = threadIdx.x;
  const int lx = lane & 7, ly = lane >> 3;
  const int tx = tile % P.widthTiles, ty = tile / P.widthTiles;
  const int px = (tx << kTileLog2) + lx, py = (ty << kTileLog2) + ly;
  const size_t pitch = (size_t)P.widthPixels * N;
  const size_t texel = (size_t)py * pitch + (size_t)tx * (8 * N) + lx;
  const bool depthOn = (P.flags & kFlagDepth) != 0;
  const bool lerpOn = (P.flags & kFlagLerp) != 0;
  const bool idOn = triid != nullptr;

  uint32_t col[N], dep[N], tid[N];
  if (P.deferredClear) {
#pragma unroll
    for (int i = 0; i < N; i++) { col[i] = P.clearColor; dep[i] = P.clearDepth; tid[i] = 0xFFFFFFFFu; }
  } else {
#pragma unroll
    for (int i = 0; i < N; i++) {
      col[i] = color[texel + i * 8];
      dep[i] = depth[texel + i * 8];
      tid[i] = idOn ? triid[texel + i * 8] : 0xFFFFFFFFu;
    }
  }

  // Tile origin relative to the viewport centre (header coordinates).
  const int32_t Tx = (tx << (kTileLog2 + kSubpixelLog2)) - (P.viewportWidth << (kSubpixelLog2 - 1));
  const int32_t Ty = (ty << (kTileLog2 + kSubpixelLog2)) - (P.viewportHeight << (kSubpixelLog2 - 1));
  const int32_t lx16 = lx << kSubpixelLog2, ly16 = ly << kSubpixelLog2;

  uint32_t tileZMax = 0;
  bool zdirty = true;
  uint64_t frags = 0;
  const uint32_t start = P.tileStart[tile];

  for (uint32_t c = 0; c < cnt; c += 64) {
    const uint32_t n = min(64u, cnt - c);
    const uint32_t mySlot = (lane < (int)n) ? tileList[start + c + lane] : 0u;
    for (uint32_t j = 0; j < n; j++) {
      const uint32_t slot = __builtin_amdgcn_readlane(mySlot, j);
      if (slot >= (uint32_t)P.maxSubtris) continue;
      const uint4 h = hdr[slot];

      // Hierarchical Z: conservative tile-level rejection (FineRaster.inl:229-240).
      if (depthOn && P.hiz) {
        if (zdirty) {
          uint32_t m = dep[0];
#pragma unroll
          for (int i = 1; i < N; i++) m = max(m, dep[i]);
          tileZMax = waveMax(m);
          zdirty = false;
        }
        if ((h.w & 0xFFFFF000u) >= tileZMax) continue;
      }

      // Edge functions in tile-relative subpixels.  |a|,|d| <= 2^16; the
      // constant term is clamped to +-2^30, the lane term |L| < 2^24 and the
      // sample term |o| < 2^21, so C - L - o keeps the exact sign in 32 bits.
      int32_t ax[3], ay[3];
      ax[0] = (int32_t)(int16_t)(h.x & 0xFFFFu) - Tx; ay[0] = (int32_t)(int16_t)(h.x >> 16) - Ty;
      ax[1] = (int32_t)(int16_t)(h.y & 0xFFFFu) - Tx; ay[1] = (int32_t)(int16_t)(h.y >> 16) - Ty;
      ax[2] = (int32_t)(int16_t)(h.z & 0xFFFFu) - Tx; ay[2] = (int32_t)(int16_t)(h.z >> 16) - Ty;
      int32_t T[3], dxk[3], dyk[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int kn = (k + 1) % 3;
        const int32_t dx = ax[kn] - ax[k], dy = ay[kn] - ay[k];
        int64_t cc = (int64_t)ax[k] * dy - (int64_t)ay[k] * dx;
        if (dy > 0 || (dy == 0 && dx <= 0)) cc -= 1;      // exclusive edge (Util.inl:100)
        cc = cc > (1ll << 30) ? (1ll << 30) : (cc < -(1ll << 30) ? -(1ll << 30) : cc);
        T[k] = (int32_t)cc - (lx16 * dy - ly16 * dx);
        dxk[k] = dx;
        dyk[k] = dy;
      }
      uint32_t cover = 0;
#pragma unroll
      for (int i = 0; i < N; i++) {
        const int32_t ox = (2 * msaaPattern(S, i) + 1) << (3 - S);
        const int32_t oy = (2 * i + 1) << (3 - S);
        bool in = true;
#pragma unroll
        for (int k = 0; k < 3; k++) in = in && (T[k] >= ox * dyk[k] - oy * dxk[k]);
        cover |= in ? (1u << i) : 0u;
      }
      const uint64_t covBallot = __ballot(cover != 0);
      if (covBallot == 0) continue;
      frags += (uint64_t)__popcll(covBallot);

      const uint4* td = tdata + (size_t)slot * 4;
      uint32_t write = cover;
      uint32_t dz[N];
      if (depthOn) {
        const uint4 z = td[0];
        const uint32_t pl = z.x * (uint32_t)(px * N) + z.y * (uint32_t)(py * N) + z.z;
#pragma unroll
        for (int i = 0; i < N; i++) {
          dz[i] = pl + z.x * (uint32_t)msaaPattern(S, i) + z.y * (uint32_t)i;
          if (dz[i] >= dep[i]) write &= ~(1u << i);        // LESS (FineRaster.inl:697)
        }
      }
      if (__ballot(write != 0) == 0) continue;

      const uint32_t src_tri = (slot < (uint32_t)P.numTris) ? slot : subtriTri[slot - (uint32_t)P.numTris];
      uint32_t src = P.constColor;
      if (write != 0 && P.shader == kShaderGouraud) {
        const uint4 t3 = td[3];
        float b0, b1, b2;
        if (lerpOn) {
          // computeBarys at the centroid (FineRaster.inl:21-44, 93-114).
          const uint4 t1 = td[1], t2 = td[2];
          const int ctr = centroidSample<S>(cover);
          const int cx = (ctr < 0) ? N : 2 * msaaPattern(S, ctr) + 1;
          const int cy = (ctr < 0) ? N : 2 * ctr + 1;
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
        const float4 c0 = *reinterpret_cast<const float4*>(verts + (size_t)t3.y * P.vertexStride + 16);
        const float4 c1 = *reinterpret_cast<const float4*>(verts + (size_t)t3.z * P.vertexStride + 16);
        const float4 c2 = *reinterpret_cast<const float4*>(verts + (size_t)t3.w * P.vertexStride + 16);
        // interpolateVarying (PixelPipe.inl:52-59) + toABGR (PixelPipe.inl:68).
        src = toABGR(fmaf(c2.x, b2, fmaf(c0.x, b0, c1.x * b1)), fmaf(c2.y, b2, fmaf(c0.y, b0, c1.y * b1)),
                     fmaf(c2.z, b2, fmaf(c0.z, b0, c1.z * b1)), fmaf(c2.w, b2, fmaf(c0.w, b0, c1.w * b1)));
      }
#pragma unroll
      for (int i = 0; i < N; i++) {
        if (write & (1u << i)) {
          if (depthOn) dep[i] = dz[i];
          col[i] = blendPixel(P.blend, src, col[i]);
          tid[i] = src_tri;
        }
      }
      if (depthOn) zdirty = true;
    }
  }

#pragma unroll
  for (int i = 0; i < N; i++) {
    color[texel + i * 8] = col[i];
    depth[texel + i * 8] = dep[i];
    if (idOn) triid[texel + i * 8] = tid[i];
  }
  if (lane == 0 && frags) atomicAdd((unsigned long long*)&P.counters->fragments, (unsigned long long)frags);
}

// Host-side launcher (one instantiation per sample count).
void launchFine(const Params& P, hipStream_t stream) {
  const int per = (P.numTiles + 7) >> 3;
  dim3 grid(per * 8), block(64);
  const uint4* hdr = reinterpret_cast<const uint4*>(P.triHeader);
  const uint4* td = reinterpret_cast<const uint4*>(P.triData);
  switch (P.samplesLog2) {
    case 0: hipLaunchKernelGGL(fineKernel<0>, grid, block, 0, stream, P, hdr, td, P.tileList, P.vertices, P.subtriTri, P.color, P.depth, P.triid); break;
    case 1: hipLaunchKernelGGL(fineKernel<1>, grid, block, 0, stream, P, hdr, td, P.tileList, P.vertices, P.subtriTri, P.color, P.depth, P.triid); break;
    case 2: hipLaunchKernelGGL(fineKernel<2>, grid, block, 0, stream, P, hdr, td, P.tileList, P.vertices, P.subtriTri, P.color, P.depth, P.triid); break;
    default: hipLaunchKernelGGL(fineKernel<3>, grid, block, 0, stream, P, hdr, td, P.tileList, P.vertices, P.subtriTri, P.color, P.depth, P.triid); break;
  }
}

}  // namespace crhip
