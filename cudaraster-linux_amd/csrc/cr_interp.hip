// Differentiable attribute interpolation over a rasterized frame.
//
// Forward: for every pixel of a single-sample triangle-id surface (the
// per-sample index of the last triangle written, 0xFFFFFFFF where none),
// interpolate C float channels of the triangle's three vertices with
// perspective-correct barycentrics at the pixel centre.  Backward: given
// dL/dout, accumulate dL/dattr and dL/dpos (clip x, y, w) of those vertices,
// with visibility fixed by the forward triangle ids (interior barycentric
// gradients only; no edge or occlusion terms).
//
// Barycentrics from homogeneous 2D coordinates: with q = (qx, qy, 1) the
// pixel centre in NDC and v_k = (x_k, y_k, w_k), the point of the triangle
// that projects to q has barycentrics b_k = c_k / (c_0 + c_1 + c_2),
// c_k = (v_{k+1} x v_{k+2}) . q.  Shearing by x' = x - qx w, y' = y - qy w
// (determinant 1) moves q to (0, 0, 1) and leaves
// c_k = x'_{k+1} y'_{k+2} - y'_{k+1} x'_{k+2}.  All of it is evaluated in
// FP64 (cheap on CDNA4 and well conditioned for sub-pixel triangles); the
// backward sums each triangle's per-pixel gradient terms in FP64 and adds
// the sums to the FP32 outputs with atomics.
#include <algorithm>

#include "cr_device.h"

namespace crhip {

struct Bary {
  double b[3];
  double xs[3], ys[3];          // sheared vertex coordinates x', y'
  double qx, qy, sum;
  bool ok;                      // false for a degenerate projected triangle (sum == 0)
};

__device__ __forceinline__ void baryAt(const float4* __restrict__ pos, const int32_t i[3], double qx, double qy,
                                       Bary& B) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float4 p = pos[i[k]];
    B.xs[k] = (double)p.x - qx * (double)p.w;
    B.ys[k] = (double)p.y - qy * (double)p.w;
  }
  double c[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
    c[k] = B.xs[k1] * B.ys[k2] - B.ys[k1] * B.xs[k2];
  }
  B.sum = c[0] + c[1] + c[2];
  B.ok = B.sum != 0.0;
  const double r = B.ok ? 1.0 / B.sum : 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) B.b[k] = c[k] * r;
  B.qx = qx;
  B.qy = qy;
}

// Looks up pixel (x, y) of the id surface; fills the vertex indices.
// Returns false for background or out-of-range ids / indices.
__device__ __forceinline__ bool pixelTriangle(const InterpGeom& G, const uint32_t* __restrict__ ids,
                                              const int3* __restrict__ tris, int x, int y, uint32_t& t,
                                              int32_t i[3]) {
  t = ids[(size_t)y * G.texW + x];
  if (t >= (uint32_t)G.numTris) return false;
  const int3 v = tris[t];
  i[0] = v.x; i[1] = v.y; i[2] = v.z;
  for (int k = 0; k < 3; k++)
    if (i[k] < 0 || (int64_t)i[k] >= G.numVertices) return false;
  return true;
}

// One lane per pixel, one grid row per image row.  vec4: C % 4 == 0 and
// attr / out 16-byte aligned, so channels move as float4.
__global__ void __launch_bounds__(256) interpForwardKernel(const InterpGeom G, const uint32_t* __restrict__ ids,
                                                           const float4* __restrict__ pos,
                                                           const float* __restrict__ attr,
                                                           const int3* __restrict__ tris, float* __restrict__ out,
                                                           int32_t* __restrict__ triImage, int vec4) {
  const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int y = (int)blockIdx.y;
  if (x >= G.width) return;
  const int64_t p = (int64_t)y * G.width + x;
  const int C = G.numChannels;
  float* o = out + p * C;
  uint32_t t;
  int32_t i[3];
  if (!pixelTriangle(G, ids, tris, x, y, t, i)) {
    if (vec4) {
      for (int c = 0; c < C; c += 4) *reinterpret_cast<float4*>(o + c) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
      for (int c = 0; c < C; c++) o[c] = 0.0f;
    }
    if (triImage) triImage[p] = -1;
    return;
  }
  const double qx = (2.0 * x + 1.0) / G.width - 1.0, qy = (2.0 * y + 1.0) / G.height - 1.0;
  Bary B;
  baryAt(pos, i, qx, qy, B);
  const float* a0 = attr + (int64_t)i[0] * C;
  const float* a1 = attr + (int64_t)i[1] * C;
  const float* a2 = attr + (int64_t)i[2] * C;
  if (vec4) {
    for (int c = 0; c < C; c += 4) {
      const float4 u = *reinterpret_cast<const float4*>(a0 + c), v = *reinterpret_cast<const float4*>(a1 + c),
                   w = *reinterpret_cast<const float4*>(a2 + c);
      *reinterpret_cast<float4*>(o + c) = make_float4((float)(B.b[0] * u.x + B.b[1] * v.x + B.b[2] * w.x),
                                                      (float)(B.b[0] * u.y + B.b[1] * v.y + B.b[2] * w.y),
                                                      (float)(B.b[0] * u.z + B.b[1] * v.z + B.b[2] * w.z),
                                                      (float)(B.b[0] * u.w + B.b[1] * v.w + B.b[2] * w.w));
    }
  } else {
    for (int c = 0; c < C; c++) o[c] = (float)(B.b[0] * a0[c] + B.b[1] * a1[c] + B.b[2] * a2[c]);
  }
  if (triImage) triImage[p] = (int32_t)t;
}

// Backward over 16x16-pixel tiles, one 256-lane workgroup each.  The tile's
// pixels are first sorted by triangle id in LDS (bitonic, 64-bit keys: id,
// then pixel), so every triangle's pixels of the tile are consecutive.
// Phase 1, one lane per sorted pixel: barycentrics, dL/db_k = sum_c g_c
// a_k,c, dL/dc_j = (dL/db_j - sum_k b_k dL/db_k) / (c_0 + c_1 + c_2), then
// through c_k = x'_{k+1} y'_{k+2} - y'_{k+1} x'_{k+2} and the shear
// x' = x - qx w, y' = y - qy w to the clip x, y, w of each corner; the
// results are staged in LDS.  The pixels then form runs of one triangle
// each; phases 2 and 3 sum every (run, corner, channel) and (run, corner,
// x/y/w) entry over the run's pixels in FP64 and add it to the output
// gradient with one FP32 atomic per entry, consecutive lanes on consecutive
// entries -- consecutive addresses within a run, which device-scope atomics
// serve several times faster than scattered ones.  A vertex receives one
// addition per (tile, triangle) run that touches it, in whatever order the
// atomics land, so the last bits of a gradient can differ between runs.
constexpr int kBwdTile = 16;
constexpr int kBwdThreads = kBwdTile * kBwdTile;
constexpr int kBwdStageC = 8;                  // channels of dL/dout staged in LDS per pixel

// Bit-reproducible accumulation: a value is a 128-bit two's-complement
// fixed-point number with 2^-64 resolution, two 64-bit words (low, high).
// Each FP64 sum is converted once (magnitude truncated to the resolution,
// |x| < 2^63) and added with integer atomics: the low word, then the high
// word with the low word's carry (only when that is not zero).  Integer
// addition is associative, so the total does not depend on the order the
// atomics land in.
__device__ __forceinline__ void fixedAdd(unsigned long long* __restrict__ acc, int64_t idx, double s) {
  const bool neg = s < 0.0;
  double m = fmin(fabs(s) * 18446744073709551616.0, 1.7014118346046923e38);   // |s| 2^64, below 2^127
  const double mh = floor(m * 5.421010862427522e-20);                          // m / 2^64
  const double ml = m - mh * 18446744073709551616.0;                           // exact
  uint64_t lo = (uint64_t)ml, hi = (uint64_t)mh;
  if (neg) {
    lo = ~lo + 1u;
    hi = ~hi + (lo == 0u ? 1u : 0u);
  }
  unsigned long long* w = acc + 2 * idx;
  const unsigned long long old = atomicAdd(w, (unsigned long long)lo);
  const uint64_t carry = (old + lo) < old ? 1u : 0u;
  const uint64_t h = hi + carry;
  if (h != 0u) atomicAdd(w + 1, (unsigned long long)h);
}

// DET: accumulate into the fixed-point words `acc` (positions V*4 values,
// then attributes V*C) instead of FP32 atomics on the outputs.
template <bool DET>
__global__ void __launch_bounds__(kBwdThreads) interpBackwardKernel(const InterpGeom G,
                                                                    const int32_t* __restrict__ triImage,
                                                                    const float4* __restrict__ pos,
                                                                    const float* __restrict__ attr,
                                                                    const int3* __restrict__ tris,
                                                                    const float* __restrict__ gradOut,
                                                                    float* __restrict__ gradPos,
                                                                    float* __restrict__ gradAttr,
                                                                    unsigned long long* __restrict__ acc) {
  __shared__ double sB[kBwdThreads][3];        // barycentrics
  __shared__ double sP[kBwdThreads][9];        // dL/d(x, y, w) per corner
  __shared__ int32_t sV[kBwdThreads][3];       // vertex indices
  __shared__ int32_t sT[kBwdThreads];          // triangle id, -1: no pixel
  __shared__ uint32_t sPix[kBwdThreads];
  __shared__ uint16_t sRun[kBwdThreads + 1];   // first pixel of each run, then the end
  __shared__ uint32_t sWaveRuns[kBwdThreads / 64];
  __shared__ uint64_t sKey[kBwdThreads];
  __shared__ float sG[kBwdThreads][kBwdStageC];   // dL/dout of the first channels, per pixel
  const int tid = threadIdx.x;
  const int C = G.numChannels;

  // ---- sort the tile's pixels by triangle id ---------------------------------
  {
    const int x = (int)blockIdx.x * kBwdTile + (tid & (kBwdTile - 1));
    const int y = (int)blockIdx.y * kBwdTile + tid / kBwdTile;
    uint32_t key = 0xFFFFFFFFu;                // no triangle: sorts last
    if (x < G.width && y < G.height) {
      const int32_t t = triImage[(int64_t)y * G.width + x];
      if (t >= 0 && t < G.numTris) key = (uint32_t)t;
    }
    sKey[tid] = ((uint64_t)key << 32) | (uint32_t)((int64_t)y * G.width + x);
  }
  __syncthreads();
  for (int k = 2; k <= kBwdThreads; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      const int o = tid ^ jj;
      if (o > tid) {
        const uint64_t a = sKey[tid], b = sKey[o];
        if ((a > b) == ((tid & k) == 0)) {
          sKey[tid] = b;
          sKey[o] = a;
        }
      }
      __syncthreads();
    }
  }

  // ---- phase 1: one lane per sorted pixel ------------------------------------
  int32_t myTri = -1;
  const uint64_t myKey = sKey[tid];
  if ((uint32_t)(myKey >> 32) != 0xFFFFFFFFu) {
    const int64_t p = (int64_t)(uint32_t)myKey;
    const int32_t t = (int32_t)(myKey >> 32);
    const int3 v = tris[t];
    const int32_t i[3] = {v.x, v.y, v.z};
    bool ok = true;
    for (int k = 0; k < 3; k++) ok = ok && i[k] >= 0 && (int64_t)i[k] < G.numVertices;
    const int x = (int)(p % G.width), y = (int)(p / G.width);
    const double qx = (2.0 * x + 1.0) / G.width - 1.0, qy = (2.0 * y + 1.0) / G.height - 1.0;
    Bary B;
    if (ok) {
      baryAt(pos, i, qx, qy, B);
      ok = B.ok;
    }
    if (ok) {
      const float* g = gradOut + p * C;
      for (int c = 0; c < min(C, kBwdStageC); c++) sG[tid][c] = g[c];
      double dB[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < 3; k++) {
        const float* a = attr + (int64_t)i[k] * C;
        double s = 0.0;
        for (int c = 0; c < C; c++) s += (double)g[c] * (double)a[c];
        dB[k] = s;
      }
      const double mean = B.b[0] * dB[0] + B.b[1] * dB[1] + B.b[2] * dB[2];
      const double rs = 1.0 / B.sum;
      double dxs[3] = {0.0, 0.0, 0.0}, dys[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double dc = (dB[k] - mean) * rs;
        const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
        dxs[k1] += dc * B.ys[k2];
        dys[k1] -= dc * B.xs[k2];
        dxs[k2] -= dc * B.ys[k1];
        dys[k2] += dc * B.xs[k1];
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        sB[tid][k] = B.b[k];
        sV[tid][k] = i[k];
        sP[tid][3 * k] = dxs[k];
        sP[tid][3 * k + 1] = dys[k];
        sP[tid][3 * k + 2] = -B.qx * dxs[k] - B.qy * dys[k];
      }
      sPix[tid] = (uint32_t)p;
      myTri = t;
    }
  }
  sT[tid] = myTri;
  __syncthreads();

  // ---- runs: a pixel starts one where its triangle differs from the
  // previous pixel's; run r covers [sRun[r], sRun[r + 1]) minus any trailing
  // pixels without a triangle (sT < 0), which are skipped.
  const bool head = myTri >= 0 && (tid == 0 || sT[tid - 1] != myTri);
  const uint64_t hm = __ballot(head);
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) sWaveRuns[wave] = (uint32_t)__popcll(hm);
  __syncthreads();
  uint32_t before = 0, numRuns = 0;
#pragma unroll
  for (int w = 0; w < kBwdThreads / 64; w++) {
    before += (w < wave) ? sWaveRuns[w] : 0u;
    numRuns += sWaveRuns[w];
  }
  if (head) sRun[before + (uint32_t)__popcll(hm & ((1ull << lane) - 1ull))] = (uint16_t)tid;
  if (tid == 0) sRun[numRuns] = (uint16_t)kBwdThreads;
  __syncthreads();

  // ---- phase 2: attribute gradients sum_q b_k g_c per (run, corner, channel)
  const int per = 3 * C;
  for (int e = tid; e < (int)numRuns * per; e += kBwdThreads) {
    const int r = e / per, rem = e - r * per, k = rem / C, c = rem - k * C;
    const int q0 = sRun[r], q1 = sRun[r + 1];
    const int32_t t = sT[q0];
    double s = 0.0;
    if (c < kBwdStageC) {
      for (int q = q0; q < q1 && sT[q] == t; q++) s += sB[q][k] * (double)sG[q][c];
    } else {
      for (int q = q0; q < q1 && sT[q] == t; q++) s += sB[q][k] * (double)gradOut[(int64_t)sPix[q] * C + c];
    }
    if (s != 0.0) {
      const int64_t o = (int64_t)sV[q0][k] * C + c;
      if (DET) fixedAdd(acc, G.numVertices * 4 + o, s);
      else unsafeAtomicAdd(&gradAttr[o], (float)s);
    }
  }
  // ---- phase 3: position gradients per (run, corner, x/y/w) ----------------
  for (int e = tid; e < (int)numRuns * 9; e += kBwdThreads) {
    const int r = e / 9, rem = e - r * 9, k = rem / 3, comp = rem - k * 3;
    const int q0 = sRun[r], q1 = sRun[r + 1];
    const int32_t t = sT[q0];
    double s = 0.0;
    for (int q = q0; q < q1 && sT[q] == t; q++) s += sP[q][rem];
    if (s != 0.0) {
      const int64_t o = (int64_t)sV[q0][k] * 4 + (comp == 2 ? 3 : comp);
      if (DET) fixedAdd(acc, o, s);
      else unsafeAtomicAdd(&gradPos[o], (float)s);
    }
  }
}

// Zeroes [0, n) of p: up to three floats until a 16-byte boundary, then
// 16-byte stores, then the tail; `lane` and `lanes` stride the grid.
__device__ __forceinline__ void zeroFloats(float* __restrict__ p, int64_t n, int64_t lane, int64_t lanes) {
  const int64_t head = min<int64_t>(n, (int64_t)(((16u - ((uintptr_t)p & 15u)) & 15u) >> 2));
  if (lane < head) p[lane] = 0.0f;
  float4* __restrict__ q = reinterpret_cast<float4*>(p + head);
  const int64_t n4 = (n - head) >> 2;
  for (int64_t k = lane; k < n4; k += lanes) q[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const int64_t tail = head + (n4 << 2);
  if (lane < n - tail) p[tail + lane] = 0.0f;
}

// Zeroes [0, na) of a and [0, nb) of b (4-byte aligned).
__global__ void __launch_bounds__(256) zeroTwoKernel(float* __restrict__ a, int64_t na, float* __restrict__ b,
                                                     int64_t nb) {
  const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, lanes = (int64_t)gridDim.x * blockDim.x;
  zeroFloats(a, na, lane, lanes);
  zeroFloats(b, nb, lane, lanes);
}

static int gridFor(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 1 << 20); }

// Zeroes n 64-bit words.
__global__ void __launch_bounds__(256) zeroWordsKernel(unsigned long long* __restrict__ w, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    w[k] = 0ull;
}

// Rounds the fixed-point totals to FP32: positions (V*4), then attributes.
__global__ void __launch_bounds__(256) fixedToFloatKernel(const unsigned long long* __restrict__ acc,
                                                          int64_t nPos, int64_t nAttr, float* __restrict__ gradPos,
                                                          float* __restrict__ gradAttr) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nPos + nAttr;
       k += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = acc[2 * k];
    const int64_t hi = (int64_t)acc[2 * k + 1];
    const double v = (double)hi + (double)lo * 5.421010862427522e-20;
    if (k < nPos) gradPos[k] = (float)v;
    else gradAttr[k - nPos] = (float)v;
  }
}

void launchInterpForward(const InterpGeom& G, const uint32_t* ids, const float* pos, const float* attr,
                         const int32_t* tris, float* out, int32_t* triImage, hipStream_t stream) {
  if (G.width <= 0 || G.height <= 0) return;
  const int vec4 = G.numChannels % 4 == 0 && ((uintptr_t)attr & 15) == 0 && ((uintptr_t)out & 15) == 0;
  hipLaunchKernelGGL(interpForwardKernel, dim3((unsigned)((G.width + 255) / 256), (unsigned)G.height), dim3(256), 0,
                     stream, G, ids, reinterpret_cast<const float4*>(pos), attr, reinterpret_cast<const int3*>(tris),
                     out, triImage, vec4);
}

void launchInterpBackward(const InterpGeom& G, const int32_t* triImage, const float* pos, const float* attr,
                          const int32_t* tris, const float* gradOut, float* gradPos, float* gradAttr,
                          unsigned long long* acc, hipStream_t stream) {
  const int64_t nPos = G.numVertices * 4, nAttr = G.numVertices * G.numChannels;
  const dim3 grid((unsigned)((G.width + kBwdTile - 1) / kBwdTile), (unsigned)((G.height + kBwdTile - 1) / kBwdTile));
  const bool work = G.width > 0 && G.height > 0 && G.numTris > 0;
  if (acc == nullptr) {
    if (nPos + nAttr > 0)
      hipLaunchKernelGGL(zeroTwoKernel, dim3(gridFor((nPos + nAttr + 3) / 4)), dim3(256), 0, stream, gradPos, nPos,
                         gradAttr, nAttr);
    if (work)
      hipLaunchKernelGGL(interpBackwardKernel<false>, grid, dim3(kBwdThreads), 0, stream, G, triImage,
                         reinterpret_cast<const float4*>(pos), attr, reinterpret_cast<const int3*>(tris), gradOut,
                         gradPos, gradAttr, acc);
    return;
  }
  if (nPos + nAttr == 0) return;
  hipLaunchKernelGGL(zeroWordsKernel, dim3(gridFor(2 * (nPos + nAttr))), dim3(256), 0, stream, acc,
                     2 * (nPos + nAttr));
  if (work)
    hipLaunchKernelGGL(interpBackwardKernel<true>, grid, dim3(kBwdThreads), 0, stream, G, triImage,
                       reinterpret_cast<const float4*>(pos), attr, reinterpret_cast<const int3*>(tris), gradOut,
                       gradPos, gradAttr, acc);
  hipLaunchKernelGGL(fixedToFloatKernel, dim3(gridFor(nPos + nAttr)), dim3(256), 0, stream, acc, nPos, nAttr,
                     gradPos, gradAttr);
}

}  // namespace crhip
