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
// gradients accumulate with FP64 atomics and are rounded to FP32 once.
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

__global__ void __launch_bounds__(256) interpForwardKernel(const InterpGeom G, const uint32_t* __restrict__ ids,
                                                           const float4* __restrict__ pos,
                                                           const float* __restrict__ attr,
                                                           const int3* __restrict__ tris, float* __restrict__ out,
                                                           int32_t* __restrict__ triImage) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)G.width * G.height) return;
  const int x = (int)(p % G.width), y = (int)(p / G.width);
  const int C = G.numChannels;
  float* o = out + p * C;
  uint32_t t;
  int32_t i[3];
  if (!pixelTriangle(G, ids, tris, x, y, t, i)) {
    for (int c = 0; c < C; c++) o[c] = 0.0f;
    if (triImage) triImage[p] = -1;
    return;
  }
  const double qx = (2.0 * x + 1.0) / G.width - 1.0, qy = (2.0 * y + 1.0) / G.height - 1.0;
  Bary B;
  baryAt(pos, i, qx, qy, B);
  const float* a0 = attr + (int64_t)i[0] * C;
  const float* a1 = attr + (int64_t)i[1] * C;
  const float* a2 = attr + (int64_t)i[2] * C;
  for (int c = 0; c < C; c++) o[c] = (float)(B.b[0] * a0[c] + B.b[1] * a1[c] + B.b[2] * a2[c]);
  if (triImage) triImage[p] = (int32_t)t;
}

// One lane per pixel.  dL/db_k = sum_c g_c a_k,c; dL/dc_j = (dL/db_j -
// sum_k b_k dL/db_k) / (c_0 + c_1 + c_2); c_k = x'_{k+1} y'_{k+2} -
// y'_{k+1} x'_{k+2}; x' = x - qx w, y' = y - qy w.
__global__ void __launch_bounds__(256) interpBackwardKernel(const InterpGeom G, const int32_t* __restrict__ triImage,
                                                            const float4* __restrict__ pos,
                                                            const float* __restrict__ attr,
                                                            const int3* __restrict__ tris,
                                                            const float* __restrict__ gradOut,
                                                            double* __restrict__ accPos,
                                                            double* __restrict__ accAttr) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)G.width * G.height) return;
  const int32_t t = triImage[p];
  if (t < 0 || t >= G.numTris) return;
  const int3 v = tris[t];
  const int32_t i[3] = {v.x, v.y, v.z};
  for (int k = 0; k < 3; k++)
    if (i[k] < 0 || (int64_t)i[k] >= G.numVertices) return;
  const int x = (int)(p % G.width), y = (int)(p / G.width);
  const int C = G.numChannels;
  const double qx = (2.0 * x + 1.0) / G.width - 1.0, qy = (2.0 * y + 1.0) / G.height - 1.0;
  Bary B;
  baryAt(pos, i, qx, qy, B);
  if (!B.ok) return;
  const float* g = gradOut + p * C;
  double dB[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 3; k++) {
    const float* a = attr + (int64_t)i[k] * C;
    double* acc = accAttr + (int64_t)i[k] * C;
    double s = 0.0;
    for (int c = 0; c < C; c++) {
      const double gc = (double)g[c];
      s += gc * (double)a[c];
      if (gc != 0.0) unsafeAtomicAdd(&acc[c], B.b[k] * gc);
    }
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
    double* acc = accPos + (int64_t)i[k] * 4;
    if (dxs[k] != 0.0) unsafeAtomicAdd(&acc[0], dxs[k]);
    if (dys[k] != 0.0) unsafeAtomicAdd(&acc[1], dys[k]);
    const double dw = -B.qx * dxs[k] - B.qy * dys[k];
    if (dw != 0.0) unsafeAtomicAdd(&acc[3], dw);
  }
}

__global__ void __launch_bounds__(256) fillZeroKernel(double* __restrict__ p, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    p[k] = 0.0;
}

__global__ void __launch_bounds__(256) roundToFloatKernel(const double* __restrict__ src, float* __restrict__ dst,
                                                          int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    dst[k] = (float)src[k];
}

static int gridFor(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 1 << 20); }

void launchInterpForward(const InterpGeom& G, const uint32_t* ids, const float* pos, const float* attr,
                         const int32_t* tris, float* out, int32_t* triImage, hipStream_t stream) {
  const int64_t n = (int64_t)G.width * G.height;
  if (n == 0) return;
  hipLaunchKernelGGL(interpForwardKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, G, ids,
                     reinterpret_cast<const float4*>(pos), attr, reinterpret_cast<const int3*>(tris), out, triImage);
}

void launchInterpBackward(const InterpGeom& G, const int32_t* triImage, const float* pos, const float* attr,
                          const int32_t* tris, const float* gradOut, double* accPos, double* accAttr,
                          float* gradPos, float* gradAttr, hipStream_t stream) {
  const int64_t nPos = G.numVertices * 4, nAttr = G.numVertices * G.numChannels;
  if (nPos > 0) hipLaunchKernelGGL(fillZeroKernel, dim3(gridFor(nPos)), dim3(256), 0, stream, accPos, nPos);
  if (nAttr > 0) hipLaunchKernelGGL(fillZeroKernel, dim3(gridFor(nAttr)), dim3(256), 0, stream, accAttr, nAttr);
  const int64_t n = (int64_t)G.width * G.height;
  if (n > 0)
    hipLaunchKernelGGL(interpBackwardKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, G, triImage,
                       reinterpret_cast<const float4*>(pos), attr, reinterpret_cast<const int3*>(tris), gradOut,
                       accPos, accAttr);
  if (nPos > 0) hipLaunchKernelGGL(roundToFloatKernel, dim3(gridFor(nPos)), dim3(256), 0, stream, accPos, gradPos, nPos);
  if (nAttr > 0)
    hipLaunchKernelGGL(roundToFloatKernel, dim3(gridFor(nAttr)), dim3(256), 0, stream, accAttr, gradAttr, nAttr);
}

}  // namespace crhip
