// The step before the path: the passthrough vertex shader.
//
// vertexShader_passthrough (test/shader/PassThrough.cu:16-35) writes each
// vertex's clip position posToClip * (modelPos, 1), with Mat4f * Vec4f as
// base/Math.hpp:990-1001 evaluates it: per row i, rr = 0, then
// rr += m(i, j) * v[j] for j = 0..3, which nvcc contracts to fused
// multiply-adds -- restated here as an explicit fmaf chain in the same
// order, so the clip positions are bit-identical.  Only the first 16 bytes
// of each output record are written (ShadedVertexBase::clipPos); the rest of
// the record (varyings) is left as it is, so the kernel can fill the
// position part of an interleaved vertex buffer that setup then reads.
//
// One lane per vertex: 12 bytes read, 16 bytes written, HBM-bound.  The
// matrix arrives as a kernel argument (scalar registers).
#include "cr_device.h"

namespace crhip {

struct Mat4 { float m[16]; };      // Mat4f memory order: m(r, c) = m[r + 4c]

__global__ void __launch_bounds__(256) vertexPassthroughKernel(const float* __restrict__ modelPos,
                                                               int64_t numVertices, const Mat4 M,
                                                               uint8_t* __restrict__ out, int64_t outStride) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= numVertices) return;
  const float in[4] = {modelPos[3 * v + 0], modelPos[3 * v + 1], modelPos[3 * v + 2], 1.0f};
  float r[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float rr = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++) rr = fmaf(M.m[i + 4 * j], in[j], rr);
    r[i] = rr;
  }
  *reinterpret_cast<float4*>(out + v * outStride) = make_float4(r[0], r[1], r[2], r[3]);
}

void launchVertexPassthrough(const float* modelPos, int64_t numVertices, const float posToClip[16], void* out,
                             int64_t outStride, hipStream_t stream) {
  if (numVertices <= 0) return;
  Mat4 M;
  for (int k = 0; k < 16; k++) M.m[k] = posToClip[k];
  hipLaunchKernelGGL(vertexPassthroughKernel, dim3((unsigned)((numVertices + 255) / 256)), dim3(256), 0, stream,
                     modelPos, numVertices, M, static_cast<uint8_t*>(out), outStride);
}

}  // namespace crhip
