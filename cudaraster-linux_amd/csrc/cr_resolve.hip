// The step after the path: MSAA resolve of a colour surface.
//
// The reference declares a resolve but leaves it commented out
// (CudaSurface.hpp:73) and the demo samples the raw multisample texture
// (test/SceneCR.cpp:319-339).  This is the box resolve: each RGBA8 channel of
// a pixel is the mean of its N samples, rounded half up,
// (sum + N/2) >> log2(N) -- integer and exact.
//
// Surface layout (CudaSurface.cpp:86-91): row y of the texture holds, for
// each 8-pixel tile column tx, sample 0 of its 8 pixels, then sample 1, ...;
// sample i of pixel (x, y) is texel y * textureW + (x >> 3) * 8N + i * 8 +
// (x & 7).  One lane per pixel: for each sample, 8 consecutive lanes read 8
// consecutive texels.  HBM-bound: 4N bytes read and 4 written per pixel.
#include "cr_device.h"

namespace crhip {

__global__ void __launch_bounds__(256) resolveKernel(const uint32_t* __restrict__ tex, int32_t textureW, int32_t width,
                                                     int32_t height, int32_t samplesLog2,
                                                     uint32_t* __restrict__ out) {
  const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int y = (int)blockIdx.y;
  if (x >= width || y >= height) return;
  const int n = 1 << samplesLog2;
  const uint32_t* row = tex + (size_t)y * textureW + (size_t)(x >> 3) * 8 * n + (x & 7);
  uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int i = 0; i < n; i++) {
    const uint32_t c = row[i * 8];
    s0 += c & 0xFFu;
    s1 += (c >> 8) & 0xFFu;
    s2 += (c >> 16) & 0xFFu;
    s3 += c >> 24;
  }
  const uint32_t h = (uint32_t)n >> 1;
  out[(size_t)y * width + x] = ((s0 + h) >> samplesLog2) | (((s1 + h) >> samplesLog2) << 8) |
                               (((s2 + h) >> samplesLog2) << 16) | (((s3 + h) >> samplesLog2) << 24);
}

void launchResolve(const uint32_t* tex, int32_t textureW, int32_t width, int32_t height, int32_t samplesLog2,
                   uint32_t* out, hipStream_t stream) {
  if (width <= 0 || height <= 0) return;
  hipLaunchKernelGGL(resolveKernel, dim3((unsigned)((width + 255) / 256), (unsigned)height), dim3(256), 0, stream,
                     tex, textureW, width, height, samplesLog2, out);
}

}  // namespace crhip
