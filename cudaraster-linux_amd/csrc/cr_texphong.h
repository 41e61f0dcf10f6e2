// FragmentShader_texPhong (test/shader/Shaders.cu:122-199) for the fine
// stage's in-order shading: perspective-correct varyings at the centroid
// (camera position, camera normal) and the texture coordinate with its
// derivatives -- analytic (computeBarys, FineRaster.inl:21-44) or, with
// RenderModeFlag_EnableQuads, numerical over the 2x2 pixel quad
// (PixelPipe.hpp:59-60) -- then the material, mip-mapped texture lookups
// (texture2D, Shaders.cu:125-148), the alpha test and evaluateLighting
// (Shaders.cu:37-48).
//
// Every operation is IEEE single precision with the fused multiply-adds
// written out, and the transcendental functions are this library's own
// (the reference builds its shaders with nvcc -use_fast_math, whose
// approximations are not restated): log2 by the atanh series on the
// mantissa, exp2 by range reduction and the Taylor series of e^z.
// Bilinear filtering is done in FP32 by the shader itself (normalised
// coordinates, wrap addressing), not by the texture unit.
#pragma once

#include "cr_device.h"

namespace crhip {

// log2: x = m 2^e, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(t),
// t = (m-1)/(m+1), atanh t = t (1 + t^2/3 + t^4/5 + t^6/7 + t^8/9).
__device__ __forceinline__ float crLog2(float x) {
  if (x != x || x < 0.0f) return __builtin_nanf("");
  if (x == 0.0f) return -__builtin_inff();
  if (x == __builtin_inff()) return __builtin_inff();
  int e = 0;
  uint32_t u = __float_as_uint(x);
  if ((u >> 23) == 0) {                 // subnormal: scale by 2^23 (exact)
    u = __float_as_uint(x * 8388608.0f);
    e = -23;
  }
  e += (int)(u >> 23) - 127;
  float m = __uint_as_float((u & 0x007FFFFFu) | 0x3F800000u);
  if (m > 1.41421354f) { m *= 0.5f; e += 1; }
  const float t = (m - 1.0f) / (m + 1.0f);
  const float t2 = t * t;
  float p = fmaf(t2, 0.111111112f, 0.142857149f);
  p = fmaf(t2, p, 0.200000003f);
  p = fmaf(t2, p, 0.333333343f);
  p = fmaf(t2, p, 1.0f);
  const float ln = (2.0f * t) * p;
  return fmaf(ln, 1.44269502f, (float)e);
}

// exp2: y = n + f, |f| <= 1/2; e^z, z = f ln 2, by its Taylor series to
// z^7/7!, scaled by 2^n.
__device__ __forceinline__ float crExp2(float y) {
  if (y != y) return y;
  if (y >= 128.0f) return __builtin_inff();
  if (y < -150.0f) return 0.0f;
  const float n = __builtin_rintf(y);
  const float f = y - n;
  const float z = f * 0.693147182f;
  float p = fmaf(z, 1.98412701e-04f, 1.38888892e-03f);
  p = fmaf(z, p, 8.33333377e-03f);
  p = fmaf(z, p, 4.16666679e-02f);
  p = fmaf(z, p, 1.66666672e-01f);
  p = fmaf(z, p, 0.5f);
  p = fmaf(z, p, 1.0f);
  p = fmaf(z, p, 1.0f);
  return __builtin_ldexpf(p, (int)n);
}

__device__ __forceinline__ float crPow(float x, float y) {
  if (y == 0.0f) return 1.0f;
  if (x == 0.0f) return y > 0.0f ? 0.0f : __builtin_inff();
  return crExp2(y * crLog2(x));
}

// Texel index of floor value f wrapped into [0, n); NaN and out-of-range
// values give 0.
__device__ __forceinline__ int32_t wrapIndex(float f, int32_t n) {
  if (!(fabsf(f) < 1073741824.0f)) return 0;
  const int32_t k = (int32_t)f % n;
  return k < 0 ? k + n : k;
}

// tex2D on the float4 atlas: normalised (u, v), wrap addressing, bilinear in
// FP32 (texel space x = u W - 1/2; x first, then y).
__device__ __forceinline__ float4 texSample(const float4* __restrict__ atlas, int32_t w, int32_t h, float u,
                                            float v) {
  const float x = fmaf(u, (float)w, -0.5f), y = fmaf(v, (float)h, -0.5f);
  const float fx = floorf(x), fy = floorf(y);
  const float a = x - fx, b = y - fy;
  const int32_t i0 = wrapIndex(fx, w), j0 = wrapIndex(fy, h);
  const int32_t i1 = i0 + 1 == w ? 0 : i0 + 1, j1 = j0 + 1 == h ? 0 : j0 + 1;
  const float4 t00 = atlas[(int64_t)j0 * w + i0], t10 = atlas[(int64_t)j0 * w + i1];
  const float4 t01 = atlas[(int64_t)j1 * w + i0], t11 = atlas[(int64_t)j1 * w + i1];
  auto bil = [&](float c00, float c10, float c01, float c11) {
    const float top = fmaf(c10 - c00, a, c00);
    const float bot = fmaf(c11 - c01, a, c01);
    return fmaf(bot - top, b, top);
  };
  return make_float4(bil(t00.x, t10.x, t01.x, t11.x), bil(t00.y, t10.y, t01.y, t11.y),
                     bil(t00.z, t10.z, t01.z, t11.z), bil(t00.w, t10.w, t01.w, t11.w));
}

// texture2D (Shaders.cu:125-148).
__device__ __forceinline__ float4 texture2D(const TexSpec& spec, const float4* __restrict__ atlas, int32_t w,
                                            int32_t h, float2 tex, float2 texDX, float2 texDY) {
  const float ax = texDX.x * spec.size[0], ay = texDX.y * spec.size[1];
  const float bx = texDY.x * spec.size[0], by = texDY.y * spec.size[1];
  const float dxlen = fmaf(ax, ax, ay * ay);
  const float dylen = fmaf(bx, bx, by * by);
  const float lod = fminf(fmaxf(crLog2(fmaxf(dxlen, dylen)) * 0.5f, 0.0f), 14.0f);
  const int li = (int)lod;
  const float4 m0 = spec.mip[li], m1 = spec.mip[li + 1];
  const float tx = tex.x - floorf(tex.x);
  const float ty = tex.y - floorf(tex.y);
  const float4 v0 = texSample(atlas, w, h, fmaf(tx, m0.x, m0.z), fmaf(ty, m0.y, m0.w));
  const float4 v1 = texSample(atlas, w, h, fmaf(tx, m1.x, m1.z), fmaf(ty, m1.y, m1.w));
  const float f = lod - (float)li;
  const float g = 1.0f - f;
  return make_float4(fmaf(v0.x, g, v1.x * f), fmaf(v0.y, g, v1.y * f), fmaf(v0.z, g, v1.z * f),
                     fmaf(v0.w, g, v1.w * f));
}

__device__ __forceinline__ float3 normalize3(float3 v) {
  const float l = fmaf(v.z, v.z, fmaf(v.y, v.y, v.x * v.x));
  const float r = __builtin_sqrtf(l);
  const float k = r != 0.0f ? 1.0f / r : 0.0f;
  return make_float3(v.x * k, v.y * k, v.z * k);
}

__device__ __forceinline__ float dot3(float3 a, float3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }

// evaluateLighting (Shaders.cu:37-48).
__device__ __forceinline__ float3 evaluateLighting(float3 cameraPos, float3 cameraNormal, const Material& m,
                                                   float3 diffuse) {
  const float3 I = normalize3(cameraPos);
  const float3 N = normalize3(cameraNormal);
  const float dotIN = dot3(I, N);
  const float s2 = dotIN * 2.0f;
  const float3 R = make_float3(fmaf(-N.x, s2, I.x), fmaf(-N.y, s2, I.y), fmaf(-N.z, s2, I.z));
  const float diffuseCoef = fmaf(fmaxf(-dotIN, 0.0f), 0.75f, 0.25f);
  const float specularCoef = crPow(fmaxf(-dot3(I, R), 0.0f), m.glossiness);
  return make_float3(fmaf(diffuseCoef, diffuse.x, specularCoef * m.specular[0]),
                     fmaf(diffuseCoef, diffuse.y, specularCoef * m.specular[1]),
                     fmaf(diffuseCoef, diffuse.z, specularCoef * m.specular[2]));
}

// Barycentric planes of one subtriangle (TriData words 4..12).
struct Planes {
  int32_t wx, wy, wb, ux, uy, ub, vx, vy, vb;
};

// computeBarys (FineRaster.inl:21-44) at half-sample position (sx, sy):
// barycentrics and their change per pixel step.
template <int S>
__device__ __forceinline__ void computeBarys(const Planes& q, int32_t sx, int32_t sy, float3& b, float3& bdx,
                                             float3& bdy) {
  const int32_t W = wadd(wadd(wmul(q.wx, sx), wmul(q.wy, sy)), q.wb);
  const int32_t U = wadd(wadd(wmul(q.ux, sx), wmul(q.uy, sy)), q.ub);
  const int32_t V = wadd(wadd(wmul(q.vx, sx), wmul(q.vy, sy)), q.vb);
  const float w = 1.0f / (float)W;
  const float u = w * (float)U;
  const float v = w * (float)V;
  b = make_float3(1.0f - u - v, u, v);
  const float wd = w * (float)(1 << (S + 1));
  const float udx = wd * fmaf(-u, (float)q.wx, (float)q.ux);
  const float udy = wd * fmaf(-u, (float)q.wy, (float)q.uy);
  const float vdx = wd * fmaf(-v, (float)q.wx, (float)q.vx);
  const float vdy = wd * fmaf(-v, (float)q.wy, (float)q.vy);
  bdx = make_float3(-udx - vdx, udx, vdx);
  bdy = make_float3(-udy - vdy, udy, vdy);
}

// interpolateVarying (PixelPipe.inl:52-59).
__device__ __forceinline__ float4 interpVarying(const float4& c0, const float4& c1, const float4& c2, float3 b) {
  return make_float4(fmaf(c2.x, b.z, fmaf(c0.x, b.x, c1.x * b.y)), fmaf(c2.y, b.z, fmaf(c0.y, b.x, c1.y * b.y)),
                     fmaf(c2.z, b.z, fmaf(c0.z, b.x, c1.z * b.y)), fmaf(c2.w, b.z, fmaf(c0.w, b.x, c1.w * b.y)));
}

__device__ __forceinline__ float4 loadVarying(const uint8_t* verts, uint32_t vi, int stride, int k) {
  return *reinterpret_cast<const float4*>(verts + (uint64_t)vi * (uint32_t)stride + 16 + 16 * k);
}

// Texture coordinate (varying 2) at the centre of pixel (px, py).
template <int S>
__device__ __forceinline__ float2 centerTex(const Planes& q, int px, int py, const float4& t0, const float4& t1,
                                            const float4& t2) {
  constexpr int N = 1 << S;
  float3 b, bdx, bdy;
  computeBarys<S>(q, px * (2 * N) + N, py * (2 * N) + N, b, bdx, bdy);
  const float4 t = interpVarying(t0, t1, t2, b);
  return make_float2(t.x, t.y);
}

// FragmentShader_texPhong::run for pixel (px, py) of source triangle tri,
// coverage centroid code (15: pixel centre).  Returns true when the
// fragment is discarded (alpha test, or a material index outside the
// table); otherwise *color is the shaded ABGR colour.
template <int S>
__device__ __forceinline__ bool shadeTexPhong(const Params& P, const Planes& q, uint32_t vi0, uint32_t vi1,
                                              uint32_t vi2, uint32_t tri, int code, int px, int py,
                                              uint32_t* color) {
  constexpr int N = 1 << S;
  const bool lerp = (P.flags & kFlagLerp) != 0;
  const int stride = P.vertexStride;
  float3 cb = make_float3(0.0f, 0.0f, 1.0f), cdx = make_float3(0.0f, 0.0f, 0.0f), cdy = cdx;
  if (lerp) {
    const int cx = (code == 15) ? N : 2 * msaaPattern(S, code) + 1;
    const int cy = (code == 15) ? N : 2 * code + 1;
    computeBarys<S>(q, px * (2 * N) + cx, py * (2 * N) + cy, cb, cdx, cdy);
  }
  const float4 p0 = loadVarying(P.vertices, vi0, stride, 0), p1 = loadVarying(P.vertices, vi1, stride, 0),
               p2 = loadVarying(P.vertices, vi2, stride, 0);
  const float4 n0 = loadVarying(P.vertices, vi0, stride, 1), n1 = loadVarying(P.vertices, vi1, stride, 1),
               n2 = loadVarying(P.vertices, vi2, stride, 1);
  const float4 t0 = loadVarying(P.vertices, vi0, stride, 2), t1 = loadVarying(P.vertices, vi1, stride, 2),
               t2 = loadVarying(P.vertices, vi2, stride, 2);
  const float4 cp = interpVarying(p0, p1, p2, cb);
  const float4 cn = interpVarying(n0, n1, n2, cb);
  float2 tex, texDX, texDY;
  if (!(P.flags & kFlagQuads)) {
    // Centroid sample, analytic derivatives.
    const float4 t = interpVarying(t0, t1, t2, cb);
    const float4 tdx = interpVarying(t0, t1, t2, cdx);
    const float4 tdy = interpVarying(t0, t1, t2, cdy);
    tex = make_float2(t.x, t.y);
    texDX = make_float2(tdx.x, tdx.y);
    texDY = make_float2(tdy.x, tdy.y);
  } else if (!lerp) {
    // Varyings of vertex 2 everywhere: nothing changes across the quad.
    const float4 t = interpVarying(t0, t1, t2, make_float3(0.0f, 0.0f, 1.0f));
    tex = make_float2(t.x, t.y);
    texDX = make_float2(0.0f, 0.0f);
    texDY = texDX;
  } else {
    // Pixel centre, numerical derivatives over the 2x2 quad: the odd
    // column's (row's) value minus the even one's; the quad partner's value
    // is the same shader run at its own pixel centre (as the reference's
    // helper lanes do).
    tex = centerTex<S>(q, px, py, t0, t1, t2);
    const float2 ox = centerTex<S>(q, px ^ 1, py, t0, t1, t2);
    const float2 oy = centerTex<S>(q, px, py ^ 1, t0, t1, t2);
    const bool oddX = (px & 1) != 0, oddY = (py & 1) != 0;
    texDX = oddX ? make_float2(tex.x - ox.x, tex.y - ox.y) : make_float2(ox.x - tex.x, ox.y - tex.y);
    texDY = oddY ? make_float2(tex.x - oy.x, tex.y - oy.y) : make_float2(oy.x - tex.x, oy.y - tex.y);
  }
  const int32_t mi = (int64_t)tri < P.numTriMaterials ? P.triMaterial[tri] : -1;
  if (mi < 0 || mi >= P.numMaterials) return true;
  const Material& M = P.materials[mi];
  float4 dc = M.diffuse;
  if (M.diffuseTexture.size[0] != 0.0f) {
    const float4 t = texture2D(M.diffuseTexture, P.atlas, P.atlasWidth, P.atlasHeight, tex, texDX, texDY);
    dc.x = t.x;
    dc.y = t.y;
    dc.z = t.z;
  }
  if (M.alphaTexture.size[0] != 0.0f)
    dc.w = texture2D(M.alphaTexture, P.atlas, P.atlasWidth, P.atlasHeight, tex, texDX, texDY).y;
  if (dc.w < 0.5f) return true;
  const float3 c = evaluateLighting(make_float3(cp.x, cp.y, cp.z), make_float3(cn.x, cn.y, cn.z), M,
                                    make_float3(dc.x, dc.y, dc.z));
  *color = toABGR(c.x, c.y, c.z, dc.w);
  return false;
}

}  // namespace crhip
