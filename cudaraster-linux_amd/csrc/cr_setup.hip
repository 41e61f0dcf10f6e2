// Stage 1: triangle setup, one lane per input triangle.
//
// Follows triangleSetupImpl (cudaraster/cuda/TriangleSetup.inl:219-417):
// frustum cull, fast path (snap + guard band), or barycentric clip against
// the frustum fanned into subtriangles; backface and between-samples culls;
// plane equations for depth and perspective-correct barycentrics.  Output is
// the per-triangle subtriangle count, the 16-byte header and 64-byte data
// record per subtriangle slot.  Clipped triangles with >= 2 visible
// subtriangles take consecutive slots from a device counter (slots beyond
// numTris), their source index goes to subtriTri, and misc holds the base.
//
// The clip polygon lives in registers (at most 9 vertices) instead of the
// reference's per-thread shared-memory array.
#include "cr_device.h"

namespace crhip {

struct I2 { int32_t x, y; };

// Util.inl:65-91 (GPU plane-equation setup).
__device__ __forceinline__ void setupPleq(float vx, float vy, float vz, int32_t v0x, int32_t v0y,
                                          int32_t d1x, int32_t d1y, int32_t d2x, int32_t d2y,
                                          float areaRcp, int samplesLog2, uint32_t out[3]) {
  float mx = fmaxf(fmaxf(vx, vy), vz);
  int32_t sh = asr32((int32_t)__float_as_uint(mx), 23) - (127 + 22);
  sh = min(max(sh, 0), 8);
  int32_t t0 = (int32_t)(cvtRziSatU32(vx) >> sh);
  int32_t t1 = wsub((int32_t)(cvtRziSatU32(vy) >> sh), t0);
  int32_t t2 = wsub((int32_t)(cvtRziSatU32(vz) >> sh), t0);

  uint32_t rcpMant = (__float_as_uint(areaRcp) & 0x007FFFFFu) | 0x00800000u;
  int32_t rcpShift = (23 + 127) - asr32((int32_t)__float_as_uint(areaRcp), 23);

  int64_t xc = wmul64((int64_t)t1 * d2y - (int64_t)t2 * d1y, (int64_t)rcpMant);
  int64_t yc = wmul64((int64_t)t2 * d1x - (int64_t)t1 * d2x, (int64_t)rcpMant);
  int shp = rcpShift - (sh + kSubpixelLog2 - samplesLog2);
  out[0] = (uint32_t)asr64(xc, shp);
  out[1] = (uint32_t)asr64(yc, shp);

  int sc = kSubpixelLog2 - samplesLog2;
  int32_t minx = min(min(d1x, d2x), 0), maxx = max(max(d1x, d2x), 0);
  int32_t miny = min(min(d1y, d2y), 0), maxy = max(max(d1y, d2y), 0);
  int32_t centerX = asr32(wadd(wadd(wmul(v0x, 2), minx), maxx), sc + 1);
  int32_t centerY = asr32(wadd(wadd(wmul(v0y, 2), miny), maxy), sc + 1);
  int32_t vcx = wsub(v0x, (int32_t)((uint32_t)centerX << sc));
  int32_t vcy = wsub(v0y, (int32_t)((uint32_t)centerY << sc));

  uint32_t z = (uint32_t)t0 << sh;
  int64_t cross = (int64_t)((uint64_t)wmul64(asr64(xc, 13), vcx) + (uint64_t)wmul64(asr64(yc, 13), vcy));
  z -= (uint32_t)asr64(cross, rcpShift - (sh + 13));
  z -= out[0] * (uint32_t)centerX + out[1] * (uint32_t)centerY;
  out[2] = z;
}

// cover8x8_selectFlips (Util.hpp:207-217); kept in the header record.
__device__ __forceinline__ uint32_t selectFlips(int32_t dx, int32_t dy) {
  uint32_t f = 0;
  if (dy > 0 || (dy == 0 && dx <= 0)) f ^= (1u << 3) ^ (1u << 2) ^ (1u << 5);
  if (dx > 0) f ^= (1u << 3) ^ (1u << 2);
  if (wabs32(dx) < wabs32(dy)) f ^= (1u << 4) ^ (1u << 2);
  return f;
}

// The draw state the clip path reads, passed to it by value: a reference
// to the kernel's Params would make every wave copy the whole argument
// block to scratch memory.
struct ClipCtx {
  int32_t viewportWidth, viewportHeight, samplesLog2, flags, shader, numTris, maxSubtris;
  uint8_t* triSubtris;
  Header* triHeader;
  TriData* triData;
  uint32_t* subtriTri;
  Counters* counters;
};

// snapTriangle (TriangleSetup.inl:19-31)
template <class PP>
__device__ __forceinline__ void snap(const PP& P, const float4& v0, const float4& v1, const float4& v2,
                                     I2& p0, I2& p1, I2& p2, float rcpW[3], I2& lo, I2& hi) {
  float sx = (float)(P.viewportWidth << (kSubpixelLog2 - 1));
  float sy = (float)(P.viewportHeight << (kSubpixelLog2 - 1));
  rcpW[0] = 1.0f / v0.w;
  rcpW[1] = 1.0f / v1.w;
  rcpW[2] = 1.0f / v2.w;
  p0.x = cvtRniSatS32(v0.x * rcpW[0] * sx); p0.y = cvtRniSatS32(v0.y * rcpW[0] * sy);
  p1.x = cvtRniSatS32(v1.x * rcpW[1] * sx); p1.y = cvtRniSatS32(v1.y * rcpW[1] * sy);
  p2.x = cvtRniSatS32(v2.x * rcpW[2] * sx); p2.y = cvtRniSatS32(v2.y * rcpW[2] * sy);
  lo.x = min(min(p0.x, p1.x), p2.x); lo.y = min(min(p0.y, p1.y), p2.y);
  hi.x = max(max(p0.x, p1.x), p2.x); hi.y = max(max(p0.y, p1.y), p2.y);
}

// prepareTriangle (TriangleSetup.inl:38-100): 0 visible, 1 backfacing,
// 2 between samples.
template <class PP>
__device__ __forceinline__ int prepare(const PP& P, I2 p0, I2 p1, I2 p2, I2 lo, I2 hi,
                                       I2& d1, I2& d2, int32_t& area) {
  int s = P.samplesLog2;
  d1.x = wsub(p1.x, p0.x); d1.y = wsub(p1.y, p0.y);
  d2.x = wsub(p2.x, p0.x); d2.y = wsub(p2.y, p0.y);
  area = wsub(wmul(d1.x, d2.y), wmul(d1.y, d2.x));
  if (area <= 0) return 1;

  int32_t ss = 1 << (kSubpixelLog2 - s);
  int32_t biasX = (P.viewportWidth << (kSubpixelLog2 - 1)) - (ss >> 1);
  int32_t biasY = (P.viewportHeight << (kSubpixelLog2 - 1)) - (ss >> 1);
  int32_t lox = wadd(wadd(lo.x, ss - 1), biasX) & -ss;
  int32_t loy = wadd(wadd(lo.y, ss - 1), biasY) & -ss;
  int32_t hix = wadd(hi.x, biasX) & -ss;
  int32_t hiy = wadd(hi.y, biasY) & -ss;
  if (lox > hix || loy > hiy) return 2;

  int32_t diff = wsub(wadd(hix, hiy), wadd(lox, loy));
  if (diff <= ss) {
    I2 t0 = {wsub(wadd(p0.x, biasX), lox), wsub(wadd(p0.y, biasY), loy)};
    I2 t1 = {wsub(wadd(p1.x, biasX), lox), wsub(wadd(p1.y, biasY), loy)};
    I2 t2 = {wsub(wadd(p2.x, biasX), lox), wsub(wadd(p2.y, biasY), loy)};
    int32_t e0 = wsub(wmul(t0.x, t1.y), wmul(t0.y, t1.x));
    int32_t e1 = wsub(wmul(t1.x, t2.y), wmul(t1.y, t2.x));
    int32_t e2 = wsub(wmul(t2.x, t0.y), wmul(t2.y, t0.x));
    if (e0 < 0 || e1 < 0 || e2 < 0) {
      if (diff == 0) return 2;
      t0 = {wsub(wadd(p0.x, biasX), hix), wsub(wadd(p0.y, biasY), hiy)};
      t1 = {wsub(wadd(p1.x, biasX), hix), wsub(wadd(p1.y, biasY), hiy)};
      t2 = {wsub(wadd(p2.x, biasX), hix), wsub(wadd(p2.y, biasY), hiy)};
      e0 = wsub(wmul(t0.x, t1.y), wmul(t0.y, t1.x));
      e1 = wsub(wmul(t1.x, t2.y), wmul(t1.y, t2.x));
      e2 = wsub(wmul(t2.x, t0.y), wmul(t2.y, t0.x));
      if (e0 < 0 || e1 < 0 || e2 < 0) return 2;
    }
  }
  return 0;
}

// setupTriangle (TriangleSetup.inl:104-215)
template <class PP>
__device__ __forceinline__ void setupTriangle(const PP& P, Header* th, TriData* td, int3 vidx,
                                              const float4& v0, const float4& v1, const float4& v2,
                                              float2 b0, float2 b1, float2 b2, I2 p0, I2 p1, I2 p2,
                                              const float rcpW[3], I2 d1, I2 d2, int32_t area) {
  int s = P.samplesLog2;
  float areaRcp = 0.0f;
  I2 wv0 = {0, 0};
  if (P.flags & (kFlagDepth | kFlagLerp)) {
    areaRcp = 1.0f / (float)area;
    wv0.x = wadd(p0.x, P.viewportWidth << (kSubpixelLog2 - 1));
    wv0.y = wadd(p0.y, P.viewportHeight << (kSubpixelLog2 - 1));
  }

  uint32_t zmin = 0, zslope = 0;
  uint32_t zp[3] = {0u, 0u, 0u};
  if (P.flags & kFlagDepth) {
    const float zcoef = (float)(kDepthMax - kDepthMin) * 0.5f;
    const float zbias = (float)(kDepthMax + kDepthMin) * 0.5f;
    float z0 = fmaf(v0.z * zcoef, rcpW[0], zbias);
    float z1 = fmaf(v1.z * zcoef, rcpW[1], zbias);
    float z2 = fmaf(v2.z * zcoef, rcpW[2], zbias);
    int32_t zv0x = wv0.x - (1 << (kSubpixelLog2 - s - 1));
    int32_t zv0y = wv0.y - (1 << (kSubpixelLog2 - s - 1));
    setupPleq(z0, z1, z2, zv0x, zv0y, d1.x, d1.y, d2.x, d2.y, areaRcp, s, zp);
    zmin = cvtRniSatU32(fminf(fminf(z0, z1), z2) - (float)lerpError(s));
    if (s != 0) {
      uint32_t tmp = wabs32((int32_t)zp[0]) + wabs32(max((int32_t)zp[1], -INT32_MAX));
      int sh = max(s - 1, 0);
      zslope = tmp << sh;
      if ((zslope >> sh) != tmp) zslope = 0xFFFFFFFFu;
    }
  }
  uint32_t wp[3] = {0u, 0u, 0u}, up[3] = {0u, 0u, 0u}, vp[3] = {0u, 0u, 0u};
  if (P.flags & kFlagLerp) {
    float wcoef = fminf(fminf(v0.w, v1.w), v2.w) * (float)kBaryMax;
    float w0 = wcoef * rcpW[0], w1 = wcoef * rcpW[1], w2 = wcoef * rcpW[2];
    setupPleq(w0, w1, w2, wv0.x, wv0.y, d1.x, d1.y, d2.x, d2.y, areaRcp, s + 1, wp);
    setupPleq(b0.x * w0, b1.x * w1, b2.x * w2, wv0.x, wv0.y, d1.x, d1.y, d2.x, d2.y, areaRcp, s + 1, up);
    setupPleq(b0.y * w0, b1.y * w1, b2.y * w2, wv0.x, wv0.y, d1.x, d1.y, d2.x, d2.y, areaRcp, s + 1, vp);
  }
  uint4* tdv = reinterpret_cast<uint4*>(td);
  tdv[0] = make_uint4(zp[0], zp[1], zp[2], zslope);
  if (P.shader != kShaderConstant) {
    // Barycentric planes and vertex indices: read only by the shaders that
    // interpolate varyings (a constant-colour draw writes the depth plane
    // alone).
    tdv[1] = make_uint4(wp[0], wp[1], wp[2], up[0]);
    tdv[2] = make_uint4(up[1], up[2], vp[0], vp[1]);
    tdv[3] = make_uint4(vp[2], (uint32_t)vidx.x, (uint32_t)vidx.y, (uint32_t)vidx.z);
  }

  uint32_t f01 = selectFlips(d1.x, d1.y);
  uint32_t f12 = selectFlips(wsub(d2.x, d1.x), wsub(d2.y, d1.y));
  uint32_t f20 = selectFlips(-d2.x, -d2.y);
  *reinterpret_cast<uint4*>(th) = make_uint4(
      ((uint32_t)p0.x & 0xFFFFu) | ((uint32_t)p0.y << 16),
      ((uint32_t)p1.x & 0xFFFFu) | ((uint32_t)p1.y << 16),
      ((uint32_t)p2.x & 0xFFFFu) | ((uint32_t)p2.y << 16),
      (zmin & 0xFFFFF000u) | (f01 << 6) | (f12 << 2) | (f20 >> 2));
}

// clipPolygonWithPlane (Util.hpp:221-250), polygon in registers.
__device__ __forceinline__ int clipPolygon(float* out, const float* in, int numIn, float v0, float v1, float v2) {
  int numOut = 0;
  if (numIn >= 3) {
    int ai = (numIn - 1) * 2;
    float av = fmaf(v2, in[ai + 1], fmaf(v1, in[ai + 0], v0));
    for (int bi = 0; bi < numIn * 2; bi += 2) {
      float bv = fmaf(v2, in[bi + 1], fmaf(v1, in[bi + 0], v0));
      if (av * bv < 0.0f) {
        float bc = av / (av - bv);
        float ac = 1.0f - bc;
        out[numOut + 0] = fmaf(in[ai + 0], ac, in[bi + 0] * bc);
        out[numOut + 1] = fmaf(in[ai + 1], ac, in[bi + 1] * bc);
        numOut += 2;
      }
      if (bv >= 0.0f) {
        out[numOut + 0] = in[bi + 0];
        out[numOut + 1] = in[bi + 1];
        numOut += 2;
      }
      ai = bi;
      av = bv;
    }
  }
  return numOut >> 1;
}

__device__ __forceinline__ float comp(const float4& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// clipTriangleWithFrustum (Util.hpp:260-286)
__device__ int clipTriangle(float bary[18], const float4& v0, const float4& v1, const float4& v2,
                            const float4& d1, const float4& d2) {
  int num = 3;
  bary[0] = 0.0f; bary[1] = 0.0f;
  bary[2] = 1.0f; bary[3] = 0.0f;
  bary[4] = 0.0f; bary[5] = 1.0f;
  for (int a = 0; a < 3; a++) {
    if ((v0.w < fabsf(comp(v0, a))) | (v1.w < fabsf(comp(v1, a))) | (v2.w < fabsf(comp(v2, a)))) {
      float temp[18];
      num = clipPolygon(temp, bary, num, v0.w + comp(v0, a), d1.w + comp(d1, a), d2.w + comp(d2, a));
      num = clipPolygon(bary, temp, num, v0.w - comp(v0, a), d1.w - comp(d1, a), d2.w - comp(d2, a));
    }
  }
  return num;
}

__device__ __forceinline__ float4 lerpClip(const float4& o, const float4& d1, const float4& d2, float b0, float b1) {
  return make_float4(fmaf(d2.x, b1, fmaf(d1.x, b0, o.x)), fmaf(d2.y, b1, fmaf(d1.y, b0, o.y)),
                     fmaf(d2.z, b1, fmaf(d1.z, b0, o.z)), fmaf(d2.w, b1, fmaf(d1.w, b0, o.w)));
}

// Slow path, kept out of line: frustum clip and fan (TriangleSetup.inl:326-412).
// Returns the number of subtriangles written (0 when the slot allocation
// overflowed) and their first slot in *baseOut.
__device__ __noinline__ int setupClipped(const ClipCtx P, int t, int3 vidx, float4 v0, float4 v1, float4 v2,
                                         int* baseOut) {
  float4 od1 = make_float4(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z, v1.w - v0.w);
  float4 od2 = make_float4(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z, v2.w - v0.w);
  float4 ov0 = v0;
  float bary[18];
  int numVerts = clipTriangle(bary, ov0, v1, v2, od1, od2);
  v0 = lerpClip(ov0, od1, od2, bary[0], bary[1]);
  v1 = lerpClip(ov0, od1, od2, bary[2], bary[3]);
  float4 tv1 = v1;
  I2 p0, p1, p2, lo, hi, d1, d2;
  float rcpW[3];
  int32_t area;

  int numSub = 0;
  for (int i = 2; i < numVerts; i++) {
    float4 vv2 = lerpClip(ov0, od1, od2, bary[i * 2], bary[i * 2 + 1]);
    snap(P, v0, v1, vv2, p0, p1, p2, rcpW, lo, hi);
    if (prepare(P, p0, p1, p2, lo, hi, d1, d2, area) == 0) numSub++;
    v1 = vv2;
  }
  P.triSubtris[t] = (uint8_t)numSub;

  int base = t;
  if (numSub > 1) {
    base = (int)atomicAdd(&P.counters->numSubtris, (uint32_t)numSub);
    P.triHeader[t].misc = (uint32_t)base;
    if (base + numSub > P.maxSubtris) {
      atomicOr(&P.counters->overflow, 1u);
      numVerts = 0;
      numSub = 0;
    }
  }
  *baseOut = base;
  v1 = tv1;
  for (int i = 2; i < numVerts; i++) {
    float4 vv2 = lerpClip(ov0, od1, od2, bary[i * 2], bary[i * 2 + 1]);
    snap(P, v0, v1, vv2, p0, p1, p2, rcpW, lo, hi);
    if (prepare(P, p0, p1, p2, lo, hi, d1, d2, area) == 0) {
      setupTriangle(P, &P.triHeader[base], &P.triData[base], vidx, v0, v1, vv2,
                    make_float2(bary[0], bary[1]), make_float2(bary[i * 2 - 2], bary[i * 2 - 1]),
                    make_float2(bary[i * 2], bary[i * 2 + 1]), p0, p1, p2, rcpW, d1, d2, area);
      if (base >= P.numTris) P.subtriTri[base - P.numTris] = (uint32_t)t;
      base++;
    }
    v1 = vv2;
  }
  return numSub;
}

// Outcome of one triangle's setup, as the reference's setup counters
// classify it (TriangleSetup.inl:255-330).
enum SetupOutcome : int { kSetupVisible = 0, kSetupFrustum, kSetupBackface, kSetupBetween, kSetupClipPath,
                          kSetupDropped };

// One triangle: returns its number of subtriangles and their first slot.
// When the triangle is set up without clipping, *hv receives the header's
// vertex words (as written to P.triHeader[t]) and *hvOk is set.  *why
// receives the SetupOutcome.
__device__ __forceinline__ int setupOne(const Params& P, int t, int* base, uint4* hv, bool* hvOk, int* why) {
  *base = t;
  *hvOk = false;
  *why = kSetupDropped;
  int3 vidx;
  const int32_t* ip = P.indices + (int64_t)t * 3;
  vidx.x = ip[0]; vidx.y = ip[1]; vidx.z = ip[2];
  // Indices outside the vertex buffer: the triangle is dropped (the
  // reference reads past the buffer; here the result is defined).
  if ((uint32_t)vidx.x >= (uint32_t)P.numVertices || (uint32_t)vidx.y >= (uint32_t)P.numVertices ||
      (uint32_t)vidx.z >= (uint32_t)P.numVertices) {
    P.triSubtris[t] = 0;
    return 0;
  }
  float4 v0 = *reinterpret_cast<const float4*>(P.vertices + (int64_t)vidx.x * P.vertexStride);
  float4 v1 = *reinterpret_cast<const float4*>(P.vertices + (int64_t)vidx.y * P.vertexStride);
  float4 v2 = *reinterpret_cast<const float4*>(P.vertices + (int64_t)vidx.z * P.vertexStride);

  // Outside view frustum => cull (:262-281).
  if ((v0.w < fabsf(v0.x)) | (v0.w < fabsf(v0.y)) | (v0.w < fabsf(v0.z))) {
    bool out = ((v0.w < +v0.x) & (v1.w < +v1.x) & (v2.w < +v2.x)) |
               ((v0.w < -v0.x) & (v1.w < -v1.x) & (v2.w < -v2.x)) |
               ((v0.w < +v0.y) & (v1.w < +v1.y) & (v2.w < +v2.y)) |
               ((v0.w < -v0.y) & (v1.w < -v1.y) & (v2.w < -v2.y)) |
               ((v0.w < +v0.z) & (v1.w < +v1.z) & (v2.w < +v2.z)) |
               ((v0.w < -v0.z) & (v1.w < -v1.z) & (v2.w < -v2.z));
    if (out) {
      P.triSubtris[t] = 0;
      *why = kSetupFrustum;
      return 0;
    }
  }

  // Inside depth range => snap; inside the guard band => fast path (:285-322).
  if ((v0.w >= fabsf(v0.z)) & (v1.w >= fabsf(v1.z)) & (v2.w >= fabsf(v2.z))) {
    I2 p0, p1, p2, lo, hi, d1, d2;
    float rcpW[3];
    int32_t area;
    snap(P, v0, v1, v2, p0, p1, p2, rcpW, lo, hi);
    int32_t loxy = min(lo.x, lo.y);
    int32_t hixy = max(hi.x, hi.y);
    const int32_t aabbLimit = (1 << (11 + kSubpixelLog2)) - 1;
    if (loxy >= -32768 && hixy <= 32767 && hixy - loxy <= aabbLimit) {
      int res = prepare(P, p0, p1, p2, lo, hi, d1, d2, area);
      P.triSubtris[t] = (res == 0) ? 1 : 0;
      *why = res == 0 ? kSetupVisible : (res == 1 ? kSetupBackface : kSetupBetween);
      if (res == 0) {
        setupTriangle(P, &P.triHeader[t], &P.triData[t], vidx, v0, v1, v2, make_float2(0.0f, 0.0f),
                      make_float2(1.0f, 0.0f), make_float2(0.0f, 1.0f), p0, p1, p2, rcpW, d1, d2, area);
        *hv = make_uint4(((uint32_t)p0.x & 0xFFFFu) | ((uint32_t)p0.y << 16),
                         ((uint32_t)p1.x & 0xFFFFu) | ((uint32_t)p1.y << 16),
                         ((uint32_t)p2.x & 0xFFFFu) | ((uint32_t)p2.y << 16), 0u);
        *hvOk = true;
      }
      return res == 0 ? 1 : 0;
    }
  }
  *why = kSetupClipPath;
  const ClipCtx C = {P.viewportWidth, P.viewportHeight, P.samplesLog2, (int32_t)P.flags, P.shader,
                     P.numTris, P.maxSubtris, P.triSubtris, P.triHeader, P.triData, P.subtriTri, P.counters};
  return setupClipped(C, t, vidx, v0, v1, v2, base);
}

// ---------------------------------------------------------------------------
// Fused setup + bin stage.
//
// One block sets up a chunk of kSetupChunk consecutive triangles (one per
// thread) and immediately bins their subtriangles (128x128-pixel bins,
// bounding-box test).  The chunk's entries go to a region of the bin queue
// reserved with one atomic; inside the region the entries of each bin are
// contiguous and in (triangle, subtriangle) order -- ranks come from per-wave
// 64-bit emit masks and popcounts.  The coarse stage reads bin b as the
// concatenation of its per-chunk segments in chunk order, which is
// submission order, so no global scan is needed.  (The reference streams
// triangles into linked 512-entry segments per (bin, stream) and merges 16
// streams, BinRaster.inl:19-469, CoarseRaster.inl:170-257.)
// ---------------------------------------------------------------------------

constexpr int kSetupChunk = kSetupChunkTris;
constexpr int kSetupWaves = kSetupChunk / 64;

struct BinBox { int32_t x0, y0, x1, y1; };   // inclusive; empty when x0 > x1

__device__ __forceinline__ BinBox binBoxOf(const Params& P, uint4 h) {
  const int32_t ox = P.viewportWidth << (kSubpixelLog2 - 1), oy = P.viewportHeight << (kSubpixelLog2 - 1);
  const int32_t x0 = (int32_t)(int16_t)(h.x & 0xFFFFu), y0 = (int32_t)(int16_t)(h.x >> 16);
  const int32_t x1 = (int32_t)(int16_t)(h.y & 0xFFFFu), y1 = (int32_t)(int16_t)(h.y >> 16);
  const int32_t x2 = (int32_t)(int16_t)(h.z & 0xFFFFu), y2 = (int32_t)(int16_t)(h.z >> 16);
  const int32_t minx = min(min(x0, x1), x2) + ox, maxx = max(max(x0, x1), x2) + ox;
  const int32_t miny = min(min(y0, y1), y2) + oy, maxy = max(max(y0, y1), y2) + oy;
  BinBox b;
  const int sh = kSubpixelLog2 + kTileLog2 + P.binLog2;
  if (maxx < 0 || maxy < 0 || minx >= (P.widthPixels << kSubpixelLog2) || miny >= (P.heightPixels << kSubpixelLog2)) {
    b.x0 = 1; b.x1 = 0; b.y0 = 1; b.y1 = 0;
    return b;
  }
  b.x0 = max(minx >> sh, 0); b.x1 = min(maxx >> sh, P.widthBins - 1);
  b.y0 = max(miny >> sh, 0); b.y1 = min(maxy >> sh, P.heightBins - 1);
  return b;
}

__device__ __forceinline__ uint64_t laneLt() {
  const uint32_t lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Exclusive scan of one value per thread over the whole block.
template <int NT>
__device__ __forceinline__ uint32_t blockScan(uint32_t v, uint32_t* sums, uint32_t& total) {
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

// Setup and binning of chunk `chunk` of one draw (one 1024-lane workgroup).
__device__ __forceinline__ void setupBinBody(const Params& P, const int chunk) {
  __shared__ uint32_t cnt[kMaxBins];
  __shared__ uint32_t run[kMaxBins];
  __shared__ uint64_t emitMask[kSetupWaves][kMaxBins];
  __shared__ uint32_t emitPre[kSetupWaves][kMaxBins];
  __shared__ uint32_t list[kSetupChunk];
  __shared__ uint32_t sums[kSetupWaves + 1];
  __shared__ uint32_t sBase;
  __shared__ uint32_t sProf[8];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = chunk * kSetupChunk + threadIdx.x;

  // ---- setup ---------------------------------------------------------------
  int base = t, k = 0, why = kSetupDropped;
  uint4 hvOwn = make_uint4(0u, 0u, 0u, 0u);   // header vertex words kept from setup (no re-read)
  bool hvOwnOk = false;
  if (t < P.numTris) k = setupOne(P, t, &base, &hvOwn, &hvOwnOk, &why);
  if (P.prof != nullptr) {
    // Profiling draws: per-outcome triangle counts of the chunk, then its
    // triangles with at least one subtriangle.
    if (threadIdx.x < 8) sProf[threadIdx.x] = 0;
    __syncthreads();
    const bool valid = t < P.numTris;
    const uint32_t nValid = (uint32_t)__popcll(__ballot(valid));
    const uint32_t nFr = (uint32_t)__popcll(__ballot(valid && why == kSetupFrustum));
    const uint32_t nBf = (uint32_t)__popcll(__ballot(valid && why == kSetupBackface));
    const uint32_t nBt = (uint32_t)__popcll(__ballot(valid && why == kSetupBetween));
    const uint32_t nCl = (uint32_t)__popcll(__ballot(valid && why == kSetupClipPath));
    const uint32_t nVis = (uint32_t)__popcll(__ballot(valid && k > 0));
    if (lane == 0) {
      atomicAdd(&sProf[0], nValid);
      atomicAdd(&sProf[1], nFr);
      atomicAdd(&sProf[2], nBf);
      atomicAdd(&sProf[3], nBt);
      atomicAdd(&sProf[4], nCl);
      atomicAdd(&sProf[5], nVis);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      profAdd(P.prof, kProfSetupViewportCull, 100ull * sProf[1], sProf[0]);
      profAdd(P.prof, kProfSetupBackfaceCull, 100ull * sProf[2], sProf[0]);
      profAdd(P.prof, kProfSetupBetweenPixelsCull, 100ull * sProf[3], sProf[0]);
      profAdd(P.prof, kProfSetupClipped, 100ull * sProf[4], sProf[0]);
      profAdd(P.prof, kProfSetupSamplesPerTri, 0, sProf[5]);
    }
  }

  for (int i = threadIdx.x; i < P.numBins; i += kSetupChunk) cnt[i] = 0;
  // Window = up to kSetupChunk consecutive subtriangles of the chunk in
  // submission order; a chunk without clipped triangles is a single window.
  const bool compact = __syncthreads_or(k > 1) != 0;
  uint32_t total;
  const uint32_t pos = blockScan<kSetupChunk>((uint32_t)k, sums, total);
  const uint32_t numWin = (total + kSetupChunk - 1) / kSetupChunk;
  if (P.chunkSubtris != nullptr && threadIdx.x == 0) P.chunkSubtris[chunk] = total;

  // Emit masks of one window: bit `lane` of emitMask[wave][bin] is set when
  // the window entry held by this thread touches the bin.
  auto buildMasks = [&](uint32_t w, uint32_t& slot, BinBox& bb, uint4& hv) -> bool {
    for (int i = threadIdx.x; i < kSetupWaves * P.numBins; i += kSetupChunk)
      emitMask[i / P.numBins][i % P.numBins] = 0ull;
    if (!compact) {
      // fast path: entry of this thread is its own subtriangle (k <= 1)
      slot = (uint32_t)base;
      bb.x0 = 1; bb.x1 = 0; bb.y0 = 1; bb.y1 = 0;
      __syncthreads();
      if (k == 1) {
        hv = hvOwnOk ? hvOwn : *reinterpret_cast<const uint4*>(&P.triHeader[slot]);
        bb = binBoxOf(P, hv);
      }
    } else {
      for (int s = 0; s < k; s++) {
        const uint32_t e = pos + (uint32_t)s;
        if (e >= w * kSetupChunk && e < (w + 1) * kSetupChunk) list[e - w * kSetupChunk] = (uint32_t)base + s;
      }
      __syncthreads();
      const uint32_t e = w * kSetupChunk + threadIdx.x;
      bb.x0 = 1; bb.x1 = 0; bb.y0 = 1; bb.y1 = 0;
      slot = 0;
      if (e < total) {
        slot = list[threadIdx.x];
        hv = *reinterpret_cast<const uint4*>(&P.triHeader[slot]);
        bb = binBoxOf(P, hv);
      }
    }
    const uint64_t bit = 1ull << lane;
    for (int y = bb.y0; y <= bb.y1; y++)
      for (int x = bb.x0; x <= bb.x1; x++)
        atomicOr((unsigned long long*)&emitMask[wave][x + y * P.widthBins], (unsigned long long)bit);
    __syncthreads();
    return true;
  };

  // ---- pass 1: per-bin counts ---------------------------------------------
  uint32_t slot0 = 0;
  BinBox bb0;
  uint4 hv0 = make_uint4(0, 0, 0, 0);
  for (uint32_t w = 0; w < numWin; w++) {
    buildMasks(w, slot0, bb0, hv0);
    for (int i = threadIdx.x; i < P.numBins; i += kSetupChunk) {
      uint32_t acc = cnt[i];
#pragma unroll
      for (int q = 0; q < kSetupWaves; q++) acc += (uint32_t)__popcll(emitMask[q][i]);
      cnt[i] = acc;
    }
    if (P.prof != nullptr) {
      // Subtriangles whose box lies in a single bin.
      const uint32_t n1 = (uint32_t)__popcll(__ballot(bb0.x0 == bb0.x1 && bb0.y0 == bb0.y1));
      if (lane == 0 && n1) atomicAdd(&sProf[6], n1);
    }
    __syncthreads();
  }

  // ---- reserve the chunk's region; publish counts and offsets ---------------
  uint32_t tot2;
  const uint32_t c = (threadIdx.x < (unsigned)P.numBins) ? cnt[threadIdx.x] : 0u;
  const uint32_t off = blockScan<kSetupChunk>(c, sums, tot2);
  if (P.prof != nullptr && threadIdx.x == 0) {
    profAdd(P.prof, kProfBinTrisPerChunk, total, 1);
    profAdd(P.prof, kProfBinTriBBArea, tot2, total);
    profAdd(P.prof, kProfBinTriSingleBin, 100ull * sProf[6], total);
  }
  if (threadIdx.x == 0) {
    uint32_t b = tot2 ? atomicAdd(&P.counters->binEntries, tot2) : 0u;
    if (tot2 && (uint64_t)b + tot2 > (uint64_t)P.maxBinEntries) {
      atomicOr(&P.counters->overflow, 2u);
      b = 0xFFFFFFFFu;
    }
    sBase = b;
    P.chunkBase[chunk] = b;
  }
  if (threadIdx.x < (unsigned)P.numBins) {
    P.binCnt[(size_t)threadIdx.x * P.numChunks + chunk] = c;
    P.binOff[(size_t)threadIdx.x * P.numChunks + chunk] = off;
    run[threadIdx.x] = off;
  }
  __syncthreads();
  const uint32_t regionBase = sBase;
  if (regionBase == 0xFFFFFFFFu || tot2 == 0) return;

  // ---- pass 2: write the entries in order ------------------------------------
  for (uint32_t w = 0; w < numWin; w++) {
    uint32_t slot = slot0;
    BinBox bb = bb0;
    uint4 hv = hv0;
    if (numWin > 1) buildMasks(w, slot, bb, hv);
    for (int i = threadIdx.x; i < P.numBins; i += kSetupChunk) {
      uint32_t acc = run[i];
#pragma unroll
      for (int q = 0; q < kSetupWaves; q++) {
        emitPre[q][i] = acc;
        acc += (uint32_t)__popcll(emitMask[q][i]);
      }
      run[i] = acc;
    }
    __syncthreads();
    const uint64_t lt = laneLt();
    for (int y = bb.y0; y <= bb.y1; y++)
      for (int x = bb.x0; x <= bb.x1; x++) {
        const int bin = x + y * P.widthBins;
        const uint32_t p = regionBase + emitPre[wave][bin] + (uint32_t)__popcll(emitMask[wave][bin] & lt);
        if (p < (uint32_t)P.maxBinEntries) P.binList[p] = make_uint4(slot, hv.x, hv.y, hv.z);
      }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kSetupChunk, 8) setupBinKernel(const Params P) { setupBinBody(P, blockIdx.x); }

// Batched draws: one workgroup per chunk of every view; blockStart[v] is the
// first workgroup of view v (numViews + 1 entries).
__global__ void __launch_bounds__(kSetupChunk, 8) setupBinBatchKernel(const Params* __restrict__ views,
                                                                   const int32_t* __restrict__ blockStart,
                                                                   const int32_t* __restrict__ order,
                                                                   int numViews) {
  const int k = findView(blockStart, numViews, (int)blockIdx.x);
  setupBinBody(views[order[k]], (int)blockIdx.x - blockStart[k]);
}

void launchSetup(const Params& P, hipStream_t stream) {
  if (P.numTris <= 0) return;
  hipLaunchKernelGGL(setupBinKernel, dim3(P.numChunks), dim3(kSetupChunk), 0, stream, P);
}

void launchSetupBatch(const Params* views, const int32_t* blockStart, const int32_t* order, int numViews,
                      int totalBlocks, hipStream_t stream) {
  if (totalBlocks <= 0) return;
  hipLaunchKernelGGL(setupBinBatchKernel, dim3(totalBlocks), dim3(kSetupChunk), 0, stream, views, blockStart,
                     order, numViews);
}

}  // namespace crhip
