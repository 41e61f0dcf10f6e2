/*
 * CPU oracle for the four-stage rasterization path: test infrastructure only
 * (see cr_oracle.h for the reference files each part restates).
 *
 * Arithmetic policy.  The reference's device code is the reference of record
 * where it differs from its host emulation (the emulation uses F64 plane
 * equations, CudaRaster.cpp:669-688, and floor(x+0.5) snapping,
 * CudaRaster.cpp:701-703).  Float expressions follow the device source with
 * the multiply-add contraction that nvcc applies by default (-fmad=true):
 * an `a*b + c` whose product feeds an add is evaluated as one fused
 * multiply-add, and for `a*b + c*d` the left product is the one fused (the
 * LLVM DAG-combine rule nvcc's NVVM back end uses).  Every fused site is an
 * explicit fmaf() below; everything else is compiled with
 * -ffp-contract=off.  PTX conversions are restated exactly:
 *   cvt.rni.sat.{s32,u32}.f32 -> round-to-nearest-even, saturate, NaN -> 0
 *   (U32)float in device code -> cvt.rzi.u32.f32: truncate, saturate, NaN -> 0
 * Integer arithmetic wraps modulo 2^32 / 2^64 as in PTX; it is done in
 * unsigned types here so that C's signed-overflow rules never apply.
 */
#define _GNU_SOURCE
#include "cr_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* Constants.hpp:21-84 */
#define SUBPIXEL_LOG2   4
#define TILE_LOG2       3
#define MAXVIEWPORT     2048
#define LERP_ERROR(s)   (2200u << (s))
#define DEPTH_MIN       LERP_ERROR(3)
#define DEPTH_MAX       (0xFFFFFFFFu - LERP_ERROR(3))
#define BARY_MAX        ((1 << (30 - SUBPIXEL_LOG2)) - 1)
#define FLIP_Y          2
#define FLIP_X          3
#define SWAP_XY         4
#define COMPL           5

/* Util.hpp:27-36: c_msaaPatterns[log2 N][sampleY] = sampleX. */
static const int32_t kPattern[4][8] = {
    {0},
    {0, 1},
    {1, 3, 0, 2},
    {7, 2, 4, 0, 6, 3, 1, 5},
};

/* ------------------------------------------------------------------ */
/* Bit-level helpers                                                   */
/* ------------------------------------------------------------------ */

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline int32_t  imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline int32_t  imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t  wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t  wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t  wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
static inline int64_t  wmul64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
static inline int64_t  wsub64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
static inline int64_t  wadd64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
/* Arithmetic shifts (PTX shr.s32 / shr.s64 on in-range amounts). */
static inline int32_t  asr32(int32_t a, int s) { return a >= 0 ? (a >> s) : ~((~a) >> s); }
static inline int64_t  asr64(int64_t a, int s) { return a >= 0 ? (a >> s) : ~((~a) >> s); }
static inline int32_t  shl32(int32_t a, int s) { return (int32_t)((uint32_t)a << s); }
/* abs() on S32 with PTX wrap (abs(INT_MIN) == INT_MIN). */
static inline uint32_t wabs32(int32_t a) { return a < 0 ? (uint32_t)0 - (uint32_t)a : (uint32_t)a; }

/* cvt.rni.sat.s32.f32 (Util.hpp:88) */
static inline int32_t cvt_rni_sat_s32(float a)
{
    if (isnan(a)) return 0;
    float r = rintf(a);
    if (r >= 2147483648.0f) return INT32_MAX;
    if (r < -2147483648.0f) return INT32_MIN;
    return (int32_t)r;
}

/* cvt.rni.sat.u32.f32 (Util.hpp:89) */
static inline uint32_t cvt_rni_sat_u32(float a)
{
    if (isnan(a)) return 0;
    float r = rintf(a);
    if (r >= 4294967296.0f) return 0xFFFFFFFFu;
    if (r <= 0.0f) return 0;
    return (uint32_t)r;
}

/* (U32)float in device code: cvt.rzi.u32.f32, which saturates in PTX. */
static inline uint32_t cvt_rzi_sat_u32(float a)
{
    if (isnan(a) || a <= 0.0f) return 0;
    if (a >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)a;
}

/* ------------------------------------------------------------------ */
/* Quantisation and blending                                           */
/* ------------------------------------------------------------------ */

/* One channel of toABGR (Util.inl:30-38): fma.rm(c, 255*2^24, 2^23) then
 * cvt.rmi.sat.u32 and the top byte.  Rounding down to float and flooring
 * cannot cross a multiple of 2^24 (those are floats), so the channel is
 * floor(255*c + 1/2) of the exact product, saturated to [0,255]; NaN -> 0.
 * 255*c + 1/2 is exact in double for every float c that does not saturate. */
static inline uint32_t abgr_channel(float c)
{
    if (isnan(c)) return 0;
    double x = (double)c * 255.0 + 0.5;
    if (!(x >= 0.0)) return 0;
    if (x >= 256.0) return 255;
    uint32_t v = (uint32_t)floor(x);
    return v > 255 ? 255 : v;
}

uint32_t cro_to_abgr(const float c[4])
{
    return abgr_channel(c[0]) | (abgr_channel(c[1]) << 8) |
           (abgr_channel(c[2]) << 16) | (abgr_channel(c[3]) << 24);
}

/* Vec4f::toABGR (Math.cpp:41-48): clamp to [0,1], scale by 2^56 in F64,
 * truncate to U64, times 255, round half up through the >>55, +1, >>1. */
static inline uint32_t host_channel(float c)
{
    float cl = c;
    if (!(cl >= 0.0f)) cl = 0.0f;   /* FW::clamp = min(max(v,lo),hi); NaN -> lo */
    if (cl > 1.0f) cl = 1.0f;
    uint64_t q = (uint64_t)((double)cl * 72057594037927936.0);
    return (uint32_t)(((q * 255u) >> 55) + 1u) >> 1;
}

uint32_t cro_to_abgr_host(const float c[4])
{
    return host_channel(c[0]) | (host_channel(c[1]) << 8) |
           (host_channel(c[2]) << 16) | (host_channel(c[3]) << 24);
}

/* blendABGR / blendABGRClamp (Util.inl:42-60) with the stock factors of
 * BlendSrcOver (PixelPipe.inl:73-77) and BlendAdditive (:81-85). */
uint32_t cro_blend(int32_t mode, uint32_t src, uint32_t dst)
{
    if (mode == CRO_BLEND_REPLACE) return src;
    if (mode == CRO_BLEND_DEPTHONLY) return dst;
    uint32_t out = 0;
    for (int ch = 0; ch < 4; ch++) {
        uint32_t s = (src >> (8 * ch)) & 0xFF;
        uint32_t d = (dst >> (8 * ch)) & 0xFF;
        uint32_t x;
        if (mode == CRO_BLEND_SRCOVER) {
            uint32_t a = src >> 24;
            x = s * a + d * (255u - a);
        } else {
            x = s * 255u + d * 255u;
            if (x > 255u * 255u) x = 255u * 255u;
        }
        out |= ((x * 0x010101u + 0x800000u) >> 24) << (8 * ch);
    }
    return out;
}

/* ------------------------------------------------------------------ */
/* Depth encoding                                                      */
/* ------------------------------------------------------------------ */

uint32_t cro_encode_depth(uint32_t depth)
{
    double v = (double)depth;
    v /= 65536.0 * 65536.0 - 1.0;
    v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    v *= (double)(DEPTH_MAX - DEPTH_MIN);
    v += (double)DEPTH_MIN;
    return (uint32_t)v;
}

uint32_t cro_decode_depth(uint32_t depth)
{
    double v = (double)depth;
    v -= (double)DEPTH_MIN;
    v *= 1.0 / (double)(DEPTH_MAX - DEPTH_MIN);
    v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    v *= 65536.0 * 65536.0 - 1.0;
    return (uint32_t)v;
}

/* CudaRaster::deferredClear (CudaRaster.cpp:178):
 * encodeDepth((U32)min((U64)(depth * 2^32), (U64)U32_MAX)). */
uint32_t cro_clear_depth(float depth)
{
    double d = (double)depth * 4294967296.0;
    uint64_t q;
    if (isnan(d) || d <= 0.0) q = 0;
    else if (d >= 18446744073709551616.0) q = UINT64_MAX;
    else q = (uint64_t)d;
    if (q > 0xFFFFFFFFull) q = 0xFFFFFFFFull;
    return cro_encode_depth((uint32_t)q);
}

/* ------------------------------------------------------------------ */
/* MSAA                                                                */
/* ------------------------------------------------------------------ */

int32_t cro_msaa_pattern(int32_t s, int32_t i)
{
    if (s < 0 || s > 3 || i < 0 || i >= (1 << s)) return -1;
    return kPattern[s][i];
}

/* selectMSAACentroid (Util.hpp:182-203) */
int32_t cro_select_centroid(int32_t s, uint32_t mask)
{
    int n = 1 << s;
    if (mask == 0 || mask == (1u << n) - 1u) return -1;
    int best = -1, bestDist = INT32_MAX;
    for (int i = 0; i < n; i++) {
        if (mask & (1u << i)) {
            int dx = kPattern[s][i] * 2 + 1 - n;
            int dy = i * 2 + 1 - n;
            int dist = dx * dx + dy * dy;
            if (dist < bestDist) { best = i; bestDist = dist; }
        }
    }
    return best;
}

/* cover8x8_selectFlips (Util.hpp:207-217) */
static uint32_t select_flips(int32_t dx, int32_t dy)
{
    uint32_t f = 0;
    if (dy > 0 || (dy == 0 && dx <= 0))
        f ^= (1u << FLIP_X) ^ (1u << FLIP_Y) ^ (1u << COMPL);
    if (dx > 0)
        f ^= (1u << FLIP_X) ^ (1u << FLIP_Y);
    if (wabs32(dx) < wabs32(dy))
        f ^= (1u << SWAP_XY) ^ (1u << FLIP_Y);
    return f;
}

/* ------------------------------------------------------------------ */
/* Plane equations (GPU variant, Util.inl:65-91)                       */
/* ------------------------------------------------------------------ */

void cro_setup_pleq(const float values[3], int32_t v0x, int32_t v0y,
                    int32_t d1x, int32_t d1y, int32_t d2x, int32_t d2y,
                    float areaRcp, int32_t samplesLog2, uint32_t out[3])
{
    float mx = fmaxf(fmaxf(values[0], values[1]), values[2]);
    int32_t sh = asr32((int32_t)fbits(mx), 23) - (127 + 22);
    sh = imin(imax(sh, 0), 8);
    int32_t t0 = (int32_t)(cvt_rzi_sat_u32(values[0]) >> sh);
    int32_t t1 = wsub((int32_t)(cvt_rzi_sat_u32(values[1]) >> sh), t0);
    int32_t t2 = wsub((int32_t)(cvt_rzi_sat_u32(values[2]) >> sh), t0);

    uint32_t rcpMant = (fbits(areaRcp) & 0x007FFFFFu) | 0x00800000u;
    int32_t rcpShift = (23 + 127) - asr32((int32_t)fbits(areaRcp), 23);

    int64_t xc = wmul64(wsub64((int64_t)t1 * d2y, (int64_t)t2 * d1y), (int64_t)rcpMant);
    int64_t yc = wmul64(wsub64((int64_t)t2 * d1x, (int64_t)t1 * d2x), (int64_t)rcpMant);
    int shp = rcpShift - (sh + SUBPIXEL_LOG2 - samplesLog2);
    out[0] = (uint32_t)asr64(xc, shp);
    out[1] = (uint32_t)asr64(yc, shp);

    int sc = SUBPIXEL_LOG2 - samplesLog2;
    int32_t minx = imin(imin(d1x, d2x), 0), maxx = imax(imax(d1x, d2x), 0);
    int32_t miny = imin(imin(d1y, d2y), 0), maxy = imax(imax(d1y, d2y), 0);
    int32_t centerX = asr32(wadd(wadd(wmul(v0x, 2), minx), maxx), sc + 1);
    int32_t centerY = asr32(wadd(wadd(wmul(v0y, 2), miny), maxy), sc + 1);
    int32_t vcx = wsub(v0x, shl32(centerX, sc));
    int32_t vcy = wsub(v0y, shl32(centerY, sc));

    uint32_t z = (uint32_t)t0 << sh;
    int64_t cross = wadd64(wmul64(asr64(xc, 13), vcx), wmul64(asr64(yc, 13), vcy));
    z -= (uint32_t)asr64(cross, rcpShift - (sh + 13));
    z -= out[0] * (uint32_t)centerX + out[1] * (uint32_t)centerY;
    out[2] = z;
}

/* ------------------------------------------------------------------ */
/* Clipping (Util.hpp:221-286)                                         */
/* ------------------------------------------------------------------ */

static int clip_polygon(float* out, const float* in, int numIn, float v0, float v1, float v2)
{
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

/* clipTriangleWithFrustum; d1 = v1 - v0, d2 = v2 - v0 componentwise. */
static int clip_triangle(float* bary, const float* v0, const float* v1, const float* v2,
                         const float* d1, const float* d2)
{
    int num = 3;
    bary[0] = 0.0f; bary[1] = 0.0f;
    bary[2] = 1.0f; bary[3] = 0.0f;
    bary[4] = 0.0f; bary[5] = 1.0f;
    for (int axis = 0; axis < 3; axis++) {
        if ((v0[3] < fabsf(v0[axis])) | (v1[3] < fabsf(v1[axis])) | (v2[3] < fabsf(v2[axis]))) {
            float temp[18];
            num = clip_polygon(temp, bary, num, v0[3] + v0[axis], d1[3] + d1[axis], d2[3] + d2[axis]);
            num = clip_polygon(bary, temp, num, v0[3] - v0[axis], d1[3] - d1[axis], d2[3] - d2[axis]);
        }
    }
    return num;
}

int32_t cro_clip_triangle(float bary[18], const float v0[4], const float v1[4], const float v2[4])
{
    float d1[4], d2[4];
    for (int k = 0; k < 4; k++) { d1[k] = v1[k] - v0[k]; d2[k] = v2[k] - v0[k]; }
    return clip_triangle(bary, v0, v1, v2, d1, d2);
}

/* ------------------------------------------------------------------ */
/* Triangle setup (TriangleSetup.inl)                                  */
/* ------------------------------------------------------------------ */

typedef struct { int32_t x, y; } i2;

/* snapTriangle (TriangleSetup.inl:19-31) */
static void snap(const cro_pipe* P, const float* v0, const float* v1, const float* v2,
                 i2* p0, i2* p1, i2* p2, float rcpW[3], i2* lo, i2* hi)
{
    float sx = (float)(P->width << (SUBPIXEL_LOG2 - 1));
    float sy = (float)(P->height << (SUBPIXEL_LOG2 - 1));
    rcpW[0] = 1.0f / v0[3];
    rcpW[1] = 1.0f / v1[3];
    rcpW[2] = 1.0f / v2[3];
    p0->x = cvt_rni_sat_s32(v0[0] * rcpW[0] * sx); p0->y = cvt_rni_sat_s32(v0[1] * rcpW[0] * sy);
    p1->x = cvt_rni_sat_s32(v1[0] * rcpW[1] * sx); p1->y = cvt_rni_sat_s32(v1[1] * rcpW[1] * sy);
    p2->x = cvt_rni_sat_s32(v2[0] * rcpW[2] * sx); p2->y = cvt_rni_sat_s32(v2[1] * rcpW[2] * sy);
    lo->x = imin(imin(p0->x, p1->x), p2->x); lo->y = imin(imin(p0->y, p1->y), p2->y);
    hi->x = imax(imax(p0->x, p1->x), p2->x); hi->y = imax(imax(p0->y, p1->y), p2->y);
}

/* prepareTriangle (TriangleSetup.inl:38-100): 0 visible, 1 backfacing,
 * 2 between samples. */
static int prepare(const cro_pipe* P, i2 p0, i2 p1, i2 p2, i2 lo, i2 hi,
                   i2* d1, i2* d2, int32_t* area)
{
    int s = P->samplesLog2;
    d1->x = wsub(p1.x, p0.x); d1->y = wsub(p1.y, p0.y);
    d2->x = wsub(p2.x, p0.x); d2->y = wsub(p2.y, p0.y);
    *area = wsub(wmul(d1->x, d2->y), wmul(d1->y, d2->x));
    if (*area <= 0) return 1;

    int32_t ss = 1 << (SUBPIXEL_LOG2 - s);
    int32_t biasX = (P->width << (SUBPIXEL_LOG2 - 1)) - (ss >> 1);
    int32_t biasY = (P->height << (SUBPIXEL_LOG2 - 1)) - (ss >> 1);
    int32_t lox = wadd(wadd(lo.x, ss - 1), biasX) & -ss;
    int32_t loy = wadd(wadd(lo.y, ss - 1), biasY) & -ss;
    int32_t hix = wadd(hi.x, biasX) & -ss;
    int32_t hiy = wadd(hi.y, biasY) & -ss;
    if (lox > hix || loy > hiy) return 2;

    int32_t diff = wsub(wadd(hix, hiy), wadd(lox, loy));
    if (diff <= ss) {
        i2 t0 = {wsub(wadd(p0.x, biasX), lox), wsub(wadd(p0.y, biasY), loy)};
        i2 t1 = {wsub(wadd(p1.x, biasX), lox), wsub(wadd(p1.y, biasY), loy)};
        i2 t2 = {wsub(wadd(p2.x, biasX), lox), wsub(wadd(p2.y, biasY), loy)};
        int32_t e0 = wsub(wmul(t0.x, t1.y), wmul(t0.y, t1.x));
        int32_t e1 = wsub(wmul(t1.x, t2.y), wmul(t1.y, t2.x));
        int32_t e2 = wsub(wmul(t2.x, t0.y), wmul(t2.y, t0.x));
        if (e0 < 0 || e1 < 0 || e2 < 0) {
            if (diff == 0) return 2;
            t0.x = wsub(wadd(p0.x, biasX), hix); t0.y = wsub(wadd(p0.y, biasY), hiy);
            t1.x = wsub(wadd(p1.x, biasX), hix); t1.y = wsub(wadd(p1.y, biasY), hiy);
            t2.x = wsub(wadd(p2.x, biasX), hix); t2.y = wsub(wadd(p2.y, biasY), hiy);
            e0 = wsub(wmul(t0.x, t1.y), wmul(t0.y, t1.x));
            e1 = wsub(wmul(t1.x, t2.y), wmul(t1.y, t2.x));
            e2 = wsub(wmul(t2.x, t0.y), wmul(t2.y, t0.x));
            if (e0 < 0 || e1 < 0 || e2 < 0) return 2;
        }
    }
    return 0;
}

/* setupTriangle (TriangleSetup.inl:104-215) */
static void setup_tri(const cro_pipe* P, cro_header* th, cro_data* td, const int32_t vidx[3],
                      const float* v0, const float* v1, const float* v2,
                      const float b0[2], const float b1[2], const float b2[2],
                      i2 p0, i2 p1, i2 p2, const float rcpW[3], i2 d1, i2 d2, int32_t area)
{
    int s = P->samplesLog2;
    uint32_t flags = P->renderModeFlags;
    float areaRcp = 0.0f;
    i2 wv0 = {0, 0};
    memset(td, 0, sizeof(*td));

    if (flags & (CRO_FLAG_DEPTH | CRO_FLAG_LERP)) {
        areaRcp = 1.0f / (float)area;
        wv0.x = wadd(p0.x, P->width << (SUBPIXEL_LOG2 - 1));
        wv0.y = wadd(p0.y, P->height << (SUBPIXEL_LOG2 - 1));
    }

    uint32_t zmin = 0, zslope = 0;
    uint32_t zp[3] = {0, 0, 0};
    if (flags & CRO_FLAG_DEPTH) {
        float zcoef = (float)(DEPTH_MAX - DEPTH_MIN) * 0.5f;
        float zbias = (float)(DEPTH_MAX + DEPTH_MIN) * 0.5f;
        float zvert[3];
        zvert[0] = fmaf(v0[2] * zcoef, rcpW[0], zbias);
        zvert[1] = fmaf(v1[2] * zcoef, rcpW[1], zbias);
        zvert[2] = fmaf(v2[2] * zcoef, rcpW[2], zbias);
        int32_t zv0x = wv0.x - (1 << (SUBPIXEL_LOG2 - s - 1));
        int32_t zv0y = wv0.y - (1 << (SUBPIXEL_LOG2 - s - 1));
        cro_setup_pleq(zvert, zv0x, zv0y, d1.x, d1.y, d2.x, d2.y, areaRcp, s, zp);
        zmin = cvt_rni_sat_u32(fminf(fminf(zvert[0], zvert[1]), zvert[2]) - (float)LERP_ERROR(s));
        if (s != 0) {
            int32_t zy = (int32_t)zp[1];
            uint32_t tmp = wabs32((int32_t)zp[0]) + wabs32(imax(zy, -INT32_MAX));
            int sh = imax(s - 1, 0);
            zslope = tmp << sh;
            if ((zslope >> sh) != tmp) zslope = 0xFFFFFFFFu;
        }
        td->zx = zp[0]; td->zy = zp[1]; td->zb = zp[2]; td->zslope = zslope;
    }

    if (flags & CRO_FLAG_LERP) {
        float wcoef = fminf(fminf(v0[3], v1[3]), v2[3]) * (float)BARY_MAX;
        float wv[3] = {wcoef * rcpW[0], wcoef * rcpW[1], wcoef * rcpW[2]};
        float uv[3] = {b0[0] * wv[0], b1[0] * wv[1], b2[0] * wv[2]};
        float vv[3] = {b0[1] * wv[0], b1[1] * wv[1], b2[1] * wv[2]};
        uint32_t wp[3], up[3], vp[3];
        cro_setup_pleq(wv, wv0.x, wv0.y, d1.x, d1.y, d2.x, d2.y, areaRcp, s + 1, wp);
        cro_setup_pleq(uv, wv0.x, wv0.y, d1.x, d1.y, d2.x, d2.y, areaRcp, s + 1, up);
        cro_setup_pleq(vv, wv0.x, wv0.y, d1.x, d1.y, d2.x, d2.y, areaRcp, s + 1, vp);
        td->wx = (int32_t)wp[0]; td->wy = (int32_t)wp[1]; td->wb = (int32_t)wp[2];
        td->ux = (int32_t)up[0]; td->uy = (int32_t)up[1]; td->ub = (int32_t)up[2];
        td->vx = (int32_t)vp[0]; td->vy = (int32_t)vp[1]; td->vb = (int32_t)vp[2];
    }
    td->vi0 = (uint32_t)vidx[0]; td->vi1 = (uint32_t)vidx[1]; td->vi2 = (uint32_t)vidx[2];

    uint32_t f01 = select_flips(d1.x, d1.y);
    uint32_t f12 = select_flips(wsub(d2.x, d1.x), wsub(d2.y, d1.y));
    uint32_t f20 = select_flips(-d2.x, -d2.y);
    th->v0x = (int16_t)p0.x; th->v0y = (int16_t)p0.y;
    th->v1x = (int16_t)p1.x; th->v1y = (int16_t)p1.y;
    th->v2x = (int16_t)p2.x; th->v2y = (int16_t)p2.y;
    th->misc = (zmin & 0xFFFFF000u) | (f01 << 6) | (f12 << 2) | (f20 >> 2);
}

/* triangleSetupImpl (TriangleSetup.inl:219-417) for one triangle. */
int32_t cro_setup_triangle(const cro_pipe* P, const float iv0[4], const float iv1[4],
                           const float iv2[4], const int32_t vidx[3],
                           cro_header hdr[7], cro_data data[7], int32_t* cull)
{
    float v0[4], v1[4], v2[4];
    memcpy(v0, iv0, 16); memcpy(v1, iv1, 16); memcpy(v2, iv2, 16);
    i2 p0, p1, p2, lo, hi, d1, d2;
    float rcpW[3];
    int32_t area;
    static const float B0[2] = {0.0f, 0.0f}, B1[2] = {1.0f, 0.0f}, B2[2] = {0.0f, 1.0f};

    /* Outside view frustum => cull (:262-281). */
    if ((v0[3] < fabsf(v0[0])) | (v0[3] < fabsf(v0[1])) | (v0[3] < fabsf(v0[2]))) {
        int out = 0;
        for (int a = 0; a < 3; a++) {
            out |= (v0[3] < +v0[a]) & (v1[3] < +v1[a]) & (v2[3] < +v2[a]);
            out |= (v0[3] < -v0[a]) & (v1[3] < -v1[a]) & (v2[3] < -v2[a]);
        }
        if (out) { *cull = 1; return 0; }
    }

    /* Inside depth range => fast path (:285-322). */
    if ((v0[3] >= fabsf(v0[2])) & (v1[3] >= fabsf(v1[2])) & (v2[3] >= fabsf(v2[2]))) {
        snap(P, v0, v1, v2, &p0, &p1, &p2, rcpW, &lo, &hi);
        int32_t loxy = imin(lo.x, lo.y);
        int32_t hixy = imax(hi.x, hi.y);
        int32_t aabbLimit = (1 << (11 + SUBPIXEL_LOG2)) - 1;
        if (loxy >= -32768 && hixy <= 32767 && hixy - loxy <= aabbLimit) {
            int res = prepare(P, p0, p1, p2, lo, hi, &d1, &d2, &area);
            if (res != 0) { *cull = (res == 1) ? 2 : 3; return 0; }
            setup_tri(P, &hdr[0], &data[0], vidx, v0, v1, v2, B0, B1, B2,
                      p0, p1, p2, rcpW, d1, d2, area);
            *cull = 0;
            return 1;
        }
    }

    /* Clip to view frustum (:326-412). */
    float ov0[4], od1[4], od2[4], bary[18];
    for (int k = 0; k < 4; k++) { ov0[k] = v0[k]; od1[k] = v1[k] - v0[k]; od2[k] = v2[k] - v0[k]; }
    int numVerts = clip_triangle(bary, ov0, v1, v2, od1, od2);
    for (int k = 0; k < 4; k++) {
        v0[k] = fmaf(od2[k], bary[1], fmaf(od1[k], bary[0], ov0[k]));
        v1[k] = fmaf(od2[k], bary[3], fmaf(od1[k], bary[2], ov0[k]));
    }
    int n = 0;
    for (int i = 2; i < numVerts; i++) {
        for (int k = 0; k < 4; k++)
            v2[k] = fmaf(od2[k], bary[i * 2 + 1], fmaf(od1[k], bary[i * 2 + 0], ov0[k]));
        snap(P, v0, v1, v2, &p0, &p1, &p2, rcpW, &lo, &hi);
        if (prepare(P, p0, p1, p2, lo, hi, &d1, &d2, &area) == 0) {
            setup_tri(P, &hdr[n], &data[n], vidx, v0, v1, v2,
                      &bary[0], &bary[i * 2 - 2], &bary[i * 2], p0, p1, p2, rcpW, d1, d2, area);
            n++;
        }
        memcpy(v1, v2, 16);
    }
    *cull = (n == 0) ? 4 : 5;
    return n;
}

/* ------------------------------------------------------------------ */
/* Coverage (R9/R10)                                                   */
/* ------------------------------------------------------------------ */

typedef struct {
    int64_t ax[3], ay[3];      /* absolute subpixel vertex positions */
    int64_t dx[3], dy[3];      /* edge k: vertex k -> vertex k+1 */
    int64_t b[3];              /* a_k x d_k, minus 1 for exclusive edges */
} edges_t;

static void make_edges(const cro_header* h, int32_t W, int32_t H, edges_t* e)
{
    int64_t ox = (int64_t)W * 8, oy = (int64_t)H * 8;
    e->ax[0] = h->v0x + ox; e->ay[0] = h->v0y + oy;
    e->ax[1] = h->v1x + ox; e->ay[1] = h->v1y + oy;
    e->ax[2] = h->v2x + ox; e->ay[2] = h->v2y + oy;
    for (int k = 0; k < 3; k++) {
        int j = (k + 1) % 3;
        e->dx[k] = e->ax[j] - e->ax[k];
        e->dy[k] = e->ay[j] - e->ay[k];
        e->b[k] = e->ax[k] * e->dy[k] - e->ay[k] * e->dx[k];
        /* Exclusive edge: dy > 0 || (dy == 0 && dx <= 0) (Util.inl:100). */
        if (e->dy[k] > 0 || (e->dy[k] == 0 && e->dx[k] <= 0)) e->b[k]--;
    }
}

static inline int covered(const edges_t* e, int64_t sx, int64_t sy)
{
    for (int k = 0; k < 3; k++)
        if (sx * e->dy[k] - sy * e->dx[k] > e->b[k]) return 0;
    return 1;
}

int32_t cro_cover_point(const cro_header* h, int32_t W, int32_t H, int64_t sx, int64_t sy)
{
    edges_t e;
    make_edges(h, W, H, &e);
    return covered(&e, sx, sy);
}

static inline uint32_t pixel_mask(const edges_t* e, int s, int32_t px, int32_t py)
{
    int n = 1 << s;
    uint32_t m = 0;
    for (int i = 0; i < n; i++) {
        int64_t sampleX = (int64_t)px * n + kPattern[s][i];
        int64_t sampleY = (int64_t)py * n + i;
        int64_t sx = (sampleX * 2 + 1) << (SUBPIXEL_LOG2 - s - 1);
        int64_t sy = (sampleY * 2 + 1) << (SUBPIXEL_LOG2 - s - 1);
        if (covered(e, sx, sy)) m |= 1u << i;
    }
    return m;
}

uint32_t cro_cover_pixel(const cro_header* h, int32_t W, int32_t H, int32_t s, int32_t px, int32_t py)
{
    edges_t e;
    make_edges(h, W, H, &e);
    return pixel_mask(&e, s, px, py);
}

/* ------------------------------------------------------------------ */
/* Whole-buffer setup                                                  */
/* ------------------------------------------------------------------ */

typedef struct {
    const cro_pipe* P;
    const uint8_t* verts;
    int64_t numVertices;
    const int32_t* idx;
    int64_t numTris;
    uint8_t* triSubtris;
    cro_header* hdr;
    cro_data* data;
    uint32_t* subtriTri;
    int64_t* clipBase;       /* per triangle: slot base for >= 2 subtris, else -1 */
    int phase;               /* 0: count, 1: write clipped */
    int64_t next;            /* work counter */
    pthread_mutex_t lock;
    int64_t cull[6];
} setup_job;

static int fetch_tri(const setup_job* J, int64_t t, float v[3][4], int32_t vidx[3])
{
    for (int k = 0; k < 3; k++) {
        vidx[k] = J->idx[t * 3 + k];
        if (vidx[k] < 0 || vidx[k] >= J->numVertices) return 0;
        memcpy(v[k], J->verts + (int64_t)vidx[k] * J->P->vertexStride, 16);
    }
    return 1;
}

#define SETUP_CHUNK 4096

static void* setup_worker(void* arg)
{
    setup_job* J = (setup_job*)arg;
    int64_t cull[6] = {0, 0, 0, 0, 0, 0};
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int64_t begin = J->next;
        J->next += SETUP_CHUNK;
        pthread_mutex_unlock(&J->lock);
        if (begin >= J->numTris) break;
        int64_t end = begin + SETUP_CHUNK < J->numTris ? begin + SETUP_CHUNK : J->numTris;
        for (int64_t t = begin; t < end; t++) {
            float v[3][4];
            int32_t vidx[3];
            cro_header h[7];
            cro_data d[7];
            int32_t c = 0, n = 0;
            if (J->phase == 1 && J->clipBase[t] < 0) continue;
            if (fetch_tri(J, t, v, vidx))
                n = cro_setup_triangle(J->P, v[0], v[1], v[2], vidx, h, d, &c);
            else
                c = 1;
            if (J->phase == 0) {
                cull[c]++;
                J->triSubtris[t] = (uint8_t)n;
                J->clipBase[t] = (n >= 2) ? 1 : -1;
                if (n == 1) { J->hdr[t] = h[0]; J->data[t] = d[0]; J->subtriTri[t] = (uint32_t)t; }
            } else {
                int64_t base = J->clipBase[t];
                J->hdr[t].misc = (uint32_t)base;
                for (int k = 0; k < n; k++) {
                    J->hdr[base + k] = h[k];
                    J->data[base + k] = d[k];
                    J->subtriTri[base + k] = (uint32_t)t;
                }
            }
        }
    }
    pthread_mutex_lock(&J->lock);
    for (int k = 0; k < 6; k++) J->cull[k] += cull[k];
    pthread_mutex_unlock(&J->lock);
    return NULL;
}

static void run_workers(void* (*fn)(void*), void* job, int nthreads)
{
    if (nthreads <= 1) { fn(job); return; }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, fn, job);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
}

static int64_t setup_all(const cro_pipe* P, const void* vertices, int64_t numVertices,
                         const int32_t* indices, int64_t numTris, uint8_t* triSubtris,
                         cro_header* hdr, cro_data* data, uint32_t* subtriTri,
                         int64_t maxSlots, int nthreads, int64_t cullOut[6])
{
    setup_job J;
    memset(&J, 0, sizeof(J));
    J.P = P; J.verts = (const uint8_t*)vertices; J.numVertices = numVertices;
    J.idx = indices; J.numTris = numTris; J.triSubtris = triSubtris;
    J.hdr = hdr; J.data = data; J.subtriTri = subtriTri;
    J.clipBase = (int64_t*)malloc(sizeof(int64_t) * (size_t)(numTris > 0 ? numTris : 1));
    pthread_mutex_init(&J.lock, NULL);
    if (numTris > maxSlots) { free(J.clipBase); return -1; }

    J.phase = 0; J.next = 0;
    run_workers(setup_worker, &J, nthreads);

    /* Deterministic allocation of clipped subtriangles in triangle order
     * (the device allocates them with an atomic counter; slot numbers are
     * not part of any output). */
    int64_t next = numTris;
    for (int64_t t = 0; t < numTris; t++) {
        if (J.clipBase[t] > 0) { J.clipBase[t] = next; next += triSubtris[t]; }
    }
    if (next > maxSlots) { free(J.clipBase); pthread_mutex_destroy(&J.lock); return -1; }
    J.phase = 1; J.next = 0;
    run_workers(setup_worker, &J, nthreads);
    if (cullOut) memcpy(cullOut, J.cull, sizeof(J.cull));
    free(J.clipBase);
    pthread_mutex_destroy(&J.lock);
    return next;
}

int64_t cro_setup(const cro_pipe* P, const void* vertices, int64_t numVertices,
                  const int32_t* indices, int64_t numTris, uint8_t* triSubtris,
                  cro_header* hdr, cro_data* data, uint32_t* subtriTri,
                  int64_t maxSlots, int32_t nthreads)
{
    return setup_all(P, vertices, numVertices, indices, numTris, triSubtris, hdr, data,
                     subtriTri, maxSlots, nthreads, NULL);
}

/* ------------------------------------------------------------------ */
/* texPhong shader (test/shader/Shaders.cu:37-48, 122-199)             */
/* ------------------------------------------------------------------ */

static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* log2 by the atanh series: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(t), t = (m-1)/(m+1), |t| <= 0.1716, atanh t = t (1 + t^2/3
 * + t^4/5 + t^6/7 + t^8/9) (the next term is below 2^-30 relative). */
float cro_log2f(float x)
{
    if (isnan(x) || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (isinf(x)) return INFINITY;
    int e = 0;
    uint32_t u = fbits(x);
    if ((u >> 23) == 0) {                 /* subnormal: scale by 2^23 (exact) */
        u = fbits(x * 8388608.0f);
        e = -23;
    }
    e += (int)(u >> 23) - 127;
    float m = bitsf((u & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421354f) { m *= 0.5f; e += 1; }
    float t = (m - 1.0f) / (m + 1.0f);
    float t2 = t * t;
    float p = fmaf(t2, 0.111111112f, 0.142857149f);
    p = fmaf(t2, p, 0.200000003f);
    p = fmaf(t2, p, 0.333333343f);
    p = fmaf(t2, p, 1.0f);
    float ln = (2.0f * t) * p;
    return fmaf(ln, 1.44269502f, (float)e);
}

/* exp2 by range reduction y = n + f, |f| <= 1/2, and the Taylor series of
 * e^z, z = f ln 2, to z^7/7! (next term below 2^-27), scaled by 2^n. */
float cro_exp2f(float y)
{
    if (isnan(y)) return y;
    if (y >= 128.0f) return INFINITY;
    if (y < -150.0f) return 0.0f;
    float n = rintf(y);
    float f = y - n;
    float z = f * 0.693147182f;
    float p = fmaf(z, 1.98412701e-04f, 1.38888892e-03f);
    p = fmaf(z, p, 8.33333377e-03f);
    p = fmaf(z, p, 4.16666679e-02f);
    p = fmaf(z, p, 1.66666672e-01f);
    p = fmaf(z, p, 0.5f);
    p = fmaf(z, p, 1.0f);
    p = fmaf(z, p, 1.0f);
    return ldexpf(p, (int)n);
}

float cro_powf(float x, float y)
{
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    return cro_exp2f(y * cro_log2f(x));
}

/* Integer texel index of floor value f, wrapped into [0, n); NaN and
 * out-of-range values give 0. */
static inline int32_t wrap_index(float f, int32_t n)
{
    if (!(fabsf(f) < 1073741824.0f)) return 0;
    int32_t k = (int32_t)f % n;
    return k < 0 ? k + n : k;
}

/* tex2D with normalised coordinates, wrap addressing and bilinear
 * filtering in FP32: texel space x = u W - 1/2, weights from the fraction,
 * x first, then y. */
void cro_tex_sample(const float* atlas, int32_t w, int32_t h, float u, float v, float out[4])
{
    float x = fmaf(u, (float)w, -0.5f), y = fmaf(v, (float)h, -0.5f);
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int32_t i0 = wrap_index(fx, w), j0 = wrap_index(fy, h);
    int32_t i1 = i0 + 1 == w ? 0 : i0 + 1, j1 = j0 + 1 == h ? 0 : j0 + 1;
    const float* t00 = atlas + 4 * ((int64_t)j0 * w + i0);
    const float* t10 = atlas + 4 * ((int64_t)j0 * w + i1);
    const float* t01 = atlas + 4 * ((int64_t)j1 * w + i0);
    const float* t11 = atlas + 4 * ((int64_t)j1 * w + i1);
    for (int c = 0; c < 4; c++) {
        float top = fmaf(t10[c] - t00[c], a, t00[c]);
        float bot = fmaf(t11[c] - t01[c], a, t01[c]);
        out[c] = fmaf(bot - top, b, top);
    }
}

/* texture2D (Shaders.cu:125-148): sqr(a) + sqr(b) fused as fma(a, a, b*b);
 * lerp(v0, v1, t) = v0 (1 - t) + v1 t (base/Math.hpp:114) fused on the left. */
void cro_texture2d(const cro_texture_spec* spec, const float* atlas, int32_t w, int32_t h,
                   const float tex[2], const float texDX[2], const float texDY[2], float out[4], float* lodOut)
{
    float ax = texDX[0] * spec->size[0], ay = texDX[1] * spec->size[1];
    float bx = texDY[0] * spec->size[0], by = texDY[1] * spec->size[1];
    float dxlen = fmaf(ax, ax, ay * ay);
    float dylen = fmaf(bx, bx, by * by);
    float lod = fminf(fmaxf(cro_log2f(fmaxf(dxlen, dylen)) * 0.5f, 0.0f), 14.0f);
    int li = (int)lod;
    const float* m0 = spec->mipLevels[li];
    const float* m1 = spec->mipLevels[li + 1];
    float tx = tex[0] - floorf(tex[0]);
    float ty = tex[1] - floorf(tex[1]);
    float v0[4], v1[4];
    cro_tex_sample(atlas, w, h, fmaf(tx, m0[0], m0[2]), fmaf(ty, m0[1], m0[3]), v0);
    cro_tex_sample(atlas, w, h, fmaf(tx, m1[0], m1[2]), fmaf(ty, m1[1], m1[3]), v1);
    float f = lod - (float)li;
    float g = 1.0f - f;
    for (int c = 0; c < 4; c++) out[c] = fmaf(v0[c], g, v1[c] * f);
    if (lodOut) *lodOut = lod;
}

/* normalize (base/Math.hpp:140-142): v * (1 * rcp(sqrt(lenSqr))), rcp(0) = 0,
 * lenSqr accumulated r += x*x from 0 (each step fused). */
static void normalize3(const float v[3], float out[3])
{
    float l = fmaf(v[2], v[2], fmaf(v[1], v[1], v[0] * v[0]));
    float r = sqrtf(l);
    float k = r != 0.0f ? 1.0f / r : 0.0f;
    for (int i = 0; i < 3; i++) out[i] = v[i] * k;
}

static inline float dot3(const float a[3], const float b[3])
{
    return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0]));
}

void cro_lighting(const float cameraPos[3], const float cameraNormal[3], const cro_material* m,
                  const float diffuse[3], float out[3])
{
    float I[3], N[3], R[3];
    normalize3(cameraPos, I);
    normalize3(cameraNormal, N);
    float dotIN = dot3(I, N);
    float s2 = dotIN * 2.0f;
    for (int i = 0; i < 3; i++) R[i] = fmaf(-N[i], s2, I[i]);
    float diffuseCoef = fmaf(fmaxf(-dotIN, 0.0f), 0.75f, 0.25f);
    float specularCoef = cro_powf(fmaxf(-dot3(I, R), 0.0f), m->glossiness);
    for (int i = 0; i < 3; i++) out[i] = fmaf(diffuseCoef, diffuse[i], specularCoef * m->specularColor[i]);
}

/* ------------------------------------------------------------------ */
/* Rasterization in submission order                                   */
/* ------------------------------------------------------------------ */

#define BAND_ROWS 16

typedef struct {
    const cro_pipe* P;
    const uint8_t* verts;
    const cro_header* hdr;
    const cro_data* data;
    const uint32_t* subtriTri;
    const int64_t* order;       /* ordered subtriangle slots */
    const int64_t* bandStart;   /* CSR over bands into bandList */
    const int64_t* bandList;    /* positions into order[] */
    int32_t numBands;
    int32_t wr, hr;             /* rounded surface size in pixels */
    uint32_t* color;
    uint32_t* depth;
    uint32_t* triid;
    int64_t nextBand;
    pthread_mutex_t lock;
    int64_t fragments;
    int64_t coveredSamples;
    int64_t zKills;
} raster_job;

/* computeBarys (FineRaster.inl:21-44): barycentrics at half-sample position
 * (sx, sy) of the W/U/V planes and their change per pixel step;
 * `(F32)p.x - u * (F32)w.x` is fused as fma(-u, w.x, p.x). */
static void compute_barys(const cro_data* td, int s, int32_t sx, int32_t sy, float b[3], float bdx[3],
                          float bdy[3])
{
    int32_t W = wadd(wadd(wmul(td->wx, sx), wmul(td->wy, sy)), td->wb);
    int32_t U = wadd(wadd(wmul(td->ux, sx), wmul(td->uy, sy)), td->ub);
    int32_t V = wadd(wadd(wmul(td->vx, sx), wmul(td->vy, sy)), td->vb);
    float w = 1.0f / (float)W;
    float u = w * (float)U;
    float v = w * (float)V;
    b[0] = 1.0f - u - v; b[1] = u; b[2] = v;
    float wd = w * (float)(1 << (s + 1));
    float udx = wd * fmaf(-u, (float)td->wx, (float)td->ux);
    float udy = wd * fmaf(-u, (float)td->wy, (float)td->uy);
    float vdx = wd * fmaf(-v, (float)td->wx, (float)td->vx);
    float vdy = wd * fmaf(-v, (float)td->wy, (float)td->vy);
    bdx[0] = -udx - vdx; bdx[1] = udx; bdx[2] = vdx;
    bdy[0] = -udy - vdy; bdy[1] = udy; bdy[2] = vdy;
}

/* interpolateVarying (PixelPipe.inl:52-59): v0*b.x + v1*b.y + v2*b.z,
 * fused as fma(v2, b.z, fma(v0, b.x, v1*b.y)). */
static void interp_varying(const raster_job* R, const cro_data* td, int k, const float b[3], float out[4])
{
    const int64_t stride = R->P->vertexStride;
    const float* c0 = (const float*)(R->verts + (int64_t)td->vi0 * stride + 16 + 16 * k);
    const float* c1 = (const float*)(R->verts + (int64_t)td->vi1 * stride + 16 + 16 * k);
    const float* c2 = (const float*)(R->verts + (int64_t)td->vi2 * stride + 16 + 16 * k);
    for (int c = 0; c < 4; c++) out[c] = fmaf(c2[c], b[2], fmaf(c0[c], b[0], c1[c] * b[1]));
}

/* Texture coordinate (varying 2) at the centre of pixel (px, py). */
static void center_tex(const raster_job* R, const cro_data* td, int64_t px, int64_t py, float out[2])
{
    int s = R->P->samplesLog2, n = 1 << s;
    float b[3], bdx[3], bdy[3], t[4];
    compute_barys(td, s, (int32_t)(px * n * 2 + n), (int32_t)(py * n * 2 + n), b, bdx, bdy);
    interp_varying(R, td, 2, b, t);
    out[0] = t[0]; out[1] = t[1];
}

/* FragmentShader_texPhong::run (Shaders.cu:152-199).  Returns 1 when the
 * fragment is discarded (alpha test, or a material index outside the
 * material table). */
static int shade_tex_phong(const raster_job* R, const cro_data* td, uint32_t tri, int64_t px, int64_t py,
                           int ctr, uint32_t* color)
{
    const cro_pipe* P = R->P;
    int s = P->samplesLog2, n = 1 << s;
    int lerp = (P->renderModeFlags & CRO_FLAG_LERP) != 0;
    float cb[3] = {0.0f, 0.0f, 1.0f}, cdx[3] = {0, 0, 0}, cdy[3] = {0, 0, 0};   /* centroid */
    if (lerp) {
        int64_t sx2 = px * n * 2 + ((ctr == -1) ? n : kPattern[s][ctr] * 2 + 1);
        int64_t sy2 = py * n * 2 + ((ctr == -1) ? n : ctr * 2 + 1);
        compute_barys(td, s, (int32_t)sx2, (int32_t)sy2, cb, cdx, cdy);
    }
    float cp[4], cn[4], t[4];
    interp_varying(R, td, 0, cb, cp);
    interp_varying(R, td, 1, cb, cn);
    float tex[2], texDX[2], texDY[2];
    if (!(P->renderModeFlags & CRO_FLAG_QUADS)) {
        /* Centroid sample, analytic derivatives. */
        interp_varying(R, td, 2, cb, t);
        tex[0] = t[0]; tex[1] = t[1];
        interp_varying(R, td, 2, cdx, t);
        texDX[0] = t[0]; texDX[1] = t[1];
        interp_varying(R, td, 2, cdy, t);
        texDY[0] = t[0]; texDY[1] = t[1];
    } else if (!lerp) {
        /* Varyings of vertex 2 everywhere: no change across the quad. */
        float z[3] = {0.0f, 0.0f, 1.0f};
        interp_varying(R, td, 2, z, t);
        tex[0] = t[0]; tex[1] = t[1];
        texDX[0] = texDX[1] = texDY[0] = texDY[1] = 0.0f;
    } else {
        /* Pixel centre, numerical derivatives over the 2x2 quad: dFdx is the
         * odd column's value minus the even column's, dFdy likewise for rows
         * (PixelPipe.hpp:59-60); the quad partners run the same shader at
         * their own pixel centres. */
        float a[2], b[2];
        center_tex(R, td, px, py, tex);
        center_tex(R, td, px & ~(int64_t)1, py, a);
        center_tex(R, td, px | 1, py, b);
        texDX[0] = b[0] - a[0]; texDX[1] = b[1] - a[1];
        center_tex(R, td, px, py & ~(int64_t)1, a);
        center_tex(R, td, px, py | 1, b);
        texDY[0] = b[0] - a[0]; texDY[1] = b[1] - a[1];
    }
    int32_t mi = P->triangleMaterialIdx ? P->triangleMaterialIdx[tri] : -1;
    if (mi < 0 || mi >= P->numMaterials) return 1;
    const cro_material* M = &P->materials[mi];
    float dc[4] = {M->diffuseColor[0], M->diffuseColor[1], M->diffuseColor[2], M->diffuseColor[3]};
    if (M->diffuseTexture.size[0] != 0.0f) {
        cro_texture2d(&M->diffuseTexture, P->atlas, P->atlasWidth, P->atlasHeight, tex, texDX, texDY, t, NULL);
        dc[0] = t[0]; dc[1] = t[1]; dc[2] = t[2];
    }
    if (M->alphaTexture.size[0] != 0.0f) {
        cro_texture2d(&M->alphaTexture, P->atlas, P->atlasWidth, P->atlasHeight, tex, texDX, texDY, t, NULL);
        dc[3] = t[1];
    }
    if (dc[3] < 0.5f) return 1;
    float lit[4];
    cro_lighting(cp, cn, M, dc, lit);
    lit[3] = dc[3];
    *color = cro_to_abgr(lit);
    return 0;
}

/* Fragment shader at pixel (px, py) with coverage centroid ctr (-1: pixel
 * centre).  Returns 1 when the fragment is discarded. */
static int shade_fragment(const raster_job* R, const cro_data* td, uint32_t tri, int64_t px, int64_t py, int ctr,
                          uint32_t* color)
{
    const cro_pipe* P = R->P;
    if (P->shader == CRO_SHADER_CONSTANT) {
        *color = P->constColor;
        return 0;
    }
    if (P->shader == CRO_SHADER_TEXPHONG) return shade_tex_phong(R, td, tri, px, py, ctr, color);
    /* GouraudShader (PixelPipe.inl:65-69) at the centroid. */
    int s = P->samplesLog2, n = 1 << s;
    float b[3] = {0.0f, 0.0f, 1.0f}, bdx[3], bdy[3];   /* no lerp: vertex 2 (FineRaster.inl:72-81) */
    if (P->renderModeFlags & CRO_FLAG_LERP) {
        int64_t sx2 = px * n * 2 + ((ctr == -1) ? n : kPattern[s][ctr] * 2 + 1);
        int64_t sy2 = py * n * 2 + ((ctr == -1) ? n : ctr * 2 + 1);
        compute_barys(td, s, (int32_t)sx2, (int32_t)sy2, b, bdx, bdy);
    }
    float c[4];
    interp_varying(R, td, 0, b, c);
    *color = cro_to_abgr(c);
    return 0;
}

/* frags[0]: fragments, frags[1]: covered samples, frags[2]: fragments whose
 * covered samples all failed the depth test. */
static void raster_subtri(raster_job* R, int64_t slot, int32_t y0, int32_t y1, int64_t frags[3])
{
    const cro_pipe* P = R->P;
    const cro_header* h = &R->hdr[slot];
    const cro_data* td = &R->data[slot];
    int s = P->samplesLog2, n = 1 << s;
    edges_t e;
    make_edges(h, P->width, P->height, &e);
    int64_t minx = e.ax[0], maxx = e.ax[0], miny = e.ay[0], maxy = e.ay[0];
    for (int k = 1; k < 3; k++) {
        if (e.ax[k] < minx) minx = e.ax[k];
        if (e.ax[k] > maxx) maxx = e.ax[k];
        if (e.ay[k] < miny) miny = e.ay[k];
        if (e.ay[k] > maxy) maxy = e.ay[k];
    }
    int64_t pxlo = minx >> 4, pxhi = maxx >> 4, pylo = miny >> 4, pyhi = maxy >> 4;
    if (pxlo < 0) pxlo = 0;
    if (pxhi > R->wr - 1) pxhi = R->wr - 1;
    if (pylo < y0) pylo = y0;
    if (pyhi > y1 - 1) pyhi = y1 - 1;
    uint32_t tri = R->subtriTri[slot];
    int64_t pitch = (int64_t)R->wr * n;
    int depthOn = (P->renderModeFlags & CRO_FLAG_DEPTH) != 0;

    for (int64_t py = pylo; py <= pyhi; py++) {
        for (int64_t px = pxlo; px <= pxhi; px++) {
            uint32_t cover = pixel_mask(&e, s, (int32_t)px, (int32_t)py);
            if (!cover) continue;
            frags[0]++;
            frags[1] += __builtin_popcount(cover);
            /* texel of sample i: row py, column tileX*8N + i*8 + (px&7) */
            int64_t base = py * pitch + (px >> 3) * 8 * n + (px & 7);
            uint32_t write = 0, dnew[8];
            for (int i = 0; i < n; i++) {
                if (!(cover & (1u << i))) continue;
                if (depthOn) {
                    uint32_t sampleX = (uint32_t)(px * n + kPattern[s][i]);
                    uint32_t sampleY = (uint32_t)(py * n + i);
                    uint32_t d = td->zx * sampleX + td->zy * sampleY + td->zb;
                    if (d >= R->depth[base + i * 8]) continue;   /* LESS (FineRaster.inl:697) */
                    dnew[i] = d;
                }
                write |= 1u << i;
            }
            if (!write) {
                frags[2]++;
                continue;
            }
            /* Shade, then write: a discarded fragment writes neither depth
             * nor colour. */
            uint32_t src;
            if (shade_fragment(R, td, tri, px, py, cro_select_centroid(s, cover), &src)) continue;
            for (int i = 0; i < n; i++) {
                if (!(write & (1u << i))) continue;
                if (depthOn) R->depth[base + i * 8] = dnew[i];
                uint32_t* cp = &R->color[base + i * 8];
                if (P->blend != CRO_BLEND_DEPTHONLY) *cp = cro_blend(P->blend, src, *cp);
                if (R->triid) R->triid[base + i * 8] = tri;
            }
        }
    }
}

static void* raster_worker(void* arg)
{
    raster_job* R = (raster_job*)arg;
    int64_t frags[3] = {0, 0, 0};
    for (;;) {
        pthread_mutex_lock(&R->lock);
        int64_t band = R->nextBand++;
        pthread_mutex_unlock(&R->lock);
        if (band >= R->numBands) break;
        int32_t y0 = (int32_t)band * BAND_ROWS;
        int32_t y1 = y0 + BAND_ROWS < R->hr ? y0 + BAND_ROWS : R->hr;
        for (int64_t k = R->bandStart[band]; k < R->bandStart[band + 1]; k++)
            raster_subtri(R, R->order[R->bandList[k]], y0, y1, frags);
    }
    pthread_mutex_lock(&R->lock);
    R->fragments += frags[0];
    R->coveredSamples += frags[1];
    R->zKills += frags[2];
    pthread_mutex_unlock(&R->lock);
    return NULL;
}

static void subtri_rows(const cro_pipe* P, const cro_header* h, int32_t hr, int32_t* r0, int32_t* r1)
{
    int64_t oy = (int64_t)P->height * 8;
    int64_t y[3] = {h->v0y + oy, h->v1y + oy, h->v2y + oy};
    int64_t lo = y[0], hi = y[0];
    for (int k = 1; k < 3; k++) { if (y[k] < lo) lo = y[k]; if (y[k] > hi) hi = y[k]; }
    lo >>= 4; hi >>= 4;
    if (lo < 0) lo = 0;
    if (hi > hr - 1) hi = hr - 1;
    *r0 = (int32_t)lo; *r1 = (int32_t)hi;
}

int32_t cro_render(const cro_pipe* P, const void* vertices, int64_t numVertices,
                   const int32_t* indices, int64_t numTris,
                   uint32_t* color, uint32_t* depth, uint32_t* triid,
                   int32_t nthreads, cro_stats* stats)
{
    if (!P || P->width < 1 || P->height < 1 || P->width > MAXVIEWPORT || P->height > MAXVIEWPORT ||
        P->samplesLog2 < 0 || P->samplesLog2 > 3 || P->vertexStride < 16 || (P->vertexStride & 15) ||
        !color || !depth || numTris < 0)
        return -1;
    if (P->shader == CRO_SHADER_GOURAUD && P->vertexStride < 32) return -1;
    if (P->shader == CRO_SHADER_TEXPHONG &&
        (P->vertexStride < 64 || (P->numMaterials > 0 && (!P->materials || !P->triangleMaterialIdx)) ||
         (P->atlas && (P->atlasWidth < 1 || P->atlasHeight < 1))))
        return -1;
    int n = 1 << P->samplesLog2;
    int32_t wr = (P->width + 7) & ~7, hr = (P->height + 7) & ~7;
    int64_t texels = (int64_t)wr * n * hr;

    if (P->deferredClear) {
        for (int64_t i = 0; i < texels; i++) { color[i] = P->clearColor; depth[i] = P->clearDepth; }
        if (triid) for (int64_t i = 0; i < texels; i++) triid[i] = 0xFFFFFFFFu;
    }

    /* Setup.  Worst case 7 subtriangles per triangle. */
    int64_t maxSlots = numTris * 7 + 1;
    uint8_t* triSubtris = (uint8_t*)malloc((size_t)(numTris > 0 ? numTris : 1));
    cro_header* hdr = (cro_header*)malloc(sizeof(cro_header) * (size_t)maxSlots);
    cro_data* data = (cro_data*)malloc(sizeof(cro_data) * (size_t)maxSlots);
    uint32_t* subtriTri = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)maxSlots);
    int64_t cull[6] = {0, 0, 0, 0, 0, 0};
    int64_t slots = setup_all(P, vertices, numVertices, indices, numTris, triSubtris, hdr, data,
                              subtriTri, maxSlots, nthreads, cull);

    /* Submission order: triangle t, then its subtriangles in fan order. */
    int64_t numSub = 0, trisVisible = 0;
    for (int64_t t = 0; t < numTris; t++) {
        numSub += triSubtris[t];
        trisVisible += triSubtris[t] != 0;
    }
    int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)(numSub + 1));
    int64_t k = 0;
    for (int64_t t = 0; t < numTris; t++) {
        int c = triSubtris[t];
        if (c == 1) order[k++] = t;
        else for (int j = 0; j < c; j++) order[k++] = (int64_t)hdr[t].misc + j;
    }
    (void)slots;

    /* Bands of BAND_ROWS pixel rows, each with its ordered subtriangle list. */
    int32_t numBands = (hr + BAND_ROWS - 1) / BAND_ROWS;
    int64_t* bandStart = (int64_t*)calloc((size_t)numBands + 1, sizeof(int64_t));
    for (int64_t j = 0; j < numSub; j++) {
        int32_t r0, r1;
        subtri_rows(P, &hdr[order[j]], hr, &r0, &r1);
        for (int32_t b = r0 / BAND_ROWS; b <= r1 / BAND_ROWS && r0 <= r1; b++) bandStart[b + 1]++;
    }
    for (int32_t b = 0; b < numBands; b++) bandStart[b + 1] += bandStart[b];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)numBands);
    memcpy(fill, bandStart, sizeof(int64_t) * (size_t)numBands);
    int64_t* bandList = (int64_t*)malloc(sizeof(int64_t) * (size_t)(bandStart[numBands] + 1));
    for (int64_t j = 0; j < numSub; j++) {
        int32_t r0, r1;
        subtri_rows(P, &hdr[order[j]], hr, &r0, &r1);
        for (int32_t b = r0 / BAND_ROWS; b <= r1 / BAND_ROWS && r0 <= r1; b++) bandList[fill[b]++] = j;
    }

    raster_job R;
    memset(&R, 0, sizeof(R));
    R.P = P; R.verts = (const uint8_t*)vertices; R.hdr = hdr; R.data = data; R.subtriTri = subtriTri;
    R.order = order; R.bandStart = bandStart; R.bandList = bandList; R.numBands = numBands;
    R.wr = wr; R.hr = hr; R.color = color; R.depth = depth; R.triid = triid;
    pthread_mutex_init(&R.lock, NULL);
    run_workers(raster_worker, &R, nthreads);
    pthread_mutex_destroy(&R.lock);

    if (stats) {
        stats->numTris = numTris;
        stats->numSubtris = numSub;
        stats->culledFrustum = cull[1];
        stats->culledBackface = cull[2];
        stats->culledBetween = cull[3];
        stats->clipped = cull[4] + cull[5];
        stats->fragments = R.fragments;
        stats->coveredSamples = R.coveredSamples;
        stats->zKills = R.zKills;
        stats->trisVisible = trisVisible;
    }
    free(triSubtris); free(hdr); free(data); free(subtriTri); free(order);
    free(bandStart); free(fill); free(bandList);
    return 0;
}

/* ---- passthrough vertex shader (test/shader/PassThrough.cu:16-35) -------- */

void cro_vertex_passthrough(const float* modelPos, int64_t numVertices, const float posToClip[16],
                            float* out, int64_t outStrideFloats) {
    for (int64_t v = 0; v < numVertices; v++) {
        const float in[4] = {modelPos[3 * v + 0], modelPos[3 * v + 1], modelPos[3 * v + 2], 1.0f};
        for (int i = 0; i < 4; i++) {
            /* MatrixBase::operator*(Vector) (base/Math.hpp:990-1001): rr = 0;
             * rr += get(i, j) * v[j], contracted to fmaf; get(i, j) = m[i + 4j]. */
            float rr = 0.0f;
            for (int j = 0; j < 4; j++) rr = fmaf(posToClip[i + 4 * j], in[j], rr);
            out[outStrideFloats * v + i] = rr;
        }
    }
}
