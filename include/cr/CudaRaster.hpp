/*
 * CudaRaster.hpp -- header-only C++ facade over the C ABI (cr.h) with the
 * reference's class and method names, so that a CudaRaster-shaped client
 * (test/SceneCR.cpp:149-195, 267-292 in the reference) compiles against it
 * with few changes.
 *
 *   FW::CudaRaster  (src/cudaraster/CudaRaster.hpp:42-190)  -> cr::CudaRaster
 *   FW::CudaSurface (src/cudaraster/CudaSurface.hpp:36-110) -> cr::CudaSurface
 *
 * Where the reference calls fail() and exits (base/Defs.hpp:107-117), the
 * facade throws cr::Error carrying the library's message.  Buffers are
 * device pointers (the reference's Buffer objects migrate data lazily; here
 * the caller owns device memory).
 */
#ifndef CR_CUDARASTER_HPP
#define CR_CUDARASTER_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "cr.h"

namespace cr {

class Error : public std::runtime_error {
 public:
  Error(cr_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
  cr_status status;
};

inline void check(cr_status s, const cr_context* ctx) {
  if (s != CR_OK) throw Error(s, cr_last_error(ctx));
}

class CudaSurface {
 public:
  enum Format { FormatRGBA8 = CR_FORMAT_RGBA8, FormatDepth32 = CR_FORMAT_DEPTH32,
                FormatTriangleId = CR_FORMAT_TRIANGLE_ID };

  // CudaSurface(const Vec2i& size, Format format, int numSamples = 1)
  CudaSurface(int width, int height, Format format, int numSamples = 1) {
    check(cr_surface_create(width, height, format, numSamples, &s_), nullptr);
  }
  ~CudaSurface() { if (s_) cr_surface_destroy(s_); }
  CudaSurface(const CudaSurface&) = delete;
  CudaSurface& operator=(const CudaSurface&) = delete;

  int getWidth() const { int v[7]; info(v); return v[0]; }
  int getHeight() const { int v[7]; info(v); return v[1]; }
  int getRoundedWidth() const { int v[7]; info(v); return v[2]; }
  int getRoundedHeight() const { int v[7]; info(v); return v[3]; }
  int getTextureWidth() const { int v[7]; info(v); return v[4]; }
  int getNumSamples() const { int v[7]; info(v); return v[5]; }
  Format getFormat() const { int v[7]; info(v); return (Format)v[6]; }

  // Replaces getCudaArray(): device address of the texels.
  void* getDevicePtr(size_t* bytes = nullptr) const {
    void* p = nullptr;
    size_t b = 0;
    check(cr_surface_get_device_ptr(s_, &p, &b), nullptr);
    if (bytes) *bytes = b;
    return p;
  }
  // Whole texture, (texture width x rounded height) u32 texels.
  std::vector<uint32_t> download() const {
    size_t b = 0;
    getDevicePtr(&b);
    std::vector<uint32_t> out(b / 4);
    check(cr_surface_download(s_, out.data(), b), nullptr);
    return out;
  }
  void upload(const std::vector<uint32_t>& texels) {
    check(cr_surface_upload(s_, texels.data(), texels.size() * 4), nullptr);
  }
  // Box MSAA resolve into width x height RGBA8 words at the device address
  // devOut (the reference's commented-out resolve, CudaSurface.hpp:73).
  void resolveToDevice(void* devOut, void* hipStream = nullptr) const {
    check(cr_surface_resolve(s_, devOut, hipStream), nullptr);
  }
  cr_surface* handle() const { return s_; }

 private:
  void info(int v[7]) const {
    int32_t w, h, rw, rh, tw, n, f;
    check(cr_surface_get_info(s_, &w, &h, &rw, &rh, &tw, &n, &f), nullptr);
    v[0] = w; v[1] = h; v[2] = rw; v[3] = rh; v[4] = tw; v[5] = n; v[6] = f;
  }
  cr_surface* s_ = nullptr;
};

class CudaRaster {
 public:
  struct Stats { float setupTime, binTime, coarseTime, fineTime; };   // CudaRaster.hpp:45-51
  // DebugParams (CudaRaster.hpp:53-67) with this library's knobs; profilingMode
  // takes the reference pipe's ProfilingMode role (CR_PROFILING_*).
  struct DebugParams {
    bool disableHiZ = false;
    bool countFragments = false;
    int coarseThreads = 0;
    int profilingMode = CR_PROFILING_DEFAULT;
  };

  CudaRaster() { check(cr_context_create(&ctx_), nullptr); }
  ~CudaRaster() { if (ctx_) cr_context_destroy(ctx_); }
  CudaRaster(const CudaRaster&) = delete;             // non-copyable, as the reference (:187-188)
  CudaRaster& operator=(const CudaRaster&) = delete;

  void init() {}                                      // the context is ready after construction

  void setSurfaces(CudaSurface* color, CudaSurface* depth) {
    check(cr_set_surfaces(ctx_, color ? color->handle() : nullptr, depth ? depth->handle() : nullptr), ctx_);
  }
  void setTriangleIdSurface(CudaSurface* ids) {
    check(cr_set_triangle_id_surface(ctx_, ids ? ids->handle() : nullptr), ctx_);
  }
  void deferredClear(const float rgba[4], float depth = 1.0f) { check(cr_deferred_clear(ctx_, rgba, depth), ctx_); }
  void deferredClear(float depth = 1.0f) {
    const float zero[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    deferredClear(zero, depth);
  }

  // setPixelPipe(CudaModule*, const String& name): the stock pipes by name
  // with the demo's compile-time choices, or a descriptor.
  void setPixelPipe(const std::string& name, int blendShader = CR_BLEND_REPLACE, int samplesLog2 = 0,
                    unsigned renderModeFlags = CR_RENDER_DEPTH | CR_RENDER_LERP) {
    cr_pipe_desc d;
    check(cr_pipe_lookup(name.c_str(), blendShader, samplesLog2, renderModeFlags, &d), ctx_);
    setPixelPipe(d);
  }
  void setPixelPipe(const cr_pipe_desc& d) { check(cr_set_pixel_pipe(ctx_, &d), ctx_); }

  void setVertexBuffer(const void* devPtr, int64_t ofs, int64_t numVertices) {
    check(cr_set_vertex_buffer(ctx_, devPtr, ofs, numVertices), ctx_);
  }
  void setIndexBuffer(const void* devPtr, int64_t ofs, int numTris) {
    check(cr_set_index_buffer(ctx_, devPtr, ofs, numTris), ctx_);
  }

  void drawTriangles(void* hipStream = nullptr) { check(cr_draw_triangles(ctx_, hipStream), ctx_); }
  // drawTriangles without device allocation (cr_draw_triangles_noalloc).
  void drawTrianglesNoAlloc(void* hipStream = nullptr) { check(cr_draw_triangles_noalloc(ctx_, hipStream), ctx_); }

  // The demo's passthrough vertex shader (test/shader/PassThrough.cu:16-35):
  // clip positions posToClip * (modelPos, 1) into the first 16 bytes of each
  // outStride-byte record of the device vertex buffer devOut.  posToClip is
  // a Mat4f's getPtr() (column-major, host memory).
  void vertexShaderPassthrough(const void* devModelPos, int64_t numVertices, const float* posToClip, void* devOut,
                               int outStride, void* hipStream = nullptr) {
    check(cr_vertex_shader_passthrough(ctx_, static_cast<const float*>(devModelPos), numVertices, posToClip, devOut,
                                       outStride, hipStream),
          ctx_);
  }

  Stats getStats() {
    cr_stats s;
    check(cr_get_stats(ctx_, &s), ctx_);
    return Stats{s.setup_time, s.bin_time, s.coarse_time, s.fine_time};
  }
  cr_counters getCounters() {
    cr_counters c;
    check(cr_get_counters(ctx_, &c), ctx_);
    return c;
  }
  std::string getProfilingInfo() {
    std::string buf(8192, '\0');
    check(cr_get_profiling_info(ctx_, &buf[0], buf.size()), ctx_);
    buf.resize(buf.find('\0'));
    return buf;
  }
  void setDebugParams(const DebugParams& p) {
    cr_debug_params d;
    d.disable_hiz = p.disableHiZ ? 1 : 0;
    d.count_fragments = p.countFragments ? 1 : 0;
    d.coarse_threads = p.coarseThreads;
    d.profiling_mode = p.profilingMode;
    check(cr_set_debug_params(ctx_, &d), ctx_);
  }
  // The texPhong shader's materials, per-triangle material indices and
  // texture atlas (what the demo puts in c_constants / t_textureAtlas).
  void setShaderConstants(const cr_shader_constants& c) { check(cr_set_shader_constants(ctx_, &c), ctx_); }
  cr_profile_counters getProfileCounters() {
    cr_profile_counters c;
    check(cr_get_profile_counters(ctx_, &c), ctx_);
    return c;
  }

  cr_context* handle() const { return ctx_; }

 private:
  cr_context* ctx_ = nullptr;
};

}  // namespace cr

#endif
