// tvk_abi.cpp -- extern "C" entry points of libtvkhip.so (include/tvk_hip.h).
// Each function forwards to the AbstrRenderer-shaped HipRenderer; C++
// exceptions never cross the boundary.
#include <algorithm>
#include <fcntl.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/tvk_hip.h"
#include "errors.hpp"
#include "kernels.hpp"
#include "memman.hpp"
#include "renderer.hpp"
#include "server.hpp"
#include "toc.hpp"

namespace tvk {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }
}  // namespace tvk

using tvk::HipRenderer;

struct tvk_renderer {
  HipRenderer impl;
  tvk_renderer(int t, bool a, bool b, bool c, bool d) : impl(t, a, b, c, d) {}
};

struct tvk_server {
  tvk::RenderServer impl;
};

struct tvk_toc {
  tvk::TocVolume impl;
};

struct tvk_lod_mlp {
  int device = -1;
  tvk::DevBuf w, b, x, y;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
};

#define TVK_GUARD(r)                                  \
  do {                                                \
    if (!(r)) {                                       \
      tvk::set_last_error("null renderer handle");    \
      return TVK_ERR_ARG;                             \
    }                                                 \
  } while (0)

#define TVK_TRY(expr)                                 \
  try {                                               \
    return (expr);                                    \
  } catch (const std::bad_alloc&) {                   \
    tvk::set_last_error("host allocation failed");    \
    return TVK_ERR_OOM;                               \
  } catch (const std::exception& e) {                 \
    tvk::set_last_error(e.what());                    \
    return TVK_ERR_ARG;                               \
  }

extern "C" {

const char* tvk_last_error(void) { return tvk::g_last_error.c_str(); }
const char* tvk_version(void) { return "tvk-hip 0.1 (gfx950)"; }

int tvk_create(int type, int pow2, int to8, int no_border, int bias_scale, tvk_renderer** out) {
  if (!out) {
    tvk::set_last_error("null output handle");
    return TVK_ERR_ARG;
  }
  *out = nullptr;
  if (type != TVK_OPENGL_RAYCASTER && type != TVK_OPENGL_GRIDLEAPER) {
    tvk::set_last_error("renderer type not provided by the HIP backend (raycaster or grid leaper only)");
    return TVK_ERR_ARG;
  }
  try {
    *out = new tvk_renderer(type, pow2 != 0, to8 != 0, no_border != 0, bias_scale != 0);
  } catch (...) {
    tvk::set_last_error("host allocation failed");
    return TVK_ERR_OOM;
  }
  return TVK_OK;
}

int tvk_destroy(tvk_renderer* r) {
  delete r;
  return TVK_OK;
}

int tvk_initialize(tvk_renderer* r, int device) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Initialize(device));
}

int tvk_resize(tvk_renderer* r, uint32_t w, uint32_t h) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Resize(w, h));
}

int tvk_set_renderer_target(tvk_renderer* r, int t) {
  TVK_GUARD(r);
  return r->impl.SetRendererTarget(t);
}

int tvk_get_renderer_target(tvk_renderer* r, int* t) {
  TVK_GUARD(r);
  if (!t) return TVK_ERR_ARG;
  *t = r->impl.RendererTarget();
  return TVK_OK;
}

int tvk_load_dataset(tvk_renderer* r, const char* path, int verify) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.LoadDataset(path, verify != 0));
}

int tvk_set_generated_volume(tvk_renderer* r, int kind, uint64_t seed, uint64_t nx, uint64_t ny,
                             uint64_t nz, uint32_t bits, uint32_t brick, uint32_t apron,
                             uint32_t n_lod) {
  TVK_GUARD(r);
  const uint64_t d[3] = {nx, ny, nz};
  TVK_TRY(r->impl.SetGeneratedVolume(kind, seed, d, bits, brick, apron, n_lod));
}

int tvk_generate_volume_gpu(int device, int kind, uint64_t seed, uint64_t nx, uint64_t ny,
                            uint64_t nz, uint32_t bits, void* out, uint64_t out_bytes) {
  if (!out || (bits != 8 && bits != 16) || nx == 0 || ny == 0 || nz == 0 ||
      !tvk::is_gpu_volume_kind(kind)) {
    tvk::set_last_error("bad device volume generator arguments");
    return TVK_ERR_ARG;
  }
  const uint64_t bpv = bits / 8, plane = nx * ny * bpv;
  if (out_bytes < plane * nz) {
    tvk::set_last_error("output buffer too small");
    return TVK_ERR_ARG;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    tvk::set_last_error("device ordinal out of range");
    return TVK_ERR_ARG;
  }
  if (hipSetDevice(device) != hipSuccess) {
    tvk::set_last_error("hipSetDevice failed");
    return TVK_ERR_HIP;
  }
  // z slabs of at most 1 GiB, generated on the device and copied back
  const uint64_t slab = std::max<uint64_t>(1, (1ull << 30) / plane);
  void* buf = nullptr;
  if (hipMalloc(&buf, std::min(slab, nz) * plane) != hipSuccess) {
    tvk::set_last_error("device allocation for the generator failed");
    return TVK_ERR_OOM;
  }
  const int64_t d[3] = {(int64_t)nx, (int64_t)ny, (int64_t)nz};
  int rc = TVK_OK;
  for (uint64_t z0 = 0; z0 < nz && rc == TVK_OK; z0 += slab) {
    const uint64_t cnt = std::min(slab, nz - z0);
    hipError_t e = tvk::launch_generate_volume(kind, seed, d, (int)bits, (int64_t)z0, (int64_t)cnt,
                                               buf, nullptr);
    if (e == hipSuccess)
      e = hipMemcpy((uint8_t*)out + z0 * plane, buf, cnt * plane, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      tvk::set_last_error(std::string("device volume generator: ") + hipGetErrorString(e));
      rc = TVK_ERR_HIP;
    }
  }
  (void)hipFree(buf);
  return rc;
}

int tvk_set_volume(tvk_renderer* r, const void* v, uint64_t nx, uint64_t ny, uint64_t nz,
                   uint32_t bits, uint32_t brick, uint32_t apron, uint32_t n_lod) {
  TVK_GUARD(r);
  const uint64_t d[3] = {nx, ny, nz};
  TVK_TRY(r->impl.SetVolume(v, d, bits, brick, apron, n_lod));
}

int tvk_set_paged_volume(tvk_renderer* r, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t bits,
                         uint32_t brick, uint32_t apron, uint32_t n_lod, tvk_brick_source_fn src,
                         void* user, const float* ranges, uint64_t pool_bytes) {
  TVK_GUARD(r);
  const uint64_t d[3] = {nx, ny, nz};
  TVK_TRY(r->impl.SetPagedVolume(d, bits, brick, apron, n_lod, src, user, ranges, pool_bytes));
}

int tvk_set_paged_flat_volume(tvk_renderer* r, const void* v, uint64_t nx, uint64_t ny, uint64_t nz,
                              uint32_t bits, uint32_t brick, uint32_t apron, uint32_t n_lod,
                              uint64_t pool_bytes) {
  TVK_GUARD(r);
  const uint64_t d[3] = {nx, ny, nz};
  TVK_TRY(r->impl.SetPagedFlatVolume(v, d, bits, brick, apron, n_lod, pool_bytes));
}

int tvk_get_lod_level_count(tvk_renderer* r, uint32_t* n) {
  TVK_GUARD(r);
  if (!n) return TVK_ERR_ARG;
  *n = (uint32_t)r->impl.LodLevelCount();
  return TVK_OK;
}

int tvk_get_bit_width(tvk_renderer* r, uint32_t* bits) {
  TVK_GUARD(r);
  if (!bits) return TVK_ERR_ARG;
  return r->impl.BitWidth(bits);
}

int tvk_get_domain_size(tvk_renderer* r, uint32_t lod, uint64_t dims[3]) {
  TVK_GUARD(r);
  if (!dims) return TVK_ERR_ARG;
  return r->impl.DomainSize(lod, dims);
}

int tvk_get_brick_layout(tvk_renderer* r, uint32_t lod, uint32_t grid[3], uint32_t* brick,
                         uint32_t* apron) {
  TVK_GUARD(r);
  if (!grid) return TVK_ERR_ARG;
  return r->impl.BrickLayout(lod, grid, brick, apron);
}

int tvk_read_lod(tvk_renderer* r, uint32_t lod, void* out, uint64_t bytes) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.ReadLod(lod, out, bytes));
}

int tvk_read_brick(tvk_renderer* r, uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz, void* out,
                   uint64_t bytes) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.ReadBrick(lod, bx, by, bz, out, bytes));
}

int tvk_toc_build(int device, const void* voxels, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t bits,
                  uint32_t max_brick, uint32_t overlap, int use_median, int clamp_to_edge,
                  uint32_t compression, uint32_t level, uint32_t layout, const char* path,
                  tvk_toc_info* info) {
  if (!path) {
    tvk::set_last_error("null path");
    return TVK_ERR_ARG;
  }
  const uint64_t d[3] = {nx, ny, nz};
  TVK_TRY(tvk::toc_build(device, voxels, d, bits, max_brick, overlap, use_median != 0, clamp_to_edge != 0,
                         compression, level, layout, path, info));
}

int tvk_toc_open(const char* path, tvk_toc** out) {
  if (!path || !out) {
    tvk::set_last_error("null argument");
    return TVK_ERR_ARG;
  }
  *out = nullptr;
  try {
    std::unique_ptr<tvk_toc> t(new tvk_toc);
    if (int e = t->impl.Open(path)) return e;
    *out = t.release();
    return TVK_OK;
  } catch (const std::bad_alloc&) {
    tvk::set_last_error("host allocation failed");
    return TVK_ERR_OOM;
  }
}

int tvk_toc_close(tvk_toc* t) {
  delete t;
  return TVK_OK;
}

int tvk_toc_get_info(tvk_toc* t, tvk_toc_info* info) {
  if (!t || !info) {
    tvk::set_last_error("null argument");
    return TVK_ERR_ARG;
  }
  t->impl.Info(info);
  return TVK_OK;
}

int tvk_toc_get_lod(tvk_toc* t, uint32_t lod, uint64_t dims[3], uint64_t grid[3]) {
  if (!t || !dims || !grid) {
    tvk::set_last_error("null argument");
    return TVK_ERR_ARG;
  }
  if (lod >= t->impl.n_lod) {
    tvk::set_last_error("LOD index out of range");
    return TVK_ERR_ARG;
  }
  for (int i = 0; i < 3; ++i) {
    dims[i] = t->impl.lods[lod].dims[i];
    grid[i] = t->impl.lods[lod].grid[i];
  }
  return TVK_OK;
}

int tvk_toc_read_brick(tvk_toc* t, uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz, void* out,
                       uint64_t out_bytes, uint32_t* compression, float* min_value, float* max_value) {
  if (!t) {
    tvk::set_last_error("null TOC handle");
    return TVK_ERR_ARG;
  }
  const tvk::TocVolume& T = t->impl;
  if (lod >= T.n_lod || bx >= T.lods[lod].grid[0] || by >= T.lods[lod].grid[1] || bz >= T.lods[lod].grid[2]) {
    tvk::set_last_error("brick index out of range");
    return TVK_ERR_ARG;
  }
  const tvk::TocEntry& E =
      T.entries[T.lods[lod].first + ((uint64_t)bz * T.lods[lod].grid[1] + by) * T.lods[lod].grid[0] + bx];
  if (compression) *compression = E.compression;
  if (min_value) *min_value = E.mn;
  if (max_value) *max_value = E.mx;
  if (!out) return TVK_OK;
  TVK_TRY(t->impl.ReadBrick(lod, bx, by, bz, out, out_bytes));
}

int tvk_toc_get_histogram_1d(tvk_toc* t, uint64_t* out, uint64_t* n_bins) {
  if (!t || !n_bins) {
    tvk::set_last_error("null argument");
    return TVK_ERR_ARG;
  }
  const auto& h = t->impl.hist1d;
  if (h.empty()) {
    tvk::set_last_error("the file has no 1D histogram block");
    return TVK_ERR_STATE;
  }
  const uint64_t cap = *n_bins;
  *n_bins = h.size();
  if (!out) return TVK_OK;
  if (cap < h.size()) {
    tvk::set_last_error("histogram buffer too small");
    return TVK_ERR_ARG;
  }
  memcpy(out, h.data(), h.size() * 8);
  return TVK_OK;
}

int tvk_toc_get_histogram_2d(tvk_toc* t, uint64_t* out, uint64_t* n_value, uint64_t* n_grad,
                             float* max_grad) {
  if (!t || !n_value || !n_grad) {
    tvk::set_last_error("null argument");
    return TVK_ERR_ARG;
  }
  const tvk::TocVolume& T = t->impl;
  if (T.hist2d.empty()) {
    tvk::set_last_error("the file has no 2D histogram block");
    return TVK_ERR_STATE;
  }
  const uint64_t cap = *n_value * *n_grad;
  *n_value = T.hist2d_x;
  *n_grad = T.hist2d_y;
  if (max_grad) *max_grad = T.max_grad;
  if (!out) return TVK_OK;
  if (cap < T.hist2d.size()) {
    tvk::set_last_error("histogram buffer too small");
    return TVK_ERR_ARG;
  }
  memcpy(out, T.hist2d.data(), T.hist2d.size() * 8);
  return TVK_OK;
}

int tvk_toc_build_raw(int device, const char* raw_path, uint64_t header_bytes, uint64_t nx, uint64_t ny,
                      uint64_t nz, uint32_t bits, uint32_t max_brick, uint32_t overlap, int use_median,
                      int clamp_to_edge, uint32_t compression, uint32_t level, uint32_t layout,
                      const char* path, tvk_toc_info* info) {
  if (!raw_path || !path || (bits != 8 && bits != 16)) {
    tvk::set_last_error("bad raw input arguments");
    return TVK_ERR_ARG;
  }
  const int fd = open(raw_path, O_RDONLY);
  if (fd < 0) {
    tvk::set_last_error(std::string("cannot open ") + raw_path);
    return TVK_ERR_IO;
  }
  struct stat st;
  const uint64_t need = nx * ny * nz * (bits / 8);
  if (fstat(fd, &st) != 0 || nx == 0 || ny == 0 || nz == 0 || (uint64_t)st.st_size < header_bytes ||
      (uint64_t)st.st_size - header_bytes < need) {
    close(fd);
    tvk::set_last_error(std::string(raw_path) + " is smaller than the volume it should hold");
    return TVK_ERR_IO;
  }
  void* map = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (map == MAP_FAILED) {
    tvk::set_last_error(std::string("cannot map ") + raw_path);
    return TVK_ERR_IO;
  }
  const uint64_t d[3] = {nx, ny, nz};
  int rc;
  try {
    rc = tvk::toc_build(device, (const uint8_t*)map + header_bytes, d, bits, max_brick, overlap,
                        use_median != 0, clamp_to_edge != 0, compression, level, layout, path, info);
  } catch (const std::bad_alloc&) {
    tvk::set_last_error("host allocation failed");
    rc = TVK_ERR_OOM;
  }
  munmap(map, (size_t)st.st_size);
  return rc;
}

static int toc_brick_source(void* user, uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz, void* slot,
                            uint64_t slot_bytes) {
  return static_cast<tvk::TocVolume*>(user)->ReadBrick(lod, bx, by, bz, slot, slot_bytes) == TVK_OK ? 0 : -1;
}

int tvk_set_toc_volume(tvk_renderer* r, tvk_toc* t, uint64_t pool_bytes) {
  TVK_GUARD(r);
  if (!t) {
    tvk::set_last_error("null TOC handle");
    return TVK_ERR_ARG;
  }
  const tvk::TocVolume& T = t->impl;
  uint64_t planned = 0;
  if (int e = r->impl.PlannedBrickCount(T.dims, T.bits, T.payload, T.overlap, T.n_lod, &planned)) return e;
  if (planned != T.entries.size()) {
    tvk::set_last_error("TOC brick table does not match its level layout");
    return TVK_ERR_FORMAT;
  }
  std::vector<float> ranges(T.entries.size() * 2);
  for (size_t i = 0; i < T.entries.size(); ++i) {
    ranges[2 * i] = T.entries[i].mn;
    ranges[2 * i + 1] = T.entries[i].mx;
  }
  TVK_TRY(r->impl.SetPagedVolume(T.dims, T.bits, T.payload, T.overlap, T.n_lod, &toc_brick_source, &t->impl,
                                 ranges.data(), pool_bytes));
}

uint64_t tvk_brick_slot_bytes(uint32_t brick_size, uint32_t apron, uint32_t bits) {
  if (brick_size < 1 || brick_size > (1u << 20) || apron > 64 || (bits != 8 && bits != 16)) return 0;
  const int S = (int)(brick_size + 2 * apron);
  return ((uint64_t)tvk::slot_block_elems(S) * (bits / 8) + 255) & ~255ull;
}

int tvk_get_brick_minmax(tvk_renderer* r, uint32_t lod, float* mins, float* maxs, uint64_t n) {
  TVK_GUARD(r);
  return r->impl.BrickMinMax(lod, mins, maxs, n);
}

int tvk_compute_histogram_1d(tvk_renderer* r, uint64_t* hist, uint64_t n_bins) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Histogram1D(hist, n_bins));
}

int tvk_compute_histogram_2d(tvk_renderer* r, uint64_t* hist, uint32_t nv, uint32_t ng,
                             float* max_grad) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Histogram2D(hist, nv, ng, max_grad));
}

int tvk_set_render_mode(tvk_renderer* r, int m) {
  TVK_GUARD(r);
  return r->impl.SetRenderMode(m);
}

int tvk_set_1d_trans(tvk_renderer* r, const float* rgba, uint32_t n) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Set1DTrans(rgba, n));
}

int tvk_set_2d_trans(tvk_renderer* r, const float* rgba, uint32_t w, uint32_t h, float mg) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Set2DTrans(rgba, w, h, mg));
}

int tvk_set_use_mip(tvk_renderer* r, int use_mip) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.SetUseMIP(use_mip != 0));
}

int tvk_get_use_mip(tvk_renderer* r, int* use_mip) {
  TVK_GUARD(r);
  if (!use_mip) {
    tvk::set_last_error("null output");
    return TVK_ERR_ARG;
  }
  *use_mip = r->impl.GetUseMIP() ? 1 : 0;
  return TVK_OK;
}

int tvk_set_mip_rotation_angle(tvk_renderer* r, float degrees) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.SetMIPRotationAngle(degrees));
}

int tvk_set_iso_value(tvk_renderer* r, float iso) {
  TVK_GUARD(r);
  return r->impl.SetIsoValue(iso);
}

int tvk_set_isosurface_color(tvk_renderer* r, float red, float green, float blue) {
  TVK_GUARD(r);
  return r->impl.SetIsosurfaceColor(red, green, blue);
}

int tvk_set_clearview_enabled(tvk_renderer* r, int enabled) {
  TVK_GUARD(r);
  return r->impl.SetClearViewEnabled(enabled != 0);
}

int tvk_set_cv_iso_value(tvk_renderer* r, float iso) {
  TVK_GUARD(r);
  return r->impl.SetCVIsoValue(iso);
}

int tvk_set_cv_color(tvk_renderer* r, float red, float green, float blue) {
  TVK_GUARD(r);
  return r->impl.SetCVColor(red, green, blue);
}

int tvk_set_cv_size(tvk_renderer* r, float size) {
  TVK_GUARD(r);
  return r->impl.SetCVSize(size);
}

int tvk_set_cv_context_scale(tvk_renderer* r, float scale) {
  TVK_GUARD(r);
  return r->impl.SetCVContextScale(scale);
}

int tvk_set_cv_border_scale(tvk_renderer* r, float scale) {
  TVK_GUARD(r);
  return r->impl.SetCVBorderScale(scale);
}

int tvk_set_cv_focus_pos(tvk_renderer* r, int32_t x, int32_t y) {
  TVK_GUARD(r);
  return r->impl.SetCVFocusPos(x, y);
}

int tvk_get_clearview(tvk_renderer* r, tvk_clearview* out) {
  TVK_GUARD(r);
  if (!out) {
    tvk::set_last_error("null output");
    return TVK_ERR_ARG;
  }
  r->impl.GetClearView(out);
  return TVK_OK;
}

int tvk_set_view_pos(tvk_renderer* r, float x, float y, float z) {
  TVK_GUARD(r);
  const float v[3] = {x, y, z};
  return r->impl.SetViewPos(v);
}

int tvk_set_view_dir(tvk_renderer* r, float x, float y, float z) {
  TVK_GUARD(r);
  const float v[3] = {x, y, z};
  return r->impl.SetViewDir(v);
}

int tvk_set_up_dir(tvk_renderer* r, float x, float y, float z) {
  TVK_GUARD(r);
  const float v[3] = {x, y, z};
  return r->impl.SetUpDir(v);
}

int tvk_set_fov(tvk_renderer* r, float fov) {
  TVK_GUARD(r);
  return r->impl.SetFoV(fov);
}

int tvk_set_ortho_view(tvk_renderer* r, int enabled, float hh) {
  TVK_GUARD(r);
  return r->impl.SetOrtho(enabled != 0, hh);
}

int tvk_set_rotation_4x4(tvk_renderer* r, const float m[16]) {
  TVK_GUARD(r);
  if (!m) return TVK_ERR_ARG;
  return r->impl.SetRotation(m);
}

int tvk_set_translation_4x4(tvk_renderer* r, const float m[16]) {
  TVK_GUARD(r);
  if (!m) return TVK_ERR_ARG;
  return r->impl.SetTranslation(m);
}

int tvk_set_rescale_factors(tvk_renderer* r, double sx, double sy, double sz) {
  TVK_GUARD(r);
  const double s[3] = {sx, sy, sz};
  return r->impl.SetRescaleFactors(s);
}

int tvk_set_sample_rate_modifier(tvk_renderer* r, float rate) {
  TVK_GUARD(r);
  return r->impl.SetSampleRateModifier(rate);
}

int tvk_set_interpolant(tvk_renderer* r, int i) {
  TVK_GUARD(r);
  return r->impl.SetInterpolant(i);
}

int tvk_set_lod_limits(tvk_renderer* r, uint32_t mn, uint32_t mx) {
  TVK_GUARD(r);
  return r->impl.SetLODLimits(mn, mx);
}

int tvk_set_bg_colors(tvk_renderer* r, const float top[3], const float bottom[3]) {
  TVK_GUARD(r);
  if (!top || !bottom) return TVK_ERR_ARG;
  return r->impl.SetBGColors(top, bottom);
}

int tvk_set_lighting_enabled(tvk_renderer* r, int on) {
  TVK_GUARD(r);
  return r->impl.SetLightingEnabled(on != 0);
}

int tvk_set_light_colors(tvk_renderer* r, const float a[4], const float d[4], const float s[4],
                         const float dir[3]) {
  TVK_GUARD(r);
  if (!a || !d || !s || !dir) return TVK_ERR_ARG;
  return r->impl.SetLightColors(a, d, s, dir);
}

int tvk_set_early_ray_termination(tvk_renderer* r, float a) {
  TVK_GUARD(r);
  return r->impl.SetEarlyRayTermination(a);
}

int tvk_set_empty_space_skipping(tvk_renderer* r, int on) {
  TVK_GUARD(r);
  return r->impl.SetEmptySpaceSkipping(on != 0);
}

int tvk_render_batch_noalloc(tvk_renderer* r, const tvk_view* views, uint32_t n, void* const* device_rgba,
                             tvk_view_stats* stats) {
  TVK_GUARD(r);
  return r->impl.RenderBatchNoAlloc(views, n, device_rgba, stats);
}

int tvk_set_clip_plane(tvk_renderer* r, const float plane[4]) {
  TVK_GUARD(r);
  return r->impl.SetClipPlane(plane);
}

int tvk_enable_clip_plane(tvk_renderer* r, int enabled) {
  TVK_GUARD(r);
  return r->impl.EnableClipPlane(enabled != 0);
}

int tvk_is_clip_plane_enabled(tvk_renderer* r, int* enabled) {
  TVK_GUARD(r);
  if (!enabled) {
    tvk::set_last_error("null output pointer");
    return TVK_ERR_ARG;
  }
  *enabled = r->impl.ClipPlaneEnabled() ? 1 : 0;
  return TVK_OK;
}

int tvk_get_clip_plane(tvk_renderer* r, float plane[4]) {
  TVK_GUARD(r);
  if (!plane) {
    tvk::set_last_error("null output pointer");
    return TVK_ERR_ARG;
  }
  r->impl.GetClipPlane(plane);
  return TVK_OK;
}

int tvk_set_blend_precision(tvk_renderer* r, int precision) {
  TVK_GUARD(r);
  return r->impl.SetBlendPrecision(precision);
}

int tvk_get_blend_precision(tvk_renderer* r, int* precision) {
  TVK_GUARD(r);
  if (!precision) {
    tvk::set_last_error("null output pointer");
    return TVK_ERR_ARG;
  }
  *precision = r->impl.BlendPrecision();
  return TVK_OK;
}

int tvk_set_perf_measures(tvk_renderer* r, uint32_t min_framerate, int low_res_intermediate,
                          float screen_res_dec_factor, float sample_dec_factor, uint32_t lod_delay) {
  TVK_GUARD(r);
  return r->impl.SetPerfMeasures(min_framerate, low_res_intermediate != 0, screen_res_dec_factor,
                                 sample_dec_factor, lod_delay);
}

int tvk_read_framebuffer(tvk_renderer* r, uint8_t* rgba, uint64_t rgba_bytes) {
  TVK_GUARD(r);
  return r->impl.ReadFramebuffer(rgba, rgba_bytes);
}

int tvk_set_kernel_tuning(tvk_renderer* r, uint32_t batch, uint32_t tile_group,
                          uint32_t xcd_split) {
  TVK_GUARD(r);
  return r->impl.SetKernelTuning(batch, tile_group, xcd_split);
}

int tvk_set_work_group_waves(tvk_renderer* r, uint32_t waves) {
  TVK_GUARD(r);
  return r->impl.SetWorkGroupWaves(waves);
}

int tvk_set_brick_staging(tvk_renderer* r, int on) {
  TVK_GUARD(r);
  return r->impl.SetBrickStaging(on != 0);
}

int tvk_set_tf_grouping(tvk_renderer* r, int on) {
  TVK_GUARD(r);
  return r->impl.SetTFGrouping(on != 0);
}

int tvk_set_ray_tile_shape(tvk_renderer* r, uint32_t tile_width) {
  TVK_GUARD(r);
  return r->impl.SetRayTileShape(tile_width);
}

int tvk_set_empty_space_cell(tvk_renderer* r, uint32_t cell) {
  TVK_GUARD(r);
  return r->impl.SetEmptySpaceCell(cell);
}

int tvk_paint(tvk_renderer* r) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Paint());
}

int tvk_check_for_redraw(tvk_renderer* r) {
  if (!r) return 0;
  return r->impl.CheckForRedraw();
}

int tvk_schedule_complete_redraw(tvk_renderer* r) {
  TVK_GUARD(r);
  return r->impl.ScheduleCompleteRedraw();
}

int tvk_capture_single_frame(tvk_renderer* r, uint8_t* rgba, uint64_t bytes, int preserve) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Capture(rgba, bytes, preserve != 0));
}

int tvk_get_progress(tvk_renderer* r, uint64_t* sf, uint32_t* wsf, uint32_t* bc, uint32_t* wb,
                     uint64_t* ml) {
  TVK_GUARD(r);
  return r->impl.Progress(sf, wsf, bc, wb, ml);
}

int tvk_get_frame_stats(tvk_renderer* r, tvk_frame_stats* out) {
  TVK_GUARD(r);
  return r->impl.FrameStats(out);
}

int tvk_get_device_framebuffer(tvk_renderer* r, void** dptr, void** stream) {
  TVK_GUARD(r);
  return r->impl.DeviceFramebuffer(dptr, stream);
}

int tvk_synchronize(tvk_renderer* r) {
  TVK_GUARD(r);
  return r->impl.Synchronize();
}

int tvk_set_1d_trans_slot(tvk_renderer* r, uint32_t slot, const float* rgba, uint32_t n) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Set1DTransSlot(slot, rgba, n));
}

int tvk_set_2d_trans_slot(tvk_renderer* r, uint32_t slot, const float* rgba, uint32_t w, uint32_t h,
                          float max_grad) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Set2DTransSlot(slot, rgba, w, h, max_grad));
}

int tvk_render_batch(tvk_renderer* r, const tvk_view* views, uint32_t n, tvk_view_stats* stats,
                     float* kernel_ms) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.RenderBatch(views, n, stats, kernel_ms));
}

int tvk_get_batch_framebuffer(tvk_renderer* r, uint32_t index, void** dptr) {
  TVK_GUARD(r);
  return r->impl.BatchFramebuffer(index, dptr);
}

int tvk_brick_histograms(tvk_renderer* r, uint32_t lod, uint32_t bins, uint32_t* out, uint64_t n) {
  TVK_GUARD(r);
  return r->impl.BrickHistograms(lod, bins, out, n);
}

int tvk_region_detail(tvk_renderer* r, float* out, uint64_t n) {
  TVK_GUARD(r);
  return r->impl.RegionDetail(out, n);
}

int tvk_set_region_lods(tvk_renderer* r, const uint8_t* levels, uint64_t n) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.SetRegionLods(levels, n));
}

int tvk_set_region_output(tvk_renderer* r, int enabled) {
  TVK_GUARD(r);
  return r->impl.SetRegionOutput(enabled != 0);
}

int tvk_read_region_output(tvk_renderer* r, int32_t* out, uint64_t n) {
  TVK_GUARD(r);
  return r->impl.ReadRegionOutput(out, n);
}

int tvk_set_region_check(tvk_renderer* r, uint32_t tile_sample, float max_rmse) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.SetRegionCheck(tile_sample, max_rmse));
}

int tvk_get_region_check(tvk_renderer* r, tvk_region_check* out) {
  TVK_GUARD(r);
  return r->impl.RegionCheckStats(out);
}

int tvk_autotune(tvk_renderer* r, const tvk_view* views, uint32_t n, uint32_t flags,
                 tvk_tuning* out) {
  TVK_GUARD(r);
  TVK_TRY(r->impl.Autotune(views, n, flags, out));
}

int tvk_server_create(int device, uint64_t arena_bytes, uint32_t lanes, tvk_server** out) {
  if (!out) {
    tvk::set_last_error("null output pointer");
    return TVK_ERR_ARG;
  }
  *out = nullptr;
  try {
    tvk_server* s = new tvk_server();
    if (int e = s->impl.Init(device, arena_bytes, lanes)) {
      delete s;
      return e;
    }
    *out = s;
    return TVK_OK;
  } catch (const std::bad_alloc&) {
    tvk::set_last_error("host allocation failed");
    return TVK_ERR_OOM;
  }
}

int tvk_server_destroy(tvk_server* s) {
  delete s;
  return TVK_OK;
}

int tvk_server_add_volume(tvk_server* s, tvk_renderer* r, uint32_t* id) {
  if (!s || !r) {
    tvk::set_last_error("null server or renderer handle");
    return TVK_ERR_ARG;
  }
  TVK_TRY(s->impl.AddVolume(&r->impl, id));
}

int tvk_server_submit(tvk_server* s, const tvk_request* q, uint64_t* ticket) {
  if (!s || !q) {
    tvk::set_last_error("null server handle or request");
    return TVK_ERR_ARG;
  }
  TVK_TRY(s->impl.Submit(*q, ticket));
}

int tvk_server_wait(tvk_server* s, uint64_t ticket, uint32_t timeout_ms, tvk_request_result* out) {
  if (!s) {
    tvk::set_last_error("null server handle");
    return TVK_ERR_ARG;
  }
  TVK_TRY(s->impl.Wait(ticket, timeout_ms, out));
}

int tvk_server_get_stats(tvk_server* s, tvk_server_stats* out) {
  if (!s) {
    tvk::set_last_error("null server handle");
    return TVK_ERR_ARG;
  }
  return s->impl.Stats(out);
}

int tvk_server_set_part_cost(tvk_server* s, uint64_t samples) {
  if (!s) {
    tvk::set_last_error("null server handle");
    return TVK_ERR_ARG;
  }
  return s->impl.SetPartCost(samples);
}

static const uint64_t kMlpWeights = 32 * 64 + 3 * 64 * 64 + 64, kMlpBiases = 4 * 64 + 1;

int tvk_lod_mlp_create(int device, const uint16_t* w, const float* b, tvk_lod_mlp** out) {
  if (!w || !b || !out) {
    tvk::set_last_error("null weights, biases or output");
    return TVK_ERR_ARG;
  }
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    tvk::set_last_error("device ordinal out of range");
    return TVK_ERR_ARG;
  }
  tvk_lod_mlp* m = new (std::nothrow) tvk_lod_mlp();
  if (!m) return TVK_ERR_OOM;
  m->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&m->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&m->ev[1]);
  std::vector<uint16_t> packed(tvk::lod_mlp_packed_elems());
  tvk::lod_mlp_pack(w, packed.data());
  if (e == hipSuccess) e = m->w.alloc(packed.size() * 2);
  if (e == hipSuccess) e = m->b.alloc(kMlpBiases * 4);
  if (e == hipSuccess) e = hipMemcpy(m->w.p, packed.data(), packed.size() * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->b.p, b, kMlpBiases * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    tvk::set_last_error(std::string("lod mlp setup: ") + hipGetErrorString(e));
    tvk_lod_mlp_destroy(m);
    return TVK_ERR_HIP;
  }
  *out = m;
  return TVK_OK;
}

int tvk_lod_mlp_predict(tvk_lod_mlp* m, const float* x, uint64_t n, float* y, float* kernel_ms) {
  if (!m || (n && (!x || !y)) || n > (1ull << 30)) {
    tvk::set_last_error("bad lod mlp arguments");
    return TVK_ERR_ARG;
  }
  if (n == 0) return TVK_OK;
  hipError_t e = hipSetDevice(m->device);
  if (e == hipSuccess && m->x.bytes < n * 32 * 4) e = m->x.alloc(n * 32 * 4);
  if (e == hipSuccess && m->y.bytes < n * 4) e = m->y.alloc(n * 4);
  if (e == hipSuccess) e = hipMemcpyAsync(m->x.p, x, n * 32 * 4, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipEventRecord(m->ev[0], m->stream);
  if (e == hipSuccess)
    e = tvk::launch_lod_mlp(m->w.as<uint16_t>(), m->b.as<float>(), m->x.as<float>(), m->y.as<float>(),
                            (int)n, m->stream);
  if (e == hipSuccess) e = hipEventRecord(m->ev[1], m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(y, m->y.p, n * 4, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e == hipSuccess && kernel_ms) e = hipEventElapsedTime(kernel_ms, m->ev[0], m->ev[1]);
  if (e != hipSuccess) {
    tvk::set_last_error(std::string("lod mlp predict: ") + hipGetErrorString(e));
    return TVK_ERR_HIP;
  }
  return TVK_OK;
}

int tvk_lod_mlp_destroy(tvk_lod_mlp* m) {
  if (!m) return TVK_OK;
  if (m->device >= 0) {
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    m->w.reset();
    m->b.reset();
    m->x.reset();
    m->y.reset();
    for (auto& ev : m->ev)
      if (ev) (void)hipEventDestroy(ev);
    if (m->stream) (void)hipStreamDestroy(m->stream);
  }
  delete m;
  return TVK_OK;
}

int tvk_set_launch_graphs(tvk_renderer* r, int enabled) {
  TVK_GUARD(r);
  return r->impl.SetLaunchGraphs(enabled != 0);
}

int tvk_set_max_usable_gpu_mem(tvk_renderer* r, uint64_t bytes) {
  TVK_GUARD(r);
  return r->impl.SetMaxUsableGPUMem(bytes);
}

int tvk_get_mem_stats(tvk_renderer* r, tvk_mem_stats* out) {
  TVK_GUARD(r);
  return r->impl.MemStats(out);
}

int tvk_set_max_usable_cpu_mem(tvk_renderer* r, uint64_t bytes) {
  TVK_GUARD(r);
  return r->impl.SetMaxUsableCPUMem(bytes);
}

int tvk_get_memman_stats(int device, tvk_memman_stats* out) {
  if (!out) {
    tvk::set_last_error("null stats pointer");
    return TVK_ERR_ARG;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    tvk::set_last_error("device ordinal out of range");
    return TVK_ERR_ARG;
  }
  tvk::MemManStats m;
  tvk::MemMan::ForDevice(device).Stats(&m);
  *out = tvk_memman_stats{m.gpu_total, m.gpu_budget, m.gpu_used, m.cpu_total, m.cpu_budget,
                          m.cpu_used, m.pool_bytes, m.pool_slots, m.pool_resident, m.pool_evictions,
                          m.chunks_freed, m.cache_bytes, m.cache_hits, m.cache_misses};
  return TVK_OK;
}

int tvk_set_brick_source_threads(tvk_renderer* r, uint32_t threads) {
  TVK_GUARD(r);
  return r->impl.SetBrickSourceThreads(threads);
}

int tvk_get_paging_stats(tvk_renderer* r, uint64_t* bytes, double* ms) {
  TVK_GUARD(r);
  if (!bytes || !ms) {
    tvk::set_last_error("null output pointer");
    return TVK_ERR_ARG;
  }
  r->impl.PagingStats(bytes, ms);
  return TVK_OK;
}

}  // extern "C"
