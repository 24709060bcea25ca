// renderer.cpp -- HipRenderer (see renderer.hpp).
#include "renderer.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>

#include "errors.hpp"
#include "kernels.hpp"
#include "uvf.hpp"

namespace tvk {

#define TVK_HIP(call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return hip_fail(e_, #call);                     \
  } while (0)

static int hip_fail(hipError_t e, const char* what) {
  set_last_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? TVK_ERR_OOM : TVK_ERR_HIP;
}

static int arg_fail(const std::string& msg) {
  set_last_error(msg);
  return TVK_ERR_ARG;
}

static int state_fail(const std::string& msg) {
  set_last_error(msg);
  return TVK_ERR_STATE;
}

hipError_t DevBuf::alloc(size_t n) {
  reset();
  if (n == 0) return hipSuccess;
  hipError_t e = hipMalloc(&p, n);
  if (e == hipSuccess) bytes = n;
  else p = nullptr;
  return e;
}

void DevBuf::reset() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

// ------------------------------------------------------------ small math
namespace {

void mat_identity(double m[16]) {
  for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0 : 0.0;
}

void mat_mul(const double a[16], const double b[16], double o[16]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
      o[i * 4 + j] = s;
    }
}

bool mat_inverse(const double m[16], double out[16]) {
  double a[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m[i * 4 + j] : (j - 4 == i ? 1.0 : 0.0);
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
    if (fabs(a[piv][c]) < 1e-300) return false;
    if (piv != c)
      for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[piv][j]);
    const double d = a[c][c];
    for (int j = 0; j < 8; ++j) a[c][j] /= d;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[i * 4 + j] = a[i][j + 4];
  return true;
}

void xform(const double m[16], const double p[3], double w, double o[3]) {
  for (int j = 0; j < 3; ++j) o[j] = p[0] * m[j] + p[1] * m[4 + j] + p[2] * m[8 + j] + w * m[12 + j];
}

void cross(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

void normalize(double a[3]) {
  const double l = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  if (l > 0)
    for (int i = 0; i < 3; ++i) a[i] /= l;
}

bool finite3(const float v[3]) { return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]); }

}  // namespace

// -------------------------------------------------------------- lifecycle
HipRenderer::HipRenderer(int type, bool pow2, bool to8, bool no_border, bool bias_scale_tf)
    : type_(type), pow2_(pow2), to8_(to8), no_border_(no_border), bias_scale_tf_(bias_scale_tf) {
  mat_identity(rot_);
  mat_identity(trans_);
}

HipRenderer::~HipRenderer() {
  if (initialized_) {
    (void)hipSetDevice(device_);
    (void)hipStreamSynchronize(stream_);
    for (auto& e : ev_)
      if (e) (void)hipEventDestroy(e);
    if (counters_h_) (void)hipHostFree(counters_h_);
    for (auto& l : levels_) l.reset();
    pool_.slots.reset();
    pool_.page_table.reset();
    pool_.brick_range.reset();
    scratch_.reset();
    cell_range_.reset();
    tf_dev_.reset();
    tf_prefix_dev_.reset();
    framebuffer_.reset();
    counters_.reset();
    if (stream_) (void)hipStreamDestroy(stream_);
  }
}

int HipRenderer::Initialize(int device) {
  if (initialized_) return state_fail("renderer already initialized");
  int n = 0;
  TVK_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return arg_fail("device ordinal out of range");
  device_ = device;
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  for (auto& e : ev_) TVK_HIP(hipEventCreate(&e));
  // [0] evaluated samples, [1] nominal samples, [2..9] per-XCD tile queues
  TVK_HIP(counters_.alloc(16 * sizeof(unsigned long long)));
  TVK_HIP(hipDeviceGetAttribute(&compute_units_, hipDeviceAttributeMultiprocessorCount, device_));
  TVK_HIP(hipHostMalloc((void**)&counters_h_, 4 * sizeof(unsigned long long), hipHostMallocDefault));
  memset(counters_h_, 0, 4 * sizeof(unsigned long long));
  size_t free_b = 0, total_b = 0;
  TVK_HIP(hipMemGetInfo(&free_b, &total_b));
  gpu_total_ = total_b;
  if (gpu_budget_ == 0) gpu_budget_ = total_b;
  initialized_ = true;
  return TVK_OK;
}

int HipRenderer::Resize(uint32_t w, uint32_t h) {
  if (!initialized_) return state_fail("resize before initialize");
  if (w == 0 || h == 0 || w > 16384 || h > 16384) return arg_fail("window size out of range");
  if (w == width_ && h == height_) return TVK_OK;
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipStreamSynchronize(stream_));
  TVK_HIP(framebuffer_.alloc((size_t)w * h * 4));
  width_ = w;
  height_ = h;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetRendererTarget(int t) {
  if (t != TVK_RT_INTERACTIVE && t != TVK_RT_CAPTURE && t != TVK_RT_HEADLESS)
    return arg_fail("unknown renderer target");
  target_ = t;
  return TVK_OK;
}

// ------------------------------------------------------------------- data
int HipRenderer::LoadDataset(const char* path, bool verify) {
  if (!initialized_) return state_fail("loadDataset before initialize");
  if (!path) return arg_fail("null path");
  UvfFile f;
  std::string err;
  int rc = uvf_read(path, verify, &f, &err);
  if (rc) {
    set_last_error(err);
    return rc;
  }
  std::vector<uint8_t> vox;
  rc = uvf_read_voxels(path, f.rasters[0], &vox, &err);
  if (rc) {
    set_last_error(err);
    return rc;
  }
  const UvfRaster& r = f.rasters[0];
  // the stored overlap is the total of both sides (UVFReader/DataSource.h:537-539)
  const uint32_t apron = (uint32_t)std::max<uint64_t>(2, r.overlap[0] / 2);
  uint64_t bs = std::max(r.dims[0], std::max(r.dims[1], r.dims[2]));
  return SetVolume(vox.data(), r.dims, r.bits, (uint32_t)bs, apron, 1);
}

int HipRenderer::SetVolume(const void* voxels, const uint64_t dims[3], uint32_t bits,
                           uint32_t brick, uint32_t apron, uint32_t n_lod) {
  if (!initialized_) return state_fail("setVolume before initialize");
  if (!voxels) return arg_fail("null voxel pointer");
  if (bits != 8 && bits != 16) return arg_fail("bits must be 8 or 16");
  for (int i = 0; i < 3; ++i)
    if (dims[i] == 0 || dims[i] > (1u << 20)) return arg_fail("volume dimension out of range");
  if (brick < 8 || brick > 1024) return arg_fail("brick size must be in [8, 1024]");
  if (apron < 1 || apron > 8) return arg_fail("apron must be in [1, 8]");
  const uint64_t S = (uint64_t)brick + 2ull * apron;
  if (S * S * S >= (1ull << 31)) return arg_fail("brick slot too large");
  const int bpv = (int)bits / 8;

  // LOD count: RasterDataBlock rule when not given (UVFReader/DataSource.h:521-529)
  int nl = (int)n_lod;
  if (nl == 0) {
    nl = 1;
    uint64_t m = std::max(dims[0], std::max(dims[1], dims[2]));
    while (m > brick) {
      m /= 2;
      ++nl;
    }
  }
  if (nl > kMaxLod) return arg_fail("too many LOD levels");
  uint64_t ld[kMaxLod][3];
  for (int i = 0; i < 3; ++i) ld[0][i] = dims[i];
  for (int l = 1; l < nl; ++l)
    for (int i = 0; i < 3; ++i) ld[l][i] = (ld[l - 1][i] + 1) / 2;

  // empty-space cell edge: 16 voxels when it divides the brick, else the brick
  const int cell = (brick % 16 == 0) ? 16 : (int)brick;
  LodInfo lods[kMaxLod];
  int total = 0, total_cells = 0;
  uint64_t level_bytes = 0;
  for (int l = 0; l < nl; ++l) {
    for (int i = 0; i < 3; ++i) {
      lods[l].dims[i] = (int32_t)ld[l][i];
      lods[l].grid[i] = (int32_t)((ld[l][i] + brick - 1) / brick);
      lods[l].to_level[i] = (float)((double)ld[l][i] / (double)dims[i]);
      lods[l].cells[i] = (int32_t)((ld[l][i] + cell - 1) / cell);
    }
    lods[l].brick_base = total;
    lods[l].cell_base = total_cells;
    total += lods[l].grid[0] * lods[l].grid[1] * lods[l].grid[2];
    total_cells += lods[l].cells[0] * lods[l].cells[1] * lods[l].cells[2];
    level_bytes += ld[l][0] * ld[l][1] * ld[l][2] * bpv;
  }
  const uint64_t slot_bytes = ((S * S * S * bpv) + 255) & ~255ull;
  const uint64_t need = level_bytes + (uint64_t)total * slot_bytes + (uint64_t)total * 12 +
                        (uint64_t)total_cells * 4 +
                        (uint64_t)total * sizeof(BrickJob) + (size_t)width_ * height_ * 4;
  if (need > gpu_budget_) {
    set_last_error("volume needs " + std::to_string(need) + " bytes of device memory, budget is " +
                   std::to_string(gpu_budget_));
    return TVK_ERR_OOM;
  }

  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipStreamSynchronize(stream_));
  has_volume_ = false;
  for (auto& l : levels_) l.reset();
  pool_.slots.reset();
  pool_.page_table.reset();
  pool_.brick_range.reset();
  pool_.n_slots = pool_.slot_bytes = pool_.resident = 0;
  pool_.page_table_h.clear();
  pool_.brick_range_h.clear();

  for (int l = 0; l < nl; ++l) TVK_HIP(levels_[l].alloc(ld[l][0] * ld[l][1] * ld[l][2] * bpv));
  TVK_HIP(hipMemcpyAsync(levels_[0].p, voxels, levels_[0].bytes, hipMemcpyHostToDevice, stream_));
  for (int l = 1; l < nl; ++l) {
    const int64_t sd[3] = {(int64_t)ld[l - 1][0], (int64_t)ld[l - 1][1], (int64_t)ld[l - 1][2]};
    TVK_HIP(launch_downsample(levels_[l - 1].p, bpv, sd, levels_[l].p, stream_));
  }

  TVK_HIP(pool_.slots.alloc((size_t)total * slot_bytes));
  TVK_HIP(pool_.page_table.alloc((size_t)total * sizeof(int32_t)));
  TVK_HIP(pool_.brick_range.alloc((size_t)total * 2 * sizeof(float)));
  pool_.n_slots = (uint64_t)total;
  pool_.slot_bytes = slot_bytes;
  std::vector<BrickJob> jobs((size_t)total);
  for (int l = 0; l < nl; ++l) {
    const LodInfo& L = lods[l];
    for (int bz = 0; bz < L.grid[2]; ++bz)
      for (int by = 0; by < L.grid[1]; ++by)
        for (int bx = 0; bx < L.grid[0]; ++bx) {
          const int id = L.brick_base + (bz * L.grid[1] + by) * L.grid[0] + bx;
          jobs[id] = BrickJob{bx, by, bz, id, (int64_t)id};
        }
  }
  DevBuf djobs;
  TVK_HIP(djobs.alloc(jobs.size() * sizeof(BrickJob)));
  TVK_HIP(hipMemcpyAsync(djobs.p, jobs.data(), djobs.bytes, hipMemcpyHostToDevice, stream_));
  TVK_HIP(scratch_.alloc((size_t)total * 2 * sizeof(uint32_t)));
  const int64_t stride = (int64_t)(slot_bytes / bpv);
  for (int l = 0; l < nl; ++l) {
    const LodInfo& L = lods[l];
    const int n = L.grid[0] * L.grid[1] * L.grid[2];
    const int64_t d[3] = {L.dims[0], L.dims[1], L.dims[2]};
    TVK_HIP(launch_extract_bricks(levels_[l].p, bpv, d, djobs.as<BrickJob>() + L.brick_base, n,
                                  pool_.slots.p, stride, (int)brick, (int)apron,
                                  pool_.brick_range.as<float>(), scratch_.as<uint32_t>(), stream_));
  }
  TVK_HIP(cell_range_.alloc((size_t)total_cells * sizeof(uint32_t)));
  for (int l = 0; l < nl; ++l) {
    const int64_t d[3] = {lods[l].dims[0], lods[l].dims[1], lods[l].dims[2]};
    TVK_HIP(launch_cell_ranges(levels_[l].p, bpv, d, cell,
                               cell_range_.as<uint32_t>() + lods[l].cell_base, stream_));
  }
  pool_.page_table_h.resize(total);
  for (int i = 0; i < total; ++i) pool_.page_table_h[i] = i;
  TVK_HIP(hipMemcpyAsync(pool_.page_table.p, pool_.page_table_h.data(), total * sizeof(int32_t),
                         hipMemcpyHostToDevice, stream_));
  pool_.brick_range_h.resize((size_t)total * 2);
  TVK_HIP(hipMemcpyAsync(pool_.brick_range_h.data(), pool_.brick_range.p, total * 2 * sizeof(float),
                         hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  pool_.resident = (uint64_t)total;
  pool_.uploaded += (uint64_t)total;

  bits_ = bits;
  for (int i = 0; i < 3; ++i) dims0_[i] = dims[i];
  n_lod_ = nl;
  brick_ = (int)brick;
  apron_ = (int)apron;
  slot_edge_ = (int)S;
  cell_ = cell;
  for (int l = 0; l < nl; ++l) lods_[l] = lods[l];
  total_bricks_ = total;
  lod_min_ = 0;
  lod_max_ = (uint32_t)(nl - 1);
  has_volume_ = true;
  if (!tf1d_.empty()) tf_dirty_ = true;
  invalidate();
  return TVK_OK;
}

int HipRenderer::BitWidth(uint32_t* bits) const {
  if (!has_volume_) return state_fail("no volume loaded");
  *bits = bits_;
  return TVK_OK;
}

int HipRenderer::DomainSize(uint32_t lod, uint64_t d[3]) const {
  if (!has_volume_) return state_fail("no volume loaded");
  if ((int)lod >= n_lod_) return arg_fail("LOD index out of range");
  for (int i = 0; i < 3; ++i) d[i] = (uint64_t)lods_[lod].dims[i];
  return TVK_OK;
}

int HipRenderer::BrickLayout(uint32_t lod, uint32_t grid[3], uint32_t* brick, uint32_t* apron) const {
  if (!has_volume_) return state_fail("no volume loaded");
  if ((int)lod >= n_lod_) return arg_fail("LOD index out of range");
  for (int i = 0; i < 3; ++i) grid[i] = (uint32_t)lods_[lod].grid[i];
  if (brick) *brick = (uint32_t)brick_;
  if (apron) *apron = (uint32_t)apron_;
  return TVK_OK;
}

int HipRenderer::ReadLod(uint32_t lod, void* out, uint64_t bytes) {
  if (!has_volume_) return state_fail("no volume loaded");
  if ((int)lod >= n_lod_) return arg_fail("LOD index out of range");
  if (!out || bytes < levels_[lod].bytes) return arg_fail("output buffer too small");
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipMemcpyAsync(out, levels_[lod].p, levels_[lod].bytes, hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

int HipRenderer::ReadBrick(uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz, void* out,
                           uint64_t bytes) {
  if (!has_volume_) return state_fail("no volume loaded");
  if ((int)lod >= n_lod_) return arg_fail("LOD index out of range");
  const LodInfo& L = lods_[lod];
  if ((int)bx >= L.grid[0] || (int)by >= L.grid[1] || (int)bz >= L.grid[2])
    return arg_fail("brick index out of range");
  const uint64_t n = (uint64_t)slot_edge_ * slot_edge_ * slot_edge_ * (bits_ / 8);
  if (!out || bytes < n) return arg_fail("output buffer too small");
  const int gid = L.brick_base + ((int)bz * L.grid[1] + (int)by) * L.grid[0] + (int)bx;
  const int slot = pool_.page_table_h[gid];
  if (slot < 0) return state_fail("brick is not resident");
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipMemcpyAsync(out, (const uint8_t*)pool_.slots.p + (uint64_t)slot * pool_.slot_bytes, n,
                         hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

int HipRenderer::BrickMinMax(uint32_t lod, float* mins, float* maxs, uint64_t n) {
  if (!has_volume_) return state_fail("no volume loaded");
  if ((int)lod >= n_lod_) return arg_fail("LOD index out of range");
  const LodInfo& L = lods_[lod];
  const uint64_t cnt = (uint64_t)L.grid[0] * L.grid[1] * L.grid[2];
  if (!mins || !maxs || n < cnt) return arg_fail("output buffers too small");
  for (uint64_t i = 0; i < cnt; ++i) {
    mins[i] = pool_.brick_range_h[2 * (L.brick_base + i)];
    maxs[i] = pool_.brick_range_h[2 * (L.brick_base + i) + 1];
  }
  return TVK_OK;
}

int HipRenderer::Histogram1D(uint64_t* hist, uint64_t n_bins) {
  if (!has_volume_) return state_fail("no volume loaded");
  const uint64_t nb = 1ull << bits_;
  if (!hist || n_bins != nb) return arg_fail("histogram must have 2^bits bins");
  TVK_HIP(hipSetDevice(device_));
  DevBuf h;
  TVK_HIP(h.alloc(nb * 8));
  const int64_t n = (int64_t)(dims0_[0] * dims0_[1] * dims0_[2]);
  TVK_HIP(launch_histogram_1d(levels_[0].p, bits_ / 8, n, h.as<unsigned long long>(), stream_));
  TVK_HIP(hipMemcpyAsync(hist, h.p, nb * 8, hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

int HipRenderer::Histogram2D(uint64_t* hist, uint32_t n_value, uint32_t n_grad, float* max_grad) {
  if (!has_volume_) return state_fail("no volume loaded");
  if (!hist || n_value == 0 || n_grad < 2 || n_value > (1u << bits_) || n_grad > 65536 ||
      (n_value & (n_value - 1)))
    return arg_fail("bad 2D histogram size (value bins: power of two <= 2^bits)");
  TVK_HIP(hipSetDevice(device_));
  DevBuf h, km;
  TVK_HIP(h.alloc((size_t)n_value * n_grad * 8));
  TVK_HIP(km.alloc(8));
  const int64_t d[3] = {(int64_t)dims0_[0], (int64_t)dims0_[1], (int64_t)dims0_[2]};
  TVK_HIP(launch_gradient_kmax(levels_[0].p, bits_ / 8, d, km.as<unsigned long long>(), stream_));
  TVK_HIP(launch_histogram_2d(levels_[0].p, bits_ / 8, d, km.as<unsigned long long>(), (int)n_value,
                              (int)n_grad, h.as<unsigned long long>(), stream_));
  unsigned long long kmax = 0;
  TVK_HIP(hipMemcpyAsync(hist, h.p, h.bytes, hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipMemcpyAsync(&kmax, km.p, 8, hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  if (max_grad) {
    const double maxval = (double)((1u << bits_) - 1);
    *max_grad = (float)(sqrt((double)kmax) / (2.0 * maxval));
  }
  return TVK_OK;
}

// ------------------------------------------------------------ mode / TFs
int HipRenderer::SetRenderMode(int m) {
  if (m != TVK_RM_1DTRANS && m != TVK_RM_2DTRANS && m != TVK_RM_ISOSURFACE)
    return arg_fail("unknown render mode");
  render_mode_ = m;
  invalidate();
  return TVK_OK;
}

int HipRenderer::Set1DTrans(const float* rgba, uint32_t n) {
  if (!rgba || n < 1 || n > 65536) return arg_fail("1D TF needs 1..65536 entries");
  for (uint64_t i = 0; i < 4ull * n; ++i)
    if (!isfinite(rgba[i])) return arg_fail("1D TF holds a non-finite value");
  tf1d_.assign(rgba, rgba + 4ull * n);
  tf_dirty_ = true;
  invalidate();
  return TVK_OK;
}

int HipRenderer::Set2DTrans(const float* rgba, uint32_t w, uint32_t h, float max_grad) {
  if (!rgba || w < 1 || h < 1 || w > 65536 || h > 4096 || (uint64_t)w * h > (1ull << 24))
    return arg_fail("2D TF size out of range");
  if (!(max_grad > 0.f) || !isfinite(max_grad)) return arg_fail("2D TF max gradient must be > 0");
  for (uint64_t i = 0; i < 4ull * w * h; ++i)
    if (!isfinite(rgba[i])) return arg_fail("2D TF holds a non-finite value");
  tf2d_.assign(rgba, rgba + 4ull * w * h);
  tf2d_w_ = w;
  tf2d_h_ = h;
  tf2d_max_grad_ = max_grad;
  tf2d_dirty_ = true;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetIsoValue(float v) {
  if (!isfinite(v)) return arg_fail("iso value must be finite");
  iso_value_ = v;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetIsosurfaceColor(float r, float g, float b) {
  iso_color_[0] = r;
  iso_color_[1] = g;
  iso_color_[2] = b;
  invalidate();
  return TVK_OK;
}

int HipRenderer::upload_tf() {
  if (tf2d_dirty_) {
    // 2D raster + prefix count over columns holding any alpha > 0
    const size_t W = tf2d_w_, H = tf2d_h_;
    std::vector<int32_t> prefix(W + 1, 0);
    for (size_t i = 0; i < W; ++i) {
      int any = 0;
      for (size_t j = 0; j < H && !any; ++j) any = tf2d_[4 * (j * W + i) + 3] > 0.f;
      prefix[i + 1] = prefix[i] + any;
    }
    if (tf2d_dev_.bytes < W * H * 16) TVK_HIP(tf2d_dev_.alloc(W * H * 16));
    if (tf2d_prefix_dev_.bytes < (W + 1) * 4) TVK_HIP(tf2d_prefix_dev_.alloc((W + 1) * 4));
    TVK_HIP(hipMemcpyAsync(tf2d_dev_.p, tf2d_.data(), W * H * 16, hipMemcpyHostToDevice, stream_));
    TVK_HIP(hipMemcpyAsync(tf2d_prefix_dev_.p, prefix.data(), (W + 1) * 4, hipMemcpyHostToDevice,
                           stream_));
    TVK_HIP(hipStreamSynchronize(stream_));
    tf2d_dirty_ = false;
  }
  if (!tf_dirty_) return TVK_OK;
  const size_t n = tf1d_.size() / 4;
  std::vector<int32_t> prefix(n + 1, 0);
  for (size_t i = 0; i < n; ++i) prefix[i + 1] = prefix[i] + (tf1d_[4 * i + 3] > 0.f ? 1 : 0);
  if (tf_dev_.bytes < n * 16) TVK_HIP(tf_dev_.alloc(n * 16));
  if (tf_prefix_dev_.bytes < (n + 1) * 4) TVK_HIP(tf_prefix_dev_.alloc((n + 1) * 4));
  TVK_HIP(hipMemcpyAsync(tf_dev_.p, tf1d_.data(), n * 16, hipMemcpyHostToDevice, stream_));
  TVK_HIP(hipMemcpyAsync(tf_prefix_dev_.p, prefix.data(), (n + 1) * 4, hipMemcpyHostToDevice, stream_));
  // the host vector may change before the copy runs: wait for it
  TVK_HIP(hipStreamSynchronize(stream_));
  tf_dirty_ = false;
  return TVK_OK;
}

// ----------------------------------------------------------------- camera
int HipRenderer::SetViewPos(const float v[3]) {
  if (!finite3(v)) return arg_fail("view position must be finite");
  for (int i = 0; i < 3; ++i) view_pos_[i] = v[i];
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetViewDir(const float v[3]) {
  if (!finite3(v) || (v[0] == 0 && v[1] == 0 && v[2] == 0)) return arg_fail("view direction must be non-zero");
  for (int i = 0; i < 3; ++i) view_dir_[i] = v[i];
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetUpDir(const float v[3]) {
  if (!finite3(v) || (v[0] == 0 && v[1] == 0 && v[2] == 0)) return arg_fail("up direction must be non-zero");
  for (int i = 0; i < 3; ++i) up_dir_[i] = v[i];
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetFoV(float deg) {
  if (!(deg > 0.f && deg < 180.f)) return arg_fail("field of view must be in (0, 180)");
  fov_ = deg;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetOrtho(bool on, float hh) {
  if (on && !(hh > 0.f && isfinite(hh))) return arg_fail("orthographic half height must be > 0");
  ortho_ = on;
  if (on) ortho_hh_ = hh;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetRotation(const float m[16]) {
  double d[16], inv[16];
  for (int i = 0; i < 16; ++i) d[i] = m[i];
  if (!mat_inverse(d, inv)) return arg_fail("rotation matrix is singular");
  memcpy(rot_, d, sizeof d);
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetTranslation(const float m[16]) {
  double d[16], inv[16];
  for (int i = 0; i < 16; ++i) d[i] = m[i];
  if (!mat_inverse(d, inv)) return arg_fail("translation matrix is singular");
  memcpy(trans_, d, sizeof d);
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetRescaleFactors(const double s[3]) {
  for (int i = 0; i < 3; ++i)
    if (!(s[i] > 0.0) || !isfinite(s[i])) return arg_fail("rescale factors must be > 0");
  for (int i = 0; i < 3; ++i) rescale_[i] = s[i];
  invalidate();
  return TVK_OK;
}

// ------------------------------------------------------- quality and LOD
int HipRenderer::SetSampleRateModifier(float r) {
  if (!(r > 0.f && r <= 64.f)) return arg_fail("sample rate modifier must be in (0, 64]");
  rate_ = r;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetInterpolant(int i) {
  if (i != TVK_NEAREST && i != TVK_LINEAR) return arg_fail("unknown interpolant");
  interp_ = i;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetLODLimits(uint32_t mn, uint32_t mx) {
  if (mn > mx) return arg_fail("LOD limits: min must be <= max");
  lod_min_ = mn;
  lod_max_ = mx;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetBGColors(const float top[3], const float bottom[3]) {
  if (!finite3(top) || !finite3(bottom)) return arg_fail("background colours must be finite");
  for (int i = 0; i < 3; ++i) {
    bg_top_[i] = top[i];
    bg_bottom_[i] = bottom[i];
  }
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetLightingEnabled(bool on) {
  lighting_ = on;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetLightColors(const float amb[4], const float dif[4], const float spec[4],
                                const float dir[3]) {
  for (int i = 0; i < 4; ++i) {
    light_[i] = amb[i];
    light_[4 + i] = dif[i];
    light_[8 + i] = spec[i];
  }
  for (int i = 0; i < 3; ++i) light_[12 + i] = dir[i];
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetEarlyRayTermination(float a) {
  if (!(a > 0.f)) return arg_fail("termination threshold must be > 0");
  term_alpha_ = a;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetEmptySpaceSkipping(bool on) {
  skip_ = on;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetKernelTuning(uint32_t batch, uint32_t tile_group, uint32_t xcd_split) {
  if (batch != 1 && batch != 2 && batch != 4 && batch != 8) return arg_fail("batch must be 1, 2, 4 or 8");
  if (tile_group < 1 || tile_group > 4096) return arg_fail("tile group must be in [1, 4096]");
  if (xcd_split != 0 && xcd_split != 1 && xcd_split != 2 && xcd_split != 4 && xcd_split != 8)
    return arg_fail("xcd split must be 0, 1, 2, 4 or 8");
  batch_ = (int)batch;
  tile_group_ = (int)tile_group;
  xcd_split_ = (int)xcd_split;
  return TVK_OK;
}

// ----------------------------------------------------------------- render
int HipRenderer::check_ready() const {
  if (!initialized_) return state_fail("paint before initialize");
  if (!has_volume_) return state_fail("paint before a dataset is loaded");
  if (width_ == 0) return state_fail("paint before resize");
  if (render_mode_ == TVK_RM_1DTRANS && tf1d_.empty()) return state_fail("paint before a 1D TF is set");
  if (render_mode_ == TVK_RM_2DTRANS && tf2d_.empty()) return state_fail("paint before a 2D TF is set");
  return TVK_OK;
}

double HipRenderer::value_max() const {
  // the largest value of the finest level (the ranges of its bricks)
  double m = 0.0;
  const LodInfo& L = lods_[0];
  const int nb = L.grid[0] * L.grid[1] * L.grid[2];
  for (int b = 0; b < nb; ++b) m = std::max(m, (double)pool_.brick_range_h[2 * (L.brick_base + b) + 1]);
  return m > 0.0 ? m : 1.0;
}

void HipRenderer::build_params(RaycastParams* p, int lod) const {
  memset(p, 0, sizeof(*p));
  double M[16], Minv[16];
  mat_mul(rot_, trans_, M);
  if (!mat_inverse(M, Minv)) mat_identity(Minv);  // setters reject singular input
  double m = 0;
  for (int i = 0; i < 3; ++i) m = std::max(m, (double)dims0_[i] * rescale_[i]);
  double vscale[3], voff[3];
  for (int i = 0; i < 3; ++i) {
    vscale[i] = m / rescale_[i];
    voff[i] = 0.5 * (double)dims0_[i];
  }
  auto vox_point = [&](const double w[3], float out[3]) {
    double pm[3];
    xform(Minv, w, 1.0, pm);
    for (int i = 0; i < 3; ++i) out[i] = (float)(pm[i] * vscale[i] + voff[i]);
  };
  auto vox_dir = [&](const double w[3], double out[3]) {
    double dm[3];
    xform(Minv, w, 0.0, dm);
    for (int i = 0; i < 3; ++i) out[i] = dm[i] * vscale[i];
  };
  double f[3] = {view_dir_[0], view_dir_[1], view_dir_[2]}, r[3], u[3];
  normalize(f);
  cross(f, up_dir_, r);
  normalize(r);
  cross(r, f, u);
  const double W = width_, H = height_, asp = W / H;
  double tmp[3], dd[3];
  if (ortho_) {
    const double hh = ortho_hh_;
    double o00[3];
    for (int i = 0; i < 3; ++i) o00[i] = view_pos_[i] + (1.0 / W - 1.0) * hh * asp * r[i] + (1.0 - 1.0 / H) * hh * u[i];
    vox_point(o00, p->eye);
    for (int i = 0; i < 3; ++i) tmp[i] = 2.0 * hh * asp / W * r[i];
    vox_dir(tmp, dd);
    for (int i = 0; i < 3; ++i) p->du[i] = (float)dd[i];
    for (int i = 0; i < 3; ++i) tmp[i] = -2.0 * hh / H * u[i];
    vox_dir(tmp, dd);
    for (int i = 0; i < 3; ++i) p->dv[i] = (float)dd[i];
    vox_dir(f, dd);
    normalize(dd);
    for (int i = 0; i < 3; ++i) p->odir[i] = (float)dd[i];
    p->ortho = 1;
  } else {
    const double th = tan(fov_ * 0.5 * M_PI / 180.0);
    vox_point(view_pos_, p->eye);
    for (int i = 0; i < 3; ++i) tmp[i] = f[i] + (1.0 / W - 1.0) * th * asp * r[i] + (1.0 - 1.0 / H) * th * u[i];
    vox_dir(tmp, dd);
    for (int i = 0; i < 3; ++i) p->dir00[i] = (float)dd[i];
    for (int i = 0; i < 3; ++i) tmp[i] = 2.0 * th * asp / W * r[i];
    vox_dir(tmp, dd);
    for (int i = 0; i < 3; ++i) p->du[i] = (float)dd[i];
    for (int i = 0; i < 3; ++i) tmp[i] = -2.0 * th / H * u[i];
    vox_dir(tmp, dd);
    for (int i = 0; i < 3; ++i) p->dv[i] = (float)dd[i];
    p->ortho = 0;
  }
  p->width = (int32_t)width_;
  p->height = (int32_t)height_;
  p->tiles_x = (p->width + kTile - 1) / kTile;
  p->tiles_y = (p->height + kTile - 1) / kTile;
  for (int i = 0; i < 3; ++i) p->box[i] = (float)dims0_[i];
  p->step = (float)(ldexp(1.0, lod) / (double)rate_);
  p->nominal_rate = rate_;
  p->term_alpha = term_alpha_;
  p->interp = interp_;
  p->batch = batch_;
  p->tile_group = tile_group_;
  p->xcd_split = xcd_split_;
  p->compute_units = compute_units_;
  p->skip = skip_ ? 1 : 0;
  p->level = lods_[lod];
  p->brick = brick_;
  p->apron = apron_;
  p->slot_edge = slot_edge_;
  p->slot_elems = (int64_t)(pool_.slot_bytes / (bits_ / 8));
  p->n_tf = (int32_t)(tf1d_.size() / 4);
  p->tf_scale = bias_scale_tf_ ? (float)((double)(p->n_tf - 1) / value_max()) : 1.0f;
  for (int i = 0; i < 3; ++i) {
    p->bg_top[i] = bg_top_[i];
    p->bg_bottom[i] = bg_bottom_[i];
  }
  p->bg_alpha = bg_alpha_;
  p->pool = pool_.slots.p;
  p->page_table = pool_.page_table.as<int32_t>();
  p->brick_range = pool_.brick_range.as<float>();
  p->cell_range = cell_range_.as<uint32_t>();
  p->cell = skip_ ? cell_ : brick_;
  p->tf = tf_dev_.as<float>();
  p->tf_prefix = tf_prefix_dev_.as<int32_t>();
  p->framebuffer = framebuffer_.as<uint32_t>();
  p->counters = counters_.as<unsigned long long>();

  // render mode, gradient units, lighting
  p->mode = render_mode_ == TVK_RM_2DTRANS ? 1 : (render_mode_ == TVK_RM_ISOSURFACE ? 2 : 0);
  p->lighting = (lighting_ && p->mode == 0) ? 1 : 0;
  const double maxval = (double)((1u << bits_) - 1);
  for (int i = 0; i < 3; ++i) p->grad_scale[i] = (float)((double)lods_[lod].to_level[i] / (2.0 * maxval));
  for (int i = 0; i < 3; ++i) {
    p->ka[i] = light_[i];
    p->kd[i] = light_[4 + i];
    p->ks[i] = light_[8 + i];
  }
  p->shininess = light_[11];
  double ldir[3] = {light_[12], light_[13], light_[14]}, lv[3];
  vox_dir(ldir, lv);
  normalize(lv);
  for (int i = 0; i < 3; ++i) p->light[i] = (float)lv[i];
  if (p->mode == 1) {
    p->lighting = lighting_ ? 1 : 0;
    p->tf2d = tf2d_dev_.as<float>();
    p->tf2d_w = (int32_t)tf2d_w_;
    p->tf2d_h = (int32_t)tf2d_h_;
    p->tf2d_row_scale = (float)((double)(tf2d_h_ - 1) / (double)tf2d_max_grad_);
    p->n_tf = (int32_t)tf2d_w_;  // the skip test runs over the value columns
    p->tf_scale = bias_scale_tf_ ? (float)((double)(tf2d_w_ - 1) / value_max()) : 1.0f;
    p->tf_prefix = tf2d_prefix_dev_.as<int32_t>();
  }
  p->iso_value = iso_value_;
  for (int i = 0; i < 3; ++i) p->iso_color[i] = iso_color_[i];
}

int HipRenderer::Paint() {
  if (int e = check_ready()) return e;
  TVK_HIP(hipSetDevice(device_));
  if (int e = upload_tf()) return e;
  const int lo = std::min<int>((int)lod_min_, n_lod_ - 1);
  const int hi = std::max(lo, std::min<int>((int)lod_max_, n_lod_ - 1));
  int lod;
  if (target_ != TVK_RT_INTERACTIVE) lod = lo;
  else if (frame_complete_) lod = lo;
  else if (next_lod_ < 0) lod = hi;
  else lod = std::max(lo, next_lod_);

  RaycastParams p;
  build_params(&p, lod);
  TVK_HIP(hipEventRecord(ev_[0], stream_));
  TVK_HIP(launch_fill_u64(counters_.as<unsigned long long>(), 10, 0ull, stream_));
  TVK_HIP(hipEventRecord(ev_[1], stream_));
  TVK_HIP(launch_raycast(p, (int)bits_ / 8, stream_));
  TVK_HIP(hipEventRecord(ev_[2], stream_));
  TVK_HIP(hipMemcpyAsync(counters_h_, counters_.p, 2 * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipEventRecord(ev_[3], stream_));

  // frame bookkeeping and the host-side part of the stats
  const LodInfo& L = lods_[lod];
  const int nb = L.grid[0] * L.grid[1] * L.grid[2];
  uint32_t skipped = 0;
  if (skip_) {
    // bricks whose whole value range is transparent (below the iso value)
    const int ntf = p.n_tf;
    std::vector<int32_t> prefix(ntf + 1, 0);
    for (int i = 0; i < ntf; ++i) {
      int any = 0;
      if (p.mode == 0) any = tf1d_[4 * i + 3] > 0.f;
      else if (p.mode == 1)
        for (uint32_t j = 0; j < tf2d_h_ && !any; ++j) any = tf2d_[4 * ((size_t)j * tf2d_w_ + i) + 3] > 0.f;
      prefix[i + 1] = prefix[i] + any;
    }
    for (int b = 0; b < nb; ++b) {
      const float mn = pool_.brick_range_h[2 * (L.brick_base + b)];
      const float mx = pool_.brick_range_h[2 * (L.brick_base + b) + 1];
      if (p.mode == 2) {
        if (mx < p.iso_value) ++skipped;
        continue;
      }
      const int i0 = (int)std::min(std::max(mn * p.tf_scale, 0.f), (float)(ntf - 1));
      const int i1 = std::min((int)std::min(std::max(mx * p.tf_scale, 0.f), (float)(ntf - 1)) + 1, ntf - 1);
      if (prefix[i1 + 1] - prefix[i0] == 0) ++skipped;
    }
  }
  stats_ = tvk_frame_stats{};
  stats_.bricks_total = (uint32_t)nb;
  stats_.bricks_skipped = skipped;
  stats_.lod = (uint32_t)lod;
  stats_.width = width_;
  stats_.height = height_;
  stats_pending_ = true;
  last_lod_ = lod;
  ++subframes_;
  if (lod <= lo) {
    frame_complete_ = true;
    next_lod_ = -1;
  } else {
    frame_complete_ = false;
    next_lod_ = lod - 1;
  }
  return TVK_OK;
}

int HipRenderer::ScheduleCompleteRedraw() {
  invalidate();
  return TVK_OK;
}

int HipRenderer::Capture(uint8_t* rgba, uint64_t bytes, bool preserve) {
  if (int e = check_ready()) return e;
  if (!rgba || bytes < (uint64_t)width_ * height_ * 4) return arg_fail("capture buffer too small");
  float top[3], bot[3];
  const float bga = bg_alpha_;
  if (preserve) {
    memcpy(top, bg_top_, sizeof top);
    memcpy(bot, bg_bottom_, sizeof bot);
    for (int i = 0; i < 3; ++i) bg_top_[i] = bg_bottom_[i] = 0.f;
    bg_alpha_ = 0.f;
    invalidate();
  }
  int rc = TVK_OK;
  for (int guard = 0; guard < kMaxLod + 1 && !frame_complete_ && rc == TVK_OK; ++guard) rc = Paint();
  if (rc == TVK_OK) {
    hipError_t e = hipMemcpyAsync(rgba, framebuffer_.p, (size_t)width_ * height_ * 4,
                                  hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) rc = hip_fail(e, "capture readback");
  }
  if (preserve) {
    memcpy(bg_top_, top, sizeof top);
    memcpy(bg_bottom_, bot, sizeof bot);
    bg_alpha_ = bga;
    invalidate();
  }
  return rc;
}

int HipRenderer::Progress(uint64_t* subframes, uint32_t* working_subframe, uint32_t* bricks,
                          uint32_t* working_brick, uint64_t* min_lod) const {
  const int lo = has_volume_ ? std::min<int>((int)lod_min_, n_lod_ - 1) : 0;
  const int hi = has_volume_ ? std::max(lo, std::min<int>((int)lod_max_, n_lod_ - 1)) : 0;
  if (subframes) *subframes = (uint64_t)(hi - lo + 1);
  if (working_subframe) *working_subframe = last_lod_ < 0 ? 0u : (uint32_t)(hi - last_lod_);
  if (bricks) *bricks = stats_.bricks_total;
  if (working_brick) *working_brick = stats_.bricks_total;  // one launch covers every brick
  if (min_lod) *min_lod = (uint64_t)lo;
  return TVK_OK;
}

int HipRenderer::FrameStats(tvk_frame_stats* out) {
  if (!out) return arg_fail("null stats pointer");
  if (stats_pending_) {
    TVK_HIP(hipSetDevice(device_));
    TVK_HIP(hipEventSynchronize(ev_[3]));
    float k = 0.f, all = 0.f;
    TVK_HIP(hipEventElapsedTime(&k, ev_[1], ev_[2]));
    TVK_HIP(hipEventElapsedTime(&all, ev_[0], ev_[3]));
    stats_.kernel_ms = k;
    stats_.paint_ms = all;
    stats_.evaluated_samples = counters_h_[0];
    stats_.nominal_samples = counters_h_[1];
    stats_pending_ = false;
  }
  *out = stats_;
  return TVK_OK;
}

int HipRenderer::DeviceFramebuffer(void** dptr, void** stream) {
  if (!initialized_ || !framebuffer_.p) return state_fail("no framebuffer (initialize + resize first)");
  if (dptr) *dptr = framebuffer_.p;
  if (stream) *stream = (void*)stream_;
  return TVK_OK;
}

int HipRenderer::Synchronize() {
  if (!initialized_) return state_fail("not initialized");
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

// ----------------------------------------------------------------- memory
int HipRenderer::SetMaxUsableGPUMem(uint64_t bytes) {
  if (bytes == 0) return arg_fail("GPU memory budget must be > 0");
  gpu_budget_ = bytes;
  return TVK_OK;
}

int HipRenderer::MemStats(tvk_mem_stats* out) {
  if (!out) return arg_fail("null stats pointer");
  *out = tvk_mem_stats{};
  out->gpu_mem = gpu_total_;
  out->gpu_budget = gpu_budget_;
  uint64_t a = framebuffer_.bytes + counters_.bytes + tf_dev_.bytes + tf_prefix_dev_.bytes +
               pool_.slots.bytes + pool_.page_table.bytes + pool_.brick_range.bytes + scratch_.bytes +
               cell_range_.bytes;
  for (auto& l : levels_) a += l.bytes;
  out->allocated_gpu_mem = a;
  out->pool_slots = pool_.n_slots;
  out->resident_bricks = pool_.resident;
  out->slot_bytes = pool_.slot_bytes;
  out->bricks_uploaded = pool_.uploaded;
  out->bricks_evicted = pool_.evicted;
  return TVK_OK;
}

}  // namespace tvk
