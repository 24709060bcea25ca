// renderer.cpp -- HipRenderer (see renderer.hpp).
#include "renderer.hpp"
#include "memman.hpp"

#include <math.h>
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <cmath>

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
  int dev = 0;
  if (hipError_t e = hipGetDevice(&dev)) return e;
  MemMan& mm = MemMan::ForDevice(dev);
  if (!mm.ReserveGPU(n)) return hipErrorOutOfMemory;  // beyond the device's budget
  hipError_t e = hipMalloc(&p, n);
  if (e == hipSuccess) {
    bytes = n;
    device = dev;
  } else {
    p = nullptr;
    mm.ReleaseGPU(n);
  }
  return e;
}

void DevBuf::reset() {
  if (p) {
    (void)hipFree(p);
    MemMan::ForDevice(device).ReleaseGPU(bytes);
  }
  p = nullptr;
  bytes = 0;
}

int LaunchSlot::alloc() {
  TVK_HIP(table.alloc(sizeof(JobTable)));
  TVK_HIP(counters.alloc(kCounterWords * sizeof(unsigned long long)));
  TVK_HIP(hipHostMalloc((void**)&table_h, sizeof(JobTable), hipHostMallocDefault));
  TVK_HIP(hipHostMalloc((void**)&counters_h, kCounterWords * sizeof(unsigned long long),
                        hipHostMallocDefault));
  memset(table_h, 0, sizeof(JobTable));
  memset(counters_h, 0, kCounterWords * sizeof(unsigned long long));
  for (auto& e : ev) TVK_HIP(hipEventCreate(&e));
  return TVK_OK;
}

void LaunchSlot::release() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  for (auto& e : ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  if (table_h) (void)hipHostFree(table_h);
  if (counters_h) (void)hipHostFree(counters_h);
  table_h = nullptr;
  counters_h = nullptr;
  table.reset();
  counters.reset();
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

double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// The job's tile rect: the tiles whose rays can reach the volume box
// [0, box].  Pixel (x, y)'s ray is eye + s (dir00 + x du + y dv), s >= 0
// (perspective) or eye + x du + y dv + s odir (orthographic), so with the
// basis M = [du dv c] (c = dir00 or odir) a point q has M^-1 (q - eye) =
// (s x, s y, s) resp. (x, y, s).  When every box corner lies in front of
// the eye plane (s > 0; always taken for orthographic), the pixels whose
// rays meet the box are the convex hull of the projected corners, so
// their bounding rectangle grown by 2 pixels (far beyond float rounding of
// the rays) holds every pixel that can hit; tiles outside it get only the
// background.  A corner on or behind the eye plane (eye inside or beside
// the box) or a degenerate basis keeps the full frame.
void set_tile_rect(RaycastParams* p) {
  p->rect_x0 = p->rect_y0 = 0;
  p->rect_w = p->tiles_x;
  p->rect_h = p->tiles_y;
  p->tile_sample = 1;
  p->tile_phase = 0;
  double c0[3], c1[3], c2[3], eye[3];
  for (int i = 0; i < 3; ++i) {
    c0[i] = p->du[i];
    c1[i] = p->dv[i];
    c2[i] = p->ortho ? p->odir[i] : p->dir00[i];
    eye[i] = p->eye[i];
  }
  double r0[3], r1[3], r2[3];  // rows of det * M^-1
  cross(c1, c2, r0);
  cross(c2, c0, r1);
  cross(c0, c1, r2);
  const double det = dot3(c0, r0);
  const double scale = sqrt(dot3(c0, c0) * dot3(c1, c1) * dot3(c2, c2));
  if (!(fabs(det) > 1e-9 * scale)) return;
  double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
  for (int k = 0; k < 8; ++k) {
    double q[3];
    for (int i = 0; i < 3; ++i) q[i] = ((k >> i) & 1 ? (double)p->box[i] : 0.0) - eye[i];
    double x = dot3(r0, q) / det, y = dot3(r1, q) / det;
    if (!p->ortho) {
      const double s = dot3(r2, q) / det;
      if (!(s > 0.0)) return;
      x /= s;
      y /= s;
    }
    if (!std::isfinite(x) || !std::isfinite(y)) return;
    xmin = std::min(xmin, x);
    xmax = std::max(xmax, x);
    ymin = std::min(ymin, y);
    ymax = std::max(ymax, y);
  }
  const double W = p->width, H = p->height;
  const double x0 = std::max(floor(xmin) - 2.0, 0.0), x1 = std::min(ceil(xmax) + 2.0, W - 1.0);
  const double y0 = std::max(floor(ymin) - 2.0, 0.0), y1 = std::min(ceil(ymax) + 2.0, H - 1.0);
  const int tw = 1 << p->tile_shift, th = 64 >> p->tile_shift;
  if (x0 > x1 || y0 > y1) {  // the box projects outside the image
    p->rect_w = p->rect_h = 0;
    return;
  }
  p->rect_x0 = (int)x0 / tw;
  p->rect_y0 = (int)y0 / th;
  p->rect_w = (int)x1 / tw + 1 - p->rect_x0;
  p->rect_h = (int)y1 / th + 1 - p->rect_y0;
}

}  // namespace

// -------------------------------------------------------------- lifecycle
HipRenderer::HipRenderer(int type, bool pow2, bool to8, bool no_border, bool bias_scale_tf)
    : type_(type), pow2_(pow2), to8_(to8), no_border_(no_border), bias_scale_tf_(bias_scale_tf) {
  mat_identity(rot_);
  mat_identity(trans_);
  for (double& t : lod_ms_) t = -1.0;
}

HipRenderer::~HipRenderer() {
  if (initialized_) {
    (void)hipSetDevice(device_);
    (void)hipStreamSynchronize(stream_);
    for (auto& sl : paint_slots_) sl.release();
    batch_slot_.release();
    check_slot_.release();
    if (check_sse_h_) (void)hipHostFree(check_sse_h_);
    for (hipEvent_t ev : check_ev_)
      if (ev) (void)hipEventDestroy(ev);
    check_copy_.reset();
    check_sse_.reset();
    batch_fb_.reset();
    release_volume();  // gives a paged volume's slots back to the device pool
    if (requests_h_) (void)hipHostFree(requests_h_);
    for (hipEvent_t ev : staging_ev_)
      if (ev) (void)hipEventDestroy(ev);
    if (upload_done_) (void)hipEventDestroy(upload_done_);
    requests_dev_.reset();
    touch_dev_.reset();
    tfs_.clear();
    for (auto& l : levels_) l.reset();
    pool_.slots.reset();
    pool_.page_table.reset();
    pool_.brick_range.reset();
    scratch_.reset();
    cell_range_.reset();
    framebuffer_.reset();
    lowres_fb_.reset();
    if (copy_stream_) (void)hipStreamDestroy(copy_stream_);
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
  TVK_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
  TVK_HIP(noalloc_scratch_init(device_));  // this device's allocation-free launch scratch
  for (auto& sl : paint_slots_)
    if (int e = sl.alloc()) return e;
  if (int e = batch_slot_.alloc()) return e;
  TVK_HIP(hipDeviceGetAttribute(&compute_units_, hipDeviceAttributeMultiprocessorCount, device_));
  size_t free_b = 0, total_b = 0;
  TVK_HIP(hipMemGetInfo(&free_b, &total_b));
  gpu_total_ = total_b;
  initialized_ = true;
  return TVK_OK;
}

int HipRenderer::Resize(uint32_t w, uint32_t h) {
  if (!initialized_) return state_fail("resize before initialize");
  if (w == 0 || h == 0 || w > 16384 || h > 16384) return arg_fail("window size out of range");
  if (w == width_ && h == height_) return TVK_OK;
  if (int e = resolve_check()) return e;
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipStreamSynchronize(stream_));
  TVK_HIP(framebuffer_.alloc((size_t)w * h * 4));
  for (double& t : lod_ms_) t = -1.0;  // timings were for the old window size
  time_pending_ = false;
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

int HipRenderer::plan_levels(const uint64_t dims[3], uint32_t bits, uint32_t brick, uint32_t apron,
                             uint32_t n_lod, LevelPlan* out) const {
  if (bits != 8 && bits != 16) return arg_fail("bits must be 8 or 16");
  for (int i = 0; i < 3; ++i)
    if (dims[i] == 0 || dims[i] > (1u << 20)) return arg_fail("volume dimension out of range");
  if (brick < 8 || brick > 1024) return arg_fail("brick size must be in [8, 1024]");
  if (apron < 1 || apron > 8) return arg_fail("apron must be in [1, 8]");
  const uint64_t S = (uint64_t)brick + 2ull * apron;
  if ((uint64_t)slot_block_elems((int)S) >= (1ull << 31)) return arg_fail("brick slot too large");
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
  for (int i = 0; i < 3; ++i) out->dims[0][i] = dims[i];
  for (int l = 1; l < nl; ++l)
    for (int i = 0; i < 3; ++i) out->dims[l][i] = (out->dims[l - 1][i] + 1) / 2;
  // empty-space cell edge: the chosen edge (default 16) when it divides the
  // brick, else 16 when that does, else the brick
  int cell = (int)brick;
  for (const int c : {cell_pref_, 16})
    if (c >= 4 && c <= (int)brick && brick % c == 0) {
      cell = c;
      break;
    }
  int total = 0, total_cells = 0;
  uint64_t level_bytes = 0;
  for (int l = 0; l < nl; ++l) {
    LodInfo& L = out->lods[l];
    for (int i = 0; i < 3; ++i) {
      L.dims[i] = (int32_t)out->dims[l][i];
      L.grid[i] = (int32_t)((out->dims[l][i] + brick - 1) / brick);
      L.to_level[i] = (float)((double)out->dims[l][i] / (double)dims[i]);
      L.cells[i] = (int32_t)((out->dims[l][i] + cell - 1) / cell);
    }
    L.brick_base = total;
    L.cell_base = total_cells;
    total += L.grid[0] * L.grid[1] * L.grid[2];
    total_cells += L.cells[0] * L.cells[1] * L.cells[2];
    level_bytes += out->dims[l][0] * out->dims[l][1] * out->dims[l][2] * bpv;
  }
  out->n_lod = nl;
  out->total_bricks = total;
  out->total_cells = total_cells;
  out->cell = cell;
  out->brick = (int)brick;
  out->level_bytes = level_bytes;
  out->slot_edge = (int)S;
  out->slot_bytes = ((uint64_t)slot_block_elems((int)S) * bpv + 255) & ~255ull;
  return TVK_OK;
}

void HipRenderer::release_volume() {
  (void)resolve_check();  // its launches read this volume
  has_volume_ = false;
  paged_ = false;
  for (double& t : lod_ms_) t = -1.0;  // the governor's timings were for the old volume
  time_pending_ = false;
  for (auto& l : levels_) l.reset();
  if (pool_client_ >= 0) {
    // the copy stream may still fill this volume's slots
    (void)hipStreamSynchronize(copy_stream_);
    MemMan::ForDevice(device_).Unregister(pool_client_);
    pool_client_ = -1;
  }
  MemMan& mm = MemMan::ForDevice(device_);
  for (void* b : staging_) (void)hipHostFree(b);
  staging_.clear();
  mm.ReleaseCPU(staging_host_bytes_ + host_levels_bytes_);
  staging_host_bytes_ = host_levels_bytes_ = 0;
  upload_stage_.clear();
  pt_dirty_ = false;
  pool_.slots.reset();
  pool_.page_table.reset();
  pool_.brick_range.reset();
  pool_.n_slots = pool_.slot_bytes = pool_.resident = 0;
  pool_.page_table_h.clear();
  pool_.brick_range_h.clear();
  pool_.brick_slot.clear();
  cell_range_.reset();
  requests_dev_.reset();
  touch_dev_.reset();
  host_levels_.clear();
  host_level0_ = nullptr;
  source_fn_ = nullptr;
  source_user_ = nullptr;
  total_cells_ = 0;
  ++cells_gen_;
}

int HipRenderer::SetVolume(const void* voxels, const uint64_t dims[3], uint32_t bits,
                           uint32_t brick, uint32_t apron, uint32_t n_lod) {
  if (!voxels) return arg_fail("null voxel pointer");
  return set_volume_impl(voxels, -1, 0, dims, bits, brick, apron, n_lod);
}

int HipRenderer::SetGeneratedVolume(int kind, uint64_t seed, const uint64_t dims[3], uint32_t bits,
                                    uint32_t brick, uint32_t apron, uint32_t n_lod) {
  if (!is_gpu_volume_kind(kind)) return arg_fail("not a device-generated volume kind");
  return set_volume_impl(nullptr, kind, seed, dims, bits, brick, apron, n_lod);
}

int HipRenderer::set_volume_impl(const void* voxels, int gen_kind, uint64_t seed,
                                 const uint64_t dims[3], uint32_t bits, uint32_t brick,
                                 uint32_t apron, uint32_t n_lod) {
  if (!initialized_) return state_fail("setVolume before initialize");
  LevelPlan plan;
  if (int e = plan_levels(dims, bits, brick, apron, n_lod, &plan)) return e;
  const int bpv = (int)bits / 8;
  const int nl = plan.n_lod, total = plan.total_bricks, total_cells = plan.total_cells;
  const uint64_t level_bytes = plan.level_bytes;
  const uint64_t (*ld)[3] = plan.dims;
  const uint64_t slot_bytes = plan.slot_bytes;
  const uint64_t need = level_bytes + (uint64_t)total * slot_bytes + (uint64_t)total * 12 +
                        (uint64_t)total_cells * 4 +
                        (uint64_t)total * sizeof(BrickJob) + (size_t)width_ * height_ * 4;
  const uint64_t budget = MemMan::ForDevice(device_).GPUBudget();
  if (need > budget) {
    set_last_error("volume needs " + std::to_string(need) + " bytes of device memory, budget is " +
                   std::to_string(budget));
    return TVK_ERR_OOM;
  }

  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipStreamSynchronize(stream_));
  release_volume();

  for (int l = 0; l < nl; ++l) TVK_HIP(levels_[l].alloc(ld[l][0] * ld[l][1] * ld[l][2] * bpv));
  if (voxels) {
    TVK_HIP(hipMemcpyAsync(levels_[0].p, voxels, levels_[0].bytes, hipMemcpyHostToDevice, stream_));
  } else {
    const int64_t d[3] = {(int64_t)dims[0], (int64_t)dims[1], (int64_t)dims[2]};
    const int64_t slab = std::max<int64_t>(1, (int64_t)((1ull << 30) / (dims[0] * dims[1])));
    for (int64_t z0 = 0; z0 < d[2]; z0 += slab) {
      uint8_t* dst = (uint8_t*)levels_[0].p + (uint64_t)z0 * dims[0] * dims[1] * bpv;
      TVK_HIP(launch_generate_volume(gen_kind, seed, d, (int)bits, z0, std::min(slab, d[2] - z0),
                                     dst, stream_));
    }
  }
  for (int l = 1; l < nl; ++l) {
    const int64_t sd[3] = {(int64_t)ld[l - 1][0], (int64_t)ld[l - 1][1], (int64_t)ld[l - 1][2]};
    TVK_HIP(launch_downsample(levels_[l - 1].p, bpv, sd, levels_[l].p, stream_));
  }

  bits_ = bits;
  for (int i = 0; i < 3; ++i) dims0_[i] = dims[i];
  apron_ = (int)apron;
  if (int e = brick_levels(plan)) return e;
  lod_min_ = 0;
  lod_max_ = (uint32_t)(nl - 1);
  has_volume_ = true;
  invalidate();
  return TVK_OK;
}

// Cut the flat levels on the device (levels_) into a fresh brick pool of
// the plan's brick size; per-brick and per-cell ranges; identity page table.
int HipRenderer::brick_levels(const LevelPlan& plan) {
  const int bpv = (int)bits_ / 8;
  const int nl = plan.n_lod, total = plan.total_bricks, total_cells = plan.total_cells;
  const int cell = plan.cell;
  const uint64_t slot_bytes = plan.slot_bytes;
  const int brick = plan.brick, apron = apron_;
  const LodInfo* lods = plan.lods;
  TVK_HIP(pool_.slots.alloc((size_t)total * slot_bytes));
  TVK_HIP(pool_.page_table.alloc((size_t)total * sizeof(uint64_t)));
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
                                  pool_.slots.p, stride, brick, apron,
                                  pool_.brick_range.as<float>(), scratch_.as<uint32_t>(), stream_));
  }
  TVK_HIP(cell_range_.alloc((size_t)total_cells * sizeof(uint32_t)));
  DevBuf cell_hi;  // the maxima while the kernel folds them in
  TVK_HIP(cell_hi.alloc((size_t)total_cells * sizeof(uint32_t)));
  for (int l = 0; l < nl; ++l) {
    const int64_t d[3] = {lods[l].dims[0], lods[l].dims[1], lods[l].dims[2]};
    TVK_HIP(launch_cell_ranges(levels_[l].p, bpv, d, cell, cell_range_.as<uint32_t>() + lods[l].cell_base,
                               cell_hi.as<uint32_t>() + lods[l].cell_base, stream_));
  }
  pool_.page_table_h.resize(total);
  for (int i = 0; i < total; ++i)
    pool_.page_table_h[i] = (uint64_t)(uintptr_t)((uint8_t*)pool_.slots.p + (uint64_t)i * slot_bytes);
  TVK_HIP(hipMemcpyAsync(pool_.page_table.p, pool_.page_table_h.data(), total * sizeof(uint64_t),
                         hipMemcpyHostToDevice, stream_));
  pool_.brick_range_h.resize((size_t)total * 2);
  TVK_HIP(hipMemcpyAsync(pool_.brick_range_h.data(), pool_.brick_range.p, total * 2 * sizeof(float),
                         hipMemcpyDeviceToHost, stream_));
  // djobs and the host vectors must outlive the copies
  TVK_HIP(hipStreamSynchronize(stream_));
  pool_.resident = (uint64_t)total;
  pool_.uploaded += (uint64_t)total;
  n_lod_ = nl;
  brick_ = brick;
  slot_edge_ = plan.slot_edge;
  cell_ = cell;
  total_cells_ = total_cells;
  ++cells_gen_;
  for (int l = 0; l < nl; ++l) lods_[l] = lods[l];
  total_bricks_ = total;
  region_lod_h_.clear();
  return upload_lods();
}

int HipRenderer::upload_lods() {
  if (!lods_dev_.p) TVK_HIP(lods_dev_.alloc(sizeof(LodInfo) * kMaxLod));
  TVK_HIP(hipMemcpyAsync(lods_dev_.p, lods_, sizeof(LodInfo) * kMaxLod, hipMemcpyHostToDevice, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

int HipRenderer::SetRegionLods(const uint8_t* levels, uint64_t n) {
  if (!has_volume_) return state_fail("setRegionLODs before a volume is set");
  if (n == 0 || !levels) {
    region_lod_h_.clear();
    ++region_version_;
    invalidate();
    return TVK_OK;
  }
  const LodInfo& L0 = lods_[0];
  const uint64_t regions = (uint64_t)L0.grid[0] * L0.grid[1] * L0.grid[2];
  if (n != regions) return arg_fail("region LOD table needs one entry per LOD-0 brick (" + std::to_string(regions) + ")");
  for (uint64_t i = 0; i < n; ++i)
    if (levels[i] >= (uint8_t)n_lod_) return arg_fail("region LOD beyond the coarsest level");
  region_lod_h_.assign(levels, levels + n);
  ++region_version_;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetRegionCheck(uint32_t tile_sample, float max_rmse) {
  if (tile_sample == 1 || tile_sample > 4096) return arg_fail("check tile sampling must be 0 (off) or 2..4096");
  if (!(max_rmse >= 0.f) || !std::isfinite(max_rmse)) return arg_fail("check RMSE bound must be finite and >= 0");
  if (int e = resolve_check()) return e;  // a frame's check runs with the settings it started with
  if (tile_sample > 0 && initialized_) {
    TVK_HIP(hipSetDevice(device_));
    if (!check_slot_.table.p)
      if (int e = check_slot_.alloc()) return e;
    if (!check_sse_.p) TVK_HIP(check_sse_.alloc(sizeof(unsigned long long)));
    if (!check_sse_h_) TVK_HIP(hipHostMalloc((void**)&check_sse_h_, sizeof(unsigned long long), hipHostMallocDefault));
    for (hipEvent_t& ev : check_ev_)
      if (!ev) TVK_HIP(hipEventCreate(&ev));
  } else if (tile_sample > 0) {
    return state_fail("region check before initialize");
  }
  check_k_ = tile_sample;
  check_max_ = max_rmse;
  check_phase_ = 0;
  check_stats_ = tvk_region_check{};
  check_timing_ = false;
  invalidate();  // the next capture renders (and checks) the frame again
  return TVK_OK;
}

int HipRenderer::RegionCheckStats(tvk_region_check* out) {
  if (!out) return arg_fail("null check stats");
  if (int e = resolve_check()) return e;
  if (check_timing_) {
    TVK_HIP(hipSetDevice(device_));
    TVK_HIP(hipEventSynchronize(check_ev_[1]));
    float ms = 0.f;
    TVK_HIP(hipEventElapsedTime(&ms, check_ev_[0], check_ev_[1]));
    check_stats_.check_ms = ms;
    check_timing_ = false;
  }
  *out = check_stats_;
  return TVK_OK;
}

// The per-frame check of a per-region frame just launched with params p:
// keep a copy of the frame, render the sampled level-0 tiles into it and
// sum the squared RGB differences over the frame (non-zero on the sampled
// tiles only).  All queued on the stream; resolve_check reads the sum.
int HipRenderer::check_launch(const RaycastParams& p) {
  const size_t px = (size_t)width_ * height_;
  if (check_copy_.bytes < px * 4) TVK_HIP(check_copy_.alloc(px * 4));
  RaycastParams q = p;
  q.region_code = tfs_[0].region_code0.as<uint8_t>();
  q.region_out = nullptr;
  TVK_HIP(hipEventRecord(check_ev_[0], stream_));
  TVK_HIP(hipMemcpyAsync(check_copy_.p, framebuffer_.p, px * 4, hipMemcpyDeviceToDevice, stream_));
  check_params_ = q;  // the fallback: the whole frame at level 0
  q.tile_sample = (int32_t)check_k_;
  q.tile_phase = (int32_t)(check_phase_++);
  if (int e = launch_jobs(check_slot_, &q, 1, launch_info(q.n_tf), false)) return e;
  TVK_HIP(launch_fill_u64(check_sse_.as<unsigned long long>(), 1, 0ull, stream_));
  TVK_HIP(launch_rgb_sse(framebuffer_.as<uint32_t>(), check_copy_.as<uint32_t>(), (int64_t)px,
                         check_sse_.as<unsigned long long>(), stream_));
  TVK_HIP(hipMemcpyAsync(check_sse_h_, check_sse_.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipEventRecord(check_ev_[2], stream_));
  check_open_ = true;
  return TVK_OK;
}

// Reads the check's error sum and, when the RMSE of the rest of the frame
// -- estimated as (k - 1) times the sampled tiles' error, each sampled tile
// standing for the k - 1 others of its run -- exceeds the bound, renders
// the whole frame at level 0.
int HipRenderer::resolve_check() {
  if (!check_open_) return TVK_OK;
  check_open_ = false;
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipEventSynchronize(check_ev_[2]));
  const double px = (double)check_params_.width * (double)check_params_.height;
  const double est = std::sqrt((double)(check_k_ - 1) * (double)*check_sse_h_ / (3.0 * px));
  check_stats_ = tvk_region_check{};
  check_stats_.rmse_estimate = (float)est;
  check_stats_.check_samples = check_slot_.counters_h[kQueues];
  if (est > (double)check_max_) {
    if (int e = launch_jobs(check_slot_, &check_params_, 1, launch_info(check_params_.n_tf), false)) return e;
    check_stats_.fell_back = 1;
  }
  TVK_HIP(hipEventRecord(check_ev_[1], stream_));
  check_timing_ = true;
  return TVK_OK;
}

int HipRenderer::SetRegionOutput(bool on) {
  region_out_on_ = on;
  return TVK_OK;
}

int HipRenderer::ReadRegionOutput(int32_t* out, uint64_t n) {
  if (!region_out_on_ || !region_out_dev_.p) return state_fail("region output is off (or nothing painted)");
  if (!out || n < (uint64_t)width_ * height_) return arg_fail("region output buffer too small");
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipMemcpyAsync(out, region_out_dev_.p, (size_t)width_ * height_ * 4, hipMemcpyDeviceToHost,
                         stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

// Region codes of TF slot `tf_slot`: the chosen level of every LOD-0 region,
// with 0x80 when the level's brick holding the region centre (the kernel's
// rule) is transparent under the 1D TF over its stored range.
int HipRenderer::prepare_regions(int tf_slot) {
  TfSet& t = tfs_[tf_slot];
  const int n = (int)t.prefix1d.size() - 1;
  const float scale = tf_scale(n);
  uint32_t scale_bits = 0;
  memcpy(&scale_bits, &scale, 4);
  uint64_t key = 1469598103934665603ull;
  for (const uint64_t v : {region_version_, t.version, cells_gen_, (uint64_t)scale_bits,
                           (uint64_t)skip_})
    key = (key ^ v) * 1099511628211ull;
  if (key == t.region_key && t.region_code.p) return TVK_OK;
  const LodInfo& L0 = lods_[0];
  const int g[3] = {L0.grid[0], L0.grid[1], L0.grid[2]};
  // the chosen levels' codes and, for the per-frame check, level 0's
  std::vector<uint8_t> code(region_lod_h_.size()), code0(region_lod_h_.size());
  const float B = (float)brick_;
  auto region_code = [&](int l, int cx, int cy, int cz) {
    const LodInfo& lv = lods_[l];
    const int c[3] = {cx, cy, cz};
    int b[3];
    for (int i = 0; i < 3; ++i) {
      const float qc = ((float)c[i] + 0.5f) * B * lv.to_level[i];
      b[i] = std::min(std::max((int)floorf(qc / B), 0), lv.grid[i] - 1);
    }
    const int gid = lv.brick_base + (b[2] * lv.grid[1] + b[1]) * lv.grid[0] + b[0];
    bool empty = false;
    if (skip_ && n > 0) {
      const float mn = pool_.brick_range_h[2 * gid], mx = pool_.brick_range_h[2 * gid + 1];
      const int i0 = (int)std::min(std::max(mn * scale, 0.f), (float)(n - 1));
      const int i1 = std::min((int)std::min(std::max(mx * scale, 0.f), (float)(n - 1)) + 1, n - 1);
      empty = t.prefix1d[i1 + 1] - t.prefix1d[i0] == 0;
    }
    return (uint8_t)(l | (empty ? 0x80 : 0));
  };
  for (int cz = 0; cz < g[2]; ++cz)
    for (int cy = 0; cy < g[1]; ++cy)
      for (int cx = 0; cx < g[0]; ++cx) {
        const size_t r = ((size_t)cz * g[1] + cy) * g[0] + cx;
        code[r] = region_code(region_lod_h_[r], cx, cy, cz);
        code0[r] = region_lod_h_[r] == 0 ? code[r] : region_code(0, cx, cy, cz);
      }
  if (t.region_code.bytes < code.size()) TVK_HIP(t.region_code.alloc(code.size()));
  if (t.region_code0.bytes < code0.size()) TVK_HIP(t.region_code0.alloc(code0.size()));
  TVK_HIP(hipMemcpyAsync(t.region_code.p, code.data(), code.size(), hipMemcpyHostToDevice, stream_));
  TVK_HIP(hipMemcpyAsync(t.region_code0.p, code0.data(), code0.size(), hipMemcpyHostToDevice, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  t.region_key = key;
  return TVK_OK;
}

int HipRenderer::Rebrick(uint32_t brick, uint32_t cell) {
  if (!has_volume_) return state_fail("rebrick before a volume is set");
  if (int e = resolve_check()) return e;  // its launches read the slots being replaced
  const int saved = cell_pref_;
  if (paged_) {
    // a paged volume over the caller's flat voxels is set up again from
    // them: new bricks, ranges and coarser levels, the same pool size (its
    // bricks are paged in afresh)
    if (!host_level0_) return state_fail("a paged volume fed by a brick source keeps its brick size");
    if (cell) cell_pref_ = (int)cell;
    const uint64_t dims[3] = {dims0_[0], dims0_[1], dims0_[2]};
    const int e = SetPagedFlatVolume(host_level0_, dims, bits_, brick, (uint32_t)apron_, (uint32_t)n_lod_,
                                     pool_bytes_req_);
    if (e) cell_pref_ = saved;
    return e;
  }
  if (cell) cell_pref_ = (int)cell;
  LevelPlan plan;
  const int e0 = plan_levels(dims0_, bits_, brick, (uint32_t)apron_, (uint32_t)n_lod_, &plan);
  if (e0) {
    cell_pref_ = saved;
    return e0;
  }
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipStreamSynchronize(stream_));
  pool_.slots.reset();
  pool_.page_table.reset();
  pool_.brick_range.reset();
  cell_range_.reset();
  if (int e = brick_levels(plan)) return e;
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
  if (paged_) {
    // the host keeps the levels of a flat paged volume
    const LodInfo& L = lods_[lod];
    const uint64_t n = (uint64_t)L.dims[0] * L.dims[1] * L.dims[2] * (bits_ / 8);
    if (!host_level0_) return state_fail("levels of a callback-sourced volume live with the source");
    if (!out || bytes < n) return arg_fail("output buffer too small");
    memcpy(out, lod == 0 ? host_level0_ : host_levels_[lod - 1].data(), n);
    return TVK_OK;
  }
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
  TVK_HIP(hipSetDevice(device_));
  DevBuf lin;
  TVK_HIP(lin.alloc(n));
  std::unique_lock<std::mutex> lk(pool_mu_);  // the slot stays this brick's while it is read
  const uint64_t addr = pool_.page_table_h[gid];
  if (!addr) {
    lk.unlock();
    if (!paged_) return state_fail("brick is not resident");
    // not resident: the brick source produces the same slot contents
    std::vector<uint8_t> tmp(pool_.slot_bytes);
    if (source_fn_(source_user_, lod, bx, by, bz, tmp.data(), pool_.slot_bytes) != 0) {
      set_last_error("brick source failed");
      return TVK_ERR_IO;
    }
    memcpy(out, tmp.data(), n);
    return TVK_OK;
  }
  // the slot's blocked layout back to an x-fastest brick, then to the host
  TVK_HIP(launch_slot_to_linear((const void*)(uintptr_t)addr, lin.p, (int)bits_ / 8, slot_edge_, stream_));
  TVK_HIP(hipMemcpyAsync(out, lin.p, n, hipMemcpyDeviceToHost, stream_));
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

int HipRenderer::BrickHistograms(uint32_t lod, uint32_t bins, uint32_t* out, uint64_t n) {
  if (!has_volume_) return state_fail("no volume loaded");
  if (paged_) return state_fail("brick histograms need an in-core volume");
  if (lod >= (uint32_t)n_lod_) return arg_fail("no such LOD level");
  if (bins < 2 || bins > 64 || (bins & (bins - 1))) return arg_fail("bins must be a power of two in [2, 64]");
  const LodInfo& L = lods_[lod];
  const uint64_t nb = (uint64_t)L.grid[0] * L.grid[1] * L.grid[2];
  if (!out || n < nb * bins) return arg_fail("output needs n_bricks x bins entries");
  TVK_HIP(hipSetDevice(device_));
  DevBuf h;
  TVK_HIP(h.alloc(nb * bins * 4));
  const int64_t d[3] = {L.dims[0], L.dims[1], L.dims[2]};
  TVK_HIP(launch_brick_histograms(levels_[lod].p, (int)bits_ / 8, d, brick_, (int)bins, h.as<uint32_t>(),
                                  stream_));
  TVK_HIP(hipMemcpyAsync(out, h.p, nb * bins * 4, hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

int HipRenderer::RegionDetail(float* out, uint64_t n) {
  if (!has_volume_) return state_fail("no volume loaded");
  if (paged_) return state_fail("region detail needs an in-core volume");
  const LodInfo& L = lods_[0];
  const uint64_t nr = (uint64_t)L.grid[0] * L.grid[1] * L.grid[2];
  if (!out || n < nr * n_lod_) return arg_fail("output needs n_regions x n_lod entries");
  TVK_HIP(hipSetDevice(device_));
  DevBuf d;
  TVK_HIP(d.alloc(nr * n_lod_ * 4));
  const void* lv[kMaxLod];
  int64_t dims[kMaxLod][3];
  for (int l = 0; l < n_lod_; ++l) {
    lv[l] = levels_[l].p;
    for (int i = 0; i < 3; ++i) dims[l][i] = lods_[l].dims[i];
  }
  TVK_HIP(launch_region_detail(lv, dims, n_lod_, (int)bits_ / 8, brick_, d.as<float>(), stream_));
  TVK_HIP(hipMemcpyAsync(out, d.p, nr * n_lod_ * 4, hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

int HipRenderer::Histogram1D(uint64_t* hist, uint64_t n_bins) {
  if (!has_volume_) return state_fail("no volume loaded");
  const uint64_t nb = 1ull << bits_;
  if (!hist || n_bins != nb) return arg_fail("histogram must have 2^bits bins");
  if (paged_) return state_fail("histograms need an in-core volume (tvk_set_volume / loadDataset)");
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
  if (paged_) return state_fail("histograms need an in-core volume (tvk_set_volume / loadDataset)");
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

int HipRenderer::Set1DTransSlot(uint32_t slot, const float* rgba, uint32_t n) {
  if (slot >= (uint32_t)kMaxTfSlots) return arg_fail("TF slot out of range (0..63)");
  if (!rgba || n < 1 || n > 65536) return arg_fail("1D TF needs 1..65536 entries");
  for (uint64_t i = 0; i < 4ull * n; ++i)
    if (!isfinite(rgba[i])) return arg_fail("1D TF holds a non-finite value");
  if (tfs_.size() <= slot) tfs_.resize(slot + 1);
  TfSet& t = tfs_[slot];
  t.tf1d.assign(rgba, rgba + 4ull * n);
  t.prefix1d.assign(n + 1, 0);
  for (uint32_t i = 0; i < n; ++i) t.prefix1d[i + 1] = t.prefix1d[i] + (rgba[4 * i + 3] > 0.f ? 1 : 0);
  t.dirty1d = true;
  ++t.version;
  if (slot == 0) invalidate();
  return TVK_OK;
}

int HipRenderer::Set2DTransSlot(uint32_t slot, const float* rgba, uint32_t w, uint32_t h,
                                float max_grad) {
  if (slot >= (uint32_t)kMaxTfSlots) return arg_fail("TF slot out of range (0..63)");
  if (!rgba || w < 1 || h < 1 || w > 65536 || h > 4096 || (uint64_t)w * h > (1ull << 24))
    return arg_fail("2D TF size out of range");
  if (!(max_grad > 0.f) || !isfinite(max_grad)) return arg_fail("2D TF max gradient must be > 0");
  for (uint64_t i = 0; i < 4ull * w * h; ++i)
    if (!isfinite(rgba[i])) return arg_fail("2D TF holds a non-finite value");
  if (tfs_.size() <= slot) tfs_.resize(slot + 1);
  TfSet& t = tfs_[slot];
  t.tf2d.assign(rgba, rgba + 4ull * w * h);
  t.w2d = w;
  t.h2d = h;
  t.max_grad = max_grad;
  // prefix count over the value columns holding any alpha > 0
  t.prefix2d.assign(w + 1, 0);
  for (size_t i = 0; i < w; ++i) {
    int any = 0;
    for (size_t j = 0; j < h && !any; ++j) any = rgba[4 * (j * w + i) + 3] > 0.f;
    t.prefix2d[i + 1] = t.prefix2d[i] + any;
  }
  t.dirty2d = true;
  ++t.version;
  if (slot == 0) invalidate();
  return TVK_OK;
}

int HipRenderer::SetIsoValue(float v) {
  if (!isfinite(v)) return arg_fail("iso value must be finite");
  iso_value_ = v;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetUseMIP(bool on) {
  use_mip_ = on;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetMIPRotationAngle(float deg) {
  if (!isfinite(deg)) return arg_fail("MIP rotation angle must be finite");
  mip_angle_ = deg;
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

int HipRenderer::SetClearViewEnabled(bool on) {
  cv_enabled_ = on;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetCVIsoValue(float v) {
  if (!isfinite(v)) return arg_fail("ClearView iso value must be finite");
  cv_iso_ = v;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetCVColor(float r, float g, float b) {
  if (!isfinite(r) || !isfinite(g) || !isfinite(b)) return arg_fail("ClearView colour must be finite");
  cv_color_[0] = r;
  cv_color_[1] = g;
  cv_color_[2] = b;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetCVSize(float v) {
  if (!isfinite(v) || v < 0.f) return arg_fail("ClearView size must be finite and >= 0");
  cv_size_ = v;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetCVContextScale(float v) {
  if (!isfinite(v) || v < 0.f) return arg_fail("ClearView context scale must be finite and >= 0");
  cv_context_ = v;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetCVBorderScale(float v) {
  if (!isfinite(v) || v < 0.f) return arg_fail("ClearView border scale must be finite and >= 0");
  cv_border_ = v;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetCVFocusPos(int32_t x, int32_t y) {
  cv_focus_[0] = x;
  cv_focus_[1] = y;
  invalidate();
  return TVK_OK;
}

void HipRenderer::GetClearView(tvk_clearview* out) const {
  out->enabled = cv_enabled_ ? 1 : 0;
  out->iso_value = cv_iso_;
  for (int i = 0; i < 3; ++i) out->color[i] = cv_color_[i];
  out->size = cv_size_;
  out->context_scale = cv_context_;
  out->border_scale = cv_border_;
  out->focus[0] = cv_focus_[0];
  out->focus[1] = cv_focus_[1];
}

float HipRenderer::iso_skip_threshold() const {
  return cv_enabled_ ? std::min(iso_value_, cv_iso_) : iso_value_;
}

int HipRenderer::upload_tf() {
  // a pending region check's fallback reads the current TF buffers
  if (int e = resolve_check()) return e;
  bool copied = false;
  for (TfSet& t : tfs_) {
    if (t.dirty2d) {
      const size_t W = t.w2d, H = t.h2d;
      if (t.tf2d_dev.bytes < W * H * 16) TVK_HIP(t.tf2d_dev.alloc(W * H * 16));
      if (t.prefix2d_dev.bytes < (W + 1) * 4) TVK_HIP(t.prefix2d_dev.alloc((W + 1) * 4));
      TVK_HIP(hipMemcpyAsync(t.tf2d_dev.p, t.tf2d.data(), W * H * 16, hipMemcpyHostToDevice, stream_));
      TVK_HIP(hipMemcpyAsync(t.prefix2d_dev.p, t.prefix2d.data(), (W + 1) * 4, hipMemcpyHostToDevice,
                             stream_));
      t.dirty2d = false;
      copied = true;
    }
    if (t.dirty1d) {
      const size_t n = t.tf1d.size() / 4;
      if (t.tf1d_dev.bytes < n * 16) TVK_HIP(t.tf1d_dev.alloc(n * 16));
      if (t.prefix1d_dev.bytes < (n + 1) * 4) TVK_HIP(t.prefix1d_dev.alloc((n + 1) * 4));
      TVK_HIP(hipMemcpyAsync(t.tf1d_dev.p, t.tf1d.data(), n * 16, hipMemcpyHostToDevice, stream_));
      TVK_HIP(hipMemcpyAsync(t.prefix1d_dev.p, t.prefix1d.data(), (n + 1) * 4, hipMemcpyHostToDevice,
                             stream_));
      t.dirty1d = false;
      copied = true;
    }
  }
  // the host vectors may change before the copies run: wait for them
  if (copied) TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

// ----------------------------------------------------------------- camera
int HipRenderer::SetViewPos(const float v[3]) {
  if (!finite3(v)) return arg_fail("view position must be finite");
  for (int i = 0; i < 3; ++i) cam_.pos[i] = v[i];
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetViewDir(const float v[3]) {
  if (!finite3(v) || (v[0] == 0 && v[1] == 0 && v[2] == 0)) return arg_fail("view direction must be non-zero");
  for (int i = 0; i < 3; ++i) cam_.dir[i] = v[i];
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetUpDir(const float v[3]) {
  if (!finite3(v) || (v[0] == 0 && v[1] == 0 && v[2] == 0)) return arg_fail("up direction must be non-zero");
  for (int i = 0; i < 3; ++i) cam_.up[i] = v[i];
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetFoV(float deg) {
  if (!(deg > 0.f && deg < 180.f)) return arg_fail("field of view must be in (0, 180)");
  cam_.fov = deg;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetOrtho(bool on, float hh) {
  if (on && !(hh > 0.f && isfinite(hh))) return arg_fail("orthographic half height must be > 0");
  cam_.ortho = on;
  if (on) cam_.ortho_hh = hh;
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

int HipRenderer::SetClipPlane(const float plane[4]) {
  if (!plane) return arg_fail("null clip plane");
  for (int i = 0; i < 4; ++i)
    if (!isfinite(plane[i])) return arg_fail("clip plane must be finite");
  if (plane[0] == 0.f && plane[1] == 0.f && plane[2] == 0.f) return arg_fail("clip plane normal must be non-zero");
  for (int i = 0; i < 4; ++i) clip_[i] = plane[i];
  invalidate();
  return TVK_OK;
}

int HipRenderer::EnableClipPlane(bool on) {
  clip_on_ = on;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetBlendPrecision(int precision) {
  if (precision != TVK_BP_8BIT && precision != TVK_BP_16BIT && precision != TVK_BP_32BIT)
    return arg_fail("blend precision must be 0 (8 bit), 1 (16 bit) or 2 (32 bit)");
  blend_ = precision;
  invalidate();
  return TVK_OK;
}

int HipRenderer::SetPerfMeasures(uint32_t min_fps, bool low_res, float screen_dec, float sample_dec,
                                 uint32_t lod_delay) {
  if (!(screen_dec >= 1.f && screen_dec <= 64.f) || !(sample_dec >= 1.f && sample_dec <= 64.f))
    return arg_fail("resolution and sample-rate decrease factors must be in [1, 64]");
  min_fps_ = min_fps;
  low_res_ = low_res;
  screen_dec_ = screen_dec;
  sample_dec_ = sample_dec;
  lod_delay_ = lod_delay;
  invalidate();
  return TVK_OK;
}

// Predicted kernel time of a subframe at `lod`: its own last measurement
// (normalised to full size and rate), else the nearest measured level's
// scaled by 2 per level (each level halves the samples along a ray); an
// intermediate subframe (lod > lo) with low-resolution intermediates costs
// the pixel and sample-rate fractions of that.  < 0: nothing measured yet.
double HipRenderer::predicted_ms(int lod, int lo) const {
  double ms = lod_ms_[lod];
  if (ms < 0.0) {
    for (int dl = 1; dl < n_lod_ && ms < 0.0; ++dl) {
      if (lod + dl < n_lod_ && lod_ms_[lod + dl] >= 0.0) ms = lod_ms_[lod + dl] * ldexp(1.0, dl);
      else if (lod - dl >= 0 && lod_ms_[lod - dl] >= 0.0) ms = lod_ms_[lod - dl] * ldexp(1.0, -dl);
    }
  }
  if (ms >= 0.0 && low_res_ && lod > lo) ms /= (double)screen_dec_ * screen_dec_ * sample_dec_;
  return ms;
}

// The finest allowed LOD whose predicted subframe time fits the frame
// budget 1000 / min_fps ms; the coarsest when nothing fits or nothing has
// been measured (the reference's default: start coarse, refine).
int HipRenderer::governor_start_lod(int lo, int hi) const {
  if (min_fps_ == 0) return hi;
  const double budget = 1000.0 / (double)min_fps_;
  for (int l = lo; l <= hi; ++l) {
    const double ms = predicted_ms(l, lo);
    if (ms < 0.0) return hi;
    if (ms <= budget) return l;
  }
  return hi;
}

void HipRenderer::record_subframe_time() {
  if (!time_pending_ || !stats_slot_ || last_lod_ < 0) return;
  time_pending_ = false;
  LaunchSlot& sl = *stats_slot_;
  if (hipEventSynchronize(sl.ev[3]) != hipSuccess) return;
  float k = 0.f;
  if (hipEventElapsedTime(&k, sl.ev[1], sl.ev[2]) != hipSuccess) return;
  const double full = (double)k / last_scale_;
  double& m = lod_ms_[last_lod_];
  m = m < 0.0 ? full : 0.5 * m + 0.5 * full;
}

int HipRenderer::ReadFramebuffer(uint8_t* rgba, uint64_t bytes) {
  if (!initialized_ || !framebuffer_.p) return state_fail("no framebuffer (initialize + resize first)");
  if (!rgba || bytes < (uint64_t)width_ * height_ * 4) return arg_fail("framebuffer readback buffer too small");
  if (int e = resolve_check()) return e;
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipMemcpyAsync(rgba, framebuffer_.p, (size_t)width_ * height_ * 4, hipMemcpyDeviceToHost, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
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

int HipRenderer::SetWorkGroupWaves(uint32_t waves) {
  if (waves != 4 && waves != 8) return arg_fail("work-group waves must be 4 or 8");
  wg_waves_ = (int)waves;
  return TVK_OK;
}

int HipRenderer::SetBrickStaging(bool on) {
  stage_ = on;
  return TVK_OK;
}

int HipRenderer::SetTFGrouping(bool on) {
  group_tfs_ = on;
  return TVK_OK;
}

int HipRenderer::SetRayTileShape(uint32_t tile_width) {
  if (tile_width != 4 && tile_width != 8 && tile_width != 16) return arg_fail("tile width must be 4, 8 or 16");
  tile_shift_ = tile_width == 4 ? 2 : (tile_width == 8 ? 3 : 4);
  return TVK_OK;
}

int HipRenderer::SetEmptySpaceCell(uint32_t cell) {
  if (cell < 4 || cell > 1024) return arg_fail("empty-space cell must be in [4, 1024]");
  cell_pref_ = (int)cell;
  return TVK_OK;
}

// ----------------------------------------------------------------- render
int HipRenderer::check_ready() const {
  if (!initialized_) return state_fail("paint before initialize");
  if (!has_volume_) return state_fail("paint before a dataset is loaded");
  if (width_ == 0) return state_fail("paint before resize");
  return check_tf(0);
}

int HipRenderer::check_tf(int slot) const {
  const bool have = slot >= 0 && slot < (int)tfs_.size();
  if (render_mode_ == TVK_RM_1DTRANS && (!have || tfs_[slot].tf1d.empty()))
    return state_fail("render before a 1D TF is set (slot " + std::to_string(slot) + ")");
  if (render_mode_ == TVK_RM_2DTRANS && (!have || tfs_[slot].tf2d.empty()))
    return state_fail("render before a 2D TF is set (slot " + std::to_string(slot) + ")");
  // the value + gradient fetch reads one voxel beyond the trilinear cell
  if ((render_mode_ == TVK_RM_2DTRANS || (lighting_ && render_mode_ == TVK_RM_1DTRANS)) &&
      has_volume_ && apron_ < 2)
    return state_fail("gradient-based rendering (2D TF, lighting) needs a brick apron >= 2");
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

// Kernel render mode of this handle's jobs: 0 1D TF, 1 2D TF, 2 isosurface,
// 3 maximum intensity projection (setUseMIP in 1D TF mode).
int HipRenderer::job_mode() const {
  if (render_mode_ == TVK_RM_2DTRANS) return 1;
  if (render_mode_ == TVK_RM_ISOSURFACE) return 2;
  return use_mip_ ? 3 : 0;
}

void HipRenderer::build_params(RaycastParams* p, int lod, const CameraState& cam_in, uint32_t width,
                               uint32_t height, int tf_slot, uint32_t* framebuffer, float rate_div) const {
  memset(p, 0, sizeof(*p));
  // MIP rotation (setMIPRotationAngle, RenderWindow.cpp:1687-1703): the view
  // turned about its up vector around the volume centre (world origin)
  CameraState cam = cam_in;
  if (job_mode() == 3 && mip_angle_ != 0.f) {
    double ax[3] = {cam.up[0], cam.up[1], cam.up[2]};
    normalize(ax);
    const double a = (double)mip_angle_ * M_PI / 180.0, ca = cos(a), sa = sin(a);
    auto rot = [&](double v[3]) {  // Rodrigues
      double kxv[3], out[3];
      cross(ax, v, kxv);
      const double kv = ax[0] * v[0] + ax[1] * v[1] + ax[2] * v[2];
      for (int i = 0; i < 3; ++i) out[i] = v[i] * ca + kxv[i] * sa + ax[i] * kv * (1.0 - ca);
      for (int i = 0; i < 3; ++i) v[i] = out[i];
    };
    rot(cam.pos);
    rot(cam.dir);
  }
  const TfSet& tfs = tfs_[tf_slot];
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
  double f[3] = {cam.dir[0], cam.dir[1], cam.dir[2]}, r[3], u[3];
  normalize(f);
  cross(f, cam.up, r);
  normalize(r);
  cross(r, f, u);
  const double W = width, H = height, asp = W / H;
  double tmp[3], dd[3];
  if (cam.ortho) {
    const double hh = cam.ortho_hh;
    double o00[3];
    for (int i = 0; i < 3; ++i) o00[i] = cam.pos[i] + (1.0 / W - 1.0) * hh * asp * r[i] + (1.0 - 1.0 / H) * hh * u[i];
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
    const double th = tan(cam.fov * 0.5 * M_PI / 180.0);
    vox_point(cam.pos, p->eye);
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
  p->width = (int32_t)width;
  p->height = (int32_t)height;
  p->tile_shift = tile_shift_;
  p->tiles_x = (p->width + (1 << tile_shift_) - 1) >> tile_shift_;
  const int tile_h = 64 >> tile_shift_;
  p->tiles_y = (p->height + tile_h - 1) / tile_h;
  for (int i = 0; i < 3; ++i) p->box[i] = (float)dims0_[i];
  set_tile_rect(p);
  p->step = (float)(ldexp(1.0, lod) * (double)rate_div / (double)rate_);
  p->nominal_rate = rate_;
  p->term_alpha = term_alpha_;
  if (clip_on_) {
    // world plane n.w + d with w = m.M (row vectors) and voxel v = m * vscale
    // + voff: the plane in voxel space is (M n) / vscale, d + M[3].n - ...
    double nm[3], dm = clip_[3];
    for (int i = 0; i < 3; ++i) nm[i] = M[i * 4 + 0] * clip_[0] + M[i * 4 + 1] * clip_[1] + M[i * 4 + 2] * clip_[2];
    for (int j = 0; j < 3; ++j) dm += M[12 + j] * clip_[j];
    double nv[3], dv = dm, len = 0.0;
    for (int i = 0; i < 3; ++i) {
      nv[i] = nm[i] / vscale[i];
      dv -= nv[i] * voff[i];
      len += nv[i] * nv[i];
    }
    len = sqrt(len);  // normalised: the ray parameter is in voxels either way
    for (int i = 0; i < 3; ++i) p->clip[i] = (float)(nv[i] / len);
    p->clip[3] = (float)(dv / len);
    p->clip_on = 1;
  }
  p->blend_bits = blend_ == TVK_BP_8BIT ? 8 : (blend_ == TVK_BP_16BIT ? 16 : 32);
  p->interp = interp_;
  p->batch = batch_;
  p->tile_group = tile_group_;
  p->xcd_split = xcd_split_;
  p->compute_units = compute_units_;
  p->skip = (skip_ && job_mode() != 3) ? 1 : 0;
  p->level = lods_[lod];
  p->brick = brick_;
  p->apron = apron_;
  p->slot_edge = slot_edge_;
  p->n_tf = (int32_t)(tfs.tf1d.size() / 4);
  p->tf_scale = tf_scale(p->n_tf);
  for (int i = 0; i < 3; ++i) {
    p->bg_top[i] = bg_top_[i];
    p->bg_bottom[i] = bg_bottom_[i];
  }
  p->bg_alpha = bg_alpha_;
  p->page_table = pool_.page_table.as<uint64_t>();
  p->brick_range = pool_.brick_range.as<float>();
  p->cell_range = cell_range_.as<uint32_t>();
  p->cell_empty = tfs.cell_empty[mode_index()].as<uint8_t>();
  p->cell = skip_ ? cell_ : brick_;
  p->tf = tfs.tf1d_dev.as<float>();
  p->tf_prefix = tfs.prefix1d_dev.as<int32_t>();
  p->framebuffer = framebuffer;

  // render mode, gradient units, lighting
  p->mode = job_mode();
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
    p->tf2d = tfs.tf2d_dev.as<float>();
    p->tf2d_w = (int32_t)tfs.w2d;
    p->tf2d_h = (int32_t)tfs.h2d;
    p->tf2d_row_scale = (float)((double)(tfs.h2d - 1) / (double)tfs.max_grad);
    p->n_tf = (int32_t)tfs.w2d;  // the skip test runs over the value columns
    p->tf_scale = tf_scale((int)tfs.w2d);
    p->tf_prefix = tfs.prefix2d_dev.as<int32_t>();
  }
  p->iso_value = iso_value_;
  for (int i = 0; i < 3; ++i) p->iso_color[i] = iso_color_[i];
  p->cv = (cv_enabled_ && p->mode == 2) ? 1 : 0;
  p->cv_iso_value = cv_iso_;
  for (int i = 0; i < 3; ++i) p->cv_color[i] = cv_color_[i];
  p->cv_focus[0] = (float)cv_focus_[0];
  p->cv_focus[1] = (float)cv_focus_[1];
  p->cv_size = cv_size_;
  p->cv_context_scale = cv_context_;
  p->cv_border_scale = cv_border_;
  p->lods = lods_dev_.as<LodInfo>();
  if (regions_active() && p->mode == 0 && !p->lighting) p->region_code = tfs.region_code.as<uint8_t>();
}

LaunchInfo HipRenderer::launch_info(int n_tf_lds) const {
  LaunchInfo li;
  li.bpv = (int)bits_ / 8;
  li.interp = interp_;
  li.mode = job_mode();
  li.lighting = (lighting_ && li.mode == 0) ? 1 : 0;
  li.batch = batch_;
  li.n_tf_lds = ((li.mode == 0 || li.mode == 3) && n_tf_lds <= 4096) ? n_tf_lds : 0;
  li.regions = (regions_active() && li.mode == 0 && !li.lighting) ? 1 : 0;
  li.compute_units = compute_units_;
  li.wg_waves = wg_waves_;
  li.stage = (stage_ && li.mode == 0 && !li.lighting && !li.regions && li.interp == 1) ? 1 : 0;
  li.group = (group_tfs_ && (li.mode == 1 || li.lighting)) ? 1 : 0;
  return li;
}

// Tile slots of a job in the queues: its rect's tiles, or one per run of
// tile_sample of them (a part: its own slots).
int job_tiles(const RaycastParams& p) {
  if (p.group_n < 0) return 0;  // a TF group member: its leader's rays render it
  if (p.tile_count > 0) return p.tile_count;  // a part of a split job
  const int all = p.rect_w * p.rect_h;
  return p.tile_sample <= 1 ? all : (all + p.tile_sample - 1) / p.tile_sample;
}

// Can jobs a and b share one march?  Every field equal except those of the
// TF, the output framebuffer and the group fields.
static bool same_but_tf(const RaycastParams& a, const RaycastParams& b) {
  RaycastParams x, y;
  memcpy(&x, &a, sizeof x);
  memcpy(&y, &b, sizeof y);
  for (RaycastParams* p : {&x, &y}) {
    p->tf2d = nullptr;
    p->tf2d_w = p->tf2d_h = 0;
    p->tf2d_row_scale = p->tf_scale = 0.f;
    p->n_tf = 0;
    p->tf = nullptr;
    p->tf_prefix = nullptr;
    p->cell_empty = nullptr;
    p->framebuffer = nullptr;
    p->group_n = 0;
    memset(p->group_jobs, 0, sizeof p->group_jobs);
  }
  return memcmp(&x, &y, sizeof x) == 0;
}

// Fill the pinned table `t` from jobs[0..n) for a device table at
// dev_table whose counters live at dev_counters; returns the tiles of the
// launch and whether some job has background pixels outside its tile rect.
// With li.group, jobs that differ only in their TF are grouped (up to
// kGroupMax, the first of them leading); *groups = the groups formed.
static int fill_job_table(JobTable* t, const RaycastParams* jobs, int n, const LaunchInfo& li,
                          int tile_group, int xcd_split, unsigned long long* dev_counters, bool* culled,
                          int* groups) {
  t->h.n_jobs = n;
  t->h.tile_group = tile_group;
  t->h.xcd_split = xcd_split;
  t->h.n_tf_lds = li.n_tf_lds;
  t->h.tf_lds = li.n_tf_lds > 0 ? jobs[0].tf : nullptr;
  t->h.counters = dev_counters;
  *groups = 0;
  for (int k = 0; k < n; ++k) {
    t->jobs[k] = jobs[k];
    t->jobs[k].group_n = 0;
    memset(t->jobs[k].group_jobs, 0, sizeof t->jobs[k].group_jobs);
  }
  if (li.group) {
    for (int k = 0; k < n; ++k) {
      RaycastParams& lead = t->jobs[k];
      if (lead.group_n != 0) continue;
      int cnt = 1;
      for (int m = k + 1; m < n && cnt < kGroupMax; ++m) {
        // compared as the caller built them (zeroed, so padding compares equal)
        if (t->jobs[m].group_n != 0 || !same_but_tf(jobs[k], jobs[m])) continue;
        lead.group_jobs[cnt - 1] = m;
        t->jobs[m].group_n = -1;
        ++cnt;
      }
      if (cnt > 1) {
        lead.group_n = cnt;
        ++*groups;
      }
    }
  }
  int total = 0;
  *culled = false;
  for (int k = 0; k < n; ++k) {
    const RaycastParams& j = t->jobs[k];
    t->tile_start[k] = total;
    total += job_tiles(j);
    *culled = *culled || j.rect_w * j.rect_h != j.tiles_x * j.tiles_y;
    // the frame-region split addresses tiles by rect position: sampled
    // jobs and parts take the round-robin runs
    if (j.tile_sample > 1 || j.tile_count > 0) t->h.xcd_split = 0;
  }
  t->tile_start[n] = total;
  t->h.total_tiles = total;
  return total;
}

int launch_job_table_at(JobTable* table_h, void* dev_table, unsigned long long* dev_counters,
                        unsigned long long* counters_h, hipEvent_t ev[4], const RaycastParams* jobs, int n,
                        const LaunchInfo& li_in, int tile_group, int xcd_split, hipStream_t stream_) {
  bool culled = false;  // some job has background tiles outside its rect
  int groups = 0;
  fill_job_table(table_h, jobs, n, li_in, tile_group, xcd_split, dev_counters, &culled, &groups);
  LaunchInfo li = li_in;
  li.group = groups > 0 ? 1 : 0;
  const int words = kQueues + kJobCounters * n;
  const size_t table_bytes = offsetof(JobTable, jobs) + (size_t)n * sizeof(RaycastParams);
  TVK_HIP(hipEventRecord(ev[0], stream_));
  TVK_HIP(hipMemcpyAsync(dev_table, table_h, table_bytes, hipMemcpyHostToDevice, stream_));
  TVK_HIP(launch_fill_u64(dev_counters, words, 0ull, stream_));
  TVK_HIP(hipEventRecord(ev[1], stream_));
  if (culled) TVK_HIP(launch_background((const JobTable*)dev_table, stream_));
  TVK_HIP(launch_raycast(li, (const JobTable*)dev_table, stream_));
  TVK_HIP(hipEventRecord(ev[2], stream_));
  if (counters_h)
    TVK_HIP(hipMemcpyAsync(counters_h, dev_counters, words * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           stream_));
  TVK_HIP(hipEventRecord(ev[3], stream_));
  return TVK_OK;
}

int launch_job_table(LaunchSlot& sl, const RaycastParams* jobs, int n, const LaunchInfo& li_in,
                     int tile_group, int xcd_split, hipStream_t stream_, bool graph) {
  // the pinned table may still be read by this slot's previous upload
  if (sl.used) TVK_HIP(hipEventSynchronize(sl.ev[3]));
  sl.used = true;
  if (!graph)
    return launch_job_table_at(sl.table_h, sl.table.p, sl.counters.as<unsigned long long>(), sl.counters_h,
                               sl.ev, jobs, n, li_in, tile_group, xcd_split, stream_);
  JobTable* t = sl.table_h;
  bool culled = false;
  int groups = 0;
  fill_job_table(t, jobs, n, li_in, tile_group, xcd_split, sl.counters.as<unsigned long long>(), &culled,
                 &groups);
  LaunchInfo li = li_in;
  li.group = groups > 0 ? 1 : 0;
  TVK_HIP(hipEventRecord(sl.ev[0], stream_));
  // Graph replay: the graph copies the whole table and all counters, so one
  // capture serves any job count; it is re-captured when the kernel
  // variant or its LDS size changes.
  const uint64_t key = 1ull | (uint64_t)li.bpv << 1 | (uint64_t)li.interp << 3 | (uint64_t)li.mode << 4 |
                       (uint64_t)li.lighting << 6 | (uint64_t)li.batch << 7 | (uint64_t)li.regions << 11 |
                       (uint64_t)li.n_tf_lds << 12 |
                       (uint64_t)(li.wg_waves == 4 ? 1 : 0) << 40 | (uint64_t)li.stage << 41 |
                       (uint64_t)li.group << 42;
  if (!sl.exec || sl.graph_key != key) {
    if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
    if (sl.graph) (void)hipGraphDestroy(sl.graph);
    sl.exec = nullptr;
    sl.graph = nullptr;
    TVK_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipMemcpyAsync(sl.table.p, t, sizeof(JobTable), hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess)
      e = launch_fill_u64(sl.counters.as<unsigned long long>(), kCounterWords, 0ull, stream_);
    if (e == hipSuccess) e = launch_background(sl.table.as<JobTable>(), stream_);
    if (e == hipSuccess) e = launch_raycast(li, sl.table.as<JobTable>(), stream_);
    if (e == hipSuccess)
      e = hipMemcpyAsync(sl.counters_h, sl.counters.p, kCounterWords * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, stream_);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(stream_, &g);
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(e, "raycast graph capture");
    }
    TVK_HIP(e2);
    sl.graph = g;
    TVK_HIP(hipGraphInstantiate(&sl.exec, sl.graph, nullptr, nullptr, 0));
    sl.graph_key = key;
  }
  TVK_HIP(hipEventRecord(sl.ev[1], stream_));
  TVK_HIP(hipGraphLaunch(sl.exec, stream_));
  TVK_HIP(hipEventRecord(sl.ev[2], stream_));
  TVK_HIP(hipEventRecord(sl.ev[3], stream_));
  return TVK_OK;
}

int HipRenderer::launch_jobs(LaunchSlot& sl, const RaycastParams* jobs, int n, const LaunchInfo& li,
                             bool graph) {
  return launch_job_table(sl, jobs, n, li, tile_group_, xcd_split_, stream_, graph);
}

float HipRenderer::tf_scale(int n_entries) const {
  return bias_scale_tf_ ? (float)((double)(n_entries - 1) / value_max()) : 1.0f;
}

int HipRenderer::prepare_cells(int tf_slot, int mode) {
  TfSet& t = tfs_[tf_slot];
  if (!skip_ || total_cells_ == 0) return TVK_OK;
  const std::vector<int32_t>& prefix = mode == 1 ? t.prefix2d : t.prefix1d;
  const int n = (int)prefix.size() - 1;
  const float scale = tf_scale(n);
  uint32_t iso_bits = 0, scale_bits = 0;
  const float iso = iso_skip_threshold();
  memcpy(&iso_bits, &iso, 4);
  memcpy(&scale_bits, &scale, 4);
  // FNV-1a over what the flags depend on
  uint64_t key = 1469598103934665603ull;
  for (const uint64_t v : {cells_gen_, (uint64_t)mode, t.version, (uint64_t)(mode == 2 ? iso_bits : 0),
                           (uint64_t)scale_bits, (uint64_t)n})
    key = (key ^ v) * 1099511628211ull;
  DevBuf& flags = t.cell_empty[mode];
  if (key == t.cell_key[mode] && flags.p) return TVK_OK;
  if (flags.bytes < (size_t)total_cells_) TVK_HIP(flags.alloc((size_t)total_cells_));
  const int32_t* prefix_dev = mode == 1 ? t.prefix2d_dev.as<int32_t>() : t.prefix1d_dev.as<int32_t>();
  if (mode != 2 && (n < 1 || !prefix_dev)) return state_fail("TF not uploaded");
  TVK_HIP(launch_cell_empty(cell_range_.as<uint32_t>(), total_cells_, prefix_dev, n, scale, mode,
                            iso, flags.as<uint8_t>(), stream_));
  t.cell_key[mode] = key;
  return TVK_OK;
}

uint32_t HipRenderer::bricks_skipped(int lod, int tf_slot) const {
  // bricks whose whole value range is transparent (below the iso value);
  // maximum intensity projection visits every brick
  if (!skip_ || job_mode() == 3) return 0;
  const TfSet& t = tfs_[tf_slot];
  const LodInfo& L = lods_[lod];
  const int nb = L.grid[0] * L.grid[1] * L.grid[2];
  const int mode = render_mode_ == TVK_RM_2DTRANS ? 1 : (render_mode_ == TVK_RM_ISOSURFACE ? 2 : 0);
  const std::vector<int32_t>& prefix = mode == 1 ? t.prefix2d : t.prefix1d;
  const int ntf = (int)prefix.size() - 1;
  const float scale = ntf > 0 ? tf_scale(ntf) : 1.0f;
  uint32_t skipped = 0;
  for (int b = 0; b < nb; ++b) {
    const float mn = pool_.brick_range_h[2 * (L.brick_base + b)];
    const float mx = pool_.brick_range_h[2 * (L.brick_base + b) + 1];
    if (mode == 2) {
      if (mx < iso_skip_threshold()) ++skipped;
      continue;
    }
    if (ntf < 1) continue;
    const int i0 = (int)std::min(std::max(mn * scale, 0.f), (float)(ntf - 1));
    const int i1 = std::min((int)std::min(std::max(mx * scale, 0.f), (float)(ntf - 1)) + 1, ntf - 1);
    if (prefix[i1 + 1] - prefix[i0] == 0) ++skipped;
  }
  return skipped;
}

int HipRenderer::Paint() {
  if (int e = check_ready()) return e;
  if (int e = resolve_check()) return e;  // the last frame's check ends before this one starts
  TVK_HIP(hipSetDevice(device_));
  if (int e = upload_tf()) return e;
  const int lo = std::min<int>((int)lod_min_, n_lod_ - 1);
  const int hi = std::max(lo, std::min<int>((int)lod_max_, n_lod_ - 1));
  const bool interactive = target_ == TVK_RT_INTERACTIVE;
  const bool governed = interactive && min_fps_ > 0;
  if (governed) record_subframe_time();
  int lod;
  if (!interactive) lod = lo;
  else if (frame_complete_) lod = lo;
  else if (next_lod_ < 0) {
    // a new frame: the governor's start LOD (the coarsest without it)
    lod = governor_start_lod(lo, hi);
    delay_left_ = lod_delay_;
  } else {
    lod = std::max(lo, next_lod_);
    if (governed && !capturing_ && delay_left_ > 0 && lod < last_lod_) {
      --delay_left_;  // LOD delay: hold the refinement for this many paints
      return TVK_OK;
    }
  }
  if (paged_) lod = std::max(lod, std::min(lod_floor_, hi));
  // intermediate subframes at a reduced resolution and sample rate
  const bool lowres = governed && low_res_ && lod > lo && !regions_active() && !region_out_on_ &&
                      (screen_dec_ > 1.f || sample_dec_ > 1.f);
  uint32_t rw = width_, rh = height_;
  float rdiv = 1.f;
  if (lowres) {
    rw = std::max<uint32_t>(1, (uint32_t)floorf((float)width_ / screen_dec_));
    rh = std::max<uint32_t>(1, (uint32_t)floorf((float)height_ / screen_dec_));
    rdiv = sample_dec_;
    if (lowres_fb_.bytes < (size_t)rw * rh * 4) TVK_HIP(lowres_fb_.alloc((size_t)rw * rh * 4));
  }

  if (int e = prepare_cells(0, mode_index())) return e;
  if (regions_active()) {
    if (render_mode_ != TVK_RM_1DTRANS || lighting_)
      return state_fail("per-region LOD renders the 1D TF without lighting");
    if (check_k_ > 0 && paged_) return state_fail("the per-frame region check renders in-core volumes");
    if (int e = prepare_regions(0)) return e;
  }
  RaycastParams p;
  build_params(&p, lod, cam_, rw, rh, 0, lowres ? lowres_fb_.as<uint32_t>() : framebuffer_.as<uint32_t>(),
               rdiv);
  if (regions_active() && region_out_on_) {
    const size_t need = (size_t)width_ * height_ * 4;
    if (region_out_dev_.bytes < need) TVK_HIP(region_out_dev_.alloc(need));
    p.region_out = region_out_dev_.as<int32_t>();
  }
  if (paged_) {
    ++frame_;
    p.requests = requests_dev_.as<uint32_t>();
    p.touch = touch_dev_.as<uint32_t>();
    p.frame = frame_;
    TVK_HIP(launch_fill_u64(requests_dev_.as<unsigned long long>(), (int64_t)(requests_dev_.bytes / 8),
                            0ull, stream_));
  }
  // a paged volume's bricks sampled from here on are this frame's (a pass
  // that repeats an incomplete one keeps the frame going)
  if (paged_ && !(last_pass_incomplete_ && lod == last_lod_)) MemMan::ForDevice(device_).BeginFrame(pool_client_);
  LaunchSlot& sl = paint_slots_[paint_cur_];
  paint_cur_ ^= 1;
  {
    std::lock_guard<std::mutex> g(pool_mu_);  // no eviction between the page table and the launch
    if (int e = sync_page_table_locked()) return e;
    if (int e = launch_jobs(sl, &p, 1, launch_info(p.n_tf), false)) return e;
  }
  if (check_k_ > 0 && regions_active())
    if (int e = check_launch(p)) return e;
  if (lowres)
    TVK_HIP(launch_upscale_nearest(lowres_fb_.as<uint32_t>(), (int)rw, (int)rh, framebuffer_.as<uint32_t>(),
                                   (int)width_, (int)height_, stream_));
  stats_slot_ = &sl;
  last_scale_ = lowres ? (double)rw * rh / ((double)width_ * height_) / (double)rdiv : 1.0;
  time_pending_ = governed;
  bool pass_complete = true;
  if (paged_) {
    // read the brick requests of this pass and page the bricks in
    TVK_HIP(hipMemcpyAsync(requests_h_, requests_dev_.p, requests_dev_.bytes, hipMemcpyDeviceToHost,
                           stream_));
    TVK_HIP(hipStreamSynchronize(stream_));
    bool overflow = false;
    if (regions_active()) {
      if (int e = service_requests(0, n_lod_ - 1, &pass_complete, &overflow)) return e;
      if (overflow) {
        set_last_error("the pool cannot hold this view's region bricks");
        return TVK_ERR_OOM;
      }
    } else {
      if (int e = service_requests(lod, &pass_complete, &overflow)) return e;
      if (overflow) {
        if (lod >= hi) {
          set_last_error("the pool cannot hold this view's bricks at the coarsest allowed LOD");
          return TVK_ERR_OOM;
        }
        lod_floor_ = lod + 1;  // the view's bricks do not fit: go coarser
      }
    }
  }

  // frame bookkeeping and the host-side part of the stats
  const LodInfo& L = lods_[lod];
  stats_ = tvk_frame_stats{};
  stats_.bricks_total = (uint32_t)(L.grid[0] * L.grid[1] * L.grid[2]);
  stats_.bricks_skipped = bricks_skipped(lod, 0);
  stats_.lod = (uint32_t)lod;
  stats_.width = width_;
  stats_.height = height_;
  stats_pending_ = true;
  last_lod_ = lod;
  last_pass_incomplete_ = !pass_complete;
  ++subframes_;
  if (!pass_complete) {
    // bricks were paged in: render this LOD again (a coarser one after an
    // overflow) before the frame counts as done
    frame_complete_ = false;
    next_lod_ = std::max(lod, std::min(lod_floor_, hi));
  } else if (lod <= std::max(lo, paged_ ? std::min(lod_floor_, hi) : lo)) {
    frame_complete_ = true;
    next_lod_ = -1;
  } else {
    frame_complete_ = false;
    next_lod_ = lod - 1;
  }
  return TVK_OK;
}

int HipRenderer::RenderBatch(const tvk_view* views, uint32_t n, tvk_view_stats* stats,
                             float* kernel_ms) {
  if (!initialized_) return state_fail("render before initialize");
  if (!has_volume_) return state_fail("render before a dataset is loaded");
  if (n == 0) return TVK_OK;
  if (!views) return arg_fail("null view array");
  if (int e = resolve_check()) return e;  // its fallback reads buffers a batch may replace
  std::vector<CameraState> cams(n);
  uint64_t fb_bytes = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (int e = view_camera(views[i], i, &cams[i])) return e;
    if (int e = check_tf((int)views[i].tf_slot)) return e;
    fb_bytes += ((uint64_t)views[i].width * views[i].height * 4 + 255) & ~255ull;
  }
  TVK_HIP(hipSetDevice(device_));
  if (int e = upload_tf()) return e;
  if (batch_fb_.bytes < fb_bytes) {
    TVK_HIP(hipStreamSynchronize(stream_));
    TVK_HIP(batch_fb_.alloc(fb_bytes));
  }
  batch_fb_offset_.assign(n, 0);
  uint64_t off = 0;
  for (uint32_t i = 0; i < n; ++i) {
    batch_fb_offset_[i] = off;
    off += ((uint64_t)views[i].width * views[i].height * 4 + 255) & ~255ull;
  }
  const int lo = std::min<int>((int)lod_min_, n_lod_ - 1);
  const int hi = std::max(lo, std::min<int>((int)lod_max_, n_lod_ - 1));
  float kms = 0.f;
  for (uint32_t first = 0; first < n; first += kMaxJobs) {
    const int cnt = (int)std::min<uint32_t>(kMaxJobs, n - first);
    // one shared 1D TF goes to LDS
    bool same_tf = true;
    for (int k = 1; k < cnt; ++k) same_tf = same_tf && views[first + k].tf_slot == views[first].tf_slot;
    const int n_tf0 = (int)(tfs_[views[first].tf_slot].tf1d.size() / 4);
    const LaunchInfo li = launch_info(same_tf ? n_tf0 : 0);
    int lod = lo;
    if (paged_) lod = std::max(lod, std::min(lod_floor_, hi));
    std::vector<RaycastParams> jobs(cnt);
    int batch_lod = -1;
    // paged volumes: a pass per set of requested bricks, as in a capture
    for (int pass = 0; pass < 4 * kMaxLod + 8; ++pass) {
      for (int k = 0; k < cnt; ++k) {
        if (int e = prepare_cells((int)views[first + k].tf_slot, mode_index())) return e;
        if (regions_active() && mode_index() == 0 && !lighting_)
          if (int e = prepare_regions((int)views[first + k].tf_slot)) return e;
      }
      for (int k = 0; k < cnt; ++k) {
        const uint32_t i = first + k;
        build_params(&jobs[k], lod, cams[i], views[i].width, views[i].height, (int)views[i].tf_slot,
                     (uint32_t*)((uint8_t*)batch_fb_.p + batch_fb_offset_[i]));
        if (paged_) {
          jobs[k].requests = requests_dev_.as<uint32_t>();
          jobs[k].touch = touch_dev_.as<uint32_t>();
          jobs[k].frame = frame_ + 1;
        }
      }
      if (paged_) {
        ++frame_;
        if (pass == 0 || lod != batch_lod) MemMan::ForDevice(device_).BeginFrame(pool_client_);
        batch_lod = lod;
        TVK_HIP(launch_fill_u64(requests_dev_.as<unsigned long long>(),
                                (int64_t)(requests_dev_.bytes / 8), 0ull, stream_));
      }
      {
        std::lock_guard<std::mutex> g(pool_mu_);  // no eviction between the page table and the launch
        if (int e = sync_page_table_locked()) return e;
        if (int e = launch_jobs(batch_slot_, jobs.data(), cnt, li, use_graphs_)) return e;
      }
      if (paged_)
        TVK_HIP(hipMemcpyAsync(requests_h_, requests_dev_.p, requests_dev_.bytes,
                               hipMemcpyDeviceToHost, stream_));
      TVK_HIP(hipEventSynchronize(batch_slot_.ev[3]));
      float k_ms = 0.f;
      TVK_HIP(hipEventElapsedTime(&k_ms, batch_slot_.ev[1], batch_slot_.ev[2]));
      kms += k_ms;
      bool complete = true, overflow = false;
      if (paged_) {
        TVK_HIP(hipStreamSynchronize(stream_));
        if (int e = regions_active() ? service_requests(0, n_lod_ - 1, &complete, &overflow)
                                     : service_requests(lod, &complete, &overflow))
          return e;
        if (overflow) {
          if (regions_active() || lod >= hi) {
            set_last_error("the pool cannot hold the batch's bricks at the coarsest allowed LOD");
            return TVK_ERR_OOM;
          }
          lod_floor_ = lod + 1;
          lod = lod + 1;
        }
      }
      if (complete) break;
      if (pass + 1 == 4 * kMaxLod + 8) {
        set_last_error("paging did not converge within the pass limit");
        return TVK_ERR_OOM;
      }
    }
    for (int k = 0; k < cnt; ++k) {
      const uint32_t i = first + k;
      if (stats) {
        stats[i].evaluated_samples = batch_slot_.counters_h[kQueues + kJobCounters * k];
        stats[i].nominal_samples = batch_slot_.counters_h[kQueues + kJobCounters * k + 1];
        stats[i].gathered_samples = batch_slot_.counters_h[kQueues + kJobCounters * k + 2];
        stats[i].lod = (uint32_t)lod;
        stats[i].reserved = 0;
      }
      if (views[i].rgba)
        TVK_HIP(hipMemcpyAsync(views[i].rgba, (uint8_t*)batch_fb_.p + batch_fb_offset_[i],
                               (size_t)views[i].width * views[i].height * 4, hipMemcpyDeviceToHost,
                               stream_));
    }
    TVK_HIP(hipStreamSynchronize(stream_));
  }
  if (kernel_ms) *kernel_ms = kms;
  return TVK_OK;
}

// Would upload_tf / prepare_cells need a new device buffer before a launch
// of these views in `mode`?  (They only copy and launch otherwise.)
bool HipRenderer::launch_ready_without_alloc(const tvk_view* views, uint32_t n, int mode) const {
  for (const TfSet& t : tfs_) {
    if (t.dirty1d) {
      const size_t m = t.tf1d.size() / 4;
      if (t.tf1d_dev.bytes < m * 16 || t.prefix1d_dev.bytes < (m + 1) * 4) return false;
    }
    if (t.dirty2d) {
      const size_t W = t.w2d, H = t.h2d;
      if (t.tf2d_dev.bytes < W * H * 16 || t.prefix2d_dev.bytes < (W + 1) * 4) return false;
    }
  }
  if (skip_ && total_cells_ > 0)
    for (uint32_t i = 0; i < n; ++i)
      if (tfs_[views[i].tf_slot].cell_empty[mode].bytes < (size_t)total_cells_) return false;
  return true;
}

int HipRenderer::RenderBatchNoAlloc(const tvk_view* views, uint32_t n, void* const* dev_rgba,
                                    tvk_view_stats* stats) {
  if (!initialized_) return state_fail("render before initialize");
  if (!has_volume_) return state_fail("render before a dataset is loaded");
  if (n > (uint32_t)kMaxJobs)
    return arg_fail("at most " + std::to_string(kMaxJobs) +
                    " views per allocation-free launch (its device scratch holds one job table)");
  if (n == 0) return TVK_OK;
  if (!views || !dev_rgba) return arg_fail("null view or framebuffer array");
  if (paged_) return state_fail("the allocation-free launch renders in-core volumes (paging takes pool memory)");
  if (regions_active()) return state_fail("per-region LOD is not available allocation-free");
  // fixed-size arrays: no heap memory either
  CameraState cams[kMaxJobs];
  RaycastParams jobs[kMaxJobs];
  for (uint32_t i = 0; i < n; ++i) {
    if (int e = view_camera(views[i], i, &cams[i])) return e;
    if (int e = check_tf((int)views[i].tf_slot)) return e;
    if (!dev_rgba[i]) return arg_fail("null device framebuffer for view " + std::to_string(i));
  }
  if (!launch_ready_without_alloc(views, n, mode_index()))
    return state_fail("a TF slot of these views is not on the device for this render mode yet: render it "
                      "once with tvk_render_batch or tvk_paint first");
  TVK_HIP(hipSetDevice(device_));
  if (int e = upload_tf()) return e;
  for (uint32_t i = 0; i < n; ++i)
    if (int e = prepare_cells((int)views[i].tf_slot, mode_index())) return e;
  JobTable* dev_table = nullptr;
  unsigned long long* dev_counters = nullptr;
  TVK_HIP(noalloc_scratch(device_, &dev_table, &dev_counters));
  const int lo = std::min<int>((int)lod_min_, n_lod_ - 1);
  bool same_tf = true;
  for (uint32_t k = 1; k < n; ++k) same_tf = same_tf && views[k].tf_slot == views[0].tf_slot;
  const LaunchInfo li = launch_info(same_tf ? (int)(tfs_[views[0].tf_slot].tf1d.size() / 4) : 0);
  for (uint32_t i = 0; i < n; ++i)
    build_params(&jobs[i], lo, cams[i], views[i].width, views[i].height, (int)views[i].tf_slot,
                 (uint32_t*)dev_rgba[i]);
  // the pinned table of the batch slot may still feed its last upload
  if (batch_slot_.used) TVK_HIP(hipEventSynchronize(batch_slot_.ev[3]));
  batch_slot_.used = true;
  if (int e = launch_job_table_at(batch_slot_.table_h, dev_table, dev_counters, stats ? batch_slot_.counters_h : nullptr,
                                  batch_slot_.ev, jobs, (int)n, li, tile_group_, xcd_split_, stream_))
    return e;
  if (stats) {
    TVK_HIP(hipEventSynchronize(batch_slot_.ev[3]));
    for (uint32_t k = 0; k < n; ++k) {
      stats[k].evaluated_samples = batch_slot_.counters_h[kQueues + kJobCounters * k];
      stats[k].nominal_samples = batch_slot_.counters_h[kQueues + kJobCounters * k + 1];
      stats[k].gathered_samples = batch_slot_.counters_h[kQueues + kJobCounters * k + 2];
      stats[k].lod = (uint32_t)lo;
      stats[k].reserved = 0;
    }
  }
  return TVK_OK;
}

int HipRenderer::SetLaunchGraphs(bool on) {
  use_graphs_ = on;
  return TVK_OK;
}

int HipRenderer::view_camera(const tvk_view& v, uint32_t i, CameraState* c) const {
  if (v.width == 0 || v.height == 0 || v.width > 16384 || v.height > 16384)
    return arg_fail("view " + std::to_string(i) + ": image size out of range");
  if (!finite3(v.view_pos) || !finite3(v.view_dir) || !finite3(v.up_dir))
    return arg_fail("view " + std::to_string(i) + ": camera must be finite");
  const float* d = v.view_dir;
  if (d[0] == 0.f && d[1] == 0.f && d[2] == 0.f) return arg_fail("view direction must be non-zero");
  if (v.ortho ? !(v.ortho_half_height > 0.f) : !(v.fov_deg > 0.f && v.fov_deg < 180.f))
    return arg_fail("view " + std::to_string(i) + ": fov / ortho extent out of range");
  for (int k = 0; k < 3; ++k) {
    c->pos[k] = v.view_pos[k];
    c->dir[k] = v.view_dir[k];
    c->up[k] = v.up_dir[k];
  }
  c->fov = v.fov_deg;
  c->ortho = v.ortho != 0;
  c->ortho_hh = v.ortho_half_height;
  return TVK_OK;
}

int HipRenderer::ServePrepare(uint32_t tf_slot, int mode) {
  if (!initialized_ || !has_volume_) return state_fail("serving needs an initialized renderer with a volume");
  if (paged_) return state_fail("the server renders in-core volumes only");
  if (mode != TVK_RM_1DTRANS && mode != TVK_RM_2DTRANS) return arg_fail("served render mode must be 1D or 2D TF");
  const int saved = render_mode_;
  render_mode_ = mode;
  int e = check_tf((int)tf_slot);
  render_mode_ = saved;
  if (e) return e;
  TVK_HIP(hipSetDevice(device_));
  if (int e2 = upload_tf()) return e2;
  if (int e2 = prepare_cells((int)tf_slot, mode == TVK_RM_2DTRANS ? 1 : 0)) return e2;
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

int HipRenderer::ServeCheck(const tvk_view& v, int mode) {
  if (!initialized_ || !has_volume_) return state_fail("serving needs an initialized renderer with a volume");
  if (mode != TVK_RM_1DTRANS && mode != TVK_RM_2DTRANS) return arg_fail("served render mode must be 1D or 2D TF");
  CameraState cam;
  if (int e = view_camera(v, 0, &cam)) return e;
  const int saved = render_mode_;
  render_mode_ = mode;
  const int e = check_tf((int)v.tf_slot);
  render_mode_ = saved;
  return e;
}

int HipRenderer::ServeBatch(const tvk_view* views, uint32_t n, int mode, tvk_view_stats* stats) {
  const int saved = render_mode_;
  render_mode_ = mode;
  float kms = 0.f;
  const int e = RenderBatch(views, n, stats, &kms);
  render_mode_ = saved;
  return e;
}

uint64_t HipRenderer::PagedWorkingSetBytes() const {
  if (!paged_) return 0;
  const int lo = std::min<int>((int)lod_min_, n_lod_ - 1);
  const LodInfo& L = lods_[lo];
  const uint64_t bricks = (uint64_t)L.grid[0] * L.grid[1] * L.grid[2];
  return std::min<uint64_t>(bricks, pool_.n_slots) * pool_.slot_bytes;
}

int HipRenderer::ServeJob(const tvk_view& v, int mode, uint32_t* fb, RaycastParams* out,
                          LaunchInfo* li) {
  CameraState cam;
  if (int e = view_camera(v, 0, &cam)) return e;
  const int saved = render_mode_;
  render_mode_ = mode;
  const int lo = std::min<int>((int)lod_min_, n_lod_ - 1);
  build_params(out, lo, cam, v.width, v.height, (int)v.tf_slot, fb);
  *li = launch_info(0);
  render_mode_ = saved;
  return TVK_OK;
}

int HipRenderer::BatchFramebuffer(uint32_t index, void** dptr) {
  if (index >= batch_fb_offset_.size()) return arg_fail("no such view in the last batch");
  if (!dptr) return arg_fail("null pointer");
  *dptr = (uint8_t*)batch_fb_.p + batch_fb_offset_[index];
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
  // a paged pool may need a paging pass per LOD on top of the subframes;
  // a capture finishes the frame without the governor's LOD delay
  capturing_ = true;
  for (int guard = 0; guard < 4 * kMaxLod + 8 && !frame_complete_ && rc == TVK_OK; ++guard) rc = Paint();
  capturing_ = false;
  if (rc == TVK_OK && !frame_complete_) {
    set_last_error("the frame did not complete within the paging pass limit");
    rc = TVK_ERR_OOM;
  }
  if (rc == TVK_OK) rc = resolve_check();
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
  if (int e = resolve_check()) return e;
  if (stats_pending_ && stats_slot_) {
    TVK_HIP(hipSetDevice(device_));
    LaunchSlot& sl = *stats_slot_;
    TVK_HIP(hipEventSynchronize(sl.ev[3]));
    float k = 0.f, all = 0.f;
    TVK_HIP(hipEventElapsedTime(&k, sl.ev[1], sl.ev[2]));
    TVK_HIP(hipEventElapsedTime(&all, sl.ev[0], sl.ev[3]));
    stats_.kernel_ms = k;
    stats_.paint_ms = all;
    stats_.evaluated_samples = sl.counters_h[kQueues];
    stats_.nominal_samples = sl.counters_h[kQueues + 1];
    stats_pending_ = false;
  }
  *out = stats_;
  return TVK_OK;
}

int HipRenderer::DeviceFramebuffer(void** dptr, void** stream) {
  if (!initialized_ || !framebuffer_.p) return state_fail("no framebuffer (initialize + resize first)");
  if (int e = resolve_check()) return e;  // a fallback render is queued before the caller reads
  if (dptr) *dptr = framebuffer_.p;
  if (stream) *stream = (void*)stream_;
  return TVK_OK;
}

int HipRenderer::Synchronize() {
  if (!initialized_) return state_fail("not initialized");
  if (int e = resolve_check()) return e;
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

// ----------------------------------------------------------------- memory
int HipRenderer::SetMaxUsableGPUMem(uint64_t bytes) {
  if (bytes == 0) return arg_fail("GPU memory budget must be > 0");
  if (!initialized_) return state_fail("the memory budget belongs to the device: initialize first");
  MemMan::ForDevice(device_).SetGPUBudget(bytes);
  return TVK_OK;
}

int HipRenderer::SetMaxUsableCPUMem(uint64_t bytes) {
  if (bytes == 0) return arg_fail("CPU memory budget must be > 0");
  if (!initialized_) return state_fail("the memory budget belongs to the device: initialize first");
  MemMan::ForDevice(device_).SetCPUBudget(bytes);
  return TVK_OK;
}

int HipRenderer::SetBrickSourceThreads(uint32_t n) {
  if (n < 1 || n > 32) return arg_fail("brick source threads must be in [1, 32]");
  source_threads_ = (int)n;
  return TVK_OK;
}

int HipRenderer::MemStats(tvk_mem_stats* out) {
  if (!out) return arg_fail("null stats pointer");
  *out = tvk_mem_stats{};
  out->gpu_mem = gpu_total_;
  out->gpu_budget = initialized_ ? MemMan::ForDevice(device_).GPUBudget() : 0;
  uint64_t a = framebuffer_.bytes + batch_fb_.bytes + pool_.slots.bytes + pool_.page_table.bytes + pool_.brick_range.bytes + scratch_.bytes +
               cell_range_.bytes;
  for (auto& l : levels_) a += l.bytes;
  for (const TfSet& t : tfs_) {
    a += t.tf1d_dev.bytes + t.prefix1d_dev.bytes + t.tf2d_dev.bytes + t.prefix2d_dev.bytes;
    for (const DevBuf& f : t.cell_empty) a += f.bytes;
  }
  for (const LaunchSlot* sl : {&paint_slots_[0], &paint_slots_[1], &batch_slot_})
    a += sl->table.bytes + sl->counters.bytes;
  a += requests_dev_.bytes + touch_dev_.bytes + lowres_fb_.bytes + region_out_dev_.bytes;
  for (const DevBuf& d : upload_stage_) a += d.bytes;
  if (pool_client_ >= 0) a += MemMan::ForDevice(device_).ClientSlots(pool_client_) * pool_.slot_bytes;
  out->allocated_gpu_mem = a;
  out->pool_slots = pool_.n_slots;
  out->resident_bricks = pool_.resident;
  out->slot_bytes = pool_.slot_bytes;
  out->bricks_uploaded = pool_.uploaded;
  out->bricks_evicted = pool_.evicted;
  return TVK_OK;
}

}  // namespace tvk
