// paging.cpp -- out-of-core volumes on the device-wide brick pool (the role
// of Tuvok's GLGridLeaper paging over GPUMemMan, ImageVis3D/UI/
// RenderWindowGL.cpp:153-163, with the memory budgets of
// ImageVis3D_Settings.cpp:556-559).
//
// A paged volume's bricks start non-resident (page table 0).  A raycast
// that needs a non-resident, non-empty brick writes requests[brick] = 1 and
// samples nothing there; every brick it does sample gets touch[brick] =
// frame.  After the launch the host stamps the sampled bricks' slots in the
// device pool (LRU), reads the requests, takes slots from the pool (MemMan:
// free slots, new chunks within the GPU budget, else the least recently
// sampled slots of any volume -- never this frame's own), and streams the
// bricks in: the brick source (or the host brick cache) fills pinned staging
// buffers on worker threads, one round of bricks while the previous round's
// copies run; each brick is copied to a device staging buffer on the copy
// stream, rearranged into its slot's blocked layout and its empty-space cell
// ranges refreshed there, and the render stream waits on the copy stream's
// event before the page table (now holding the new slots) goes up and the
// next pass runs.  When a view needs more bricks of a LOD than the volume
// may hold, the frame falls back to the next coarser LOD.  Empty bricks (by
// the per-brick ranges the source provides) are skipped without ever being
// requested.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <numeric>

#include "errors.hpp"
#include "host_volume.hpp"
#include "kernels.hpp"
#include "memman.hpp"
#include "renderer.hpp"

namespace tvk {

#define TVK_HIP(call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      set_last_error(std::string(#call) + ": " + hipGetErrorString(e_));  \
      return e_ == hipErrorOutOfMemory ? TVK_ERR_OOM : TVK_ERR_HIP;       \
    }                                                                     \
  } while (0)

namespace {
constexpr uint64_t kStagingTarget = 128ull << 20;  // pinned bytes per staging round
}  // namespace

int HipRenderer::SetPagedVolume(const uint64_t dims[3], uint32_t bits, uint32_t brick,
                                uint32_t apron, uint32_t n_lod, BrickSourceFn fn, void* user,
                                const float* ranges, uint64_t pool_bytes) {
  if (!initialized_) {
    set_last_error("setPagedVolume before initialize");
    return TVK_ERR_STATE;
  }
  if (!fn) {
    set_last_error("null brick source");
    return TVK_ERR_ARG;
  }
  LevelPlan plan;
  if (int e = plan_levels(dims, bits, brick, apron, n_lod, &plan)) return e;
  const uint64_t quota = pool_bytes / plan.slot_bytes;
  if (quota < 1) {
    set_last_error("brick pool smaller than one brick slot");
    return TVK_ERR_OOM;
  }
  const int total = plan.total_bricks;
  const uint64_t S = (uint64_t)plan.slot_edge;
  const uint64_t brick_bytes = S * S * S * (bits / 8);  // one brick as the source gives it
  TVK_HIP(hipSetDevice(device_));
  TVK_HIP(hipStreamSynchronize(stream_));
  release_volume();

  const uint32_t maxval = (1u << bits) - 1;
  pool_.brick_range_h.resize((size_t)total * 2);
  for (int i = 0; i < total; ++i) {
    pool_.brick_range_h[2 * i] = ranges ? ranges[2 * i] : 0.f;
    pool_.brick_range_h[2 * i + 1] = ranges ? ranges[2 * i + 1] : (float)maxval;
  }
  // cells start with their brick's range; refined when the brick arrives
  cell_init_h_.assign((size_t)plan.total_cells, 0u);
  const int per = (int)brick / plan.cell;
  for (int l = 0; l < plan.n_lod; ++l) {
    const LodInfo& L = plan.lods[l];
    for (int cz = 0; cz < L.cells[2]; ++cz)
      for (int cy = 0; cy < L.cells[1]; ++cy)
        for (int cx = 0; cx < L.cells[0]; ++cx) {
          const int b = L.brick_base + ((cz / per) * L.grid[1] + cy / per) * L.grid[0] + cx / per;
          const uint32_t lo = (uint32_t)pool_.brick_range_h[2 * b];
          const uint32_t hi = (uint32_t)pool_.brick_range_h[2 * b + 1];
          cell_init_h_[L.cell_base + (cz * L.cells[1] + cy) * L.cells[0] + cx] = lo | (hi << 16);
        }
  }

  TVK_HIP(pool_.page_table.alloc((size_t)total * sizeof(uint64_t)));
  TVK_HIP(pool_.brick_range.alloc((size_t)total * 2 * sizeof(float)));
  TVK_HIP(cell_range_.alloc((size_t)plan.total_cells * sizeof(uint32_t)));
  const size_t req_words = ((size_t)total + 1) / 2;  // cleared as u64 words
  TVK_HIP(requests_dev_.alloc(req_words * 8));
  TVK_HIP(touch_dev_.alloc((size_t)total * sizeof(uint32_t)));
  pool_.n_slots = quota;
  pool_.slot_bytes = plan.slot_bytes;
  pool_bytes_req_ = pool_bytes;
  pool_.resident = 0;
  pool_.page_table_h.assign((size_t)total, 0);
  pool_.brick_slot.assign((size_t)total, -1);
  TVK_HIP(hipMemcpy(pool_.page_table.p, pool_.page_table_h.data(), (size_t)total * 8,
                    hipMemcpyHostToDevice));
  TVK_HIP(hipMemcpy(pool_.brick_range.p, pool_.brick_range_h.data(), (size_t)total * 8,
                    hipMemcpyHostToDevice));
  TVK_HIP(hipMemcpy(cell_range_.p, cell_init_h_.data(), cell_init_h_.size() * 4,
                    hipMemcpyHostToDevice));
  TVK_HIP(hipMemset(touch_dev_.p, 0, touch_dev_.bytes));
  if (requests_h_count_ < (uint64_t)total) {
    if (requests_h_) (void)hipHostFree(requests_h_);
    requests_h_ = nullptr;
    TVK_HIP(hipHostMalloc((void**)&requests_h_, (size_t)total * 8, hipHostMallocDefault));
    requests_h_count_ = (uint64_t)total;
  }
  // staging: two rounds of pinned buffers, one round of device buffers
  MemMan& mm = MemMan::ForDevice(device_);
  const int round = (int)std::max<uint64_t>(2, std::min<uint64_t>(32, kStagingTarget / brick_bytes));
  const uint64_t host_bytes = 2ull * round * brick_bytes;
  if (!mm.ReserveCPU(host_bytes)) {
    set_last_error("the CPU memory budget cannot hold the brick staging buffers");
    return TVK_ERR_OOM;
  }
  staging_host_bytes_ = host_bytes;
  staging_.assign((size_t)2 * round, nullptr);
  for (void*& b : staging_) TVK_HIP(hipHostMalloc(&b, brick_bytes, hipHostMallocDefault));
  upload_stage_.clear();
  upload_stage_.resize((size_t)round);
  for (DevBuf& d : upload_stage_) TVK_HIP(d.alloc(brick_bytes));
  staging_round_ = round;
  for (hipEvent_t& ev : staging_ev_)
    if (!ev) TVK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  if (!upload_done_) TVK_HIP(hipEventCreateWithFlags(&upload_done_, hipEventDisableTiming));

  bits_ = bits;
  for (int i = 0; i < 3; ++i) dims0_[i] = dims[i];
  n_lod_ = plan.n_lod;
  brick_ = (int)brick;
  apron_ = (int)apron;
  slot_edge_ = plan.slot_edge;
  cell_ = plan.cell;
  total_cells_ = plan.total_cells;
  ++cells_gen_;
  region_lod_h_.clear();
  for (int l = 0; l < plan.n_lod; ++l) lods_[l] = plan.lods[l];
  if (int e = upload_lods()) return e;
  total_bricks_ = total;
  lod_min_ = 0;
  lod_max_ = (uint32_t)(plan.n_lod - 1);
  lod_floor_ = 0;
  source_fn_ = fn;
  source_user_ = user;
  // bricks of a flat host volume are cheap to cut again; other sources
  // (files, decompression, callbacks) keep theirs in the host cache
  cache_bricks_ = fn != &HipRenderer::flat_source;
  pool_client_ = mm.Register(this, plan.slot_bytes, quota);
  paged_ = true;
  has_volume_ = true;
  invalidate();
  return TVK_OK;
}

int HipRenderer::flat_source(void* user, uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz,
                             void* slot, uint64_t slot_bytes) {
  HipRenderer* r = static_cast<HipRenderer*>(user);
  if ((int)lod >= r->n_lod_) return -1;
  const LodInfo& L = r->lods_[lod];
  const void* level = lod == 0 ? r->host_level0_ : r->host_levels_[lod - 1].data();
  const int64_t d[3] = {L.dims[0], L.dims[1], L.dims[2]};
  (void)slot_bytes;
  host_extract_brick(level, (int)r->bits_ / 8, d, r->brick_, r->apron_, (int)bx, (int)by, (int)bz,
                     slot);
  return 0;
}

int HipRenderer::SetPagedFlatVolume(const void* voxels, const uint64_t dims[3], uint32_t bits,
                                    uint32_t brick, uint32_t apron, uint32_t n_lod,
                                    uint64_t pool_bytes) {
  if (!voxels) {
    set_last_error("null voxel pointer");
    return TVK_ERR_ARG;
  }
  LevelPlan plan;
  if (int e = plan_levels(dims, bits, brick, apron, n_lod, &plan)) return e;
  const int bpv = (int)bits / 8;
  uint64_t coarse_bytes = 0;
  for (int l = 1; l < plan.n_lod; ++l) coarse_bytes += plan.dims[l][0] * plan.dims[l][1] * plan.dims[l][2] * bpv;
  if (!initialized_) {
    set_last_error("setPagedFlatVolume before initialize");
    return TVK_ERR_STATE;
  }
  MemMan& mm = MemMan::ForDevice(device_);
  if (!mm.ReserveCPU(coarse_bytes)) {
    set_last_error("the CPU memory budget cannot hold the volume's coarser levels (" +
                   std::to_string(coarse_bytes) + " bytes)");
    return TVK_ERR_OOM;
  }
  // coarser levels and per-brick ranges on host threads
  std::vector<std::vector<uint8_t>> levels((size_t)plan.n_lod - 1);
  std::vector<float> ranges((size_t)plan.total_bricks * 2);
  for (int l = 0; l < plan.n_lod; ++l) {
    const int64_t d[3] = {(int64_t)plan.dims[l][0], (int64_t)plan.dims[l][1], (int64_t)plan.dims[l][2]};
    const void* src = l == 0 ? voxels : levels[l - 1].data();
    if (l + 1 < plan.n_lod) {
      levels[l].resize(plan.dims[l + 1][0] * plan.dims[l + 1][1] * plan.dims[l + 1][2] * bpv);
      host_downsample(src, bpv, d, levels[l].data());
    }
    host_brick_ranges(src, bpv, d, (int)brick, (int)apron, plan.lods[l].grid,
                      ranges.data() + 2 * plan.lods[l].brick_base);
  }
  if (int e = SetPagedVolume(dims, bits, brick, apron, (uint32_t)plan.n_lod, &HipRenderer::flat_source,
                             this, ranges.data(), pool_bytes)) {
    mm.ReleaseCPU(coarse_bytes);
    return e;
  }
  host_level0_ = voxels;
  host_levels_ = std::move(levels);
  host_levels_bytes_ = coarse_bytes;
  return TVK_OK;
}

void HipRenderer::PoolEvict(int gid) {
  std::lock_guard<std::mutex> g(pool_mu_);
  // a queued raycast may still read the slot
  (void)hipStreamSynchronize(stream_);
  if (gid < 0 || (size_t)gid >= pool_.page_table_h.size()) return;
  if (pool_.page_table_h[gid]) {
    pool_.page_table_h[gid] = 0;
    --pool_.resident;
    ++pool_.evicted;
  }
  pool_.brick_slot[gid] = -1;
  pt_dirty_ = true;
}

int HipRenderer::sync_page_table_locked() {
  if (!pt_dirty_ || !paged_) return TVK_OK;
  TVK_HIP(hipMemcpyAsync(pool_.page_table.p, pool_.page_table_h.data(), pool_.page_table_h.size() * 8,
                         hipMemcpyHostToDevice, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));  // the host copy may change once the lock is released
  pt_dirty_ = false;
  return TVK_OK;
}

int HipRenderer::touch_sampled() {
  std::vector<uint32_t> touch((size_t)total_bricks_);
  TVK_HIP(hipMemcpy(touch.data(), touch_dev_.p, touch.size() * 4, hipMemcpyDeviceToHost));
  std::vector<int64_t> used;
  {
    std::lock_guard<std::mutex> g(pool_mu_);
    for (int b = 0; b < total_bricks_; ++b)
      if (touch[b] == frame_ && pool_.brick_slot[b] >= 0) used.push_back(pool_.brick_slot[b]);
  }
  if (!used.empty()) MemMan::ForDevice(device_).Touch(pool_client_, used.data(), used.size());
  return TVK_OK;
}

int HipRenderer::upload_bricks(const int* gids, const int* lods, void* const* addrs, size_t n) {
  MemMan& mm = MemMan::ForDevice(device_);
  const int bpv = (int)bits_ / 8;
  const uint64_t bb = (uint64_t)slot_edge_ * slot_edge_ * slot_edge_ * bpv;
  const size_t R = (size_t)staging_round_;
  const auto t0 = std::chrono::steady_clock::now();
  // per-cell folding words of the cell-range kernels (copy-stream order)
  const int per = brick_ / cell_;
  if (slot_cell_tmp_per_ != per) {
    const size_t need = 2ull * per * per * per * sizeof(uint32_t);
    if (slot_cell_tmp_.bytes < need) {
      TVK_HIP(hipStreamSynchronize(copy_stream_));  // earlier bricks' kernels may hold it
      TVK_HIP(slot_cell_tmp_.alloc(need));
    }
    TVK_HIP(launch_slot_cell_tmp_init(slot_cell_tmp_.as<uint32_t>(), per, copy_stream_));
    slot_cell_tmp_per_ = per;
  }
  std::atomic<int> failed{-1};
  for (size_t first = 0, r = 0; first < n; first += R, ++r) {
    const size_t cnt = std::min(R, n - first);
    const int set = (int)(r & 1);
    // this staging set's copies from two rounds ago are done
    if (r >= 2) TVK_HIP(hipEventSynchronize(staging_ev_[set]));
    mm.ParallelFor(cnt, source_threads_, [&](size_t j) {
      const size_t i = first + j;
      const int gid = gids[i], lod = lods[i];
      void* buf = staging_[(size_t)set * R + j];
      if (cache_bricks_ && mm.CacheGet(pool_client_, gid, buf, bb)) return;
      const LodInfo& L = lods_[lod];
      const int local = gid - L.brick_base;
      const int bx = local % L.grid[0], by = (local / L.grid[0]) % L.grid[1],
                bz = local / (L.grid[0] * L.grid[1]);
      if (source_fn_(source_user_, (uint32_t)lod, (uint32_t)bx, (uint32_t)by, (uint32_t)bz, buf, bb) != 0) {
        failed = gid;
        return;
      }
      if (cache_bricks_) mm.CachePut(pool_client_, gid, buf, bb);
    });
    if (failed >= 0) {
      set_last_error("brick source failed for brick " + std::to_string(failed.load()));
      return TVK_ERR_IO;
    }
    // x-fastest brick -> device staging -> the slot's blocked layout, and
    // the brick's exact empty-space cell ranges, in copy-stream order
    for (size_t j = 0; j < cnt; ++j) {
      const size_t i = first + j;
      const LodInfo& L = lods_[lods[i]];
      const int local = gids[i] - L.brick_base;
      const int b[3] = {local % L.grid[0], (local / L.grid[0]) % L.grid[1], local / (L.grid[0] * L.grid[1])};
      TVK_HIP(hipMemcpyAsync(upload_stage_[j].p, staging_[(size_t)set * R + j], bb, hipMemcpyHostToDevice,
                             copy_stream_));
      TVK_HIP(launch_linear_to_slot(upload_stage_[j].p, addrs[i], bpv, slot_edge_, copy_stream_));
      TVK_HIP(launch_slot_cell_ranges(addrs[i], bpv, slot_edge_, apron_, brick_, cell_, b, L.cells,
                                      cell_range_.as<uint32_t>() + L.cell_base, slot_cell_tmp_.as<uint32_t>(),
                                      copy_stream_));
    }
    TVK_HIP(hipEventRecord(staging_ev_[set], copy_stream_));
  }
  // the render stream runs nothing that reads these slots before they land
  TVK_HIP(hipEventRecord(upload_done_, copy_stream_));
  TVK_HIP(hipStreamWaitEvent(stream_, upload_done_, 0));
  paging_bytes_ += (uint64_t)n * bb;
  paging_ms_ += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  ++cells_gen_;  // the cells of these bricks now hold their exact ranges
  return TVK_OK;
}

int HipRenderer::service_requests(int lod, bool* complete, bool* overflow) {
  return service_requests(lod, lod, complete, overflow);
}

int HipRenderer::service_requests(int lod_lo, int lod_hi, bool* complete, bool* overflow) {
  *complete = true;
  *overflow = false;
  if (int e = touch_sampled()) return e;
  std::vector<int> want, want_lod;
  for (int l = lod_lo; l <= lod_hi; ++l) {
    const LodInfo& L = lods_[l];
    const int nb = L.grid[0] * L.grid[1] * L.grid[2];
    for (int b = 0; b < nb; ++b)
      if (requests_h_[L.brick_base + b]) {
        want.push_back(L.brick_base + b);
        want_lod.push_back(l);
      }
  }
  if (want.empty()) return TVK_OK;
  *complete = false;
  MemMan& mm = MemMan::ForDevice(device_);
  std::vector<int64_t> slots(want.size());
  std::vector<void*> addrs(want.size());
  // may take back this volume's least recent bricks (PoolEvict) or others'
  const size_t got = mm.Acquire(pool_client_, want.data(), want.size(), slots.data(), addrs.data());
  if (got < want.size()) *overflow = true;
  int rc = upload_bricks(want.data(), want_lod.data(), addrs.data(), got);
  if (rc == TVK_OK) {
    std::lock_guard<std::mutex> g(pool_mu_);
    for (size_t i = 0; i < got; ++i) {
      pool_.page_table_h[want[i]] = (uint64_t)(uintptr_t)addrs[i];
      pool_.brick_slot[want[i]] = slots[i];
    }
    pool_.resident += got;
    pool_.uploaded += got;
    pt_dirty_ = true;
    // after the copy stream's event (upload_bricks made stream_ wait on it)
    rc = sync_page_table_locked();
  } else {
    // queued copies may still write the slots: let them land first
    (void)hipStreamSynchronize(copy_stream_);
    for (size_t i = 0; i < got; ++i) mm.Free(pool_client_, slots[i]);
  }
  if (rc == TVK_OK) mm.Commit(pool_client_, slots.data(), got);
  return rc;
}

}  // namespace tvk
