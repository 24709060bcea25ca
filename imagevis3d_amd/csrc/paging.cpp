// paging.cpp -- out-of-core volumes on a fixed HBM brick pool (the role of
// Tuvok's GLGridLeaper paging over GPUMemMan, ImageVis3D/UI/RenderWindowGL.cpp
// :153-163 and ImageVis3D_Settings.cpp:557-559 for the memory budget).
//
// The pool holds `pool_bytes / slot_bytes` brick slots.  Bricks start
// non-resident (page table -1).  A raycast that needs a non-resident,
// non-empty brick writes requests[brick] = 1 and samples nothing there;
// every brick it does sample gets touch[brick] = frame.  After the launch
// the host reads the requests, fills bricks from the brick source into
// pinned staging buffers, copies them to the device and rearranges them
// (blocked slot layout) into free slots or into the least recently touched
// slots not used by this frame, refreshes those bricks'
// empty-space cell ranges on the device, and asks for another pass.  When
// a view needs more bricks of a LOD than the pool holds, the frame falls
// back to the next coarser LOD.  Empty bricks (by the per-brick ranges the
// source provides) are skipped without ever being requested.
#include <string.h>

#include <algorithm>
#include <numeric>

#include "errors.hpp"
#include "host_volume.hpp"
#include "kernels.hpp"
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
  const uint64_t n_slots = pool_bytes / plan.slot_bytes;
  if (n_slots < 1) {
    set_last_error("brick pool smaller than one brick slot");
    return TVK_ERR_OOM;
  }
  const int total = plan.total_bricks;
  const uint64_t S = (uint64_t)plan.slot_edge;
  const uint64_t brick_bytes = S * S * S * (bits / 8);  // one brick as the source gives it
  const uint64_t need = n_slots * plan.slot_bytes + brick_bytes + (uint64_t)total * 20 +
                        (uint64_t)plan.total_cells * 4 + (uint64_t)width_ * height_ * 4;
  if (need > gpu_budget_) {
    set_last_error("paged pool needs " + std::to_string(need) + " bytes of device memory, budget is " +
                   std::to_string(gpu_budget_));
    return TVK_ERR_OOM;
  }
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

  TVK_HIP(pool_.slots.alloc(n_slots * plan.slot_bytes));
  TVK_HIP(upload_stage_.alloc(brick_bytes));
  TVK_HIP(pool_.page_table.alloc((size_t)total * sizeof(int32_t)));
  TVK_HIP(pool_.brick_range.alloc((size_t)total * 2 * sizeof(float)));
  TVK_HIP(cell_range_.alloc((size_t)plan.total_cells * sizeof(uint32_t)));
  const size_t req_words = ((size_t)total + 1) / 2;  // cleared as u64 words
  TVK_HIP(requests_dev_.alloc(req_words * 8));
  TVK_HIP(touch_dev_.alloc((size_t)total * sizeof(uint32_t)));
  pool_.n_slots = n_slots;
  pool_.slot_bytes = plan.slot_bytes;
  pool_.resident = 0;
  pool_.page_table_h.assign((size_t)total, -1);
  pool_.slot_owner.assign((size_t)n_slots, -1);
  pool_.slot_touch.assign((size_t)n_slots, 0u);
  TVK_HIP(hipMemcpy(pool_.page_table.p, pool_.page_table_h.data(), (size_t)total * 4,
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
  for (int i = 0; i < kStaging; ++i) {
    if (staging_[i]) (void)hipHostFree(staging_[i]);
    staging_[i] = nullptr;
    TVK_HIP(hipHostMalloc(&staging_[i], plan.slot_bytes, hipHostMallocDefault));
    if (!staging_ev_[i]) TVK_HIP(hipEventCreate(&staging_ev_[i]));
    TVK_HIP(hipEventRecord(staging_ev_[i], stream_));
  }

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
                             this, ranges.data(), pool_bytes))
    return e;
  host_level0_ = voxels;
  host_levels_ = std::move(levels);
  return TVK_OK;
}

int HipRenderer::upload_brick(int gid, int lod, int bx, int by, int bz, int slot, int st) {
  TVK_HIP(hipEventSynchronize(staging_ev_[st]));  // the staging buffer's last copy is done
  if (source_fn_(source_user_, (uint32_t)lod, (uint32_t)bx, (uint32_t)by, (uint32_t)bz, staging_[st],
                 pool_.slot_bytes) != 0) {
    set_last_error("brick source failed for LOD " + std::to_string(lod) + " brick " + std::to_string(gid));
    return TVK_ERR_IO;
  }
  uint8_t* dst = (uint8_t*)pool_.slots.p + (uint64_t)slot * pool_.slot_bytes;
  // the source's x-fastest brick -> device staging -> the slot's blocked
  // layout (copies and kernels are ordered on the one stream)
  const uint64_t n = (uint64_t)slot_edge_ * slot_edge_ * slot_edge_ * (bits_ / 8);
  TVK_HIP(hipMemcpyAsync(upload_stage_.p, staging_[st], n, hipMemcpyHostToDevice, stream_));
  TVK_HIP(hipEventRecord(staging_ev_[st], stream_));
  TVK_HIP(launch_linear_to_slot(upload_stage_.p, dst, (int)bits_ / 8, slot_edge_, stream_));
  const LodInfo& L = lods_[lod];
  const int b[3] = {bx, by, bz};
  TVK_HIP(launch_slot_cell_ranges(dst, (int)bits_ / 8, slot_edge_, apron_, brick_, cell_, b, L.cells,
                                  cell_range_.as<uint32_t>() + L.cell_base, stream_));
  ++cells_gen_;  // the cells of this brick now hold their exact ranges
  return TVK_OK;
}

int HipRenderer::service_requests(int lod, bool* complete, bool* overflow) {
  return service_requests(lod, lod, complete, overflow);
}

int HipRenderer::service_requests(int lod_lo, int lod_hi, bool* complete, bool* overflow) {
  *complete = true;
  *overflow = false;
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
  // which resident bricks did this frame sample?
  std::vector<uint32_t> touch((size_t)total_bricks_);
  TVK_HIP(hipMemcpy(touch.data(), touch_dev_.p, touch.size() * 4, hipMemcpyDeviceToHost));
  std::vector<int> cand;  // free slots, then LRU slots not used by this frame
  std::vector<int> used;
  for (uint64_t s = 0; s < pool_.n_slots; ++s) {
    const int owner = pool_.slot_owner[s];
    if (owner < 0) cand.push_back((int)s);
    else {
      pool_.slot_touch[s] = std::max(pool_.slot_touch[s], touch[owner]);
      if (pool_.slot_touch[s] < frame_) used.push_back((int)s);
    }
  }
  std::stable_sort(used.begin(), used.end(),
                   [&](int a, int b) { return pool_.slot_touch[a] < pool_.slot_touch[b]; });
  cand.insert(cand.end(), used.begin(), used.end());
  if (want.size() > cand.size()) *overflow = true;
  const size_t n_up = std::min(want.size(), cand.size());
  for (size_t i = 0; i < n_up; ++i) {
    const int gid = want[i], s = cand[i], lod = want_lod[i];
    const LodInfo& L = lods_[lod];
    const int local = gid - L.brick_base;
    const int bx = local % L.grid[0], by = (local / L.grid[0]) % L.grid[1],
              bz = local / (L.grid[0] * L.grid[1]);
    if (pool_.slot_owner[s] >= 0) {
      pool_.page_table_h[pool_.slot_owner[s]] = -1;
      ++pool_.evicted;
      --pool_.resident;
    }
    if (int e = upload_brick(gid, lod, bx, by, bz, s, (int)(i % kStaging))) return e;
    pool_.page_table_h[gid] = s;
    pool_.slot_owner[s] = gid;
    pool_.slot_touch[s] = frame_;
    ++pool_.uploaded;
    ++pool_.resident;
  }
  TVK_HIP(hipMemcpyAsync(pool_.page_table.p, pool_.page_table_h.data(), pool_.page_table_h.size() * 4,
                         hipMemcpyHostToDevice, stream_));
  TVK_HIP(hipStreamSynchronize(stream_));
  return TVK_OK;
}

}  // namespace tvk
