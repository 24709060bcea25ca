// autotune.cpp -- HipRenderer::Autotune: coordinate descent over brick
// size, empty-space cell edge and launch shape, timed with the renderer's
// own batched launches (HIP events around each raycast launch).
#include <algorithm>
#include <vector>

#include "errors.hpp"
#include "renderer.hpp"

namespace tvk {

namespace {

struct Setting {
  uint32_t brick, cell, batch, tile_width, tile_group, wg_waves, stage;
};

}  // namespace

int HipRenderer::Autotune(const tvk_view* views, uint32_t n, uint32_t flags, tvk_tuning* out) {
  if (!has_volume_) {
    set_last_error("autotune before a volume is set");
    return TVK_ERR_STATE;
  }
  if (!views || n == 0) {
    set_last_error("autotune needs at least one view");
    return TVK_ERR_ARG;
  }
  // images stay on the device while tuning
  std::vector<tvk_view> v(views, views + n);
  for (auto& x : v) x.rgba = nullptr;
  int tried = 0;
  // median of three timed repeats after one warm-up (graph capture, paging)
  auto time_now = [&](float* ms) -> int {
    float t[3];
    if (int e = RenderBatch(v.data(), n, nullptr, &t[0])) return e;
    for (float& x : t)
      if (int e = RenderBatch(v.data(), n, nullptr, &x)) return e;
    std::sort(t, t + 3);
    *ms = t[1];
    ++tried;
    return TVK_OK;
  };
  Setting cur{(uint32_t)brick_, (uint32_t)cell_, (uint32_t)batch_, 1u << tile_shift_,
              (uint32_t)tile_group_, (uint32_t)wg_waves_, stage_ ? 1u : 0u};
  auto apply = [&](const Setting& s, bool rebrick) -> int {
    if (rebrick) {
      if (int e = Rebrick(s.brick, s.cell)) return e;
    }
    if (int e = SetKernelTuning(s.batch, s.tile_group, (uint32_t)xcd_split_)) return e;
    if (int e = SetWorkGroupWaves(s.wg_waves)) return e;
    if (int e = SetBrickStaging(s.stage != 0)) return e;
    return SetRayTileShape(s.tile_width);
  };
  float best = 0.f;
  if (int e = time_now(&best)) return e;
  const float baseline = best;

  auto try_values = [&](const std::vector<uint32_t>& values, uint32_t Setting::*field,
                        bool rebrick) -> int {
    for (uint32_t val : values) {
      if (cur.*field == val) continue;
      Setting s = cur;
      s.*field = val;
      if (field == &Setting::brick) {
        // keep the cell edge when it divides the new brick, else the brick
        if (s.brick % s.cell != 0 || s.cell > s.brick) s.cell = std::min<uint32_t>(s.brick, 16);
      }
      if (int e = apply(s, rebrick)) {
        // an unusable candidate (e.g. out of device memory): restore, go on
        if (int e2 = apply(cur, rebrick)) return e2;
        (void)e;
        continue;
      }
      float ms = 0.f;
      if (int e = time_now(&ms)) return e;
      if (ms < best) {
        best = ms;
        cur = s;
      }
    }
    return apply(cur, rebrick);
  };

  const uint64_t maxdim = std::max(dims0_[0], std::max(dims0_[1], dims0_[2]));
  // Launch shape first (cheap to try), then the bricking with that shape,
  // then the launch shape again for the chosen bricks: tuning the bricks
  // under the starting launch shape can settle on bricks that are only best
  // for that shape.
  // samples per gather batch: 2, 4 or 8 for the 1D march; for the gradient
  // marches (2D TF, or 1D TF with lighting) 1 or 2 per lane, or 4 = runs of
  // 4 samples per ray gathered across lanes; the other modes' batch is
  // fixed
  const int mode = job_mode();
  const bool grad = mode == 1 || (mode == 0 && lighting_);
  const std::vector<uint32_t> batches =
      grad ? std::vector<uint32_t>{1u, 2u, 4u}
           : (mode == 0 && !regions_active() ? std::vector<uint32_t>{2u, 4u, 8u} : std::vector<uint32_t>{});
  auto tune_launch = [&]() -> int {
    if (int e = try_values(batches, &Setting::batch, false)) return e;
    if (int e = try_values({4u, 8u, 16u}, &Setting::tile_width, false)) return e;
    if (int e = try_values({1u, 2u, 4u, 8u, 16u}, &Setting::tile_group, false)) return e;
    if (int e = try_values({4u, 8u}, &Setting::wg_waves, false)) return e;
    return try_values({0u, 1u}, &Setting::stage, false);
  };
  if (flags & TVK_TUNE_LAUNCH)
    if (int e = tune_launch()) return e;
  // a paged volume re-bricks from the caller's flat voxels (a brick
  // source fixes the brick size); each candidate is timed after a warm-up
  // batch has paged its bricks in
  const bool rebrickable = !paged_ || host_level0_ != nullptr;
  const Setting before = cur;
  std::vector<uint32_t> bricks;
  for (uint32_t b : {32u, 64u, 128u, 256u})
    if (b <= std::max<uint64_t>(32, maxdim)) bricks.push_back(b);
  if ((flags & TVK_TUNE_BRICK) && rebrickable)
    if (int e = try_values(bricks, &Setting::brick, true)) return e;
  // cells, then the launch shape again; a second round when the launch
  // shape moved (the best cell size depends on it: coordinate descent
  // from one start settled 5% away from the best on the paged 2D-TF batch)
  for (int round = 0; round < 2; ++round) {
    const Setting start = cur;
    if ((flags & TVK_TUNE_CELL) && rebrickable) {
      std::vector<uint32_t> cells;
      for (uint32_t c = 8; c <= cur.brick; c *= 2)
        if (cur.brick % c == 0) cells.push_back(c);
      if (std::find(cells.begin(), cells.end(), cur.brick) == cells.end()) cells.push_back(cur.brick);
      // each cell size with both work-group sizes: large cells (long
      // unskipped runs) and small ones favour different occupancies, and
      // tuning them one at a time settled up to 6% away from the best pair
      // on the paged 2D-TF batch
      const uint32_t wgs[2] = {cur.wg_waves, cur.wg_waves == 4 ? 8u : 4u};
      const Setting base = cur;
      uint32_t bricked = cur.cell;  // the cell size the volume is set up with
      for (uint32_t c : cells) {
        for (int k = 0; k < ((flags & TVK_TUNE_LAUNCH) ? 2 : 1); ++k) {
          Setting s = base;
          s.cell = c;
          s.wg_waves = wgs[k];
          if (c == base.cell && k == 0) continue;  // the current setting, already timed
          const bool rebrick = s.cell != bricked;
          bricked = s.cell;
          if (int e = apply(s, rebrick)) {
            if (int e2 = apply(cur, true)) return e2;
            bricked = cur.cell;
            (void)e;
            continue;
          }
          float ms = 0.f;
          if (int e = time_now(&ms)) return e;
          if (ms < best) {
            best = ms;
            cur = s;
          }
        }
      }
      if (int e = apply(cur, true)) return e;
    }
    if (!(flags & TVK_TUNE_LAUNCH)) break;
    if (round == 0 && cur.brick == before.brick && cur.cell == before.cell) break;
    if (int e = tune_launch()) return e;
    const bool moved = cur.batch != start.batch || cur.tile_width != start.tile_width ||
                       cur.tile_group != start.tile_group || cur.wg_waves != start.wg_waves ||
                       cur.stage != start.stage;
    if (!moved || !(flags & TVK_TUNE_CELL) || !rebrickable) break;
  }
  // the brick sizes once more with the tuned cells (cells as large as their
  // brick grow with it): the first pass timed each brick size with the
  // starting cells, which settled the paged 2D-TF batch on 128^3 bricks and
  // cells in one run (126 ms) and on 256^3 ones (121 ms) in another
  if ((flags & TVK_TUNE_BRICK) && (flags & TVK_TUNE_CELL) && rebrickable) {
    const Setting tuned = cur;
    Setting set = cur;  // the bricking the volume is set up with
    for (uint32_t b : bricks) {
      if (b == tuned.brick) continue;
      Setting s = tuned;
      s.brick = b;
      s.cell = tuned.cell == tuned.brick ? b : tuned.cell;
      if (s.cell > s.brick || s.brick % s.cell != 0) continue;
      set = s;
      if (int e = apply(s, true)) {
        if (int e2 = apply(cur, true)) return e2;
        set = cur;
        (void)e;
        continue;
      }
      float ms = 0.f;
      if (int e = time_now(&ms)) return e;
      if (ms < best) {
        best = ms;
        cur = s;
      }
    }
    if (set.brick != cur.brick || set.cell != cur.cell) {
      // set the chosen bricking up again, and page its bricks in (a paged
      // volume) with one batch, as every timed setting was
      if (int e = apply(cur, true)) return e;
      if (int e = RenderBatch(v.data(), n, nullptr, nullptr)) return e;
    }
  }
  if (out) {
    *out = tvk_tuning{};
    out->brick_size = (uint32_t)brick_;
    out->cell = (uint32_t)cell_;
    out->samples_per_batch = (uint32_t)batch_;
    out->tile_width = 1u << tile_shift_;
    out->tile_group = (uint32_t)tile_group_;
    out->xcd_split = (uint32_t)xcd_split_;
    out->wg_waves = (uint32_t)wg_waves_;
    out->brick_staging = stage_ ? 1u : 0u;
    out->kernel_ms = best;
    out->baseline_ms = baseline;
    out->candidates = (uint32_t)tried;
  }
  return TVK_OK;
}

}  // namespace tvk
