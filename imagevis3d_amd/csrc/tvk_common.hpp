// tvk_common.hpp -- plain structs shared by the host renderer and the HIP
// kernels of the MI355X Tuvok backend.
//
// Device data layout (HBM):
//   brick pool   one allocation of fixed slots; a slot holds one brick of
//                payload B^3 plus an apron of `apron` voxels on every side,
//                S^3 voxels (S = B+2a), bpv bytes per voxel, in the blocked
//                slot layout below.  Slots are 256-byte aligned.
//   page table   uint64 per brick of every LOD (LOD-major, then z, y, x):
//                device address of the brick's slot, or 0 when the brick
//                is not resident.  Slots live in the renderer's in-core pool
//                or in any chunk of the device-wide paged pool.
//   brick ranges float2 (min, max) per brick over its stored slot.
//   cell ranges  uint32 (min | max << 16) per empty-space cell of edge C
//                (C divides the brick size), over the voxels
//                [c*C - 1, (c+1)*C] that trilinear samples in the cell read.
//   cell flags   uint8 per cell and TF slot: 1 when the TF is transparent
//                over the cell's range (rebuilt when the TF, mode or the
//                cell ranges change).
//   TF           n_tf float4 (RGBA) + int32 prefix count of alpha > 0.
//   framebuffer  W*H uint32 RGBA8 (r in the low byte), row 0 at the top.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define TVK_HD __host__ __device__
#else
#define TVK_HD
#endif

namespace tvk {

// Blocked slot layout.  The slot's S^3 voxels are stored in 4x4x4 blocks,
// ceil(S/4) per axis, blocks x fastest, then y, then z.  Inside a block the
// 16 rows of x (y, then z within the block) are 5 voxels long: the row's 4
// voxels and the voxel after them in x (the first voxel of the next block's
// row; past the slot's last voxel, the last voxel again).  So the two
// x-neighbours (x, x+1) of any voxel are adjacent in memory: a trilinear
// sample is 4 two-voxel loads, and the 64 lanes of a wave, whose samples
// lie in a small patch of the volume, touch few cache lines per load (the
// x-fastest layout spreads the same patch over one line per row).  Costs
// 1.25x the brick's voxels.
constexpr int kSlotBlockElems = 80;  // 16 rows of 5

TVK_HD inline int slot_blocks(int S) { return (S + 3) >> 2; }
TVK_HD inline int64_t slot_block_elems(int S) {
  const int64_t n = slot_blocks(S);
  return n * n * n * kSlotBlockElems;
}

// Element offsets of slot voxel (x, y, z) = sx(x) + sy(y) + sz(z).
struct SlotGeo {
  int S;   // slot edge B + 2a
  int by;  // elements per block step in y (blocks per row x 80)
  int bz;  // elements per block step in z
};
TVK_HD inline SlotGeo slot_geo(int S) {
  const int n = slot_blocks(S);
  return SlotGeo{S, n * kSlotBlockElems, n * n * kSlotBlockElems};
}
TVK_HD inline int slot_ox(int x) { return (x >> 2) * kSlotBlockElems + (x & 3); }
TVK_HD inline int slot_oy(const SlotGeo& g, int y) { return (y >> 2) * g.by + (y & 3) * 5; }
TVK_HD inline int slot_oz(const SlotGeo& g, int z) { return (z >> 2) * g.bz + (z & 3) * 20; }
TVK_HD inline int64_t slot_offset(const SlotGeo& g, int x, int y, int z) {
  return (int64_t)slot_ox(x) + slot_oy(g, y) + slot_oz(g, z);
}

constexpr int kMaxLod = 16;
constexpr int kTile = 8;  // ray tile edge: one 64-lane wavefront = 8x8 rays

struct LodInfo {
  int32_t dims[3];      // voxels of this level
  int32_t grid[3];      // bricks of this level
  int32_t brick_base;   // index of the level's first brick in the page table
  float to_level[3];    // level texel coords per LOD-0 voxel (N_l / N_0)
  int32_t cells[3];     // empty-space cells of this level (edge = cell size)
  int32_t cell_base;    // index of the level's first cell in the cell-range grid
};

// One job of a raycast launch (an entry of the device job table).
struct RaycastParams {
  // Camera in LOD-0 voxel space.  Perspective: ray origin `eye`, direction
  // normalize(dir00 + x*du + y*dv) for pixel (x, y).  Orthographic: origin
  // eye + x*du + y*dv, direction odir.  Pixel-centre offsets are folded in.
  float eye[3];
  float dir00[3];
  float du[3];
  float dv[3];
  float odir[3];
  int32_t ortho;
  int32_t width, height;
  int32_t tile_shift;   // ray tile = (1 << tile_shift) x (64 >> tile_shift) rays
  int32_t tiles_x, tiles_y;
  // The tiles whose rays can reach the volume box: the rectangle
  // [rect_x0, rect_x0 + rect_w) x [rect_y0, rect_y0 + rect_h) of the tile
  // grid (the box's projected bounding rectangle plus a margin).  Only
  // these enter the raycast's tile queues; the background kernel writes
  // the job's other pixels.
  int32_t rect_x0, rect_y0, rect_w, rect_h;
  // Tile sampling: with tile_sample = k > 1 the job renders one tile of
  // each run of k consecutive rect tiles, at an offset hashed from the run
  // index and tile_phase; the other pixels keep what the framebuffer holds.
  int32_t tile_sample, tile_phase;
  float box[3];         // LOD-0 dims (the volume box is [0, box])
  float step;           // sample spacing in LOD-0 voxels (2^lod / rate)
  float nominal_rate;   // user rate: nominal samples per LOD-0 voxel
  float term_alpha;     // early ray termination threshold
  // Clip plane in LOD-0 voxel space: points with clip.xyz . p + clip.w > 0
  // are cut away (the ray interval is clipped to the kept half-space).
  float clip[4];
  int32_t clip_on;
  // Blend precision of the compositing target: 8 (unorm8), 16 (half) or 32
  // (float).  Below 32 the running composite (C, A) is rounded to it each
  // time the ray enters another brick of the level and at the end, as a
  // per-brick blend into a target of that precision would (1D / 2D TF).
  int32_t blend_bits;
  int32_t interp;       // 0 nearest, 1 trilinear
  int32_t batch;        // samples gathered per step of the march (1, 2, 4, 8)
  int32_t tile_group;   // consecutive ray tiles dealt to one XCD in turn
  int32_t xcd_split;    // 0: round-robin tile runs; rx: rx x (8/rx) frame regions
  int32_t compute_units;  // CUs of the device (sizes the persistent grid)
  int32_t skip;         // empty-space skipping on/off
  LodInfo level;        // the level being sampled
  int32_t brick;        // payload edge B
  int32_t apron;        // apron a
  int32_t slot_edge;    // S = B + 2a
  int32_t n_tf;
  float tf_scale;
  float bg_top[3], bg_bottom[3];
  float bg_alpha;
  // render mode: 0 = 1D TF, 1 = 2D TF, 2 = isosurface; lighting for 0/1
  int32_t mode;
  int32_t lighting;
  float grad_scale[3];  // gradient units: (N_l/N_0) / (2 (2^bits - 1)) per axis
  float ka[3], kd[3], ks[3], shininess;
  float light[3];       // unit light direction in voxel space
  const float* tf2d;    // tf2d_w x tf2d_h float4, value fastest
  int32_t tf2d_w, tf2d_h;
  float tf2d_row_scale; // (tf2d_h - 1) / max gradient
  float iso_value;      // raw data units
  float iso_color[3];
  // ClearView (mode 2, cv != 0): the iso_value surface is the context; in a
  // focus disc around cv_focus (pixel) it fades to its silhouettes and the
  // first cv_iso_value crossing behind it shows in cv_color.  Focus weight
  // w = clamp(1 - cv_size * r, 0, 1), r = distance from cv_focus in image
  // heights; context opacity (1 - w) + w clamp(cv_context_scale
  // (1 - |N.V|)^cv_border_scale, 0, 1).
  int32_t cv;
  float cv_iso_value;
  float cv_color[3];
  float cv_focus[2];
  float cv_size, cv_context_scale, cv_border_scale;
  const uint64_t* page_table;   // per brick: device address of its slot, 0 when not resident
  const float* brick_range;     // 2 floats per brick
  const uint32_t* cell_range;   // per cell: min | max << 16 (raw values)
  const uint8_t* cell_empty;    // per cell: 1 when this job's TF is transparent over it
  int32_t cell;                 // march unit edge in level voxels (divides brick)
  const float* tf;              // 4 floats per entry
  const int32_t* tf_prefix;     // n_tf + 1
  // paged pools: a ray that needs a non-resident brick sets requests[brick]
  // = 1; every brick a ray samples gets touch[brick] = frame (LRU).  Both
  // are null for in-core volumes.
  uint32_t* requests;
  uint32_t* touch;
  uint32_t frame;
  uint32_t* framebuffer;
  // per-region LOD (1D TF only): code per LOD-0 brick (level | 0x80 when
  // transparent), all levels' geometry, and an optional per-pixel output of
  // the region that added the most opacity.  region_code null: uniform LOD.
  const uint8_t* region_code;
  const LodInfo* lods;
  int32_t* region_out;
  // TF groups (gradient launches): jobs that differ only in their TF share
  // one march.  group_n > 1: this job leads a group of group_n jobs, the
  // others being jobs group_jobs[0 .. group_n - 2] of the table; the
  // leader's rays composite every member's TF and write every member's
  // framebuffer.  group_n < 0: a member (no tiles of its own).  0: alone.
  int32_t group_n;
  int32_t group_jobs[3];
  // Parts (the render server's split of a heavy request over launches):
  // tile_count > 0 renders only the job's tile slots [tile_first,
  // tile_first + tile_count) -- slots as the queues count them (rect tiles,
  // or runs of tile_sample) -- and a part with tile_first > 0 leaves the
  // pixels outside the tile rect to the part that starts at slot 0.  Both
  // 0: every slot.
  int32_t tile_first;
  int32_t tile_count;
};
constexpr int kGroupMax = 4;

// Launch-wide part of a raycast launch.  One launch renders any number of
// jobs (views x TFs x volumes that share a kernel variant); their ray tiles
// form one index space that the persistent kernel's waves drain.
constexpr int kMaxJobs = 64;
constexpr int kQueues = 8;  // one tile queue per XCD
struct JobTableHeader {
  int32_t n_jobs;
  int32_t total_tiles;
  int32_t tile_group;   // consecutive ray tiles dealt to one XCD in turn
  int32_t xcd_split;    // one-job launches: 0 round-robin runs; rx: rx x (8/rx) regions
  int32_t n_tf_lds;     // > 0: every job reads this 1D TF, staged into LDS
  int32_t pad_[3];
  const float* tf_lds;
  // [0, 8) per-XCD tile queue heads, then per job j: [8 + 3j] evaluated
  // samples, [9 + 3j] nominal samples, [10 + 3j] gathered samples (those
  // whose voxels the march loaded: a TF group's walk is counted on its
  // leader and 0 on its members; a job marched alone gathers the samples
  // it evaluates)
  unsigned long long* counters;
};

// The job table as the kernel reads it from device memory.
struct JobTable {
  JobTableHeader h;
  int32_t tile_start[kMaxJobs + 8];  // prefix sum of the jobs' tile counts
  RaycastParams jobs[kMaxJobs];
};
constexpr int kJobCounters = 3;
constexpr int kCounterWords = kQueues + kJobCounters * kMaxJobs;

// One brick to cut out of a flat level into a pool slot.
struct BrickJob {
  int32_t bx, by, bz;
  int32_t brick_id;  // global brick index (page-table index)
  int64_t slot;      // destination slot
};

}  // namespace tvk
