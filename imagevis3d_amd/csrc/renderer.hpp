// renderer.hpp -- HipRenderer: the C++ side of the MI355X Tuvok backend.
//
// It keeps the shape of Tuvok's AbstrRenderer as the ImageVis3D tree uses
// it through Lua (ImageVis3D/UI/RenderWindow.cpp, LuaScripts/
// BatchRenderer.lua): a state object with setters for render mode, TFs,
// camera, sampling rate and LOD limits, and a Paint() that renders one
// subframe.  The GPU memory manager (GPUMemMan) role is played by BrickPool.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <string>
#include <vector>

#include <mutex>

#include "../../include/tvk_hip.h"
#include "kernels.hpp"
#include "memman.hpp"
#include "tvk_common.hpp"

namespace tvk {

// Owning device allocation, accounted against its device's memory budget
// (MemMan): an allocation beyond the budget first frees paged-pool chunks
// and fails with hipErrorOutOfMemory when that is not enough.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int device = -1;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), device(o.device) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      bytes = o.bytes;
      device = o.device;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { reset(); }
  hipError_t alloc(size_t n);
  void reset();
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// A volume's bricks on the device: in-core volumes own one allocation of
// slots (every brick resident); paged volumes hold slots of the device-wide
// pool (MemMan), at most n_slots of them.  The page table holds each
// brick's slot address (0: not resident).
struct BrickPool {
  DevBuf slots;            // in-core: n_slots * slot_bytes
  DevBuf page_table;       // uint64 per brick
  DevBuf brick_range;      // float2 per brick
  uint64_t n_slots = 0;
  uint64_t slot_bytes = 0;
  uint64_t resident = 0;
  uint64_t uploaded = 0;
  uint64_t evicted = 0;
  std::vector<uint64_t> page_table_h;
  std::vector<float> brick_range_h;
  std::vector<int64_t> brick_slot;  // paged: the brick's slot id in the device pool, -1
};

// Geometry of a bricked LOD pyramid (RasterDataBlock layout).
struct LevelPlan {
  uint64_t dims[kMaxLod][3];
  LodInfo lods[kMaxLod];
  int n_lod = 0, total_bricks = 0, total_cells = 0, cell = 16, slot_edge = 0, brick = 0;
  uint64_t level_bytes = 0, slot_bytes = 0;
};

// One transfer-function slot: a 1D TF and/or a 2D TF with the prefix
// counts of entries (columns) whose alpha is > 0 (the empty-space test).
struct TfSet {
  std::vector<float> tf1d;
  std::vector<int32_t> prefix1d;
  bool dirty1d = false;
  DevBuf tf1d_dev, prefix1d_dev;
  std::vector<float> tf2d;
  std::vector<int32_t> prefix2d;
  uint32_t w2d = 0, h2d = 0;
  float max_grad = 1.f;
  bool dirty2d = false;
  DevBuf tf2d_dev, prefix2d_dev;
  uint64_t version = 0;   // bumped by every change of this slot
  DevBuf region_code;     // per LOD-0 region: chosen level | 0x80 when TF-empty
  DevBuf region_code0;    // the same with every region at level 0
  uint64_t region_key = 0;
  DevBuf cell_empty[3];      // per render mode: u8 per empty-space cell of every level
  uint64_t cell_key[3] = {};  // what each cell_empty was built for
};

// A camera (Tuvok's view position / direction / up, FoV or ortho extent).
struct CameraState {
  double pos[3] = {0.0, 0.0, 2.0};
  double dir[3] = {0.0, 0.0, -1.0};
  double up[3] = {0.0, 1.0, 0.0};
  double fov = 50.0;
  bool ortho = false;
  double ortho_hh = 0.5;
};

// Device and pinned host resources of one raycast launch: the job table,
// the counters and, for batched launches, the captured graph replaying
// [table upload, counter clear, raycast, counter readback].
struct LaunchSlot {
  DevBuf table;                                 // JobTable
  JobTable* table_h = nullptr;                  // pinned staging copy
  DevBuf counters;                              // kCounterWords u64
  unsigned long long* counters_h = nullptr;     // pinned copy after the launch
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // start, kernel in, kernel out, done
  bool used = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t graph_key = 0;
  int alloc();
  void release();
};

// Tile slots a job puts in the launch's queues: its rect's tiles, or one
// per run of tile_sample of them; a part's tile_count; 0 for a TF group
// member.
int job_tiles(const RaycastParams& p);

// Fill `sl`'s job table from jobs[0..n) and launch it on `stream` (graph
// replay when `graph`; one graph per kernel variant, re-captured on change).
int launch_job_table(LaunchSlot& sl, const RaycastParams* jobs, int n, const LaunchInfo& li,
                     int tile_group, int xcd_split, hipStream_t stream, bool graph);
// The eager sequence on explicit buffers: fill the pinned table_h, copy it
// to dev_table, clear dev_counters, background + raycast, and (when
// counters_h) copy the counters back; ev = start, kernel in, kernel out, done.
int launch_job_table_at(JobTable* table_h, void* dev_table, unsigned long long* dev_counters,
                        unsigned long long* counters_h, hipEvent_t ev[4], const RaycastParams* jobs, int n,
                        const LaunchInfo& li, int tile_group, int xcd_split, hipStream_t stream);

// Out-of-core brick source (see tvk_brick_source_fn in tvk_hip.h).
typedef int (*BrickSourceFn)(void* user, uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz,
                             void* slot, uint64_t slot_bytes);

class HipRenderer : public PoolClient {
 public:
  HipRenderer(int type, bool pow2, bool to8, bool no_border, bool bias_scale_tf);
  ~HipRenderer();

  int Initialize(int device);
  int Resize(uint32_t w, uint32_t h);
  int SetRendererTarget(int t);
  int RendererTarget() const { return target_; }

  int LoadDataset(const char* path, bool verify);
  int SetVolume(const void* voxels, const uint64_t dims[3], uint32_t bits, uint32_t brick,
                uint32_t apron, uint32_t n_lod);
  int SetGeneratedVolume(int kind, uint64_t seed, const uint64_t dims[3], uint32_t bits,
                         uint32_t brick, uint32_t apron, uint32_t n_lod);
  int SetPagedVolume(const uint64_t dims[3], uint32_t bits, uint32_t brick, uint32_t apron,
                     uint32_t n_lod, BrickSourceFn fn, void* user, const float* brick_ranges,
                     uint64_t pool_bytes);
  int SetPagedFlatVolume(const void* voxels, const uint64_t dims[3], uint32_t bits, uint32_t brick,
                         uint32_t apron, uint32_t n_lod, uint64_t pool_bytes);
  // bricks of every level of the pyramid SetPagedVolume would build
  int PlannedBrickCount(const uint64_t dims[3], uint32_t bits, uint32_t brick, uint32_t apron,
                        uint32_t n_lod, uint64_t* count) const {
    LevelPlan plan;
    if (int e = plan_levels(dims, bits, brick, apron, n_lod, &plan)) return e;
    *count = (uint64_t)plan.total_bricks;
    return 0;
  }
  int LodLevelCount() const { return n_lod_; }
  int BitWidth(uint32_t* bits) const;
  int DomainSize(uint32_t lod, uint64_t dims[3]) const;
  int BrickLayout(uint32_t lod, uint32_t grid[3], uint32_t* brick, uint32_t* apron) const;
  int ReadLod(uint32_t lod, void* out, uint64_t bytes);
  int ReadBrick(uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz, void* out, uint64_t bytes);
  int BrickMinMax(uint32_t lod, float* mins, float* maxs, uint64_t n);
  int Histogram1D(uint64_t* hist, uint64_t n_bins);
  int BrickHistograms(uint32_t lod, uint32_t bins, uint32_t* out, uint64_t n);
  int RegionDetail(float* out, uint64_t n);
  int Histogram2D(uint64_t* hist, uint32_t n_value, uint32_t n_grad, float* max_grad);

  int SetRenderMode(int m);
  int Set1DTrans(const float* rgba, uint32_t n) { return Set1DTransSlot(0, rgba, n); }
  int Set2DTrans(const float* rgba, uint32_t w, uint32_t h, float max_grad) {
    return Set2DTransSlot(0, rgba, w, h, max_grad);
  }
  int Set1DTransSlot(uint32_t slot, const float* rgba, uint32_t n);
  int Set2DTransSlot(uint32_t slot, const float* rgba, uint32_t w, uint32_t h, float max_grad);
  int SetIsoValue(float v);
  int SetUseMIP(bool on);
  bool GetUseMIP() const { return use_mip_; }
  int SetMIPRotationAngle(float deg);
  int SetIsosurfaceColor(float r, float g, float b);
  // ClearView (isosurface mode): focus surface, colour, focus disc, context
  // and border scales; the focus position is in pixels, row 0 at the top
  int SetClearViewEnabled(bool on);
  int SetCVIsoValue(float v);
  int SetCVColor(float r, float g, float b);
  int SetCVSize(float s);
  int SetCVContextScale(float s);
  int SetCVBorderScale(float s);
  int SetCVFocusPos(int32_t x, int32_t y);
  void GetClearView(tvk_clearview* out) const;

  int SetViewPos(const float v[3]);
  int SetViewDir(const float v[3]);
  int SetUpDir(const float v[3]);
  int SetFoV(float deg);
  int SetOrtho(bool on, float half_height);
  int SetRotation(const float m[16]);
  int SetTranslation(const float m[16]);
  int SetRescaleFactors(const double s[3]);

  int SetSampleRateModifier(float r);
  int SetInterpolant(int i);
  int SetLODLimits(uint32_t mn, uint32_t mx);
  int SetBGColors(const float top[3], const float bottom[3]);
  int SetLightingEnabled(bool on);
  int SetLightColors(const float amb[4], const float dif[4], const float spec[4], const float dir[3]);
  int SetEarlyRayTermination(float a);
  // Clip plane (a, b, c, d) in world space: points with a x + b y + c z + d
  // > 0 are cut away while it is enabled.
  int SetClipPlane(const float plane[4]);
  int EnableClipPlane(bool on);
  bool ClipPlaneEnabled() const { return clip_on_; }
  void GetClipPlane(float plane[4]) const {
    for (int i = 0; i < 4; ++i) plane[i] = clip_[i];
  }
  // Compositing target precision: TVK_BP_8BIT, TVK_BP_16BIT, TVK_BP_32BIT.
  int SetBlendPrecision(int precision);
  int BlendPrecision() const { return blend_; }
  // Interactive frame-time governor (setPerfMeasures): min_fps 0 turns it
  // off (every interactive frame then starts at the coarsest allowed LOD).
  int SetPerfMeasures(uint32_t min_fps, bool low_res, float screen_dec, float sample_dec,
                      uint32_t lod_delay);
  // the framebuffer as the last paint left it (no further painting)
  int ReadFramebuffer(uint8_t* rgba, uint64_t bytes);
  int SetEmptySpaceSkipping(bool on);
  int SetKernelTuning(uint32_t batch, uint32_t tile_group, uint32_t xcd_split);
  int SetWorkGroupWaves(uint32_t waves);  // 4 or 8 (variants without an LDS TF)
  int SetBrickStaging(bool on);          // 1D march over LDS-staged cells
  int SetTFGrouping(bool on);            // one march per TF group (gradient launches)
  int SetRayTileShape(uint32_t tile_width);
  // Per-region LOD (1D TF): one level per LOD-0 brick; n == 0 turns it off.
  int SetRegionLods(const uint8_t* levels, uint64_t n);
  int SetRegionOutput(bool on);
  int ReadRegionOutput(int32_t* out, uint64_t n);
  // Per-frame check of per-region LOD frames against level 0: every paint
  // also renders one ray tile in each run of `tile_sample` at level 0 into
  // the frame, estimates the frame's remaining RGB RMSE (1/255 units) from
  // those tiles and, when it exceeds max_rmse, renders the whole frame at
  // level 0.  tile_sample 0 turns it off.
  int SetRegionCheck(uint32_t tile_sample, float max_rmse);
  int RegionCheckStats(tvk_region_check* out);
  // In-core volumes: cut the device-resident levels into bricks of another
  // size / empty-space cell edge (cell 0 keeps the current preference).
  int Rebrick(uint32_t brick, uint32_t cell);
  int Autotune(const tvk_view* views, uint32_t n, uint32_t flags, tvk_tuning* out);
  int SetEmptySpaceCell(uint32_t cell);

  int SetLaunchGraphs(bool on);

  int Paint();
  // Many views (camera, size, TF slot) of this volume in one launch.
  int RenderBatch(const tvk_view* views, uint32_t n, tvk_view_stats* stats, float* kernel_ms);
  int BatchFramebuffer(uint32_t index, void** dptr);
  // RenderBatch into caller device framebuffers with no allocation: the job
  // table and counters in the device's fixed scratch (tvk_render_batch_noalloc)
  int RenderBatchNoAlloc(const tvk_view* views, uint32_t n, void* const* dev_rgba, tvk_view_stats* stats);

  // Serving (RenderServer): ServePrepare checks the TF slot for `mode`,
  // uploads TFs and builds the slot's empty-cell flags (device work on this
  // renderer's stream is finished when it returns); ServeJob builds the job
  // of one view in `mode` (finest LOD limit) and its kernel variant.
  int ServePrepare(uint32_t tf_slot, int mode);
  int ServeJob(const tvk_view& v, int mode, uint32_t* fb, RaycastParams* out, LaunchInfo* li);
  int Device() const { return device_; }
  // Serving paged volumes: ServeCheck validates a request (view, TF slot,
  // mode); ServeBatch renders views in `mode` with the paging loop of
  // RenderBatch; PagedWorkingSetBytes is the pool memory a batch may hold
  // (the finest allowed LOD's bricks, at most the volume's share).
  bool Paged() const { return paged_; }
  int ServeCheck(const tvk_view& v, int mode);
  int ServeBatch(const tvk_view* views, uint32_t n, int mode, tvk_view_stats* stats);
  uint64_t PagedWorkingSetBytes() const;
  int CheckForRedraw() const { return frame_complete_ ? 0 : 1; }
  int ScheduleCompleteRedraw();
  int Capture(uint8_t* rgba, uint64_t bytes, bool preserve_transparency);
  int Progress(uint64_t* subframes, uint32_t* working_subframe, uint32_t* bricks,
               uint32_t* working_brick, uint64_t* min_lod) const;
  int FrameStats(tvk_frame_stats* out);
  int DeviceFramebuffer(void** dptr, void** stream);
  int Synchronize();

  // budgets of this renderer's device (shared by every renderer there)
  int SetMaxUsableGPUMem(uint64_t bytes);
  int SetMaxUsableCPUMem(uint64_t bytes);
  int MemStats(tvk_mem_stats* out);
  // threads that run a paged volume's brick source (1: one at a time)
  int SetBrickSourceThreads(uint32_t n);
  void PagingStats(uint64_t* bytes, double* ms) const {
    *bytes = paging_bytes_;
    *ms = paging_ms_;
  }
  // the device-wide pool takes back brick gid's slot (MemMan callback)
  void PoolEvict(int gid) override;

 private:
  int check_ready() const;
  int check_tf(int slot) const;
  bool launch_ready_without_alloc(const tvk_view* views, uint32_t n, int mode) const;
  int view_camera(const tvk_view& v, uint32_t index, CameraState* c) const;
  int upload_tf();
  void build_params(RaycastParams* p, int lod, const CameraState& cam, uint32_t w, uint32_t h,
                    int tf_slot, uint32_t* framebuffer, float rate_div = 1.f) const;
  // governor: the interactive frame's first LOD, and the time model behind it
  int governor_start_lod(int lo, int hi) const;
  double predicted_ms(int lod, int lo) const;
  void record_subframe_time();
  LaunchInfo launch_info(int n_tf_lds) const;
  // fill slot's table from jobs[0..n) and launch (graph replay if enabled)
  int launch_jobs(LaunchSlot& slot, const RaycastParams* jobs, int n, const LaunchInfo& li,
                  bool graph);
  uint32_t bricks_skipped(int lod, int tf_slot) const;
  float tf_scale(int n_entries) const;
  bool regions_active() const { return !region_lod_h_.empty(); }
  int prepare_regions(int tf_slot);
  int upload_lods();
  // (re)build the empty flags of TF slot `tf_slot` for the current mode
  int prepare_cells(int tf_slot, int mode);
  int job_mode() const;
  int mode_index() const {
    return render_mode_ == TVK_RM_2DTRANS ? 1 : (render_mode_ == TVK_RM_ISOSURFACE ? 2 : 0);
  }
  double value_max() const;
  int plan_levels(const uint64_t dims[3], uint32_t bits, uint32_t brick, uint32_t apron,
                  uint32_t n_lod, LevelPlan* out) const;
  void release_volume();
  int set_volume_impl(const void* voxels, int gen_kind, uint64_t seed, const uint64_t dims[3],
                      uint32_t bits, uint32_t brick, uint32_t apron, uint32_t n_lod);
  int brick_levels(const LevelPlan& plan);
  // paged pools: upload the bricks the last raycast asked for; *complete is
  // false when another pass is needed, *overflow when the pool cannot hold
  // the bricks of this LOD for this view
  int service_requests(int lod, bool* complete, bool* overflow);
  int service_requests(int lod_lo, int lod_hi, bool* complete, bool* overflow);
  // fill bricks gids[0..n) from the source (or the host cache) on worker
  // threads and stream them into their slots on the copy stream
  int upload_bricks(const int* gids, const int* lods, void* const* addrs, size_t n);
  // the page table to the device if evictions changed it (pool_mu_ held)
  int sync_page_table_locked();
  // the LRU stamps of the bricks the last pass sampled
  int touch_sampled();
  static int flat_source(void* user, uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz,
                         void* slot, uint64_t slot_bytes);
  void invalidate() {
    frame_complete_ = false;
    next_lod_ = -1;
    lod_floor_ = 0;
  }

  // construction flags (tuvok.renderer.new)
  int type_;
  bool pow2_, to8_, no_border_, bias_scale_tf_;

  // device / context
  bool initialized_ = false;
  int device_ = -1;
  hipStream_t stream_ = nullptr;
  uint32_t width_ = 0, height_ = 0;
  DevBuf framebuffer_;
  LaunchSlot paint_slots_[2];  // Paint alternates between two
  int paint_cur_ = 0;
  LaunchSlot batch_slot_;
  DevBuf batch_fb_;
  std::vector<uint64_t> batch_fb_offset_;  // per view of the last batch
  bool use_graphs_ = true;  // batched launches replay captured graphs
  uint64_t gpu_total_ = 0, gpu_budget_ = 0;
  int compute_units_ = 0;

  // volume
  bool has_volume_ = false;
  uint32_t bits_ = 8;
  uint64_t dims0_[3] = {0, 0, 0};
  int n_lod_ = 0;
  int brick_ = 0, apron_ = 0, slot_edge_ = 0;
  LodInfo lods_[kMaxLod];
  int total_bricks_ = 0;
  DevBuf levels_[kMaxLod];
  BrickPool pool_;
  DevBuf scratch_;
  DevBuf cell_range_;  // empty-space cell ranges of every level
  int cell_ = 16;      // empty-space cell edge (divides the brick size)
  int cell_pref_ = 16; // requested cell edge for the next volume
  int total_cells_ = 0;
  DevBuf lods_dev_;                    // LodInfo of every level
  std::vector<uint8_t> region_lod_h_;  // per LOD-0 brick (empty: uniform LOD)
  uint64_t region_version_ = 0;
  bool region_out_on_ = false;
  DevBuf region_out_dev_;              // int32 per pixel of the framebuffer
  // per-frame region check: launched by Paint, resolved (estimate read,
  // fallback launched) by the next call that needs the frame
  int check_launch(const RaycastParams& p);
  int resolve_check();
  uint32_t check_k_ = 0;               // tile sampling (0: off)
  float check_max_ = 1.f;              // RGB RMSE bound, 1/255 units
  uint32_t check_phase_ = 0;
  DevBuf check_copy_;                  // the frame before the check's tiles
  DevBuf check_sse_;                   // u64 squared-error sum
  unsigned long long* check_sse_h_ = nullptr;  // pinned
  LaunchSlot check_slot_;
  hipEvent_t check_ev_[3] = {nullptr, nullptr, nullptr};  // start, end, sum copied
  RaycastParams check_params_{};       // the frame's params (for the fallback)
  bool check_open_ = false;            // launched, not resolved
  bool check_timing_ = false;          // check_stats_.check_ms not read yet
  tvk_region_check check_stats_{};
  uint64_t cells_gen_ = 1;  // bumped whenever cell ranges change

  // paged (out-of-core) pools: GridLeaper-style requests + LRU
  bool paged_ = false;
  BrickSourceFn source_fn_ = nullptr;
  void* source_user_ = nullptr;
  const void* host_level0_ = nullptr;             // caller-owned LOD 0 (flat source)
  uint64_t pool_bytes_req_ = 0;                   // the pool size the caller asked for
  std::vector<std::vector<uint8_t>> host_levels_;  // LOD >= 1 (flat source)
  DevBuf requests_dev_, touch_dev_;               // u32 per brick
  uint32_t* requests_h_ = nullptr;                // pinned
  uint64_t requests_h_count_ = 0;
  // upload pipeline: two sets of pinned staging buffers (one filled by the
  // source threads while the other's copies run), device staging buffers
  // the copies land in before the layout kernel moves them into slots
  std::vector<void*> staging_;
  hipEvent_t staging_ev_[2] = {nullptr, nullptr};
  std::vector<DevBuf> upload_stage_;
  DevBuf slot_cell_tmp_;        // per-cell min / max words of one brick's cell-range pass
  int slot_cell_tmp_per_ = 0;   // cells per brick edge they are set up for
  int staging_round_ = 0;              // bricks per round
  uint64_t staging_host_bytes_ = 0;    // accounted as host memory
  hipStream_t copy_stream_ = nullptr;  // brick uploads
  hipEvent_t upload_done_ = nullptr;
  int pool_client_ = -1;               // MemMan client id of a paged volume
  std::mutex pool_mu_;                 // page table vs evictions from other threads
  bool pt_dirty_ = false;              // page_table_h changed since its upload
  bool cache_bricks_ = false;          // keep source bricks in the host cache
  int source_threads_ = 8;
  uint64_t host_levels_bytes_ = 0;     // host copies of paged levels (accounted)
  uint64_t paging_bytes_ = 0;          // bricks uploaded (bytes) since create
  double paging_ms_ = 0.0;             // host time spent paging them
  uint32_t frame_ = 0;
  int lod_floor_ = 0;       // coarsest-needed fallback when the pool overflows
  std::vector<uint32_t> cell_init_h_;  // brick-range initial cell ranges

  // transfer functions: slot 0 is the one Set1DTrans / Set2DTrans set
  static constexpr int kMaxTfSlots = 64;
  int render_mode_ = TVK_RM_1DTRANS;
  std::vector<TfSet> tfs_ = std::vector<TfSet>(1);
  float iso_value_ = 0.f;
  bool use_mip_ = false;     // maximum intensity projection (1D TF mode)
  float mip_angle_ = 0.f;    // MIP view rotation about the up vector, degrees
  float iso_color_[3] = {1.f, 1.f, 1.f};
  bool cv_enabled_ = false;
  float cv_iso_ = 0.f;
  float cv_color_[3] = {1.f, 0.f, 0.f};
  float cv_size_ = 5.5f, cv_context_ = 1.f, cv_border_ = 60.f;
  int32_t cv_focus_[2] = {0, 0};
  // lowest iso value the isosurface march looks for (the context's, or the
  // focus's when ClearView is on and it is lower): cells below it are skipped
  float iso_skip_threshold() const;

  // camera
  CameraState cam_;
  double rot_[16], trans_[16];
  double rescale_[3] = {1.0, 1.0, 1.0};

  // quality / LOD
  float rate_ = 1.0f;
  int interp_ = TVK_LINEAR;
  uint32_t lod_min_ = 0, lod_max_ = 0;
  float bg_top_[3] = {0.f, 0.f, 0.f}, bg_bottom_[3] = {0.f, 0.f, 0.f};
  float bg_alpha_ = 1.f;
  bool lighting_ = false;
  // ambient rgba, diffuse rgba, specular rgb + shininess, light direction
  float light_[15] = {0.1f, 0.1f, 0.1f, 1.f, 0.9f, 0.9f, 0.9f, 1.f,
                      0.3f, 0.3f, 0.3f, 16.f, 0.f, 0.f, 1.f};
  float term_alpha_ = 0.99f;
  float clip_[4] = {0.f, 0.f, -1.f, 0.f};  // keeps z >= 0 of the world when enabled
  bool clip_on_ = false;
  int blend_ = TVK_BP_32BIT;
  bool skip_ = true;
  int batch_ = 8;
  int tile_group_ = 4;
  int xcd_split_ = 0;
  int wg_waves_ = 4;
  bool stage_ = false;  // LDS-staged 1D march
  bool group_tfs_ = true;  // gradient launches group views that differ only in TF
  int tile_shift_ = 3;  // 8 x 8 ray tiles
  int target_ = TVK_RT_INTERACTIVE;

  // frame-time governor (setPerfMeasures)
  uint32_t min_fps_ = 0;
  bool low_res_ = false;
  float screen_dec_ = 2.f, sample_dec_ = 2.f;
  uint32_t lod_delay_ = 0, delay_left_ = 0;
  double lod_ms_[kMaxLod];             // measured full-size subframe ms per LOD (< 0: none)
  double last_scale_ = 1.0;            // pixel x rate fraction of the last subframe
  bool time_pending_ = false;          // the last subframe's time is not recorded yet
  bool capturing_ = false;             // inside Capture: no LOD delay
  bool last_pass_incomplete_ = false;  // the last paint paged bricks in: its LOD repeats
  DevBuf lowres_fb_;                   // reduced-resolution intermediate subframes

  // progress
  bool frame_complete_ = false;
  int next_lod_ = -1;
  int last_lod_ = -1;
  uint64_t subframes_ = 0;
  tvk_frame_stats stats_{};
  bool stats_pending_ = false;
  LaunchSlot* stats_slot_ = nullptr;
};

}  // namespace tvk
