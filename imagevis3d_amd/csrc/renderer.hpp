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

#include "../../include/tvk_hip.h"
#include "tvk_common.hpp"

namespace tvk {

// Owning device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  hipError_t alloc(size_t n);
  void reset();
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Device brick pool + page table + per-brick value ranges (GPUMemMan).
struct BrickPool {
  DevBuf slots;            // n_slots * slot_bytes
  DevBuf page_table;       // int32 per brick
  DevBuf brick_range;      // float2 per brick
  uint64_t n_slots = 0;
  uint64_t slot_bytes = 0;
  uint64_t resident = 0;
  uint64_t uploaded = 0;
  uint64_t evicted = 0;
  std::vector<int32_t> page_table_h;
  std::vector<float> brick_range_h;
};

class HipRenderer {
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
  int LodLevelCount() const { return n_lod_; }
  int BitWidth(uint32_t* bits) const;
  int DomainSize(uint32_t lod, uint64_t dims[3]) const;
  int BrickLayout(uint32_t lod, uint32_t grid[3], uint32_t* brick, uint32_t* apron) const;
  int ReadLod(uint32_t lod, void* out, uint64_t bytes);
  int ReadBrick(uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz, void* out, uint64_t bytes);
  int BrickMinMax(uint32_t lod, float* mins, float* maxs, uint64_t n);
  int Histogram1D(uint64_t* hist, uint64_t n_bins);
  int Histogram2D(uint64_t* hist, uint32_t n_value, uint32_t n_grad, float* max_grad);

  int SetRenderMode(int m);
  int Set1DTrans(const float* rgba, uint32_t n);
  int Set2DTrans(const float* rgba, uint32_t w, uint32_t h, float max_grad);
  int SetIsoValue(float v);
  int SetIsosurfaceColor(float r, float g, float b);

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
  int SetEmptySpaceSkipping(bool on);
  int SetKernelTuning(uint32_t batch, uint32_t tile_group, uint32_t xcd_split);

  int Paint();
  int CheckForRedraw() const { return frame_complete_ ? 0 : 1; }
  int ScheduleCompleteRedraw();
  int Capture(uint8_t* rgba, uint64_t bytes, bool preserve_transparency);
  int Progress(uint64_t* subframes, uint32_t* working_subframe, uint32_t* bricks,
               uint32_t* working_brick, uint64_t* min_lod) const;
  int FrameStats(tvk_frame_stats* out);
  int DeviceFramebuffer(void** dptr, void** stream);
  int Synchronize();

  int SetMaxUsableGPUMem(uint64_t bytes);
  int MemStats(tvk_mem_stats* out);

 private:
  int check_ready() const;
  int upload_tf();
  void build_params(RaycastParams* p, int lod) const;
  double value_max() const;
  void invalidate() { frame_complete_ = false; next_lod_ = -1; }

  // construction flags (tuvok.renderer.new)
  int type_;
  bool pow2_, to8_, no_border_, bias_scale_tf_;

  // device / context
  bool initialized_ = false;
  int device_ = -1;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t width_ = 0, height_ = 0;
  DevBuf framebuffer_;
  DevBuf counters_;
  unsigned long long* counters_h_ = nullptr;  // pinned
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

  // transfer functions
  int render_mode_ = TVK_RM_1DTRANS;
  std::vector<float> tf1d_;
  bool tf_dirty_ = false;
  DevBuf tf_dev_, tf_prefix_dev_;
  float iso_value_ = 0.f;
  float iso_color_[3] = {1.f, 1.f, 1.f};
  std::vector<float> tf2d_;
  uint32_t tf2d_w_ = 0, tf2d_h_ = 0;
  float tf2d_max_grad_ = 1.f;
  bool tf2d_dirty_ = false;
  DevBuf tf2d_dev_, tf2d_prefix_dev_;

  // camera
  double view_pos_[3] = {0.0, 0.0, 2.0};
  double view_dir_[3] = {0.0, 0.0, -1.0};
  double up_dir_[3] = {0.0, 1.0, 0.0};
  double fov_ = 50.0;
  bool ortho_ = false;
  double ortho_hh_ = 0.5;
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
  bool skip_ = true;
  int batch_ = 8;
  int tile_group_ = 4;
  int xcd_split_ = 0;
  int target_ = TVK_RT_INTERACTIVE;

  // progress
  bool frame_complete_ = false;
  int next_lod_ = -1;
  int last_lod_ = -1;
  uint64_t subframes_ = 0;
  tvk_frame_stats stats_{};
  bool stats_pending_ = false;
};

}  // namespace tvk
