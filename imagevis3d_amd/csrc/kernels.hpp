// kernels.hpp -- host-side launchers of the HIP kernels (gfx950).
// All launchers are asynchronous on `stream` and return the hipError_t of
// the launch.
#pragma once
#include <hip/hip_runtime_api.h>

#include "tvk_common.hpp"

namespace tvk {

// Level l+1 from level l by 2x2x2 integer mean: floor(sum / count) over the
// existing children, or (median) the lower median of the existing children.
// dst dims are ceil(src / 2).
hipError_t launch_downsample(const void* src, int bpv, const int64_t src_dims[3], void* dst,
                             hipStream_t stream, bool median = false);

// Cut `n_jobs` bricks of payload B with apron a out of a flat level into
// pool slots (blocked slot layout, clamp-to-edge outside the volume) and
// write each brick's (min, max) over the slot to brick_range[2*brick_id].
// Slot j starts at element j * slot_stride of the pool; scratch_minmax
// holds 2 u32 per job.
hipError_t launch_extract_bricks(const void* level, int bpv, const int64_t dims[3],
                                 const BrickJob* jobs, int n_jobs, void* pool,
                                 int64_t slot_stride, int B, int apron, float* brick_range,
                                 uint32_t* scratch_minmax, hipStream_t stream);

// Overlap-inclusive x-fastest bricks (S = P + 2o per axis) of a flat level
// for the TOC builder: job j (its `slot` field = index into `out` and into
// minmax) is brick (bx, by, bz); outside the level the edge voxel repeats
// (clamp) or 0 is stored.  minmax gets 2 u32 (min, max) per job.
hipError_t launch_extract_linear_bricks(const void* level, int bpv, const int64_t dims[3],
                                        const BrickJob* jobs, int n_jobs, void* out, int P, int o,
                                        bool clamp, uint32_t* minmax, hipStream_t stream);

// An x-fastest S^3 brick (src) into one slot of the blocked slot layout,
// and one slot back into an x-fastest S^3 brick (dst).
hipError_t launch_linear_to_slot(const void* src, void* slot, int bpv, int S, hipStream_t stream);
hipError_t launch_slot_to_linear(const void* slot, void* dst, int bpv, int S, hipStream_t stream);

// Empty-space cell ranges of one flat level: out[(cz*ncy + cy)*ncx + cx] =
// min | max << 16 over voxels [c*C - 1, (c+1)*C] (clamped to the level).
hipError_t launch_cell_ranges(const void* level, int bpv, const int64_t dims[3], int C,
                              uint32_t* out, uint32_t* tmp_hi, hipStream_t stream);

// Cell ranges of the cells inside one uploaded brick, computed from its pool
// slot (slot edge S = B + 2*apron; `cells` = the level's cell grid; out =
// the level's cell-range array).  tmp: 2 (B/C)^3 words set up once by
// launch_slot_cell_tmp_init and left set up by every call; calls that share
// it run in one stream.
hipError_t launch_slot_cell_ranges(const void* slot, int bpv, int S, int apron, int B, int C,
                                   const int brick[3], const int cells[3], uint32_t* out,
                                   uint32_t* tmp, hipStream_t stream);
hipError_t launch_slot_cell_tmp_init(uint32_t* tmp, int per, hipStream_t stream);

// 1D histogram of raw values, n_bins = 2^(8*bpv), u64 counts (cleared here).
hipError_t launch_histogram_1d(const void* vol, int bpv, int64_t n, unsigned long long* hist,
                               hipStream_t stream);

// Squared-gradient pass of the 2D histogram: K = integer squared gradient
// magnitude in units of 1/(2*maxval); kmax receives max K (cleared here).
hipError_t launch_gradient_kmax(const void* vol, int bpv, const int64_t dims[3],
                                unsigned long long* kmax, hipStream_t stream);

// Binning pass: hist[v_bin * n_grad + g_bin] (cleared here) with
// g_bin = floor((n_grad-1) * sqrt(K / kmax)) decided in exact integers.
hipError_t launch_histogram_2d(const void* vol, int bpv, const int64_t dims[3],
                               const unsigned long long* kmax, int n_value, int n_grad,
                               unsigned long long* hist, hipStream_t stream);

// Per-brick histograms of a flat level (payload only): bins in {2, 4, ...,
// 64} equal bins over the value range; out = n_bricks x bins u32.
hipError_t launch_brick_histograms(const void* level, int bpv, const int64_t dims[3], int B, int bins,
                                   uint32_t* out, hipStream_t stream);

// Mean |LOD-0 voxel - its level-l ancestor| per LOD-0 brick region of edge
// B and level (out = n_regions x n_lod floats, level 0 = 0).
hipError_t launch_region_detail(const void* const* levels, const int64_t (*dims)[3], int n_lod, int bpv,
                                int B, float* out, hipStream_t stream);

// Zero-fill a u64 buffer (a kernel, so the call is safe under graph capture).
hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v,
                           hipStream_t stream);

// Empty flags (u8) of n_cells empty-space cells for one TF (mode 0/1: TF
// prefix counts over n_tf entries or columns; mode 2: iso value).
hipError_t launch_cell_empty(const uint32_t* cell_range, int64_t n_cells, const int32_t* prefix,
                             int n_tf, float tf_scale, int mode, float iso_value, uint8_t* out,
                             hipStream_t stream);

// Device-side synthetic volumes (TVK_VOL_FBM_PLANES, _NOISE_TUBES,
// _FBM_SPHERES): z slab [z0, z0 + nz) of a dims volume into `out`.
bool is_gpu_volume_kind(int kind);
hipError_t launch_generate_volume(int kind, uint64_t seed, const int64_t dims[3], int bits,
                                  int64_t z0, int64_t nz, void* out, hipStream_t stream);

// Learned-LOD error predictor (bf16 MLP 32 -> 4 x 64 -> 1 on MFMA).
// lod_mlp_pack turns the weights -- bf16 [in][out] per layer, then the 64
// output weights -- into the kernel's LDS image (lod_mlp_packed_elems()
// elements), once per predictor; launch_lod_mlp takes the packed weights,
// b = f32 biases (4 x 64 + 1), x = n x 32 f32 inputs, y = n outputs.
size_t lod_mlp_packed_elems();
void lod_mlp_pack(const uint16_t* w, uint16_t* packed);
hipError_t launch_lod_mlp(const uint16_t* w_packed, const float* b, const float* x, float* y, int n,
                          hipStream_t stream);

// Kernel variant of a raycast launch: every job of the table shares it.
struct LaunchInfo {
  int bpv = 1;          // bytes per voxel of every job's volume
  int interp = 1;       // 0 nearest, 1 trilinear
  int mode = 0;         // 0 1D TF, 1 2D TF, 2 isosurface
  int lighting = 0;     // 1D TF with lighting
  int batch = 8;        // samples per gather batch (1D TF without lighting)
  int n_tf_lds = 0;     // > 0: the table's shared 1D TF goes to LDS
  int regions = 0;      // 1D TF with per-region LOD (jobs carry region codes)
  int compute_units = 0;
  int wg_waves = 4;      // waves per work-group of variants without an LDS TF (4 or 8)
  int stage = 0;        // 1D TF without lighting, trilinear: voxels of each cell staged in LDS
  // Gradient launches (2D TF, or 1D TF with lighting): 1 lets the table
  // group views that differ only in their TF into one march (launch_info);
  // the launch keeps it at 1 only when the filled table holds a group.
  int group = 0;
};

// One raycast launch over a device job table `tab` (its header counters
// must be zeroed first).
hipError_t launch_raycast(const LaunchInfo& li, const JobTable* tab, hipStream_t stream);

// Background of the pixels outside each job's tile rect, for every job of
// the device table `tab` (a fixed grid of kBackgroundBlocks work-groups that
// strides over all jobs' pixels).
constexpr int kBackgroundBlocks = 1024;
hipError_t launch_background(const JobTable* tab, hipStream_t stream);
// The allocation-free raycast's device scratch (noalloc.hip): init once per
// device with that device current; then the device's job table and counters.
hipError_t noalloc_scratch_init(int device);
hipError_t noalloc_scratch(int device, JobTable** table, unsigned long long** counters);
// *out += sum over the n pixels of two RGBA8 images of the squared RGB
// differences (integer units)
hipError_t launch_rgb_sse(const uint32_t* a, const uint32_t* b, int64_t n, unsigned long long* out,
                          hipStream_t stream);
// dst (dw x dh) from src (sw x sh) by nearest pixel: dst(x, y) = src(x sw / dw, y sh / dh)
hipError_t launch_upscale_nearest(const uint32_t* src, int sw, int sh, uint32_t* dst, int dw, int dh,
                                  hipStream_t stream);

}  // namespace tvk
