/*
 * tvk_hip.h -- C ABI of the MI355X (gfx950) renderer backend for Tuvok.
 *
 * This is the drop-in boundary for Tuvok's raycasting + brick/LOD hot path.
 * Every entry point replaces a method of Tuvok's AbstrRenderer (reached from
 * Lua through `cexec("<renderer>.<method>")`) or of the IO/UVF layer that
 * feeds it.  The reference interface each one stands in for is cited as
 * `file:line` of the ImageVis3D tree (Tuvok itself is an un-vendored
 * submodule, so its call sites are the evidence of the interface).
 *
 * Conventions
 *  - Plain C: opaque handle, int status (TVK_OK == 0, negative on error),
 *    caller-owned output buffers, no C++ or torch types.
 *  - One handle == one AbstrRenderer instance == one HIP stream on one
 *    device.  A handle is thread-compatible (not thread-safe), like the
 *    single-threaded Tuvok renderer (ImageVis3D/UI/ImageVis3D.cpp:976-983).
 *  - tvk_last_error() returns a message for the last failed call on this
 *    thread (the log-macro channel of Tuvok's T_ERROR).
 *  - Matrices are float[16], row-major, applied to row vectors
 *    (p' = p * M, translation in elements 12..14), which is the layout of
 *    Tuvok's FLOATMATRIX4 passed by setRotation4x4/setTranslation4x4.
 *  - World space: the volume's bounding box is centred at the origin and
 *    scaled so its largest extent (dims * rescale factors) is 1.  This is
 *    the space of setViewPos/setViewDir/setUpDir.
 *  - Images are RGBA8, row 0 at the top, W*H*4 bytes.
 */
#ifndef TVK_HIP_H
#define TVK_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum tvk_status {
  TVK_OK = 0,
  TVK_ERR_ARG = -1,     /* invalid argument (bool false / LuaError)        */
  TVK_ERR_HIP = -2,     /* a HIP runtime call failed                       */
  TVK_ERR_STATE = -3,   /* call out of order (e.g. paint before a volume)  */
  TVK_ERR_IO = -4,      /* file could not be read                          */
  TVK_ERR_OOM = -5,     /* brick pool / device memory budget exhausted     */
  TVK_ERR_FORMAT = -6,  /* malformed or unsupported UVF / TF file          */
  TVK_ERR_CHECKSUM = -7 /* UVF checksum mismatch (UVFReader/main.cpp:151)  */
};

/* EVolumeRendererType (ImageVis3D/UI/SettingsDlg.cpp:486-517,
 * RenderWindowGL.cpp:176-188).  This backend implements the two raycasters;
 * both map to the HIP raycaster, GRIDLEAPER additionally pages bricks on
 * demand from the host (out-of-core). */
enum tvk_renderer_type {
  TVK_OPENGL_SBVR = 0,
  TVK_OPENGL_RAYCASTER = 1,
  TVK_DIRECTX_SBVR = 2,
  TVK_DIRECTX_RAYCASTER = 3,
  TVK_OPENGL_2DSBVR = 4,
  TVK_DIRECTX_2DSBVR = 5,
  TVK_OPENGL_GRIDLEAPER = 6,
  TVK_DIRECTX_GRIDLEAPER = 7
};

/* AbstrRenderer::ERenderMode (RenderWindow.cpp:2051-2059). */
enum tvk_render_mode { TVK_RM_1DTRANS = 0, TVK_RM_2DTRANS = 1, TVK_RM_ISOSURFACE = 2 };

/* AbstrRenderer::ERendererTarget (DebugScriptWindow.cpp:219-221,
 * LuaScripts/BatchRenderer.lua:34,37). */
enum tvk_target { TVK_RT_INTERACTIVE = 0, TVK_RT_CAPTURE = 1, TVK_RT_HEADLESS = 2 };

/* Interpolant (RenderWindow.cpp:437-446). */
enum tvk_interpolant { TVK_NEAREST = 0, TVK_LINEAR = 1 };

/* AbstrRenderer::EBlendPrecision (SettingsDlg.cpp:247-251: 0 = 8 bit,
 * 1 = 16 bit, 2 = 32 bit; ImageVis3D_Settings.cpp:574 applies it). */
enum tvk_blend_precision { TVK_BP_8BIT = 0, TVK_BP_16BIT = 1, TVK_BP_32BIT = 2 };

typedef struct tvk_renderer tvk_renderer;

/* Per-paint statistics (the progress counters of RenderWindow.cpp:453-486
 * plus the measurement counters of this backend). */
typedef struct tvk_frame_stats {
  uint64_t evaluated_samples; /* samples fetched+composited (after skipping) */
  uint64_t nominal_samples;   /* sum over pixels of floor(chord * rate) on LOD 0 */
  uint32_t bricks_total;      /* bricks of the rendered LOD                  */
  uint32_t bricks_skipped;    /* bricks empty under the current TF           */
  uint32_t lod;               /* LOD index rendered by this paint            */
  uint32_t width, height;
  float kernel_ms;            /* raycast kernel time (HIP events)            */
  float paint_ms;             /* whole paint() on the device (HIP events)    */
} tvk_frame_stats;

/* Brick-pool / GPUMemMan accounting (ImageVis3D_DebugWindow.cpp:172-206). */
typedef struct tvk_mem_stats {
  uint64_t gpu_mem;            /* device memory total (SysInfo GetGPUMem)     */
  uint64_t gpu_budget;         /* SetMaxUsableGPUMem                          */
  uint64_t allocated_gpu_mem;  /* bytes held by this renderer's pool + FB     */
  uint64_t pool_slots;         /* brick slots in the pool                     */
  uint64_t resident_bricks;    /* slots holding a brick                       */
  uint64_t slot_bytes;         /* bytes per slot (tvk_brick_slot_bytes)       */
  uint64_t bricks_uploaded;    /* host->device brick uploads since create     */
  uint64_t bricks_evicted;     /* LRU evictions since create                  */
} tvk_mem_stats;

/* --------------------------------------------------------------- errors */
const char* tvk_last_error(void);
const char* tvk_version(void);

/* -------------------------------------------------------------- lifecycle */
/* tuvok.renderer.new(type, useOnlyPowerOfTwo, downSampleTo8Bits,
 * disableBorder, biasAndScaleTF) -- RenderWindowGL.cpp:97-100,
 * LuaScripts/BatchRenderer.lua:18-22.  Only TVK_OPENGL_RAYCASTER and
 * TVK_OPENGL_GRIDLEAPER are accepted; other types return TVK_ERR_ARG. */
int tvk_create(int renderer_type, int use_only_pow2, int downsample_to_8bits,
               int disable_border, int bias_and_scale_tf, tvk_renderer** out);
/* renderer.cleanup() + deleteClass (RenderWindow.cpp:1590-1602). */
int tvk_destroy(tvk_renderer* r);
/* renderer.initialize(context) -- RenderWindowGL.cpp:361-362.  The context
 * is a HIP device ordinal; creates the stream, events and framebuffer. */
int tvk_initialize(tvk_renderer* r, int device);
/* renderer.resize(UINTVECTOR2) -- RenderWindow.cpp:2446. */
int tvk_resize(tvk_renderer* r, uint32_t width, uint32_t height);
/* setRendererTarget/getRendererTarget -- RenderWindow.cpp:151-163. */
int tvk_set_renderer_target(tvk_renderer* r, int target);
int tvk_get_renderer_target(tvk_renderer* r, int* target);

/* ------------------------------------------------------------------ data */
/* renderer.loadDataset(file) -- RenderWindowGL.cpp:111.  Reads a UVF v2
 * file with an uncompressed RasterDataBlock (doc/import.adoc:904-1102),
 * verifies its MD5 when `verify` is set (UVFReader/main.cpp:67-68,151),
 * and builds the device brick pool from it. */
int tvk_load_dataset(tvk_renderer* r, const char* uvf_path, int verify);
/* In-memory replacement of loadDataset for a flat scalar volume, x fastest:
 * bricks it into the device pool with `apron` voxels of overlap per side
 * and builds `n_lod` levels by 2x2x2 integer mean (RasterDataBlock::
 * FlatDataToBrickedLOD + CombineAverage, UVFReader/DataSource.h:550-570).
 * bits is 8 or 16.  n_lod == 0 picks the RasterDataBlock rule
 * (1 + halvings until the largest dimension fits one brick,
 * UVFReader/DataSource.h:521-529). */
int tvk_set_volume(tvk_renderer* r, const void* voxels, uint64_t nx, uint64_t ny,
                   uint64_t nz, uint32_t bits, uint32_t brick_size, uint32_t apron,
                   uint32_t n_lod);
/* Out-of-core datasets on a fixed HBM brick pool (GLGridLeaper paging,
 * RenderWindowGL.cpp:153-163; the data side is UVFDataset::GetBrick,
 * doc/import.adoc:761-775).  A brick source fills one stored brick:
 * (B+2a)^3 voxels of LOD `lod`, brick (bx,by,bz), x fastest, clamp-to-edge
 * outside the level, the same contents tvk_read_brick returns; it returns 0
 * on success.  It is called from the painting thread. */
typedef int (*tvk_brick_source_fn)(void* user, uint32_t lod, uint32_t bx, uint32_t by,
                                   uint32_t bz, void* slot, uint64_t slot_bytes);
/* Paged volume from a brick source.  brick_ranges: 2 floats (min, max) per
 * brick of every LOD (LOD-major, then z, y, x) over its stored voxels, or
 * NULL (no empty-space skipping before a brick is resident).  The pool
 * holds pool_bytes / slot_bytes bricks; paint() pages bricks in on demand
 * (LRU eviction), needs more passes while bricks arrive (checkForRedraw
 * stays 1) and falls back to a coarser LOD when one view's bricks exceed
 * the pool. */
int tvk_set_paged_volume(tvk_renderer* r, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t bits,
                         uint32_t brick_size, uint32_t apron, uint32_t n_lod,
                         tvk_brick_source_fn source, void* user, const float* brick_ranges,
                         uint64_t pool_bytes);
/* Paged volume whose LOD 0 stays in caller memory (kept valid until the
 * volume is replaced); coarser levels and brick ranges are built on host
 * threads with the in-core rules. */
int tvk_set_paged_flat_volume(tvk_renderer* r, const void* voxels, uint64_t nx, uint64_t ny,
                              uint64_t nz, uint32_t bits, uint32_t brick_size, uint32_t apron,
                              uint32_t n_lod, uint64_t pool_bytes);
/* dataset.getLODLevelCount / getDomainSize(lod) -- ImageVis3D_Progress.cpp:104-117. */
int tvk_get_lod_level_count(tvk_renderer* r, uint32_t* n_lod);
int tvk_get_domain_size(tvk_renderer* r, uint32_t lod, uint64_t dims[3]);
/* dataset.getBitWidth -- RenderWindow.cpp:2124-2132 (8 or 16). */
int tvk_get_bit_width(tvk_renderer* r, uint32_t* bits);
/* dataset.getBrickCount(lod) and the brick grid -- BlockInfo.h:49-93. */
int tvk_get_brick_layout(tvk_renderer* r, uint32_t lod, uint32_t grid[3],
                         uint32_t* brick_size, uint32_t* apron);
/* Read back one LOD level as a flat volume (x fastest), bpv bytes/voxel:
 * the device pyramid that RasterDataBlock::FlatDataToBrickedLOD builds. */
int tvk_read_lod(tvk_renderer* r, uint32_t lod, void* out, uint64_t out_bytes);
/* Read back one stored brick slot including its apron ((B+2a)^3 voxels, x
 * fastest): UVFDataset::GetBrick(BrickKey) (doc/import.adoc:761-775). */
int tvk_read_brick(tvk_renderer* r, uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz,
                   void* out, uint64_t out_bytes);
/* Per-brick min/max of one LOD (MaxMinDataBlock, BlockInfo.h:255-271),
 * over the brick's stored voxels (payload + apron), in brick-grid order. */
int tvk_get_brick_minmax(tvk_renderer* r, uint32_t lod, float* mins, float* maxs,
                         uint64_t n_bricks);
/* Device bytes of one brick-pool slot for bricks of payload brick_size and
 * the given apron at 8 or 16 bits (GPUMemMan's per-brick texture size;
 * pool budgets hold pool_bytes / this many bricks).  A slot keeps the
 * (B+2a)^3 voxels in 4x4x4 blocks whose x-rows carry one extra voxel
 * (1.25x the voxels), 256-byte aligned; 0 for invalid arguments. */
uint64_t tvk_brick_slot_bytes(uint32_t brick_size, uint32_t apron, uint32_t bits);

/* ------------------------------------------------- TOC (UVF v3) volumes */
/* Bricked multi-resolution volume files: TOCBlock::FlatDataToBrickedLOD
 * (UVFReader/DataSource.h:475-491, options of ImageVis3D_Settings.cpp:
 * 516-548 and uvfconvert, UVFReader/main.cpp:51-52,86-109) and the TOC
 * queries GetLoDCount / GetMaxBrickSize / GetLODDomainSize / GetBrickCount /
 * GetBrickInfo (UVFReader/BlockInfo.h:38-96).
 *
 * A file is a UVF version-3 container (master header and one data block,
 * doc/import.adoc:904-1102) whose TOC block stores every LOD as bricks of
 * max_brick^3 voxels that include `overlap` voxels on each side (payload
 * max_brick - 2*overlap; ExtendedOctree's convention), x fastest.  LOD l+1
 * halves LOD l (ceil) by the 2x2x2 integer mean (RasterDataBlock's rule)
 * or, with use_median, the lower median of the existing children; levels
 * are added until the largest dimension fits one payload.  Voxels outside
 * the volume repeat the edge voxel (clamp_to_edge) or are 0.  Bricks are
 * compressed one by one -- compression 0 none, 1 zlib, 2 lzma (xz), 3 lz4
 * (block), 4 bzip2, uvfconvert's numbering, level 1..10 -- and stored in
 * the order of `layout`: 0 scanline, 1 Morton, 2 Hilbert, 3 random.  The
 * block payload's byte layout is this library's (ExtendedOctree's is not in
 * the reference tree).  The pyramid and the bricks are built on the GPU,
 * compression runs on host threads. */
typedef struct tvk_toc tvk_toc;
typedef struct tvk_toc_info {
  uint64_t dims[3];
  uint32_t bits, max_brick, overlap, n_lod;
  uint32_t use_median, clamp_to_edge, compression, level, layout;
  uint32_t reserved;
  uint64_t n_bricks;
  uint64_t stored_bytes;  /* bytes of all stored bricks */
  uint64_t raw_bytes;     /* the same bricks uncompressed */
} tvk_toc_info;
int tvk_toc_build(int device, const void* voxels, uint64_t nx, uint64_t ny, uint64_t nz,
                  uint32_t bits, uint32_t max_brick, uint32_t overlap, int use_median,
                  int clamp_to_edge, uint32_t compression, uint32_t level, uint32_t layout,
                  const char* path, tvk_toc_info* info);
int tvk_toc_open(const char* path, tvk_toc** out);
int tvk_toc_close(tvk_toc* t);
int tvk_toc_get_info(tvk_toc* t, tvk_toc_info* info);
/* GetLODDomainSize(lod) and GetBrickCount(lod). */
int tvk_toc_get_lod(tvk_toc* t, uint32_t lod, uint64_t dims[3], uint64_t grid[3]);
/* GetBrickInfo + the brick's voxels (max_brick^3 * bpv bytes, decompressed)
 * and its value range; any pointer may be NULL. */
int tvk_toc_read_brick(tvk_toc* t, uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz, void* out,
                       uint64_t out_bytes, uint32_t* compression, float* min_value,
                       float* max_value);
/* The file's Histogram1DDataBlock / Histogram2DDataBlock (DataSource.h:
 * 595-636; layouts doc/import.adoc:1021-1056), written by tvk_toc_build from
 * the LOD-0 voxels with the c60-pinned rules: 2^bits value bins; 256 value x
 * 256 gradient-magnitude bins (value slowest) and the maximum gradient.
 * *n_bins / *n_value, *n_grad hold the capacity of `out` on input and the
 * stored sizes on output (out may be NULL to query them); TVK_ERR_STATE
 * when the file has no such block. */
int tvk_toc_get_histogram_1d(tvk_toc* t, uint64_t* out, uint64_t* n_bins);
int tvk_toc_get_histogram_2d(tvk_toc* t, uint64_t* out, uint64_t* n_value, uint64_t* n_grad,
                             float* max_grad);
/* uvfconvert's input side (UVFReader/main.cpp:143-147,372-414): the same
 * build from a raw file of nx*ny*nz voxels (x fastest, little endian) that
 * starts `header_bytes` into the file. */
int tvk_toc_build_raw(int device, const char* raw_path, uint64_t header_bytes, uint64_t nx,
                      uint64_t ny, uint64_t nz, uint32_t bits, uint32_t max_brick, uint32_t overlap,
                      int use_median, int clamp_to_edge, uint32_t compression, uint32_t level,
                      uint32_t layout, const char* path, tvk_toc_info* info);
/* Page a TOC volume into the handle's HBM brick pool (the out-of-core
 * GridLeaper path of tvk_set_paged_volume with the TOC as brick source:
 * brick payload and apron = the TOC's payload and overlap).  `t` must stay
 * open while the volume is in use. */
int tvk_set_toc_volume(tvk_renderer* r, tvk_toc* t, uint64_t pool_bytes);
/* Histogram1DDataBlock::Compute over LOD 0 (UVFReader/DataSource.h:595-610):
 * n_bins = 2^bits (u64 counts). */
int tvk_compute_histogram_1d(tvk_renderer* r, uint64_t* hist, uint64_t n_bins);
/* Histogram2DDataBlock::Compute over LOD 0 (DataSource.h:611-636): value x
 * gradient-magnitude bins, 256 x 256 (value bin = raw value scaled to 256
 * bins, X slowest), gradient of v/(2^bits-1) by central differences.
 * max_grad receives the maximum gradient magnitude. */
int tvk_compute_histogram_2d(tvk_renderer* r, uint64_t* hist, uint32_t n_value_bins,
                             uint32_t n_grad_bins, float* max_grad);

/* ------------------------------------------------------- mode and TFs */
/* setRenderMode/getRenderMode -- RenderWindow.cpp:2051-2053. */
int tvk_set_render_mode(tvk_renderer* r, int mode);
/* TransferFunction1D colour data (Q1DTransferFunction.cpp:254-258) and
 * tuvok.gpu.changed1DTrans (Q1DTransferFunction.cpp:465-469): n RGBA floats.
 * Data value v indexes entry v * (n-1) / vmax when biasAndScaleTF is set
 * (vmax = the largest value of the volume, so the TF spans the data range
 * of e.g. 12-bit data in 16-bit voxels), otherwise entry v (clamped to n-1). */
int tvk_set_1d_trans(tvk_renderer* r, const float* rgba, uint32_t n);
/* TransferFunction2D raster (Q2DTransferFunction.cpp:426-470): w (value) x h
 * (gradient magnitude) RGBA floats, value fastest; max_grad is the gradient
 * magnitude that maps to the last row. */
int tvk_set_2d_trans(tvk_renderer* r, const float* rgba, uint32_t w, uint32_t h,
                     float max_grad);
/* setIsoValue -- RenderWindow.cpp:2117-2147 (raw data units). */
int tvk_set_iso_value(tvk_renderer* r, float iso);
/* setUseMIP / getUseMIP -- RenderWindow.cpp:281-289: maximum intensity
 * projection in 1D TF mode: each pixel is the 1D TF at the largest sample
 * value along its ray (every sample visited, no early termination), over
 * the background.  ImageVis3D offers it on its 2D views (slice regions,
 * not built here); this backend applies it to the raycast view. */
int tvk_set_use_mip(tvk_renderer* r, int use_mip);
int tvk_get_use_mip(tvk_renderer* r, int* use_mip);
/* setMIPRotationAngle -- RenderWindow.cpp:1687-1703 (CaptureMIPFrame): MIP
 * views turned by `degrees` about the camera's up vector around the volume
 * centre. */
int tvk_set_mip_rotation_angle(tvk_renderer* r, float degrees);
/* setIsosurfaceColor -- RenderWindow.cpp:2329. */
int tvk_set_isosurface_color(tvk_renderer* r, float red, float green, float blue);

/* ClearView in isosurface mode: setClearViewEnabled (RenderWindow.cpp:
 * 2348-2356), setCVIsoValue (:2155-2167), setCVSize (:2169-2180),
 * setCVContextScale (:2182-2193), setCVBorderScale (:2195-2206),
 * setCVColor (:2337-2346), setCVFocusPos (:2359-2370, the mouse position
 * while dragging, :908-912); getters getClearViewEnabled, getCVIsoValue,
 * getCVSize, getCVContextScale, getCVBorderScale, getCVColor (:326-360,
 * 2418-2422) come back together from tvk_get_clearview.  The iso-value
 * surface is the context; around the focus position it fades to its
 * silhouettes and the first crossing of the ClearView iso value behind it
 * shows, shaded in the ClearView colour:
 *   w  = clamp(1 - size * r, 0, 1), r = pixel distance from the focus in
 *        image heights (size 0: the whole image is in focus)
 *   ac = (1 - w) + w * clamp(context_scale * (1 - |N.V|)^border_scale, 0, 1)
 *   colour = ac * context + (1 - ac) * focus (no focus surface: ac *
 *   context over the background, opacity ac)
 * N = the context surface's unit gradient, V = towards the eye.  The
 * reference's ClearView shaders are not in its tree; this model is this
 * library's.  Defaults: off, iso 0, red, size 5.5, context 1, border 60,
 * focus (0, 0) -- inside the ranges of the reference's sliders
 * (ImageVis3D.cpp:543-553: size 0..10, context 0..9.9, border 0..99).
 * Sizes and scales must be >= 0. */
typedef struct tvk_clearview {
  int32_t enabled;
  float iso_value;     /* raw data units */
  float color[3];
  float size;
  float context_scale;
  float border_scale;
  int32_t focus[2];    /* pixel x, y (row 0 at the top) */
} tvk_clearview;
int tvk_set_clearview_enabled(tvk_renderer* r, int enabled);
int tvk_set_cv_iso_value(tvk_renderer* r, float iso);
int tvk_set_cv_color(tvk_renderer* r, float red, float green, float blue);
int tvk_set_cv_size(tvk_renderer* r, float size);
int tvk_set_cv_context_scale(tvk_renderer* r, float scale);
int tvk_set_cv_border_scale(tvk_renderer* r, float scale);
int tvk_set_cv_focus_pos(tvk_renderer* r, int32_t x, int32_t y);
int tvk_get_clearview(tvk_renderer* r, tvk_clearview* out);

/* ---------------------------------------------------------------- camera */
/* setViewPos/setViewDir/setUpDir -- RenderWindow.cpp:1033-1082 (world space). */
int tvk_set_view_pos(tvk_renderer* r, float x, float y, float z);
int tvk_set_view_dir(tvk_renderer* r, float x, float y, float z);
int tvk_set_up_dir(tvk_renderer* r, float x, float y, float z);
/* setFoV -- RenderWindow.cpp:2105-2115 (vertical, degrees). */
int tvk_set_fov(tvk_renderer* r, float fov_degrees);
/* setOrthoViewEnabled -- RenderWindow.cpp:1699-1700.  half_height is the
 * world-space half height of the orthographic view volume. */
int tvk_set_ortho_view(tvk_renderer* r, int enabled, float half_height);
/* renderRegion3D setRotation4x4/setTranslation4x4 -- RenderWindow.cpp:166-182. */
int tvk_set_rotation_4x4(tvk_renderer* r, const float m[16]);
int tvk_set_translation_4x4(tvk_renderer* r, const float m[16]);
/* setRescaleFactors -- RenderWindow.cpp:592-599 (voxel aspect). */
int tvk_set_rescale_factors(tvk_renderer* r, double sx, double sy, double sz);

/* ------------------------------------------------------- quality and LOD */
/* setSampleRateModifier -- RenderWindow.cpp:2090-2103: samples per voxel of
 * the LOD being rendered. */
int tvk_set_sample_rate_modifier(tvk_renderer* r, float rate);
/* setInterpolant -- RenderWindow.cpp:437-440. */
int tvk_set_interpolant(tvk_renderer* r, int interpolant);
/* setLODLimits -- RenderWindow.cpp:447-451: finest (min) and coarsest (max)
 * LOD index a paint may use (0 = full resolution). */
int tvk_set_lod_limits(tvk_renderer* r, uint32_t min_lod, uint32_t max_lod);
/* setBGColors -- RenderWindow.cpp:401-406 (vertical gradient, top/bottom). */
int tvk_set_bg_colors(tvk_renderer* r, const float top[3], const float bottom[3]);
/* setLightingEnabled / setLightColors -- RenderWindow.cpp:2076,408-415.
 * Two-sided Blinn-Phong on the TF colour c with the central-difference
 * gradient normal: c*(ambient + diffuse*|N.L|) + specular*|N.H|^s, where
 * s = specular[3] (shininess) and light_dir points towards the light in
 * world space.  Defaults: ambient 0.1, diffuse 0.9, specular 0.3, s = 16,
 * light_dir (0, 0, 1).  Lighting applies to the 1D and 2D TF modes; the
 * isosurface mode is always lit. */
int tvk_set_lighting_enabled(tvk_renderer* r, int enabled);
int tvk_set_light_colors(tvk_renderer* r, const float ambient[4], const float diffuse[4],
                         const float specular[4], const float light_dir[3]);
/* Early ray termination threshold on accumulated opacity (default 0.99). */
int tvk_set_early_ray_termination(tvk_renderer* r, float alpha_threshold);
/* Empty-space skipping on per-brick min/max (default on); the image is the
 * same either way, only the work changes. */
int tvk_set_empty_space_skipping(tvk_renderer* r, int enabled);
/* setClipPlane -- RenderWindow.cpp:1920-1927 (the region's ExtendedPlane,
 * already in world space, RenderWindow.cpp:1852-1877): plane (a, b, c, d);
 * points with a x + b y + c z + d > 0 are cut away while it is enabled.
 * enableClipPlane -- RenderWindow.cpp:2229-2251; isClipPlaneEnabled --
 * RenderWindow.cpp:613-617; getClipPlane -- RenderWindow.cpp:667-670.
 * Clipping shortens each ray to the kept half-space (samples restart at
 * the plane); nominal sample counts still use the whole box chord.  The
 * default plane (0, 0, -1, 0) keeps z >= 0; clipping starts disabled. */
int tvk_set_clip_plane(tvk_renderer* r, const float plane[4]);
int tvk_enable_clip_plane(tvk_renderer* r, int enabled);
int tvk_is_clip_plane_enabled(tvk_renderer* r, int* enabled);
int tvk_get_clip_plane(tvk_renderer* r, float plane[4]);
/* setBlendPrecision -- RenderWindow.cpp:1618-1623 (TVK_BP_*; default
 * TVK_BP_32BIT).  Below 32 bit the running composite of the 1D / 2D TF
 * modes is rounded to the target's precision (unorm8: round to nearest of
 * x*255; half float: round to nearest even) each time a ray enters another
 * brick of the rendered level and once at the end, as the per-brick blend
 * into an 8- or 16-bit target rounds it.  Isosurface and MIP frames do not
 * blend and are unaffected; so is the per-region LOD march. */
int tvk_set_blend_precision(tvk_renderer* r, int precision);
int tvk_get_blend_precision(tvk_renderer* r, int* precision);
/* setPerfMeasures -- RenderWindow.cpp:1625-1640 (applied by
 * ImageVis3D_Settings.cpp:575 with both decrease factors 2 and the LOD
 * delay in redraw-timer ticks).  The interactive frame-time governor: a new
 * interactive frame (after any change) starts at the finest allowed LOD
 * whose predicted subframe time fits 1000 / min_framerate ms -- predicted
 * from the kernel times this renderer measured per LOD, each level about
 * half the work of the next finer one -- and, with nothing measured yet or
 * nothing fitting, at the coarsest; then refines one LOD per paint.  With
 * low_res_intermediate, every subframe coarser than the finest allowed LOD
 * is rendered at width / screen_res_dec_factor x height /
 * screen_res_dec_factor and sample rate / sample_dec_factor, then enlarged
 * to the window by nearest pixel (dst(x, y) = src(x*w/W, y*h/H)).
 * lod_delay: paints that return without rendering before each refinement
 * (the user may still be interacting).  min_framerate 0 (the default) turns
 * the governor off.  Capture and headless targets always render the finest
 * allowed LOD at full resolution and rate; a capture skips the delay. */
int tvk_set_perf_measures(tvk_renderer* r, uint32_t min_framerate, int low_res_intermediate,
                          float screen_res_dec_factor, float sample_dec_factor, uint32_t lod_delay);
/* The framebuffer as the last paint left it (the subframe the reference
 * presents after each paint, RenderWindow.cpp:2453-2502), without painting
 * further.  rgba: W*H*4 bytes, row 0 at the top. */
int tvk_read_framebuffer(tvk_renderer* r, uint8_t* rgba, uint64_t rgba_bytes);

/* Kernel launch tuning (no counterpart in the reference; chosen by the
 * autotuner): samples gathered per march step (1, 2, 4 or 8; default 8:
 * the 1D march takes 2, 4 or 8 (1 as 8); the gradient marches of the 2D TF
 * and of lighting take 1 or 2 samples per lane, and any other value
 * gathers runs of 4 consecutive samples of a ray on 4 lanes (the default);
 * the other modes take 4);
 * how 8x8 ray tiles are dealt to the 8 XCDs: xcd_split 0 deals runs of
 * tile_group (1..4096) consecutive tiles round-robin, xcd_split rx in
 * {1, 2, 4, 8} gives each XCD one region of an rx x (8/rx) split of the
 * frame.  Idle XCDs steal tiles either way.  The image does not depend on
 * any of them. */
int tvk_set_kernel_tuning(tvk_renderer* r, uint32_t samples_per_batch, uint32_t tile_group,
                          uint32_t xcd_split);
/* Waves per work-group (4 or 8; default 4) of the raycast variants that do
 * not stage the 1D TF in LDS (2D TF, lighting without LDS TF, isosurface,
 * MIP, mixed-TF launches); no counterpart in the reference, chosen by the
 * autotuner (TVK_TUNE_LAUNCH).  Those variants fit 2-3 waves per SIMD, so
 * 4-wave work-groups put up to 12 waves on a CU and 8-wave ones 8: more
 * waves hide more latency on cache-friendly volumes and thrash the caches
 * on very large ones.  The image does not depend on it. */
int tvk_set_work_group_waves(tvk_renderer* r, uint32_t waves);
/* LDS staging of the 1D march (no counterpart in the reference; chosen by
 * the autotuner, TVK_TUNE_LAUNCH): with on = 1, 1D-TF launches without
 * lighting and with trilinear sampling march each wave's rays in rounds
 * over 8^3 cells of the brick slots, the voxels of each cell copied into
 * LDS by the whole wave (contiguous block loads) and read from there,
 * instead of four x-pair gathers per sample.  The image and the sample
 * counts do not depend on it.  Default 0. */
int tvk_set_brick_staging(tvk_renderer* r, int on);
/* TF groups (no counterpart in the reference): with on = 1 (the default), a
 * launch in a gradient mode (2D TF, or 1D TF with lighting) walks the rays
 * of views that differ only in their TF slot (same camera, image size,
 * volume, LOD and settings; up to 4 per group) once, gathering each
 * sample's value and gradient once and compositing every view's TF.  Each
 * view's image and sample counts equal its own launch's bit for bit; 0
 * renders every view on its own. */
int tvk_set_tf_grouping(tvk_renderer* r, int on);
/* Ray tile shape of one 64-lane wavefront (no counterpart in the
 * reference; chosen by the autotuner): tile_width x (64 / tile_width) rays,
 * tile_width in {4, 8, 16} (default 8).  The image does not depend on it. */
int tvk_set_ray_tile_shape(tvk_renderer* r, uint32_t tile_width);
/* Empty-space cell edge in voxels of the rendered LOD (no counterpart in the
 * reference): 4..brick size, a divisor of the brick size (default 16, or the
 * brick size when that is smaller).  Takes effect at the next setVolume /
 * loadDataset.  The image does not depend on it. */
int tvk_set_empty_space_cell(tvk_renderer* r, uint32_t cell);

/* ---------------------------------------------------------------- render */
/* renderer.paint() -- RenderWindow.cpp:2464, BatchRenderer.lua:35.  Renders
 * one subframe into the device framebuffer.  In RT_CAPTURE/RT_HEADLESS the
 * finest allowed LOD is rendered directly; in RT_INTERACTIVE successive
 * paints refine from the coarsest to the finest allowed LOD. */
int tvk_paint(tvk_renderer* r);
/* checkForRedraw -- RenderWindow.cpp:392,1605-1610: 1 while finer subframes
 * remain, 0 when the frame is complete. */
int tvk_check_for_redraw(tvk_renderer* r);
/* scheduleCompleteRedraw -- RenderWindow.cpp:385-393. */
int tvk_schedule_complete_redraw(tvk_renderer* r);
/* captureSingleFrame(file, preserveTransparency) readback part --
 * RenderWindow.cpp:1642-1671, BatchRenderer.lua:38.  Renders until the
 * frame is complete and copies W*H*4 bytes to `rgba`.  With
 * preserve_transparency the background is black with alpha = opacity. */
int tvk_capture_single_frame(tvk_renderer* r, uint8_t* rgba, uint64_t rgba_bytes,
                             int preserve_transparency);
/* Progress counters -- RenderWindow.cpp:453-486. */
int tvk_get_progress(tvk_renderer* r, uint64_t* subframe_count, uint32_t* working_subframe,
                     uint32_t* brick_count, uint32_t* working_brick, uint64_t* min_lod_index);
int tvk_get_frame_stats(tvk_renderer* r, tvk_frame_stats* out);
/* Device pointer of the RGBA8 framebuffer and the handle's HIP stream, for
 * callers that keep frames on the device. */
int tvk_get_device_framebuffer(tvk_renderer* r, void** dptr, void** hip_stream);
/* Block until this handle's stream has drained. */
int tvk_synchronize(tvk_renderer* r);

/* -------------------------------------------------------------- batching */
/* Batched rendering (no counterpart in the reference, which renders one
 * view per paint; BatchRenderer.lua:22-41 loops paint + capture per image).
 * Many views of the handle's volume -- each with its own camera, image size
 * and transfer-function slot -- are rendered by one persistent kernel
 * launch whose waves drain the ray tiles of all views; all other state
 * (render mode, lighting, sampling rate, background, region transforms,
 * finest LOD limit) is the handle's.  Each view's image equals what
 * tvk_capture_single_frame gives for the same camera and TF, bit for bit. */
typedef struct tvk_view {
  float view_pos[3];
  float view_dir[3];
  float up_dir[3];
  float fov_deg;            /* perspective vertical field of view        */
  int32_t ortho;            /* 1: orthographic with ortho_half_height     */
  float ortho_half_height;
  uint32_t width, height;
  uint32_t tf_slot;         /* transfer function slot (0 = setXDTrans)    */
  uint8_t* rgba;            /* host RGBA8 output (W*H*4 bytes) or NULL     */
} tvk_view;

typedef struct tvk_view_stats {
  uint64_t evaluated_samples;
  uint64_t nominal_samples;
  uint32_t lod;
  uint32_t reserved;
  /* samples whose voxels the view's march gathered: its evaluated samples
   * when it is marched alone; for a TF group (tvk_set_tf_grouping) the
   * group's shared walk, counted on its first view and 0 on the others */
  uint64_t gathered_samples;
} tvk_view_stats;

/* Transfer-function slots 0..63 for batched views (slot 0 is the one
 * tvk_set_1d_trans / tvk_set_2d_trans set); same conventions as those. */
int tvk_set_1d_trans_slot(tvk_renderer* r, uint32_t slot, const float* rgba, uint32_t n_entries);
int tvk_set_2d_trans_slot(tvk_renderer* r, uint32_t slot, const float* rgba, uint32_t width,
                          uint32_t height, float max_grad);
/* Render n views (any n; up to 64 share one launch).  `stats` (may be NULL)
 * receives n entries; `kernel_ms` (may be NULL) the summed raycast launch
 * time.  Views with rgba == NULL stay on the device
 * (tvk_get_batch_framebuffer).  Paged volumes are paged in as for a
 * capture. */
int tvk_render_batch(tvk_renderer* r, const tvk_view* views, uint32_t n, tvk_view_stats* stats,
                     float* kernel_ms);
/* Device pointer of view `index` of the last batch (RGBA8, row 0 at top). */
int tvk_get_batch_framebuffer(tvk_renderer* r, uint32_t index, void** dptr);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  tvk_render_batch for n views rendered into
 * caller-provided device framebuffers (device_rgba[i]: width*height*4 bytes
 * of RGBA8, row 0 at the top; views[i].rgba is ignored) on the renderer's
 * stream, allocating nothing: the launch's job table and counters live in a
 * fixed device scratch array of 64 KiB per GPU, which each GPU the library
 * runs on holds until the process ends.  Limits: n <= 64 (TVK_ERR_ARG
 * beyond); in-core volumes; no per-region LOD; every TF slot the views use
 * must already be on the device for the current render mode (render it once
 * with tvk_render_batch or tvk_paint), else TVK_ERR_STATE.  stats NULL:
 * returns once the work is queued (tvk_synchronize waits for it); else
 * waits and fills the per-view sample counts.  Every call on a GPU shares
 * that GPU's scratch, which keeps the last call's table: calls whose work
 * overlaps on one GPU (other threads, or other renderers' streams before
 * their work is done) race. */
int tvk_render_batch_noalloc(tvk_renderer* r, const tvk_view* views, uint32_t n, void* const* device_rgba,
                             tvk_view_stats* stats);
/* ---------------------------------------------------------- per-region LOD */
/* Per-region LOD (no direct counterpart in the reference, whose LOD choice
 * is per frame: setLODLimits, RenderWindow.cpp:447-451; this is what the
 * learned LOD selector drives).  levels[r] is the LOD level used for LOD-0
 * brick region r (x fastest, one entry per LOD-0 brick); the ray samples
 * that region from that level with step 2^level / rate, its segments
 * restarting at the region entry.  Applies to the 1D TF without lighting;
 * n = 0 turns it off. */
int tvk_set_region_lods(tvk_renderer* r, const uint8_t* levels, uint64_t n);
/* With region output on, each paint also records per pixel the LOD-0
 * region that added the most opacity (-1: none); read it back with
 * tvk_read_region_output (W*H int32). */
int tvk_set_region_output(tvk_renderer* r, int enabled);
int tvk_read_region_output(tvk_renderer* r, int32_t* out, uint64_t n);
/* Per-frame check of per-region LOD frames (no counterpart in the
 * reference; it holds each learned-LOD frame to an error bound at run
 * time).  With tile_sample = k >= 2, every paint of a per-region frame also
 * renders one ray tile in each run of k at level 0 into the frame (those
 * tiles become exact), estimates the RGB RMSE of the rest of the frame
 * against the all-level-0 frame as sqrt((k-1) * SSE / (3 * W * H)) in
 * 1/255 units (SSE: the sampled tiles' squared differences in 8-bit
 * units) and, above max_rmse, renders the whole frame at level 0.  The
 * paint stays asynchronous: the next call that needs the frame (paint,
 * capture, read/device framebuffer, frame stats, synchronize, resize)
 * reads the estimate and queues the fallback.  The sampled tiles' offsets
 * change from frame to frame and start over at each call of
 * tvk_set_region_check.  In-core volumes only
 * (TVK_ERR_STATE at paint on a paged one).  tile_sample 0 turns it off;
 * 1 or > 4096 is TVK_ERR_ARG. */
typedef struct tvk_region_check {
  float rmse_estimate;     /* the last checked frame's estimate (1/255)  */
  uint32_t fell_back;      /* 1: that frame was rendered at level 0      */
  float check_ms;          /* device time of the check and any fallback  */
  uint32_t reserved;
  uint64_t check_samples;  /* samples of the check's level-0 tiles       */
} tvk_region_check;
int tvk_set_region_check(tvk_renderer* r, uint32_t tile_sample, float max_rmse);
int tvk_get_region_check(tvk_renderer* r, tvk_region_check* out);

/* Per-brick value histograms of LOD level `lod` (payload voxels, `bins` =
 * 2..64 equal bins over [0, 2^bits)): n_bricks x bins counts, bricks in
 * page-table order (the learned-LOD features; in-core volumes). */
int tvk_brick_histograms(tvk_renderer* r, uint32_t lod, uint32_t bins, uint32_t* out, uint64_t n);

/* Per LOD-0 brick region and LOD level l: mean |v0 - v_l| between each
 * LOD-0 voxel and its level-l ancestor (x >> l, y >> l, z >> l) -- the
 * detail a region loses at level l (learned-LOD feature; in-core volumes).
 * out: n_regions x n_lod floats (level 0 = 0), regions x fastest. */
int tvk_region_detail(tvk_renderer* r, float* out, uint64_t n);

/* Learned-LOD error predictor (no counterpart in the reference): a bf16 MLP
 * 32 inputs -> 4 hidden layers of 64 (ReLU) -> 1 output evaluated on the
 * MFMA units with FP32 accumulation.  weights_bf16: layer-major, each layer
 * [in][out] row-major (32x64, then 3 x 64x64), then the 64 output weights
 * (32*64 + 3*64*64 + 64 values, bfloat16 bits); biases: 4 x 64 + 1 floats.
 * predict: n rows of 32 floats (host) -> n floats (host); kernel_ms (may be
 * NULL) gets the kernel time. */
typedef struct tvk_lod_mlp tvk_lod_mlp;
int tvk_lod_mlp_create(int device, const uint16_t* weights_bf16, const float* biases,
                       tvk_lod_mlp** out);
int tvk_lod_mlp_predict(tvk_lod_mlp* m, const float* inputs, uint64_t n, float* outputs,
                        float* kernel_ms);
int tvk_lod_mlp_destroy(tvk_lod_mlp* m);

/* Learned-LOD frame planning on the host (Tuvok's per-frame LOD choice,
 * RenderWindow.cpp:447-451, made per LOD-0 brick region).  Both run on
 * host threads and touch no device.
 *
 * tvk_lod_features: per LOD-0 region r (x fastest over the brick grid
 * ceil(dims / edge)) its projected footprint in pixels (screen box of the
 * region's corners, clipped to the view; 0 behind the camera) and centre
 * distance from the eye, then the predictor's 32 inputs for every coarser
 * level l = 1 .. n_lod-1 into x[((l-1) * R + r) * 32 + k]:
 *   l/8, log2(1 + footprint)/24, hist[r][0..16), moments[r][0..8),
 *   distance, detail[r][l], detail[r][l] * moments[r][5], zeros.
 * detail may be NULL (inputs 27-28 zero); dist_out may be NULL.  The view
 * is perspective (ortho views are rejected); world box centred at 0 with
 * largest extent 1. */
int tvk_lod_features(const tvk_view* view, const uint64_t dims[3], uint32_t edge, uint32_t n_lod,
                     const double* hist16, const double* moments8, const double* detail,
                     float* x, double* area_out, double* dist_out);

/* tvk_lod_select_octree: octree cut of the brick pyramid from predicted
 * log mean-squared errors pred[(l-1) * R + r] (levels 1 .. n_levels-1):
 * start from the coarsest level's bricks and repeatedly split the brick
 * whose children cut the predicted SSE (sum over its regions of
 * exp(pred) x footprint) the most per added brick, ties to the lower
 * (level, z, y, x), until the SSE fits budget_sse or the next split would
 * pass mem_budget bytes at brick_bytes per brick (0: no cap).  Writes one
 * level per region (x fastest) and the predicted SSE and brick bytes. */
int tvk_lod_select_octree(const float* pred, const double* area, const uint32_t grid[3],
                          uint32_t n_levels, double budget_sse, uint64_t brick_bytes,
                          uint64_t mem_budget, uint8_t* levels, double* sse_out,
                          uint64_t* bytes_out);

/* ------------------------------------------------------------ autotuner */
/* Launch autotuner (no counterpart in the reference).  Times the given
 * views (tvk_render_batch, images stay on the device) under candidate
 * settings by coordinate descent and leaves the fastest applied:
 *   TVK_TUNE_BRICK  brick size in {32, 64, 128, 256} (in-core volumes
 *                   and paged volumes set from flat voxels, re-bricked
 *                   from them), tried with the starting cells and again
 *                   with the tuned ones,
 *   TVK_TUNE_CELL   empty-space cell edge (divisors of the brick size),
 *   TVK_TUNE_LAUNCH samples per batch, ray tile width, XCD tile group.
 * Cell and launch settings leave the images bit-identical; a different
 * brick size changes rounding in the slot-local coordinates (tested:
 * at most 1/255 per channel). */
enum { TVK_TUNE_BRICK = 1, TVK_TUNE_CELL = 2, TVK_TUNE_LAUNCH = 4 };
typedef struct tvk_tuning {
  uint32_t brick_size;
  uint32_t cell;
  uint32_t samples_per_batch;
  uint32_t tile_width;
  uint32_t tile_group;
  uint32_t xcd_split;
  float kernel_ms;          /* best summed launch time over the views   */
  float baseline_ms;        /* the same views under the starting settings */
  uint32_t candidates;      /* settings timed                            */
  uint32_t wg_waves;        /* tvk_set_work_group_waves                  */
  uint32_t brick_staging;   /* tvk_set_brick_staging                     */
} tvk_tuning;
int tvk_autotune(tvk_renderer* r, const tvk_view* views, uint32_t n, uint32_t flags,
                 tvk_tuning* out);

/* --------------------------------------------------------------- serving */
/* Render server (no counterpart in the reference, where each paint serves
 * one view of one renderer): a worker thread batches requests as they
 * arrive into persistent-kernel launches over several volumes and overlaps
 * launches on up to `lanes` HIP streams.  In-core volumes keep their bricks
 * resident; their requests are admitted when a launch lane and output
 * memory (device arena + pinned host arena of `arena_bytes` each) are free.
 * Requests that share a kernel variant (bytes per voxel, interpolant,
 * render mode, lighting) share a launch, oldest first, up to 64 requests
 * and -- after the first -- about 1e9 estimated samples (tile-rect rays x
 * the volume's longest axis in steps), so light requests are not held
 * until a heavy one finishes.  Paged volumes (tvk_set_paged_volume and
 * friends) page through the device-wide brick pool: each has a worker that
 * renders its requests in batches of up to 64 of one render mode with the
 * renderer's multi-pass paging, on the renderer's own streams, beside the
 * lanes; a batch is admitted by brick-pool memory -- when its working set
 * (the bricks of the finest allowed LOD, at most the volume's pool share)
 * fits the pool beside the paged batches already running (GPU budget less
 * the device memory held outside the pool); one batch alone always runs.
 * Bricks stay resident between batches until the pool needs their slots.
 * A request's image is the same whether it runs alone or in a mix.
 * Registered renderers must not be changed or used by other threads while
 * the server runs. */
typedef struct tvk_server tvk_server;
typedef struct tvk_request {
  uint32_t volume;        /* id from tvk_server_add_volume                */
  int32_t render_mode;    /* TVK_RM_1DTRANS or TVK_RM_2DTRANS             */
  tvk_view view;          /* camera, size, TF slot; rgba = host output    */
} tvk_request;
typedef struct tvk_request_result {
  int32_t status;         /* TVK_OK or the error of this request          */
  uint32_t launch_jobs;   /* jobs in the launch that rendered it          */
  uint64_t evaluated_samples;
  uint64_t nominal_samples;
  double queue_ms;        /* submit -> launch (of its first part)         */
  double latency_ms;      /* submit -> image in host memory               */
  uint32_t parts;         /* launches it was rendered in (1: not split)   */
  uint32_t reserved_;
} tvk_request_result;
typedef struct tvk_server_stats {
  uint64_t submitted;
  uint64_t completed;
  uint64_t launches;
  uint64_t max_jobs_per_launch;
  uint64_t max_lanes_busy;
  uint64_t nominal_samples;
  uint64_t paged_launches;    /* batches of paged volumes rendered            */
  uint64_t pool_waits;        /* batches that waited for device pool memory   */
  uint64_t max_pool_reserved; /* largest pool bytes held by running batches   */
  uint64_t split_requests;    /* requests rendered in parts                   */
  uint64_t parts_launched;    /* jobs launched (one per part of a request)    */
} tvk_server_stats;
int tvk_server_create(int device, uint64_t arena_bytes, uint32_t lanes, tvk_server** out);
int tvk_server_destroy(tvk_server* s);
int tvk_server_add_volume(tvk_server* s, tvk_renderer* r, uint32_t* volume_id);
int tvk_server_submit(tvk_server* s, const tvk_request* req, uint64_t* ticket);
/* Wait up to timeout_ms (0 = forever) for a request; on completion copies
 * its image to req.view.rgba (when not NULL), fills `out` and retires the
 * ticket.  Returns TVK_ERR_STATE on time-out (the ticket stays valid). */
int tvk_server_wait(tvk_server* s, uint64_t ticket, uint32_t timeout_ms, tvk_request_result* out);
int tvk_server_get_stats(tvk_server* s, tvk_server_stats* out);
/* Requests of more than `samples` estimated samples (as above) render in
 * parts: their tile slots split evenly over ceil(estimate / samples)
 * launches (at most 64, and at most one per tile), one part on the GPU at
 * a time, so that the requests behind a heavy one run between its parts;
 * the image is copied out after the last part.  Images and sample counts
 * do not depend on it.  0 restores the default, 4e9 (about 20 ms of
 * raycasting per part); UINT64_MAX never splits.  Applies to requests
 * submitted afterwards. */
int tvk_server_set_part_cost(tvk_server* s, uint64_t samples);

/* Replay batched launches from captured HIP graphs (default on): table
 * upload, counter clear, raycast and counter readback are one graph per
 * kernel variant, re-captured when the variant changes.  Off runs the same
 * sequence eagerly.  The images do not depend on it. */
int tvk_set_launch_graphs(tvk_renderer* r, int enabled);

/* ---------------------------------------------------------------- memory */
/* One memory manager per device, shared by every renderer and render
 * server on it (Tuvok's single GPUMemMan with SysInfo's budgets):
 *  - a GPU budget over all device memory the library allocates there (an
 *    allocation beyond it first frees paged-pool chunks, then fails with
 *    TVK_ERR_OOM); default 90% of the device;
 *  - one paged brick pool: paged volumes of all renderers take slots from
 *    it, chunk by chunk within the budget; when full, the least recently
 *    sampled brick of any volume is paged out (never one its own volume
 *    sampled in the frame being completed), and a volume's pool_bytes caps
 *    its own share;
 *  - a CPU budget over host memory held for bricks (staging buffers, the
 *    coarser levels of a flat paged volume, and a host cache of bricks
 *    produced by file or callback sources, so that a brick paged out and
 *    needed again is not read and decompressed again); default half of the
 *    host's memory.
 * SysInfo()->SetMaxUsableGPUMem / SetMaxUsableCPUMem + MemMan()->
 * MemSizesChanged() (ImageVis3D_Settings.cpp:556-559): set the budgets of
 * the renderer's device. */
int tvk_set_max_usable_gpu_mem(tvk_renderer* r, uint64_t bytes);
int tvk_set_max_usable_cpu_mem(tvk_renderer* r, uint64_t bytes);
/* MemMan()->GetGPUMem/GetAllocatedGPUMem (ImageVis3D_DebugWindow.cpp:172-206):
 * this renderer's share (its buffers and its slots in the device pool). */
int tvk_get_mem_stats(tvk_renderer* r, tvk_mem_stats* out);
/* The device's memory manager as a whole: GetGPUMem, GetAllocatedGPUMem,
 * GetCPUMem, GetAllocatedCPUMem (ImageVis3D_DebugWindow.cpp:172-206) plus
 * the paged pool and host cache.  Valid once a renderer has initialized
 * on the device. */
typedef struct tvk_memman_stats {
  uint64_t gpu_mem, gpu_budget, gpu_allocated;   /* device total, budget, allocated */
  uint64_t cpu_mem, cpu_budget, cpu_allocated;   /* host total, budget, held for bricks */
  uint64_t pool_bytes, pool_slots, pool_resident; /* paged pool chunks, slots, slots in use */
  uint64_t pool_evictions, chunks_freed;
  uint64_t cache_bytes, cache_hits, cache_misses; /* host brick cache */
} tvk_memman_stats;
int tvk_get_memman_stats(int device, tvk_memman_stats* out);
/* Threads that run a paged volume's brick source in parallel (default 8;
 * 1 for a source that must not be called concurrently).  No counterpart in
 * the reference (its bricks load on the rendering thread). */
int tvk_set_brick_source_threads(tvk_renderer* r, uint32_t threads);
/* Paging since the volume was set: bytes of bricks streamed into the pool
 * and the host milliseconds spent producing and issuing them (sources,
 * cache, copies).  No counterpart in the reference (measurement). */
int tvk_get_paging_stats(tvk_renderer* r, uint64_t* bytes, double* ms);

/* ------------------------------------------------------ synthetic data */
/* Seeded synthetic scalar volumes: UVFReader's GenerateVolumeData
 * (UVFReader/DataSource.h:335-396, kinds CT_CONST_VALUE / CT_RANDOM /
 * CT_SPHERE at :30-35) plus the benchmark workloads.  Host code only (no
 * device needed); x fastest, bits 8 or 16.  For TVK_VOL_CONST the value is
 * seed modulo 2^bits. */
enum tvk_volume_kind {
  TVK_VOL_CONST = 0,
  TVK_VOL_RANDOM = 1,
  TVK_VOL_SPHERE = 2,
  TVK_VOL_BLOBS = 3,          /* 60 Gaussian blobs on a sphere + noise     */
  TVK_VOL_PERLIN_SPHERES = 4, /* 5-octave Perlin fBm in [0,4095] + spheres */
  /* generated on the device (tvk_generate_volume_gpu,
   * tvk_set_generated_volume): */
  TVK_VOL_FBM_PLANES = 5,     /* value-noise fBm + 12 layered slabs        */
  TVK_VOL_NOISE_TUBES = 6,    /* multi-scale noise + 64 thin tubes          */
  TVK_VOL_FBM_SPHERES = 7     /* fBm + 20 solid spheres                     */
};
int tvk_generate_volume(int kind, uint64_t seed, uint64_t nx, uint64_t ny, uint64_t nz,
                        uint32_t bits, void* out, uint64_t out_bytes);
/* Device-generated kinds (5..7) into host memory, made on `device` in z
 * slabs and copied back. */
int tvk_generate_volume_gpu(int device, int kind, uint64_t seed, uint64_t nx, uint64_t ny,
                            uint64_t nz, uint32_t bits, void* out, uint64_t out_bytes);
/* setVolume with a device-generated volume: level 0 is generated straight
 * into HBM (no host copy), then downsampled and bricked as tvk_set_volume. */
int tvk_set_generated_volume(tvk_renderer* r, int kind, uint64_t seed, uint64_t nx, uint64_t ny,
                             uint64_t nz, uint32_t bits, uint32_t brick_size, uint32_t apron,
                             uint32_t n_lod);

#ifdef __cplusplus
}
#endif

#endif /* TVK_HIP_H */
