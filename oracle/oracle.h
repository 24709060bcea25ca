/*
 * oracle.h -- CPU restatement of Tuvok's raycasting + brick/LOD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or as the timed CPU baseline -- never as the product path.
 *
 * Reference anchors (the Tuvok sources themselves are an un-vendored
 * submodule, absent from the reference tree, so the GLSL raycaster cannot be
 * compiled or run; see the notes on each function):
 *  - render model: the GLRaycaster parameters exposed through Lua
 *    (ImageVis3D/UI/RenderWindow.cpp:2090-2103 sample rate, :2051-2059 render
 *    mode, :437-440 interpolant, :1033-1082 camera, :166-182 region matrices,
 *    :401-406 background, :1642-1671 capture) -- the shader maths is
 *    "parity unpinned" against the reference and pinned here by closed
 *    forms and invariants instead (tests/test_oracle_render.py).
 *  - bricking / LOD: UVFReader/DataSource.h:502-570 (brick size, overlap,
 *    2x LOD decimation, LOD count rule), CombineAverage rounding unpinned.
 *  - histograms: UVFReader/DataSource.h:595-636, pinned exactly by the
 *    blocks stored in DemoData/c60.uvf.
 */
#ifndef TVK_ORACLE_H
#define TVK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_render_params {
  /* volume actually sampled (one LOD level), x fastest */
  const void* voxels;
  int32_t bits;            /* 8 or 16 */
  int32_t nx, ny, nz;      /* dims of the sampled level */
  int32_t n0x, n0y, n0z;   /* LOD-0 dims: they define the world box */
  int32_t lod;             /* level index: step = 2^lod / rate LOD-0 voxels */
  double rescale[3];       /* setRescaleFactors */
  /* 1D transfer function, n_tf RGBA floats; entry = value * tf_scale */
  const float* tf;
  int32_t n_tf;
  double tf_scale;
  /* camera (world space), region matrix (row-vector convention) */
  double view_pos[3], view_dir[3], up_dir[3];
  double fov_deg;
  int32_t ortho;
  double ortho_half_height;
  double model[16];
  /* image + quality */
  int32_t width, height;
  double rate;             /* samples per voxel of the sampled level */
  double term_alpha;       /* early ray termination threshold; >1 disables */
  int32_t interp;          /* 0 nearest, 1 trilinear */
  double bg_top[3], bg_bottom[3], bg_alpha;
  /* brick-granular empty-space skipping (sample counting only) */
  int32_t brick_size;      /* payload B of the sampled level, 0 = no bricks */
  int32_t apron;
  /* render mode: 0 = 1D TF, 1 = 2D TF (value x gradient magnitude),
   * 2 = isosurface (first hit), 3 = maximum intensity projection (the
   * largest sample value along the ray through the 1D TF, over the
   * background; every sample visited, no early termination) */
  int32_t mode;
  const float* tf2d;       /* tf2d_w x tf2d_h RGBA floats, value fastest */
  int32_t tf2d_w, tf2d_h;
  double max_grad;         /* gradient magnitude of the last TF row */
  /* Phong lighting (modes 0/1 when lighting != 0; always for mode 2) */
  int32_t lighting;
  double ambient[3], diffuse[3], specular[3], shininess;
  double light_dir[3];     /* world space, towards the light */
  /* isosurface */
  double iso_value;        /* raw data units */
  double iso_color[3];
  /* ClearView (mode 2 with cv_enabled != 0).  The iso_value surface is the
   * context; inside a screen-space focus around cv_focus it turns
   * transparent except at its silhouettes and reveals the focus surface,
   * the first cv_iso_value crossing at or after the context hit:
   *   w  = clamp(1 - cv_size * r, 0, 1), r = |pixel - cv_focus| / height
   *   ac = (1 - w) + w * clamp(cv_context_scale * (1 - |N.V|)^cv_border_scale, 0, 1)
   *   C  = ac * context + (1 - ac) * focus (A = 1); no focus hit: C = ac *
   *   context, A = ac
   * with N the context's unit gradient and V = -ray.  This is this
   * library's model (the reference's ClearView shaders are absent). */
  int32_t cv_enabled;
  double cv_iso_value;     /* raw data units */
  double cv_color[3];
  double cv_focus[2];      /* pixel (x, y), row 0 at the top */
  double cv_size;
  double cv_context_scale;
  double cv_border_scale;
  /* Clip plane (setClipPlane, RenderWindow.cpp:1920-1927), world space:
   * points with a x + b y + c z + d > 0 are cut away; each ray's interval
   * is clipped to the kept half-space and its samples start there. */
  int32_t clip_enabled;
  double clip_plane[4];
  /* Blend precision (setBlendPrecision, RenderWindow.cpp:1618-1623): 8, 16
   * or 32.  Below 32 the running composite (C, A) of modes 0 / 1 is rounded
   * to unorm8 (floor(x * 255 + 0.5) / 255) or to binary16 (round to nearest
   * even) whenever the next sample lies in another brick of edge
   * blend_brick (the level's brick payload) and once at the end. */
  int32_t blend_bits;
  int32_t blend_brick;
} oracle_render_params;

/* The blend target's rounding of one value (bits 8 or 16; 32: unchanged). */
double oracle_blend_quantize(double x, int32_t bits);

/* FP64 reference raycast.  out_rgba: W*H*4 doubles (composited over the
 * background, [0,1]); out_samples: per-pixel evaluated samples (may be NULL);
 * out_flags: per-pixel bits, 1 = the early-termination decision was within
 * 1e-5 of the threshold, 2 = a nearest-neighbour sample lay within 1e-4 of
 * a voxel face, 4 = an isosurface sample lay within 1e-2 of the iso value
 * (may be NULL).  Returns 0 on success.
 *
 * Gradient (2D TF axis and lighting): central differences of the trilinear
 * interpolant one level voxel apart, g_a = (f(q+e_a) - f(q-e_a)) /
 * (2 (2^bits-1)) * N_l/N_0 (units: normalised value per LOD-0 voxel).
 * Lighting: two-sided Blinn-Phong, c (ka + kd |N.L|) + ks |N.H|^shininess,
 * N = g / sqrt(|g|^2 + 1e-8), V = -ray, L = light direction, H = (L+V)/|L+V|,
 * all in voxel space.  Isosurface: first sample with value >= iso, hit
 * refined linearly against the previous sample, shaded iso colour, opaque. */
/* Per-region LOD render (1D TF): levels[l] with level_dims[3l..3l+2],
 * region_lod per LOD-0 region of edge region_edge (x fastest); p supplies
 * camera, TF and quality (its voxels / dims / lod are not used). */
int oracle_render_regions(const oracle_render_params* p, int32_t n_levels,
                          const void* const* levels, const int32_t* level_dims,
                          int32_t region_edge, const uint8_t* region_lod, double* out_rgba,
                          uint32_t* out_samples, uint8_t* out_flags);
int oracle_render_1d(const oracle_render_params* p, double* out_rgba,
                     uint32_t* out_samples, uint8_t* out_flags);

/* Nominal samples: sum over pixels of floor(chord * rate) through the
 * LOD-0 box (the LOD-independent work unit of the throughput metric). */
uint64_t oracle_nominal_samples(const oracle_render_params* p);

/* FP32 multi-threaded CPU baseline of the same model (OpenMP over 8x8 pixel
 * tiles, brick-granular empty-space skipping).  out_rgba8: W*H*4 bytes.
 * Returns evaluated samples. */
uint64_t baseline_render_1d_f32(const oracle_render_params* p, uint8_t* out_rgba8,
                                int32_t n_threads);

/* LOD level l+1 from level l: each output voxel is floor(sum / count) over
 * its existing 2x2x2 children (count = 8 inside, fewer at odd edges).
 * Output dims are ceil(n/2). */
void oracle_downsample(const void* src, int32_t bits, int32_t nx, int32_t ny, int32_t nz,
                       void* dst);
/* Lower median of the existing children of each 2x2x2 group (TOC builder's
 * median option). */
void oracle_downsample_median(const void* src, int32_t bits, int32_t nx, int32_t ny, int32_t nz,
                              void* dst);

/* Copy brick (bx,by,bz) of payload B with `apron` voxels per side into a
 * (B+2a)^3 slot; voxels outside the volume replicate the nearest edge
 * voxel (clamp-to-edge). */
void oracle_extract_brick(const void* src, int32_t bits, int32_t nx, int32_t ny, int32_t nz,
                          int32_t B, int32_t apron, int32_t bx, int32_t by, int32_t bz,
                          void* slot);

/* min/max over a brick's stored slot (payload + apron). */
void oracle_brick_minmax(const void* src, int32_t bits, int32_t nx, int32_t ny, int32_t nz,
                         int32_t B, int32_t apron, int32_t bx, int32_t by, int32_t bz,
                         double* mn, double* mx);

/* Histogram1DDataBlock: counts per raw value, n_bins = 2^bits. */
void oracle_histogram_1d(const void* src, int32_t bits, uint64_t n_voxels, uint64_t* hist);

/* Histogram2DDataBlock: gradient = central differences of v/(2^bits-1)
 * ((f[i+1]-f[i-1])/2 inside, one-sided f[1]-f[0] at the borders), in
 * float; bin_y = floor(|g| / max|g| * (n_grad-1)); bin_x = value scaled to
 * n_value bins.  hist is n_value x n_grad, X slowest.  Returns max|g|. */
float oracle_histogram_2d(const void* src, int32_t bits, int32_t nx, int32_t ny, int32_t nz,
                          uint32_t n_value, uint32_t n_grad, uint64_t* hist);

#ifdef __cplusplus
}
#endif
#endif
