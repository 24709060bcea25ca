// kernels_raycast.hip -- ray-brick traversal + front-to-back compositing
// (the work of Tuvok's GLRaycaster / GLGridLeaper fragment shaders,
// reached through AbstrRenderer::Paint, ImageVis3D/UI/RenderWindow.cpp:2464).
//
// Mapping to CDNA4
//  - One 64-lane wavefront renders one 8x8 ray tile, so the 64 rays of a
//    wave march through neighbouring voxels and their gathers share L1/L2
//    lines.
//  - The kernel is persistent: one grid of work-groups fills the chip and
//    every wave pulls tiles from per-XCD queues (work-group w runs on XCD
//    w%8).  Tiles are cut into runs of `tile_group` neighbours and run g
//    belongs to XCD g%8, so each XCD's L2 sees neighbouring rays while all
//    XCDs get an even share of the frame; a wave whose own queue is empty
//    steals from the other XCDs' queues, which absorbs the long-ray tail.
//    Only the tiles inside the volume box's projected rectangle enter the
//    queues (a plain fill kernel writes the other pixels' background: each
//    queue ticket is a device-scope atomic, and a frame's empty tiles used
//    to be dealt at the atomics' rate), and their rows are dealt
//    centre-out, long rays first.
//  - The 1D TF is staged once per work-group into LDS (up to 4096 RGBA
//    entries = 64 KB), so the two TF reads per sample are LDS reads and the
//    vector-memory path only carries voxel gathers.
//  - Each ray walks a grid of march units of the chosen LOD: empty-space
//    cells (default 16^3 voxels, always inside one brick) when skipping is
//    on, whole bricks otherwise.  A cell is skipped when the TF is
//    transparent over the value range its samples can read (exact: such
//    samples add nothing); otherwise the ray looks up the brick's pool slot
//    in the page table and marches the cell's samples with trilinear
//    fetches from the slot, whose apron keeps every fetch brick-local.
//    Samples are gathered BATCH at a time so a wave keeps several gathers in
//    flight.  Rays stop at the early-termination opacity.
//
// The sample positions, segment lengths, opacity correction and blending
// follow the model stated in oracle/oracle.c (the FP64 oracle it is tested
// against); positions are global per ray, so the image does not depend on
// the brick size.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace tvk {
namespace {

constexpr float kInf = __builtin_huge_valf();
constexpr int kWaves = 8;                  // waves per work-group
constexpr int kThreads = 64 * kWaves;
constexpr int kLdsTfEntries = 4096;        // 64 KB of float4

__device__ inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ inline int clampi(int x, int lo, int hi) { return min(max(x, lo), hi); }

__device__ inline uint32_t pack_rgba8(float r, float g, float b, float a) {
  const uint32_t R = (uint32_t)floorf(clampf(r, 0.f, 1.f) * 255.f + 0.5f);
  const uint32_t G = (uint32_t)floorf(clampf(g, 0.f, 1.f) * 255.f + 0.5f);
  const uint32_t B = (uint32_t)floorf(clampf(b, 0.f, 1.f) * 255.f + 0.5f);
  const uint32_t A = (uint32_t)floorf(clampf(a, 0.f, 1.f) * 255.f + 0.5f);
  return R | (G << 8) | (B << 16) | (A << 24);
}

// Relaxed device-scope atomic add on global memory.  Called by every lane
// of a wave with a wave-uniform address, it is reduced by the compiler to
// one atomic per wave; each lane gets the old value plus the sum of the
// lower lanes' operands.
typedef __attribute__((address_space(1))) unsigned long long global_u64;
__device__ inline unsigned long long wave_atomic_add(unsigned long long* p, unsigned long long v) {
  return __hip_atomic_fetch_add((global_u64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Voxel loads go through global-address-space pointers: the pool pointer
// arrives in a generic pointer field of the job table, and a generic load
// would compile to a FLAT instruction, which also counts against the LDS
// wait counter that the TF lookups in LDS wait on.
typedef __attribute__((address_space(1))) const uint8_t global_u8;
typedef __attribute__((address_space(1))) const uint16_t global_u16;
typedef __attribute__((address_space(1))) const uint16_t __attribute__((aligned(1))) global_u16_a1;
typedef __attribute__((address_space(1))) const uint32_t __attribute__((aligned(2))) global_u32_a2;

// Load / store of one element through a global pointer.
template <typename T>
__device__ inline T gload(const T* p) {
  return *(__attribute__((address_space(1))) const T*)p;
}
template <typename T>
__device__ inline void gstore(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}

__device__ inline float ld(const uint8_t* p) { return (float)*(global_u8*)p; }
__device__ inline float ld(const uint16_t* p) { return (float)*(global_u16*)p; }

// Two x-neighbours (p[0], p[1]) with one load: a dword for 16-bit voxels, a
// short for 8-bit ones (unaligned global loads are legal on gfx950), which
// halves the vector-memory instructions of a trilinear fetch.
__device__ inline void ld2(const uint16_t* p, float& a, float& b) {
  const uint32_t v = *(global_u32_a2*)p;
  a = (float)(v & 0xffffu);
  b = (float)(v >> 16);
}
__device__ inline void ld2(const uint8_t* p, float& a, float& b) {
  const uint16_t v = *(global_u16_a1*)p;
  a = (float)(v & 0xffu);
  b = (float)(v >> 8);
}

// One sample from a brick slot (blocked slot layout) at slot-local texel
// coordinate (gx, gy, gz): trilinear over the 8 surrounding voxels -- four
// loads of an x-pair, one per (y, z) row -- or the nearest voxel.  Indices
// are clamped to the slot, so no fetch leaves the slot.
template <typename T, bool LINEAR>
__device__ inline float fetch(const T* __restrict__ base, const SlotGeo& G, float gx, float gy,
                              float gz) {
  const int S = G.S;
  if (LINEAR) {
    const int ix = clampi((int)floorf(gx), 0, S - 2);
    const int iy = clampi((int)floorf(gy), 0, S - 2);
    const int iz = clampi((int)floorf(gz), 0, S - 2);
    const float wx = clampf(gx - (float)ix, 0.f, 1.f);
    const float wy = clampf(gy - (float)iy, 0.f, 1.f);
    const float wz = clampf(gz - (float)iz, 0.f, 1.f);
    const T* p = base + slot_ox(ix);
    const int y0 = slot_oy(G, iy), y1 = slot_oy(G, iy + 1);
    const int z0 = slot_oz(G, iz), z1 = slot_oz(G, iz + 1);
    float c000, c100, c010, c110, c001, c101, c011, c111;
    ld2(p + (y0 + z0), c000, c100);
    ld2(p + (y1 + z0), c010, c110);
    ld2(p + (y0 + z1), c001, c101);
    ld2(p + (y1 + z1), c011, c111);
    const float c00 = fmaf(c100 - c000, wx, c000);
    const float c10 = fmaf(c110 - c010, wx, c010);
    const float c01 = fmaf(c101 - c001, wx, c001);
    const float c11 = fmaf(c111 - c011, wx, c011);
    const float c0 = fmaf(c10 - c00, wy, c00);
    const float c1 = fmaf(c11 - c01, wy, c01);
    return fmaf(c1 - c0, wz, c0);
  }
  const int ix = clampi((int)floorf(gx), 0, S - 1);
  const int iy = clampi((int)floorf(gy), 0, S - 1);
  const int iz = clampi((int)floorf(gz), 0, S - 1);
  return ld(base + slot_offset(G, ix, iy, iz));
}

__device__ inline float trilerp(const float c[2][2][2], float wx, float wy, float wz) {
  const float c00 = fmaf(c[0][0][1] - c[0][0][0], wx, c[0][0][0]);
  const float c10 = fmaf(c[0][1][1] - c[0][1][0], wx, c[0][1][0]);
  const float c01 = fmaf(c[1][0][1] - c[1][0][0], wx, c[1][0][0]);
  const float c11 = fmaf(c[1][1][1] - c[1][1][0], wx, c[1][1][0]);
  const float c0 = fmaf(c10 - c00, wy, c00);
  const float c1 = fmaf(c11 - c01, wy, c01);
  return fmaf(c1 - c0, wz, c0);
}

// Trilinear value at slot-local texel coordinate (gx, gy, gz) and the
// central differences of the trilinear interpolant at +-1 voxel along each
// axis (unscaled).  The seven interpolations share their weights, so each
// difference is the interpolation of voxel differences over the 2x2x2
// cell; the 32 voxels they read (the cell and one voxel beyond it along
// each axis) come in 16 x-pair loads.  The cell index is clamped so that
// every read stays in the slot (apron >= 2).
template <typename T>
__device__ inline void fetch_value_grad(const T* __restrict__ base, const SlotGeo& G, float gx,
                                        float gy, float gz, float& val, float g[3]) {
  const int S = G.S;
  const int ix = clampi((int)floorf(gx), 1, S - 3);
  const int iy = clampi((int)floorf(gy), 1, S - 3);
  const int iz = clampi((int)floorf(gz), 1, S - 3);
  const float wx = clampf(gx - (float)ix, 0.f, 1.f);
  const float wy = clampf(gy - (float)iy, 0.f, 1.f);
  const float wz = clampf(gz - (float)iz, 0.f, 1.f);
  const int xm = slot_ox(ix - 1), x0 = slot_ox(ix), xp = slot_ox(ix + 1);
  const int yo[4] = {slot_oy(G, iy - 1), slot_oy(G, iy), slot_oy(G, iy + 1), slot_oy(G, iy + 2)};
  const int zo[4] = {slot_oz(G, iz - 1), slot_oz(G, iz), slot_oz(G, iz + 1), slot_oz(G, iz + 2)};
  float row[2][2][4];  // [z][y][x - 1 .. x + 2]
  float ya[2][2][2];   // [z][y - 1, y + 2][x, x + 1]
  float za[2][2][2];   // [z - 1, z + 2][y][x, x + 1]
#pragma unroll
  for (int z = 0; z < 2; ++z)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const T* q = base + (yo[1 + y] + zo[1 + z]);
      ld2(q + xm, row[z][y][0], row[z][y][1]);
      ld2(q + xp, row[z][y][2], row[z][y][3]);
    }
#pragma unroll
  for (int z = 0; z < 2; ++z) {
    ld2(base + (x0 + yo[0] + zo[1 + z]), ya[z][0][0], ya[z][0][1]);
    ld2(base + (x0 + yo[3] + zo[1 + z]), ya[z][1][0], ya[z][1][1]);
  }
#pragma unroll
  for (int y = 0; y < 2; ++y) {
    ld2(base + (x0 + yo[1 + y] + zo[0]), za[0][y][0], za[0][y][1]);
    ld2(base + (x0 + yo[1 + y] + zo[3]), za[1][y][0], za[1][y][1]);
  }
  float c[2][2][2], dx[2][2][2], dy[2][2][2], dz[2][2][2];
#pragma unroll
  for (int z = 0; z < 2; ++z)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        c[z][y][x] = row[z][y][x + 1];
        dx[z][y][x] = row[z][y][x + 2] - row[z][y][x];
        dy[z][y][x] = (y == 0 ? row[z][1][x + 1] : ya[z][1][x]) - (y == 0 ? ya[z][0][x] : row[z][0][x + 1]);
        dz[z][y][x] = (z == 0 ? row[1][y][x + 1] : za[1][y][x]) - (z == 0 ? za[0][y][x] : row[0][y][x + 1]);
      }
  val = trilerp(c, wx, wy, wz);
  g[0] = trilerp(dx, wx, wy, wz);
  g[1] = trilerp(dy, wx, wy, wz);
  g[2] = trilerp(dz, wx, wy, wz);
}

// Render modes of the kernel template: 1D TF without a gradient, any mode
// that needs the gradient (1D TF with lighting, 2D TF), isosurface, 1D TF
// with a per-region LOD.
// kModeMIP: maximum intensity projection -- the largest sample value along
// the ray through the 1D TF; every sample is visited (a TF-transparent cell
// may hold the maximum) and rays do not terminate early.
// kMode1DStaged: the 1D march with the voxels of each 8^3 slot cell staged
// in LDS (render_tile_staged).
// kModeGradGroup: the gradient march of a TF group -- one walk per ray
// compositing up to kGroupMax TFs (render_tile_group).
constexpr int kMode1D = 0, kModeGrad = 1, kModeIso = 2, kModeRegions = 3, kModeMIP = 4, kMode1DStaged = 5,
              kModeGradGroup = 6;

// Device address of brick gid's slot from the page table (null: the brick
// is not resident).  Slots may live in any allocation of the device pool.
template <typename T>
__device__ inline const T* brick_base(const RaycastParams& P, int gid) {
  return (const T*)gload(P.page_table + gid);
}

// Sample at level texel coordinate q through the page table (clamp to
// edge): used off the march's current brick (isosurface refinement and
// normal), where the neighbouring brick may hold the point.
template <typename T, bool LINEAR>
__device__ inline float fetch_level(const RaycastParams& P, float qx, float qy, float qz) {
  const LodInfo& lv = P.level;
  const float q[3] = {qx, qy, qz};
  int b[3];
  float g[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float qi = clampf(q[i], 0.f, (float)lv.dims[i]);
    b[i] = clampi((int)floorf(qi / (float)P.brick), 0, lv.grid[i] - 1);
    g[i] = qi - (float)(b[i] * P.brick - P.apron) - (LINEAR ? 0.5f : 0.f);
  }
  const int gid = lv.brick_base + (b[2] * lv.grid[1] + b[1]) * lv.grid[0] + b[0];
  const T* base = brick_base<T>(P, gid);
  if (!base) {  // paged pool: request the brick; the frame is re-rendered
    if (P.requests) P.requests[gid] = 1u;
    return 0.f;
  }
  if (P.touch) P.touch[gid] = P.frame;
  return fetch<T, LINEAR>(base, slot_geo(P.slot_edge), g[0], g[1], g[2]);
}

// TF entry i: from LDS when the launch staged the TF there, else through a
// global pointer.
typedef float native_f4 __attribute__((ext_vector_type(4)));
__device__ inline float4 gload_f4(const float4* p) {
  const native_f4 v = *(__attribute__((address_space(1))) const native_f4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}

template <bool TF_LDS>
__device__ inline float4 tf_entry(const float4* __restrict__ tf4, int i) {
  if constexpr (TF_LDS) return tf4[i];
  else return gload_f4(tf4 + i);
}

// 1D TF: linear between the two entries around value * tf_scale.
template <bool TF_LDS>
__device__ inline void tf1d_lookup(const RaycastParams& P, const float4* __restrict__ tf4, float v,
                                   float c[4]) {
  const float x = clampf(v * P.tf_scale, 0.f, (float)(P.n_tf - 1));
  const int i = (int)x;
  const float f = x - (float)i;
  const float4 e0 = tf_entry<TF_LDS>(tf4, i), e1 = tf_entry<TF_LDS>(tf4, min(i + 1, P.n_tf - 1));
  c[0] = fmaf(e1.x - e0.x, f, e0.x);
  c[1] = fmaf(e1.y - e0.y, f, e0.y);
  c[2] = fmaf(e1.z - e0.z, f, e0.z);
  c[3] = fmaf(e1.w - e0.w, f, e0.w);
}

// 2D TF (value x gradient magnitude), bilinear, value fastest.
__device__ inline void tf2d_lookup(const RaycastParams& P, float v, float gm, float c[4]) {
  const int W = P.tf2d_w, H = P.tf2d_h;
  const float x = clampf(v * P.tf_scale, 0.f, (float)(W - 1));
  const float y = clampf(gm * P.tf2d_row_scale, 0.f, (float)(H - 1));
  const int i = (int)x, j = (int)y;
  const float fx = x - (float)i, fy = y - (float)j;
  const int i1 = min(i + 1, W - 1), j1 = min(j + 1, H - 1);
  const float4* t = (const float4*)P.tf2d;
  const float4 a = gload_f4(t + (j * W + i)), b = gload_f4(t + (j * W + i1));
  const float4 e = gload_f4(t + (j1 * W + i)), f = gload_f4(t + (j1 * W + i1));
  const float ax[4] = {a.x, a.y, a.z, a.w}, bx[4] = {b.x, b.y, b.z, b.w};
  const float ex[4] = {e.x, e.y, e.z, e.w}, fx4[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float r0 = fmaf(bx[k] - ax[k], fx, ax[k]);
    const float r1 = fmaf(fx4[k] - ex[k], fx, ex[k]);
    c[k] = fmaf(r1 - r0, fy, r0);
  }
}

// Two-sided Blinn-Phong: c*(ka + kd|N.L|) + ks|N.H|^shininess with
// N = g / sqrt(|g|^2 + 1e-8); Lh = light direction and half vector.
__device__ inline void shade(const RaycastParams& P, const float g[3], const float Lh[6], float c[4]) {
  const float s = 1.f / sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + 1e-8f);
  const float nl = fabsf((g[0] * Lh[0] + g[1] * Lh[1] + g[2] * Lh[2]) * s);
  const float nh = fabsf((g[0] * Lh[3] + g[1] * Lh[4] + g[2] * Lh[5]) * s);
  const float sp = nh > 0.f ? exp2f(P.shininess * log2f(nh)) : 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) c[i] = c[i] * fmaf(P.kd[i], nl, P.ka[i]) + P.ks[i] * sp;
}

// Isosurface hit at sample kh (value vh >= iso, sample kh-1 below it): the
// crossing refined linearly against sample kh-1 (re-fetched from the level,
// so the result does not depend on skipping), `col` shaded there with the
// central-difference gradient into c.  Returns |N.V| (N the unit gradient,
// V = -d): ClearView's silhouette term of the context surface.
template <typename T, bool LINEAR>
__device__ inline float iso_shade(const RaycastParams& P, const LodInfo& lv, const float o[3],
                                  const float d[3], const float Lh[6], int kh, int n, float t0,
                                  float step, float lastlen, float vh, float iso, const float col[3],
                                  float c[4]) {
  const float lenh = (kh == n - 1) ? lastlen : step;
  const float th0 = fmaf((float)kh, step, t0) + 0.5f * lenh;
  float th = th0;
  if (kh > 0) {
    const float tp = fmaf((float)(kh - 1), step, t0) + 0.5f * step;
    const float vp = fetch_level<T, LINEAR>(P, (o[0] + tp * d[0]) * lv.to_level[0],
                                            (o[1] + tp * d[1]) * lv.to_level[1],
                                            (o[2] + tp * d[2]) * lv.to_level[2]);
    if (vp < iso && vh > vp) th = tp + (iso - vp) / (vh - vp) * (th0 - tp);
  }
  const float qx = (o[0] + th * d[0]) * lv.to_level[0];
  const float qy = (o[1] + th * d[1]) * lv.to_level[1];
  const float qz = (o[2] + th * d[2]) * lv.to_level[2];
  float gh[3];
  gh[0] = (fetch_level<T, true>(P, qx + 1.f, qy, qz) - fetch_level<T, true>(P, qx - 1.f, qy, qz)) * P.grad_scale[0];
  gh[1] = (fetch_level<T, true>(P, qx, qy + 1.f, qz) - fetch_level<T, true>(P, qx, qy - 1.f, qz)) * P.grad_scale[1];
  gh[2] = (fetch_level<T, true>(P, qx, qy, qz + 1.f) - fetch_level<T, true>(P, qx, qy, qz - 1.f)) * P.grad_scale[2];
  c[0] = col[0];
  c[1] = col[1];
  c[2] = col[2];
  c[3] = 1.f;
  shade(P, gh, Lh, c);
  const float s = 1.f / sqrtf(gh[0] * gh[0] + gh[1] * gh[1] + gh[2] * gh[2] + 1e-8f);
  return fabsf((gh[0] * d[0] + gh[1] * d[1] + gh[2] * d[2]) * s);
}

// Render the 64 rays of ray tile `tile` (lane = ray within the 8x8 tile).
// `tf4` points at the 1D TF (LDS or global).
// Per-region LOD march (1D TF): the ray walks the LOD-0 brick grid; region
// r is rendered from level l = region_code[r] & 15 with step 2^l / rate,
// its own segments restarting at the region entry (the last one partial),
// or skipped when region_code[r] & 0x80 marks the level-l brick as
// transparent under the TF.  Returns the region that added the most
// opacity (-1 if none) for error attribution.
template <typename T, bool LINEAR, int BATCH, bool TF_LDS>
__device__ inline int march_regions(const RaycastParams& P, const float4* __restrict__ tf4,
                                    const float o[3], const float d[3], const float inv[3], float t0,
                                    float t1, float& C0, float& C1, float& C2, float& A,
                                    uint32_t& n_eval) {
  const LodInfo& L0 = P.lods[0];
  const float B = (float)P.brick;
  const float tfmax = (float)(P.n_tf - 1);
  const float base_step = P.step;
  int dominant = -1;
  float best = 0.f;
  float t = t0;
  bool live = true;
  while (t < t1 && live) {
    const float tm = t + 0.5f * fminf(base_step, t1 - t);
    int c[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) c[i] = clampi((int)floorf(fmaf(tm, d[i], o[i]) / B), 0, L0.grid[i] - 1);
    float tb = t1;
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (d[i] != 0.f) {
        const float face = (float)(d[i] > 0.f ? c[i] + 1 : c[i]) * B;
        tb = fminf(tb, (face - o[i]) * inv[i]);
      }
    tb = fminf(fmaxf(tb, t + 1e-3f), t1);  // always advance
    const int region = (c[2] * L0.grid[1] + c[1]) * L0.grid[0] + c[0];
    const uint32_t code = gload(P.region_code + region);
    if (!(code & 0x80u)) {
      const int l = (int)(code & 15u);
      const LodInfo& lv = P.lods[l];
      // the level-l brick holding the region (from its centre)
      int b[3];
      float oo[3], od[3], g0[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float qc = ((float)c[i] + 0.5f) * B * lv.to_level[i];
        b[i] = clampi((int)floorf(qc / B), 0, lv.grid[i] - 1);
        oo[i] = o[i] * lv.to_level[i];
        od[i] = d[i] * lv.to_level[i];
        g0[i] = oo[i] - (float)(b[i] * P.brick - P.apron) - (LINEAR ? 0.5f : 0.f);
      }
      const int gid = lv.brick_base + (b[2] * lv.grid[1] + b[1]) * lv.grid[0] + b[0];
      const T* __restrict__ base = brick_base<T>(P, gid);
      if (!base) {
        if (P.requests) P.requests[gid] = 1u;
      } else {
        if (P.touch) P.touch[gid] = P.frame;
        const SlotGeo G = slot_geo(P.slot_edge);
        const float step = ldexpf(base_step, l);
        const float span = tb - t;
        const int m = max((int)ceilf(span / step), 1);
        const float last = fmaxf(span - (float)(m - 1) * step, 0.f);
        const float A_in = A;
        for (int k = 0; k < m && live; k += BATCH) {
          float v[BATCH];
#pragma unroll
          for (int j = 0; j < BATCH; ++j) {
            const int kk = min(k + j, m - 1);
            const float len = kk == m - 1 ? last : step;
            const float ts = fmaf((float)kk, step, t) + 0.5f * len;
            v[j] = fetch<T, LINEAR>(base, G, fmaf(ts, od[0], g0[0]), fmaf(ts, od[1], g0[1]),
                                    fmaf(ts, od[2], g0[2]));
          }
#pragma unroll
          for (int j = 0; j < BATCH; ++j) {
            if (k + j < m && live) {
              ++n_eval;
              const float x = clampf(v[j] * P.tf_scale, 0.f, tfmax);
              const int ti = (int)x;
              const float fr = x - (float)ti;
              const float4 e0 = tf_entry<TF_LDS>(tf4, ti);
              const float4 e1 = tf_entry<TF_LDS>(tf4, min(ti + 1, P.n_tf - 1));
              const float a = fmaf(e1.w - e0.w, fr, e0.w);
              if (a > 0.f) {
                const float len = (k + j == m - 1) ? last : step;
                const float ac = a >= 1.f ? 1.f : 1.f - exp2f(len * log2f(1.f - a));
                const float w = (1.f - A) * ac;
                C0 = fmaf(w, fmaf(e1.x - e0.x, fr, e0.x), C0);
                C1 = fmaf(w, fmaf(e1.y - e0.y, fr, e0.y), C1);
                C2 = fmaf(w, fmaf(e1.z - e0.z, fr, e0.z), C2);
                A += w;
                live = A < P.term_alpha;
              }
            }
          }
        }
        if (A - A_in > best) {
          best = A - A_in;
          dominant = region;
        }
      }
    }
    t = tb;
  }
  return dominant;
}

// Premultiplied ray colour (C, A) over row py's background gradient, as the
// framebuffer's RGBA8.  A ray that misses the volume has C = A = 0 and gets
// the background itself; the background kernel calls this with zeros too,
// so both paths write the same bytes (explicit fmaf: no contraction choice
// left to the compiler).
__device__ inline uint32_t over_background(const RaycastParams& P, int py, float C0, float C1, float C2,
                                           float A) {
  const float wy = 1.f - ((float)py + 0.5f) / (float)P.height;
  const float bg0 = fmaf(P.bg_top[0] - P.bg_bottom[0], wy, P.bg_bottom[0]);
  const float bg1 = fmaf(P.bg_top[1] - P.bg_bottom[1], wy, P.bg_bottom[1]);
  const float bg2 = fmaf(P.bg_top[2] - P.bg_bottom[2], wy, P.bg_bottom[2]);
  const float T1 = 1.f - A;
  return pack_rgba8(fmaf(T1, bg0, C0), fmaf(T1, bg1, C1), fmaf(T1, bg2, C2), fmaf(T1, P.bg_alpha, A));
}

// Clip the ray interval [t0, t1] to the kept side of the clip plane
// (clip.xyz . p + clip.w <= 0); a ray parallel to the plane is kept or
// lost whole.
__device__ inline void clip_ray(const RaycastParams& P, const float o[3], const float d[3], float& t0,
                                float& t1, bool& hit) {
  const float num = fmaf(P.clip[0], o[0], fmaf(P.clip[1], o[1], fmaf(P.clip[2], o[2], P.clip[3])));
  const float den = fmaf(P.clip[0], d[0], fmaf(P.clip[1], d[1], P.clip[2] * d[2]));
  if (den == 0.f) {
    if (num > 0.f) hit = false;
    return;
  }
  const float tc = -num / den;
  if (den > 0.f) t1 = fminf(t1, tc);
  else t0 = fmaxf(t0, tc);
}

// The running composite rounded to the blend target's precision: unorm8
// (round to nearest of x * 255) or half float (round to nearest even).
__device__ inline float blend_quantize(float x, int bits) {
  if (bits == 8) return floorf(fmaf(clampf(x, 0.f, 1.f), 255.f, 0.5f)) / 255.f;
  return (float)(_Float16)x;
}
__device__ inline void blend_round(int bits, float& C0, float& C1, float& C2, float& A) {
  C0 = blend_quantize(C0, bits);
  C1 = blend_quantize(C1, bits);
  C2 = blend_quantize(C2, bits);
  A = blend_quantize(A, bits);
}

// The k-th tile row of an h-row rect in centre-out order: c, c+1, c-1,
// c+2, c-2, ... (c = h/2), then the rest of the longer side.  The rows
// through the middle of the box's projection hold the longest rays, so they
// are dealt first and the rows near the rect's edges, whose rays are short,
// end the frame (in top-down order the middle rows started about a third
// of the way in and a few of their long tiles ran on alone at the end).
// A bijection of [0, h).
__device__ inline int centre_out_row(int k, int h) {
  const int c = h >> 1, below = c, above = h - 1 - c, m = min(below, above);
  if (k == 0) return c;
  if (k <= 2 * m) return (k & 1) ? c + ((k + 1) >> 1) : c - (k >> 1);
  return above > below ? c + (k - m) : c - (k - m);
}

// A lane's pixel and ray: origin o, unit direction d (LOD-0 voxel space),
// 1/d per axis (inf on axis-parallel rays) and the interval [t0, t1] inside
// the volume box after the clip plane; hit false when the ray misses it.
struct Ray {
  int px, py;
  float o[3], d[3], inv[3];
  float t0, t1;
  bool hit;
};

// The ray of `lane` in ray tile `tile` of job P (row-major over the job's
// tile rect, rows taken centre-out).  Returns false for a lane with no
// pixel (beyond the image, or a tile that tile sampling leaves out); adds
// the nominal samples of the ray's chord through the box to n_nom.
__device__ inline bool setup_ray(const RaycastParams& P, int tile, int lane, Ray& r, uint32_t& n_nom) {
  if (P.tile_sample > 1) {
    // one tile of each run of tile_sample, at an offset hashed from the
    // run and the phase (no stripes when the rect width shares a factor
    // with tile_sample)
    const uint32_t h = ((uint32_t)tile * 0x9E3779B1u + (uint32_t)P.tile_phase * 0x85EBCA77u) >> 16;
    tile = tile * P.tile_sample + (int)(h % (uint32_t)P.tile_sample);
    if (tile >= P.rect_w * P.rect_h) return false;
  }
  const int tx = P.rect_x0 + tile % P.rect_w, ty = P.rect_y0 + centre_out_row(tile / P.rect_w, P.rect_h);
  const int sh = P.tile_shift;
  const int px = (tx << sh) + (lane & ((1 << sh) - 1));
  const int py = ty * (64 >> sh) + (lane >> sh);
  if (px >= P.width || py >= P.height) return false;
  r.px = px;
  r.py = py;
  float(&o)[3] = r.o;
  float(&d)[3] = r.d;
  float(&inv)[3] = r.inv;
  const float fx = (float)px, fy = (float)py;
  if (P.ortho) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      o[i] = P.eye[i] + fx * P.du[i] + fy * P.dv[i];
      d[i] = P.odir[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      o[i] = P.eye[i];
      d[i] = P.dir00[i] + fx * P.du[i] + fy * P.dv[i];
    }
    const float l = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) d[i] = d[i] / l;
  }
  // ray vs LOD-0 box [0, box]; axis-parallel rays use an inside test
  float t0 = -kInf, t1 = kInf;
  bool hit = true;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (d[i] == 0.f) {
      inv[i] = kInf;
      if (o[i] < 0.f || o[i] > P.box[i]) hit = false;
    } else {
      inv[i] = 1.f / d[i];
      float ta = (0.f - o[i]) * inv[i], tb = (P.box[i] - o[i]) * inv[i];
      if (ta > tb) {
        const float s = ta;
        ta = tb;
        tb = s;
      }
      t0 = fmaxf(t0, ta);
      t1 = fminf(t1, tb);
    }
  }
  t0 = fmaxf(t0, 0.f);
  // nominal samples count the chord through the box, before clipping
  if (hit && t1 > t0) n_nom += (uint32_t)floorf((t1 - t0) * P.nominal_rate);
  if (P.clip_on && hit) clip_ray(P, o, d, t0, t1, hit);
  r.t0 = t0;
  r.t1 = t1;
  r.hit = hit;
  return true;
}

template <typename T, bool LINEAR, int BATCH, int MODE, bool TF_LDS>
__device__ inline void render_tile(const RaycastParams& P, const float4* __restrict__ tf4, int tile,
                                   int lane, uint32_t& n_eval, uint32_t& n_nom) {
  Ray r;
  if (!setup_ray(P, tile, lane, r, n_nom)) return;
  const int px = r.px, py = r.py;
  const float(&o)[3] = r.o;
  const float(&d)[3] = r.d;
  const float(&inv)[3] = r.inv;
  const float t0 = r.t0, t1 = r.t1;
  const bool hit = r.hit;
  const float L = t1 - t0;
  float C0 = 0.f, C1 = 0.f, C2 = 0.f, A = 0.f;
  float vmax = -1.f;  // MIP: largest sample value so far (-1: none)
  // ClearView: this pixel's focus weight; after the context hit (stage 1)
  // its shaded colour and opacity
  float wf = 0.f, cv_ac = 1.f, cc[3] = {0.f, 0.f, 0.f};
  int cv_stage = 0;
  if (MODE == kModeIso && P.cv) {
    const float dx = (float)px - P.cv_focus[0], dy = (float)py - P.cv_focus[1];
    wf = clampf(1.f - P.cv_size * (sqrtf(dx * dx + dy * dy) / (float)P.height), 0.f, 1.f);
  }
  if (MODE == kModeRegions) {
    int dominant = -1;
    if (hit && L > 0.f) {
      dominant = march_regions<T, LINEAR, BATCH, TF_LDS>(P, tf4, o, d, inv, t0, t1, C0, C1, C2, A, n_eval);
    }
    if (P.region_out) P.region_out[(int64_t)py * P.width + px] = dominant;
  } else if (hit && L > 0.f) {
    const float step = P.step;
    const int n = (int)ceilf(L / step);
    const float lastlen = fmaxf(L - (float)(n - 1) * step, 0.f);
    const LodInfo& lv = P.level;
    // The ray walks march units of P.cell level voxels: the empty-space
    // cells when skipping (a cell lies inside one brick), else the bricks.
    const int U = P.cell;
    const int per_brick = P.brick / U;
    const float invU = 1.f / (float)U;
    float od[3], oo[3], ul0[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      od[i] = d[i] * lv.to_level[i];
      oo[i] = o[i] * lv.to_level[i];
      ul0[i] = (float)U / lv.to_level[i];  // unit edge in LOD-0 voxels
    }
    const int ucount[3] = {(lv.dims[0] + U - 1) / U, (lv.dims[1] + U - 1) / U,
                           (lv.dims[2] + U - 1) / U};
    const float tfmax = (float)(P.n_tf - 1);
    const SlotGeo G = slot_geo(P.slot_edge);
    float Lh[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // light direction, half vector
    if (MODE != kMode1D) {
      float h[3], hl = 0.f;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        Lh[i] = P.light[i];
        h[i] = P.light[i] - d[i];
        hl += h[i] * h[i];
      }
      hl = sqrtf(hl);
      if (hl > 1e-12f)
#pragma unroll
        for (int i = 0; i < 3; ++i) Lh[3 + i] = h[i] / hl;
    }
    constexpr bool kBlends = MODE == kMode1D || MODE == kModeGrad;
    int blend_gid = -1;  // brick of the last blend into the target
    int k = 0;
    if constexpr (MODE == kModeGrad || MODE == kMode1D) {
      // The 1D and gradient marches, flattened: each pass of the loop first sets up
      // the next march unit for the lanes whose unit ended (skipping empty
      // and non-resident ones), then gathers one batch of samples for every
      // lane inside a unit.  A nested per-unit sample loop would hold the
      // wave at every unit boundary until its lane with the most samples
      // left in its unit is done; here a lane whose unit ends goes on with
      // its next unit in the next pass.  Samples, their order and the
      // arithmetic are those of the nested form.
      int kend = 0;
      float g0[3] = {0.f, 0.f, 0.f};
      const T* __restrict__ base = nullptr;
      bool live = true;
      while (k < n) {
        if (k >= kend) {
          const float len0 = (k == n - 1) ? lastlen : step;
          const float tk = fmaf((float)k, step, t0) + 0.5f * len0;
          int c[3];
#pragma unroll
          for (int i = 0; i < 3; ++i)
            c[i] = clampi((int)floorf(fmaf(tk, od[i], oo[i]) * invU), 0, ucount[i] - 1);
          float tb = kInf;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            if (d[i] != 0.f) {
              const float face = (float)(d[i] > 0.f ? c[i] + 1 : c[i]) * ul0[i];
              tb = fminf(tb, (face - o[i]) * inv[i]);
            }
          }
          kend = min(max((int)ceilf((tb - t0) / step - 0.5f), k + 1), n);
          int b[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) b[i] = c[i] / per_brick;
          const int gid = lv.brick_base + (b[2] * lv.grid[1] + b[1]) * lv.grid[0] + b[0];
          const uint8_t empty =
              P.skip ? gload(P.cell_empty + (lv.cell_base + (c[2] * ucount[1] + c[1]) * ucount[0] + c[0])) : 0;
          base = brick_base<T>(P, gid);
          if (P.blend_bits < 32 && gid != blend_gid) {
            blend_round(P.blend_bits, C0, C1, C2, A);
            blend_gid = gid;
          }
          if (empty) {
            k = kend;
            continue;
          }
          if (!base) {
            if (P.requests) P.requests[gid] = 1u;
            k = kend;
            continue;
          }
          if (P.touch) P.touch[gid] = P.frame;
#pragma unroll
          for (int i = 0; i < 3; ++i)
            g0[i] = oo[i] - (float)(b[i] * P.brick - P.apron) - (LINEAR ? 0.5f : 0.f);
        }
        if constexpr (MODE == kMode1D) {
          // gather BATCH samples (clamped to the unit's last sample) so their
          // loads are in flight together, then look each up in the TF and
          // composite in order
          float v[BATCH];
#pragma unroll
          for (int j = 0; j < BATCH; ++j) {
            const int kk = min(k + j, kend - 1);
            const float lenj = (kk == n - 1) ? lastlen : step;
            const float t = fmaf((float)kk, step, t0) + 0.5f * lenj;
            v[j] = fetch<T, LINEAR>(base, G, fmaf(t, od[0], g0[0]), fmaf(t, od[1], g0[1]),
                                    fmaf(t, od[2], g0[2]));
          }
#pragma unroll
          for (int j = 0; j < BATCH; ++j) {
            if (k + j < kend && live) {
              ++n_eval;
              const float x = clampf(v[j] * P.tf_scale, 0.f, tfmax);
              const int ti = (int)x;
              const float fr = x - (float)ti;
              const float4 e0 = tf_entry<TF_LDS>(tf4, ti);
              const float4 e1 = tf_entry<TF_LDS>(tf4, min(ti + 1, P.n_tf - 1));
              const float a = fmaf(e1.w - e0.w, fr, e0.w);
              if (a > 0.f) {
                const float lenj = (k + j == n - 1) ? lastlen : step;
                const float ac = a >= 1.f ? 1.f : 1.f - exp2f(lenj * log2f(1.f - a));
                const float w = (1.f - A) * ac;
                C0 = fmaf(w, fmaf(e1.x - e0.x, fr, e0.x), C0);
                C1 = fmaf(w, fmaf(e1.y - e0.y, fr, e0.y), C1);
                C2 = fmaf(w, fmaf(e1.z - e0.z, fr, e0.z), C2);
                A += w;
                live = A < P.term_alpha;
              }
            }
          }
          k = live ? min(k + BATCH, kend) : n;
          continue;
        }
        float v[BATCH], len[BATCH], g[BATCH][3];
#pragma unroll
        for (int j = 0; j < BATCH; ++j) {
          const int kk = min(k + j, kend - 1);
          len[j] = (kk == n - 1) ? lastlen : step;
          const float t = fmaf((float)kk, step, t0) + 0.5f * len[j];
          const float gx = fmaf(t, od[0], g0[0]), gy = fmaf(t, od[1], g0[1]), gz = fmaf(t, od[2], g0[2]);
          float dd[3];
          if (LINEAR) {
            fetch_value_grad<T>(base, G, gx, gy, gz, v[j], dd);
          } else {
            v[j] = fetch<T, false>(base, G, gx, gy, gz);
            const float lx = gx - 0.5f, ly = gy - 0.5f, lz = gz - 0.5f;
            dd[0] = fetch<T, true>(base, G, lx + 1.f, ly, lz) - fetch<T, true>(base, G, lx - 1.f, ly, lz);
            dd[1] = fetch<T, true>(base, G, lx, ly + 1.f, lz) - fetch<T, true>(base, G, lx, ly - 1.f, lz);
            dd[2] = fetch<T, true>(base, G, lx, ly, lz + 1.f) - fetch<T, true>(base, G, lx, ly, lz - 1.f);
          }
          g[j][0] = dd[0] * P.grad_scale[0];
          g[j][1] = dd[1] * P.grad_scale[1];
          g[j][2] = dd[2] * P.grad_scale[2];
        }
#pragma unroll
        for (int j = 0; j < BATCH; ++j) {
          if (k + j < kend && live) {
            ++n_eval;
            float cs[4];
            if (P.mode == 1) {
              const float gm = sqrtf(g[j][0] * g[j][0] + g[j][1] * g[j][1] + g[j][2] * g[j][2]);
              tf2d_lookup(P, v[j], gm, cs);
            } else {
              tf1d_lookup<TF_LDS>(P, tf4, v[j], cs);
            }
            if (cs[3] > 0.f) {
              if (P.lighting) shade(P, g[j], Lh, cs);
              const float ac = cs[3] >= 1.f ? 1.f : 1.f - exp2f(len[j] * log2f(1.f - cs[3]));
              const float w = (1.f - A) * ac;
              C0 = fmaf(w, cs[0], C0);
              C1 = fmaf(w, cs[1], C1);
              C2 = fmaf(w, cs[2], C2);
              A += w;
              live = A < P.term_alpha;
            }
          }
        }
        k = live ? min(k + BATCH, kend) : n;
      }
      k = n;  // the generic loop below has nothing left
    }
    while (k < n) {
      const float len0 = (k == n - 1) ? lastlen : step;
      const float tk = fmaf((float)k, step, t0) + 0.5f * len0;
      int c[3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
        c[i] = clampi((int)floorf(fmaf(tk, od[i], oo[i]) * invU), 0, ucount[i] - 1);
      float tb = kInf;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        if (d[i] != 0.f) {
          const float face = (float)(d[i] > 0.f ? c[i] + 1 : c[i]) * ul0[i];
          tb = fminf(tb, (face - o[i]) * inv[i]);
        }
      }
      int kend = (int)ceilf((tb - t0) / step - 0.5f);
      kend = min(max(kend, k + 1), n);
      int b[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) b[i] = c[i] / per_brick;
      const int gid = lv.brick_base + (b[2] * lv.grid[1] + b[1]) * lv.grid[0] + b[0];
      // the cell's empty flag and the brick's slot are independent loads
      const uint8_t empty =
          (MODE != kModeMIP && P.skip)
              ? gload(P.cell_empty + (lv.cell_base + (c[2] * ucount[1] + c[1]) * ucount[0] + c[0]))
              : 0;
      const T* __restrict__ base = brick_base<T>(P, gid);
      if (kBlends && P.blend_bits < 32 && gid != blend_gid) {
        blend_round(P.blend_bits, C0, C1, C2, A);
        blend_gid = gid;
      }
      if (empty) {
        k = kend;
        continue;
      }
      if (!base) {  // not resident: ask the host for it, sample nothing
        if (P.requests) P.requests[gid] = 1u;
        k = kend;
        continue;
      }
      if (P.touch) P.touch[gid] = P.frame;
      float g0[3];  // slot-local texel coordinate at t = 0
#pragma unroll
      for (int i = 0; i < 3; ++i)
        g0[i] = oo[i] - (float)(b[i] * P.brick - P.apron) - (LINEAR ? 0.5f : 0.f);
      bool live = true;
      if (MODE == kModeIso) {
        // first sample at or above the iso value; refine against the
        // previous sample, shade at the refined point
        for (; k < kend && live; k += BATCH) {
          float v[BATCH];
#pragma unroll
          for (int j = 0; j < BATCH; ++j) {
            const int kk = min(k + j, kend - 1);
            const float lenj = (kk == n - 1) ? lastlen : step;
            const float t = fmaf((float)kk, step, t0) + 0.5f * lenj;
            v[j] = fetch<T, LINEAR>(base, G, fmaf(t, od[0], g0[0]), fmaf(t, od[1], g0[1]),
                                    fmaf(t, od[2], g0[2]));
          }
#pragma unroll
          for (int j = 0; j < BATCH; ++j) {
            if (k + j < kend && live) {
              ++n_eval;
              // stage 0: the context surface; with ClearView and the pixel
              // in the focus (wf > 0), stage 1 then looks for the focus
              // surface from the same sample on
              if (cv_stage == 0 && v[j] >= P.iso_value) {
                float c[4];
                const float nv = iso_shade<T, LINEAR>(P, lv, o, d, Lh, k + j, n, t0, step, lastlen, v[j],
                                                      P.iso_value, P.iso_color, c);
                if (wf > 0.f) {
                  cc[0] = c[0];
                  cc[1] = c[1];
                  cc[2] = c[2];
                  const float x = 1.f - nv, b = P.cv_border_scale;
                  const float pw = x > 0.f ? exp2f(b * log2f(x)) : (b == 0.f ? 1.f : 0.f);
                  cv_ac = fmaf(wf, clampf(P.cv_context_scale * pw, 0.f, 1.f) - 1.f, 1.f);
                  cv_stage = 1;
                } else {
                  C0 = c[0];
                  C1 = c[1];
                  C2 = c[2];
                  A = 1.f;
                  live = false;
                }
              }
              if (cv_stage == 1 && v[j] >= P.cv_iso_value) {
                float c[4];
                (void)iso_shade<T, LINEAR>(P, lv, o, d, Lh, k + j, n, t0, step, lastlen, v[j],
                                           P.cv_iso_value, P.cv_color, c);
                C0 = fmaf(cv_ac, cc[0] - c[0], c[0]);
                C1 = fmaf(cv_ac, cc[1] - c[1], c[1]);
                C2 = fmaf(cv_ac, cc[2] - c[2], c[2]);
                A = 1.f;
                cv_stage = 2;
                live = false;
              }
            }
          }
        }
        k = live ? kend : n;
        continue;
      }
      // Batched 1D march: gather BATCH samples (clamped to the brick's last
      // sample) so their loads are in flight together, then look each up in
      // the TF and composite in order (TF entries are read per sample, which
      // keeps the register count down).
      for (; k < kend; k += BATCH) {
        float v[BATCH];
#pragma unroll
        for (int j = 0; j < BATCH; ++j) {
          const int kk = min(k + j, kend - 1);
          const float lenj = (kk == n - 1) ? lastlen : step;
          const float t = fmaf((float)kk, step, t0) + 0.5f * lenj;
          v[j] = fetch<T, LINEAR>(base, G, fmaf(t, od[0], g0[0]), fmaf(t, od[1], g0[1]),
                                  fmaf(t, od[2], g0[2]));
        }
#pragma unroll
        for (int j = 0; j < BATCH; ++j) {
          if (MODE == kModeMIP) {
            if (k + j < kend) {
              ++n_eval;
              vmax = fmaxf(vmax, v[j]);
            }
            continue;
          }
          if (k + j < kend && live) {
            ++n_eval;
            const float x = clampf(v[j] * P.tf_scale, 0.f, tfmax);
            const int ti = (int)x;
            const float fr = x - (float)ti;
            const float4 e0 = tf_entry<TF_LDS>(tf4, ti);
            const float4 e1 = tf_entry<TF_LDS>(tf4, min(ti + 1, P.n_tf - 1));
            const float a = fmaf(e1.w - e0.w, fr, e0.w);
            if (a > 0.f) {
              const float lenj = (k + j == n - 1) ? lastlen : step;
              const float ac = a >= 1.f ? 1.f : 1.f - exp2f(lenj * log2f(1.f - a));
              const float w = (1.f - A) * ac;
              C0 = fmaf(w, fmaf(e1.x - e0.x, fr, e0.x), C0);
              C1 = fmaf(w, fmaf(e1.y - e0.y, fr, e0.y), C1);
              C2 = fmaf(w, fmaf(e1.z - e0.z, fr, e0.z), C2);
              A += w;
              live = A < P.term_alpha;
            }
          }
        }
        if (!live) break;
      }
      k = live ? kend : n;
    }
    if (kBlends && P.blend_bits < 32) blend_round(P.blend_bits, C0, C1, C2, A);
  }
  if (MODE == kModeIso && cv_stage == 1) {  // context hit, no focus surface behind it
    C0 = cv_ac * cc[0];
    C1 = cv_ac * cc[1];
    C2 = cv_ac * cc[2];
    A = cv_ac;
  }
  if (MODE == kModeMIP && vmax >= 0.f) {
    float c[4];
    tf1d_lookup<TF_LDS>(P, tf4, vmax, c);
    C0 = c[0] * c[3];
    C1 = c[1] * c[3];
    C2 = c[2] * c[3];
    A = c[3];
  }
  gstore(P.framebuffer + ((int64_t)py * P.width + px), over_background(P, py, C0, C1, C2, A));
}

// The march unit of sample k on a TF group's ray (render_tile_group and
// render_tile_group_runs), as render_tile's walk finds it: its end kend,
// the live members whose TF keeps it (need), its brick slot `base` and the
// slot-local texel origin g0.  Rounds each live member's blend when the
// brick changes (a member's own march stops rounding once it has
// terminated), asks for a non-resident brick and touches a resident one
// that some member needs.  Returns whether the unit has samples to gather.
template <typename T, bool LINEAR>
__device__ inline bool group_unit(const RaycastParams& P, const RaycastParams* const (&Q)[kGroupMax], int F,
                                  const Ray& r, const float (&od)[3], const float (&oo)[3],
                                  const float (&ul0)[3], const int (&ucount)[3], float invU, int per_brick,
                                  int k, int n, float t0, float step, float lastlen, uint32_t live, int& kend,
                                  uint32_t& need, const T* __restrict__& base, float (&g0)[3], int& blend_gid,
                                  float (&C)[kGroupMax][3], float (&A)[kGroupMax]) {
  const LodInfo& lv = P.level;
  const float len0 = (k == n - 1) ? lastlen : step;
  const float tk = fmaf((float)k, step, t0) + 0.5f * len0;
  int c[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) c[i] = clampi((int)floorf(fmaf(tk, od[i], oo[i]) * invU), 0, ucount[i] - 1);
  float tb = kInf;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (r.d[i] != 0.f) {
      const float face = (float)(r.d[i] > 0.f ? c[i] + 1 : c[i]) * ul0[i];
      tb = fminf(tb, (face - r.o[i]) * r.inv[i]);
    }
  }
  kend = min(max((int)ceilf((tb - t0) / step - 0.5f), k + 1), n);
  int b[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i] = c[i] / per_brick;
  const int gid = lv.brick_base + (b[2] * lv.grid[1] + b[1]) * lv.grid[0] + b[0];
  const int64_t cidx = lv.cell_base + (c[2] * ucount[1] + c[1]) * ucount[0] + c[0];
  need = 0;
#pragma unroll
  for (int f = 0; f < kGroupMax; ++f)
    if (f < F && (live >> f & 1u) && !(P.skip && gload(Q[f]->cell_empty + cidx))) need |= 1u << f;
  base = brick_base<T>(P, gid);
  if (P.blend_bits < 32 && gid != blend_gid) {
#pragma unroll
    for (int f = 0; f < kGroupMax; ++f)
      if (live >> f & 1u) blend_round(P.blend_bits, C[f][0], C[f][1], C[f][2], A[f]);
    blend_gid = gid;
  }
  if (!need) return false;
  if (!base) {
    if (P.requests) P.requests[gid] = 1u;
    return false;
  }
  if (P.touch) P.touch[gid] = P.frame;
#pragma unroll
  for (int i = 0; i < 3; ++i) g0[i] = oo[i] - (float)(b[i] * P.brick - P.apron) - (LINEAR ? 0.5f : 0.f);
  return true;
}

// Gradient march of a TF group (kModeGradGroup): jobs that differ only in
// their TF (same camera, image, volume, level, rate and shading) walk each
// ray once.  Every sample's value and gradient are gathered once and each
// member's TF is looked up, shaded and composited into its own (C, A);
// member f evaluates exactly the samples its own march would -- those of
// the march units its TF does not skip, up to its own early termination --
// in the same order and with the same arithmetic, so every member's image
// and sample count equal its own render's bit for bit.  The walk skips a
// unit only when it is empty for every member still compositing, and ends
// when all have terminated.  A job alone (group_n <= 1) takes the same
// path with one TF.  n_eval[f] counts member f's samples, n_walk the
// samples of the walk that some member evaluates (the union of theirs).
template <typename T, bool LINEAR, int BATCH>
__device__ inline void render_tile_group(const JobTable* __restrict__ tab, const RaycastParams& P, int tile,
                                         int lane, uint32_t (&n_eval)[kGroupMax], uint32_t& n_walk,
                                         uint32_t& n_nom) {
  Ray r;
  if (!setup_ray(P, tile, lane, r, n_nom)) return;
  const int F = P.group_n > 1 ? P.group_n : 1;
  const RaycastParams* Q[kGroupMax];
  Q[0] = &P;
#pragma unroll
  for (int f = 1; f < kGroupMax; ++f) Q[f] = f < F ? &tab->jobs[P.group_jobs[f - 1]] : &P;
  float C[kGroupMax][3], A[kGroupMax];
#pragma unroll
  for (int f = 0; f < kGroupMax; ++f) C[f][0] = C[f][1] = C[f][2] = A[f] = 0.f;
  const float t0 = r.t0, L = r.t1 - r.t0;
  if (r.hit && L > 0.f) {
    const float step = P.step;
    const int n = (int)ceilf(L / step);
    const float lastlen = fmaxf(L - (float)(n - 1) * step, 0.f);
    const LodInfo& lv = P.level;
    const int U = P.cell;
    const int per_brick = P.brick / U;
    const float invU = 1.f / (float)U;
    float od[3], oo[3], ul0[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      od[i] = r.d[i] * lv.to_level[i];
      oo[i] = r.o[i] * lv.to_level[i];
      ul0[i] = (float)U / lv.to_level[i];
    }
    const int ucount[3] = {(lv.dims[0] + U - 1) / U, (lv.dims[1] + U - 1) / U,
                           (lv.dims[2] + U - 1) / U};
    const SlotGeo G = slot_geo(P.slot_edge);
    float Lh[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // light direction, half vector
    {
      float h[3], hl = 0.f;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        Lh[i] = P.light[i];
        h[i] = P.light[i] - r.d[i];
        hl += h[i] * h[i];
      }
      hl = sqrtf(hl);
      if (hl > 1e-12f)
#pragma unroll
        for (int i = 0; i < 3; ++i) Lh[3 + i] = h[i] / hl;
    }
    uint32_t live = (1u << F) - 1;  // members still compositing
    uint32_t need = 0;              // live members whose TF keeps the current unit
    int k = 0, kend = 0, blend_gid = -1;
    float g0[3] = {0.f, 0.f, 0.f};
    const T* __restrict__ base = nullptr;
    while (k < n) {
      if (k >= kend && !group_unit<T, LINEAR>(P, Q, F, r, od, oo, ul0, ucount, invU, per_brick, k, n, t0, step,
                                              lastlen, live, kend, need, base, g0, blend_gid, C, A)) {
        k = kend;
        continue;
      }
      float v[BATCH], len[BATCH], g[BATCH][3];
#pragma unroll
      for (int j = 0; j < BATCH; ++j) {
        const int kk = min(k + j, kend - 1);
        len[j] = (kk == n - 1) ? lastlen : step;
        const float t = fmaf((float)kk, step, t0) + 0.5f * len[j];
        const float gx = fmaf(t, od[0], g0[0]), gy = fmaf(t, od[1], g0[1]), gz = fmaf(t, od[2], g0[2]);
        float dd[3];
        if (LINEAR) {
          fetch_value_grad<T>(base, G, gx, gy, gz, v[j], dd);
        } else {
          v[j] = fetch<T, false>(base, G, gx, gy, gz);
          const float lx = gx - 0.5f, ly = gy - 0.5f, lz = gz - 0.5f;
          dd[0] = fetch<T, true>(base, G, lx + 1.f, ly, lz) - fetch<T, true>(base, G, lx - 1.f, ly, lz);
          dd[1] = fetch<T, true>(base, G, lx, ly + 1.f, lz) - fetch<T, true>(base, G, lx, ly - 1.f, lz);
          dd[2] = fetch<T, true>(base, G, lx, ly, lz + 1.f) - fetch<T, true>(base, G, lx, ly, lz - 1.f);
        }
        g[j][0] = dd[0] * P.grad_scale[0];
        g[j][1] = dd[1] * P.grad_scale[1];
        g[j][2] = dd[2] * P.grad_scale[2];
      }
#pragma unroll
      for (int j = 0; j < BATCH; ++j) {
        if (k + j >= kend || !need) continue;
        ++n_walk;  // a sample some member evaluates
        const float gm = sqrtf(g[j][0] * g[j][0] + g[j][1] * g[j][1] + g[j][2] * g[j][2]);
#pragma unroll
        for (int f = 0; f < kGroupMax; ++f) {
          if (!(need >> f & 1u)) continue;
          ++n_eval[f];
          float cs[4];
          if (P.mode == 1) tf2d_lookup(*Q[f], v[j], gm, cs);
          else tf1d_lookup<false>(*Q[f], (const float4*)Q[f]->tf, v[j], cs);
          if (cs[3] > 0.f) {
            if (P.lighting) shade(P, g[j], Lh, cs);
            const float ac = cs[3] >= 1.f ? 1.f : 1.f - exp2f(len[j] * log2f(1.f - cs[3]));
            const float w = (1.f - A[f]) * ac;
            C[f][0] = fmaf(w, cs[0], C[f][0]);
            C[f][1] = fmaf(w, cs[1], C[f][1]);
            C[f][2] = fmaf(w, cs[2], C[f][2]);
            A[f] += w;
            if (!(A[f] < P.term_alpha)) {
              live &= ~(1u << f);
              need &= ~(1u << f);
            }
          }
        }
      }
      k = need ? min(k + BATCH, kend) : (live ? kend : n);
    }
    if (P.blend_bits < 32)
#pragma unroll
      for (int f = 0; f < kGroupMax; ++f) blend_round(P.blend_bits, C[f][0], C[f][1], C[f][2], A[f]);
  }
  const int64_t pix = (int64_t)r.py * P.width + r.px;
#pragma unroll
  for (int f = 0; f < kGroupMax; ++f)
    if (f < F) gstore(Q[f]->framebuffer + pix, over_background(P, r.py, C[f][0], C[f][1], C[f][2], A[f]));
}

// LDS-staged 1D march (kMode1DStaged: 1D TF without lighting, trilinear).
// The wave marches its tile's rays in rounds.  In each round every lane
// walks its ray over empty-space cells (skipping empty and non-resident
// ones, exactly as render_tile does) to its next sample that needs voxels;
// the wave then copies the 8^3 slot cell of the first such lane -- the
// voxels [8c, 8c + 8] along each axis, 9^3 of them, from the 2 x 3 x 3
// blocks of the blocked slot layout that hold them -- into its LDS buffer
// with all 64 lanes, and every lane whose next sample lies in that cell
// marches through the cell reading the 8 trilinear corners from LDS.  The
// samples, weights, TF lookups and compositing are those of render_tile's
// 1D march, so the image is the same bit for bit; what changes is that the
// voxels come in as contiguous block loads shared by the lanes instead of
// four scattered x-pair gathers per sample.
constexpr int kStageEdge = 9;                  // cell of 8 + 1 voxel for the trilinear upper corner
constexpr int kStageElems = 736;               // 9^3 = 729, padded
constexpr int kStageRuns = 9;                  // (y, z) block pairs of a staged cell
constexpr int kStageRunElems = 2 * kSlotBlockElems;  // two x-blocks per run

template <typename T>
__device__ inline void stage_cell(const T* __restrict__ base, const SlotGeo& G, int cx, int cy, int cz,
                                  int lane, T* __restrict__ sbuf) {
  const int nb = slot_blocks(G.S);
  const int bx0 = 2 * cx, by0 = 2 * cy, bz0 = 2 * cz;
  constexpr int EPW = 4 / (int)sizeof(T);            // elements per dword
  constexpr int RUN_DW = kStageRunElems / EPW;        // dwords per run
  constexpr int TOTAL = kStageRuns * RUN_DW;          // dwords of the cell's blocks
  constexpr int PER_LANE = (TOTAL + 63) / 64;
  // every load first (independent, in flight together), then the LDS writes
  uint32_t w[PER_LANE];
  int src[PER_LANE];  // first element of the dword in its block (z * 20 + y * 5 + x), -1: none
  int blk[PER_LANE];  // staged block: xb | yb << 2 | zb << 4
#pragma unroll
  for (int m = 0; m < PER_LANE; ++m) {
    const int dw = lane + 64 * m;
    const int r = dw / RUN_DW, q = dw - r * RUN_DW;
    const int yb = r % 3, zb = r / 3;
    const int e0 = q * EPW;                          // first element of this dword in the run
    const int xb = e0 / kSlotBlockElems;             // 0 or 1 (a dword never straddles blocks)
    // dwords past the cell's blocks or beyond the slot (never sampled) are not loaded
    const bool ok = dw < TOTAL && bx0 + xb < nb && by0 + yb < nb && bz0 + zb < nb;
    src[m] = ok ? e0 - xb * kSlotBlockElems : -1;
    blk[m] = xb | (yb << 2) | (zb << 4);
    w[m] = 0u;
    if (ok) {
      const int64_t off = (int64_t)(bx0 + xb) * kSlotBlockElems + (int64_t)(by0 + yb) * G.by +
                          (int64_t)(bz0 + zb) * G.bz + src[m];
      w[m] = *(__attribute__((address_space(1))) const uint32_t*)(base + off);
    }
  }
#pragma unroll
  for (int m = 0; m < PER_LANE; ++m) {
    if (src[m] < 0) continue;
    const int xb = blk[m] & 3, yb = (blk[m] >> 2) & 3, zb = blk[m] >> 4;
#pragma unroll
    for (int i = 0; i < EPW; ++i) {
      const int j = src[m] + i;  // element in the block: z * 20 + y * 5 + x
      const int zi = j / 20, yi = (j - zi * 20) / 5, xi = j - zi * 20 - yi * 5;
      const int X = xb * 4 + xi, Y = yb * 4 + yi, Z = zb * 4 + zi;
      if (X < kStageEdge && Y < kStageEdge && Z < kStageEdge)
        sbuf[(Z * kStageEdge + Y) * kStageEdge + X] = (T)(w[m] >> (8 * sizeof(T) * i));
    }
  }
}

// LDS writes of one wave visible to the same wave's later reads (and the
// reads done before the next round's writes).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, bool TF_LDS>
__device__ inline void render_tile_staged(const RaycastParams& P, const float4* __restrict__ tf4, int tile,
                                          int lane, uint32_t& n_eval, uint32_t& n_nom, T* __restrict__ sbuf) {
  if (P.tile_sample > 1) {
    const uint32_t h = ((uint32_t)tile * 0x9E3779B1u + (uint32_t)P.tile_phase * 0x85EBCA77u) >> 16;
    tile = tile * P.tile_sample + (int)(h % (uint32_t)P.tile_sample);
    if (tile >= P.rect_w * P.rect_h) return;  // wave-uniform
  }
  const int tx = P.rect_x0 + tile % P.rect_w, ty = P.rect_y0 + centre_out_row(tile / P.rect_w, P.rect_h);
  const int sh = P.tile_shift;
  const int px = (tx << sh) + (lane & ((1 << sh) - 1));
  const int py = ty * (64 >> sh) + (lane >> sh);
  // every lane stays in the rounds (the staging needs all 64); lanes off
  // the image only help load
  const bool on_image = px < P.width && py < P.height;
  float o[3], d[3];
  const float fx = (float)px, fy = (float)py;
  if (P.ortho) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      o[i] = P.eye[i] + fx * P.du[i] + fy * P.dv[i];
      d[i] = P.odir[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      o[i] = P.eye[i];
      d[i] = P.dir00[i] + fx * P.du[i] + fy * P.dv[i];
    }
    const float l = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) d[i] = d[i] / l;
  }
  float t0 = -kInf, t1 = kInf, inv[3];
  bool hit = on_image;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (d[i] == 0.f) {
      inv[i] = kInf;
      if (o[i] < 0.f || o[i] > P.box[i]) hit = false;
    } else {
      inv[i] = 1.f / d[i];
      float ta = (0.f - o[i]) * inv[i], tb = (P.box[i] - o[i]) * inv[i];
      if (ta > tb) {
        const float s = ta;
        ta = tb;
        tb = s;
      }
      t0 = fmaxf(t0, ta);
      t1 = fminf(t1, tb);
    }
  }
  t0 = fmaxf(t0, 0.f);
  if (hit && t1 > t0) n_nom += (uint32_t)floorf((t1 - t0) * P.nominal_rate);
  if (P.clip_on && hit) clip_ray(P, o, d, t0, t1, hit);
  const float L = t1 - t0;
  float C0 = 0.f, C1 = 0.f, C2 = 0.f, A = 0.f;
  const float step = P.step;
  const int n = (hit && L > 0.f) ? (int)ceilf(L / step) : 0;
  const float lastlen = fmaxf(L - (float)(n - 1) * step, 0.f);
  const LodInfo& lv = P.level;
  const int U = P.cell;
  const int per_brick = P.brick / U;
  const float invU = 1.f / (float)U;
  float od[3], oo[3], ul0[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    od[i] = d[i] * lv.to_level[i];
    oo[i] = o[i] * lv.to_level[i];
    ul0[i] = (float)U / lv.to_level[i];
  }
  const int ucount[3] = {(lv.dims[0] + U - 1) / U, (lv.dims[1] + U - 1) / U, (lv.dims[2] + U - 1) / U};
  const float tfmax = (float)(P.n_tf - 1);
  const SlotGeo G = slot_geo(P.slot_edge);
  const int Smax = G.S - 2;
  int k = 0, kend = 0, gid = -1, blend_gid = -1;
  const T* __restrict__ base = nullptr;
  float g0[3] = {0.f, 0.f, 0.f};
  bool done = n == 0, live = true;
  while (true) {
    // walk to the next sample that needs voxels (a resident, non-empty unit)
    bool need = false;
    while (!done) {
      if (k >= kend) {
        if (k >= n) {
          done = true;
          break;
        }
        const float len0 = (k == n - 1) ? lastlen : step;
        const float tk = fmaf((float)k, step, t0) + 0.5f * len0;
        int c[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          c[i] = clampi((int)floorf(fmaf(tk, od[i], oo[i]) * invU), 0, ucount[i] - 1);
        float tb = kInf;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          if (d[i] != 0.f) {
            const float face = (float)(d[i] > 0.f ? c[i] + 1 : c[i]) * ul0[i];
            tb = fminf(tb, (face - o[i]) * inv[i]);
          }
        }
        kend = min(max((int)ceilf((tb - t0) / step - 0.5f), k + 1), n);
        int b[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) b[i] = c[i] / per_brick;
        gid = lv.brick_base + (b[2] * lv.grid[1] + b[1]) * lv.grid[0] + b[0];
        const uint8_t empty =
            P.skip ? gload(P.cell_empty + (lv.cell_base + (c[2] * ucount[1] + c[1]) * ucount[0] + c[0])) : 0;
        base = brick_base<T>(P, gid);
        if (P.blend_bits < 32 && gid != blend_gid) {
          blend_round(P.blend_bits, C0, C1, C2, A);
          blend_gid = gid;
        }
        if (empty) {
          k = kend;
          continue;
        }
        if (!base) {
          if (P.requests) P.requests[gid] = 1u;
          k = kend;
          continue;
        }
        if (P.touch) P.touch[gid] = P.frame;
#pragma unroll
        for (int i = 0; i < 3; ++i) g0[i] = oo[i] - (float)(b[i] * P.brick - P.apron) - 0.5f;
      }
      need = true;
      break;
    }
    const unsigned long long want = __ballot(need);
    if (want == 0ull) break;  // wave-uniform: every lane is done
    // this lane's slot cell (8^3 of slot texels) at sample k
    int cell = -1;
    float tnext = kInf;
    if (need) {
      const float len0 = (k == n - 1) ? lastlen : step;
      const float tk = fmaf((float)k, step, t0) + 0.5f * len0;
      const int ix = clampi((int)floorf(fmaf(tk, od[0], g0[0])), 0, Smax) >> 3;
      const int iy = clampi((int)floorf(fmaf(tk, od[1], g0[1])), 0, Smax) >> 3;
      const int iz = clampi((int)floorf(fmaf(tk, od[2], g0[2])), 0, Smax) >> 3;
      cell = (iz << 16) | (iy << 8) | ix;
      tnext = tk;
    }
    // the leader is the lane furthest behind along its ray: the tile's rays
    // then cross the cells front to back together, and the lanes that
    // share the leader's cell march through it in the same round
    float tmin = tnext;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) tmin = fminf(tmin, __shfl_xor(tmin, o, 64));
    const unsigned long long behind = __ballot(need && tnext == tmin);
    const int lead = __builtin_ctzll(behind ? behind : want);  // (want when no t compares equal)
    const int lead_cell = __builtin_amdgcn_readlane(cell, lead);
    const int lead_gid = __builtin_amdgcn_readlane(gid, lead);
    const uint64_t bp = (uint64_t)(uintptr_t)base;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)bp, lead);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(bp >> 32), lead);
    const T* lead_base = (const T*)(uintptr_t)(((uint64_t)hi << 32) | lo);
    const int cx = lead_cell & 0xff, cy = (lead_cell >> 8) & 0xff, cz = lead_cell >> 16;
    stage_cell<T>(lead_base, G, cx, cy, cz, lane, sbuf);
    wave_lds_sync();
    if (need && cell == lead_cell && gid == lead_gid) {
      const int ox = 8 * cx, oy = 8 * cy, oz = 8 * cz;
      while (k < kend) {
        const float lenj = (k == n - 1) ? lastlen : step;
        const float t = fmaf((float)k, step, t0) + 0.5f * lenj;
        const float gx = fmaf(t, od[0], g0[0]), gy = fmaf(t, od[1], g0[1]), gz = fmaf(t, od[2], g0[2]);
        const int ix = clampi((int)floorf(gx), 0, Smax);
        const int iy = clampi((int)floorf(gy), 0, Smax);
        const int iz = clampi((int)floorf(gz), 0, Smax);
        if ((ix >> 3) != cx || (iy >> 3) != cy || (iz >> 3) != cz) break;  // next round
        const float wx = clampf(gx - (float)ix, 0.f, 1.f);
        const float wy = clampf(gy - (float)iy, 0.f, 1.f);
        const float wz = clampf(gz - (float)iz, 0.f, 1.f);
        const T* q = sbuf + (((iz - oz) * kStageEdge + (iy - oy)) * kStageEdge + (ix - ox));
        const float c000 = (float)q[0], c100 = (float)q[1];
        const float c010 = (float)q[kStageEdge], c110 = (float)q[kStageEdge + 1];
        const float c001 = (float)q[kStageEdge * kStageEdge], c101 = (float)q[kStageEdge * kStageEdge + 1];
        const float c011 = (float)q[kStageEdge * kStageEdge + kStageEdge];
        const float c111 = (float)q[kStageEdge * kStageEdge + kStageEdge + 1];
        const float c00 = fmaf(c100 - c000, wx, c000);
        const float c10 = fmaf(c110 - c010, wx, c010);
        const float c01 = fmaf(c101 - c001, wx, c001);
        const float c11 = fmaf(c111 - c011, wx, c011);
        const float c0 = fmaf(c10 - c00, wy, c00);
        const float c1 = fmaf(c11 - c01, wy, c01);
        const float v = fmaf(c1 - c0, wz, c0);
        ++n_eval;
        const float x = clampf(v * P.tf_scale, 0.f, tfmax);
        const int ti = (int)x;
        const float fr = x - (float)ti;
        const float4 e0 = tf_entry<TF_LDS>(tf4, ti);
        const float4 e1 = tf_entry<TF_LDS>(tf4, min(ti + 1, P.n_tf - 1));
        const float a = fmaf(e1.w - e0.w, fr, e0.w);
        ++k;
        if (a > 0.f) {
          const float ac = a >= 1.f ? 1.f : 1.f - exp2f(lenj * log2f(1.f - a));
          const float w = (1.f - A) * ac;
          C0 = fmaf(w, fmaf(e1.x - e0.x, fr, e0.x), C0);
          C1 = fmaf(w, fmaf(e1.y - e0.y, fr, e0.y), C1);
          C2 = fmaf(w, fmaf(e1.z - e0.z, fr, e0.z), C2);
          A += w;
          live = A < P.term_alpha;
          if (!live) {
            done = true;
            break;
          }
        }
      }
    }
    wave_lds_sync();  // the cell's reads end before the next round's writes
  }
  if (P.blend_bits < 32) blend_round(P.blend_bits, C0, C1, C2, A);
  if (on_image) gstore(P.framebuffer + ((int64_t)py * P.width + px), over_background(P, py, C0, C1, C2, A));
}

// Empty flag per empty-space cell for one TF: the TF entries (1D) or
// columns (2D) that the cell's value range reaches hold no alpha > 0, or
// (isosurface) the cell's maximum is below the iso value.  Such cells
// contribute nothing to any ray, so the raycast skips them.
__global__ void cell_empty_kernel(const uint32_t* __restrict__ cell_range, int64_t n_cells,
                                  const int32_t* __restrict__ prefix, int n_tf, float tf_scale,
                                  int mode, float iso_value, uint8_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cells) return;
  const uint32_t r = cell_range[i];
  bool empty;
  if (mode == kModeIso) {
    empty = (float)(r >> 16) < iso_value;
  } else {
    const float tfmax = (float)(n_tf - 1);
    const int i0 = (int)clampf((float)(r & 0xffffu) * tf_scale, 0.f, tfmax);
    const int i1 = min((int)clampf((float)(r >> 16) * tf_scale, 0.f, tfmax) + 1, n_tf - 1);
    empty = prefix[i1 + 1] - prefix[i0] == 0;
  }
  out[i] = empty ? 1 : 0;
}

// Background of every pixel outside its job's tile rect (such rays miss the
// volume box): the bytes render_tile writes for a missed ray, and no
// dominant region.  Every work-group strides over every job's pixels, so
// one grid size serves any job count (graph replay).  Coalesced stores, no
// queue atomics.
__global__ __launch_bounds__(256) void background_kernel(const JobTable* __restrict__ tab) {
  const int n_jobs = tab->h.n_jobs;
  for (int job = 0; job < n_jobs; ++job) {
    const RaycastParams& P = tab->jobs[job];
    if (P.tile_first > 0) continue;  // a later part of a split job
    const int64_t npix = (int64_t)P.width * P.height;
    const int th = 64 >> P.tile_shift;
    const int x0 = P.rect_x0 << P.tile_shift, x1 = (P.rect_x0 + P.rect_w) << P.tile_shift;
    const int y0 = P.rect_y0 * th, y1 = (P.rect_y0 + P.rect_h) * th;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < npix; i += (int64_t)gridDim.x * 256) {
      const int py = (int)(i / P.width), px = (int)(i - (int64_t)py * P.width);
      if (px >= x0 && px < x1 && py >= y0 && py < y1) continue;
      gstore(P.framebuffer + i, over_background(P, py, 0.f, 0.f, 0.f, 0.f));
      if (P.region_out) gstore(P.region_out + i, (int32_t)-1);
    }
  }
}

// Sum over pixels of the squared RGB differences of two RGBA8 images
// (integer units), folded into *out (cleared by the caller).
__global__ __launch_bounds__(256) void rgb_sse_kernel(const uint32_t* __restrict__ a,
                                                      const uint32_t* __restrict__ b, int64_t n,
                                                      unsigned long long* __restrict__ out) {
  unsigned long long acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const uint32_t x = gload(a + i), y = gload(b + i);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int d = (int)((x >> (8 * c)) & 255u) - (int)((y >> (8 * c)) & 255u);
      acc += (unsigned long long)(d * d);
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) wave_atomic_add(out, acc);  // one atomic per wave
}

// Nearest-pixel enlargement of a reduced-resolution subframe into the
// framebuffer (the governor's low-resolution intermediates): integer
// arithmetic, so the host can restate it exactly.
__global__ __launch_bounds__(256) void upscale_nearest_kernel(const uint32_t* __restrict__ src, int sw,
                                                              int sh, uint32_t* __restrict__ dst, int dw,
                                                              int dh) {
  const int64_t n = (int64_t)dw * dh;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int y = (int)(i / dw), x = (int)(i - (int64_t)y * dw);
    const int sx = (int)((int64_t)x * sw / dw), sy = (int)((int64_t)y * sh / dh);
    gstore(dst + i, gload(src + (int64_t)sy * sw + sx));
  }
}

// Gradient march of a TF group in runs (kModeGradGroup at a gather batch of
// kRun): the group march above, with the samples gathered by lanes other
// than the ray's.  Rays about 2 voxels apart (1080p over a 2048-voxel
// volume) touch a cache line per lane in every voxel gather; here each
// lane walks its own ray to its next unit with samples to gather, as
// render_tile_group does, and then the wave gathers in 4 passes: in pass
// p, lane l takes sample (l & 3) of the run of ray 16p + (l >> 2) -- kRun
// consecutive samples, 2/3 of a voxel apart, of one ray per 4 lanes, which
// share their lines -- looks up every member's TF, shades, and leaves
// (colour, opacity-corrected alpha) per member in the wave's LDS slab
// (kRun passes of 64 / kRun rays in general);
// the ray's own lane then composites its run member by member in sample
// order.  Sample positions, values, lookups and the compositing arithmetic
// are those of render_tile_group, and a member's samples and termination
// are its own, so images and counts are the same bit for bit.  Every lane
// takes part in the passes, rays finished or not; the loop ends when no
// lane has a run left (a wave-wide ballot).
constexpr int kRun = 4;
template <typename T, bool LINEAR>
__device__ inline void render_tile_group_runs(const JobTable* __restrict__ tab, const RaycastParams& P,
                                              int tile, int lane, uint32_t (&n_eval)[kGroupMax],
                                              uint32_t& n_walk, uint32_t& n_nom, float4* __restrict__ slab) {
  Ray r = {};
  const bool mine = setup_ray(P, tile, lane, r, n_nom);
  const int F = P.group_n > 1 ? P.group_n : 1;
  const RaycastParams* Q[kGroupMax];
  Q[0] = &P;
#pragma unroll
  for (int f = 1; f < kGroupMax; ++f) Q[f] = f < F ? &tab->jobs[P.group_jobs[f - 1]] : &P;
  float C[kGroupMax][3], A[kGroupMax];
#pragma unroll
  for (int f = 0; f < kGroupMax; ++f) C[f][0] = C[f][1] = C[f][2] = A[f] = 0.f;
  const float step = P.step;
  const float t0 = r.t0, L = r.t1 - r.t0;
  const int n = (mine && r.hit && L > 0.f) ? (int)ceilf(L / step) : 0;
  const float lastlen = n > 0 ? fmaxf(L - (float)(n - 1) * step, 0.f) : 0.f;
  const LodInfo& lv = P.level;
  const int U = P.cell;
  const int per_brick = P.brick / U;
  const float invU = 1.f / (float)U;
  float od[3], oo[3], ul0[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    od[i] = r.d[i] * lv.to_level[i];
    oo[i] = r.o[i] * lv.to_level[i];
    ul0[i] = (float)U / lv.to_level[i];
  }
  const int ucount[3] = {(lv.dims[0] + U - 1) / U, (lv.dims[1] + U - 1) / U, (lv.dims[2] + U - 1) / U};
  const SlotGeo G = slot_geo(P.slot_edge);
  uint32_t live = n > 0 ? (1u << F) - 1 : 0u;  // members still compositing
  uint32_t need = 0;                           // live members whose TF keeps the current unit
  int k = 0, kend = 0, blend_gid = -1;
  float g0[3] = {0.f, 0.f, 0.f};
  const T* __restrict__ base = nullptr;
  while (true) {
    // this lane's ray: on to the next unit with samples to gather
    bool ready = false;
    while (live && k < n) {
      if (k < kend) {
        ready = true;
        break;
      }
      if (!group_unit<T, LINEAR>(P, Q, F, r, od, oo, ul0, ucount, invU, per_brick, k, n, t0, step, lastlen, live,
                                 kend, need, base, g0, blend_gid, C, A))
        k = kend;
    }
    if (__ballot(ready) == 0) break;  // wave-uniform: every lane reaches it
    const int cnt = ready ? min(kRun, kend - k) : 0;
    const uint32_t run_need = ready ? need : 0u;
    const int ib0 = (int)(uint32_t)(uint64_t)base, ib1 = (int)(uint32_t)((uint64_t)base >> 32);
    constexpr int kRays = 64 / kRun;  // rays whose runs a pass gathers
    for (int p = 0; p < kRun; ++p) {
      // this lane gathers sample q of the run of ray `src`
      const int src = kRays * p + lane / kRun, q = lane % kRun;
      const int s_cnt = __shfl(cnt, src, 64);
      const uint32_t s_need = (uint32_t)__shfl((int)run_need, src, 64);
      const int s_k = __shfl(k, src, 64), s_n = __shfl(n, src, 64);
      const float s_t0 = __shfl(t0, src, 64), s_last = __shfl(lastlen, src, 64);
      float s_od[3], s_g0[3], s_d[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        s_od[i] = __shfl(od[i], src, 64);
        s_g0[i] = __shfl(g0[i], src, 64);
        s_d[i] = __shfl(r.d[i], src, 64);
      }
      const uint64_t sb = (uint64_t)(uint32_t)__shfl(ib0, src, 64) | (uint64_t)(uint32_t)__shfl(ib1, src, 64) << 32;
      if (q < s_cnt) {
        const T* __restrict__ sbase = (const T*)sb;
        const int kk = s_k + q;
        const float len = (kk == s_n - 1) ? s_last : step;
        const float t = fmaf((float)kk, step, s_t0) + 0.5f * len;
        const float gx = fmaf(t, s_od[0], s_g0[0]), gy = fmaf(t, s_od[1], s_g0[1]), gz = fmaf(t, s_od[2], s_g0[2]);
        float v, dd[3];
        if (LINEAR) {
          fetch_value_grad<T>(sbase, G, gx, gy, gz, v, dd);
        } else {
          v = fetch<T, false>(sbase, G, gx, gy, gz);
          const float lx = gx - 0.5f, ly = gy - 0.5f, lz = gz - 0.5f;
          dd[0] = fetch<T, true>(sbase, G, lx + 1.f, ly, lz) - fetch<T, true>(sbase, G, lx - 1.f, ly, lz);
          dd[1] = fetch<T, true>(sbase, G, lx, ly + 1.f, lz) - fetch<T, true>(sbase, G, lx, ly - 1.f, lz);
          dd[2] = fetch<T, true>(sbase, G, lx, ly, lz + 1.f) - fetch<T, true>(sbase, G, lx, ly, lz - 1.f);
        }
        float g[3] = {dd[0] * P.grad_scale[0], dd[1] * P.grad_scale[1], dd[2] * P.grad_scale[2]};
        const float gm = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        // the source ray's light direction and half vector, as its own march sets them up
        float Lh[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        {
          float h[3], hl = 0.f;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            Lh[i] = P.light[i];
            h[i] = P.light[i] - s_d[i];
            hl += h[i] * h[i];
          }
          hl = sqrtf(hl);
          if (hl > 1e-12f)
#pragma unroll
            for (int i = 0; i < 3; ++i) Lh[3 + i] = h[i] / hl;
        }
#pragma unroll
        for (int f = 0; f < kGroupMax; ++f) {
          if (!(s_need >> f & 1u)) continue;
          float cs[4];
          if (P.mode == 1) tf2d_lookup(*Q[f], v, gm, cs);
          else tf1d_lookup<false>(*Q[f], (const float4*)Q[f]->tf, v, cs);
          float ac = 0.f;  // compositing with 0 leaves (C, A) as skipping the sample does
          if (cs[3] > 0.f) {
            if (P.lighting) shade(P, g, Lh, cs);
            ac = cs[3] >= 1.f ? 1.f : 1.f - exp2f(len * log2f(1.f - cs[3]));
          }
          slab[f * 64 + lane] = make_float4(cs[0], cs[1], cs[2], ac);
        }
      }
      wave_lds_sync();
      if (lane / kRays == p && cnt > 0) {
        // this lane's ray was gathered in this pass by lanes kRun (lane - kRays p) + q
        const int w0 = kRun * (lane - kRays * p);
        int most = 0;  // samples of the run some member evaluates
#pragma unroll
        for (int f = 0; f < kGroupMax; ++f) {
          if (!(run_need >> f & 1u)) continue;
          int e = 0;
          for (int j = 0; j < cnt && (live >> f & 1u); ++j) {
            const float4 x = slab[f * 64 + w0 + j];
            ++e;
            const float w = (1.f - A[f]) * x.w;
            C[f][0] = fmaf(w, x.x, C[f][0]);
            C[f][1] = fmaf(w, x.y, C[f][1]);
            C[f][2] = fmaf(w, x.z, C[f][2]);
            A[f] += w;
            if (!(A[f] < P.term_alpha)) {
              live &= ~(1u << f);
              need &= ~(1u << f);
            }
          }
          n_eval[f] += (uint32_t)e;
          most = max(most, e);
        }
        n_walk += (uint32_t)most;
      }
      wave_lds_sync();  // the slab is the next pass's
    }
    if (ready) k = live ? (need ? k + cnt : kend) : n;
  }
  if (n > 0 && P.blend_bits < 32)
#pragma unroll
    for (int f = 0; f < kGroupMax; ++f) blend_round(P.blend_bits, C[f][0], C[f][1], C[f][2], A[f]);
  if (!mine) return;
  const int64_t pix = (int64_t)r.py * P.width + r.px;
#pragma unroll
  for (int f = 0; f < kGroupMax; ++f)
    if (f < F) gstore(Q[f]->framebuffer + pix, over_background(P, r.py, C[f][0], C[f][1], C[f][2], A[f]));
}

// Persistent raycast over a job table: every wave loops over ray tiles
// taken from the per-XCD queues until all queues are drained.  Tile t of
// the launch belongs to job j with tile_start[j] <= t < tile_start[j + 1].
template <typename T, bool LINEAR, int BATCH, bool TF_LDS, int MODE>
__global__ __launch_bounds__(kThreads, 2) void raycast_kernel(const JobTable* __restrict__ tab) {
  const int kT = (int)blockDim.x;  // 256 or 512 threads (see launch_one)
  extern __shared__ float4 s_tf[];  // n_tf_lds entries when TF_LDS
  __shared__ unsigned long long s_count[kJobCounters * kMaxJobs];  // per job: evaluated, nominal, gathered
  const JobTableHeader& H = tab->h;
  const float4* tf_shared = (const float4*)H.tf_lds;
  if (TF_LDS) {
    for (int i = (int)threadIdx.x; i < H.n_tf_lds; i += kT) s_tf[i] = tf_shared[i];
    tf_shared = s_tf;
  }
  for (int i = (int)threadIdx.x; i < kJobCounters * kMaxJobs; i += kT) s_count[i] = 0;
  __syncthreads();
  const int lane = (int)(threadIdx.x & 63);
  const int xcd = (int)(blockIdx.x & 7);
  unsigned long long* queues = H.counters;
  const int n_jobs = H.n_jobs;
  const int total = H.total_tiles;
  const int G = H.tile_group;
  // xcd_split regions apply to one-job launches only
  const int split = n_jobs == 1 ? H.xcd_split : 0;
  const int tiles_x0 = tab->jobs[0].rect_w, tiles_y0 = tab->jobs[0].rect_h;
  // per-lane sample counts of the job the wave is on; added to the
  // work-group's LDS counters when the wave moves to another job and at the
  // end, then the work-group adds its counters to the job's global ones
  int cur_job = -1;
  uint32_t n_eval = 0, n_nom = 0;
  uint32_t n_evalg[kGroupMax] = {0, 0, 0, 0};  // kModeGradGroup: per member of the job's group
  uint32_t n_walk = 0;                          // kModeGradGroup: samples the group's walk gathered
  // adds the wave's counts of cur_job (and of its group's members) to the
  // work-group's counters
  auto flush = [&]() {
    if constexpr (MODE == kModeGradGroup) {
      const RaycastParams& Pc = tab->jobs[cur_job];
      const int F = Pc.group_n > 1 ? Pc.group_n : 1;
#pragma unroll
      for (int f = 0; f < kGroupMax; ++f) {
        if (f < F) {
          const int j = f == 0 ? cur_job : Pc.group_jobs[f > 0 ? f - 1 : 0];
          atomicAdd(&s_count[kJobCounters * j], (unsigned long long)n_evalg[f]);
          atomicAdd(&s_count[kJobCounters * j + 1], (unsigned long long)n_nom);
        }
        n_evalg[f] = 0;
      }
      atomicAdd(&s_count[kJobCounters * cur_job + 2], (unsigned long long)n_walk);
      n_walk = 0;
    } else {
      atomicAdd(&s_count[kJobCounters * cur_job], (unsigned long long)n_eval);
      atomicAdd(&s_count[kJobCounters * cur_job + 1], (unsigned long long)n_nom);
      atomicAdd(&s_count[kJobCounters * cur_job + 2], (unsigned long long)n_eval);
    }
  };
  for (int q = 0; q < kQueues; ++q) {
    const int x = (xcd + q) & 7;  // own queue first, then steal
    // Queue x enumerates either runs of G tiles dealt round-robin
    // (split == 0) or, with split = rx in {1,2,4,8}, the tiles of region x
    // of an rx by (8/rx) split of the frame in row-major order.
    int rx0 = 0, rw = 1, ry0 = 0, slots;
    if (split == 0) {
      const int ngroups = (total + G - 1) / G;
      slots = ((ngroups + 7) / 8) * G;
    } else {
      const int rx = split, ry = 8 / split;
      const int i = x % rx, jr = x / rx;
      rx0 = i * tiles_x0 / rx;
      rw = (i + 1) * tiles_x0 / rx - rx0;
      ry0 = jr * tiles_y0 / ry;
      const int rh = (jr + 1) * tiles_y0 / ry - ry0;
      slots = rw * rh;
    }
    while (true) {
      // Every lane takes part in the queue atomic (lanes other than 0 add
      // 0; the compiler reduces it to one atomic per wave) and all branches
      // in this loop are wave-uniform: no lane-0-only block in the source
      // that the compiler could split off the loop around the wave-wide
      // broadcast of the ticket.
      const unsigned long long ticket = wave_atomic_add(&queues[x], lane == 0 ? 1ull : 0ull);
      const int pos = __builtin_amdgcn_readfirstlane((int)__shfl((long long)ticket, 0, 64));
      if (pos >= slots) break;
      int tile;
      if (split == 0) tile = ((pos / G) * 8 + x) * G + (pos % G);
      else tile = (ry0 + pos / rw) * tiles_x0 + rx0 + pos % rw;
      if (tile >= total) continue;
      // the job of this tile (wave-uniform binary search over the prefix)
      int lo = 0, hi = n_jobs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab->tile_start[mid] <= tile) lo = mid;
        else hi = mid - 1;
      }
      const int job = __builtin_amdgcn_readfirstlane(lo);
      const RaycastParams& P = tab->jobs[job];
      const float4* tf4 = TF_LDS ? tf_shared : (const float4*)P.tf;
      if (job != cur_job) {  // wave-uniform
        if (cur_job >= 0) flush();
        cur_job = job;
        n_eval = n_nom = 0;
      }
      // the job's own tile slot (a part's slots start at tile_first)
      const int slot = tile - tab->tile_start[job] + P.tile_first;
      if constexpr (MODE == kMode1DStaged) {
        // this wave's staging buffer follows the LDS TF
        T* sbuf = (T*)(s_tf + (TF_LDS ? H.n_tf_lds : 0)) + (threadIdx.x >> 6) * kStageElems;
        render_tile_staged<T, TF_LDS>(P, tf4, slot, lane, n_eval, n_nom, sbuf);
      } else if constexpr (MODE == kModeGradGroup && BATCH == kRun) {
        // this wave's slab follows the work-group's counters in LDS
        float4* slab = s_tf + (threadIdx.x >> 6) * 64 * kGroupMax;
        render_tile_group_runs<T, LINEAR>(tab, P, slot, lane, n_evalg, n_walk, n_nom,
                                          slab);
      } else if constexpr (MODE == kModeGradGroup) {
        render_tile_group<T, LINEAR, BATCH>(tab, P, slot, lane, n_evalg, n_walk, n_nom);
      } else {
        render_tile<T, LINEAR, BATCH, MODE, TF_LDS>(P, tf4, slot, lane, n_eval, n_nom);
      }
    }
  }
  if (cur_job >= 0) flush();
  __syncthreads();
  for (int i = (int)threadIdx.x; i < kJobCounters * n_jobs; i += kT)
    if (s_count[i] != 0) atomicAdd(&H.counters[kQueues + i], s_count[i]);
}

template <typename T, bool LINEAR, int BATCH, bool TF_LDS, int MODE>
hipError_t launch_one(const LaunchInfo& li, const JobTable* tab, hipStream_t stream) {
  auto kern = raycast_kernel<T, LINEAR, BATCH, TF_LDS, MODE>;
  size_t lds = TF_LDS ? (size_t)li.n_tf_lds * sizeof(float4) : 0;
  int per_cu = 0;
  // Work-group size: 8 waves when the 1D TF is staged in LDS (64 KB per
  // work-group, so two fit a CU's 160 KB); otherwise li.wg_waves (4 or 8).
  // The variants without an LDS TF take 114-173 VGPRs (2-4 waves per SIMD;
  // the gradient marches 114-128 at 1 sample per batch): 8-wave
  // work-groups fit a CU once or twice, 4-wave ones up to four times,
  // which is faster on cache-friendly volumes and slower where more waves
  // in flight thrash the caches (the autotuner picks).
  const int kT = TF_LDS ? kThreads : 64 * (li.wg_waves == 4 ? 4 : 8);
  if (MODE == kMode1DStaged) lds += (size_t)(kT / 64) * kStageElems * sizeof(T);  // per-wave cell buffers
  if (MODE == kModeGradGroup && BATCH == kRun) lds += (size_t)(kT / 64) * 64 * kGroupMax * sizeof(float4);
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kT, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  int grid = li.compute_units * per_cu;
  grid = ((grid + 7) / 8) * 8;
  kern<<<grid, kT, lds, stream>>>(tab);
  return hipGetLastError();
}

// The gradient variants (2D TF, or 1D TF with lighting): TF groups, or one
// march per view with the 1D TF in LDS when the launch shares one.
template <typename T, bool LINEAR, int BATCH>
hipError_t launch_grad(const LaunchInfo& li, const JobTable* tab, hipStream_t stream) {
  if (li.group) return launch_one<T, LINEAR, BATCH, false, kModeGradGroup>(li, tab, stream);
  return li.mode == 0 && li.n_tf_lds > 0 ? launch_one<T, LINEAR, BATCH, true, kModeGrad>(li, tab, stream)
                                         : launch_one<T, LINEAR, BATCH, false, kModeGrad>(li, tab, stream);
}

template <typename T, bool LINEAR>
hipError_t launch_typed(const LaunchInfo& li, const JobTable* tab, hipStream_t stream) {
  if (li.regions)
    return li.n_tf_lds > 0 ? launch_one<T, LINEAR, 4, true, kModeRegions>(li, tab, stream)
                           : launch_one<T, LINEAR, 4, false, kModeRegions>(li, tab, stream);
  if (li.mode == 2) return launch_one<T, LINEAR, 4, false, kModeIso>(li, tab, stream);
  if (li.mode == 3)
    return li.n_tf_lds > 0 ? launch_one<T, LINEAR, 4, true, kModeMIP>(li, tab, stream)
                           : launch_one<T, LINEAR, 4, false, kModeMIP>(li, tab, stream);
  // the gradient marches: li.batch 1 or 2 samples per gather batch, each
  // lane its own ray; otherwise runs of kRun samples per ray gathered
  // across lanes (the TF-group march, which takes lone views as groups of
  // one)
  if (li.mode == 1 || li.lighting) {
    if (li.batch == 1) return launch_grad<T, LINEAR, 1>(li, tab, stream);
    if (li.batch == 2) return launch_grad<T, LINEAR, 2>(li, tab, stream);
    return launch_one<T, LINEAR, kRun, false, kModeGradGroup>(li, tab, stream);
  }
  const bool lds = li.n_tf_lds > 0;
  if (li.stage && LINEAR && !li.lighting)
    return lds ? launch_one<T, true, 1, true, kMode1DStaged>(li, tab, stream)
               : launch_one<T, true, 1, false, kMode1DStaged>(li, tab, stream);
  switch (li.batch) {
    case 2:
      return lds ? launch_one<T, LINEAR, 2, true, kMode1D>(li, tab, stream)
                 : launch_one<T, LINEAR, 2, false, kMode1D>(li, tab, stream);
    case 4:
      return lds ? launch_one<T, LINEAR, 4, true, kMode1D>(li, tab, stream)
                 : launch_one<T, LINEAR, 4, false, kMode1D>(li, tab, stream);
    default:
      return lds ? launch_one<T, LINEAR, 8, true, kMode1D>(li, tab, stream)
                 : launch_one<T, LINEAR, 8, false, kMode1D>(li, tab, stream);
  }
}

}  // namespace

hipError_t launch_cell_empty(const uint32_t* cell_range, int64_t n_cells, const int32_t* prefix,
                             int n_tf, float tf_scale, int mode, float iso_value, uint8_t* out,
                             hipStream_t stream) {
  if (n_cells <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((n_cells + 255) / 256);
  cell_empty_kernel<<<blocks, 256, 0, stream>>>(cell_range, n_cells, prefix, n_tf, tf_scale,
                                                mode == 2 ? kModeIso : mode, iso_value, out);
  return hipGetLastError();
}

hipError_t launch_rgb_sse(const uint32_t* a, const uint32_t* b, int64_t n, unsigned long long* out,
                          hipStream_t stream) {
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(rgb_sse_kernel, dim3(blocks), dim3(256), 0, stream, a, b, n, out);
  return hipGetLastError();
}

hipError_t launch_upscale_nearest(const uint32_t* src, int sw, int sh, uint32_t* dst, int dw, int dh,
                                  hipStream_t stream) {
  const int64_t n = (int64_t)dw * dh;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(upscale_nearest_kernel, dim3(blocks), dim3(256), 0, stream, src, sw, sh, dst, dw, dh);
  return hipGetLastError();
}

hipError_t launch_background(const JobTable* tab, hipStream_t stream) {
  background_kernel<<<kBackgroundBlocks, 256, 0, stream>>>(tab);
  return hipGetLastError();
}

hipError_t launch_raycast(const LaunchInfo& li, const JobTable* tab, hipStream_t stream) {
  if (li.bpv == 1)
    return li.interp ? launch_typed<uint8_t, true>(li, tab, stream)
                     : launch_typed<uint8_t, false>(li, tab, stream);
  return li.interp ? launch_typed<uint16_t, true>(li, tab, stream)
                   : launch_typed<uint16_t, false>(li, tab, stream);
}

}  // namespace tvk
