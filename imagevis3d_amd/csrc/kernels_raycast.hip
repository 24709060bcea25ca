// kernels_raycast.hip -- ray-brick traversal + front-to-back compositing
// (the work of Tuvok's GLRaycaster / GLGridLeaper fragment shaders,
// reached through AbstrRenderer::Paint, ImageVis3D/UI/RenderWindow.cpp:2464).
//
// Mapping to CDNA4: one 64-lane wavefront renders one 8x8 ray tile, so the
// 64 rays of a wave march through neighbouring voxels and their gathers hit
// the same L1/L2 lines.  Work-groups are remapped so each of the 8 XCDs
// renders one contiguous band of tiles (its own L2 holds the bricks that
// band sees).  Each ray walks the brick grid of the chosen LOD; per brick it
// looks up the pool slot in the page table, skips the brick when the 1D TF
// is transparent over the brick's [min, max] (exact: such samples add
// nothing), and otherwise marches the brick's samples with trilinear
// fetches from the slot, whose apron makes every fetch brick-local.  Rays
// stop at the early-termination opacity.
//
// The sample positions, segment lengths, opacity correction and blending
// follow the model stated in oracle/oracle.c (the FP64 oracle it is tested
// against); positions are global per ray, so the image does not depend on
// the brick size.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tvk {
namespace {

constexpr float kInf = __builtin_huge_valf();

__device__ inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ inline int clampi(int x, int lo, int hi) { return min(max(x, lo), hi); }

__device__ inline uint32_t pack_rgba8(float r, float g, float b, float a) {
  const uint32_t R = (uint32_t)floorf(clampf(r, 0.f, 1.f) * 255.f + 0.5f);
  const uint32_t G = (uint32_t)floorf(clampf(g, 0.f, 1.f) * 255.f + 0.5f);
  const uint32_t B = (uint32_t)floorf(clampf(b, 0.f, 1.f) * 255.f + 0.5f);
  const uint32_t A = (uint32_t)floorf(clampf(a, 0.f, 1.f) * 255.f + 0.5f);
  return R | (G << 8) | (B << 16) | (A << 24);
}

template <typename T>
__device__ inline float ld(const T* p) {
  return (float)*p;
}

template <typename T, bool LINEAR>
__global__ __launch_bounds__(64) void raycast1d_kernel(const RaycastParams P) {
  // XCD-aware tile order: the dispatcher deals work-groups round-robin to
  // the 8 XCDs, so work-group w runs on XCD w%8; give that XCD the
  // contiguous tile band [xcd*per_xcd, (xcd+1)*per_xcd).
  const int ntiles = P.tiles_x * P.tiles_y;
  const int per_xcd = (ntiles + 7) >> 3;
  const int tile = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (tile >= ntiles) return;  // uniform per wave
  const int tx = tile % P.tiles_x, ty = tile / P.tiles_x;
  const int px = tx * kTile + (int)(threadIdx.x & 7);
  const int py = ty * kTile + (int)(threadIdx.x >> 3);

  uint32_t n_eval = 0, n_nom = 0;
  if (px < P.width && py < P.height) {
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
    // ray vs LOD-0 box [0, box]; axis-parallel rays use an inside test
    float t0 = -kInf, t1 = kInf, inv[3];
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
    const float L = t1 - t0;
    float C0 = 0.f, C1 = 0.f, C2 = 0.f, A = 0.f;
    if (hit && L > 0.f) {
      n_nom = (uint32_t)floorf(L * P.nominal_rate);
      const float step = P.step;
      const int n = (int)ceilf(L / step);
      const float lastlen = fmaxf(L - (float)(n - 1) * step, 0.f);
      const LodInfo& lv = P.level;
      const float invB = 1.f / (float)P.brick;
      float od[3], oo[3], bl0[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        od[i] = d[i] * lv.to_level[i];
        oo[i] = o[i] * lv.to_level[i];
        bl0[i] = (float)P.brick / lv.to_level[i];  // brick edge in LOD-0 voxels
      }
      const float4* __restrict__ tf4 = (const float4*)P.tf;
      const float tfmax = (float)(P.n_tf - 1);
      const int S = P.slot_edge;
      const int S2 = S * S;
      int k = 0;
      while (k < n) {
        const float len0 = (k == n - 1) ? lastlen : step;
        const float tk = fmaf((float)k, step, t0) + 0.5f * len0;
        int b[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          b[i] = clampi((int)floorf(fmaf(tk, od[i], oo[i]) * invB), 0, lv.grid[i] - 1);
        float tb = kInf;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          if (d[i] != 0.f) {
            const float face = (float)(d[i] > 0.f ? b[i] + 1 : b[i]) * bl0[i];
            tb = fminf(tb, (face - o[i]) * inv[i]);
          }
        }
        int kend = (int)ceilf((tb - t0) / step - 0.5f);
        kend = min(max(kend, k + 1), n);
        const int gid = lv.brick_base + (b[2] * lv.grid[1] + b[1]) * lv.grid[0] + b[0];
        const int slot = P.page_table[gid];
        bool empty = slot < 0;
        if (!empty && P.skip) {
          const float2 r = ((const float2*)P.brick_range)[gid];
          const int i0 = (int)clampf(r.x * P.tf_scale, 0.f, tfmax);
          const int i1 = min((int)clampf(r.y * P.tf_scale, 0.f, tfmax) + 1, P.n_tf - 1);
          empty = (P.tf_prefix[i1 + 1] - P.tf_prefix[i0]) == 0;
        }
        if (empty) {
          k = kend;
          continue;
        }
        const T* __restrict__ base = (const T*)P.pool + (int64_t)slot * P.slot_elems;
        float g0[3];  // slot-local texel coordinate at t = 0
#pragma unroll
        for (int i = 0; i < 3; ++i)
          g0[i] = oo[i] - (float)(b[i] * P.brick - P.apron) - (LINEAR ? 0.5f : 0.f);
        for (; k < kend; ++k) {
          const float len = (k == n - 1) ? lastlen : step;
          const float t = fmaf((float)k, step, t0) + 0.5f * len;
          const float gx = fmaf(t, od[0], g0[0]);
          const float gy = fmaf(t, od[1], g0[1]);
          const float gz = fmaf(t, od[2], g0[2]);
          float v;
          if (LINEAR) {
            const float fx0 = floorf(gx), fy0 = floorf(gy), fz0 = floorf(gz);
            const int ix = clampi((int)fx0, 0, S - 2);
            const int iy = clampi((int)fy0, 0, S - 2);
            const int iz = clampi((int)fz0, 0, S - 2);
            const float wx = clampf(gx - (float)ix, 0.f, 1.f);
            const float wy = clampf(gy - (float)iy, 0.f, 1.f);
            const float wz = clampf(gz - (float)iz, 0.f, 1.f);
            const T* p = base + (iz * S2 + iy * S + ix);
            const float c000 = ld(p), c100 = ld(p + 1);
            const float c010 = ld(p + S), c110 = ld(p + S + 1);
            const float c001 = ld(p + S2), c101 = ld(p + S2 + 1);
            const float c011 = ld(p + S2 + S), c111 = ld(p + S2 + S + 1);
            const float c00 = fmaf(c100 - c000, wx, c000);
            const float c10 = fmaf(c110 - c010, wx, c010);
            const float c01 = fmaf(c101 - c001, wx, c001);
            const float c11 = fmaf(c111 - c011, wx, c011);
            const float c0 = fmaf(c10 - c00, wy, c00);
            const float c1 = fmaf(c11 - c01, wy, c01);
            v = fmaf(c1 - c0, wz, c0);
          } else {
            const int ix = clampi((int)floorf(gx), 0, S - 1);
            const int iy = clampi((int)floorf(gy), 0, S - 1);
            const int iz = clampi((int)floorf(gz), 0, S - 1);
            v = ld(base + (iz * S2 + iy * S + ix));
          }
          ++n_eval;
          const float x = clampf(v * P.tf_scale, 0.f, tfmax);
          const int ti = (int)x;
          const float fr = x - (float)ti;
          const int tj = min(ti + 1, P.n_tf - 1);
          const float4 e0 = tf4[ti], e1 = tf4[tj];
          const float a = fmaf(e1.w - e0.w, fr, e0.w);
          if (a > 0.f) {
            const float ac = a >= 1.f ? 1.f : 1.f - exp2f(len * log2f(1.f - a));
            const float w = (1.f - A) * ac;
            C0 = fmaf(w, fmaf(e1.x - e0.x, fr, e0.x), C0);
            C1 = fmaf(w, fmaf(e1.y - e0.y, fr, e0.y), C1);
            C2 = fmaf(w, fmaf(e1.z - e0.z, fr, e0.z), C2);
            A += w;
            if (A >= P.term_alpha) {
              k = n;
              break;
            }
          }
        }
      }
    }
    const float wy = 1.f - ((float)py + 0.5f) / (float)P.height;
    const float bg0 = P.bg_bottom[0] + (P.bg_top[0] - P.bg_bottom[0]) * wy;
    const float bg1 = P.bg_bottom[1] + (P.bg_top[1] - P.bg_bottom[1]) * wy;
    const float bg2 = P.bg_bottom[2] + (P.bg_top[2] - P.bg_bottom[2]) * wy;
    const float T1 = 1.f - A;
    P.framebuffer[(int64_t)py * P.width + px] =
        pack_rgba8(fmaf(T1, bg0, C0), fmaf(T1, bg1, C1), fmaf(T1, bg2, C2), fmaf(T1, P.bg_alpha, A));
  }
  // one atomic per wave for each counter
  unsigned long long e = n_eval, m = n_nom;
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    m += __shfl_xor(m, off, 64);
  }
  if (threadIdx.x == 0) {
    atomicAdd(&P.counters[0], e);
    atomicAdd(&P.counters[1], m);
  }
}

}  // namespace

hipError_t launch_raycast_1d(const RaycastParams& p, int bpv, hipStream_t stream) {
  const int ntiles = p.tiles_x * p.tiles_y;
  const int grid = 8 * ((ntiles + 7) / 8);
  if (bpv == 1) {
    if (p.interp)
      raycast1d_kernel<uint8_t, true><<<grid, 64, 0, stream>>>(p);
    else
      raycast1d_kernel<uint8_t, false><<<grid, 64, 0, stream>>>(p);
  } else {
    if (p.interp)
      raycast1d_kernel<uint16_t, true><<<grid, 64, 0, stream>>>(p);
    else
      raycast1d_kernel<uint16_t, false><<<grid, 64, 0, stream>>>(p);
  }
  return hipGetLastError();
}

}  // namespace tvk
