/*
 * baseline.c -- FP32 multi-threaded CPU raycaster: the same render model as
 * the FP64 oracle (oracle.c) and the HIP kernels, written for host cores
 * (OpenMP over 8x8 pixel tiles, brick-granular empty-space skipping).  It is
 * the timed CPU baseline of the benchmark.  TEST INFRASTRUCTURE ONLY.
 */
#include "oracle_common.h"
#ifdef _OPENMP
#include <omp.h>
#endif

/* Built twice (Makefile): once for AVX2 + FMA and once for AVX-512; the
 * entry point baseline_render_1d_f32 (baseline_dispatch.c) picks the one
 * the host's cores run. */
#ifndef BASELINE_FN
#define BASELINE_FN baseline_render_1d_f32_avx2
#endif

uint64_t BASELINE_FN(const oracle_render_params* p, uint8_t* out8, int32_t n_threads) {
  cam_ctx c;
  if (cam_setup(p, &c)) return 0;
  int nb[3] = {1, 1, 1};
  uint8_t* empty = p->brick_size > 0 ? oracle_brick_empty_flags(p, nb) : NULL;
  const int W = p->width, H = p->height;
  const int tiles_x = (W + 7) / 8, tiles_y = (H + 7) / 8;
  const float n0f[3] = {(float)p->n0x, (float)p->n0y, (float)p->n0z};
  const float lsc[3] = {(float)p->nx / p->n0x, (float)p->ny / p->n0y, (float)p->nz / p->n0z};
  const float step = (float)(ldexp(1.0, p->lod) / p->rate);
  const int nxl = p->nx, nyl = p->ny, nzl = p->nz;
  const int64_t sy = nxl, sz = (int64_t)nxl * nyl;
  const float tfs = (float)p->tf_scale, tfmax = (float)(p->n_tf - 1);
  const float thr = (float)p->term_alpha;
  const int B = p->brick_size;
  uint64_t total = 0;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
#endif
  for (int tile = 0; tile < tiles_x * tiles_y; ++tile) {
    int tx = tile % tiles_x, ty = tile / tiles_x;
    for (int ly = 0; ly < 8; ++ly)
      for (int lx = 0; lx < 8; ++lx) {
        int x = tx * 8 + lx, y = ty * 8 + ly;
        if (x >= W || y >= H) continue;
        double od[3], dd[3];
        cam_ray(p, &c, x, y, od, dd);
        float o[3] = {(float)od[0], (float)od[1], (float)od[2]};
        float d[3] = {(float)dd[0], (float)dd[1], (float)dd[2]};
        float C0 = 0, C1 = 0, C2 = 0, A = 0;
        float ta = -INFINITY, tb = INFINITY;
        int hit = 1;
        for (int i = 0; i < 3; ++i) {
          if (d[i] == 0.0f) { if (o[i] < 0.0f || o[i] > n0f[i]) hit = 0; continue; }
          float inv = 1.0f / d[i];
          float t0 = (0.0f - o[i]) * inv, t1 = (n0f[i] - o[i]) * inv;
          if (t0 > t1) { float t = t0; t0 = t1; t1 = t; }
          ta = t0 > ta ? t0 : ta;
          tb = t1 < tb ? t1 : tb;
        }
        if (hit && tb > ta) {
          if (ta < 0) ta = 0;
          float L = tb - ta;
          int n = L > 0 ? (int)ceilf(L / step) : 0;
          for (int k = 0; k < n; ++k) {
            float len = (k == n - 1) ? L - (float)(n - 1) * step : step;
            float t = ta + (float)k * step + 0.5f * len;
            float q0 = (o[0] + t * d[0]) * lsc[0], q1 = (o[1] + t * d[1]) * lsc[1], q2 = (o[2] + t * d[2]) * lsc[2];
            if (empty) {
              int b0 = clampi((int)floorf(q0 / B), 0, nb[0] - 1);
              int b1 = clampi((int)floorf(q1 / B), 0, nb[1] - 1);
              int b2 = clampi((int)floorf(q2 / B), 0, nb[2] - 1);
              if (empty[(b2 * nb[1] + b1) * nb[0] + b0]) {
                /* jump to the last sample inside this empty brick: its exit
                 * face along the ray (in LOD-0 voxels) */
                const int bb[3] = {b0, b1, b2};
                float texit = INFINITY;
                for (int i = 0; i < 3; ++i) {
                  if (d[i] == 0.0f) continue;
                  const float face = (float)((d[i] > 0.0f ? bb[i] + 1 : bb[i]) * B) / lsc[i];
                  const float tf = (face - o[i]) / d[i];
                  texit = tf < texit ? tf : texit;
                }
                /* samples before index ceil(x) - 1 have their midpoints
                 * inside; stop one short so rounding never skips a sample
                 * of the next brick (the rest are tested one by one) */
                const int kl = (int)ceilf((texit - ta) / step - 0.5f) - 2;
                if (kl > k) k = kl < n - 1 ? kl : n - 1;
                continue;
              }
            }
            float v;
            if (p->interp) {
              float g0 = q0 - 0.5f, g1 = q1 - 0.5f, g2 = q2 - 0.5f;
              float f0 = floorf(g0), f1 = floorf(g1), f2 = floorf(g2);
              int i0 = (int)f0, j0 = (int)f1, k0 = (int)f2;
              float w0 = g0 - f0, w1 = g1 - f1, w2 = g2 - f2;
              int ia = clampi(i0, 0, nxl - 1), ib = clampi(i0 + 1, 0, nxl - 1);
              int ja = clampi(j0, 0, nyl - 1), jb = clampi(j0 + 1, 0, nyl - 1);
              int ka = clampi(k0, 0, nzl - 1), kb = clampi(k0 + 1, 0, nzl - 1);
              float c000 = vox_f(p->voxels, p->bits, ia + ja * sy + ka * sz);
              float c100 = vox_f(p->voxels, p->bits, ib + ja * sy + ka * sz);
              float c010 = vox_f(p->voxels, p->bits, ia + jb * sy + ka * sz);
              float c110 = vox_f(p->voxels, p->bits, ib + jb * sy + ka * sz);
              float c001 = vox_f(p->voxels, p->bits, ia + ja * sy + kb * sz);
              float c101 = vox_f(p->voxels, p->bits, ib + ja * sy + kb * sz);
              float c011 = vox_f(p->voxels, p->bits, ia + jb * sy + kb * sz);
              float c111 = vox_f(p->voxels, p->bits, ib + jb * sy + kb * sz);
              float c00 = c000 + (c100 - c000) * w0, c10 = c010 + (c110 - c010) * w0;
              float c01 = c001 + (c101 - c001) * w0, c11 = c011 + (c111 - c011) * w0;
              float cc0 = c00 + (c10 - c00) * w1, cc1 = c01 + (c11 - c01) * w1;
              v = cc0 + (cc1 - cc0) * w2;
            } else {
              int ia = clampi((int)floorf(q0), 0, nxl - 1), ja = clampi((int)floorf(q1), 0, nyl - 1),
                  ka = clampi((int)floorf(q2), 0, nzl - 1);
              v = vox_f(p->voxels, p->bits, ia + ja * sy + ka * sz);
            }
            ++total;
            float xx = v * tfs;
            xx = xx < 0 ? 0 : (xx > tfmax ? tfmax : xx);
            int ti = (int)xx;
            float tf = xx - (float)ti;
            int tj = ti + 1 < p->n_tf ? ti + 1 : p->n_tf - 1;
            const float* e0 = p->tf + 4 * ti;
            const float* e1 = p->tf + 4 * tj;
            float a = e0[3] + (e1[3] - e0[3]) * tf;
            if (a <= 0.0f) continue;
            float ac = a >= 1.0f ? 1.0f : 1.0f - exp2f(len * log2f(1.0f - a));
            float w = (1.0f - A) * ac;
            C0 += w * (e0[0] + (e1[0] - e0[0]) * tf);
            C1 += w * (e0[1] + (e1[1] - e0[1]) * tf);
            C2 += w * (e0[2] + (e1[2] - e0[2]) * tf);
            A += w;
            if (A >= thr) break;
          }
        }
        float wy = 1.0f - (y + 0.5f) / H;
        float bg0 = (float)(p->bg_bottom[0] + (p->bg_top[0] - p->bg_bottom[0]) * wy);
        float bg1 = (float)(p->bg_bottom[1] + (p->bg_top[1] - p->bg_bottom[1]) * wy);
        float bg2 = (float)(p->bg_bottom[2] + (p->bg_top[2] - p->bg_bottom[2]) * wy);
        float r = C0 + (1 - A) * bg0, g = C1 + (1 - A) * bg1, b = C2 + (1 - A) * bg2;
        float al = A + (1 - A) * (float)p->bg_alpha;
        uint8_t* px = out8 + 4 * ((int64_t)y * W + x);
        px[0] = (uint8_t)floorf(fminf(fmaxf(r, 0.f), 1.f) * 255.f + 0.5f);
        px[1] = (uint8_t)floorf(fminf(fmaxf(g, 0.f), 1.f) * 255.f + 0.5f);
        px[2] = (uint8_t)floorf(fminf(fmaxf(b, 0.f), 1.f) * 255.f + 0.5f);
        px[3] = (uint8_t)floorf(fminf(fmaxf(al, 0.f), 1.f) * 255.f + 0.5f);
      }
  }
  free(empty);
  return total;
}

