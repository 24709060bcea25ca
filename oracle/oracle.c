/*
 * oracle.c -- FP64 CPU restatement of the Tuvok raycast + brick/LOD path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); the
 * product path never loads it.  See oracle.h for the reference anchors and
 * for which functions are pinned by what.
 *
 * Render model (shared with the HIP kernels, stated once here):
 *  world   : volume box centred at 0, largest extent (dims*rescale) = 1.
 *  voxel   : LOD-0 voxel i covers [i, i+1); its centre is i + 0.5.  A point
 *            p in model space maps to v = p * (m / s) + N0 / 2 with
 *            m = max_j(N0_j * s_j) (GL texture convention, clamp-to-edge).
 *  rays    : pixel centres, row 0 at the top; perspective through the eye
 *            with vertical FoV, or orthographic with a stated half height.
 *            Rays start at t >= 0 (eye / image plane) and are clipped to the
 *            LOD-0 box; t is voxel-space length.
 *  samples : step D = 2^lod / rate; the chord L is cut into n = ceil(L/D)
 *            segments, the last one of length L-(n-1)D; each segment is
 *            sampled at its midpoint and its opacity is corrected with its
 *            own length: a' = 1 - (1-a)^len (len in LOD-0 voxels).  The
 *            composite is therefore continuous in the entry/exit distances.
 *  level   : sample position in level texel space q = v * N_l / N0.
 *  TF      : x = clamp(value * tf_scale, 0, n-1), linear between entries.
 *  blend   : front to back, C += (1-A) a' c, A += (1-A) a'; stop once
 *            A >= term_alpha.  Pixel = C + (1-A) bg, alpha A + (1-A) bg_a.
 *  bg      : vertical gradient, top colour at the top row.
 */
#include "oracle_common.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static double sample_lin(const oracle_render_params* p, const double q[3]) {
  int n[3] = {p->nx, p->ny, p->nz};
  int i0[3], i1[3];
  double f[3];
  for (int a = 0; a < 3; ++a) {
    double g = q[a] - 0.5;
    double fl = floor(g);
    f[a] = g - fl;
    i0[a] = clampi((int)fl, 0, n[a] - 1);
    i1[a] = clampi((int)fl + 1, 0, n[a] - 1);
  }
  int64_t sx = 1, sy = n[0], sz = (int64_t)n[0] * n[1];
  double c000 = vox_at(p->voxels, p->bits, i0[0] * sx + i0[1] * sy + i0[2] * sz);
  double c100 = vox_at(p->voxels, p->bits, i1[0] * sx + i0[1] * sy + i0[2] * sz);
  double c010 = vox_at(p->voxels, p->bits, i0[0] * sx + i1[1] * sy + i0[2] * sz);
  double c110 = vox_at(p->voxels, p->bits, i1[0] * sx + i1[1] * sy + i0[2] * sz);
  double c001 = vox_at(p->voxels, p->bits, i0[0] * sx + i0[1] * sy + i1[2] * sz);
  double c101 = vox_at(p->voxels, p->bits, i1[0] * sx + i0[1] * sy + i1[2] * sz);
  double c011 = vox_at(p->voxels, p->bits, i0[0] * sx + i1[1] * sy + i1[2] * sz);
  double c111 = vox_at(p->voxels, p->bits, i1[0] * sx + i1[1] * sy + i1[2] * sz);
  double c00 = c000 + (c100 - c000) * f[0];
  double c10 = c010 + (c110 - c010) * f[0];
  double c01 = c001 + (c101 - c001) * f[0];
  double c11 = c011 + (c111 - c011) * f[0];
  double c0 = c00 + (c10 - c00) * f[1];
  double c1 = c01 + (c11 - c01) * f[1];
  return c0 + (c1 - c0) * f[2];
}

static double sample_nn(const oracle_render_params* p, const double q[3], int* near_face) {
  int n[3] = {p->nx, p->ny, p->nz};
  int i[3];
  for (int a = 0; a < 3; ++a) {
    double fl = floor(q[a]);
    double fr = q[a] - fl;
    if (fr < 1e-4 || fr > 1.0 - 1e-4) *near_face = 1;
    i[a] = clampi((int)fl, 0, n[a] - 1);
  }
  return vox_at(p->voxels, p->bits, i[0] + (int64_t)n[0] * (i[1] + (int64_t)n[1] * i[2]));
}

static void tf_lookup(const oracle_render_params* p, double v, double rgba[4]) {
  double x = clampd(v * p->tf_scale, 0.0, (double)(p->n_tf - 1));
  int i = (int)floor(x);
  double f = x - i;
  int j = i + 1 < p->n_tf ? i + 1 : p->n_tf - 1;
  for (int c = 0; c < 4; ++c) rgba[c] = (double)p->tf[4 * i + c] * (1.0 - f) + (double)p->tf[4 * j + c] * f;
}

/* 2D TF lookup, bilinear: x = value * tf_scale (columns), y = |g| / max_grad
 * * (rows - 1), both clamped to the raster (value fastest). */
static void tf2d_lookup(const oracle_render_params* p, double v, double gmag, double rgba[4]) {
  const int W = p->tf2d_w, H = p->tf2d_h;
  double x = clampd(v * p->tf_scale, 0.0, (double)(W - 1));
  double y = p->max_grad > 0 ? clampd(gmag / p->max_grad * (H - 1), 0.0, (double)(H - 1)) : 0.0;
  int i = (int)floor(x), j = (int)floor(y);
  double fx = x - i, fy = y - j;
  int i1 = i + 1 < W ? i + 1 : W - 1, j1 = j + 1 < H ? j + 1 : H - 1;
  for (int c = 0; c < 4; ++c) {
    double a = p->tf2d[4 * ((int64_t)j * W + i) + c], b = p->tf2d[4 * ((int64_t)j * W + i1) + c];
    double e = p->tf2d[4 * ((int64_t)j1 * W + i) + c], f = p->tf2d[4 * ((int64_t)j1 * W + i1) + c];
    double r0 = a + (b - a) * fx, r1 = e + (f - e) * fx;
    rgba[c] = r0 + (r1 - r0) * fy;
  }
}

/* Gradient at level texel position q: central differences of the trilinear
 * interpolant one level voxel apart, in value/(2^bits-1) per LOD-0 voxel. */
static void gradient(const oracle_render_params* p, const double q[3], double g[3]) {
  const double maxval = p->bits == 8 ? 255.0 : 65535.0;
  const double lsc[3] = {(double)p->nx / p->n0x, (double)p->ny / p->n0y, (double)p->nz / p->n0z};
  for (int a = 0; a < 3; ++a) {
    double qp[3] = {q[0], q[1], q[2]}, qm[3] = {q[0], q[1], q[2]};
    qp[a] += 1.0;
    qm[a] -= 1.0;
    g[a] = (sample_lin(p, qp) - sample_lin(p, qm)) / (2.0 * maxval) * lsc[a];
  }
}

/* Two-sided Blinn-Phong on colour c: c*(ka + kd|N.L|) + ks|N.H|^shininess
 * with N = g / sqrt(|g|^2 + eps^2) (eps = 1e-4: flat regions fade to the
 * ambient term instead of taking a noisy direction). */
static void shade(const oracle_render_params* p, const double g[3], const double V[3],
                  const double L[3], double rgb[3]) {
  const double eps = 1e-4;
  double s = 1.0 / sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + eps * eps);
  double N[3] = {g[0] * s, g[1] * s, g[2] * s};
  double H[3] = {L[0] + V[0], L[1] + V[1], L[2] + V[2]};
  double hl = sqrt(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
  if (hl > 1e-12) for (int i = 0; i < 3; ++i) H[i] /= hl;
  double nl = fabs(N[0] * L[0] + N[1] * L[1] + N[2] * L[2]);
  double nh = fabs(N[0] * H[0] + N[1] * H[1] + N[2] * H[2]);
  double sp = pow(nh, p->shininess);
  for (int i = 0; i < 3; ++i) rgb[i] = rgb[i] * (p->ambient[i] + p->diffuse[i] * nl) + p->specular[i] * sp;
}

/* Isosurface hit at sample k (value v >= iso, ray parameter t): the hit
 * refined linearly against sample k-1 (re-fetched, so the result does not
 * depend on skipping), `color` shaded there into rgb.  Returns |N.V| of the
 * unit gradient N = g / sqrt(|g|^2 + 1e-8) (ClearView's context term). */
static double iso_hit(const oracle_render_params* p, const double o[3], const double d[3],
                      const double lsc[3], double t0, double step, int64_t k, double t, double v,
                      double iso, const double color[3], const double V[3], const double Lv[3],
                      double rgb[3], uint8_t* flags) {
  double th = t;
  if (k > 0) {
    int nf = 0;
    double tp = t0 + ((double)(k - 1) + 0.5) * step, qp[3];
    for (int i = 0; i < 3; ++i) qp[i] = (o[i] + tp * d[i]) * lsc[i];
    double vp = p->interp ? sample_lin(p, qp) : sample_nn(p, qp, &nf);
    if (fabs(vp - iso) < 1e-2) *flags |= 4;
    if (vp < iso && v > vp) th = tp + (iso - vp) / (v - vp) * (t - tp);
  }
  double qh[3], g[3];
  for (int i = 0; i < 3; ++i) qh[i] = (o[i] + th * d[i]) * lsc[i];
  gradient(p, qh, g);
  for (int i = 0; i < 3; ++i) rgb[i] = color[i];
  shade(p, g, V, Lv, rgb);
  const double s = 1.0 / sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + 1e-8);
  return fabs((g[0] * V[0] + g[1] * V[1] + g[2] * V[2]) * s);
}

/* TF-empty flag per brick of the sampled level: brick empty iff every TF
 * entry (1D) or column (2D, any row) its [min, max] range can reach has
 * alpha == 0 (exact skip); for isosurfaces, iff its max is below the iso
 * value. */
uint8_t* oracle_brick_empty_flags(const oracle_render_params* p, int nb[3]) {
  int B = p->brick_size, a = p->apron;
  nb[0] = (p->nx + B - 1) / B; nb[1] = (p->ny + B - 1) / B; nb[2] = (p->nz + B - 1) / B;
  int64_t cnt = (int64_t)nb[0] * nb[1] * nb[2];
  uint8_t* e = (uint8_t*)malloc(cnt);
  const int ncol = p->mode == 1 ? p->tf2d_w : p->n_tf;
  uint8_t* col = (uint8_t*)calloc(ncol > 0 ? ncol : 1, 1);
  for (int i = 0; i < ncol; ++i) {
    if (p->mode == 1) {
      for (int j = 0; j < p->tf2d_h; ++j)
        if (p->tf2d[4 * ((int64_t)j * p->tf2d_w + i) + 3] > 0.0f) { col[i] = 1; break; }
    } else if (p->mode == 0) {
      col[i] = p->tf[4 * i + 3] > 0.0f;
    }
  }
  for (int bz = 0; bz < nb[2]; ++bz)
    for (int by = 0; by < nb[1]; ++by)
      for (int bx = 0; bx < nb[0]; ++bx) {
        double mn, mx;
        oracle_brick_minmax(p->voxels, p->bits, p->nx, p->ny, p->nz, B, a, bx, by, bz, &mn, &mx);
        int empty = 1;
        if (p->mode == 2) {
          /* with ClearView both surfaces are searched for */
          const double lowest =
              p->cv_enabled && p->cv_iso_value < p->iso_value ? p->cv_iso_value : p->iso_value;
          empty = mx < lowest;
        } else {
          double x0 = clampd(mn * p->tf_scale, 0, ncol - 1), x1 = clampd(mx * p->tf_scale, 0, ncol - 1);
          int i0 = (int)floor(x0), i1 = (int)floor(x1) + 1;
          if (i1 > ncol - 1) i1 = ncol - 1;
          for (int i = i0; i <= i1; ++i)
            if (col[i]) { empty = 0; break; }
        }
        e[(bz * nb[1] + by) * nb[0] + bx] = (uint8_t)empty;
      }
  free(col);
  return e;
}

/* IEEE 754 binary16, round to nearest even: 11 significant bits for
 * |x| >= 2^-14, multiples of 2^-24 below (subnormals); the composite stays
 * far below the binary16 overflow. */
static double quantize_half(double x) {
  const double ax = fabs(x);
  if (ax == 0.0) return x;
  if (ax < ldexp(1.0, -14)) return nearbyint(x * ldexp(1.0, 24)) * ldexp(1.0, -24);
  int e;
  const double m = frexp(x, &e); /* x = m 2^e, 0.5 <= |m| < 1 */
  return ldexp(nearbyint(ldexp(m, 11)), e - 11);
}

double oracle_blend_quantize(double x, int32_t bits) {
  if (bits == 8) return floor(clampd(x, 0.0, 1.0) * 255.0 + 0.5) / 255.0;
  if (bits == 16) return quantize_half(x);
  return x;
}

static void blend_round(int bits, double C[3], double* A) {
  for (int i = 0; i < 3; ++i) C[i] = oracle_blend_quantize(C[i], bits);
  *A = oracle_blend_quantize(*A, bits);
}

int oracle_render_1d(const oracle_render_params* p, double* out, uint32_t* out_samples,
                     uint8_t* out_flags) {
  cam_ctx c;
  if (cam_setup(p, &c)) return -1;
  double clip[4] = {0, 0, 0, 0};
  if (p->clip_enabled) clip_plane_voxel(p, &c, clip);
  const int blends = p->blend_bits > 0 && p->blend_bits < 32 && (p->mode == 0 || p->mode == 1);
  int bnb[3] = {1, 1, 1};
  if (blends) {
    const int ln[3] = {p->nx, p->ny, p->nz};
    for (int i = 0; i < 3; ++i) bnb[i] = (ln[i] + p->blend_brick - 1) / p->blend_brick;
  }
  int nb[3] = {0, 0, 0};
  /* MIP (mode 3) visits every sample: a TF-transparent region may hold the
   * maximum, so the TF-based empty-space flags do not apply */
  uint8_t* empty = p->brick_size > 0 && p->mode != 3 ? oracle_brick_empty_flags(p, nb) : NULL;
  const double lo[3] = {0, 0, 0};
  const double hi[3] = {p->n0x, p->n0y, p->n0z};
  const double lsc[3] = {(double)p->nx / p->n0x, (double)p->ny / p->n0y, (double)p->nz / p->n0z};
  const double step = ldexp(1.0, p->lod) / p->rate;
  const int grad = p->mode == 1 || p->mode == 2 || (p->lighting && p->mode != 3);
  /* light direction in voxel space (a direction: linear part only) */
  double Lv[3] = {0, 0, 1};
  if (grad) {
    double lm[3];
    xform(c.minv, p->light_dir, 0.0, lm);
    for (int i = 0; i < 3; ++i) Lv[i] = lm[i] * c.vscale[i];
    norm3(Lv);
  }
  const double iso_eps = 1e-2;
  for (int y = 0; y < p->height; ++y) {
    double wy = 1.0 - (y + 0.5) / p->height;
    double bg[3];
    for (int i = 0; i < 3; ++i) bg[i] = p->bg_bottom[i] + (p->bg_top[i] - p->bg_bottom[i]) * wy;
    for (int x = 0; x < p->width; ++x) {
      double o[3], d[3];
      cam_ray(p, &c, x, y, o, d);
      const double V[3] = {-d[0], -d[1], -d[2]};
      double C[3] = {0, 0, 0}, A = 0;
      uint32_t ns = 0;
      uint8_t flags = 0;
      double t0, t1;
      double vmax = 0.0;
      int any = 0;
      /* ClearView: focus weight of this pixel, context colour and opacity */
      int cv_stage = 0;
      double cctx[3] = {0, 0, 0}, ac = 1.0, wf = 0.0;
      if (p->mode == 2 && p->cv_enabled) {
        const double r = hypot(x - p->cv_focus[0], y - p->cv_focus[1]) / p->height;
        wf = 1.0 - p->cv_size * r;
        wf = wf < 0 ? 0 : (wf > 1 ? 1 : wf);
      }
      if (box_hit(o, d, lo, hi, &t0, &t1)) {
        if (t0 < 0) t0 = 0;
        if (p->clip_enabled && !clip_interval(clip, o, d, &t0, &t1)) t1 = t0;
        double L = t1 - t0;
        if (L > 0) {
          int64_t n = (int64_t)ceil(L / step);
          int64_t blend_b = -1;
          for (int64_t k = 0; k < n; ++k) {
            double len = (k == n - 1) ? L - (double)(n - 1) * step : step;
            double t = t0 + (double)k * step + 0.5 * len;
            double q[3];
            for (int i = 0; i < 3; ++i) q[i] = (o[i] + t * d[i]) * lsc[i];
            if (blends) {
              /* the composite is blended into the target brick by brick */
              int64_t b[3];
              for (int i = 0; i < 3; ++i) b[i] = clampi((int)floor(q[i] / p->blend_brick), 0, bnb[i] - 1);
              const int64_t bi = (b[2] * bnb[1] + b[1]) * bnb[0] + b[0];
              if (bi != blend_b) {
                blend_round(p->blend_bits, C, &A);
                blend_b = bi;
              }
            }
            if (empty) {
              int b[3];
              for (int i = 0; i < 3; ++i) b[i] = clampi((int)floor(q[i] / p->brick_size), 0, nb[i] - 1);
              if (empty[(b[2] * nb[1] + b[1]) * nb[0] + b[0]]) continue;
            }
            int nf = 0;
            double v = p->interp ? sample_lin(p, q) : sample_nn(p, q, &nf);
            if (nf) flags |= 2;
            ++ns;
            if (p->mode == 3) {
              /* maximum intensity projection: the largest sample value */
              if (!any || v > vmax) vmax = v;
              any = 1;
              continue;
            }
            if (p->mode == 2) {
              /* first sample at or above the iso value, refined against
               * the previous sample; with ClearView and the pixel in the
               * focus, then the first sample from there on at or above
               * the focus iso value */
              if (cv_stage == 0) {
                if (fabs(v - p->iso_value) < iso_eps) flags |= 4;
                if (v < p->iso_value) continue;
                double rgb[3];
                const double nv = iso_hit(p, o, d, lsc, t0, step, k, t, v, p->iso_value,
                                          p->iso_color, V, Lv, rgb, &flags);
                if (wf <= 0.0) {
                  for (int i = 0; i < 3; ++i) C[i] = rgb[i];
                  A = 1.0;
                  break;
                }
                for (int i = 0; i < 3; ++i) cctx[i] = rgb[i];
                const double x = 1.0 - nv;
                double sil = p->cv_context_scale * pow(x, p->cv_border_scale);
                sil = sil < 0 ? 0 : (sil > 1 ? 1 : sil);
                ac = (1.0 - wf) + wf * sil;
                cv_stage = 1;
              }
              if (fabs(v - p->cv_iso_value) < iso_eps) flags |= 4;
              if (v < p->cv_iso_value) continue;
              double rgb[3];
              iso_hit(p, o, d, lsc, t0, step, k, t, v, p->cv_iso_value, p->cv_color, V, Lv, rgb,
                      &flags);
              for (int i = 0; i < 3; ++i) C[i] = ac * cctx[i] + (1.0 - ac) * rgb[i];
              A = 1.0;
              cv_stage = 2;
              break;
            }
            double rgba[4], g[3] = {0, 0, 0};
            if (grad) gradient(p, q, g);
            if (p->mode == 1) tf2d_lookup(p, v, sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]), rgba);
            else tf_lookup(p, v, rgba);
            double a = rgba[3];
            if (a <= 0.0) continue;
            if (p->lighting) shade(p, g, V, Lv, rgba);
            double ac = a >= 1.0 ? 1.0 : 1.0 - pow(1.0 - a, len);
            double w = (1.0 - A) * ac;
            for (int i = 0; i < 3; ++i) C[i] += w * rgba[i];
            A += w;
            if (fabs(A - p->term_alpha) < 1e-5) flags |= 1;
            if (A >= p->term_alpha) break;
          }
          if (blends) blend_round(p->blend_bits, C, &A);
        }
      }
      if (cv_stage == 1) {  /* context hit, no focus surface behind it */
        for (int i = 0; i < 3; ++i) C[i] = ac * cctx[i];
        A = ac;
      }
      if (p->mode == 3 && any) {
        /* the maximum through the 1D TF, over the background */
        double rgba[4];
        tf_lookup(p, vmax, rgba);
        for (int i = 0; i < 3; ++i) C[i] = rgba[i] * rgba[3];
        A = rgba[3];
      }
      double* px = out + 4 * ((int64_t)y * p->width + x);
      for (int i = 0; i < 3; ++i) px[i] = C[i] + (1.0 - A) * bg[i];
      px[3] = A + (1.0 - A) * p->bg_alpha;
      if (out_samples) out_samples[(int64_t)y * p->width + x] = ns;
      if (out_flags) out_flags[(int64_t)y * p->width + x] = flags;
    }
  }
  free(empty);
  return 0;
}

/* Per-region LOD model (1D TF, no lighting): the ray walks the grid of
 * LOD-0 regions of edge `region_edge` (the LOD-0 brick size); region r is
 * sampled from level l = region_lod[r] with step 2^l / rate, its segments
 * restarting at the region entry (last one partial, midpoints, opacity
 * correction by segment length in LOD-0 voxels).  A region whose level-l
 * brick (the brick holding the region centre, stored range over payload +
 * apron) is TF-transparent contributes nothing and is not counted.  The
 * region boundary of a ray is found at the midpoint of a short probe step
 * and crossed with at least 1e-3 voxel of progress, as the kernel does. */
int oracle_render_regions(const oracle_render_params* p, int32_t n_levels,
                          const void* const* levels, const int32_t* level_dims,
                          int32_t region_edge, const uint8_t* region_lod, double* out,
                          uint32_t* out_samples, uint8_t* out_flags) {
  cam_ctx c;
  if (cam_setup(p, &c)) return -1;
  if (n_levels < 1 || n_levels > 16 || region_edge < 1) return -1;
  double clip[4] = {0, 0, 0, 0};
  if (p->clip_enabled) clip_plane_voxel(p, &c, clip);
  oracle_render_params lp[16];
  uint8_t* empty[16];
  int nb[16][3];
  for (int l = 0; l < n_levels; ++l) {
    lp[l] = *p;
    lp[l].voxels = levels[l];
    lp[l].nx = level_dims[3 * l];
    lp[l].ny = level_dims[3 * l + 1];
    lp[l].nz = level_dims[3 * l + 2];
    lp[l].brick_size = region_edge;
    lp[l].mode = 0;
    empty[l] = oracle_brick_empty_flags(&lp[l], nb[l]);
  }
  const double E = region_edge;
  const int g0[3] = {(p->n0x + region_edge - 1) / region_edge, (p->n0y + region_edge - 1) / region_edge,
                     (p->n0z + region_edge - 1) / region_edge};
  const double lo[3] = {0, 0, 0};
  const double hi[3] = {p->n0x, p->n0y, p->n0z};
  const double base_step = 1.0 / p->rate;
  for (int y = 0; y < p->height; ++y) {
    double wy = 1.0 - (y + 0.5) / p->height;
    double bg[3];
    for (int i = 0; i < 3; ++i) bg[i] = p->bg_bottom[i] + (p->bg_top[i] - p->bg_bottom[i]) * wy;
    for (int x = 0; x < p->width; ++x) {
      double o[3], d[3];
      cam_ray(p, &c, x, y, o, d);
      double C[3] = {0, 0, 0}, A = 0;
      uint32_t ns = 0;
      uint8_t flags = 0;
      double t0, t1;
      if (box_hit(o, d, lo, hi, &t0, &t1)) {
        if (t0 < 0) t0 = 0;
        if (p->clip_enabled && !clip_interval(clip, o, d, &t0, &t1)) t1 = t0;
        double t = t0;
        int live = t1 > t0;
        while (t < t1 && live) {
          const double tm = t + 0.5 * fmin(base_step, t1 - t);
          int cc[3];
          for (int i = 0; i < 3; ++i) cc[i] = clampi((int)floor((o[i] + tm * d[i]) / E), 0, g0[i] - 1);
          double tb = t1;
          for (int i = 0; i < 3; ++i)
            if (d[i] != 0.0) {
              const double face = (d[i] > 0 ? cc[i] + 1 : cc[i]) * E;
              tb = fmin(tb, (face - o[i]) / d[i]);
            }
          tb = fmin(fmax(tb, t + 1e-3), t1);
          const int l = region_lod[((int64_t)cc[2] * g0[1] + cc[1]) * g0[0] + cc[0]] & 15;
          if (l >= n_levels) return -1;
          const oracle_render_params* q = &lp[l];
          const double lsc[3] = {(double)q->nx / p->n0x, (double)q->ny / p->n0y, (double)q->nz / p->n0z};
          int b[3];
          for (int i = 0; i < 3; ++i) {
            const double qc = (cc[i] + 0.5) * E * lsc[i];
            b[i] = clampi((int)floor(qc / E), 0, nb[l][i] - 1);
          }
          if (!empty[l][((int64_t)b[2] * nb[l][1] + b[1]) * nb[l][0] + b[0]]) {
            const double step = ldexp(base_step, l);
            const double span = tb - t;
            int64_t m = (int64_t)ceil(span / step);
            if (m < 1) m = 1;
            const double last = fmax(span - (double)(m - 1) * step, 0.0);
            for (int64_t k = 0; k < m; ++k) {
              const double len = k == m - 1 ? last : step;
              const double ts = t + (double)k * step + 0.5 * len;
              double qq[3];
              for (int i = 0; i < 3; ++i) qq[i] = (o[i] + ts * d[i]) * lsc[i];
              int nf = 0;
              const double v = p->interp ? sample_lin(q, qq) : sample_nn(q, qq, &nf);
              if (nf) flags |= 2;
              ++ns;
              double rgba[4];
              tf_lookup(q, v, rgba);
              const double a = rgba[3];
              if (a <= 0.0) continue;
              const double ac = a >= 1.0 ? 1.0 : 1.0 - pow(1.0 - a, len);
              const double w = (1.0 - A) * ac;
              for (int i = 0; i < 3; ++i) C[i] += w * rgba[i];
              A += w;
              if (fabs(A - p->term_alpha) < 1e-5) flags |= 1;
              if (A >= p->term_alpha) {
                live = 0;
                break;
              }
            }
          }
          t = tb;
        }
      }
      double* px = out + 4 * ((int64_t)y * p->width + x);
      for (int i = 0; i < 3; ++i) px[i] = C[i] + (1.0 - A) * bg[i];
      px[3] = A + (1.0 - A) * p->bg_alpha;
      if (out_samples) out_samples[(int64_t)y * p->width + x] = ns;
      if (out_flags) out_flags[(int64_t)y * p->width + x] = flags;
    }
  }
  for (int l = 0; l < n_levels; ++l) free(empty[l]);
  return 0;
}

uint64_t oracle_nominal_samples(const oracle_render_params* p) {
  cam_ctx c;
  if (cam_setup(p, &c)) return 0;
  const double lo[3] = {0, 0, 0};
  const double hi[3] = {p->n0x, p->n0y, p->n0z};
  uint64_t total = 0;
  for (int y = 0; y < p->height; ++y)
    for (int x = 0; x < p->width; ++x) {
      double o[3], d[3], t0, t1;
      cam_ray(p, &c, x, y, o, d);
      if (!box_hit(o, d, lo, hi, &t0, &t1)) continue;
      if (t0 < 0) t0 = 0;
      if (t1 > t0) total += (uint64_t)floor((t1 - t0) * p->rate);
    }
  return total;
}

/* ------------------------------------------------------- bricking / LOD */
void oracle_downsample(const void* src, int32_t bits, int32_t nx, int32_t ny, int32_t nz, void* dst) {
  int mx = (nx + 1) / 2, my = (ny + 1) / 2, mz = (nz + 1) / 2;
  for (int z = 0; z < mz; ++z)
    for (int y = 0; y < my; ++y)
      for (int x = 0; x < mx; ++x) {
        uint64_t sum = 0, cnt = 0;
        for (int dz = 0; dz < 2; ++dz)
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
              int sx = 2 * x + dx, sy = 2 * y + dy, sz = 2 * z + dz;
              if (sx >= nx || sy >= ny || sz >= nz) continue;
              int64_t i = sx + (int64_t)nx * (sy + (int64_t)ny * sz);
              sum += bits == 8 ? ((const uint8_t*)src)[i] : ((const uint16_t*)src)[i];
              ++cnt;
            }
        int64_t o = x + (int64_t)mx * (y + (int64_t)my * z);
        if (bits == 8) ((uint8_t*)dst)[o] = (uint8_t)(sum / cnt);
        else ((uint16_t*)dst)[o] = (uint16_t)(sum / cnt);
      }
}

/* Median variant of the 2x2x2 reduction (the TOC builder's UseMedian
 * option, ImageVis3D_Settings.cpp:217,521; Tuvok's median rule is in the
 * absent ExtendedOctree source): the lower median of the existing children,
 * element (n-1)/2 of the n sorted values. */
void oracle_downsample_median(const void* src, int32_t bits, int32_t nx, int32_t ny, int32_t nz,
                              void* dst) {
  int mx = (nx + 1) / 2, my = (ny + 1) / 2, mz = (nz + 1) / 2;
  for (int z = 0; z < mz; ++z)
    for (int y = 0; y < my; ++y)
      for (int x = 0; x < mx; ++x) {
        uint32_t v[8];
        int n = 0;
        for (int dz = 0; dz < 2; ++dz)
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
              int sx = 2 * x + dx, sy = 2 * y + dy, sz = 2 * z + dz;
              if (sx >= nx || sy >= ny || sz >= nz) continue;
              int64_t i = sx + (int64_t)nx * (sy + (int64_t)ny * sz);
              v[n++] = bits == 8 ? ((const uint8_t*)src)[i] : ((const uint16_t*)src)[i];
            }
        /* selection of the (n-1)/2-th smallest by repeated minimum */
        for (int k = 0; k <= (n - 1) / 2; ++k) {
          int m = k;
          for (int j = k + 1; j < n; ++j)
            if (v[j] < v[m]) m = j;
          uint32_t t = v[k];
          v[k] = v[m];
          v[m] = t;
        }
        int64_t o = x + (int64_t)mx * (y + (int64_t)my * z);
        if (bits == 8) ((uint8_t*)dst)[o] = (uint8_t)v[(n - 1) / 2];
        else ((uint16_t*)dst)[o] = (uint16_t)v[(n - 1) / 2];
      }
}

void oracle_extract_brick(const void* src, int32_t bits, int32_t nx, int32_t ny, int32_t nz,
                          int32_t B, int32_t a, int32_t bx, int32_t by, int32_t bz, void* slot) {
  int S = B + 2 * a;
  for (int z = 0; z < S; ++z)
    for (int y = 0; y < S; ++y)
      for (int x = 0; x < S; ++x) {
        int gx = clampi(bx * B - a + x, 0, nx - 1);
        int gy = clampi(by * B - a + y, 0, ny - 1);
        int gz = clampi(bz * B - a + z, 0, nz - 1);
        int64_t i = gx + (int64_t)nx * (gy + (int64_t)ny * gz);
        int64_t o = x + (int64_t)S * (y + (int64_t)S * z);
        if (bits == 8) ((uint8_t*)slot)[o] = ((const uint8_t*)src)[i];
        else ((uint16_t*)slot)[o] = ((const uint16_t*)src)[i];
      }
}

void oracle_brick_minmax(const void* src, int32_t bits, int32_t nx, int32_t ny, int32_t nz,
                         int32_t B, int32_t a, int32_t bx, int32_t by, int32_t bz, double* mn,
                         double* mx) {
  /* clamp-to-edge replication adds no new values, so the slot's range is
   * the range of the in-volume part of [b*B - a, (b+1)*B + a). */
  int x0 = clampi(bx * B - a, 0, nx - 1), x1 = clampi((bx + 1) * B + a - 1, 0, nx - 1);
  int y0 = clampi(by * B - a, 0, ny - 1), y1 = clampi((by + 1) * B + a - 1, 0, ny - 1);
  int z0 = clampi(bz * B - a, 0, nz - 1), z1 = clampi((bz + 1) * B + a - 1, 0, nz - 1);
  double lo = INFINITY, hi = -INFINITY;
  for (int z = z0; z <= z1; ++z)
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) {
        double v = vox_at(src, bits, x + (int64_t)nx * (y + (int64_t)ny * z));
        if (v < lo) lo = v;
        if (v > hi) hi = v;
      }
  *mn = lo;
  *mx = hi;
}

void oracle_histogram_1d(const void* src, int32_t bits, uint64_t n, uint64_t* hist) {
  uint64_t nb = (uint64_t)1 << bits;
  memset(hist, 0, nb * sizeof(uint64_t));
  for (uint64_t i = 0; i < n; ++i) hist[bits == 8 ? ((const uint8_t*)src)[i] : ((const uint16_t*)src)[i]]++;
}

static inline float vnorm(const void* src, int bits, int64_t i) {
  volatile float s = bits == 8 ? 255.0f : 65535.0f;
  return vox_f(src, bits, i) / s;
}

/* One-sided at the borders, central inside (numpy.gradient, unit spacing). */
static inline float grad1(const void* src, int bits, int64_t i, int c, int n, int64_t stride) {
  if (n < 2) return 0.0f;
  if (c == 0) return vnorm(src, bits, i + stride) - vnorm(src, bits, i);
  if (c == n - 1) return vnorm(src, bits, i) - vnorm(src, bits, i - stride);
  return (vnorm(src, bits, i + stride) - vnorm(src, bits, i - stride)) * 0.5f;
}

static inline float grad_mag(const void* src, int bits, int nx, int ny, int nz, int x, int y, int z) {
  int64_t i = x + (int64_t)nx * (y + (int64_t)ny * z);
  volatile float gx = grad1(src, bits, i, x, nx, 1);
  volatile float gy = grad1(src, bits, i, y, ny, nx);
  volatile float gz = grad1(src, bits, i, z, nz, (int64_t)nx * ny);
  volatile float sx = gx * gx, sy = gy * gy, sz = gz * gz;
  volatile float s = sx + sy;
  s = s + sz;
  return sqrtf(s);
}

float oracle_histogram_2d(const void* src, int32_t bits, int32_t nx, int32_t ny, int32_t nz,
                          uint32_t nv, uint32_t ng, uint64_t* hist) {
  float gmax = 0.0f;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        float g = grad_mag(src, bits, nx, ny, nz, x, y, z);
        if (g > gmax) gmax = g;
      }
  memset(hist, 0, (size_t)nv * ng * sizeof(uint64_t));
  int shift = 0;
  while (((uint64_t)nv << shift) < ((uint64_t)1 << bits)) ++shift;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        int64_t i = x + (int64_t)nx * (y + (int64_t)ny * z);
        uint32_t v = bits == 8 ? ((const uint8_t*)src)[i] : ((const uint16_t*)src)[i];
        uint32_t bx = v >> shift;
        float g = grad_mag(src, bits, nx, ny, nz, x, y, z);
        volatile float r = gmax > 0 ? g / gmax : 0.0f;
        volatile float rs = r * (float)(ng - 1);
        uint32_t by = (uint32_t)floorf(rs);
        if (by > ng - 1) by = ng - 1;
        hist[(uint64_t)bx * ng + by]++;
      }
  return gmax;
}
