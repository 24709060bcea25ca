/*
 * oracle_common.h -- helpers shared by the FP64 oracle and the FP32 CPU
 * baseline (camera rays, box test).  TEST INFRASTRUCTURE ONLY.
 */
#ifndef TVK_ORACLE_COMMON_H
#define TVK_ORACLE_COMMON_H
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ utilities */
static inline double vox_at(const void* v, int bits, int64_t idx) {
  return bits == 8 ? (double)((const uint8_t*)v)[idx] : (double)((const uint16_t*)v)[idx];
}
static inline float vox_f(const void* v, int bits, int64_t idx) {
  return bits == 8 ? (float)((const uint8_t*)v)[idx] : (float)((const uint16_t*)v)[idx];
}
static inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
static inline double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

static inline void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void norm3(double a[3]) {
  double l = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  if (l > 0) { a[0] /= l; a[1] /= l; a[2] /= l; }
}

/* 4x4 inverse (row-major), Gauss-Jordan with partial pivoting. */
static inline int inv4(const double m[16], double out[16]) {
  double a[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m[i * 4 + j] : (j - 4 == i ? 1.0 : 0.0);
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
    if (fabs(a[piv][c]) < 1e-300) return -1;
    if (piv != c)
      for (int j = 0; j < 8; ++j) { double t = a[c][j]; a[c][j] = a[piv][j]; a[piv][j] = t; }
    double d = a[c][c];
    for (int j = 0; j < 8; ++j) a[c][j] /= d;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      double f = a[r][c];
      for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[i * 4 + j] = a[i][j + 4];
  return 0;
}

/* Row-vector transform: p' = [p,w] * M. */
static inline void xform(const double m[16], const double p[3], double w, double o[3]) {
  for (int j = 0; j < 3; ++j) o[j] = p[0] * m[0 * 4 + j] + p[1] * m[1 * 4 + j] + p[2] * m[2 * 4 + j] + w * m[3 * 4 + j];
}

/* Ray setup shared by the oracle and the CPU baseline: origin and unit
 * direction in LOD-0 voxel space for pixel (x, y). */
typedef struct cam_ctx {
  double minv[16];
  double f[3], r[3], u[3];
  double tanh_, aspect;
  double vscale[3], voff[3];
} cam_ctx;

static inline int cam_setup(const oracle_render_params* p, cam_ctx* c) {
  if (inv4(p->model, c->minv)) return -1;
  for (int i = 0; i < 3; ++i) c->f[i] = p->view_dir[i];
  norm3(c->f);
  cross3(c->f, p->up_dir, c->r);
  norm3(c->r);
  cross3(c->r, c->f, c->u);
  c->tanh_ = tan(p->fov_deg * 0.5 * 3.14159265358979323846 / 180.0);
  c->aspect = (double)p->width / (double)p->height;
  double n0[3] = {p->n0x, p->n0y, p->n0z};
  double m = 0;
  for (int i = 0; i < 3; ++i) if (n0[i] * p->rescale[i] > m) m = n0[i] * p->rescale[i];
  for (int i = 0; i < 3; ++i) { c->vscale[i] = m / p->rescale[i]; c->voff[i] = 0.5 * n0[i]; }
  return 0;
}

static inline void cam_ray(const oracle_render_params* p, const cam_ctx* c, int x, int y,
                    double o[3], double d[3]) {
  double ndx = 2.0 * (x + 0.5) / p->width - 1.0;
  double ndy = 1.0 - 2.0 * (y + 0.5) / p->height;
  double ow[3], dw[3];
  if (p->ortho) {
    double hh = p->ortho_half_height;
    for (int i = 0; i < 3; ++i) {
      ow[i] = p->view_pos[i] + ndx * hh * c->aspect * c->r[i] + ndy * hh * c->u[i];
      dw[i] = c->f[i];
    }
  } else {
    for (int i = 0; i < 3; ++i) {
      ow[i] = p->view_pos[i];
      dw[i] = c->f[i] + ndx * c->tanh_ * c->aspect * c->r[i] + ndy * c->tanh_ * c->u[i];
    }
  }
  double om[3], dm[3];
  xform(c->minv, ow, 1.0, om);
  xform(c->minv, dw, 0.0, dm);
  for (int i = 0; i < 3; ++i) { o[i] = om[i] * c->vscale[i] + c->voff[i]; d[i] = dm[i] * c->vscale[i]; }
  norm3(d);
}

/* Slab test against [lo, hi]^3 (per axis); axis-parallel rays use an
 * explicit inside test instead of infinities. */
static inline int box_hit(const double o[3], const double d[3], const double lo[3], const double hi[3],
                   double* t0, double* t1) {
  double a = -INFINITY, b = INFINITY;
  for (int i = 0; i < 3; ++i) {
    if (d[i] == 0.0) {
      if (o[i] < lo[i] || o[i] > hi[i]) return 0;
      continue;
    }
    double inv = 1.0 / d[i];
    double ta = (lo[i] - o[i]) * inv, tb = (hi[i] - o[i]) * inv;
    if (ta > tb) { double t = ta; ta = tb; tb = t; }
    if (ta > a) a = ta;
    if (tb < b) b = tb;
  }
  *t0 = a; *t1 = b;
  return b > a;
}


/* The world clip plane (p->clip_plane) in LOD-0 voxel space.  World
 * points are w = m.M (row vectors, M = p->model) with voxel coordinates
 * v = m * vscale + voff, so n.w + d = (M n / vscale) . v + d + M[3].n -
 * sum_i (M n)_i voff_i / vscale_i. */
static inline void clip_plane_voxel(const oracle_render_params* p, const cam_ctx* c, double out[4]) {
  const double* M = p->model;
  const double* n = p->clip_plane;
  double dv = n[3];
  for (int j = 0; j < 3; ++j) dv += M[12 + j] * n[j];
  for (int i = 0; i < 3; ++i) {
    out[i] = (M[i * 4 + 0] * n[0] + M[i * 4 + 1] * n[1] + M[i * 4 + 2] * n[2]) / c->vscale[i];
    dv -= out[i] * c->voff[i];
  }
  out[3] = dv;
}

/* Clip [t0, t1] to the kept side (pl . p <= 0) of a voxel-space plane;
 * returns 0 when a ray parallel to the plane lies wholly on the cut side. */
static inline int clip_interval(const double pl[4], const double o[3], const double d[3], double* t0,
                                double* t1) {
  const double num = pl[0] * o[0] + pl[1] * o[1] + pl[2] * o[2] + pl[3];
  const double den = pl[0] * d[0] + pl[1] * d[1] + pl[2] * d[2];
  if (den == 0.0) return num <= 0.0;
  const double tc = -num / den;
  if (den > 0.0) {
    if (tc < *t1) *t1 = tc;
  } else if (tc > *t0) {
    *t0 = tc;
  }
  return 1;
}

/* TF-empty flag per brick of the sampled level (oracle.c). */
uint8_t* oracle_brick_empty_flags(const oracle_render_params* p, int nb[3]);

#endif
