// lod_select.cpp -- learned-LOD frame planning on the host: per-region
// footprint / distance / predictor inputs (tvk_lod_features) and the octree
// cut of the brick pyramid under error and brick budgets
// (tvk_lod_select_octree).  This is the per-region form of Tuvok's
// per-frame LOD choice between the setLODLimits bounds
// (ImageVis3D/UI/RenderWindow.cpp:447-451).
//
// Conventions: regions are the LOD-0 bricks, x fastest over the grid
// ceil(dims / edge); a level-l brick (bz, by, bx) covers the regions whose
// index shifted right by l equals it (power-of-two pyramid).  World space
// is the volume box centred at 0 with largest extent 1, the space of the
// renderer's camera.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <queue>
#include <vector>

#include "../../include/tvk_hip.h"
#include "errors.hpp"

namespace {

constexpr int kInputs = 32;

struct Vec3 {
  double x, y, z;
};
inline Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline Vec3 scale(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double norm(Vec3 a) { return sqrt(dot(a, a)); }

// A level-l brick of the octree cut, ordered for the split queue: larger
// gain first, ties to the lower (level, z, y, x).
struct Node {
  double gain;
  int l, z, y, x;
};
struct NodeAfter {
  bool operator()(const Node& a, const Node& b) const {
    if (a.gain != b.gain) return a.gain < b.gain;
    if (a.l != b.l) return a.l > b.l;
    if (a.z != b.z) return a.z > b.z;
    if (a.y != b.y) return a.y > b.y;
    return a.x > b.x;
  }
};

}  // namespace

extern "C" {

int tvk_lod_features(const tvk_view* view, const uint64_t dims[3], uint32_t edge, uint32_t n_lod,
                     const double* hist16, const double* moments8, const double* detail, float* x,
                     double* area_out, double* dist_out) {
  if (!view || !dims || !edge || n_lod < 2 || n_lod > 16 || !hist16 || !moments8 || !x ||
      !area_out || view->ortho || !view->width || !view->height || !dims[0] || !dims[1] ||
      !dims[2]) {
    tvk::set_last_error("tvk_lod_features: bad arguments");
    return TVK_ERR_ARG;
  }
  const int64_t g[3] = {(int64_t)((dims[0] + edge - 1) / edge), (int64_t)((dims[1] + edge - 1) / edge),
                        (int64_t)((dims[2] + edge - 1) / edge)};
  const int64_t R = g[0] * g[1] * g[2];
  const double D[3] = {(double)dims[0], (double)dims[1], (double)dims[2]};
  const double m = std::max(D[0], std::max(D[1], D[2]));
  const Vec3 eye{view->view_pos[0], view->view_pos[1], view->view_pos[2]};
  Vec3 f{view->view_dir[0], view->view_dir[1], view->view_dir[2]};
  const Vec3 up{view->up_dir[0], view->up_dir[1], view->up_dir[2]};
  if (norm(f) == 0.0) {
    tvk::set_last_error("tvk_lod_features: zero view direction");
    return TVK_ERR_ARG;
  }
  f = scale(f, 1.0 / norm(f));
  Vec3 r = cross(f, up);
  if (norm(r) == 0.0) {
    tvk::set_last_error("tvk_lod_features: up direction parallel to the view direction");
    return TVK_ERR_ARG;
  }
  r = scale(r, 1.0 / norm(r));
  const Vec3 u = cross(r, f);
  const double W = view->width, H = view->height;
  const double th = tan(view->fov_deg * (M_PI / 180.0) / 2.0), asp = W / H;
  for (int64_t i = 0; i < R; ++i) {
    const int64_t c[3] = {i % g[0], (i / g[0]) % g[1], i / (g[0] * g[1])};
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = (double)(c[a] * (int64_t)edge);
      hi[a] = std::min(lo[a] + (double)edge, D[a]);
    }
    bool ok = true;
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int k = 0; k < 8; ++k) {
      const double p[3] = {(k & 1) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1], (k & 4) ? hi[2] : lo[2]};
      const Vec3 w = sub(Vec3{(p[0] - D[0] / 2) / m, (p[1] - D[1] / 2) / m, (p[2] - D[2] / 2) / m}, eye);
      const double z = dot(w, f);
      ok = ok && z > 1e-3;
      const double zs = std::max(z, 1e-3);
      const double sx = dot(w, r) / (zs * th * asp), sy = dot(w, u) / (zs * th);
      x0 = std::min(x0, sx);
      x1 = std::max(x1, sx);
      y0 = std::min(y0, sy);
      y1 = std::max(y1, sy);
    }
    auto clip1 = [](double v) { return std::min(std::max(v, -1.0), 1.0); };
    double area = (clip1(x1) - clip1(x0)) * 0.5 * W * (clip1(y1) - clip1(y0)) * 0.5 * H;
    area = ok ? std::max(area, 0.0) : 0.0;
    const Vec3 centre{((lo[0] + hi[0]) / 2 - D[0] / 2) / m, ((lo[1] + hi[1]) / 2 - D[1] / 2) / m,
                      ((lo[2] + hi[2]) / 2 - D[2] / 2) / m};
    const double dist = norm(sub(centre, eye));
    area_out[i] = area;
    if (dist_out) dist_out[i] = dist;
    const float foot = (float)(log2(1.0 + area) / 24.0);
    for (uint32_t l = 1; l < n_lod; ++l) {
      float* row = x + ((int64_t)(l - 1) * R + i) * kInputs;
      row[0] = (float)(l / 8.0);
      row[1] = foot;
      for (int b = 0; b < 16; ++b) row[2 + b] = (float)hist16[i * 16 + b];
      for (int b = 0; b < 8; ++b) row[18 + b] = (float)moments8[i * 8 + b];
      row[26] = (float)dist;
      if (detail) {
        const double d = detail[i * n_lod + l];
        row[27] = (float)d;
        row[28] = (float)(d * moments8[i * 8 + 5]);
      } else {
        row[27] = row[28] = 0.f;
      }
      for (int k = 29; k < kInputs; ++k) row[k] = 0.f;
    }
  }
  return TVK_OK;
}

int tvk_lod_select_octree(const float* pred, const double* area, const uint32_t grid[3],
                          uint32_t n_levels, double budget_sse, uint64_t brick_bytes,
                          uint64_t mem_budget, uint8_t* levels, double* sse_out,
                          uint64_t* bytes_out) {
  if (!pred || !area || !grid || !levels || n_levels < 2 || n_levels > 16 || !grid[0] || !grid[1] ||
      !grid[2]) {
    tvk::set_last_error("tvk_lod_select_octree: bad arguments");
    return TVK_ERR_ARG;
  }
  try {
    const int gx = (int)grid[0], gy = (int)grid[1], gz = (int)grid[2];
    const int64_t R = (int64_t)gx * gy * gz;
    const int L = (int)n_levels;
    const uint64_t bb = std::max<uint64_t>(brick_bytes, 1);
    // E[l][bz][by][bx]: predicted SSE of level-l brick = sum over its
    // regions of exp(pred) x footprint (level 0 costs nothing)
    std::vector<int> ex(L), ey(L), ez(L);
    std::vector<std::vector<double>> E(L);
    for (int l = 0; l < L; ++l) {
      const int s = 1 << l;
      ex[l] = (gx + s - 1) / s;
      ey[l] = (gy + s - 1) / s;
      ez[l] = (gz + s - 1) / s;
      E[l].assign((size_t)ex[l] * ey[l] * ez[l], 0.0);
      if (l == 0) continue;
      const float* p = pred + (int64_t)(l - 1) * R;
      for (int z = 0; z < gz; ++z)
        for (int y = 0; y < gy; ++y)
          for (int x = 0; x < gx; ++x) {
            const int64_t i = ((int64_t)z * gy + y) * gx + x;
            E[l][((size_t)(z >> l) * ey[l] + (y >> l)) * ex[l] + (x >> l)] += exp((double)p[i]) * area[i];
          }
    }
    auto at = [&](int l, int z, int y, int x) -> double& {
      return E[l][((size_t)z * ey[l] + y) * ex[l] + x];
    };
    // the cut: one flag per brick of each level
    std::vector<std::vector<uint8_t>> active(L);
    for (int l = 0; l < L; ++l) active[l].assign(E[l].size(), 0);
    const int top = L - 1;
    double total = 0.0;
    int64_t n_bricks = 0;
    for (size_t i = 0; i < E[top].size(); ++i) {
      active[top][i] = 1;
      total += E[top][i];
      ++n_bricks;
    }
    auto children = [&](int l, int z, int y, int x, Node* out) {
      int n = 0;
      for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
          for (int dx = 0; dx < 2; ++dx) {
            const int cz = 2 * z + dz, cy = 2 * y + dy, cx = 2 * x + dx;
            if (cz < ez[l - 1] && cy < ey[l - 1] && cx < ex[l - 1]) out[n++] = Node{0.0, l - 1, cz, cy, cx};
          }
      return n;
    };
    auto gain = [&](int l, int z, int y, int x, Node* ch, int* nch) {
      *nch = children(l, z, y, x, ch);
      double g = at(l, z, y, x);
      for (int k = 0; k < *nch; ++k) g -= at(ch[k].l, ch[k].z, ch[k].y, ch[k].x);
      return g / (double)std::max(*nch - 1, 1);
    };
    std::priority_queue<Node, std::vector<Node>, NodeAfter> heap;
    Node ch[8];
    int nch = 0;
    // bricks with error left are split in order of gain per added brick; a
    // split that gains nothing now may still open finer levels that do
    auto push = [&](int l, int z, int y, int x) {
      if (l == 0 || !(at(l, z, y, x) > 0.0)) return;
      Node tmp[8];
      int nt = 0;
      heap.push(Node{gain(l, z, y, x, tmp, &nt), l, z, y, x});
    };
    for (int z = 0; z < ez[top]; ++z)
      for (int y = 0; y < ey[top]; ++y)
        for (int x = 0; x < ex[top]; ++x) push(top, z, y, x);
    while (!heap.empty() && total > budget_sse) {
      const Node b = heap.top();
      heap.pop();
      gain(b.l, b.z, b.y, b.x, ch, &nch);
      if (mem_budget && (uint64_t)(n_bricks + nch - 1) * bb > mem_budget) break;
      active[b.l][((size_t)b.z * ey[b.l] + b.y) * ex[b.l] + b.x] = 0;
      total -= at(b.l, b.z, b.y, b.x);
      for (int k = 0; k < nch; ++k) {
        active[ch[k].l][((size_t)ch[k].z * ey[ch[k].l] + ch[k].y) * ex[ch[k].l] + ch[k].x] = 1;
        total += at(ch[k].l, ch[k].z, ch[k].y, ch[k].x);
      }
      for (int k = 0; k < nch; ++k) push(ch[k].l, ch[k].z, ch[k].y, ch[k].x);
      n_bricks += nch - 1;
    }
    for (int z = 0; z < gz; ++z)
      for (int y = 0; y < gy; ++y)
        for (int x = 0; x < gx; ++x) {
          int l = 0;
          while (l < L && !active[l][((size_t)(z >> l) * ey[l] + (y >> l)) * ex[l] + (x >> l)]) ++l;
          levels[((int64_t)z * gy + y) * gx + x] = (uint8_t)(l < L ? l : 0);
        }
    if (sse_out) *sse_out = std::max(total, 0.0);
    if (bytes_out) *bytes_out = (uint64_t)n_bricks * bb;
    return TVK_OK;
  } catch (const std::bad_alloc&) {
    tvk::set_last_error("tvk_lod_select_octree: host allocation failed");
    return TVK_ERR_OOM;
  }
}

}  // extern "C"
