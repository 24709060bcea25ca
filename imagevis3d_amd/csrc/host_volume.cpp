// host_volume.cpp -- host-side halves of the out-of-core (paged) path: the
// LOD pyramid, per-brick value ranges and brick extraction for a volume that
// stays in host memory (the data behind UVFDataset::GetBrick,
// doc/import.adoc:761-775).  They use the same rules as the GPU kernels of
// kernels_brick.hip (floor mean over existing children, clamp-to-edge
// aprons) so a paged volume renders exactly like an in-core one.
#include "host_volume.hpp"

#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

namespace tvk {
namespace {

template <typename F>
void parallel_rows(int64_t rows, F&& f) {
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 16u));
  if (rows < (int64_t)nt) nt = (unsigned)std::max<int64_t>(1, rows);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (int64_t r = t; r < rows; r += nt) f(r);
    });
  for (auto& x : th) x.join();
}

template <typename T>
void downsample_t(const T* src, const int64_t d[3], T* dst) {
  const int64_t mx = (d[0] + 1) / 2, my = (d[1] + 1) / 2, mz = (d[2] + 1) / 2;
  parallel_rows(my * mz, [&](int64_t row) {
    const int64_t y = row % my, z = row / my;
    for (int64_t x = 0; x < mx; ++x) {
      uint32_t sum = 0, cnt = 0;
      for (int dz = 0; dz < 2; ++dz) {
        const int64_t sz = 2 * z + dz;
        if (sz >= d[2]) continue;
        for (int dy = 0; dy < 2; ++dy) {
          const int64_t sy = 2 * y + dy;
          if (sy >= d[1]) continue;
          const T* r = src + (sz * d[1] + sy) * d[0];
          for (int dx = 0; dx < 2; ++dx) {
            const int64_t sx = 2 * x + dx;
            if (sx >= d[0]) continue;
            sum += r[sx];
            ++cnt;
          }
        }
      }
      dst[(z * my + y) * mx + x] = (T)(sum / cnt);
    }
  });
}

template <typename T>
void extract_t(const T* level, const int64_t d[3], int B, int a, int bx, int by, int bz, T* slot) {
  const int S = B + 2 * a;
  for (int z = 0; z < S; ++z) {
    const int64_t gz = std::min<int64_t>(std::max<int64_t>((int64_t)bz * B - a + z, 0), d[2] - 1);
    for (int y = 0; y < S; ++y) {
      const int64_t gy = std::min<int64_t>(std::max<int64_t>((int64_t)by * B - a + y, 0), d[1] - 1);
      const T* row = level + (gz * d[1] + gy) * d[0];
      T* out = slot + ((int64_t)z * S + y) * S;
      const int64_t x0 = (int64_t)bx * B - a;
      if (x0 >= 0 && x0 + S <= d[0]) {
        memcpy(out, row + x0, (size_t)S * sizeof(T));
      } else {
        for (int x = 0; x < S; ++x)
          out[x] = row[std::min<int64_t>(std::max<int64_t>(x0 + x, 0), d[0] - 1)];
      }
    }
  }
}

template <typename T>
void ranges_t(const T* level, const int64_t d[3], int B, int a, const int g[3], float* out) {
  const int64_t nb = (int64_t)g[0] * g[1] * g[2];
  parallel_rows(nb, [&](int64_t b) {
    const int bx = (int)(b % g[0]), by = (int)((b / g[0]) % g[1]), bz = (int)(b / ((int64_t)g[0] * g[1]));
    const int64_t x0 = std::max<int64_t>((int64_t)bx * B - a, 0), x1 = std::min<int64_t>((int64_t)(bx + 1) * B + a, d[0]);
    const int64_t y0 = std::max<int64_t>((int64_t)by * B - a, 0), y1 = std::min<int64_t>((int64_t)(by + 1) * B + a, d[1]);
    const int64_t z0 = std::max<int64_t>((int64_t)bz * B - a, 0), z1 = std::min<int64_t>((int64_t)(bz + 1) * B + a, d[2]);
    uint32_t lo = 0xffffffffu, hi = 0;
    for (int64_t z = z0; z < z1; ++z)
      for (int64_t y = y0; y < y1; ++y) {
        const T* r = level + (z * d[1] + y) * d[0];
        for (int64_t x = x0; x < x1; ++x) {
          lo = std::min<uint32_t>(lo, r[x]);
          hi = std::max<uint32_t>(hi, r[x]);
        }
      }
    out[2 * b] = (float)lo;
    out[2 * b + 1] = (float)hi;
  });
}

}  // namespace

void host_downsample(const void* src, int bpv, const int64_t dims[3], void* dst) {
  if (bpv == 1) downsample_t((const uint8_t*)src, dims, (uint8_t*)dst);
  else downsample_t((const uint16_t*)src, dims, (uint16_t*)dst);
}

void host_extract_brick(const void* level, int bpv, const int64_t dims[3], int B, int apron, int bx,
                        int by, int bz, void* slot) {
  if (bpv == 1) extract_t((const uint8_t*)level, dims, B, apron, bx, by, bz, (uint8_t*)slot);
  else extract_t((const uint16_t*)level, dims, B, apron, bx, by, bz, (uint16_t*)slot);
}

void host_brick_ranges(const void* level, int bpv, const int64_t dims[3], int B, int apron,
                       const int grid[3], float* out) {
  if (bpv == 1) ranges_t((const uint8_t*)level, dims, B, apron, grid, out);
  else ranges_t((const uint16_t*)level, dims, B, apron, grid, out);
}

}  // namespace tvk
