// volume_gen.cpp -- seeded synthetic scalar volumes, the data sources of the
// benchmark workloads and tests (the role of UVFReader's GenerateVolumeData,
// UVFReader/DataSource.h:335-396, whose const/random/sphere kinds are kept).
//
// Every generator is deterministic in (kind, seed, dims) and evaluated in
// double precision on host threads (z slabs).  Per-voxel randomness comes
// from a counter-based hash of (seed, voxel index), so the result does not
// depend on the thread count.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/tvk_hip.h"
#include "errors.hpp"

namespace tvk {
namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Sequential generator for the few global parameters (blob centres ...).
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(splitmix64(seed ^ 0x5eedull)) {}
  uint64_t next() { return s = splitmix64(s); }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
  }
};

inline double hash_uniform(uint64_t seed, uint64_t idx) {
  return (double)(splitmix64(splitmix64(seed) ^ idx) >> 11) * (1.0 / 9007199254740992.0);
}

template <typename F>
void parallel_z(uint64_t nz, F&& f) {
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 16u));
  if (nz < nt) nt = (unsigned)nz;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (uint64_t z = t; z < nz; z += nt) f(z);
    });
  for (auto& x : th) x.join();
}

// Improved Perlin noise (Perlin, "Improving Noise", SIGGRAPH 2002) over a
// seeded permutation of 0..255.
struct Perlin {
  int p[512];
  explicit Perlin(uint64_t seed) {
    int q[256];
    for (int i = 0; i < 256; ++i) q[i] = i;
    Rng r(seed ^ 0x9e71ull);
    for (int i = 255; i > 0; --i) {
      const int j = (int)(r.next() % (uint64_t)(i + 1));
      std::swap(q[i], q[j]);
    }
    for (int i = 0; i < 512; ++i) p[i] = q[i & 255];
  }
  static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
  static double lerp(double t, double a, double b) { return a + t * (b - a); }
  static double grad(int h, double x, double y, double z) {
    h &= 15;
    const double u = h < 8 ? x : y;
    const double v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
  }
  double noise(double x, double y, double z) const {
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
    x -= fx;
    y -= fy;
    z -= fz;
    const double u = fade(x), v = fade(y), w = fade(z);
    const int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
    const int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    return lerp(w,
                lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                     lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(p[AB + 1], x, y - 1, z - 1), grad(p[BB + 1], x - 1, y - 1, z - 1))));
  }
  double fbm(double x, double y, double z, int octaves, double lac, double gain) const {
    double s = 0, a = 1, f = 1;
    for (int o = 0; o < octaves; ++o) {
      s += a * noise(x * f, y * f, z * f);
      a *= gain;
      f *= lac;
    }
    return s;
  }
};

template <typename T>
inline T clamp_round(double v, double maxv) {
  v = floor(v + 0.5);
  if (v < 0) v = 0;
  if (v > maxv) v = maxv;
  return (T)v;
}

template <typename T>
int generate(int kind, uint64_t seed, uint64_t nx, uint64_t ny, uint64_t nz, T* out) {
  const double maxv = (double)((1ull << (8 * sizeof(T))) - 1);
  const uint64_t sxy = nx * ny;
  switch (kind) {
    case TVK_VOL_CONST: {
      const T v = (T)(seed % ((uint64_t)maxv + 1));
      parallel_z(nz, [&](uint64_t z) { std::fill(out + z * sxy, out + (z + 1) * sxy, v); });
      return TVK_OK;
    }
    case TVK_VOL_RANDOM: {
      parallel_z(nz, [&](uint64_t z) {
        for (uint64_t i = z * sxy; i < (z + 1) * sxy; ++i) out[i] = (T)(splitmix64(splitmix64(seed) ^ i) >> (64 - 8 * sizeof(T)));
      });
      return TVK_OK;
    }
    case TVK_VOL_SPHERE: {
      // solid ball of value maxv, radius 0.4 * min dim, centred
      const double cx = 0.5 * (nx - 1), cy = 0.5 * (ny - 1), cz = 0.5 * (nz - 1);
      const double r = 0.4 * (double)std::min(nx, std::min(ny, nz));
      parallel_z(nz, [&](uint64_t z) {
        for (uint64_t y = 0; y < ny; ++y)
          for (uint64_t x = 0; x < nx; ++x) {
            const double dx = x - cx, dy = y - cy, dz = z - cz;
            out[z * sxy + y * nx + x] = (dx * dx + dy * dy + dz * dz <= r * r) ? (T)maxv : (T)0;
          }
      });
      return TVK_OK;
    }
    case TVK_VOL_BLOBS: {
      // 60 isotropic Gaussian blobs (sigma 2 voxels, peak 200) centred on a
      // sphere of radius 20 voxels (scaled with the volume) around the
      // centre, plus uniform noise in [0, 8).
      Rng r(seed);
      const double scale = (double)std::min(nx, std::min(ny, nz)) / 64.0;
      const double c0[3] = {0.5 * (nx - 1), 0.5 * (ny - 1), 0.5 * (nz - 1)};
      double cen[60][3];
      for (int b = 0; b < 60; ++b) {
        double u[3] = {r.normal(), r.normal(), r.normal()};
        double l = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        if (l < 1e-12) l = 1;
        for (int i = 0; i < 3; ++i) cen[b][i] = c0[i] + 20.0 * scale * u[i] / l;
      }
      const double s2 = 2.0 * (2.0 * scale) * (2.0 * scale);
      parallel_z(nz, [&](uint64_t z) {
        for (uint64_t y = 0; y < ny; ++y)
          for (uint64_t x = 0; x < nx; ++x) {
            double v = 0;
            for (int b = 0; b < 60; ++b) {
              const double dx = x - cen[b][0], dy = y - cen[b][1], dz = z - cen[b][2];
              v += 200.0 * exp(-(dx * dx + dy * dy + dz * dz) / s2);
            }
            const uint64_t i = z * sxy + y * nx + x;
            v += 8.0 * hash_uniform(seed, i);
            out[i] = clamp_round<T>(v, maxv);
          }
      });
      return TVK_OK;
    }
    case TVK_VOL_PERLIN_SPHERES: {
      // 5-octave improved-Perlin fBm (lacunarity 2, gain 0.5, 4 base cycles
      // across the largest dimension), affine-scaled to [0, 4095] by its
      // min/max over the volume, plus 20 solid spheres (radius 10-60
      // voxels scaled with the volume, value 3000-4000).
      const Perlin pn(seed);
      const double f0 = 4.0 / (double)std::max(nx, std::max(ny, nz));
      std::vector<float> tmp(nx * ny * nz);
      parallel_z(nz, [&](uint64_t z) {
        for (uint64_t y = 0; y < ny; ++y)
          for (uint64_t x = 0; x < nx; ++x)
            tmp[z * sxy + y * nx + x] = (float)pn.fbm((x + 0.5) * f0, (y + 0.5) * f0, (z + 0.5) * f0, 5, 2.0, 0.5);
      });
      float lo = tmp[0], hi = tmp[0];
      for (float v : tmp) {
        lo = std::min(lo, v);
        hi = std::max(hi, v);
      }
      const double top = std::min(4095.0, maxv);
      const double sc = hi > lo ? top / ((double)hi - (double)lo) : 0.0;
      Rng r(seed ^ 0x5fe7e5ull);
      const double dscale = (double)std::max(nx, std::max(ny, nz)) / 512.0;
      struct Ball { double c[3], r; double v; };
      std::vector<Ball> balls(20);
      for (auto& b : balls) {
        b.c[0] = r.uniform() * nx;
        b.c[1] = r.uniform() * ny;
        b.c[2] = r.uniform() * nz;
        b.r = (10.0 + 50.0 * r.uniform()) * dscale;
        b.v = std::min(maxv, (double)(3000 + (int)(r.next() % 1001)));
      }
      parallel_z(nz, [&](uint64_t z) {
        for (uint64_t y = 0; y < ny; ++y)
          for (uint64_t x = 0; x < nx; ++x) {
            const uint64_t i = z * sxy + y * nx + x;
            double v = ((double)tmp[i] - (double)lo) * sc;
            for (const auto& b : balls) {
              const double dx = x + 0.5 - b.c[0], dy = y + 0.5 - b.c[1], dz = z + 0.5 - b.c[2];
              if (dx * dx + dy * dy + dz * dz <= b.r * b.r) v = b.v;
            }
            out[i] = clamp_round<T>(v, maxv);
          }
      });
      return TVK_OK;
    }
    default:
      set_last_error("unknown synthetic volume kind");
      return TVK_ERR_ARG;
  }
}

}  // namespace
}  // namespace tvk

extern "C" int tvk_generate_volume(int kind, uint64_t seed, uint64_t nx, uint64_t ny, uint64_t nz,
                                   uint32_t bits, void* out, uint64_t out_bytes) {
  if (!out || (bits != 8 && bits != 16) || nx == 0 || ny == 0 || nz == 0) {
    tvk::set_last_error("bad volume generator arguments");
    return TVK_ERR_ARG;
  }
  if (out_bytes < nx * ny * nz * (bits / 8)) {
    tvk::set_last_error("output buffer too small");
    return TVK_ERR_ARG;
  }
  try {
    if (bits == 8) return tvk::generate<uint8_t>(kind, seed, nx, ny, nz, (uint8_t*)out);
    return tvk::generate<uint16_t>(kind, seed, nx, ny, nz, (uint16_t*)out);
  } catch (...) {
    tvk::set_last_error("host allocation failed");
    return TVK_ERR_OOM;
  }
}
