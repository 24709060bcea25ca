// volume_gen_gpu.hip -- seeded synthetic volumes generated on the device,
// for the workloads too large to synthesise on host threads in a benchmark
// run (2048^2 x 1024 and 4096^3 voxels, the serving pool's 2048^3):
//
//   TVK_VOL_FBM_PLANES   value-noise fBm (5 octaves, lacunarity 2, gain 0.5,
//                        base period 256 voxels) scaled to [0, 0.6 maxval],
//                        plus 12 layered slabs |n.x - d| < thickness (seeded
//                        unit normals, offsets, 2..6 voxel thickness, values
//                        0.68..0.98 maxval);
//   TVK_VOL_NOISE_TUBES  multi-scale value noise (6 octaves, base period
//                        512) scaled to [0, 0.55 maxval], plus 64 thin tubes
//                        (segments between seeded points, radius 1..3
//                        voxels, values 0.78..1.0 maxval);
//   TVK_VOL_FBM_SPHERES  fBm (4 octaves, base period 64) in [0, 0.6 maxval]
//                        plus 20 solid spheres (radius 2..12% of the largest
//                        dimension, values 0.75..0.98 maxval).
//
// Value noise: lattice values hash(seed, octave, i, j, k) in [0, 1),
// trilinear with the quintic fade 6t^5 - 15t^4 + 10t^3.  Every voxel is a
// pure function of (kind, seed, dims, voxel index), so slabs can be made in
// any order; the features (slabs, tubes, spheres) are drawn on the host
// from the seed with splitmix64.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/tvk_hip.h"
#include "kernels.hpp"

namespace tvk {
namespace {

constexpr int kMaxFeatures = 64;

struct Feature {
  float a[3], b[3];  // plane: normal, (offset, thickness, -); tube: endpoints; sphere: centre
  float r;           // tube / sphere radius
  float value;
};

struct GenArgs {
  int kind;
  int octaves;
  float base_period;
  float noise_scale;  // fBm in [0, 1) -> [0, noise_scale]
  uint32_t seed_lo, seed_hi;
  int n_features;
  float maxval;
  int64_t dims[3];
  int64_t z0;         // first z of this slab
  Feature f[kMaxFeatures];
};

__host__ __device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__device__ inline float lattice(uint32_t s, int i, int j, int k) {
  const uint32_t h = mix32(s ^ mix32((uint32_t)i * 0x9e3779b1u ^ mix32((uint32_t)j * 0x85ebca77u ^
                                                                         mix32((uint32_t)k * 0xc2b2ae3du))));
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

__device__ inline float fade(float t) { return t * t * t * (t * (t * 6.f - 15.f) + 10.f); }

__device__ float value_noise(uint32_t s, float x, float y, float z) {
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  const int i = (int)fx, j = (int)fy, k = (int)fz;
  const float u = fade(x - fx), v = fade(y - fy), w = fade(z - fz);
  const float c000 = lattice(s, i, j, k), c100 = lattice(s, i + 1, j, k);
  const float c010 = lattice(s, i, j + 1, k), c110 = lattice(s, i + 1, j + 1, k);
  const float c001 = lattice(s, i, j, k + 1), c101 = lattice(s, i + 1, j, k + 1);
  const float c011 = lattice(s, i, j + 1, k + 1), c111 = lattice(s, i + 1, j + 1, k + 1);
  const float x00 = c000 + (c100 - c000) * u, x10 = c010 + (c110 - c010) * u;
  const float x01 = c001 + (c101 - c001) * u, x11 = c011 + (c111 - c011) * u;
  const float y0 = x00 + (x10 - x00) * v, y1 = x01 + (x11 - x01) * v;
  return y0 + (y1 - y0) * w;
}

template <typename T>
__global__ void generate_kernel(const GenArgs A, T* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t nx = A.dims[0], ny = A.dims[1];
  const int64_t x = i % nx, y = (i / nx) % ny, z = A.z0 + i / (nx * ny);
  const float px = (float)x + 0.5f, py = (float)y + 0.5f, pz = (float)z + 0.5f;
  // fBm, normalised by the sum of the octave amplitudes
  float sum = 0.f, amp = 1.f, norm = 0.f, f = 1.f / A.base_period;
  for (int o = 0; o < A.octaves; ++o) {
    const uint32_t s = mix32(A.seed_lo ^ mix32(A.seed_hi + 0x51ed27u * (uint32_t)(o + 1)));
    sum += amp * value_noise(s, px * f, py * f, pz * f);
    norm += amp;
    amp *= 0.5f;
    f *= 2.f;
  }
  float v = sum / norm * A.noise_scale;
  for (int k = 0; k < A.n_features; ++k) {
    const Feature& F = A.f[k];
    bool in = false;
    if (A.kind == TVK_VOL_FBM_PLANES) {
      const float d = F.a[0] * px + F.a[1] * py + F.a[2] * pz - F.b[0];
      in = fabsf(d) < F.b[1];
    } else if (A.kind == TVK_VOL_NOISE_TUBES) {
      const float ab[3] = {F.b[0] - F.a[0], F.b[1] - F.a[1], F.b[2] - F.a[2]};
      const float ap[3] = {px - F.a[0], py - F.a[1], pz - F.a[2]};
      const float l2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
      float t = (ap[0] * ab[0] + ap[1] * ab[1] + ap[2] * ab[2]) / fmaxf(l2, 1e-12f);
      t = fminf(fmaxf(t, 0.f), 1.f);
      const float d0 = ap[0] - t * ab[0], d1 = ap[1] - t * ab[1], d2 = ap[2] - t * ab[2];
      in = d0 * d0 + d1 * d1 + d2 * d2 < F.r * F.r;
    } else {
      const float d0 = px - F.a[0], d1 = py - F.a[1], d2 = pz - F.a[2];
      in = d0 * d0 + d1 * d1 + d2 * d2 < F.r * F.r;
    }
    if (in) v = F.value;
  }
  out[i] = (T)fminf(fmaxf(floorf(v + 0.5f), 0.f), A.maxval);
}

uint64_t splitmix(uint64_t& s) {
  uint64_t x = (s += 0x9e3779b97f4a7c15ull);
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

double unif(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

bool is_gpu_volume_kind(int kind) {
  return kind == TVK_VOL_FBM_PLANES || kind == TVK_VOL_NOISE_TUBES || kind == TVK_VOL_FBM_SPHERES;
}

hipError_t launch_generate_volume(int kind, uint64_t seed, const int64_t dims[3], int bits,
                                  int64_t z0, int64_t nz, void* out, hipStream_t stream) {
  GenArgs A{};
  A.kind = kind;
  A.seed_lo = (uint32_t)seed;
  A.seed_hi = (uint32_t)(seed >> 32);
  A.maxval = (float)((1u << bits) - 1);
  for (int i = 0; i < 3; ++i) A.dims[i] = dims[i];
  A.z0 = z0;
  uint64_t s = seed ^ 0x7a11ull * (uint64_t)(kind + 1);
  const double m = (double)std::max(dims[0], std::max(dims[1], dims[2]));
  auto point = [&](float p[3]) {
    for (int i = 0; i < 3; ++i) p[i] = (float)(unif(s) * (double)dims[i]);
  };
  if (kind == TVK_VOL_FBM_PLANES) {
    A.octaves = 5;
    A.base_period = 256.f;
    A.noise_scale = 0.6f * A.maxval;
    A.n_features = 12;
    for (int k = 0; k < A.n_features; ++k) {
      Feature& F = A.f[k];
      double nrm[3], l = 0;
      for (double& c : nrm) {
        c = unif(s) * 2.0 - 1.0;
        l += c * c;
      }
      l = sqrt(l > 1e-12 ? l : 1.0);
      float c[3];
      point(c);
      double off = 0;
      for (int i = 0; i < 3; ++i) {
        F.a[i] = (float)(nrm[i] / l);
        off += F.a[i] * c[i];
      }
      F.b[0] = (float)off;
      F.b[1] = (float)(2.0 + 4.0 * unif(s));
      F.value = (float)((0.68 + 0.30 * unif(s)) * A.maxval);
    }
  } else if (kind == TVK_VOL_NOISE_TUBES) {
    A.octaves = 6;
    A.base_period = 512.f;
    A.noise_scale = 0.55f * A.maxval;
    A.n_features = 64;
    for (int k = 0; k < A.n_features; ++k) {
      Feature& F = A.f[k];
      point(F.a);
      point(F.b);
      F.r = (float)(1.0 + 2.0 * unif(s));
      F.value = (float)((0.78 + 0.22 * unif(s)) * A.maxval);
    }
  } else if (kind == TVK_VOL_FBM_SPHERES) {
    A.octaves = 4;
    A.base_period = 64.f;
    A.noise_scale = 0.6f * A.maxval;
    A.n_features = 20;
    for (int k = 0; k < A.n_features; ++k) {
      Feature& F = A.f[k];
      point(F.a);
      F.r = (float)((0.02 + 0.10 * unif(s)) * m);
      F.value = (float)((0.75 + 0.23 * unif(s)) * A.maxval);
    }
  } else {
    return hipErrorInvalidValue;
  }
  const int64_t n = dims[0] * dims[1] * nz;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (bits == 8)
    generate_kernel<uint8_t><<<blocks, 256, 0, stream>>>(A, (uint8_t*)out, n);
  else
    generate_kernel<uint16_t><<<blocks, 256, 0, stream>>>(A, (uint16_t*)out, n);
  return hipGetLastError();
}

}  // namespace tvk
