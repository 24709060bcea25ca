// kernels_brick.hip -- HBM-streaming kernels of the UVF bricking / LOD path
// (RasterDataBlock::FlatDataToBrickedLOD + CombineAverage + SimpleMaxMin,
// UVFReader/DataSource.h:550-570; Histogram1D/2DDataBlock::Compute,
// UVFReader/DataSource.h:595-636).  Integer in, integer out: every result
// is bit-exact with the CPU twins in oracle/oracle.c.
//
// All of these are byte/integer streaming work bounded by HBM bandwidth:
// threads walk rows of x so each wavefront reads and writes contiguous
// runs, and reductions go through wave shuffles before one atomic per
// workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace tvk {
namespace {

constexpr int kThreads = 256;

// 2x2x2 reduction of the existing children of each parent voxel: integer
// mean floor(sum / count) (RasterDataBlock's CombineAverage rule), or with
// MEDIAN the lower median -- element (count-1)/2 of the sorted children
// (the UseMedian option of the TOC builder).
template <typename T, bool MEDIAN>
__global__ __launch_bounds__(kThreads) void downsample_kernel(const T* __restrict__ src, int64_t nx,
                                                              int64_t ny, int64_t nz,
                                                              T* __restrict__ dst, int64_t mx,
                                                              int64_t my, int64_t mz) {
  const int64_t x = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (x >= mx) return;
  // rows of a level stay below 2^31 for every supported volume, so the
  // per-row index split uses 32-bit division.
  const uint32_t rows = (uint32_t)(my * mz), umy = (uint32_t)my;
  const int64_t x0 = 2 * x, x1 = 2 * x + 1 < nx ? 2 * x + 1 : -1;
  for (uint32_t row = blockIdx.y; row < rows; row += gridDim.y) {
    const int64_t y = row % umy, z = row / umy;
    const int64_t ys[2] = {2 * y, 2 * y + 1 < ny ? 2 * y + 1 : -1};
    const int64_t zs[2] = {2 * z, 2 * z + 1 < nz ? 2 * z + 1 : -1};
    uint32_t sum = 0, cnt = 0;
    uint32_t v[8];
#pragma unroll
    for (int dz = 0; dz < 2; ++dz) {
      if (zs[dz] < 0) continue;
#pragma unroll
      for (int dy = 0; dy < 2; ++dy) {
        if (ys[dy] < 0) continue;
        const T* r = src + (zs[dz] * ny + ys[dy]) * nx;
        if (MEDIAN) {
          v[cnt++] = r[x0];
          if (x1 >= 0) v[cnt++] = r[x1];
        } else {
          sum += r[x0];
          ++cnt;
          if (x1 >= 0) {
            sum += r[x1];
            ++cnt;
          }
        }
      }
    }
    uint32_t out;
    if (MEDIAN) {
      // insertion sort of the (at most 8) children
      for (uint32_t i = 1; i < cnt; ++i) {
        const uint32_t key = v[i];
        uint32_t j = i;
        while (j > 0 && v[j - 1] > key) {
          v[j] = v[j - 1];
          --j;
        }
        v[j] = key;
      }
      out = v[(cnt - 1) / 2];
    } else {
      out = sum / cnt;
    }
    dst[(z * my + y) * mx + x] = (T)out;
  }
}

__device__ inline uint32_t wave_min(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}
__device__ inline uint32_t wave_max(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}
__device__ inline unsigned long long wave_max64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

__device__ inline int64_t clamp64(int64_t v, int64_t lo, int64_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Slot-local voxel of element i of a slot in the blocked slot layout
// (tvk_common.hpp); the overlap voxel past a block row's end and padding
// past the slot's edge repeat the slot's last voxel along that axis.
__device__ inline void slot_element_voxel(uint32_t i, int S, int& x, int& y, int& z) {
  const uint32_t n = (uint32_t)slot_blocks(S);
  const uint32_t b = i / kSlotBlockElems, r = i % kSlotBlockElems;
  const uint32_t row = r / 5, e = r % 5;
  x = min((int)((b % n) * 4 + e), S - 1);
  y = min((int)(((b / n) % n) * 4 + (row & 3)), S - 1);
  z = min((int)((b / (n * n)) * 4 + (row >> 2)), S - 1);
}

// grid.x: chunks of one slot, grid.y: brick job.  Writes the slot in
// element order (coalesced), reading the level voxel each element holds.
template <typename T>
__global__ __launch_bounds__(kThreads) void extract_bricks_kernel(
    const T* __restrict__ level, int64_t nx, int64_t ny, int64_t nz,
    const BrickJob* __restrict__ jobs, T* __restrict__ pool, int64_t slot_stride, int64_t slot_elems,
    int B, int apron, uint32_t* __restrict__ minmax) {
  const BrickJob job = jobs[blockIdx.y];
  const int S = B + 2 * apron;
  const int64_t ox = (int64_t)job.bx * B - apron;
  const int64_t oy = (int64_t)job.by * B - apron;
  const int64_t oz = (int64_t)job.bz * B - apron;
  T* __restrict__ slot = pool + job.slot * slot_stride;
  const int64_t chunk = (slot_elems + gridDim.x - 1) / gridDim.x;
  const int64_t beg = chunk * blockIdx.x;
  const int64_t end = beg + chunk < slot_elems ? beg + chunk : slot_elems;
  uint32_t lo = 0xffffffffu, hi = 0;
  for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) {
    int x, y, z;
    slot_element_voxel((uint32_t)i, S, x, y, z);  // slot_elems < 2^31 (checked by the host)
    const int64_t gx = clamp64(ox + x, 0, nx - 1);
    const int64_t gy = clamp64(oy + y, 0, ny - 1);
    const int64_t gz = clamp64(oz + z, 0, nz - 1);
    const T v = level[(gz * ny + gy) * nx + gx];
    slot[i] = v;
    lo = min(lo, (uint32_t)v);
    hi = max(hi, (uint32_t)v);
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&minmax[2 * blockIdx.y], lo);
    atomicMax(&minmax[2 * blockIdx.y + 1], hi);
  }
}

// Overlap-inclusive bricks of a flat level, x fastest (the TOC builder's
// stored bricks): job j's brick covers level voxels [b*P - o, b*P - o + S)
// per axis (S = P + 2o); voxels outside the level repeat the nearest edge
// voxel (clamp) or are 0.  Brick j goes to out + j * S^3 (j = the job's
// slot field); its (min, max) over the stored voxels to minmax[2j], [2j+1].
// grid.x: chunks of one brick, grid.y: job.
template <typename T>
__global__ __launch_bounds__(kThreads) void extract_linear_bricks_kernel(
    const T* __restrict__ level, int64_t nx, int64_t ny, int64_t nz,
    const BrickJob* __restrict__ jobs, T* __restrict__ out, int P, int o, int clamp,
    uint32_t* __restrict__ minmax) {
  const BrickJob job = jobs[blockIdx.y];
  const int64_t S = P + 2 * o;
  const int64_t n = S * S * S;
  const int64_t ox = (int64_t)job.bx * P - o, oy = (int64_t)job.by * P - o, oz = (int64_t)job.bz * P - o;
  T* __restrict__ dst = out + job.slot * n;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t beg = chunk * blockIdx.x;
  const int64_t end = beg + chunk < n ? beg + chunk : n;
  uint32_t lo = 0xffffffffu, hi = 0;
  for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) {
    const int64_t gx = ox + i % S, gy = oy + (i / S) % S, gz = oz + i / (S * S);
    T v;
    if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz)
      v = level[(gz * ny + gy) * nx + gx];
    else if (clamp)
      v = level[(clamp64(gz, 0, nz - 1) * ny + clamp64(gy, 0, ny - 1)) * nx + clamp64(gx, 0, nx - 1)];
    else
      v = (T)0;
    dst[i] = v;
    lo = min(lo, (uint32_t)v);
    hi = max(hi, (uint32_t)v);
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&minmax[2 * job.slot], lo);
    atomicMax(&minmax[2 * job.slot + 1], hi);
  }
}

// An x-fastest S^3 brick into the blocked slot layout, and back.
template <typename T>
__global__ __launch_bounds__(kThreads) void linear_to_slot_kernel(const T* __restrict__ src,
                                                                  T* __restrict__ slot, int S,
                                                                  int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  slot_element_voxel((uint32_t)i, S, x, y, z);
  slot[i] = src[((int64_t)z * S + y) * S + x];
}

template <typename T>
__global__ __launch_bounds__(kThreads) void slot_to_linear_kernel(const T* __restrict__ slot,
                                                                  T* __restrict__ dst, int S,
                                                                  int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % S), y = (int)((i / S) % S), z = (int)(i / ((int64_t)S * S));
  dst[i] = slot[slot_offset(slot_geo(S), x, y, z)];
}

// Many waves per empty-space cell: the cell's neighbourhood [c*C - 1,
// (c+1)*C] (clamped) is cut into rows of x; a wave takes max(1, 64 / w)
// rows at a time with its lanes over x, the work-groups along grid.y split
// the rows among them, and each work-group folds its minimum and maximum
// into the cell with one atomic each (lo starts at ~0, hi at 0).  The work
// per launch does not depend on the cell size.  Work-groups along grid.x
// take cells blockIdx.x, blockIdx.x + gridDim.x, ... (a 4096^3 level has
// 16.7 M cells of 16^3, more work-groups than one launch may have).
template <typename T>
__global__ __launch_bounds__(kThreads) void cell_range_rows_kernel(const T* __restrict__ level,
                                                                   int64_t nx, int64_t ny, int64_t nz,
                                                                   int C, int ncx, int ncy, int64_t n_cells,
                                                                   uint32_t* __restrict__ lo_out,
                                                                   uint32_t* __restrict__ hi_out) {
  __shared__ uint32_t s_lo[kThreads / 64], s_hi[kThreads / 64];
  const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int waves = (int)gridDim.y * (kThreads / 64);
  const int wg = (int)blockIdx.y * (kThreads / 64) + wave;
  // the loop bound depends on blockIdx only: every thread of the group runs
  // the same iterations, so the barriers below are reached by all of them
  for (int64_t cell = blockIdx.x; cell < n_cells; cell += gridDim.x) {
    const int64_t cx = cell % ncx, cy = (cell / ncx) % ncy, cz = cell / ((int64_t)ncx * ncy);
    const int64_t x0 = clamp64(cx * C - 1, 0, nx - 1), x1 = clamp64((cx + 1) * C, 0, nx - 1);
    const int64_t y0 = clamp64(cy * C - 1, 0, ny - 1), y1 = clamp64((cy + 1) * C, 0, ny - 1);
    const int64_t z0 = clamp64(cz * C - 1, 0, nz - 1), z1 = clamp64((cz + 1) * C, 0, nz - 1);
    const int w = (int)(x1 - x0 + 1), h = (int)(y1 - y0 + 1), dz = (int)(z1 - z0 + 1);
    const int rows = h * dz;
    const int rpw = w < 64 ? 64 / w : 1;            // rows per wave step
    const int lrow = w < 64 ? lane / w : 0;          // this lane's row in the step
    const int lx = w < 64 ? lane % w : lane;         // and its first x
    const int xstep = w < 64 ? w : 64;
    uint32_t lo = 0xffffffffu, hi = 0;
    if (lrow < rpw) {
      for (int r = wg * rpw + lrow; r < rows; r += waves * rpw) {
        const int y = r % h, z = r / h;
        const T* __restrict__ row = level + ((z0 + z) * ny + (y0 + y)) * nx + x0;
        for (int x = lx; x < w; x += xstep) {
          const uint32_t v = row[x];
          lo = min(lo, v);
          hi = max(hi, v);
        }
      }
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    if (lane == 0) {
      s_lo[wave] = lo;
      s_hi[wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < kThreads / 64; ++k) {
        lo = min(lo, s_lo[k]);
        hi = max(hi, s_hi[k]);
      }
      atomicMin(lo_out + cell, lo);
      atomicMax(hi_out + cell, hi);
    }
    __syncthreads();  // s_lo / s_hi are the next cell's
  }
}

__global__ __launch_bounds__(kThreads) void fill_u32_kernel(uint32_t* __restrict__ p, int64_t n,
                                                            uint32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    p[i] = v;
}

// out = lo | hi << 16 (16-bit voxel values)
__global__ __launch_bounds__(kThreads) void pack_ranges_kernel(uint32_t* __restrict__ out,
                                                               const uint32_t* __restrict__ hi, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    out[i] |= hi[i] << 16;
}

// Cell ranges of one brick from its pool slot (paged pools, after upload).
// Cell c of the brick covers level voxels [c*C - 1, (c+1)*C], clamped into
// the slot, whose clamp-to-edge apron gives the values clamping to the level
// would.  Work-group (c, g) folds every gridDim.y-th row step of cell c (a
// wave takes 64/w rows of w < 64 voxels at once, else one row) into
// lo[c] / hi[c] with atomics, so a large cell spreads over many CUs
// (one wave per 256^3 cell ran 62 ms per brick); slot_cell_pack_kernel then
// writes the packed ranges and resets lo / hi for the next brick.
template <typename T>
__global__ __launch_bounds__(kThreads) void slot_cell_range_kernel(
    const T* __restrict__ slot, int S, int apron, int B, int C, int bx, int by, int bz,
    int ncx, int ncy, int ncz, uint32_t* __restrict__ lo_out, uint32_t* __restrict__ hi_out) {
  const int per = B / C;
  const int c = (int)blockIdx.x;  // < per^3
  const int cx = bx * per + c % per, cy = by * per + (c / per) % per, cz = bz * per + c / (per * per);
  if (cx >= ncx || cy >= ncy || cz >= ncz) return;  // whole work-group
  const int x0 = max(cx * C - 1 - (bx * B - apron), 0), x1 = min((cx + 1) * C - (bx * B - apron), S - 1);
  const int y0 = max(cy * C - 1 - (by * B - apron), 0), y1 = min((cy + 1) * C - (by * B - apron), S - 1);
  const int z0 = max(cz * C - 1 - (bz * B - apron), 0), z1 = min((cz + 1) * C - (bz * B - apron), S - 1);
  const int w = x1 - x0 + 1, h = y1 - y0 + 1, dz = z1 - z0 + 1;
  const int rows = h * dz;
  const int lane = (int)(threadIdx.x & 63);
  const int wave = (int)blockIdx.y * (kThreads / 64) + (int)(threadIdx.x >> 6);
  const int waves = (int)gridDim.y * (kThreads / 64);
  const int rpw = w < 64 ? 64 / w : 1;     // rows per wave step
  const int lrow = w < 64 ? lane / w : 0;  // this lane's row in the step
  const int lx = w < 64 ? lane % w : lane;  // and its first x
  const int xstep = w < 64 ? w : 64;
  const SlotGeo G = slot_geo(S);
  uint32_t lo = 0xffffffffu, hi = 0;
  if (lrow < rpw) {
    for (int r = wave * rpw + lrow; r < rows; r += waves * rpw) {
      const int y = y0 + r % h, z = z0 + r / h;
      const int64_t yz = slot_oy(G, y) + slot_oz(G, z);
      for (int x = x0 + lx; x <= x1; x += xstep) {
        const uint32_t v = slot[yz + slot_ox(x)];
        lo = min(lo, v);
        hi = max(hi, v);
      }
    }
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  if (lane == 0 && hi >= lo) {  // a wave with no rows folds nothing
    atomicMin(lo_out + c, lo);
    atomicMax(hi_out + c, hi);
  }
}

// The brick's cells' ranges into the level's cell-range array, lo / hi
// reset for the next brick.
__global__ __launch_bounds__(kThreads) void slot_cell_pack_kernel(uint32_t* __restrict__ lo,
                                                                  uint32_t* __restrict__ hi, int per,
                                                                  int bx, int by, int bz, int ncx,
                                                                  int ncy, int ncz,
                                                                  uint32_t* __restrict__ out) {
  const int c = (int)(blockIdx.x * kThreads + threadIdx.x);
  if (c >= per * per * per) return;
  const int cx = bx * per + c % per, cy = by * per + (c / per) % per, cz = bz * per + c / (per * per);
  if (cx < ncx && cy < ncy && cz < ncz) out[((int64_t)cz * ncy + cy) * ncx + cx] = lo[c] | (hi[c] << 16);
  lo[c] = 0xffffffffu;
  hi[c] = 0u;
}

__global__ void minmax_init_kernel(uint32_t* mm, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    mm[2 * i] = 0xffffffffu;
    mm[2 * i + 1] = 0u;
  }
}

__global__ void minmax_finish_kernel(const uint32_t* mm, const BrickJob* jobs, int n,
                                     float* brick_range) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int id = jobs[i].brick_id;
    brick_range[2 * id] = (float)mm[2 * i];
    brick_range[2 * id + 1] = (float)mm[2 * i + 1];
  }
}

__global__ void fill_u64_kernel(unsigned long long* p, int64_t n, unsigned long long v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// u8: privatised 256-bin histogram in LDS, one u64 atomic per bin per block.
__global__ __launch_bounds__(kThreads) void hist1d_u8_kernel(const uint8_t* __restrict__ v,
                                                             int64_t n,
                                                             unsigned long long* hist) {
  __shared__ uint32_t bins[256];
  bins[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    atomicAdd(&bins[v[i]], 1u);
  __syncthreads();
  if (bins[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)bins[threadIdx.x]);
}

// u16: 65536 bins do not fit a block's LDS; count in global memory.
__global__ __launch_bounds__(kThreads) void hist1d_u16_kernel(const uint16_t* __restrict__ v,
                                                              int64_t n,
                                                              unsigned long long* hist) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    atomicAdd(&hist[v[i]], 1ull);
}

// Integer squared gradient in units of 1/(2*maxval): central differences
// v[c+1]-v[c-1] inside, 2*(one-sided difference) at the borders (the
// numpy.gradient definition that reproduces the 2D histogram stored in
// DemoData/c60.uvf).
template <typename T>
__device__ inline long long grad_k(const T* __restrict__ v, int64_t i, int64_t c, int64_t n,
                                   int64_t stride) {
  if (n < 2) return 0;
  if (c == 0) return 2ll * ((long long)v[i + stride] - (long long)v[i]);
  if (c == n - 1) return 2ll * ((long long)v[i] - (long long)v[i - stride]);
  return (long long)v[i + stride] - (long long)v[i - stride];
}

template <typename T>
__device__ inline unsigned long long grad_K(const T* __restrict__ v, int64_t nx, int64_t ny,
                                            int64_t nz, int64_t x, int64_t y, int64_t z) {
  const int64_t i = (z * ny + y) * nx + x;
  const long long gx = grad_k(v, i, x, nx, 1);
  const long long gy = grad_k(v, i, y, ny, nx);
  const long long gz = grad_k(v, i, z, nz, nx * ny);
  return (unsigned long long)(gx * gx + gy * gy + gz * gz);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void gradient_kmax_kernel(const T* __restrict__ v,
                                                                 int64_t nx, int64_t ny,
                                                                 int64_t nz,
                                                                 unsigned long long* kmax) {
  unsigned long long m = 0;
  const int64_t x = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (x < nx) {
    const uint32_t rows = (uint32_t)(ny * nz), uny = (uint32_t)ny;
    for (uint32_t row = blockIdx.y; row < rows; row += gridDim.y) {
      const unsigned long long k = grad_K(v, nx, ny, nz, x, row % uny, row / uny);
      m = k > m ? k : m;
    }
  }
  m = wave_max64(m);
  if ((threadIdx.x & 63) == 0 && m) atomicMax(kmax, m);
}

// The 2D histogram bin of voxel (x, y, z): value >> shift along the value
// axis, b = max{ b : b^2 kmax <= F^2 K } along the gradient axis (F =
// n_grad - 1, K and kmax squared integer gradients), decided exactly in
// integers after a float estimate.
template <typename T>
__device__ inline int64_t hist2d_bin(const T* __restrict__ v, int64_t nx, int64_t ny, int64_t nz,
                                     int64_t x, int64_t y, int64_t z, unsigned long long kmax, int shift,
                                     int n_grad) {
  const unsigned long long F = (unsigned long long)(n_grad - 1);
  const unsigned long long K = grad_K(v, nx, ny, nz, x, y, z);
  long long b = 0;
  if (kmax > 0) {
    b = (long long)floorf((float)F * sqrtf((float)K / (float)kmax));
    if (b < 0) b = 0;
    if (b > (long long)F) b = (long long)F;
    const unsigned long long rhs = F * F * K;
    while (b < (long long)F && (unsigned long long)(b + 1) * (unsigned long long)(b + 1) * kmax <= rhs) ++b;
    while (b > 0 && (unsigned long long)b * (unsigned long long)b * kmax > rhs) --b;
  }
  const int64_t vb = (int64_t)(v[(z * ny + y) * nx + x] >> shift);
  return vb * n_grad + b;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void hist2d_kernel(const T* __restrict__ v, int64_t nx,
                                                          int64_t ny, int64_t nz,
                                                          const unsigned long long* kmax_p,
                                                          int shift, int n_grad,
                                                          unsigned long long* hist) {
  const int64_t x = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (x >= nx) return;
  const unsigned long long kmax = *kmax_p;
  const uint32_t rows = (uint32_t)(ny * nz), uny = (uint32_t)ny;
  for (uint32_t row = blockIdx.y; row < rows; row += gridDim.y) {
    const int64_t y = row % uny, z = row / uny;
    atomicAdd(&hist[hist2d_bin(v, nx, ny, nz, x, y, z, kmax, shift, n_grad)], 1ull);
  }
}

// Privatised histograms: the work-group counts in LDS, 16-bit counters two
// to a 32-bit word (up to 65536 bins = 128 KiB), over whole rows of x and
// at most 65535 voxels so that no counter carries into its neighbour, then
// adds its non-zero bins to the global u64 histogram once.
constexpr int kHistThreads = 512;
constexpr int64_t kHistMaxVoxels = 65535;

__device__ inline void lds_count(uint32_t* words, int64_t bin) {
  atomicAdd(&words[bin >> 1], (bin & 1) ? 0x10000u : 1u);
}

__device__ inline void lds_flush(const uint32_t* words, int nwords, unsigned long long* hist) {
  for (int k = (int)threadIdx.x; k < nwords; k += kHistThreads) {
    const uint32_t wv = words[k];
    if (wv & 0xffffu) atomicAdd(&hist[2 * k], (unsigned long long)(wv & 0xffffu));
    if (wv >> 16) atomicAdd(&hist[2 * k + 1], (unsigned long long)(wv >> 16));
  }
}

template <typename T>
__global__ __launch_bounds__(kHistThreads) void hist2d_lds_kernel(const T* __restrict__ v, int64_t nx,
                                                                  int64_t ny, int64_t nz,
                                                                  const unsigned long long* kmax_p,
                                                                  int shift, int n_grad, int nbins,
                                                                  int rows_per_block,
                                                                  unsigned long long* hist) {
  extern __shared__ uint32_t words[];
  const int nwords = nbins >> 1;
  for (int k = (int)threadIdx.x; k < nwords; k += kHistThreads) words[k] = 0;
  __syncthreads();
  const unsigned long long kmax = *kmax_p;
  const int64_t rows = ny * nz;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  for (int64_t row = r0; row < r1; ++row) {
    const int64_t y = row % ny, z = row / ny;
    for (int64_t x = threadIdx.x; x < nx; x += kHistThreads)
      lds_count(words, hist2d_bin(v, nx, ny, nz, x, y, z, kmax, shift, n_grad));
  }
  __syncthreads();
  lds_flush(words, nwords, hist);
}

__global__ __launch_bounds__(kHistThreads) void hist1d_u16_lds_kernel(const uint16_t* __restrict__ v,
                                                                      int64_t n,
                                                                      unsigned long long* hist) {
  extern __shared__ uint32_t words[];
  constexpr int nwords = 65536 / 2;
  for (int k = (int)threadIdx.x; k < nwords; k += kHistThreads) words[k] = 0;
  __syncthreads();
  const int64_t i0 = (int64_t)blockIdx.x * kHistMaxVoxels;
  const int64_t i1 = i0 + kHistMaxVoxels < n ? i0 + kHistMaxVoxels : n;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kHistThreads) lds_count(words, v[i]);
  __syncthreads();
  lds_flush(words, nwords, hist);
}

// Dynamic LDS of `bytes` for kernel `k` on the current device, if it allows it.
template <typename K>
bool lds_allowed(K k, size_t bytes) {
  int dev = 0, max_lds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return false;
  if ((size_t)max_lds < bytes) return false;
  return hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess;
}

inline dim3 row_grid(int64_t nx, int64_t rows) {
  const int64_t gy = rows < 65535 ? rows : 65535;
  return dim3((unsigned)((nx + kThreads - 1) / kThreads), (unsigned)(gy > 0 ? gy : 1), 1);
}

}  // namespace

hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v,
                           hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + kThreads - 1) / kThreads;
  if (blocks > 65536) blocks = 65536;
  fill_u64_kernel<<<(unsigned)blocks, kThreads, 0, stream>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_downsample(const void* src, int bpv, const int64_t d[3], void* dst,
                             hipStream_t stream, bool median) {
  const int64_t mx = (d[0] + 1) / 2, my = (d[1] + 1) / 2, mz = (d[2] + 1) / 2;
  const dim3 grid = row_grid(mx, my * mz);
  if (bpv == 1) {
    if (median)
      downsample_kernel<uint8_t, true><<<grid, kThreads, 0, stream>>>(
          (const uint8_t*)src, d[0], d[1], d[2], (uint8_t*)dst, mx, my, mz);
    else
      downsample_kernel<uint8_t, false><<<grid, kThreads, 0, stream>>>(
          (const uint8_t*)src, d[0], d[1], d[2], (uint8_t*)dst, mx, my, mz);
  } else {
    if (median)
      downsample_kernel<uint16_t, true><<<grid, kThreads, 0, stream>>>(
          (const uint16_t*)src, d[0], d[1], d[2], (uint16_t*)dst, mx, my, mz);
    else
      downsample_kernel<uint16_t, false><<<grid, kThreads, 0, stream>>>(
          (const uint16_t*)src, d[0], d[1], d[2], (uint16_t*)dst, mx, my, mz);
  }
  return hipGetLastError();
}

hipError_t launch_extract_bricks(const void* level, int bpv, const int64_t d[3],
                                 const BrickJob* jobs, int n_jobs, void* pool,
                                 int64_t slot_stride, int B, int apron, float* brick_range,
                                 uint32_t* scratch, hipStream_t stream) {
  const int64_t slot_elems = slot_block_elems(B + 2 * apron);
  if (n_jobs <= 0) return hipSuccess;
  minmax_init_kernel<<<(n_jobs + 255) / 256, 256, 0, stream>>>(scratch, n_jobs);
  int64_t chunks = (slot_elems + 8191) / 8192;
  if (chunks > 1024) chunks = 1024;
  if (chunks < 1) chunks = 1;
  for (int j0 = 0; j0 < n_jobs; j0 += 65535) {
    const int nj = n_jobs - j0 < 65535 ? n_jobs - j0 : 65535;
    const dim3 grid((unsigned)chunks, (unsigned)nj, 1);
    if (bpv == 1)
      extract_bricks_kernel<uint8_t><<<grid, kThreads, 0, stream>>>(
          (const uint8_t*)level, d[0], d[1], d[2], jobs + j0, (uint8_t*)pool, slot_stride, slot_elems, B, apron,
          scratch + 2 * j0);
    else
      extract_bricks_kernel<uint16_t><<<grid, kThreads, 0, stream>>>(
          (const uint16_t*)level, d[0], d[1], d[2], jobs + j0, (uint16_t*)pool, slot_stride,
          slot_elems, B, apron, scratch + 2 * j0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  minmax_finish_kernel<<<(n_jobs + 255) / 256, 256, 0, stream>>>(scratch, jobs, n_jobs,
                                                                  brick_range);
  return hipGetLastError();
}

hipError_t launch_extract_linear_bricks(const void* level, int bpv, const int64_t d[3],
                                        const BrickJob* jobs, int n_jobs, void* out, int P, int o,
                                        bool clamp, uint32_t* minmax, hipStream_t stream) {
  if (n_jobs <= 0) return hipSuccess;
  minmax_init_kernel<<<(n_jobs + 255) / 256, 256, 0, stream>>>(minmax, n_jobs);
  const int64_t S = P + 2 * o;
  int64_t chunks = (S * S * S + 8191) / 8192;
  if (chunks > 1024) chunks = 1024;
  if (chunks < 1) chunks = 1;
  for (int j0 = 0; j0 < n_jobs; j0 += 65535) {
    const int nj = n_jobs - j0 < 65535 ? n_jobs - j0 : 65535;
    const dim3 grid((unsigned)chunks, (unsigned)nj, 1);
    if (bpv == 1)
      extract_linear_bricks_kernel<uint8_t><<<grid, kThreads, 0, stream>>>(
          (const uint8_t*)level, d[0], d[1], d[2], jobs + j0, (uint8_t*)out, P, o, clamp ? 1 : 0, minmax);
    else
      extract_linear_bricks_kernel<uint16_t><<<grid, kThreads, 0, stream>>>(
          (const uint16_t*)level, d[0], d[1], d[2], jobs + j0, (uint16_t*)out, P, o, clamp ? 1 : 0, minmax);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_linear_to_slot(const void* src, void* slot, int bpv, int S, hipStream_t stream) {
  const int64_t n = slot_block_elems(S);
  const unsigned blocks = (unsigned)((n + kThreads - 1) / kThreads);
  if (bpv == 1)
    linear_to_slot_kernel<uint8_t><<<blocks, kThreads, 0, stream>>>((const uint8_t*)src, (uint8_t*)slot, S, n);
  else
    linear_to_slot_kernel<uint16_t><<<blocks, kThreads, 0, stream>>>((const uint16_t*)src, (uint16_t*)slot,
                                                                      S, n);
  return hipGetLastError();
}

hipError_t launch_slot_to_linear(const void* slot, void* dst, int bpv, int S, hipStream_t stream) {
  const int64_t n = (int64_t)S * S * S;
  const unsigned blocks = (unsigned)((n + kThreads - 1) / kThreads);
  if (bpv == 1)
    slot_to_linear_kernel<uint8_t><<<blocks, kThreads, 0, stream>>>((const uint8_t*)slot, (uint8_t*)dst, S, n);
  else
    slot_to_linear_kernel<uint16_t><<<blocks, kThreads, 0, stream>>>((const uint16_t*)slot, (uint16_t*)dst,
                                                                      S, n);
  return hipGetLastError();
}

hipError_t launch_cell_ranges(const void* level, int bpv, const int64_t d[3], int C, uint32_t* out,
                              uint32_t* tmp_hi, hipStream_t stream) {
  const int ncx = (int)((d[0] + C - 1) / C), ncy = (int)((d[1] + C - 1) / C),
            ncz = (int)((d[2] + C - 1) / C);
  const int64_t cells = (int64_t)ncx * ncy * ncz;
  const unsigned fill_blocks = (unsigned)std::min<int64_t>((cells + kThreads - 1) / kThreads, 65536);
  fill_u32_kernel<<<fill_blocks, kThreads, 0, stream>>>(out, cells, 0xffffffffu);
  fill_u32_kernel<<<fill_blocks, kThreads, 0, stream>>>(tmp_hi, cells, 0u);
  // work-groups per cell: about 8 row steps per wave over the largest cell
  const int64_t w = C + 2, rows = w * w;
  const int64_t rpw = w < 64 ? 64 / w : 1;
  const int64_t per_group = (kThreads / 64) * rpw * 8;
  const unsigned gy = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (rows + per_group - 1) / per_group), 65535);
  // at most 2^20 work-groups along x (a launch counts each dimension's
  // work-items in 32 bits: one work-group per cell of a 4096^3 level would
  // be 2^32 of them); more cells loop inside the kernel
  const unsigned gx = (unsigned)std::min<int64_t>(cells, (int64_t)1 << 20);
  const dim3 grid(gx, gy, 1);
  if (bpv == 1)
    cell_range_rows_kernel<uint8_t><<<grid, kThreads, 0, stream>>>((const uint8_t*)level, d[0], d[1], d[2], C,
                                                                   ncx, ncy, cells, out, tmp_hi);
  else
    cell_range_rows_kernel<uint16_t><<<grid, kThreads, 0, stream>>>((const uint16_t*)level, d[0], d[1], d[2], C,
                                                                    ncx, ncy, cells, out, tmp_hi);
  pack_ranges_kernel<<<fill_blocks, kThreads, 0, stream>>>(out, tmp_hi, cells);
  return hipGetLastError();
}

hipError_t launch_slot_cell_ranges(const void* slot, int bpv, int S, int apron, int B, int C,
                                   const int brick[3], const int cells[3], uint32_t* out,
                                   uint32_t* tmp, hipStream_t stream) {
  const int per = B / C;
  const int n = per * per * per;
  // work-groups per cell: about 8 row steps per wave over the cell
  const int64_t w = C + 2, rows = w * w;
  const int64_t rpw = w < 64 ? 64 / w : 1;
  const int64_t per_group = (kThreads / 64) * rpw * 8;
  const unsigned gy = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (rows + per_group - 1) / per_group), 65535);
  const dim3 grid((unsigned)n, gy, 1);
  uint32_t* lo = tmp;
  uint32_t* hi = tmp + n;
  if (bpv == 1)
    slot_cell_range_kernel<uint8_t><<<grid, kThreads, 0, stream>>>(
        (const uint8_t*)slot, S, apron, B, C, brick[0], brick[1], brick[2], cells[0], cells[1], cells[2], lo, hi);
  else
    slot_cell_range_kernel<uint16_t><<<grid, kThreads, 0, stream>>>(
        (const uint16_t*)slot, S, apron, B, C, brick[0], brick[1], brick[2], cells[0], cells[1], cells[2], lo,
        hi);
  slot_cell_pack_kernel<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, stream>>>(
      lo, hi, per, brick[0], brick[1], brick[2], cells[0], cells[1], cells[2], out);
  return hipGetLastError();
}

hipError_t launch_slot_cell_tmp_init(uint32_t* tmp, int per, hipStream_t stream) {
  const int64_t n = (int64_t)per * per * per;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, 65536);
  fill_u32_kernel<<<blocks, kThreads, 0, stream>>>(tmp, n, 0xffffffffu);
  fill_u32_kernel<<<blocks, kThreads, 0, stream>>>(tmp + n, n, 0u);
  return hipGetLastError();
}

// Per-brick value histogram of a flat level: one work-group per brick of
// payload B (no apron), `bins` equal bins over [0, 2^(8 bpv)), LDS counts.
template <typename T>
__global__ __launch_bounds__(kThreads) void brick_hist_kernel(const T* __restrict__ level, int64_t nx,
                                                              int64_t ny, int64_t nz, int B, int gx,
                                                              int gy, int bins, int shift,
                                                              uint32_t* __restrict__ out) {
  __shared__ uint32_t h[64];
  for (int i = (int)threadIdx.x; i < bins; i += kThreads) h[i] = 0;
  __syncthreads();
  const int b = (int)blockIdx.x;
  const int bx = b % gx, by = (b / gx) % gy, bz = b / (gx * gy);
  const int64_t x0 = (int64_t)bx * B, y0 = (int64_t)by * B, z0 = (int64_t)bz * B;
  const int w = (int)min<int64_t>(B, nx - x0), hh = (int)min<int64_t>(B, ny - y0),
            d = (int)min<int64_t>(B, nz - z0);
  const int64_t cnt = (int64_t)w * hh * d;
  for (int64_t i = threadIdx.x; i < cnt; i += kThreads) {
    const int x = (int)(i % w), y = (int)((i / w) % hh), z = (int)(i / ((int64_t)w * hh));
    const uint32_t v = level[((z0 + z) * ny + (y0 + y)) * nx + (x0 + x)];
    atomicAdd(&h[v >> shift], 1u);
  }
  __syncthreads();
  for (int i = (int)threadIdx.x; i < bins; i += kThreads) out[(int64_t)b * bins + i] = h[i];
}

// Per LOD-0 brick region and level l >= 1: mean |v0(x) - v_l(x >> l)| over
// the region's voxels (each LOD-0 voxel against its level-l ancestor), the
// detail a region loses when it is rendered from level l.  One work-group
// per region; out[region * n_lod + l], out[.. + 0] = 0.
struct DetailLevels {
  const void* level[16];
  int64_t dims[16][3];
};

template <typename T>
__global__ __launch_bounds__(kThreads) void region_detail_kernel(const DetailLevels D, int n_lod, int B,
                                                                 int gx, int gy, float* __restrict__ out) {
  __shared__ float red[kThreads / 64][16];
  const int r = (int)blockIdx.x;
  const int bx = r % gx, by = (r / gx) % gy, bz = r / (gx * gy);
  const int64_t nx = D.dims[0][0], ny = D.dims[0][1], nz = D.dims[0][2];
  const int64_t x0 = (int64_t)bx * B, y0 = (int64_t)by * B, z0 = (int64_t)bz * B;
  const int w = (int)min<int64_t>(B, nx - x0), h = (int)min<int64_t>(B, ny - y0),
            d = (int)min<int64_t>(B, nz - z0);
  const int64_t cnt = (int64_t)w * h * d;
  float acc[16];
  for (int l = 0; l < 16; ++l) acc[l] = 0.f;
  const T* v0 = (const T*)D.level[0];
  for (int64_t i = threadIdx.x; i < cnt; i += kThreads) {
    const int64_t x = x0 + i % w, y = y0 + (i / w) % h, z = z0 + i / ((int64_t)w * h);
    const float a = (float)v0[(z * ny + y) * nx + x];
    for (int l = 1; l < n_lod; ++l) {
      const T* vl = (const T*)D.level[l];
      const int64_t lx = x >> l, ly = y >> l, lz = z >> l;
      acc[l] += fabsf(a - (float)vl[(lz * D.dims[l][1] + ly) * D.dims[l][0] + lx]);
    }
  }
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
  for (int l = 1; l < n_lod; ++l) {
    float v = acc[l];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][l] = v;
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)n_lod) {
    float s = 0.f;
    if (threadIdx.x > 0)
      for (int k = 0; k < kThreads / 64; ++k) s += red[k][threadIdx.x];
    out[(int64_t)r * n_lod + threadIdx.x] = s / (float)max<int64_t>(cnt, 1);
  }
}

hipError_t launch_region_detail(const void* const* levels, const int64_t (*dims)[3], int n_lod, int bpv,
                                int B, float* out, hipStream_t stream) {
  DetailLevels D{};
  for (int l = 0; l < n_lod && l < 16; ++l) {
    D.level[l] = levels[l];
    for (int i = 0; i < 3; ++i) D.dims[l][i] = dims[l][i];
  }
  const int gx = (int)((dims[0][0] + B - 1) / B), gy = (int)((dims[0][1] + B - 1) / B),
            gz = (int)((dims[0][2] + B - 1) / B);
  const unsigned blocks = (unsigned)(gx * gy * gz);
  if (bpv == 1)
    region_detail_kernel<uint8_t><<<blocks, kThreads, 0, stream>>>(D, n_lod, B, gx, gy, out);
  else
    region_detail_kernel<uint16_t><<<blocks, kThreads, 0, stream>>>(D, n_lod, B, gx, gy, out);
  return hipGetLastError();
}

hipError_t launch_brick_histograms(const void* level, int bpv, const int64_t d[3], int B, int bins,
                                   uint32_t* out, hipStream_t stream) {
  const int gx = (int)((d[0] + B - 1) / B), gy = (int)((d[1] + B - 1) / B),
            gz = (int)((d[2] + B - 1) / B);
  int shift = 0;
  while ((1 << (8 * bpv - shift)) > bins) ++shift;
  const unsigned blocks = (unsigned)(gx * gy * gz);
  if (bpv == 1)
    brick_hist_kernel<uint8_t><<<blocks, kThreads, 0, stream>>>((const uint8_t*)level, d[0], d[1], d[2],
                                                                 B, gx, gy, bins, shift, out);
  else
    brick_hist_kernel<uint16_t><<<blocks, kThreads, 0, stream>>>((const uint16_t*)level, d[0], d[1],
                                                                  d[2], B, gx, gy, bins, shift, out);
  return hipGetLastError();
}

hipError_t launch_histogram_1d(const void* vol, int bpv, int64_t n, unsigned long long* hist,
                               hipStream_t stream) {
  const int64_t nb = bpv == 1 ? 256 : 65536;
  hipError_t e = launch_fill_u64(hist, nb, 0ull, stream);
  if (e != hipSuccess) return e;
  int64_t blocks = (n + kThreads * 16 - 1) / (kThreads * 16);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  if (bpv == 1) {
    hist1d_u8_kernel<<<(unsigned)blocks, kThreads, 0, stream>>>((const uint8_t*)vol, n, hist);
    return hipGetLastError();
  }
  const size_t lds = 65536 * 2;
  if (lds_allowed(hist1d_u16_lds_kernel, lds)) {
    const int64_t groups = (n + kHistMaxVoxels - 1) / kHistMaxVoxels;
    hist1d_u16_lds_kernel<<<(unsigned)groups, kHistThreads, lds, stream>>>((const uint16_t*)vol, n, hist);
    if (hipGetLastError() == hipSuccess) return hipSuccess;
  }
  hist1d_u16_kernel<<<(unsigned)blocks, kThreads, 0, stream>>>((const uint16_t*)vol, n, hist);
  return hipGetLastError();
}

hipError_t launch_gradient_kmax(const void* vol, int bpv, const int64_t d[3],
                                unsigned long long* kmax, hipStream_t stream) {
  hipError_t e = launch_fill_u64(kmax, 1, 0ull, stream);
  if (e != hipSuccess) return e;
  const dim3 grid = row_grid(d[0], d[1] * d[2]);
  if (bpv == 1)
    gradient_kmax_kernel<uint8_t><<<grid, kThreads, 0, stream>>>((const uint8_t*)vol, d[0], d[1],
                                                                 d[2], kmax);
  else
    gradient_kmax_kernel<uint16_t><<<grid, kThreads, 0, stream>>>((const uint16_t*)vol, d[0],
                                                                  d[1], d[2], kmax);
  return hipGetLastError();
}

hipError_t launch_histogram_2d(const void* vol, int bpv, const int64_t d[3],
                               const unsigned long long* kmax, int n_value, int n_grad,
                               unsigned long long* hist, hipStream_t stream) {
  hipError_t e = launch_fill_u64(hist, (int64_t)n_value * n_grad, 0ull, stream);
  if (e != hipSuccess) return e;
  int shift = 0;
  while (((int64_t)n_value << shift) < ((int64_t)1 << (8 * bpv))) ++shift;
  // privatised in LDS when the bins fit 128 KiB and a row fits one group
  const int64_t nbins = (int64_t)n_value * n_grad;
  const size_t lds = (size_t)nbins * 2;
  if (nbins % 2 == 0 && nbins <= 65536 && d[0] <= kHistMaxVoxels) {
    const int rows_per_block = (int)(kHistMaxVoxels / d[0]);
    const int64_t groups = (d[1] * d[2] + rows_per_block - 1) / rows_per_block;
    bool ok = false;
    if (bpv == 1 && lds_allowed(hist2d_lds_kernel<uint8_t>, lds)) {
      hist2d_lds_kernel<uint8_t><<<(unsigned)groups, kHistThreads, lds, stream>>>(
          (const uint8_t*)vol, d[0], d[1], d[2], kmax, shift, n_grad, (int)nbins, rows_per_block, hist);
      ok = hipGetLastError() == hipSuccess;
    } else if (bpv == 2 && lds_allowed(hist2d_lds_kernel<uint16_t>, lds)) {
      hist2d_lds_kernel<uint16_t><<<(unsigned)groups, kHistThreads, lds, stream>>>(
          (const uint16_t*)vol, d[0], d[1], d[2], kmax, shift, n_grad, (int)nbins, rows_per_block, hist);
      ok = hipGetLastError() == hipSuccess;
    }
    if (ok) return hipSuccess;
  }
  const dim3 grid = row_grid(d[0], d[1] * d[2]);
  if (bpv == 1)
    hist2d_kernel<uint8_t><<<grid, kThreads, 0, stream>>>((const uint8_t*)vol, d[0], d[1], d[2],
                                                          kmax, shift, n_grad, hist);
  else
    hist2d_kernel<uint16_t><<<grid, kThreads, 0, stream>>>((const uint16_t*)vol, d[0], d[1], d[2],
                                                           kmax, shift, n_grad, hist);
  return hipGetLastError();
}

}  // namespace tvk
