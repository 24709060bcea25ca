// lod_mlp.hip -- the learned-LOD error predictor: a bf16 MLP with FP32
// accumulation on the MFMA units (v_mfma_f32_16x16x16_bf16).
//
// Network: 32 inputs -> 4 hidden layers of 64 (ReLU) -> 1 output, weights
// in bf16, biases and accumulation in FP32.  A work-group of 4 waves stages
// all weights in LDS once, transposed to [out][in] so that a lane's MFMA B
// fragment (4 consecutive inputs of one output column) is one 8-byte LDS
// read; each wave then evaluates 64 rows as 4 row tiles of 16, so one B
// fragment feeds 4 MFMAs.  Each layer is (64 x K) x (K x 64) in 16x16x16
// tiles; the ReLU'd activations go back to the wave's LDS rows as bf16 to
// become the next layer's A operand.  The 64 -> 1 head is a dot product
// on the vector ALUs, one row per lane.
//
// MFMA 16x16x16 operand layout (64 lanes, 4 elements per lane):
//   A[i][k]: lane i + 16 * (k / 4), element k % 4
//   B[k][j]: lane j + 16 * (k / 4), element k % 4
//   D[i][j]: lane j + 16 * (i / 4), element i % 4
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "kernels.hpp"

namespace tvk {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 32, kHid = 64, kLayers = 4;
constexpr int kRowTiles = 4;                   // row tiles of 16 per wave
constexpr int kRowsPerWave = 16 * kRowTiles;   // 64
constexpr int kWavesPerBlock = 4;
constexpr int kPad = 4;                        // LDS row padding (bf16 elements)
constexpr int kActStride = kHid + kPad;        // 68: 136-byte rows, 8-byte aligned
// transposed weights: layer 0 is 64 rows of kIn + pad, layers 1..3 64 rows of kHid + pad
constexpr int kW0Stride = kIn + kPad, kWStride = kHid + kPad;
constexpr int kWtElems = kHid * kW0Stride + (kLayers - 1) * kHid * kWStride;

// round-to-nearest-even float -> bf16 bits
__device__ inline uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Weights in the kernel's LDS image (lod_mlp_pack): per layer the [out][in]
// transpose with rows padded by kPad, then w_out; biases (f32): b0..b3 64
// each, b_out 1.
struct MlpArgs {
  const uint16_t* w;   // kPackedElems
  const float* b;      // 4*64 + 1
  const float* x;      // n x 32 inputs
  float* y;            // n outputs
  int n;
};

constexpr int kPackedElems = kWtElems + kHid;  // transposed layers, then w_out

__global__ __launch_bounds__(64 * kWavesPerBlock) void lod_mlp_kernel(const MlpArgs a) {
  __shared__ __attribute__((aligned(16))) uint16_t s_w[kPackedElems];
  __shared__ __attribute__((aligned(16))) uint16_t s_act[kWavesPerBlock][kRowsPerWave][kActStride];
  __shared__ float s_b[kLayers * kHid + 1];
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  // the packed weights (already in this LDS layout) and biases, once per block
  {
    const uint4* src = (const uint4*)a.w;
    uint4* dst = (uint4*)s_w;
    for (int e = tid; e < kPackedElems * 2 / 16; e += 64 * kWavesPerBlock) dst[e] = src[e];
    for (int e = tid; e < kLayers * kHid + 1; e += 64 * kWavesPerBlock) s_b[e] = a.b[e];
  }
  const uint16_t* s_wo = s_w + kWtElems;
  uint16_t(*act)[kActStride] = s_act[wave];
  const int li = lane & 15, lk = (lane >> 4) * 4;  // A row / B column, k offset
  const int n_groups = (a.n + kRowsPerWave * kWavesPerBlock - 1) / (kRowsPerWave * kWavesPerBlock);
  for (int grp = (int)blockIdx.x; grp < n_groups; grp += (int)gridDim.x) {
    // this wave's 64 input rows as bf16; rows past n are zero
    const int row0 = (grp * kWavesPerBlock + wave) * kRowsPerWave;
    __syncthreads();  // the previous group's output reads are done
    {
      // one row per lane: its 32 inputs as 8 x 16-byte loads issued together
      const int gr = row0 + lane;
      float4 v[kIn / 4];
#pragma unroll
      for (int q = 0; q < kIn / 4; ++q) v[q] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gr < a.n) {  // rows past n stay zero and are never read from x
        const float4* xr = (const float4*)(a.x + (int64_t)gr * kIn);
#pragma unroll
        for (int q = 0; q < kIn / 4; ++q) v[q] = xr[q];
      }
#pragma unroll
      for (int q = 0; q < kIn / 4; ++q) {
        bf16x4 b;
        b[0] = (short)to_bf16(v[q].x);
        b[1] = (short)to_bf16(v[q].y);
        b[2] = (short)to_bf16(v[q].z);
        b[3] = (short)to_bf16(v[q].w);
        *(bf16x4*)&act[lane][4 * q] = b;
      }
    }
    __syncthreads();
    int woff = 0;
    for (int layer = 0; layer < kLayers; ++layer) {
      const int K = layer == 0 ? kIn : kHid, stride = layer == 0 ? kW0Stride : kWStride;
      const uint16_t* Wt = s_w + woff;
      f32x4 acc[kRowTiles][kHid / 16];
#pragma unroll
      for (int r = 0; r < kRowTiles; ++r)
#pragma unroll
        for (int t = 0; t < kHid / 16; ++t) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
      for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 B[kHid / 16];
#pragma unroll
        for (int t = 0; t < kHid / 16; ++t) B[t] = *(const bf16x4*)&Wt[(t * 16 + li) * stride + k0 + lk];
#pragma unroll
        for (int r = 0; r < kRowTiles; ++r) {
          const bf16x4 A = *(const bf16x4*)&act[r * 16 + li][k0 + lk];
#pragma unroll
          for (int t = 0; t < kHid / 16; ++t)
            acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, B[t], acc[r][t], 0, 0, 0);
        }
      }
      __syncthreads();  // every lane has read its A operands of this layer
      const float* bias = s_b + layer * kHid;
#pragma unroll
      for (int r = 0; r < kRowTiles; ++r)
#pragma unroll
        for (int t = 0; t < kHid / 16; ++t)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int i = r * 16 + (lane >> 4) * 4 + e, j = t * 16 + (lane & 15);
            act[i][j] = to_bf16(fmaxf(acc[r][t][e] + bias[j], 0.f));
          }
      __syncthreads();
      woff += kHid * stride;
    }
    // output layer on the vector ALUs: one row per lane
    const int gr = row0 + lane;
    float s = s_b[kLayers * kHid];
    for (int j = 0; j < kHid; j += 4) {
      const bf16x4 v = *(const bf16x4*)&act[lane][j];
#pragma unroll
      for (int e = 0; e < 4; ++e) s = fmaf(from_bf16((uint16_t)v[e]), from_bf16(s_wo[j + e]), s);
    }
    if (gr < a.n) a.y[gr] = s;
  }
}

}  // namespace

size_t lod_mlp_packed_elems() { return (size_t)kPackedElems; }

void lod_mlp_pack(const uint16_t* w, uint16_t* packed) {
  for (int i = 0; i < kPackedElems; ++i) packed[i] = 0;
  int src = 0, dst = 0;
  for (int layer = 0; layer < kLayers; ++layer) {
    const int K = layer == 0 ? kIn : kHid, stride = layer == 0 ? kW0Stride : kWStride;
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < kHid; ++j) packed[dst + j * stride + k] = w[src + k * kHid + j];
    src += K * kHid;
    dst += kHid * stride;
  }
  for (int j = 0; j < kHid; ++j) packed[kWtElems + j] = w[src + j];
}

hipError_t launch_lod_mlp(const uint16_t* w_packed, const float* b, const float* x, float* y, int n,
                          hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  MlpArgs a{w_packed, b, x, y, n};
  const int rows_per_block = kRowsPerWave * kWavesPerBlock;
  const int groups = (n + rows_per_block - 1) / rows_per_block;
  // at most two work-groups per CU (LDS), each looping over an equal share
  // of the row groups, so the weight staging is paid once per work-group
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
  }
  const int per_block = (groups + 2 * cus - 1) / (2 * cus);
  const unsigned blocks = (unsigned)((groups + per_block - 1) / per_block);
  lod_mlp_kernel<<<blocks, 64 * kWavesPerBlock, 0, stream>>>(a);
  return hipGetLastError();
}

}  // namespace tvk
