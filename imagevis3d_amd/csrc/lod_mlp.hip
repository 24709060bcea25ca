// lod_mlp.hip -- the learned-LOD error predictor: a bf16 MLP with FP32
// accumulation on the MFMA units (v_mfma_f32_16x16x16_bf16).
//
// Network: 32 inputs -> 4 hidden layers of 64 (ReLU) -> 1 output, weights
// in bf16, biases and accumulation in FP32.  One wavefront evaluates 16
// rows: each layer is a (16 x K) x (K x 64) product made of 16x16x16 MFMA
// tiles; the ReLU'd activations go through LDS (as bf16) to become the A
// operand of the next layer.  The last layer (64 -> 1) is a dot product on
// the vector ALUs.
//
// MFMA 16x16x16 operand layout (64 lanes, 4 elements per lane):
//   A[i][k]: lane i + 16 * (k / 4), element k % 4
//   B[k][j]: lane j + 16 * (k / 4), element k % 4
//   D[i][j]: lane j + 16 * (i / 4), element i % 4
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace tvk {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 32, kHid = 64, kLayers = 4;
constexpr int kRows = 16;          // rows per wave
constexpr int kWavesPerBlock = 4;

// round-to-nearest-even float -> bf16 bits
__device__ inline uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Weights (bf16 bits), row-major [in][out] per layer: W0 32x64, W1..W3
// 64x64, w_out 64; biases (f32): b0..b3 64 each, b_out 1.
struct MlpArgs {
  const uint16_t* w;   // 32*64 + 3*64*64 + 64
  const float* b;      // 4*64 + 1
  const float* x;      // n x 32 inputs
  float* y;            // n outputs
  int n;
};

__global__ __launch_bounds__(64 * kWavesPerBlock) void lod_mlp_kernel(const MlpArgs a) {
  __shared__ uint16_t s_act[kWavesPerBlock][kRows][kHid];
  const int lane = (int)(threadIdx.x & 63);
  const int wave = (int)(threadIdx.x >> 6);
  const int row0 = ((int)blockIdx.x * kWavesPerBlock + wave) * kRows;
  uint16_t(*act)[kHid] = s_act[wave];
  // stage this wave's 16 input rows (bf16) into LDS; rows past n are zero
  for (int e = lane; e < kRows * kIn; e += 64) {
    const int r = e / kIn, c = e % kIn;
    const int gr = row0 + r;
    act[r][c] = to_bf16(gr < a.n ? a.x[(int64_t)gr * kIn + c] : 0.f);
  }
  __syncthreads();
  const uint16_t* W = a.w;
  int K = kIn;
  const int ai = lane & 15, ak = (lane >> 4) * 4;  // A: row ai, k offset ak
  const int bj = lane & 15, bk = (lane >> 4) * 4;  // B: column bj, k offset bk
  for (int layer = 0; layer < kLayers; ++layer) {
    f32x4 acc[kHid / 16];
#pragma unroll
    for (int t = 0; t < kHid / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
      bf16x4 A;
#pragma unroll
      for (int e = 0; e < 4; ++e) A[e] = (short)act[ai][k0 + ak + e];
#pragma unroll
      for (int t = 0; t < kHid / 16; ++t) {
        bf16x4 Bv;
#pragma unroll
        for (int e = 0; e < 4; ++e) Bv[e] = (short)W[(k0 + bk + e) * kHid + t * 16 + bj];
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, Bv, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();  // every lane has read its A operands of this layer
    const float* bias = a.b + layer * kHid;
#pragma unroll
    for (int t = 0; t < kHid / 16; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = (lane >> 4) * 4 + e, j = t * 16 + (lane & 15);
        act[i][j] = to_bf16(fmaxf(acc[t][e] + bias[j], 0.f));
      }
    __syncthreads();
    W += K * kHid;
    K = kHid;
  }
  // output layer on the vector ALUs: lanes 0..15 take one row each
  if (lane < kRows) {
    const int gr = row0 + lane;
    float s = a.b[kLayers * kHid];
    for (int j = 0; j < kHid; ++j)
      s = fmaf(__uint_as_float((uint32_t)act[lane][j] << 16), __uint_as_float((uint32_t)W[j] << 16), s);
    if (gr < a.n) a.y[gr] = s;
  }
}

}  // namespace

hipError_t launch_lod_mlp(const uint16_t* w, const float* b, const float* x, float* y, int n,
                          hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  MlpArgs a{w, b, x, y, n};
  const int rows_per_block = kRows * kWavesPerBlock;
  const unsigned blocks = (unsigned)((n + rows_per_block - 1) / rows_per_block);
  lod_mlp_kernel<<<blocks, 64 * kWavesPerBlock, 0, stream>>>(a);
  return hipGetLastError();
}

}  // namespace tvk
