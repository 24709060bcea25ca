/*
 * baseline_dispatch.c -- the CPU baseline's entry point: the AVX-512 build
 * of baseline.c when the host's cores run AVX-512 (F, BW, VL), else the
 * AVX2 + FMA build.  TEST INFRASTRUCTURE ONLY.
 */
#include "oracle.h"

uint64_t baseline_render_1d_f32_avx2(const oracle_render_params* p, uint8_t* out8, int32_t n_threads);
uint64_t baseline_render_1d_f32_avx512(const oracle_render_params* p, uint8_t* out8, int32_t n_threads);

int baseline_uses_avx512(void) {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
         __builtin_cpu_supports("avx512vl");
}

uint64_t baseline_render_1d_f32(const oracle_render_params* p, uint8_t* out8, int32_t n_threads) {
  return baseline_uses_avx512() ? baseline_render_1d_f32_avx512(p, out8, n_threads)
                                : baseline_render_1d_f32_avx2(p, out8, n_threads);
}
