// fsk_microbench.hip - issue-rate microbenchmarks for the roofline of the pair
// kernels: FP32 MFMA 16x16x4, BF16 MFMA 16x16x32, v_exp_f32, v_fma_f32, an
// MFMA + exp mix that shows whether the two pipes overlap, and a plain HBM
// copy (known bytes per launch: the calibration point for the FETCH_SIZE /
// WRITE_SIZE counters and the measured HBM bandwidth).
//
// Every kernel runs independent dependency chains per lane in a loop and
// writes one value per thread to the caller's buffer so nothing is removed
// by the compiler. Rates are whole-chip operations per second.
#include "fsk_common.h"

namespace fsk {

typedef __bf16 bf16x8_mb __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(256) void mb_mfma_f32(float *out, int iters) {
  const float a = 1.0f + threadIdx.x * 1e-7f, b = 0.999f;
  f4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
  for (int i = 0; i < iters; ++i) {
    c0 = mfma_f32(a, b, c0);
    c1 = mfma_f32(a, b, c1);
    c2 = mfma_f32(a, b, c2);
    c3 = mfma_f32(a, b, c3);
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = c0[0] + c1[1] + c2[2] + c3[3];
}

__global__ __launch_bounds__(256) void mb_mfma_bf16(float *out, int iters) {
  bf16x8_mb a, b;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = (__bf16)(1.0f + k * 0.01f);
    b[k] = (__bf16)(0.5f);
  }
  f4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
  for (int i = 0; i < iters; ++i) {
    c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c1, 0, 0, 0);
    c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c2, 0, 0, 0);
    c3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c3, 0, 0, 0);
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = c0[0] + c1[1] + c2[2] + c3[3];
}

// The exp and FMA rates run 32 independent chains per lane with the trip
// unrolled four times (128 instructions per trip): short loop bodies (8
// instructions a trip) measured up to 25 % below this on the same GPU, an
// instruction-fetch effect that the pair kernels' long loop bodies do not see.
constexpr int kMbChains = 32;
constexpr int kMbOpsPerIter = kMbChains * 4;

__global__ __launch_bounds__(256) void mb_exp(float *out, int iters) {
  float v[kMbChains];
#pragma unroll
  for (int k = 0; k < kMbChains; ++k) v[k] = (threadIdx.x + k) * 1e-3f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < kMbChains; ++k) v[k] = fexp2(-v[k]);
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kMbChains; ++k) s += v[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// FMA operands in registers, accumulating in place: the 4-byte
// v_fmac_f32_e32 that the pair kernels issue.
__device__ __forceinline__ float mb_fa() { return 0.999f + threadIdx.x * 1e-9f; }
__device__ __forceinline__ float mb_fb() { return 1e-3f + threadIdx.x * 1e-12f; }

__global__ __launch_bounds__(256) void mb_fma(float *out, int iters) {
  float v[kMbChains];
  const float a = mb_fa(), b = mb_fb();
#pragma unroll
  for (int k = 0; k < kMbChains; ++k) v[k] = (threadIdx.x + k) * 1e-3f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < kMbChains; ++k) v[k] = fmaf(a, b, v[k]);
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kMbChains; ++k) s += v[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 4 FP32 MFMAs (4 x 2048 flop) and 16 exps per lane per iteration: the
// pair-kernel ratio (one 16x16 tile = 256 pairs = 4 exps per lane per 4 MFMAs
// for the cross term, times 4 row tiles).
__global__ __launch_bounds__(256) void mb_mix(float *out, int iters) {
  const float a = 1.0f + threadIdx.x * 1e-7f, b = 0.999f;
  f4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
  float v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = (threadIdx.x + k) * 1e-3f;
#pragma unroll 4
  for (int i = 0; i < iters; ++i) {
    c0 = mfma_f32(a, b, c0);
    c1 = mfma_f32(a, b, c1);
    c2 = mfma_f32(a, b, c2);
    c3 = mfma_f32(a, b, c3);
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = fexp2(-v[k]);
  }
  float s = c0[0] + c1[1] + c2[2] + c3[3];
#pragma unroll
  for (int k = 0; k < 16; ++k) s += v[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// As mb_mix with 16 FMAs instead of 16 exps per lane per iteration: whether
// plain VALU work overlaps the MFMA pipe where transcendentals may not.
__global__ __launch_bounds__(256) void mb_mix_fma(float *out, int iters) {
  const float a = 1.0f + threadIdx.x * 1e-7f, b = 0.999f;
  f4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
  float v[16];
  const float fa = mb_fa(), fb = mb_fb();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = (threadIdx.x + k) * 1e-3f;
#pragma unroll 4
  for (int i = 0; i < iters; ++i) {
    c0 = mfma_f32(a, b, c0);
    c1 = mfma_f32(a, b, c1);
    c2 = mfma_f32(a, b, c2);
    c3 = mfma_f32(a, b, c3);
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = fmaf(fa, fb, v[k]);
  }
  float s = c0[0] + c1[1] + c2[2] + c3[3];
#pragma unroll
  for (int k = 0; k < 16; ++k) s += v[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 4 BF16 MFMAs 16x16x32 (4 x 16384 flop) with 16 exps (EXP) or 16 FMAs per
// lane per iteration: whether the 16-bit MFMA issue overlaps the VALU, as the
// 16-bit score kernels would need for their MFMA time to hide.
template <bool EXP>
__global__ __launch_bounds__(256) void mb_mix_bf16(float *out, int iters) {
  bf16x8_mb a, b;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = (__bf16)(1.0f + k * 0.01f + threadIdx.x * 1e-6f);
    b[k] = (__bf16)(0.5f);
  }
  f4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
  float v[16];
  const float fa = mb_fa(), fb = mb_fb();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = (threadIdx.x + k) * 1e-3f;
#pragma unroll 4
  for (int i = 0; i < iters; ++i) {
    c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c1, 0, 0, 0);
    c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c2, 0, 0, 0);
    c3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c3, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = EXP ? fexp2(-v[k]) : fmaf(fa, fb, v[k]);
  }
  float s = c0[0] + c1[1] + c2[2] + c3[3];
#pragma unroll
  for (int k = 0; k < 16; ++k) s += v[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 8 exps and 8 FMAs per lane per iteration on independent chains: whether the
// transcendental issue overlaps plain VALU work.
__global__ __launch_bounds__(256) void mb_mix_exp_fma(float *out, int iters) {
  float v[8], f[8];
  const float fa = mb_fa(), fb = mb_fb();
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = (threadIdx.x + k) * 1e-3f;
    f[k] = (threadIdx.x + k) * 2e-3f;
  }
#pragma unroll 4
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      v[k] = fexp2(-v[k]);
      f[k] = fmaf(fa, fb, f[k]);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += v[k] + f[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// dst[i] = src[i] over n4 float4s; grid-stride, 16-byte loads and stores.
__global__ __launch_bounds__(256) void mb_copy(const f4 *__restrict__ src, f4 *__restrict__ dst,
                                               int64_t n4) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += stride)
    dst[i] = src[i];
}

}  // namespace fsk

using namespace fsk;

// which: 0 = FP32 MFMA (flop/s), 1 = BF16 MFMA (flop/s), 2 = v_exp_f32 (exp/s),
// 3 = v_fma_f32 (fma/s), 4 = mix (returns seconds per iteration-block, see header),
// 5 = copy of the first half of `out` to its second half, `iters` launches
// (bytes read + written per second), 6 = mix with FMAs (seconds, as 4),
// 7 / 8 = 4 BF16 MFMAs with 16 exps / 16 FMAs per iteration (seconds),
// 9 = 8 exps with 8 FMAs per iteration (seconds).
extern "C" int fsk_microbench(int which, int blocks, int iters, float *out, double *result,
                              void *stream) {
  if (which < 0 || which > 9 || blocks <= 0 || iters <= 0 || !out || !result ||
      (which == 5 && (reinterpret_cast<uintptr_t>(out) & 15))) {
    set_error("fsk_microbench: bad arguments");
    return FSK_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    set_error("fsk_microbench: event creation failed");
    return FSK_ELAUNCH;
  }
  (void)hipEventRecord(e0, s);
  switch (which) {
    case 0: hipLaunchKernelGGL(mb_mfma_f32, dim3(blocks), dim3(256), 0, s, out, iters); break;
    case 1: hipLaunchKernelGGL(mb_mfma_bf16, dim3(blocks), dim3(256), 0, s, out, iters); break;
    case 2: hipLaunchKernelGGL(mb_exp, dim3(blocks), dim3(256), 0, s, out, iters); break;
    case 3: hipLaunchKernelGGL(mb_fma, dim3(blocks), dim3(256), 0, s, out, iters); break;
    case 4: hipLaunchKernelGGL(mb_mix, dim3(blocks), dim3(256), 0, s, out, iters); break;
    case 6: hipLaunchKernelGGL(mb_mix_fma, dim3(blocks), dim3(256), 0, s, out, iters); break;
    case 7: hipLaunchKernelGGL(mb_mix_bf16<true>, dim3(blocks), dim3(256), 0, s, out, iters); break;
    case 8: hipLaunchKernelGGL(mb_mix_bf16<false>, dim3(blocks), dim3(256), 0, s, out, iters); break;
    case 9: hipLaunchKernelGGL(mb_mix_exp_fma, dim3(blocks), dim3(256), 0, s, out, iters); break;
    default: {
      // blocks * 256 floats = blocks * 64 float4; each half is blocks * 32 float4
      const int64_t n4 = int64_t(blocks) * 32;
      const f4 *src = reinterpret_cast<const f4 *>(out);
      f4 *dst = reinterpret_cast<f4 *>(out) + n4;
      for (int i = 0; i < iters; ++i)
        hipLaunchKernelGGL(mb_copy, dim3(4096), dim3(256), 0, s, src, dst, n4);
    } break;
  }
  (void)hipEventRecord(e1, s);
  int st = launch_status("fsk_microbench");
  if (st == FSK_OK && hipEventSynchronize(e1) != hipSuccess) st = FSK_ELAUNCH;
  float ms = 0.f;
  if (st == FSK_OK) (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != FSK_OK) return st;
  const double waves = double(blocks) * 4.0;
  const double sec = ms * 1e-3;
  switch (which) {
    case 0: *result = waves * iters * 4.0 * 2048.0 / sec; break;
    case 1: *result = waves * iters * 4.0 * 16384.0 / sec; break;
    case 2: *result = waves * 64.0 * iters * kMbOpsPerIter / sec; break;
    case 3: *result = waves * 64.0 * iters * kMbOpsPerIter / sec; break;
    case 5: *result = double(blocks) * 1024.0 * iters / sec; break;  // 2 x half-buffer bytes
    default: *result = sec; break;
  }
  return FSK_OK;
}
