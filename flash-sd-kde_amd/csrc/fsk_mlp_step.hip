// fsk_mlp_step.hip - the row-local part of one training step of the
// 16-H-H-H-16 SiLU surrogate (H = 256) in ONE launch.
//
// Everything in a training step except the weight gradients depends on the
// batch rows one at a time: gathering a row, the four forward layers, the
// MSE gradient and the data-gradient chain back to layer 1 need only that
// row and the weights. So a workgroup takes 32 batch rows and runs them
// through all seven GEMMs (4 forward, 3 data-gradient), reading only the BF16
// weights (from L2) between stages. Eight waves split each layer's 256
// columns (32 each, two per SIMD to hide the weight loads' latency); the BF16
// activations pass between layers through LDS (34 KB) and each wave keeps
// the pre-activations of its own columns in registers, which is where the
// data-gradient GEMM that reads them puts the same elements. It writes what the weight gradients need
// - the gathered inputs (FP32), the activations h1..h3 as the BF16 values
// the next layer multiplies (half the bytes of FP32; the weight-gradient GEMM
// rounds its inputs to BF16 anyway) and the output gradients dz1..dz4 (FP32:
// the bias gradients sum them) - and the step then finishes with one multi-layer
// weight-gradient launch and the Adam launch: three launches per step where
// the layer-by-layer form needs ten, each of which costs several
// microseconds of launch and drain on its own.
//
// The arithmetic is that of the layer kernels (fsk_mlp.hip): BF16 operands
// (activations rounded as they are stored), MFMA 16x16x32 BF16 accumulating
// over K in the same order, FP32 bias / SiLU / silu' epilogues, so a step
// gives the same values as the layer-by-layer launches bit for bit.
#include <type_traits>

#include "fsk_common.h"
#include "fsk_lowp.h"

namespace fsk {

// batch rows per workgroup (32: every CU gets one workgroup of a batch of
// 8192 rows, and each reads the weights once per stage)
constexpr int RS_ROWS = 32;
constexpr int RS_WAVES = 8;              // waves per workgroup
constexpr int RS_NB = 256 / (16 * RS_WAVES);  // 16-column tiles per wave in a hidden layer
constexpr int RS_RX = RS_ROWS / 16;      // 16-row MFMA tiles per workgroup
constexpr int RS_H = 256;                // hidden width
constexpr int RS_D = 16;                 // input / output width
constexpr int RS_ACT = RS_H + 8;         // u16 per LDS row of a BF16 activation
constexpr int RS_SMALL = RS_D + 8;       // u16 per LDS row of a 16-wide BF16 block

__device__ __forceinline__ float rs_sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.f + __expf(-z));
}
__device__ __forceinline__ float rs_silu(float z) { return z * rs_sigmoid(z); }
__device__ __forceinline__ float rs_silu_grad(float z) {
  const float s = rs_sigmoid(z);
  return s * (1.f + z * (1.f - s));
}

// acc (this wave's part of the 32 x N block) = A (32 x K, BF16, LDS, row
// stride lda) . W^T with W (N x K) BF16 row-major in global memory. For
// N = 256 wave w owns columns [16 NB w, 16 NB (w+1)), NB = RS_NB; for N = 16
// wave 0 computes the block alone. Lane (g, c) of tile (x, y) holds
// rows 16x + 4g + r, column col0 + 16y + c.
template <int K, int N>
__device__ __forceinline__ bool rs_gemm(const unsigned short *A, int lda,
                                        const unsigned short *__restrict__ W, int wave, int g,
                                        int c, f4 (&acc)[RS_RX][N == RS_H ? RS_NB : 1], int &col0) {
  constexpr int NB = N == RS_H ? RS_NB : 1;
  constexpr int KS = (K + 31) / 32;
  col0 = wave * NB * 16;
#pragma unroll
  for (int x = 0; x < RS_RX; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};
  if (col0 >= N) return false;  // wave-uniform
  const u16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int k = 32 * ks + 8 * g;
    u16x8 a[RS_RX], b[NB];
#pragma unroll
    for (int x = 0; x < RS_RX; ++x)
      a[x] = k < K ? *reinterpret_cast<const u16x8 *>(A + (16 * x + c) * lda + k) : zero8;
#pragma unroll
    for (int y = 0; y < NB; ++y)
      b[y] = k < K ? *reinterpret_cast<const u16x8 *>(W + (col0 + 16 * y + c) * K + k) : zero8;
#pragma unroll
    for (int x = 0; x < RS_RX; ++x)
#pragma unroll
      for (int y = 0; y < NB; ++y)
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf16x8, a[x]), __builtin_bit_cast(bf16x8, b[y]), acc[x][y], 0, 0,
            0);
  }
  return true;
}

struct RowsStepArgs {
  const float *x, *t;  // normalised inputs / targets, n x 16
  int64_t n;
  const int *idx;      // steps x batch row indices
  int64_t batch, steps;
  const float *table;  // steps x 3 schedule
  const int64_t *counter;
  float *sched;
  const unsigned short *w[4];   // forward operands, N x K
  const unsigned short *wt[4];  // data-gradient operands, K x N
  const float *bias[4];
  float *xb;      // batch x 16
  unsigned short *h[3];  // batch x H, layers 1..3, BF16
  float *dz[4];   // batch x H (layers 1..3), batch x 16 (layer 4)
  float mse_scale;
  double *loss;
};

__global__ __launch_bounds__(64 * RS_WAVES) void mlp_rows_step_kernel(RowsStepArgs a) {
  __shared__ __align__(16) unsigned short act[2][RS_ROWS * RS_ACT];  // BF16 ping-pong
  __shared__ __align__(16) unsigned short small[RS_ROWS * RS_SMALL];  // x0, then dz4 (BF16)
  __shared__ __align__(16) float ts[RS_ROWS * RS_D];                  // targets
  __shared__ double sl[RS_WAVES];
  // z1..z3 of this wave's columns, in the accumulator layout
  f4 zr[3][RS_RX][RS_NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t r0 = int64_t(blockIdx.x) * RS_ROWS;
  int64_t step = *a.counter;
  step = step < a.steps ? step : a.steps - 1;
  if (blockIdx.x == 0 && tid < 3) a.sched[tid] = a.table[step * 3 + tid];

  // ---- gather: the first 4 RS_ROWS threads the inputs, the next 4 RS_ROWS the
  // targets (one float4 each)
  {
    constexpr int PER = RS_ROWS * 4;  // one float4 per thread and half
    const int half = tid / PER, e = tid % PER, row = e >> 2, q = e & 3;
    const int64_t br = r0 + row < a.batch ? r0 + row : a.batch - 1;
    const int64_t j = a.idx[step * a.batch + br];
    if (half == 0) {
      const f4 v = ld4(a.x + j * RS_D + 4 * q);
      if (r0 + row < a.batch) *reinterpret_cast<f4 *>(a.xb + (r0 + row) * RS_D + 4 * q) = v;
      u16x4 w;
#pragma unroll
      for (int i = 0; i < 4; ++i) w[i] = to16<FSK_PREC_BF16>(v[i]);
      *reinterpret_cast<u16x4 *>(small + row * RS_SMALL + 4 * q) = w;
    } else if (half == 1) {
      *reinterpret_cast<f4 *>(ts + row * RS_D + 4 * q) = ld4(a.t + j * RS_D + 4 * q);
    }
  }
  __syncthreads();

  int col0;
  // ---- forward, hidden layers: z_l = A W_l^T + b_l, h_l = silu(z_l)
  auto hidden = [&](const unsigned short *A, int lda, int l, unsigned short *out, auto kdim) {
    constexpr int K = decltype(kdim)::value;
    f4 acc[RS_RX][RS_NB];
    if (rs_gemm<K, RS_H>(A, lda, a.w[l], wave, g, c, acc, col0)) {
#pragma unroll
      for (int y = 0; y < RS_NB; ++y) {
        const int nn = col0 + 16 * y + c;
        const float bv = a.bias[l][nn];
#pragma unroll
        for (int x = 0; x < RS_RX; ++x)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = 16 * x + 4 * g + r;
            const float z = acc[x][y][r] + bv;
            const unsigned short hv = to16<FSK_PREC_BF16>(rs_silu(z));
            zr[l][x][y][r] = z;
            out[row * RS_ACT + nn] = hv;
            if (r0 + row < a.batch) a.h[l][(r0 + row) * RS_H + nn] = hv;
          }
      }
    }
    __syncthreads();
  };
  hidden(small, RS_SMALL, 0, act[0], std::integral_constant<int, RS_D>{});
  hidden(act[0], RS_ACT, 1, act[1], std::integral_constant<int, RS_H>{});
  hidden(act[1], RS_ACT, 2, act[0], std::integral_constant<int, RS_H>{});

  // ---- output layer with the MSE gradient: d = z - t, dz4 = scale d
  {
    f4 acc[RS_RX][1];
    double lsum = 0.0;
    if (rs_gemm<RS_H, RS_D>(act[0], RS_ACT, a.w[3], wave, g, c, acc, col0)) {
      const int nn = c;
      const float bv = a.bias[3][nn];
#pragma unroll
      for (int x = 0; x < RS_RX; ++x)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * x + 4 * g + r;
          const float d = acc[x][0][r] + bv - ts[row * RS_D + nn];
          const float dzv = a.mse_scale * d;
          small[row * RS_SMALL + nn] = to16<FSK_PREC_BF16>(dzv);
          if (r0 + row < a.batch) {
            a.dz[3][(r0 + row) * RS_D + nn] = dzv;
            lsum += double(d) * d;
          }
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) lsum += __shfl_xor(lsum, o);
    if (lane == 0) sl[wave] = lsum;
    __syncthreads();
    // only wave 0 computes the 16-wide output layer; the others add 0
    if (tid == 0 && a.loss) atomicAdd(a.loss, sl[0]);
  }

  // ---- data gradients: dz_{l-1} = (dz_l W_l) * silu'(z_{l-1}), l = 4, 3, 2
  auto backward = [&](const unsigned short *A, int lda, int l, unsigned short *out, auto kdim) {
    constexpr int K = decltype(kdim)::value;  // width of dz_l (the reduction)
    f4 acc[RS_RX][RS_NB];
    if (rs_gemm<K, RS_H>(A, lda, a.wt[l], wave, g, c, acc, col0)) {
#pragma unroll
      for (int y = 0; y < RS_NB; ++y) {
        const int nn = col0 + 16 * y + c;
#pragma unroll
        for (int x = 0; x < RS_RX; ++x)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = 16 * x + 4 * g + r;
            const float v = acc[x][y][r] * rs_silu_grad(zr[l - 1][x][y][r]);
            if (out) out[row * RS_ACT + nn] = to16<FSK_PREC_BF16>(v);
            if (r0 + row < a.batch) a.dz[l - 1][(r0 + row) * RS_H + nn] = v;
          }
      }
    }
    __syncthreads();
  };
  backward(small, RS_SMALL, 3, act[1], std::integral_constant<int, RS_D>{});
  backward(act[1], RS_ACT, 2, act[0], std::integral_constant<int, RS_H>{});
  backward(act[0], RS_ACT, 1, nullptr, std::integral_constant<int, RS_H>{});
}

}  // namespace fsk

using namespace fsk;

extern "C" int fsk_mlp_rows_step(const fsk_mlp_rows_step_args *p, void *stream) {
  const char *what = "fsk_mlp_rows_step";
  if (!p || !p->x || !p->t || p->n <= 0 || !p->idx || p->batch <= 0 || p->steps <= 0 ||
      !p->table || !p->counter || !p->sched || !p->xb || p->hidden != RS_H ||
      p->in_dim != RS_D || p->out_dim != RS_D) {
    set_error("%s: needs the 16-%d-%d-%d-16 network and non-null buffers", what, RS_H, RS_H,
              RS_H);
    return FSK_EINVAL;
  }
  RowsStepArgs a;
  a.x = p->x;
  a.t = p->t;
  a.n = p->n;
  a.idx = p->idx;
  a.batch = p->batch;
  a.steps = p->steps;
  a.table = p->table;
  a.counter = p->counter;
  a.sched = p->sched;
  for (int l = 0; l < 4; ++l) {
    if (!p->w_bf16[l] || !p->wt_bf16[l] || !p->bias[l] || !p->dz[l] || (l < 3 && !p->h[l]) ||
        (reinterpret_cast<uintptr_t>(p->w_bf16[l]) & 15) ||
        (reinterpret_cast<uintptr_t>(p->wt_bf16[l]) & 15)) {
      set_error("%s: layer %d: null or misaligned buffer", what, l);
      return FSK_EINVAL;
    }
    a.w[l] = p->w_bf16[l];
    a.wt[l] = p->wt_bf16[l];
    a.bias[l] = p->bias[l];
    a.dz[l] = p->dz[l];
    if (l < 3) a.h[l] = p->h[l];
  }
  a.xb = p->xb;
  a.mse_scale = float(p->mse_scale);
  a.loss = p->loss;
  hipLaunchKernelGGL(mlp_rows_step_kernel, dim3(unsigned(cdiv(p->batch, RS_ROWS))),
                     dim3(64 * RS_WAVES), 0, as_stream(stream), a);
  return launch_status(what);
}
