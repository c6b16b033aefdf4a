// fsk_mlp.hip - forward / backward / optimizer kernels of the learned score
// surrogate (a small SiLU MLP with BF16 weights and FP32 accumulation).
//
//   linear_fwd : Y = act(X . Wt^T + b)   X (M x K) FP32, Wt (N x K) BF16,
//                v_mfma_f32_16x16x32_bf16, FP32 accumulate, bias + SiLU fused,
//                pre-activation Z optionally kept for the backward pass.
//                The same kernel gives the data gradient dX = dZ . W by passing
//                the transposed BF16 copy of W; act 2 multiplies that by
//                silu'(Z_prev) so it yields the previous layer's dZ directly.
//   wgrad      : dW = dZ^T X and db = sum_m dZ over split batch ranges into
//                partials (deterministic; the optimizer sums them). Chunks of
//                64 batch rows are loaded coalesced (float4 per lane), rounded
//                to BF16 and written transposed into LDS so the MFMA operands
//                (batch index contiguous) are 16-byte LDS reads.
//   adam       : Adam on FP32 master weights; writes the BF16 copy and the
//                transposed BF16 copy used by the data-gradient GEMM
//   mse        : dY = scale (Y - T), loss += sum (Y - T)^2 (FP64)
// Requirements: K and N multiples of 8 (row slices are 16-byte loads).
#include "fsk_common.h"
#include "fsk_lowp.h"

namespace fsk {

// sigmoid with the hardware reciprocal (1 ulp) instead of a full division
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.f + __expf(-z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid(z);
  return s * (1.f + z * (1.f - s));
}

__device__ __forceinline__ u16x8 bf16x8_from(const float *p) {
  const f4 a = ld4(p), b = ld4(p + 4);
  u16x8 r;
  r[0] = to16<FSK_PREC_BF16>(a[0]);
  r[1] = to16<FSK_PREC_BF16>(a[1]);
  r[2] = to16<FSK_PREC_BF16>(a[2]);
  r[3] = to16<FSK_PREC_BF16>(a[3]);
  r[4] = to16<FSK_PREC_BF16>(b[0]);
  r[5] = to16<FSK_PREC_BF16>(b[1]);
  r[6] = to16<FSK_PREC_BF16>(b[2]);
  r[7] = to16<FSK_PREC_BF16>(b[3]);
  return r;
}

// Workgroup tile 64 x 64, wave tile 32 x 32 (2 x 2 MFMA blocks), K step 32.
// ACT 0: Y = Z; 1: Y = silu(Z), Z optionally stored; 2: Y = Z * silu'(Zin)
// with Zin read from the Z pointer (row stride N).
// Loads are unconditional with clamped row / column indices (rows >= M and
// columns >= N only feed outputs that are not stored; K is a multiple of 8,
// so a lane's 8 K values are all valid or all past K and zeroed by a select),
// and the next K step's operands are loaded before the current step's MFMAs:
// no branch or wait sits between a load and the work that hides it.
template <int ACT, bool STORE_Z>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(const float *__restrict__ X, int64_t ldx,
                                                       int64_t M, int K,
                                                       const unsigned short *__restrict__ Wt,
                                                       const float *__restrict__ bias, int N,
                                                       float *__restrict__ Z,
                                                       float *__restrict__ Y, int64_t ldy) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t m0 = int64_t(blockIdx.x) * 64 + (wave & 1) * 32;
  const int n0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const float *xr[2];
  const unsigned short *wr[2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    const int64_t row = m0 + 16 * rb + c;
    xr[rb] = X + (row < M ? row : M - 1) * ldx;
  }
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    const int n = n0 + 16 * cb + c;
    wr[cb] = Wt + int64_t(n < N ? n : N - 1) * K;
  }
  f4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  // raw operands of one K step: A as FP32 (converted to BF16 when consumed,
  // so a prefetched load is not waited on before the current MFMAs), B as BF16
  struct Raw {
    f4 a[2][2];
    u16x8 b[2];
  };
  auto load = [&](int k0, Raw &r) {
    const int kc = (k0 + 8 * g < K) ? k0 + 8 * g : 0;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      r.a[rb][0] = ld4(xr[rb] + kc);
      r.a[rb][1] = ld4(xr[rb] + kc + 4);
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) r.b[cb] = *reinterpret_cast<const u16x8 *>(wr[cb] + kc);
  };
  const u16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
  auto consume = [&](int k0, const Raw &r) {
    const bool kok = k0 + 8 * g < K;
    u16x8 A[2], B[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      u16x8 v;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        v[e] = to16<FSK_PREC_BF16>(r.a[rb][0][e]);
        v[4 + e] = to16<FSK_PREC_BF16>(r.a[rb][1][e]);
      }
      A[rb] = kok ? v : zero8;
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) B[cb] = kok ? r.b[cb] : zero8;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb)
        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf16x8, A[rb]), __builtin_bit_cast(bf16x8, B[cb]), acc[rb][cb], 0,
            0, 0);
  };
  // two register buffers used in turn (no copies between them, so a
  // prefetched load is only waited on where it is consumed)
  Raw r0, r1;
  load(0, r0);
  for (int k0 = 0; k0 < K; k0 += 64) {
    if (k0 + 32 < K) load(k0 + 32, r1);
    consume(k0, r0);
    if (k0 + 32 >= K) break;
    if (k0 + 64 < K) load(k0 + 64, r0);
    consume(k0 + 32, r1);
  }
  const bool full = (m0 + 32 <= M) && (n0 + 32 <= N);  // wave-uniform
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      const int n = n0 + 16 * cb + c;
      const float bv = (bias && n < N) ? bias[n] : 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + 16 * rb + 4 * g + r;
        if (full || (row < M && n < N)) {
          const float z = acc[rb][cb][r] + bv;
          if (ACT == 2) {
            Y[row * ldy + n] = z * silu_grad(Z[row * N + n]);
          } else {
            if (STORE_Z) Z[row * N + n] = z;
            Y[row * ldy + n] = ACT ? silu(z) : z;
          }
        }
      }
    }
}

// dW_part[s] (N x K) = sum_{m in split s} bf16(dZ[m, n]) bf16(X[m, k]) and,
// from the workgroups of the first K block, db_part[s][n] = sum dZ[m, n] in
// FP32. Workgroup tile 64 (n) x 64 (k), wave tile 32 x 32, batch chunk 64.
// Thread (q, b) = (tid >> 4, tid & 15) loads rows 4q..4q+3, columns 4b..4b+3
// of the chunk (dZ and X), so it holds a 4 x 4 block that it writes
// transposed as four 8-byte LDS stores. The next chunk's global loads are
// issued before the current chunk's MFMAs. split_len is a multiple of 64.
constexpr int WG_LDS_STRIDE = 72;  // u16 per LDS row: 64 batch rows + 8 pad (144 B, 16 B aligned)

// Four rows m..m+3 of one 4-column slice; rows past `me` (loaded from row
// me - 1, valid since me >= 1) and columns past the matrix (the caller points
// p at column 0 for them) are zeroed by selects, so the loads themselves are
// unconditional and in bounds.
__device__ __forceinline__ void wgrad_load(const float *__restrict__ p, int64_t ld, int64_t m,
                                           int64_t me, bool col_ok, f4 (&v)[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const bool ok = col_ok && m + j < me;
    const f4 x = ld4(p + (m + j < me ? m + j : me - 1) * ld);
    v[j] = ok ? x : f4{0.f, 0.f, 0.f, 0.f};
  }
}

__device__ __forceinline__ void wgrad_stage(unsigned short *sT, int b, int q, const f4 (&v)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    u16x4 w;
    w[0] = to16<FSK_PREC_BF16>(v[0][i]);
    w[1] = to16<FSK_PREC_BF16>(v[1][i]);
    w[2] = to16<FSK_PREC_BF16>(v[2][i]);
    w[3] = to16<FSK_PREC_BF16>(v[3][i]);
    *reinterpret_cast<u16x4 *>(sT + (4 * b + i) * WG_LDS_STRIDE + 4 * q) = w;
  }
}

__global__ __launch_bounds__(256) void mlp_wgrad_kernel(const float *__restrict__ dZ, int64_t M,
                                                        int N, const float *__restrict__ X,
                                                        int64_t ldx, int K, int64_t split_len,
                                                        float *__restrict__ dW_part,
                                                        float *__restrict__ db_part) {
  __shared__ __align__(16) unsigned short sA[64 * WG_LDS_STRIDE];  // dZ^T chunk [n][m]
  __shared__ __align__(16) unsigned short sB[64 * WG_LDS_STRIDE];  // X^T chunk [k][m]
  __shared__ float sbias[16][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int q = tid >> 4, b = tid & 15;
  const int n0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
  const int wn = (wave & 1) * 32, wk = (wave >> 1) * 32;
  const int64_t s = blockIdx.z;
  const int64_t mb = s * split_len;
  const int64_t me = (mb + split_len < M) ? mb + split_len : M;
  const bool a_ok = n0 + 4 * b < N, b_ok = k0 + 4 * b < K;  // N, K multiples of 8
  const float *pa = dZ + (a_ok ? n0 + 4 * b : 0);
  const float *pb = X + (b_ok ? k0 + 4 * b : 0);
  f4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};
  f4 bsum = {0.f, 0.f, 0.f, 0.f};
  f4 va[4], vb[4];
  wgrad_load(pa, N, mb + 4 * q, me, a_ok, va);
  wgrad_load(pb, ldx, mb + 4 * q, me, b_ok, vb);
  for (int64_t m0 = mb; m0 < me; m0 += 64) {
#pragma unroll
    for (int j = 0; j < 4; ++j) bsum += va[j];
    __syncthreads();  // previous chunk's LDS reads are done
    wgrad_stage(sA, b, q, va);
    wgrad_stage(sB, b, q, vb);
    __syncthreads();
    if (m0 + 64 < me) {
      wgrad_load(pa, N, m0 + 64 + 4 * q, me, a_ok, va);
      wgrad_load(pb, ldx, m0 + 64 + 4 * q, me, b_ok, vb);
    }
#pragma unroll
    for (int kk = 0; kk < 64; kk += 32) {
      u16x8 A[2], B[2];
#pragma unroll
      for (int x = 0; x < 2; ++x)
        A[x] = *reinterpret_cast<const u16x8 *>(sA + (wn + 16 * x + c) * WG_LDS_STRIDE + kk + 8 * g);
#pragma unroll
      for (int y = 0; y < 2; ++y)
        B[y] = *reinterpret_cast<const u16x8 *>(sB + (wk + 16 * y + c) * WG_LDS_STRIDE + kk + 8 * g);
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(bf16x8, A[x]), __builtin_bit_cast(bf16x8, B[y]), acc[x][y], 0, 0,
              0);
    }
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int k = k0 + wk + 16 * y + c;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn + 16 * x + 4 * g + r;
        if (n < N && k < K) dW_part[(s * N + n) * K + k] = acc[x][y][r];
      }
    }
  if (blockIdx.y == 0) {  // bias gradient: fixed-order sum over the 16 row groups
#pragma unroll
    for (int i = 0; i < 4; ++i) sbias[q][4 * b + i] = bsum[i];
    __syncthreads();
    if (tid < 64 && n0 + tid < N) {
      float t = 0.f;
#pragma unroll
      for (int j = 0; j < 16; ++j) t += sbias[j][tid];
      db_part[s * N + n0 + tid] = t;
    }
  }
}

__global__ void mlp_adam_kernel(float *__restrict__ w, float *__restrict__ mom,
                                float *__restrict__ vel, const float *__restrict__ part, int S,
                                int64_t count, int cols, float lr, float b1, float b2, float eps,
                                float bc1, float bc2, unsigned short *__restrict__ wb,
                                unsigned short *__restrict__ wtb, int rows,
                                const float *__restrict__ sched) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= count) return;
  if (sched) {  // schedule kept on the device: [lr, 1 - beta1^t, 1 - beta2^t]
    lr = sched[0];
    bc1 = sched[1];
    bc2 = sched[2];
  }
  float gsum = 0.f;
#pragma unroll 8
  for (int s = 0; s < S; ++s) gsum += part[int64_t(s) * count + e];
  const float m = b1 * mom[e] + (1.f - b1) * gsum;
  const float v = b2 * vel[e] + (1.f - b2) * gsum * gsum;
  mom[e] = m;
  vel[e] = v;
  const float upd = lr * (m / bc1) / (sqrtf(v / bc2) + eps);
  const float nw = w[e] - upd;
  w[e] = nw;
  if (wb) wb[e] = to16<FSK_PREC_BF16>(nw);
  if (wtb) {
    const int64_t r = e / cols, cc = e % cols;
    wtb[cc * rows + r] = to16<FSK_PREC_BF16>(nw);
  }
}

// Batched form: block b of the grid belongs to tensor k when
// block_start[k] <= b < block_start[k + 1] (blocks never straddle tensors).
struct AdamBatch {
  fsk_adam_tensor t[FSK_ADAM_MAX_TENSORS];
  int64_t block_start[FSK_ADAM_MAX_TENSORS + 1];
  int n;
};

__global__ void mlp_adam_multi_kernel(AdamBatch batch, float lr, float b1, float b2, float eps,
                                      float bc1, float bc2, const float *__restrict__ sched) {
  int k = 0;
  while (k + 1 < batch.n && int64_t(blockIdx.x) >= batch.block_start[k + 1]) ++k;
  const fsk_adam_tensor &T = batch.t[k];
  const int64_t e = (int64_t(blockIdx.x) - batch.block_start[k]) * blockDim.x + threadIdx.x;
  if (e >= T.count) return;
  if (sched) {
    lr = sched[0];
    bc1 = sched[1];
    bc2 = sched[2];
  }
  float gsum = 0.f;
#pragma unroll 8
  for (int s = 0; s < T.splits; ++s) gsum += T.grad_part[int64_t(s) * T.count + e];
  const float m = b1 * T.m[e] + (1.f - b1) * gsum;
  const float v = b2 * T.v[e] + (1.f - b2) * gsum * gsum;
  T.m[e] = m;
  T.v[e] = v;
  const float nw = T.w[e] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
  T.w[e] = nw;
  if (T.w_bf16) T.w_bf16[e] = to16<FSK_PREC_BF16>(nw);
  if (T.wt_bf16) T.wt_bf16[(e % T.cols) * T.rows + e / T.cols] = to16<FSK_PREC_BF16>(nw);
}

__global__ void mlp_cast_kernel(const float *__restrict__ w, int64_t count, int rows, int cols,
                                unsigned short *__restrict__ wb, unsigned short *__restrict__ wtb) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= count) return;
  const unsigned short b = to16<FSK_PREC_BF16>(w[e]);
  if (wb) wb[e] = b;
  if (wtb) wtb[(e % cols) * rows + e / cols] = b;
}

__global__ void mlp_mse_kernel(const float *__restrict__ y, const float *__restrict__ t,
                               int64_t count, float scale, float *__restrict__ dy,
                               double *__restrict__ loss) {
  __shared__ double red[256];
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float d = 0.f;
  if (e < count) {
    d = y[e] - t[e];
    dy[e] = scale * d;
  }
  red[threadIdx.x] = double(d) * d;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

}  // namespace fsk

using namespace fsk;

static int mlp_dims_ok(const char *what, int64_t M, int K, int N) {
  if (M <= 0 || K <= 0 || N <= 0 || (K & 7) || (N & 7)) {
    set_error("%s: need M > 0 and K, N positive multiples of 8 (M=%lld K=%d N=%d)", what,
              (long long)M, K, N);
    return FSK_EINVAL;
  }
  return FSK_OK;
}

extern "C" int fsk_mlp_linear_fwd(const float *x, int64_t ldx, int64_t M, int K,
                                  const unsigned short *wt_bf16, const float *bias, int N, int act,
                                  float *z, float *y, int64_t ldy, void *stream) {
  const char *what = "fsk_mlp_linear_fwd";
  int st = mlp_dims_ok(what, M, K, N);
  if (st != FSK_OK) return st;
  if (!x || !wt_bf16 || !y || ldx < K || (ldx & 3) || ldy < N || act < 0 || act > 2 ||
      (act == 2 && (!z || bias)) ||
      (reinterpret_cast<uintptr_t>(x) & 15) || (reinterpret_cast<uintptr_t>(wt_bf16) & 15)) {
    set_error("%s: bad pointers, strides or activation", what);
    return FSK_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  dim3 grid(unsigned(cdiv(M, 64)), unsigned(cdiv(N, 64)));
  if (act == 2) {
    hipLaunchKernelGGL((mlp_gemm_kernel<2, false>), grid, dim3(256), 0, s, x, ldx, M, K, wt_bf16,
                       bias, N, z, y, ldy);
  } else if (act) {
    if (z)
      hipLaunchKernelGGL((mlp_gemm_kernel<1, true>), grid, dim3(256), 0, s, x, ldx, M, K, wt_bf16,
                         bias, N, z, y, ldy);
    else
      hipLaunchKernelGGL((mlp_gemm_kernel<1, false>), grid, dim3(256), 0, s, x, ldx, M, K, wt_bf16,
                         bias, N, z, y, ldy);
  } else {
    if (z)
      hipLaunchKernelGGL((mlp_gemm_kernel<0, true>), grid, dim3(256), 0, s, x, ldx, M, K, wt_bf16,
                         bias, N, z, y, ldy);
    else
      hipLaunchKernelGGL((mlp_gemm_kernel<0, false>), grid, dim3(256), 0, s, x, ldx, M, K, wt_bf16,
                         bias, N, z, y, ldy);
  }
  return launch_status(what);
}

extern "C" int fsk_mlp_linear_bwd_weight(const float *dz, int64_t M, int N, const float *x,
                                         int64_t ldx, int K, int splits, float *dw_part,
                                         float *db_part, void *stream) {
  const char *what = "fsk_mlp_linear_bwd_weight";
  int st = mlp_dims_ok(what, M, K, N);
  if (st != FSK_OK) return st;
  if (!dz || !x || !dw_part || !db_part || splits <= 0 || splits > 1024 || ldx < K || (ldx & 3) ||
      (reinterpret_cast<uintptr_t>(x) & 15) || (reinterpret_cast<uintptr_t>(dz) & 15)) {
    set_error("%s: bad arguments (pointers 16-byte aligned, ldx a multiple of 4)", what);
    return FSK_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  const int64_t split_len = round_up(cdiv(M, splits), 64);
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(unsigned(cdiv(N, 64)), unsigned(cdiv(K, 64)), splits),
                     dim3(256), 0, s, dz, M, N, x, ldx, K, split_len, dw_part, db_part);
  return launch_status(what);
}

extern "C" int fsk_mlp_adam(float *w, float *m, float *v, const float *grad_part, int splits,
                            int64_t count, int rows, int cols, double lr, double beta1,
                            double beta2, double eps, int step, unsigned short *w_bf16,
                            unsigned short *wt_bf16, const float *sched, void *stream) {
  if (!w || !m || !v || !grad_part || splits <= 0 || count <= 0 || (step <= 0 && !sched) ||
      int64_t(rows) * cols != count) {
    set_error("fsk_mlp_adam: bad arguments");
    return FSK_EINVAL;
  }
  const float bc1 = step > 0 ? float(1.0 - pow(beta1, step)) : 1.f;
  const float bc2 = step > 0 ? float(1.0 - pow(beta2, step)) : 1.f;
  hipLaunchKernelGGL(mlp_adam_kernel, dim3(cdiv(count, 256)), dim3(256), 0, as_stream(stream), w, m,
                     v, grad_part, splits, count, cols, float(lr), float(beta1), float(beta2),
                     float(eps), bc1, bc2, w_bf16, wt_bf16, rows, sched);
  return launch_status("fsk_mlp_adam");
}

extern "C" int fsk_mlp_adam_multi(const fsk_adam_tensor *tensors, int n_tensors, double lr,
                                  double beta1, double beta2, double eps, int step,
                                  const float *sched, void *stream) {
  const char *what = "fsk_mlp_adam_multi";
  if (!tensors || n_tensors <= 0 || n_tensors > FSK_ADAM_MAX_TENSORS || (step <= 0 && !sched)) {
    set_error("%s: need 1..%d tensors and step > 0 or a schedule", what, FSK_ADAM_MAX_TENSORS);
    return FSK_EINVAL;
  }
  AdamBatch b{};
  b.n = n_tensors;
  b.block_start[0] = 0;
  for (int k = 0; k < n_tensors; ++k) {
    const fsk_adam_tensor &T = tensors[k];
    if (!T.w || !T.m || !T.v || !T.grad_part || T.splits <= 0 || T.count <= 0 ||
        int64_t(T.rows) * T.cols != T.count) {
      set_error("%s: bad tensor %d", what, k);
      return FSK_EINVAL;
    }
    b.t[k] = T;
    b.block_start[k + 1] = b.block_start[k] + cdiv(T.count, 256);
  }
  const float bc1 = step > 0 ? float(1.0 - pow(beta1, step)) : 1.f;
  const float bc2 = step > 0 ? float(1.0 - pow(beta2, step)) : 1.f;
  hipLaunchKernelGGL(mlp_adam_multi_kernel, dim3(unsigned(b.block_start[n_tensors])), dim3(256), 0,
                     as_stream(stream), b, float(lr), float(beta1), float(beta2), float(eps), bc1,
                     bc2, sched);
  return launch_status(what);
}

extern "C" int fsk_mlp_cast_bf16(const float *w, int rows, int cols, unsigned short *w_bf16,
                                 unsigned short *wt_bf16, void *stream) {
  if (!w || rows <= 0 || cols <= 0) {
    set_error("fsk_mlp_cast_bf16: bad arguments");
    return FSK_EINVAL;
  }
  const int64_t count = int64_t(rows) * cols;
  hipLaunchKernelGGL(mlp_cast_kernel, dim3(cdiv(count, 256)), dim3(256), 0, as_stream(stream), w,
                     count, rows, cols, w_bf16, wt_bf16);
  return launch_status("fsk_mlp_cast_bf16");
}

extern "C" int fsk_mlp_mse(const float *y, const float *t, int64_t count, double scale, float *dy,
                           double *loss, void *stream) {
  if (!y || !t || !dy || count <= 0) {
    set_error("fsk_mlp_mse: bad arguments");
    return FSK_EINVAL;
  }
  hipLaunchKernelGGL(mlp_mse_kernel, dim3(cdiv(count, 256)), dim3(256), 0, as_stream(stream), y, t,
                     count, float(scale), dy, loss);
  return launch_status("fsk_mlp_mse");
}
