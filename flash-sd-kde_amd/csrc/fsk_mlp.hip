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

// Workgroup tile 64 rows x 128 columns, 4 waves of 32 x 64 (2 x 4 MFMA
// blocks), K step 32. Each K step's operands are loaded once per workgroup
// with coalesced 16-byte loads (X as FP32, rounded to BF16 on the way into
// LDS; Wt as BF16) into one of two LDS buffers, while the MFMAs read the
// other: the next step's global loads are issued before the current step's
// LDS reads and MFMAs and only waited on when stored. Rows >= M and columns
// >= N load clamped (in-bounds) rows and feed outputs that are not stored;
// K is a multiple of 8 and K slots past K are zero.
// ACT 0: Y = Z; 1: Y = silu(Z), Z optionally stored; 2: Y = Z * silu'(Zin)
// with Zin read from the Z pointer (row stride N).
constexpr int GM_BM = 64, GM_BN = 128, GM_BK = 32;
constexpr int GM_LDS = GM_BK + 8;  // u16 per LDS row (80 B: 16-byte aligned, staggered banks)

template <int ACT, bool STORE_Z>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(const float *__restrict__ X, int64_t ldx,
                                                       int64_t M, int K,
                                                       const unsigned short *__restrict__ Wt,
                                                       const float *__restrict__ bias, int N,
                                                       float *__restrict__ Z,
                                                       float *__restrict__ Y, int64_t ldy) {
  __shared__ __align__(16) unsigned short sA[2][GM_BM * GM_LDS];
  __shared__ __align__(16) unsigned short sB[2][GM_BN * GM_LDS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t bm = int64_t(blockIdx.x) * GM_BM;
  const int bn = blockIdx.y * GM_BN;
  // global -> LDS mapping: A 64 rows x 8 float4 (2 per thread), B 128 rows x 4 u16x8 (2 per thread)
  const int ar = tid >> 3, ac = (tid & 7) * 4;  // A: rows ar and ar + 32, k offset ac
  const int br = tid >> 2, bc = (tid & 3) * 8;  // B: rows br and br + 64, k offset bc
  const float *xa[2];
  const unsigned short *wb[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int64_t row = bm + ar + 32 * i;
    xa[i] = X + (row < M ? row : M - 1) * ldx;
    const int n = bn + br + 64 * i;
    wb[i] = Wt + int64_t(n < N ? n : N - 1) * K;
  }
  struct Raw {
    f4 a[2];
    u16x8 b[2];
  };
  const u16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
  auto load = [&](int k0, Raw &r) {
    const bool aok = k0 + ac < K, bok = k0 + bc < K;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      r.a[i] = ld4(xa[i] + (aok ? k0 + ac : 0));
      r.b[i] = *reinterpret_cast<const u16x8 *>(wb[i] + (bok ? k0 + bc : 0));
      if (!aok) r.a[i] = f4{0.f, 0.f, 0.f, 0.f};
      if (!bok) r.b[i] = zero8;
    }
  };
  auto store = [&](int buf, const Raw &r) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      u16x4 w;
#pragma unroll
      for (int e = 0; e < 4; ++e) w[e] = to16<FSK_PREC_BF16>(r.a[i][e]);
      *reinterpret_cast<u16x4 *>(sA[buf] + (ar + 32 * i) * GM_LDS + ac) = w;
      *reinterpret_cast<u16x8 *>(sB[buf] + (br + 64 * i) * GM_LDS + bc) = r.b[i];
    }
  };
  const int wr = (wave & 1) * 32, wc = (wave >> 1) * 64;  // this wave's 32 x 64 sub-tile
  f4 acc[2][4];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};
  Raw r;
  load(0, r);
  store(0, r);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += GM_BK) {
    const bool more = k0 + GM_BK < K;  // uniform
    if (more) load(k0 + GM_BK, r);
    u16x8 A[2], B[4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
      A[x] = *reinterpret_cast<const u16x8 *>(sA[buf] + (wr + 16 * x + c) * GM_LDS + 8 * g);
#pragma unroll
    for (int y = 0; y < 4; ++y)
      B[y] = *reinterpret_cast<const u16x8 *>(sB[buf] + (wc + 16 * y + c) * GM_LDS + 8 * g);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y)
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf16x8, A[x]), __builtin_bit_cast(bf16x8, B[y]), acc[x][y], 0, 0,
            0);
    if (more) store(buf ^ 1, r);
    __syncthreads();
    buf ^= 1;
  }
  const int64_t m0 = bm + wr;
  const int n0 = bn + wc;
  const bool full = (m0 + 32 <= M) && (n0 + 64 <= N);  // wave-uniform
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int n = n0 + 16 * y + c;
      const float bv = (bias && n < N) ? bias[n] : 0.f;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int64_t row = m0 + 16 * x + 4 * g + rr;
        if (full || (row < M && n < N)) {
          const float z = acc[x][y][rr] + bv;
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

// Whole-K GEMM for K <= 256 (every layer of the 16-256-256-256-16 surrogate):
// the K loop of mlp_gemm_kernel waits on one global->LDS round trip per 32-wide
// K step, which at K = 256 leaves each GEMM latency-bound. Here every thread
// first issues all of its global loads at once into registers - 16 float4 of
// the workgroup's 64 x K slice of X (coalesced along rows) and 8 u16x8 of the
// 64 x K weight slice - then rounds X to BF16 and stores both slices in LDS
// (67.6 KB: two workgroups per CU), and after one barrier the 4 waves (2 x 2
// sub-tiles of 32 x 32) run their MFMAs from LDS back to back. Tasks
// (row tile, column tile) go out in XCD-grouped order so the column tiles of a
// row tile read its X rows from one L2.
// ACT 0..2 as mlp_gemm_kernel; ACT 3 (output layer of a training step): with
// d = Z - T, Y = scale d and loss += sum d^2 (FP64), Z not stored - the loss
// and its gradient fused into the output GEMM.
constexpr int FK_BM = 64, FK_BN = 64, FK_KMAX = 256;
constexpr int FK_LDS = FK_KMAX + 8;  // u16 per LDS row (528 B: 4-bank stagger between rows)
constexpr int FK_A4 = FK_BM * FK_KMAX / 4 / 256;  // float4 loads of X per thread (16)
constexpr int FK_B8 = FK_BN * FK_KMAX / 8 / 256;  // u16x8 loads of W per thread (8)

// KT: K as a compile-time constant (16 or 256, the surrogate's widths: the
// load-index divisions become shifts), or 0 for any K <= 256 at run time.
constexpr int FK_LDS_U16 = (FK_BM + FK_BN) * FK_LDS;  // both slices (67.6 KB)

// One (row tile, column tile) task of the whole-K GEMM; lds holds FK_LDS_U16
// u16 and sl 4 doubles (shared with the weight-gradient body in the fused
// backward kernel). Returns at once for task >= ntasks (uniform per workgroup).
template <int ACT, bool STORE_Z, int KT>
__device__ __forceinline__ void fullk_body(
    int64_t task, unsigned short *__restrict__ lds, double *__restrict__ sl,
    const float *__restrict__ X, int64_t ldx, int64_t M, int K_rt,
    const unsigned short *__restrict__ Wt, const float *__restrict__ bias, int N,
    float *__restrict__ Z, float *__restrict__ Y, int64_t ldy, int ntiles_n, int64_t ntasks,
    const float *__restrict__ T, float mse_scale, double *__restrict__ loss) {
  unsigned short *sA = lds, *sB = lds + FK_BM * FK_LDS;
  const int K = KT ? KT : K_rt;
  if (task >= ntasks) return;  // uniform per workgroup
  const int64_t bm = (task / ntiles_n) * FK_BM;
  const int bn = int(task % ntiles_n) * FK_BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int ksteps = (K + 31) >> 5;
  const int k4 = K >> 2, k8 = K >> 3;  // float4 / u16x8 chunks per row (K % 8 == 0)
  // 1. all global loads of the workgroup in flight at once
  f4 ra[FK_A4];
  u16x8 rb[FK_B8];
#pragma unroll
  for (int i = 0; i < FK_A4; ++i) {
    const int e = tid + 256 * i;
    const int r = e / k4, q = e - r * k4;
    const int64_t row = bm + r < M ? bm + r : M - 1;
    if (e < FK_BM * k4) ra[i] = ld4(X + row * ldx + 4 * q);
  }
#pragma unroll
  for (int i = 0; i < FK_B8; ++i) {
    const int e = tid + 256 * i;
    const int r = e / k8, q = e - r * k8;
    const int n = bn + r < N ? bn + r : N - 1;
    if (e < FK_BN * k8) rb[i] = *reinterpret_cast<const u16x8 *>(Wt + int64_t(n) * K + 8 * q);
  }
  // 2. BF16 X and W slices into LDS; zero K slots up to the last 32-wide step
#pragma unroll
  for (int i = 0; i < FK_A4; ++i) {
    const int e = tid + 256 * i;
    const int r = e / k4, q = e - r * k4;
    if (e < FK_BM * k4) {
      u16x4 w;
#pragma unroll
      for (int j = 0; j < 4; ++j) w[j] = to16<FSK_PREC_BF16>(ra[i][j]);
      *reinterpret_cast<u16x4 *>(sA + r * FK_LDS + 4 * q) = w;
    }
  }
#pragma unroll
  for (int i = 0; i < FK_B8; ++i) {
    const int e = tid + 256 * i;
    const int r = e / k8, q = e - r * k8;
    if (e < FK_BN * k8) *reinterpret_cast<u16x8 *>(sB + r * FK_LDS + 8 * q) = rb[i];
  }
  const int kpad = ksteps * 4 - k8;  // u16x8 chunks of zeros after K in each row
  const u16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
  const int kdiv = kpad > 0 ? kpad : 1;  // the loop is empty when kpad = 0 (K = 256)
  for (int e = tid; e < (FK_BM + FK_BN) * kpad; e += 256) {
    const int r = e / kdiv, q = k8 + (e - r * kdiv);
    unsigned short *dst = r < FK_BM ? sA + r * FK_LDS : sB + (r - FK_BM) * FK_LDS;
    *reinterpret_cast<u16x8 *>(dst + 8 * q) = zero8;
  }
  __syncthreads();
  // 3. MFMAs from LDS
  const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
  f4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < FK_KMAX / 32; ++ks) {
    if (ks < ksteps) {  // uniform
      u16x8 A[2], B[2];
#pragma unroll
      for (int x = 0; x < 2; ++x)
        A[x] = *reinterpret_cast<const u16x8 *>(sA + (wr + 16 * x + c) * FK_LDS + 32 * ks + 8 * g);
#pragma unroll
      for (int y = 0; y < 2; ++y)
        B[y] = *reinterpret_cast<const u16x8 *>(sB + (wc + 16 * y + c) * FK_LDS + 32 * ks + 8 * g);
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(bf16x8, A[x]), __builtin_bit_cast(bf16x8, B[y]), acc[x][y], 0,
              0, 0);
    }
  }
  const int64_t m0 = bm + wr;
  const int n0 = bn + wc;
  double lsum = 0.0;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int n = n0 + 16 * y + c;
      const float bv = (bias && n < N) ? bias[n] : 0.f;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int64_t row = m0 + 16 * x + 4 * g + rr;
        if (row < M && n < N) {
          const float z = acc[x][y][rr] + bv;
          if (ACT == 3) {
            const float d = z - T[row * N + n];
            Y[row * ldy + n] = mse_scale * d;
            lsum += double(d) * d;
          } else if (ACT == 2) {
            Y[row * ldy + n] = z * silu_grad(Z[row * N + n]);
          } else {
            if (STORE_Z) Z[row * N + n] = z;
            Y[row * ldy + n] = ACT ? silu(z) : z;
          }
        }
      }
    }
  if (ACT == 3) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) lsum += __shfl_xor(lsum, o);
    if (lane == 0) sl[wave] = lsum;
    __syncthreads();
    if (tid == 0 && loss) atomicAdd(loss, (sl[0] + sl[1]) + (sl[2] + sl[3]));
  }
}

template <int ACT, bool STORE_Z, int KT>
__global__ __launch_bounds__(256) void mlp_gemm_fullk_kernel(
    const float *__restrict__ X, int64_t ldx, int64_t M, int K_rt,
    const unsigned short *__restrict__ Wt, const float *__restrict__ bias, int N,
    float *__restrict__ Z, float *__restrict__ Y, int64_t ldy, int ntiles_n, int64_t ntasks,
    int64_t per_xcd, const float *__restrict__ T, float mse_scale, double *__restrict__ loss) {
  __shared__ __align__(16) unsigned short lds[FK_LDS_U16];
  __shared__ double sl[4];
  fullk_body<ACT, STORE_Z, KT>(xcd_task(blockIdx.x, per_xcd), lds, sl, X, ldx, M, K_rt, Wt, bias,
                               N, Z, Y, ldy, ntiles_n, ntasks, T, mse_scale, loss);
}

template <int ACT, bool STORE_Z>
static void launch_fullk(const float *x, int64_t ldx, int64_t M, int K, const unsigned short *wt,
                         const float *bias, int N, float *z, float *y, int64_t ldy,
                         const float *t, float scale, double *loss, hipStream_t s) {
  const int ntn = int(cdiv(N, FK_BN));
  const int64_t ntasks = cdiv(M, FK_BM) * ntn;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
#define FSK_FULLK(KT)                                                                          \
  hipLaunchKernelGGL((mlp_gemm_fullk_kernel<ACT, STORE_Z, KT>), dim3(grid), dim3(256), 0, s, x, \
                     ldx, M, K, wt, bias, N, z, y, ldy, ntn, ntasks, per_xcd, t, scale, loss)
  if (K == 256)
    FSK_FULLK(256);
  else if (K == 16)
    FSK_FULLK(16);
  else
    FSK_FULLK(0);
#undef FSK_FULLK
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

// BF16 layer inputs: 4 rows x 4 values (8 bytes each) of a batch chunk.
__device__ __forceinline__ void wgrad_load_u16(const unsigned short *__restrict__ p, int64_t ld,
                                               int64_t m, int64_t me, bool col_ok,
                                               u16x4 (&v)[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const bool ok = col_ok && m + j < me;
    const u16x4 x = *reinterpret_cast<const u16x4 *>(p + (m + j < me ? m + j : me - 1) * ld);
    v[j] = ok ? x : u16x4{0, 0, 0, 0};
  }
}

__device__ __forceinline__ void wgrad_stage_u16(unsigned short *sT, int b, int q,
                                                const u16x4 (&v)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const u16x4 w = {v[0][i], v[1][i], v[2][i], v[3][i]};
    *reinterpret_cast<u16x4 *>(sT + (4 * b + i) * WG_LDS_STRIDE + 4 * q) = w;
  }
}

// LDS of one weight-gradient task: dZ^T and X^T chunks and the bias partials.
constexpr int WG_LDS_U16 = 2 * 64 * WG_LDS_STRIDE + 2 * 16 * 64;
static_assert(WG_LDS_U16 <= FK_LDS_U16, "the fused backward kernel shares one LDS buffer");

// Weight gradient of the (n tile nb, k tile kb, batch split s) task; X is
// FP32 (rounded to BF16 as it is staged) or, with XB, already BF16.
template <bool XB = false>
__device__ __forceinline__ void wgrad_body(int nb, int kb, int64_t s,
                                           unsigned short *__restrict__ lds,
                                           const float *__restrict__ dZ, int64_t M, int N,
                                           const void *__restrict__ Xv, int64_t ldx, int K,
                                           int64_t split_len, float *__restrict__ dW_part,
                                           float *__restrict__ db_part) {
  unsigned short *sA = lds;                           // dZ^T chunk [n][m]
  unsigned short *sB = lds + 64 * WG_LDS_STRIDE;      // X^T chunk [k][m]
  float(*sbias)[64] = reinterpret_cast<float(*)[64]>(lds + 2 * 64 * WG_LDS_STRIDE);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int q = tid >> 4, b = tid & 15;
  const int n0 = nb * 64, k0 = kb * 64;
  const int wn = (wave & 1) * 32, wk = (wave >> 1) * 32;
  const int64_t mb = s * split_len;
  const int64_t me = (mb + split_len < M) ? mb + split_len : M;
  const bool a_ok = n0 + 4 * b < N, b_ok = k0 + 4 * b < K;  // N, K multiples of 8
  const float *pa = dZ + (a_ok ? n0 + 4 * b : 0);
  const float *pb = static_cast<const float *>(Xv) + (b_ok ? k0 + 4 * b : 0);
  const unsigned short *pbh = static_cast<const unsigned short *>(Xv) + (b_ok ? k0 + 4 * b : 0);
  f4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};
  f4 bsum = {0.f, 0.f, 0.f, 0.f};
  f4 va[4], vb[4];
  u16x4 vh[4];
  wgrad_load(pa, N, mb + 4 * q, me, a_ok, va);
  if (XB)
    wgrad_load_u16(pbh, ldx, mb + 4 * q, me, b_ok, vh);
  else
    wgrad_load(pb, ldx, mb + 4 * q, me, b_ok, vb);
  for (int64_t m0 = mb; m0 < me; m0 += 64) {
#pragma unroll
    for (int j = 0; j < 4; ++j) bsum += va[j];
    __syncthreads();  // previous chunk's LDS reads are done
    wgrad_stage(sA, b, q, va);
    if (XB)
      wgrad_stage_u16(sB, b, q, vh);
    else
      wgrad_stage(sB, b, q, vb);
    __syncthreads();
    if (m0 + 64 < me) {
      wgrad_load(pa, N, m0 + 64 + 4 * q, me, a_ok, va);
      if (XB)
        wgrad_load_u16(pbh, ldx, m0 + 64 + 4 * q, me, b_ok, vh);
      else
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
  if (kb == 0) {  // bias gradient: fixed-order sum over the 16 row groups
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

__global__ __launch_bounds__(256) void mlp_wgrad_kernel(const float *__restrict__ dZ, int64_t M,
                                                        int N, const float *__restrict__ X,
                                                        int64_t ldx, int K, int64_t split_len,
                                                        float *__restrict__ dW_part,
                                                        float *__restrict__ db_part) {
  __shared__ __align__(16) unsigned short lds[WG_LDS_U16];
  wgrad_body(blockIdx.x, blockIdx.y, blockIdx.z, lds, dZ, M, N, X, ldx, K, split_len, dW_part,
             db_part);
}

// Backward of one hidden layer in one launch: workgroups [0, grid_d) run the
// data gradient dX = (dZ W) * silu'(Z_prev) (whole-K tasks, XCD-grouped),
// the rest the weight-gradient tasks (n tile, k tile, split) - the two are
// independent given dZ, so they need not be separate launches.
template <int KT>
__global__ __launch_bounds__(256) void mlp_bwd_kernel(
    const float *__restrict__ dZ, int64_t M, int N, const float *__restrict__ X, int64_t ldx,
    int K, int64_t split_len, float *__restrict__ dW_part, float *__restrict__ db_part,
    const unsigned short *__restrict__ Wt, const float *__restrict__ Zp, float *__restrict__ dX,
    int64_t grid_d, int ntn_d, int64_t ntasks_d, int64_t per_xcd_d, int ntn_w, int ntk_w) {
  __shared__ __align__(16) unsigned short lds[FK_LDS_U16];
  __shared__ double sl[4];
  const int64_t b = blockIdx.x;
  if (b < grid_d) {
    // dX (M x K) = dZ (M x N) . W (N x K): the whole-K GEMM over N with Wt = W^T (K x N)
    fullk_body<2, false, KT>(xcd_task(b, per_xcd_d), lds, sl, dZ, N, M, N, Wt, nullptr, K,
                             const_cast<float *>(Zp), dX, K, ntn_d, ntasks_d, nullptr, 0.f,
                             nullptr);
  } else {
    const int64_t t = b - grid_d;
    const int nb = int(t % ntn_w);
    const int kb = int((t / ntn_w) % ntk_w);
    wgrad_body(nb, kb, t / (int64_t(ntn_w) * ntk_w), lds, dZ, M, N, X, ldx, K, split_len, dW_part,
               db_part);
  }
}

// Weight gradients of up to 4 layers in one launch: block b belongs to layer
// k when start[k] <= b < start[k + 1]; within it (n tile, k tile, split).
struct WgradBatch {
  fsk_mlp_wgrad_layer l[4];
  int64_t split_len[4];
  int ntn[4], ntk[4];
  int64_t start[5];
  int n;
  int64_t M;
};

__global__ __launch_bounds__(256) void mlp_wgrad_multi_kernel(WgradBatch wb) {
  __shared__ __align__(16) unsigned short lds[WG_LDS_U16];
  int k = 0;
  while (k + 1 < wb.n && int64_t(blockIdx.x) >= wb.start[k + 1]) ++k;
  const fsk_mlp_wgrad_layer &L = wb.l[k];
  const int64_t t = int64_t(blockIdx.x) - wb.start[k];
  const int nb = int(t % wb.ntn[k]);
  const int kb = int((t / wb.ntn[k]) % wb.ntk[k]);
  const int64_t s = t / (int64_t(wb.ntn[k]) * wb.ntk[k]);
  if (L.x_bf16)  // uniform over the workgroup
    wgrad_body<true>(nb, kb, s, lds, L.dz, wb.M, L.N, L.x, L.ldx, L.K, wb.split_len[k], L.dw_part,
                     L.db_part);
  else
    wgrad_body<false>(nb, kb, s, lds, L.dz, wb.M, L.N, L.x, L.ldx, L.K, wb.split_len[k],
                      L.dw_part, L.db_part);
}

__global__ void mlp_adam_kernel(float *__restrict__ w, float *__restrict__ mom,
                                float *__restrict__ vel, const float *__restrict__ part, int S,
                                int64_t count, int cols, float lr, float b1, float omb1, float b2,
                                float omb2, float eps, float bc1, float bc2,
                                unsigned short *__restrict__ wb,
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
  const float m = b1 * mom[e] + omb1 * gsum;
  const float v = b2 * vel[e] + omb2 * gsum * gsum;
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

// omb1 / omb2 = 1 - beta1 / 1 - beta2, formed in double on the host: FP32
// 1 - float(0.999) would be 1.3e-5 off 0.001 in relative terms.
__global__ void mlp_adam_multi_kernel(AdamBatch batch, float lr, float b1, float omb1, float b2,
                                      float omb2, float eps, float bc1, float bc2,
                                      const float *__restrict__ sched, int64_t *counter) {
  // the training step's device counter (read by the next step's gather; this
  // step's schedule entry was copied to sched before the step began)
  if (counter && blockIdx.x == 0 && threadIdx.x == 0) *counter += 1;
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
  const float m = b1 * T.m[e] + omb1 * gsum;
  const float v = b2 * T.v[e] + omb2 * gsum * gsum;
  T.m[e] = m;
  T.v[e] = v;
  const float nw = T.w[e] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
  T.w[e] = nw;
  if (T.w_bf16) T.w_bf16[e] = to16<FSK_PREC_BF16>(nw);
  if (T.wt_bf16) T.wt_bf16[(e % T.cols) * T.rows + e / T.cols] = to16<FSK_PREC_BF16>(nw);
}

// Minibatch of training step c = min(*counter, steps - 1): rows idx[c][r] of
// x (n x dx) and t (n x dt) into xb, tb, and the step's schedule entry
// table[c] (3 floats) into sched. One thread per (row, float4 of x or t).
__global__ void mlp_gather_kernel(const float *__restrict__ x, const float *__restrict__ t, int dx,
                                  int dt, const int *__restrict__ idx, int64_t batch,
                                  const float *__restrict__ table, int64_t steps,
                                  const int64_t *__restrict__ counter, float *__restrict__ xb,
                                  float *__restrict__ tb, float *__restrict__ sched) {
  int64_t c = *counter;
  c = c < steps ? c : steps - 1;
  const int per = (dx + dt) / 4;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e == 0) {
    sched[0] = table[c * 3];
    sched[1] = table[c * 3 + 1];
    sched[2] = table[c * 3 + 2];
  }
  if (e >= batch * per) return;
  const int64_t r = e / per;
  const int p = int(e % per) * 4;
  const int64_t j = idx[c * batch + r];
  if (p < dx)
    *reinterpret_cast<f4 *>(xb + r * dx + p) = ld4(x + j * dx + p);
  else
    *reinterpret_cast<f4 *>(tb + r * dt + (p - dx)) = ld4(t + j * dt + (p - dx));
}

__global__ void mlp_step_advance_kernel(int64_t *counter) {
  if (threadIdx.x == 0) *counter += 1;
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
  if (K <= FK_KMAX) {  // every load issued up front (the surrogate's layers)
    if (act == 2)
      launch_fullk<2, false>(x, ldx, M, K, wt_bf16, bias, N, z, y, ldy, nullptr, 0.f, nullptr, s);
    else if (act == 1 && z)
      launch_fullk<1, true>(x, ldx, M, K, wt_bf16, bias, N, z, y, ldy, nullptr, 0.f, nullptr, s);
    else if (act == 1)
      launch_fullk<1, false>(x, ldx, M, K, wt_bf16, bias, N, z, y, ldy, nullptr, 0.f, nullptr, s);
    else if (z)
      launch_fullk<0, true>(x, ldx, M, K, wt_bf16, bias, N, z, y, ldy, nullptr, 0.f, nullptr, s);
    else
      launch_fullk<0, false>(x, ldx, M, K, wt_bf16, bias, N, z, y, ldy, nullptr, 0.f, nullptr, s);
    return launch_status(what);
  }
  dim3 grid(unsigned(cdiv(M, GM_BM)), unsigned(cdiv(N, GM_BN)));
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

extern "C" int fsk_mlp_linear_fwd_mse(const float *x, int64_t ldx, int64_t M, int K,
                                      const unsigned short *wt_bf16, const float *bias, int N,
                                      const float *t, double scale, float *dy, double *loss,
                                      void *stream) {
  const char *what = "fsk_mlp_linear_fwd_mse";
  int st = mlp_dims_ok(what, M, K, N);
  if (st != FSK_OK) return st;
  if (!x || !wt_bf16 || !t || !dy || ldx < K || (ldx & 3) || K > FK_KMAX ||
      (reinterpret_cast<uintptr_t>(x) & 15) || (reinterpret_cast<uintptr_t>(wt_bf16) & 15)) {
    set_error("%s: bad pointers or strides, or K = %d above %d", what, K, FK_KMAX);
    return FSK_EINVAL;
  }
  launch_fullk<3, false>(x, ldx, M, K, wt_bf16, bias, N, nullptr, dy, N, t, float(scale), loss,
                         as_stream(stream));
  return launch_status(what);
}

extern "C" int fsk_mlp_linear_bwd(const float *dz, int64_t M, int N, const float *x, int64_t ldx,
                                  int K, int splits, float *dw_part, float *db_part,
                                  const unsigned short *wt_bf16, const float *z_prev, float *dx,
                                  void *stream) {
  const char *what = "fsk_mlp_linear_bwd";
  int st = mlp_dims_ok(what, M, K, N);
  if (st != FSK_OK) return st;
  if (!dz || !x || !dw_part || !db_part || !wt_bf16 || !z_prev || !dx || splits <= 0 ||
      splits > 1024 || ldx < K || (ldx & 3) || N > FK_KMAX ||
      (reinterpret_cast<uintptr_t>(x) & 15) || (reinterpret_cast<uintptr_t>(dz) & 15) ||
      (reinterpret_cast<uintptr_t>(wt_bf16) & 15)) {
    set_error("%s: bad arguments (16-byte aligned pointers, ldx a multiple of 4, N <= %d)", what,
              FK_KMAX);
    return FSK_EINVAL;
  }
  const int ntn_d = int(cdiv(K, FK_BN));
  const int64_t ntasks_d = cdiv(M, FK_BM) * ntn_d;
  int64_t per_xcd_d;
  const int64_t grid_d = xcd_grid(ntasks_d, &per_xcd_d);
  const int ntn_w = int(cdiv(N, 64)), ntk_w = int(cdiv(K, 64));
  const int64_t grid = grid_d + int64_t(ntn_w) * ntk_w * splits;
  const int64_t split_len = round_up(cdiv(M, splits), 64);
  hipStream_t s = as_stream(stream);
#define FSK_BWD(KT)                                                                             \
  hipLaunchKernelGGL((mlp_bwd_kernel<KT>), dim3(unsigned(grid)), dim3(256), 0, s, dz, M, N, x,  \
                     ldx, K, split_len, dw_part, db_part, wt_bf16, z_prev, dx, grid_d, ntn_d,   \
                     ntasks_d, per_xcd_d, ntn_w, ntk_w)
  if (N == 256)
    FSK_BWD(256);
  else if (N == 16)
    FSK_BWD(16);
  else
    FSK_BWD(0);
#undef FSK_BWD
  return launch_status(what);
}

extern "C" int fsk_mlp_wgrad_multi(const fsk_mlp_wgrad_layer *layers, int n_layers, int64_t M,
                                   void *stream) {
  const char *what = "fsk_mlp_wgrad_multi";
  if (!layers || n_layers <= 0 || n_layers > 4 || M <= 0) {
    set_error("%s: need 1..4 layers and M > 0", what);
    return FSK_EINVAL;
  }
  WgradBatch wb{};
  wb.n = n_layers;
  wb.M = M;
  wb.start[0] = 0;
  for (int k = 0; k < n_layers; ++k) {
    const fsk_mlp_wgrad_layer &L = layers[k];
    int st = mlp_dims_ok(what, M, L.K, L.N);
    if (st != FSK_OK) return st;
    if (!L.dz || !L.x || !L.dw_part || !L.db_part || L.splits <= 0 || L.splits > 1024 ||
        L.ldx < L.K || (L.ldx & 3) || (reinterpret_cast<uintptr_t>(L.x) & (L.x_bf16 ? 7 : 15)) ||
        (reinterpret_cast<uintptr_t>(L.dz) & 15)) {
      set_error("%s: layer %d: bad arguments", what, k);
      return FSK_EINVAL;
    }
    wb.l[k] = L;
    wb.split_len[k] = round_up(cdiv(M, L.splits), 64);
    wb.ntn[k] = int(cdiv(L.N, 64));
    wb.ntk[k] = int(cdiv(L.K, 64));
    wb.start[k + 1] = wb.start[k] + int64_t(wb.ntn[k]) * wb.ntk[k] * L.splits;
  }
  hipLaunchKernelGGL(mlp_wgrad_multi_kernel, dim3(unsigned(wb.start[n_layers])), dim3(256), 0,
                     as_stream(stream), wb);
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
                     v, grad_part, splits, count, cols, float(lr), float(beta1),
                     float(1.0 - beta1), float(beta2), float(1.0 - beta2), float(eps), bc1, bc2,
                     w_bf16, wt_bf16, rows, sched);
  return launch_status("fsk_mlp_adam");
}

extern "C" int fsk_mlp_adam_multi_step(const fsk_adam_tensor *tensors, int n_tensors, double lr,
                                       double beta1, double beta2, double eps, int step,
                                       const float *sched, int64_t *counter, void *stream) {
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
                     as_stream(stream), b, float(lr), float(beta1), float(1.0 - beta1),
                     float(beta2), float(1.0 - beta2), float(eps), bc1, bc2, sched, counter);
  return launch_status(what);
}

extern "C" int fsk_mlp_adam_multi(const fsk_adam_tensor *tensors, int n_tensors, double lr,
                                  double beta1, double beta2, double eps, int step,
                                  const float *sched, void *stream) {
  return fsk_mlp_adam_multi_step(tensors, n_tensors, lr, beta1, beta2, eps, step, sched, nullptr,
                                 stream);
}

extern "C" int fsk_mlp_gather_batch(const float *x, const float *t, int64_t n, int dx, int dt,
                                    const int *idx, int64_t batch, const float *table,
                                    int64_t steps, const int64_t *counter, float *xb, float *tb,
                                    float *sched, void *stream) {
  const char *what = "fsk_mlp_gather_batch";
  if (!x || !t || !idx || !table || !counter || !xb || !tb || !sched || n <= 0 || batch <= 0 ||
      steps <= 0 || dx <= 0 || dt <= 0 || (dx & 3) || (dt & 3) ||
      (reinterpret_cast<uintptr_t>(x) & 15) || (reinterpret_cast<uintptr_t>(t) & 15) ||
      (reinterpret_cast<uintptr_t>(xb) & 15) || (reinterpret_cast<uintptr_t>(tb) & 15)) {
    set_error("%s: bad arguments (widths multiples of 4, 16-byte aligned rows)", what);
    return FSK_EINVAL;
  }
  const int64_t threads = batch * ((dx + dt) / 4);
  hipLaunchKernelGGL(mlp_gather_kernel, dim3(unsigned(cdiv(threads, 256))), dim3(256), 0,
                     as_stream(stream), x, t, dx, dt, idx, batch, table, steps, counter, xb, tb,
                     sched);
  return launch_status(what);
}

extern "C" int fsk_mlp_step_advance(int64_t *counter, void *stream) {
  if (!counter) {
    set_error("fsk_mlp_step_advance: null counter");
    return FSK_EINVAL;
  }
  hipLaunchKernelGGL(mlp_step_advance_kernel, dim3(1), dim3(64), 0, as_stream(stream), counter);
  return launch_status("fsk_mlp_step_advance");
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
