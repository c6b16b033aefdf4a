// fsk_batched.hip - leave-one-out KDE log-likelihood of many datasets over a
// bandwidth grid, in one persistent launch.
//
// For dataset d (n points) and bandwidth h_b:
//   LL_b = 1/n sum_i log( 1/(n-1) sum_{j != i} (2 pi)^-8 h_b^-16 exp(-|x_i - x_j|^2 / 2h_b^2) )
// Each task = (dataset, block of 16*WAVES rows) and runs two phases over all
// points of its dataset:
//   1. dmin_i = min_{j != i} |x_i - x_j|^2            (one v_min per pair)
//   2. S_ib   = sum_{j != i} exp2(c_b (|x_i - x_j|^2 - dmin_i)),  c_b = -log2(e)/(2 h_b^2)
// so log sum_j exp(-d2/2h^2) = c_b dmin_i ln 2 + log S_ib exactly (log-sum-exp
// with the maximum known in advance): every term is <= 1, the nearest
// neighbour contributes 1, and no bandwidth of the grid can underflow.
// One distance feeds all n_bw (<= 32) exponentials; the kernel is exp-bound,
// so the cross term stays FP32 MFMA (it is < 10 % of the issue slots).
// Each dataset is first centred on its own FP64 mean (distances are
// translation invariant; the norm terms, whose cancellation sets the FP32
// error of |x_i|^2 + |x_j|^2 - 2 x_i.x_j, shrink to the dataset's spread), and
// the MFMA accumulator starts from |x_j|^2 alone, |x_i|^2 being added after
// the cross term.
// Tasks come from an atomic work queue, largest datasets first; per task the
// row sums of log S_ib and of dmin_i are reduced in LDS and added to FP64
// per-(dataset, bandwidth) accumulators; a finish kernel forms LL.
#include "fsk_common.h"
#include "fsk_plan.h"

namespace fsk {

constexpr int kLooMaxSets = 64;
constexpr int kLooMaxBw = 32;
constexpr int kLooFlush = 16;
constexpr float kLooFar = 1.0e30f;

struct LooMeta {
  int n_sets, n_bw;
  int64_t ntasks;
  int64_t off[kLooMaxSets + 1];      // dataset row offsets (original order)
  int64_t task_end[kLooMaxSets];     // cumulative task count in processing order
  int order[kLooMaxSets];            // processing order (largest first)
};

// One workgroup per dataset: its FP64 mean (fixed summation order: per
// thread over a strided row set, then a fixed LDS tree), then the centred
// copy xc (16 floats per row, one rounding of the FP64 difference) and the
// squared norms of the centred rows.
__global__ __launch_bounds__(256) void loo_center_kernel(const float *__restrict__ x, int64_t ld,
                                                         LooMeta meta, float *__restrict__ xc,
                                                         float *__restrict__ nrm) {
  __shared__ double red[256][kD + 1];
  __shared__ double mu[kD];
  const int d = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t beg = meta.off[d], end = meta.off[d + 1];
  double acc[kD];
#pragma unroll
  for (int f = 0; f < kD; ++f) acc[f] = 0.0;
  for (int64_t j = beg + tid; j < end; j += 256) {
#pragma unroll
    for (int f = 0; f < kD; f += 4) {
      const f4 v = ld4(x + j * ld + f);
      acc[f] += v[0];
      acc[f + 1] += v[1];
      acc[f + 2] += v[2];
      acc[f + 3] += v[3];
    }
  }
#pragma unroll
  for (int f = 0; f < kD; ++f) red[tid][f] = acc[f];
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int f = 0; f < kD; ++f) red[tid][f] += red[tid + w][f];
    }
    __syncthreads();
  }
  if (tid < kD) mu[tid] = red[0][tid] / double(end - beg);
  __syncthreads();
  for (int64_t j = beg + tid; j < end; j += 256) {
    double s = 0.0;
#pragma unroll
    for (int f = 0; f < kD; f += 4) {
      const f4 v = ld4(x + j * ld + f);
      f4 o;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        o[e] = float(double(v[e]) - mu[f + e]);
        s += double(o[e]) * double(o[e]);
      }
      *reinterpret_cast<f4 *>(xc + j * kD + f) = o;
    }
    nrm[j] = float(s);
  }
}

template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void loo_grid_kernel(
    const float *__restrict__ x, int64_t ld, const float *__restrict__ nrm,
    const double *__restrict__ bw, LooMeta meta, int *__restrict__ queue,
    double *__restrict__ sumlog, double *__restrict__ sumdmin) {
  constexpr int ROWS = 16 * WAVES;
  __shared__ int task_sh;
  __shared__ float coef_sh[kLooMaxBw];
  __shared__ double red_sh[WAVES][kLooMaxBw + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int nbw = meta.n_bw;

  for (;;) {
    __syncthreads();
    if (tid == 0) task_sh = atomicAdd(queue, 1);
    __syncthreads();
    const int64_t t = task_sh;
    if (t >= meta.ntasks) break;
    // decode (scalar): which dataset, which row block
    int k = 0;
    while (t >= meta.task_end[k]) ++k;
    const int d = meta.order[k];
    const int64_t blk = t - (k ? meta.task_end[k - 1] : 0);
    const int64_t beg = meta.off[d], end = meta.off[d + 1];
    if (tid < nbw) {
      const double h = bw[d * kLooMaxBw + tid];
      coef_sh[tid] = float(-kLog2e / (2.0 * h * h));
    }
    __syncthreads();
    float cf[kLooMaxBw];  // wave-uniform: kept in scalar registers
#pragma unroll
    for (int b = 0; b < kLooMaxBw; ++b)
      cf[b] = __builtin_bit_cast(
          float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, b < nbw ? coef_sh[b] : 0.f)));

    const int64_t rbase = beg + blk * ROWS + wave * 16;  // this wave's 16 rows
    const int64_t i = rbase + c;
    const bool row_ok = i < end;
    const int64_t ic = row_ok ? i : end - 1;
    const f4 qv = ld4(x + ic * ld + 4 * g) * -2.0f;
    const float qa = nrm[ic];

    // ---- phase 1: nearest-neighbour distance (j != i) ----
    float dmin = kLooFar;
    for (int64_t j0 = beg; j0 < end; j0 += kTile) {
      int64_t ja = j0 + c;
      if (ja > end - 1) ja = end - 1;
      const f4 xa = ld4(x + ja * ld + 4 * g);
      f4 bb;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t jb = j0 + 4 * g + r;
        bb[r] = jb < end ? nrm[jb] : kLooFar;
      }
      f4 dd = bb;
#pragma unroll
      for (int s = 0; s < 4; ++s) dd = mfma_f32(xa[s], qv[s], dd);
      dd += qa;
      if (j0 == rbase) {  // diagonal tile: drop j == i
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (4 * g + r == c) dd[r] = kLooFar;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) dmin = fminf(dmin, dd[r]);
    }
    dmin = fminf(dmin, __shfl_xor(dmin, 16));
    dmin = fminf(dmin, __shfl_xor(dmin, 32));
    dmin = fmaxf(dmin, 0.f);

    // ---- phase 2: S_ib over the bandwidth grid ----
    float S[kLooMaxBw];
    double S64[kLooMaxBw];
#pragma unroll
    for (int b = 0; b < kLooMaxBw; ++b) {
      S[b] = 0.f;
      S64[b] = 0.0;
    }
    const float qad = qa - dmin;
    int tc = 0;
    for (int64_t j0 = beg; j0 < end; j0 += kTile) {
      int64_t ja = j0 + c;
      if (ja > end - 1) ja = end - 1;
      const f4 xa = ld4(x + ja * ld + 4 * g);
      f4 bb;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t jb = j0 + 4 * g + r;
        bb[r] = jb < end ? nrm[jb] : kLooFar;
      }
      f4 dd = bb;
#pragma unroll
      for (int s = 0; s < 4; ++s) dd = mfma_f32(xa[s], qv[s], dd);
      dd += qad;
      if (j0 == rbase) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (4 * g + r == c) dd[r] = kLooFar;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float delta = fmaxf(dd[r], 0.f);
#pragma unroll
        for (int b = 0; b < kLooMaxBw; ++b) S[b] += fexp2(delta * cf[b]);  // cf = 0 past n_bw
      }
      if (++tc == kLooFlush) {
        tc = 0;
#pragma unroll
        for (int b = 0; b < kLooMaxBw; ++b) {
          S64[b] += S[b];
          S[b] = 0.f;
        }
      }
    }
    // ---- per-task reduction: sum over rows of log S_ib and of dmin_i ----
#pragma unroll
    for (int b = 0; b < kLooMaxBw; ++b) {
      double v = sum_groups(S64[b] + double(S[b]));
      v = (row_ok && b < nbw) ? log(v) : 0.0;
      // sum over the 16 rows of the wave (lanes c, group 0 holds row c)
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      v += __shfl_xor(v, 4);
      v += __shfl_xor(v, 8);
      if (lane == 0) red_sh[wave][b] = v;
    }
    double dm = row_ok ? double(dmin) : 0.0;
    if (g != 0) dm = 0.0;
    dm += __shfl_xor(dm, 1);
    dm += __shfl_xor(dm, 2);
    dm += __shfl_xor(dm, 4);
    dm += __shfl_xor(dm, 8);
    if (lane == 0) red_sh[wave][kLooMaxBw] = dm;
    __syncthreads();
    if (tid <= kLooMaxBw) {
      double s = 0.0;
      for (int w = 0; w < WAVES; ++w) s += red_sh[w][tid];
      if (tid < nbw) atomicAdd(&sumlog[d * kLooMaxBw + tid], s);
      if (tid == kLooMaxBw) atomicAdd(&sumdmin[d], s);
    }
  }
}

__global__ void loo_finish_kernel(const double *__restrict__ sumlog,
                                  const double *__restrict__ sumdmin,
                                  const double *__restrict__ bw, LooMeta meta,
                                  double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= meta.n_sets * meta.n_bw) return;
  const int d = e / meta.n_bw, b = e % meta.n_bw;
  const double n = double(meta.off[d + 1] - meta.off[d]);
  const double h = bw[d * kLooMaxBw + b];
  out[e] = sumlog[d * kLooMaxBw + b] / n - sumdmin[d] / (2.0 * h * h * n) - log(n - 1.0) -
           0.5 * kD * log(2.0 * M_PI) - kD * log(h);
}

__global__ void loo_pack_bw_kernel(const double *__restrict__ bw_in, int n_sets, int n_bw,
                                   double *__restrict__ bw_out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_sets * kLooMaxBw) return;
  const int d = e / kLooMaxBw, b = e % kLooMaxBw;
  bw_out[e] = b < n_bw ? bw_in[d * n_bw + b] : 1.0;
}

struct LooLayout {
  int *queue;
  float *xc;  // rows centred on their dataset's mean
  float *nrm;
  double *bw, *sumlog, *sumdmin;
};

static size_t loo_layout(void *ws, int64_t N, int n_sets, LooLayout *L) {
  Carve cv(ws);
  L->queue = cv.take<int>(64);
  L->xc = cv.take<float>(N * kD);
  L->nrm = cv.take<float>(N);
  L->bw = cv.take<double>(int64_t(n_sets) * kLooMaxBw);
  L->sumlog = cv.take<double>(int64_t(n_sets) * kLooMaxBw);
  L->sumdmin = cv.take<double>(n_sets);
  return cv.used;
}

}  // namespace fsk

using namespace fsk;

extern "C" size_t fsk_loo_workspace_bytes(const int64_t *offsets, int n_sets, int n_bw) {
  if (!offsets || n_sets <= 0 || n_sets > kLooMaxSets || n_bw <= 0 || n_bw > kLooMaxBw) return 0;
  LooLayout L;
  return loo_layout(nullptr, offsets[n_sets], n_sets, &L);
}

extern "C" int fsk_loo_bandwidth_grid(const float *data, int64_t ld, const int64_t *offsets,
                                      int n_sets, const double *bandwidths, int n_bw,
                                      int waves_per_wg, int wg_per_cu, void *ws,
                                      size_t ws_bytes, double *out, void *stream) {
  const char *what = "fsk_loo_bandwidth_grid";
  if (!offsets || n_sets <= 0 || n_sets > kLooMaxSets || n_bw <= 0 || n_bw > kLooMaxBw ||
      !bandwidths || !out) {
    set_error("%s: need 1..%d datasets, 1..%d bandwidths, non-null pointers", what, kLooMaxSets,
              kLooMaxBw);
    return FSK_EINVAL;
  }
  if (offsets[0] != 0) {
    set_error("%s: offsets[0] must be 0", what);
    return FSK_EINVAL;
  }
  for (int d = 0; d < n_sets; ++d) {
    if (offsets[d + 1] - offsets[d] < 2) {
      set_error("%s: dataset %d has %lld points; leave-one-out needs at least 2", what, d,
                (long long)(offsets[d + 1] - offsets[d]));
      return FSK_EINVAL;
    }
  }
  int st;
  if ((st = check_points16(what, data, ld, offsets[n_sets])) != FSK_OK) return st;
  if (waves_per_wg != 4 && waves_per_wg != 8) {
    set_error("%s: waves_per_wg must be 4 or 8", what);
    return FSK_EINVAL;
  }
  if (wg_per_cu < 1 || wg_per_cu > 8) {
    set_error("%s: wg_per_cu must be in 1..8", what);
    return FSK_EINVAL;
  }
  const int64_t N = offsets[n_sets];
  LooLayout L;
  const size_t need = loo_layout(ws, N, n_sets, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  // plan on the host: largest datasets first (longest tasks start first)
  LooMeta meta;
  meta.n_sets = n_sets;
  meta.n_bw = n_bw;
  for (int d = 0; d <= n_sets; ++d) meta.off[d] = offsets[d];
  for (int d = 0; d < n_sets; ++d) meta.order[d] = d;
  for (int a = 0; a < n_sets; ++a)
    for (int b = a + 1; b < n_sets; ++b) {
      const int64_t na = offsets[meta.order[a] + 1] - offsets[meta.order[a]];
      const int64_t nb = offsets[meta.order[b] + 1] - offsets[meta.order[b]];
      if (nb > na) {
        const int tmp = meta.order[a];
        meta.order[a] = meta.order[b];
        meta.order[b] = tmp;
      }
    }
  const int rows = 16 * waves_per_wg;
  int64_t acc = 0;
  for (int k = 0; k < n_sets; ++k) {
    const int d = meta.order[k];
    acc += cdiv(offsets[d + 1] - offsets[d], rows);
    meta.task_end[k] = acc;
  }
  for (int k = n_sets; k < kLooMaxSets; ++k) meta.task_end[k] = acc;
  meta.ntasks = acc;

  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(fill_f32, dim3(1), dim3(64), 0, s, reinterpret_cast<float *>(L.queue),
                     int64_t(64), 0.f);
  hipLaunchKernelGGL(fill_f64, dim3(cdiv(int64_t(n_sets) * kLooMaxBw, 256)), dim3(256), 0, s,
                     L.sumlog, int64_t(n_sets) * kLooMaxBw, 0.0);
  hipLaunchKernelGGL(fill_f64, dim3(1), dim3(64), 0, s, L.sumdmin, int64_t(n_sets), 0.0);
  hipLaunchKernelGGL(loo_pack_bw_kernel, dim3(cdiv(int64_t(n_sets) * kLooMaxBw, 256)), dim3(256),
                     0, s, bandwidths, n_sets, n_bw, L.bw);
  hipLaunchKernelGGL(loo_center_kernel, dim3(n_sets), dim3(256), 0, s, data, ld, meta, L.xc,
                     L.nrm);
  static int cus_cached = 0;  // compute units of the current device (queried once)
  if (cus_cached == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cus > 0)
      cus_cached = cus;
    else
      cus_cached = 256;
  }
  const int cus = cus_cached;
  int64_t grid = int64_t(cus) * wg_per_cu;
  if (grid > meta.ntasks) grid = meta.ntasks;
  prof_begin(FSK_OP_LOO_GRID, s);
  if (waves_per_wg == 4)
    hipLaunchKernelGGL(loo_grid_kernel<4>, dim3(grid), dim3(256), 0, s, L.xc, int64_t(kD), L.nrm, L.bw, meta,
                       L.queue, L.sumlog, L.sumdmin);
  else
    hipLaunchKernelGGL(loo_grid_kernel<8>, dim3(grid), dim3(512), 0, s, L.xc, int64_t(kD), L.nrm, L.bw, meta,
                       L.queue, L.sumlog, L.sumdmin);
  prof_end(FSK_OP_LOO_GRID, s);
  if ((st = launch_status(what)) != FSK_OK) return st;
  hipLaunchKernelGGL(loo_finish_kernel, dim3(cdiv(int64_t(n_sets) * n_bw, 256)), dim3(256), 0, s,
                     L.sumlog, L.sumdmin, L.bw, meta, out);
  return launch_status(what);
}
