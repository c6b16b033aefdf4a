// fsk_kde1.hip - 1-D Gaussian KDE sums and density-derivative sums (VALU).
//
// Data points are pre-scaled by k = sqrt(log2(e) / (2 h^2)) (x' = k x, one
// rounding); with u = fma(q, k, -x') (the query's product enters unrounded)
// the kernel value is exp(-(q-x)^2/(2h^2)) = exp2(-u^2): one FMA, one
// multiply, one exp2 and one add per pair. The Laplacian factor
// (1.5 - ln2 u^2) multiplies each pair before the sum, as the reference's
// phi * (1 + 1/2 - z^2/2) does (flash_sd_kde/reference.py:200-213), so the
// sum carries no cancellation between two separately rounded totals. A data
// chunk is staged in LDS and read as wave-wide broadcasts; each thread
// carries k1dQPT queries so every LDS read feeds several pairs.
#include "fsk_common.h"
#include "fsk_plan.h"

namespace fsk {

// Padded data points sit this far away: u^2 ~ 1e36 stays finite and exp2(-u^2) = 0.
constexpr float kFar1d = 1.0e18f;

__global__ void prep_scale1(const float *__restrict__ x, int64_t n, int64_t n_pad, float k,
                            float *__restrict__ out) {
  int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n_pad) return;
  out[j] = j < n ? x[j] * k : kFar1d;
}

// MODE 0: plain sums; MODE 1: sums of phi (1.5 - ln2 u^2);
// MODE 2: pdf sums and derivative sums sum(-z phi) (two outputs);
// MODE 3: sums of phi z^2 = phi 2 ln2 u^2 (the Laplacian correction's moment).
template <int MODE, bool ATOMIC>
__global__ __launch_bounds__(256) void kde1_kernel(const float *__restrict__ xs, int64_t n_pad,
                                                   const float *__restrict__ qs, int64_t m,
                                                   int64_t split_len, int64_t qblocks,
                                                   int64_t ntasks, int64_t per_xcd,
                                                   float kscale, float zscale,
                                                   float *__restrict__ out,
                                                   float *__restrict__ out2, int64_t ld_out) {
  __shared__ f4 chunk[k1dChunk / 4];
  const int64_t task = xcd_task(blockIdx.x, per_xcd);
  if (task >= ntasks) return;
  const int64_t split = task / qblocks;
  const int64_t qb = task - split * qblocks;
  const int tid = threadIdx.x;
  float q[k1dQPT], acc[k1dQPT], acc2[k1dQPT];
  double a64[k1dQPT], b64[k1dQPT];
#pragma unroll
  for (int k = 0; k < k1dQPT; ++k) {
    int64_t i = qb * k1dRowsPerWG + k * 256 + tid;
    q[k] = qs[i < m ? i : m - 1];
    acc[k] = acc2[k] = 0.f;
    a64[k] = b64[k] = 0.0;
  }
  const int64_t j_begin = split * split_len;
  int64_t j_end = j_begin + split_len;
  if (j_end > n_pad) j_end = n_pad;
  for (int64_t j0 = j_begin; j0 < j_end; j0 += k1dChunk) {
    __syncthreads();
    reinterpret_cast<float *>(chunk)[tid] = xs[j0 + tid];
    reinterpret_cast<float *>(chunk)[tid + 256] = xs[j0 + tid + 256];
    reinterpret_cast<float *>(chunk)[tid + 512] = xs[j0 + tid + 512];
    reinterpret_cast<float *>(chunk)[tid + 768] = xs[j0 + tid + 768];
    __syncthreads();
    for (int jj = 0; jj < k1dChunk / 4; ++jj) {
      const f4 xv = chunk[jj];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
#pragma unroll
        for (int k = 0; k < k1dQPT; ++k) {
          const float u = fmaf(q[k], kscale, -xv[s]);
          const float t = u * u;
          const float e = fexp2(-t);
          if (MODE == 1)
            acc[k] = fmaf(e, fmaf(t, -float(kLn2), 1.5f), acc[k]);
          else if (MODE == 3)
            acc[k] = fmaf(e, t, acc[k]);
          else
            acc[k] += e;
          if (MODE == 2) acc2[k] = fmaf(e, u, acc2[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < k1dQPT; ++k) {
      a64[k] += acc[k];
      b64[k] += acc2[k];
      acc[k] = acc2[k] = 0.f;
    }
  }
#pragma unroll
  for (int k = 0; k < k1dQPT; ++k) {
    int64_t i = qb * k1dRowsPerWG + k * 256 + tid;
    if (i >= m) continue;
    const double v = MODE == 3 ? 2.0 * kLn2 * a64[k] : a64[k];
    float v2 = 0.f;
    // sum(-z phi) with z = u / sqrt(log2 e / 2): -z = -u * sqrt(2 ln 2)
    if (MODE == 2) v2 = float(-b64[k] * double(zscale));
    if (ATOMIC) {
      atomicAdd(out + i, float(v));
      if (MODE == 2) atomicAdd(out2 + i, v2);
    } else {
      out[split * ld_out + i] = float(v);
      if (MODE == 2) out2[split * ld_out + i] = v2;
    }
  }
}

struct Kde1Layout {
  float *xs, *part, *part2;
};

static size_t kde1_layout(void *ws, int64_t n, int64_t m, bool partials, bool two,
                          Kde1Layout *L) {
  Carve cv(ws);
  SplitPlan p = kde1_plan(n, m);
  L->xs = cv.take<float>(round_up(n, k1dChunk));
  L->part = partials ? cv.take<float>(p.splits * m) : nullptr;
  L->part2 = (partials && two) ? cv.take<float>(p.splits * m) : nullptr;
  return cv.used;
}

size_t kde1_ws_bytes(int64_t n, int64_t m, bool partials, bool two) {
  Kde1Layout L;
  return kde1_layout(nullptr, n, m, partials, two, &L);
}

static int kde1_run(const char *what, int mode, bool atomic, const float *data, int64_t n,
                    const float *query, int64_t m, double h, void *ws, size_t ws_bytes,
                    float *out, float *out2, void *stream) {
  if (!data || !query || !out || n <= 0 || m <= 0 || (mode == 2 && !out2)) {
    set_error("%s: bad arguments (n=%lld m=%lld)", what, (long long)n, (long long)m);
    return FSK_EINVAL;
  }
  if (!(h > 0.0) || !isfinite(h)) {
    set_error("%s: bandwidth must be positive, got %g", what, h);
    return FSK_EINVAL;
  }
  Kde1Layout L;
  size_t need = kde1_layout(ws, n, m, !atomic, mode == 2, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const double k = sqrt(kLog2e / (2.0 * h * h));
  const int64_t n_pad = round_up(n, k1dChunk);
  hipLaunchKernelGGL(prep_scale1, dim3(cdiv(n_pad, 256)), dim3(256), 0, s, data, n, n_pad,
                     float(k), L.xs);
  SplitPlan p = kde1_plan(n, m);
  const int64_t ntasks = p.qblocks * p.splits;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
  // z = (q - x)/h = u / sqrt(log2(e)/2)  ->  -z = -u * sqrt(2 ln 2)
  const float zscale = float(sqrt(2.0 * kLn2));
  float *o1 = atomic ? out : L.part;
  float *o2 = atomic ? out2 : L.part2;
  if (atomic) {
    hipLaunchKernelGGL(fill_f32, dim3(cdiv(m, 256)), dim3(256), 0, s, out, m, 0.f);
    if (mode == 2) hipLaunchKernelGGL(fill_f32, dim3(cdiv(m, 256)), dim3(256), 0, s, out2, m, 0.f);
  }
  const int op = mode == 2 ? FSK_OP_KDE1_SCORE : (atomic ? FSK_OP_KDE1_ATOMIC : FSK_OP_KDE1_SPLITK);
  prof_begin(op, s);
#define FSK_K1(MODE, AT)                                                                      \
  hipLaunchKernelGGL((kde1_kernel<MODE, AT>), dim3(grid), dim3(256), 0, s, L.xs, n_pad, query, m, \
                     p.split_len, p.qblocks, ntasks, per_xcd, float(k), zscale, o1, o2, m)
  if (atomic) {
    if (mode == 0) FSK_K1(0, true);
    else if (mode == 1) FSK_K1(1, true);
    else if (mode == 3) FSK_K1(3, true);
    else FSK_K1(2, true);
  } else {
    if (mode == 0) FSK_K1(0, false);
    else if (mode == 1) FSK_K1(1, false);
    else if (mode == 3) FSK_K1(3, false);
    else FSK_K1(2, false);
  }
#undef FSK_K1
  prof_end(op, s);
  int st = launch_status(what);
  if (st != FSK_OK || atomic) return st;
  if ((st = fsk_reduce_partials_f64(L.part, p.splits, m, m, out, stream)) != FSK_OK) return st;
  if (mode == 2) st = fsk_reduce_partials_f64(L.part2, p.splits, m, m, out2, stream);
  return st;
}

}  // namespace fsk

using namespace fsk;

extern "C" int fsk_kde1_splitk(const float *data, int64_t n, const float *query, int64_t m,
                               double bandwidth, unsigned flags, void *ws, size_t ws_bytes,
                               float *out, void *stream) {
  if ((flags & ~(FSK_FLAG_LAPLACIAN | FSK_FLAG_PHI_S)) ||
      (flags & (FSK_FLAG_LAPLACIAN | FSK_FLAG_PHI_S)) == (FSK_FLAG_LAPLACIAN | FSK_FLAG_PHI_S)) {
    set_error("fsk_kde1_splitk: flags 0x%x not available on the 1-D path", flags);
    return FSK_EUNSUPPORTED;
  }
  return kde1_run("fsk_kde1_splitk", (flags & FSK_FLAG_LAPLACIAN) ? 1 : ((flags & FSK_FLAG_PHI_S) ? 3 : 0), false, data, n,
                  query, m, bandwidth, ws, ws_bytes, out, nullptr, stream);
}

extern "C" int fsk_kde1_atomic(const float *data, int64_t n, const float *query, int64_t m,
                               double bandwidth, unsigned flags, void *ws, size_t ws_bytes,
                               float *out, void *stream) {
  if ((flags & ~(FSK_FLAG_LAPLACIAN | FSK_FLAG_PHI_S)) ||
      (flags & (FSK_FLAG_LAPLACIAN | FSK_FLAG_PHI_S)) == (FSK_FLAG_LAPLACIAN | FSK_FLAG_PHI_S)) {
    set_error("fsk_kde1_atomic: flags 0x%x not available on the 1-D path", flags);
    return FSK_EUNSUPPORTED;
  }
  return kde1_run("fsk_kde1_atomic", (flags & FSK_FLAG_LAPLACIAN) ? 1 : ((flags & FSK_FLAG_PHI_S) ? 3 : 0), true, data, n, query,
                  m, bandwidth, ws, ws_bytes, out, nullptr, stream);
}

extern "C" int fsk_kde1_score(const float *data, int64_t n, const float *query, int64_t m,
                              double bandwidth, void *ws, size_t ws_bytes, float *pdf_sum,
                              float *deriv_sum, void *stream) {
  return kde1_run("fsk_kde1_score", 2, false, data, n, query, m, bandwidth, ws, ws_bytes,
                  pdf_sum, deriv_sum, stream);
}
