// fsk_ref64.hip - FP64 reference kernels for measuring the error of the FP32
// and reduced-precision forms at sizes the CPU oracle cannot reach.
//
// Direct differences sum_f (q_f - x_f)^2 in FP64 (no norm expansion), FP64
// exp, FP64 accumulation. One thread per query / row; data staged through LDS.
//   ref64_logdensity: log p(q) by log-sum-exp, per (query, split) running max,
//                     FP64 combine (the semantics of the reference error
//                     suite's oracle, experiments/error_suite_16d/reference.py:32-82)
//   ref64_score     : pdf[i] = sum_j phi_ij, w[i,:] = sum_j phi_ij x_j over all j.
#include "fsk_common.h"

namespace fsk {

constexpr int kRefChunk = 128;

__global__ __launch_bounds__(256) void ref64_logdensity_kernel(
    const double *__restrict__ x, int64_t n, const double *__restrict__ q, int64_t m,
    double inv_2h2, int64_t split_len, int64_t qblocks, double2 *__restrict__ part) {
  __shared__ double xs[kRefChunk][kD + 1];
  const int64_t split = blockIdx.x / qblocks;
  const int64_t qb = blockIdx.x - split * qblocks;
  const int tid = threadIdx.x;
  const int64_t i = qb * 256 + tid;
  const int64_t ic = i < m ? i : m - 1;
  double qq[kD];
#pragma unroll
  for (int f = 0; f < kD; ++f) qq[f] = q[ic * kD + f];
  double mx = -INFINITY, s = 0.0;
  const int64_t j_begin = split * split_len;
  const int64_t j_end = (j_begin + split_len < n) ? j_begin + split_len : n;
  for (int64_t j0 = j_begin; j0 < j_end; j0 += kRefChunk) {
    const int cnt = int((j_end - j0) < kRefChunk ? (j_end - j0) : kRefChunk);
    __syncthreads();
    for (int e = tid; e < cnt * kD; e += 256) xs[e / kD][e % kD] = x[(j0 + e / kD) * kD + e % kD];
    __syncthreads();
    for (int jj = 0; jj < cnt; ++jj) {
      double d2 = 0.0;
#pragma unroll
      for (int f = 0; f < kD; ++f) {
        const double t = qq[f] - xs[jj][f];
        d2 = fma(t, t, d2);
      }
      const double e = -d2 * inv_2h2;
      if (e > mx) {
        s = s * exp(mx - e) + 1.0;
        mx = e;
      } else {
        s += exp(e - mx);
      }
    }
  }
  if (i < m) part[split * m + i] = make_double2(mx, s);
}

__global__ void ref64_lse_combine_kernel(const double2 *__restrict__ part, int64_t S, int64_t m,
                                         double log_const, double *__restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double M = -INFINITY;
  for (int64_t s = 0; s < S; ++s) M = fmax(M, part[s * m + i].x);
  double acc = 0.0;
  for (int64_t s = 0; s < S; ++s) {
    const double2 p = part[s * m + i];
    if (p.y > 0.0) acc += p.y * exp(p.x - M);
  }
  out[i] = M + log(acc) + log_const;
}

__global__ __launch_bounds__(256) void ref64_score_kernel(const double *__restrict__ x, int64_t n,
                                                          double inv_2h2, double *__restrict__ pdf,
                                                          double *__restrict__ w) {
  __shared__ double xs[kRefChunk][kD + 1];
  const int tid = threadIdx.x;
  const int64_t i = int64_t(blockIdx.x) * 256 + tid;
  const int64_t ic = i < n ? i : n - 1;
  double xi[kD], acc[kD];
#pragma unroll
  for (int f = 0; f < kD; ++f) {
    xi[f] = x[ic * kD + f];
    acc[f] = 0.0;
  }
  double p = 0.0;
  for (int64_t j0 = 0; j0 < n; j0 += kRefChunk) {
    const int cnt = int((n - j0) < kRefChunk ? (n - j0) : kRefChunk);
    __syncthreads();
    for (int e = tid; e < cnt * kD; e += 256) xs[e / kD][e % kD] = x[(j0 + e / kD) * kD + e % kD];
    __syncthreads();
    for (int jj = 0; jj < cnt; ++jj) {
      double d2 = 0.0;
#pragma unroll
      for (int f = 0; f < kD; ++f) {
        const double t = xi[f] - xs[jj][f];
        d2 = fma(t, t, d2);
      }
      const double e = exp(-d2 * inv_2h2);
      p += e;
#pragma unroll
      for (int f = 0; f < kD; ++f) acc[f] = fma(e, xs[jj][f], acc[f]);
    }
  }
  if (i < n) {
    pdf[i] = p;
#pragma unroll
    for (int f = 0; f < kD; ++f) w[i * kD + f] = acc[f];
  }
}

}  // namespace fsk

using namespace fsk;

extern "C" size_t fsk_ref64_workspace_bytes(int64_t n, int64_t m) {
  if (n <= 0 || m <= 0) return 0;
  const int64_t split_len = 16384;
  return size_t(cdiv(n, split_len)) * size_t(m) * sizeof(double2);
}

extern "C" int fsk_ref64_logdensity16(const double *data, int64_t n, const double *query,
                                      int64_t m, double bandwidth, void *ws, size_t ws_bytes,
                                      double *out, void *stream) {
  if (!data || !query || !out || n <= 0 || m <= 0 || !(bandwidth > 0.0)) {
    set_error("fsk_ref64_logdensity16: bad arguments");
    return FSK_EINVAL;
  }
  const int64_t split_len = 16384;
  const int64_t S = cdiv(n, split_len);
  if (!ws || ws_bytes < fsk_ref64_workspace_bytes(n, m)) {
    set_error("fsk_ref64_logdensity16: workspace too small");
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const int64_t qblocks = cdiv(m, 256);
  double2 *part = static_cast<double2 *>(ws);
  hipLaunchKernelGGL(ref64_logdensity_kernel, dim3(qblocks * S), dim3(256), 0, s, data, n, query, m,
                     1.0 / (2.0 * bandwidth * bandwidth), split_len, qblocks, part);
  const double log_const = -log(double(n)) - kD * log(bandwidth) - 0.5 * kD * log(2.0 * M_PI);
  hipLaunchKernelGGL(ref64_lse_combine_kernel, dim3(cdiv(m, 256)), dim3(256), 0, s, part, S, m,
                     log_const, out);
  return launch_status("fsk_ref64_logdensity16");
}

extern "C" int fsk_ref64_score16(const double *data, int64_t n, double bandwidth, double *pdf,
                                 double *w, void *stream) {
  if (!data || !pdf || !w || n <= 0 || !(bandwidth > 0.0)) {
    set_error("fsk_ref64_score16: bad arguments");
    return FSK_EINVAL;
  }
  hipLaunchKernelGGL(ref64_score_kernel, dim3(cdiv(n, 256)), dim3(256), 0, as_stream(stream), data,
                     n, 1.0 / (2.0 * bandwidth * bandwidth), pdf, w);
  return launch_status("fsk_ref64_score16");
}
