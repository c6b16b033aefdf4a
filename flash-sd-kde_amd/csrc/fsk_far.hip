// fsk_far.hip - far-block bitmap pre-pass (frame, boxes, bits); the bound
// and the layout are described in fsk_far.h.
#include "fsk_far.h"
#include "fsk_lowp.h"
#include "fsk_plan.h"

namespace fsk {

// A double from lane perm(l) of the same 16-lane row, by DPP on both halves
// (CTRL: 0xB1 / 0x4E swap lanes 1 / 2 apart in each quad, 0x141 mirrors each
// half-row, 0x140 the row).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp(int(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(int(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Sum over the 16 lanes of a lane group (every lane gets the same total:
// each step adds two values that their partner lane adds in the other order).
__device__ __forceinline__ double sum16(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  return v + dpp_f64<0x140>(v);
}

// One workgroup of 256: the covariance of a strided sample of the centred
// data in FP64 (fixed order), then 40 rounds of block power iteration with
// modified Gram-Schmidt for its kBoxAxes leading eigenvectors, on the first
// wave: lane f (of each 16-lane group) holds coordinate f of every iterate,
// and the dot products are 16-lane sums. frame[a * 16 + f]; a direction that
// vanishes (rank-deficient data) is left zero, which keeps the rows
// orthonormal or zero. sd (nullable): the sample's standard deviation along
// each row, sqrt(q^T C q / count).
__global__ __launch_bounds__(256) void frame_kernel(const float *__restrict__ x, int64_t ld,
                                                    int64_t n, const double *__restrict__ mu_sum,
                                                    double inv_n, float *__restrict__ frame,
                                                    float *__restrict__ sd) {
  __shared__ float pts[256 * kD];
  __shared__ double cov[kD * kD];
  const int t = threadIdx.x;
  const int64_t cnt = n < kFrameSample ? n : kFrameSample;
  const int64_t stride = n / cnt;
  const int ei = t < kD * kD ? t / kD : 0, ej = t < kD * kD ? t % kD : 0;
  double acc = 0.0;
  for (int64_t c0 = 0; c0 < cnt; c0 += 256) {
    __syncthreads();
    for (int e = t; e < 256 * kD; e += 256) {
      const int64_t k = c0 + e / kD;
      const int f = e % kD;
      pts[e] = k < cnt ? float(double(x[k * stride * ld + f]) - mu_sum[f] * inv_n) : 0.f;
    }
    __syncthreads();
    if (t < kD * kD)
      for (int k = 0; k < 256; ++k) acc += double(pts[k * kD + ei]) * double(pts[k * kD + ej]);
  }
  if (t < kD * kD) cov[t] = acc;
  __syncthreads();
  if (t >= 64) return;
  const int f = t & (kD - 1);
  double crow[kD], q[kBoxAxes];
#pragma unroll
  for (int g = 0; g < kD; ++g) crow[g] = cov[f * kD + g];
#pragma unroll
  for (int a = 0; a < kBoxAxes; ++a) q[a] = f == a ? 1.0 : 0.01 * double((7 * f + 3 * a) % 11 + 1);
  for (int it = 0; it < 40; ++it) {
    double z[kBoxAxes];
#pragma unroll
    for (int a = 0; a < kBoxAxes; ++a) {
      double v = 0.0;
#pragma unroll
      for (int g = 0; g < kD; ++g) v += crow[g] * __shfl(q[a], g, 16);
      z[a] = v;
    }
#pragma unroll
    for (int a = 0; a < kBoxAxes; ++a) {
#pragma unroll
      for (int b = 0; b < a; ++b) z[a] -= sum16(z[a] * z[b]) * z[b];
      const double nn = sum16(z[a] * z[a]);
      z[a] *= nn > 1e-300 && nn < 1e300 ? 1.0 / sqrt(nn) : 0.0;
    }
#pragma unroll
    for (int a = 0; a < kBoxAxes; ++a) q[a] = z[a];
  }
  if (t < kD)
#pragma unroll
    for (int a = 0; a < kBoxAxes; ++a) frame[a * kD + f] = float(q[a]);
  if (sd) {
#pragma unroll
    for (int a = 0; a < kBoxAxes; ++a) {
      double v = 0.0;
#pragma unroll
      for (int g = 0; g < kD; ++g) v += crow[g] * __shfl(q[a], g, 16);
      const double lam = sum16(q[a] * v) / double(cnt);
      if (t == 0) sd[a] = float(sqrt(lam > 0.0 ? lam : 0.0));
    }
  }
}

// Boxes of consecutive groups of g points (tile pairs: 32; row blocks: 64):
// the lowest and highest projection on each frame axis, the largest
// c |x - mu|^2 (log2 units) over the group's points j < n, and the group's
// threshold: thr, or with lower (per point) the smallest lower[j] - margin.
// A group with no points gets an empty box (lo = +inf, hi = -inf), which
// every bound finds far.
__global__ void box_kernel(const float *__restrict__ x, int64_t ld, int64_t n, int64_t groups,
                           int g, const double *__restrict__ mu_sum, double inv_n, double c,
                           const float *__restrict__ frame, float thr,
                           const float *__restrict__ lower, float margin,
                           float *__restrict__ boxes) {
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= groups) return;
  float lo[kBoxAxes], hi[kBoxAxes], nrm = 0.f, t = lower ? INFINITY : thr;
  for (int a = 0; a < kBoxAxes; ++a) {
    lo[a] = INFINITY;
    hi[a] = -INFINITY;
  }
  for (int64_t j = b * g; j < b * g + g && j < n; ++j) {
    float y[kD];
    double s2 = 0.0;
    for (int f = 0; f < kD; ++f) {
      const double v = double(x[j * ld + f]) - mu_sum[f] * inv_n;
      y[f] = float(v);
      s2 += v * v;
    }
    for (int a = 0; a < kBoxAxes; ++a) {
      float pa = 0.f;
      for (int f = 0; f < kD; ++f) pa += frame[a * kD + f] * y[f];
      lo[a] = fminf(lo[a], pa);
      hi[a] = fmaxf(hi[a], pa);
    }
    nrm = fmaxf(nrm, float(c * s2));
    if (lower) t = fminf(t, lower[j] - margin);
  }
  float *o = boxes + b * kBoxFloats;
  for (int a = 0; a < kBoxAxes; ++a) {
    o[a] = lo[a];
    o[kBoxAxes + a] = hi[a];
  }
  o[kBoxNorm] = nrm;
  o[kBoxThr] = t;
}

// bits[rb * words + p / 64] bit p % 64 = 1 when every exponent of row block
// rb against tile pair p is provably <= the row block's threshold (or p has
// no points), for the
// `waves` row blocks of workgroup g; row groups * waves + g holds their AND
// (far for every wave of the workgroup). One thread per (g, p); a wave
// covers 64 consecutive p of one g.
__global__ void far_bits_kernel(const float *__restrict__ row_boxes,
                                const float *__restrict__ pair_boxes, int64_t npairs,
                                int64_t words, int64_t groups, int waves, float c,
                                unsigned long long *__restrict__ bits) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= groups * words * 64) return;  // a multiple of 64: whole waves exit
  // g is the same on the whole wave: scalar loads of the row boxes
  const int64_t g = __builtin_amdgcn_readfirstlane(int(e / (words * 64)));
  const int64_t p = e % (words * 64);
  float plo[kBoxAxes], phi[kBoxAxes], pn = 0.f;
  const bool real = p < npairs;
  for (int a = 0; a < kBoxAxes; ++a) {
    plo[a] = real ? pair_boxes[p * kBoxFloats + a] : 0.f;
    phi[a] = real ? pair_boxes[p * kBoxFloats + kBoxAxes + a] : 0.f;
  }
  if (real) pn = pair_boxes[p * kBoxFloats + kBoxNorm];
  unsigned long long all = ~0ull;
  for (int k = 0; k < waves; ++k) {
    const float *r = row_boxes + (g * waves + k) * kBoxFloats;
    bool far = true;
    if (real) {
      float d2 = 0.f;
      for (int a = 0; a < kBoxAxes; ++a) {
        const float gap = fmaxf(fmaxf(plo[a] - r[kBoxAxes + a], r[a] - phi[a]), 0.f);
        d2 += gap * gap;
      }
      const float bound = -c * d2 + 0x1p-10f * (r[kBoxNorm] + pn) + 1.f;
      far = !(bound > r[kBoxThr]);
    }
    const unsigned long long m = __builtin_amdgcn_ballot_w64(far);
    all &= m;
    if ((threadIdx.x & 63) == 0) bits[(g * waves + k) * words + p / 64] = m;
  }
  if ((threadIdx.x & 63) == 0) bits[(groups * waves + g) * words + p / 64] = all;
}

void launch_far_bits(const float *data, int64_t ld, int64_t n, const float *rows, int64_t ldr,
                     int64_t m, const double *mu_sum, double inv_n, double c, int64_t npairs,
                     int64_t groups, int waves, float skip_t, const float *row_lower,
                     float margin, float *frame, float *pair_boxes, float *row_boxes,
                     unsigned long long *bits, hipStream_t s) {
  const int64_t row_blocks = groups * waves;
  const int64_t words = far_words(npairs);
  hipLaunchKernelGGL(frame_kernel, dim3(1), dim3(256), 0, s, data, ld, n, mu_sum, inv_n, frame,
                     (float *)nullptr);
  hipLaunchKernelGGL(box_kernel, dim3(cdiv(npairs, 256)), dim3(256), 0, s, data, ld, n, npairs,
                     kFarPairPts, mu_sum, inv_n, c, frame, 0.f, (const float *)nullptr, 0.f,
                     pair_boxes);
  hipLaunchKernelGGL(box_kernel, dim3(cdiv(row_blocks, 256)), dim3(256), 0, s, rows, ldr, m,
                     row_blocks, kFarRows, mu_sum, inv_n, c, frame, skip_t, row_lower, margin,
                     row_boxes);
  hipLaunchKernelGGL(far_bits_kernel, dim3(cdiv(groups * words * 64, 256)), dim3(256), 0, s,
                     row_boxes, pair_boxes, npairs, words, groups, waves, float(c), bits);
}

// Morton code of each point: its coordinates along the frame's first three
// rows (the leading axis first), each quantized to `bits` bits over the mean
// +- 4 sd along the row (points beyond fall into the end cells); bit b of the
// coordinate along row r goes to code bit 3b + (2 - r). A row whose sd is
// below 1e-6 of the leading one (rank-deficient data: its coordinates are
// rounding noise) contributes zeros.
__global__ void spatial_codes_kernel(const float *__restrict__ x, int64_t ld, int64_t n,
                                     const double *__restrict__ mu_sum, double inv_n,
                                     const float *__restrict__ frame, const float *__restrict__ sd,
                                     int bits, unsigned *__restrict__ codes) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float y[kD];
#pragma unroll
  for (int f = 0; f < kD; ++f) y[f] = float(double(x[j * ld + f]) - mu_sum[f] * inv_n);
  const float scale = float((1 << bits) - 1);
  unsigned code = 0;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float p = 0.f;
#pragma unroll
    for (int f = 0; f < kD; ++f) p += frame[r * kD + f] * y[f];
    const float half = fmaxf(4.f * sd[r], 1e-30f);
    const unsigned c = sd[r] > 1e-6f * sd[0]
                           ? unsigned(fminf(fmaxf((p + half) * (scale / (2.f * half)), 0.f), scale))
                           : 0u;
    for (int b = 0; b < bits; ++b) code |= ((c >> b) & 1u) << (3 * b + 2 - r);
  }
  codes[j] = code;
}

size_t spatial_ws_bytes() {
  Carve cv(nullptr);
  cv.take<double>(kMeanWs);
  cv.take<float>(kBoxAxes * kD);
  cv.take<float>(kBoxAxes);
  return cv.used;
}

}  // namespace fsk

extern "C" int fsk_spatial_codes16(const float *x, int64_t ld, int64_t n, int bits,
                                   uint32_t *codes, void *workspace, size_t workspace_bytes,
                                   void *stream) {
  using namespace fsk;
  if (!x || !codes || n <= 0 || ld < kD || bits < 1 || bits > 10) {
    set_error("fsk_spatial_codes16: bad arguments (n=%lld ld=%lld bits=%d)", (long long)n,
              (long long)ld, bits);
    return FSK_EINVAL;
  }
  const size_t need = spatial_ws_bytes();
  if (!workspace || workspace_bytes < need) {
    set_error("fsk_spatial_codes16: workspace of %zu bytes needed, got %zu", need,
              workspace_bytes);
    return FSK_EWORKSPACE;
  }
  Carve cv(workspace);
  double *mu = cv.take<double>(kMeanWs);
  float *frame = cv.take<float>(kBoxAxes * kD);
  float *sd = cv.take<float>(kBoxAxes);
  hipStream_t s = as_stream(stream);
  launch_mean16(x, ld, n, mu, s);
  hipLaunchKernelGGL(frame_kernel, dim3(1), dim3(256), 0, s, x, ld, n, mu, 1.0 / double(n), frame,
                     sd);
  hipLaunchKernelGGL(spatial_codes_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, x, ld, n, mu,
                     1.0 / double(n), frame, sd, bits, codes);
  return launch_status("fsk_spatial_codes16");
}
