// fsk_score16_valu.hip - 16-D empirical-score accumulators with the cross term
// on the vector ALU (FSK_PREC_F32_VALU): the FP32-VALU form of the precision
// configuration, beside the FP32 MFMA form (fsk_score16.hip) and the 16-bit
// MFMA forms (fsk_score16_lp.hip).
//
// One thread per row. Data points (centred on the data mean) are staged in
// LDS chunks and read back as wave-wide broadcasts. Per pair, in FP32:
//   D   = |x_j|^2 + sum_f (-2 q_f) x_jf        (16 FMAs, from the data-side norm)
//   t   = c D + c|q|^2                          (one FMA; c = -log2(e)/(2h^2))
//   phi = exp2(min(t, 0)),  pdf += phi,  w_f += phi x_jf   (16 FMAs)
// the operation order of the FP32 MFMA KDE (fsk_kde16.hip), on FMAs instead
// of v_mfma_f32_16x16x4_f32. FP32 sums over one chunk, FP64 across chunks,
// FP32 partials per data split and the FP64 finish of the 16-bit forms
// (score_lp_finish_kernel, which adds mu pdf back).
#include "fsk_common.h"
#include "fsk_lowp.h"
#include "fsk_plan.h"

namespace fsk {

constexpr int kValuScoreChunk = 128;  // data points per LDS round
constexpr int kValuScoreRows = 256;   // rows (threads) per workgroup

__global__ __launch_bounds__(kValuScoreRows) void score16_valu_kernel(
    const float *__restrict__ xc, const float *__restrict__ bn, int64_t n,
    const float *__restrict__ rc, const float *__restrict__ qa, int64_t m, float coef,
    float pad, int64_t split_len, int64_t qblocks, int64_t ntasks, int64_t per_xcd,
    float *__restrict__ pdf_part, float *__restrict__ w_part) {
  __shared__ f4 xs[kValuScoreChunk][kD / 4];
  __shared__ float bs[kValuScoreChunk];
  const int64_t task = xcd_task(blockIdx.x, per_xcd);
  if (task >= ntasks) return;
  const int64_t split = task / qblocks;
  const int64_t qb = task - split * qblocks;
  const int tid = threadIdx.x;
  const int64_t i = qb * kValuScoreRows + tid;
  const int64_t ic = i < m ? i : m - 1;
  float q[kD];
#pragma unroll
  for (int f = 0; f < kD; f += 4) {
    const f4 v = ld4(rc + ic * kD + f) * -2.f;
    q[f] = v[0];
    q[f + 1] = v[1];
    q[f + 2] = v[2];
    q[f + 3] = v[3];
  }
  const float a = qa[ic];
  float p = 0.f, w[kD];
  double p64 = 0.0, w64[kD];
#pragma unroll
  for (int f = 0; f < kD; ++f) {
    w[f] = 0.f;
    w64[f] = 0.0;
  }
  const int64_t j_begin = split * split_len;
  int64_t j_end = j_begin + split_len;
  if (j_end > n) j_end = n;
  for (int64_t j0 = j_begin; j0 < j_end; j0 += kValuScoreChunk) {
    __syncthreads();
    // 128 points x 4 float4 = 512 slots, two per thread; rows past n are
    // zero points with the pad norm (c pad = -1e30: phi = 0)
    for (int e = tid; e < kValuScoreChunk * 4; e += kValuScoreRows) {
      const int64_t j = j0 + (e >> 2);
      xs[e >> 2][e & 3] = j < j_end ? ld4(xc + j * kD + 4 * (e & 3)) : f4{0.f, 0.f, 0.f, 0.f};
    }
    if (tid < kValuScoreChunk) bs[tid] = j0 + tid < j_end ? bn[j0 + tid] : pad;
    __syncthreads();
    for (int jj = 0; jj < kValuScoreChunk; ++jj) {
      float x[kD];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const f4 xv = xs[jj][v];
        x[4 * v] = xv[0];
        x[4 * v + 1] = xv[1];
        x[4 * v + 2] = xv[2];
        x[4 * v + 3] = xv[3];
      }
      float d = bs[jj];
#pragma unroll
      for (int f = 0; f < kD; ++f) d = fmaf(q[f], x[f], d);
      const float e = fexp2_le1(fmaf(d, coef, a));
      p += e;
#pragma unroll
      for (int f = 0; f < kD; ++f) w[f] = fmaf(e, x[f], w[f]);
    }
    p64 += p;
    p = 0.f;
#pragma unroll
    for (int f = 0; f < kD; ++f) {
      w64[f] += w[f];
      w[f] = 0.f;
    }
  }
  if (i < m) {
    pdf_part[split * m + i] = float(p64);
#pragma unroll
    for (int f = 0; f < kD; ++f) w_part[(split * m + i) * kD + f] = float(w64[f]);
  }
}

struct ValuScoreLayout {
  double *mu;
  float *xc, *rc, *bn, *qa, *pdf_part, *w_part;
};

static SplitPlan score16_valu_plan(int64_t n, int64_t m) {
  int64_t cap = kOrderedPartialCap / (m * 17 * 4 > 0 ? m * 17 * 4 : 1);
  if (cap < 1) cap = 1;
  return make_split_plan(n, m, kValuScoreRows, kValuScoreChunk, 2048, cap);
}

// same_rows: the rows are the data set itself (one centred copy serves both)
static size_t valu_layout(void *ws, int64_t n, int64_t m, bool same_rows, ValuScoreLayout *L) {
  Carve cv(ws);
  const SplitPlan p = score16_valu_plan(n, m);
  L->mu = cv.take<double>(kMeanWs);
  L->xc = cv.take<float>(round_up(n, kTile) * kD);
  L->rc = same_rows ? nullptr : cv.take<float>(m * kD);
  L->bn = cv.take<float>(round_up(n, kNormPad));
  L->qa = cv.take<float>(round_up(m, kNormPad));
  L->pdf_part = cv.take<float>(p.splits * m);
  L->w_part = cv.take<float>(p.splits * m * kD);
  return cv.used;
}

size_t score16_valu_ws_bytes(int64_t n, int64_t m, bool same_rows) {
  ValuScoreLayout L;
  return valu_layout(nullptr, n, m, same_rows, &L);
}

int score16_valu_run(const char *what, const float *rows, int64_t ldr, int64_t m,
                     const float *data, int64_t ld, int64_t n, double h, bool same_rows, void *ws,
                     size_t ws_bytes, float *pdf, float *w, int64_t ldw_row, int64_t ldw_col,
                     void *stream) {
  ValuScoreLayout L;
  const size_t need = valu_layout(ws, n, m, same_rows, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const double coef = exp2_coef(h);
  const double inv_n = 1.0 / double(n);
  const int64_t n16 = round_up(n, kTile), n_norm = round_up(n, kNormPad);
  const int64_t m_norm = round_up(m, kNormPad);
  const float pad = float(fmin(1e30 / -coef, 1e38));
  launch_mean16(data, ld, n, L.mu, s);
  hipLaunchKernelGGL(center16_kernel, dim3(cdiv(n16 * kD, 256)), dim3(256), 0, s, data, ld, n,
                     n16, L.mu, inv_n, L.xc);
  const float *rc = L.xc;
  if (!same_rows) {
    hipLaunchKernelGGL(center16_kernel, dim3(cdiv(m * kD, 256)), dim3(256), 0, s, rows, ldr, m,
                       m, L.mu, inv_n, L.rc);
    rc = L.rc;
  }
  hipLaunchKernelGGL(prep_norms16, dim3(cdiv(n_norm, 256)), dim3(256), 0, s, L.xc, int64_t(kD), n,
                     n_norm, 1.0, pad, L.bn);
  hipLaunchKernelGGL(prep_norms16, dim3(cdiv(m_norm, 256)), dim3(256), 0, s, rc, int64_t(kD), m,
                     m_norm, coef, kPadNeg, L.qa);
  const SplitPlan p = score16_valu_plan(n, m);
  const int64_t ntasks = p.qblocks * p.splits;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
  prof_begin(FSK_OP_SCORE16_ORDERED, s);
  hipLaunchKernelGGL(score16_valu_kernel, dim3(grid), dim3(kValuScoreRows), 0, s, L.xc, L.bn, n,
                     rc, L.qa, m, float(coef), pad, p.split_len, p.qblocks, ntasks, per_xcd,
                     L.pdf_part, L.w_part);
  prof_end(FSK_OP_SCORE16_ORDERED, s);
  int st = launch_status(what);
  if (st != FSK_OK) return st;
  hipLaunchKernelGGL(score_lp_finish_kernel, dim3(cdiv(m * kD, 256)), dim3(256), 0, s,
                     L.pdf_part, L.w_part, p.splits, m, L.mu, inv_n, 1.0, pdf, w, ldw_row,
                     ldw_col);
  return launch_status(what);
}

}  // namespace fsk
