// fsk_nd.hip - Gaussian KDE sums, empirical-score accumulators and the FP64
// debias for any dimension d <= 64, by zero padding to ND = ceil(d/16)
// slices of 16 features.
//
// Zero coordinates leave every pairwise distance unchanged, so the padded sums
// are the d-dimensional sums exactly. The points are centred on the data mean
// into a padded copy (DP = 16 ND floats per row), the MFMA cross term runs
// over all 4 ND k-steps of the slices in one accumulator chain started from
// |x_j|^2, and t = c D + c|q_i|^2 is one FMA, as in the 16-D KDE
// (fsk_kde16.hip). The score runs the row side of the ordered split-K kernel
// with one phi . X_J accumulator per slice; per-split FP32 partials are
// combined in FP64. The 16-D kernels stay the ones used at d = 16.
#include "fsk_common.h"
#include "fsk_plan.h"
#include "fsk_tile.h"

namespace fsk {

constexpr int kNdMaxSlices = FSK_ND_MAX_DIM / kD;
constexpr int kNdMeanBlocks = 256;
constexpr int kNdFlush = 32;  // tiles between FP64 flushes of the FP32 sums

inline int nd_slices(int d) { return (d + kD - 1) / kD; }
// Row tiles per wave: the score keeps ND W accumulators per row tile.
template <int ND>
constexpr int nd_score_qt() { return ND == 1 ? 8 : (ND == 2 ? 4 : 2); }
constexpr int kNdKdeQT = 4;

// Column sums of the first d features, FP64, fixed order: block b reduces a
// strided row set (4 row lanes x 64 feature lanes) into part[b][f]; the
// finish adds the blocks in order.
__global__ __launch_bounds__(256) void mean_nd_kernel(const float *__restrict__ x, int64_t ld,
                                                      int64_t n, int d, double *__restrict__ part) {
  __shared__ double red[4][64];
  const int f = threadIdx.x & 63, lane_row = threadIdx.x >> 6;
  double acc = 0.0;
  if (f < d)
    for (int64_t j = int64_t(blockIdx.x) * 4 + lane_row; j < n; j += int64_t(gridDim.x) * 4)
      acc += double(x[j * ld + f]);
  red[lane_row][f] = acc;
  __syncthreads();
  if (lane_row == 0)
    part[int64_t(blockIdx.x) * 64 + f] = (red[0][f] + red[1][f]) + (red[2][f] + red[3][f]);
}

__global__ void mean_nd_finish_kernel(const double *__restrict__ part, int nblocks, int d,
                                      double *__restrict__ sum) {
  const int f = threadIdx.x;
  if (f >= 64) return;
  double s = 0.0;
  if (f < d)
    for (int b = 0; b < nblocks; ++b) s += part[b * 64 + f];
  sum[f] = s;
}

// xc[j][f] = x[j][f] - mu_f for j < n, f < d; zero elsewhere (DP columns, n_pad rows).
__global__ void center_nd_kernel(const float *__restrict__ x, int64_t ld, int64_t n, int d,
                                 int DP, int64_t n_pad, const double *__restrict__ mu_sum,
                                 double inv_n, float *__restrict__ xc) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_pad * DP) return;
  const int64_t j = e / DP;
  const int f = int(e % DP);
  xc[e] = (j < n && f < d) ? float(double(x[j * ld + f]) - mu_sum[f] * inv_n) : 0.f;
}

// out[j] = coef |xc_j|^2 over DP columns (FP64, one rounding); rows >= n get pad.
__global__ void norms_nd_kernel(const float *__restrict__ xc, int DP, int64_t n, int64_t n_pad,
                                double coef, float pad, float *__restrict__ out) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n_pad) return;
  if (j >= n) {
    out[j] = pad;
    return;
  }
  double s = 0.0;
  for (int f = 0; f < DP; f += 4) {
    const f4 v = ld4(xc + j * DP + f);
    s += double(v[0]) * v[0] + double(v[1]) * v[1] + double(v[2]) * v[2] + double(v[3]) * v[3];
  }
  out[j] = float(coef * s);
}

template <int ND>
struct NdTile {
  f4 xa[ND];   // features 16 s + 4g .. +3 of point j0 + c
  f4 bb;       // |x|^2 of points j0 + 4g .. +3
  float xt[ND][4];  // feature 16 s + c of point j0 + 4g + r
};

// Data rows are the padded centred copy (DP floats, rows up to a multiple of
// the tile, norms up to a multiple of kNormPad): every tile read is in bounds.
template <int ND, bool WITH_XT>
__device__ __forceinline__ void nd_load(const float *__restrict__ xc, const float *__restrict__ bn,
                                        int64_t j0, int g, int c, NdTile<ND> &t) {
  constexpr int DP = kD * ND;
#pragma unroll
  for (int s = 0; s < ND; ++s) t.xa[s] = ld4(xc + (j0 + c) * DP + kD * s + 4 * g);
  t.bb = ld4(bn + j0 + 4 * g);
  if (WITH_XT) {
#pragma unroll
    for (int s = 0; s < ND; ++s)
#pragma unroll
      for (int r = 0; r < 4; ++r) t.xt[s][r] = xc[(j0 + 4 * g + r) * DP + kD * s + c];
  }
}

template <int ND, int QT>
__device__ __forceinline__ void nd_cross(const f4 (&qv)[QT][ND], const NdTile<ND> &t,
                                         f4 (&d)[QT]) {
#pragma unroll
  for (int R = 0; R < QT; ++R) d[R] = mfma_f32(t.xa[0][0], qv[R][0][0], t.bb);
#pragma unroll
  for (int s = 0; s < ND; ++s) {
#pragma unroll
    for (int k = (s == 0 ? 1 : 0); k < 4; ++k) {
#pragma unroll
      for (int R = 0; R < QT; ++R) d[R] = mfma_f32(t.xa[s][k], qv[R][s][k], d[R]);
    }
  }
}

// One workgroup = 4 waves x QT row tiles; task = (row block, data split).
// MODE 0: KDE sums; MODE 1: KDE sums with the Laplacian factor
// (1 + d/2 - |q-x|^2/(2h^2)); MODE 2: score row side (pdf and W).
template <int ND, int QT, int MODE>
__global__ __launch_bounds__(kWG) void nd_pair_kernel(
    const float *__restrict__ xc, const float *__restrict__ bn, int64_t n_data_pad,
    const float *__restrict__ qc, const float *__restrict__ qa, int64_t m, float coef,
    float lap_base, int64_t split_len, int64_t qblocks, int64_t ntasks, int64_t per_xcd,
    float *__restrict__ out_part, float *__restrict__ w_part) {
  constexpr int DP = kD * ND;
  constexpr bool LAPL = MODE == 1, SCORE = MODE == 2;
  const int64_t task = xcd_task(blockIdx.x, per_xcd);
  if (task >= ntasks) return;
  const int64_t split = task / qblocks;
  const int64_t qb = task - split * qblocks;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t row0 = qb * (64 * QT) + int64_t(wave) * (16 * QT);

  f4 qv[QT][ND];
  float qav[QT], acc[QT], accT[QT];
  double acc64[QT], accT64[QT];
  f4 w[SCORE ? QT : 1][SCORE ? ND : 1];
  d4 w64[SCORE ? QT : 1][SCORE ? ND : 1];
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    int64_t i = row0 + 16 * R + c;
    if (i > m - 1) i = m - 1;
#pragma unroll
    for (int s = 0; s < ND; ++s) qv[R][s] = ld4(qc + i * DP + kD * s + 4 * g) * -2.f;
    qav[R] = qa[i];
    acc[R] = accT[R] = 0.f;
    acc64[R] = accT64[R] = 0.0;
  }
  if (SCORE) {
#pragma unroll
    for (int R = 0; R < QT; ++R)
#pragma unroll
      for (int s = 0; s < ND; ++s) {
        w[R][s] = f4{0.f, 0.f, 0.f, 0.f};
        w64[R][s] = d4{0.0, 0.0, 0.0, 0.0};
      }
  }
  const int64_t j_begin = split * split_len;
  int64_t j_end = j_begin + split_len;
  if (j_end > n_data_pad) j_end = n_data_pad;
  int tc = 0;
  for (int64_t j0 = j_begin; j0 < j_end; j0 += kTile) {
    NdTile<ND> t;
    nd_load<ND, SCORE>(xc, bn, j0, g, c, t);
    f4 d[QT];
    nd_cross<ND, QT>(qv, t, d);
    kernel_tile<QT, LAPL, true>(d, qav, coef, acc, accT);
    if (SCORE) {
#pragma unroll
      for (int s = 0; s < ND; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int R = 0; R < QT; ++R) w[R][s] = mfma_f32(d[R][r], t.xt[s][r], w[R][s]);
    }
    if (++tc == kNdFlush) {
      tc = 0;
#pragma unroll
      for (int R = 0; R < QT; ++R) {
        acc64[R] += acc[R];
        acc[R] = 0.f;
        if (LAPL) {
          accT64[R] += accT[R];
          accT[R] = 0.f;
        }
        if (SCORE) {
#pragma unroll
          for (int s = 0; s < ND; ++s) {
#pragma unroll
            for (int r = 0; r < 4; ++r) w64[R][s][r] += w[R][s][r];
            w[R][s] = f4{0.f, 0.f, 0.f, 0.f};
          }
        }
      }
    }
  }
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    double v = acc64[R] + acc[R];
    if (LAPL) v = double(lap_base) * v + kLn2 * (accT64[R] + accT[R]);
    v = sum_groups(v);
    const int64_t i = row0 + 16 * R + c;
    if (g == 0 && i < m) out_part[split * m + i] = float(v);
    if (SCORE) {
#pragma unroll
      for (int s = 0; s < ND; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t iw = row0 + 16 * R + 4 * g + r;
          if (iw < m) w_part[(split * m + iw) * DP + kD * s + c] = float(w64[R][s][r] + w[R][s][r]);
        }
    }
  }
}

// pdf[i] = sum_s pdf_part; w[i, f] = sum_s w_part + mu_f pdf[i] for f < d (FP64).
__global__ void score_nd_finish_kernel(const float *__restrict__ pdf_part,
                                       const float *__restrict__ w_part, int64_t S, int64_t n,
                                       int d, int DP, const double *__restrict__ mu_sum,
                                       double inv_n, float *__restrict__ pdf,
                                       float *__restrict__ w, int64_t ldw_row, int64_t ldw_col) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * d) return;
  const int64_t i = e / d;
  const int f = int(e % d);
  double p = 0.0, s = 0.0;
  for (int64_t k = 0; k < S; ++k) {
    p += double(pdf_part[k * n + i]);
    s += double(w_part[(k * n + i) * DP + f]);
  }
  w[i * ldw_row + f * ldw_col] = float(s + mu_sum[f] * inv_n * p);
  if (f == 0) pdf[i] = float(p);
}

__global__ void debias_nd_f64_kernel(const float *__restrict__ x, int64_t ldx, int64_t n, int d,
                                     const float *__restrict__ pdf, const float *__restrict__ w,
                                     int64_t ldw_row, int64_t ldw_col, double inv_h2,
                                     double delta, double eps, float *__restrict__ out,
                                     int64_t ldo) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * d) return;
  const int64_t i = e / d;
  const int f = int(e % d);
  const double xv = x[i * ldx + f];
  const double s = (double(w[i * ldw_row + f * ldw_col]) / (double(pdf[i]) + eps) - xv) * inv_h2;
  out[i * ldo + f] = float(xv + delta * s);
}

struct NdLayout {
  double *mu;     // 64 sums, then kNdMeanBlocks x 64 block parts
  float *xc, *qc; // centred padded copies
  float *bn, *qa;
  float *part, *w_part;
};

static SplitPlan nd_plan(int64_t n, int64_t m, int ND, bool score) {
  const int qt = score ? (ND == 1 ? 8 : (ND == 2 ? 4 : 2)) : kNdKdeQT;
  int64_t cap = 1 << 20;
  if (score) {
    cap = kOrderedPartialCap / (m * (kD * ND + 1) * 4 > 0 ? m * (kD * ND + 1) * 4 : 1);
    if (cap < 1) cap = 1;
  }
  return make_split_plan(n, m, 64 * qt, kTile, score ? 2048 : 1024, cap);
}

// rows = data for the score (m = n, one centred copy serves both sides)
static size_t nd_layout(void *ws, int64_t n, int64_t m, int d, bool score, NdLayout *L) {
  const int ND = nd_slices(d), DP = kD * ND;
  Carve cv(ws);
  SplitPlan p = nd_plan(n, m, ND, score);
  L->mu = cv.take<double>(64 * (1 + kNdMeanBlocks));
  L->xc = cv.take<float>(round_up(n, kNormPad) * DP);
  L->qc = score ? nullptr : cv.take<float>(m * DP);
  L->bn = cv.take<float>(round_up(n, kNormPad));
  L->qa = cv.take<float>(round_up(m, kNormPad));
  L->part = cv.take<float>(p.splits * m);
  L->w_part = score ? cv.take<float>(p.splits * m * DP) : nullptr;
  return cv.used;
}

static int check_nd(const char *what, const float *p, int64_t ld, int64_t rows, int d) {
  if (d < 1 || d > FSK_ND_MAX_DIM) {
    set_error("%s: dimension must be in 1..%d, got %d", what, FSK_ND_MAX_DIM, d);
    return FSK_EINVAL;
  }
  if (rows <= 0 || !p) {
    set_error("%s: need a non-null array with at least one row", what);
    return FSK_EINVAL;
  }
  if (ld < d) {
    set_error("%s: row stride %lld below the dimension %d", what, (long long)ld, d);
    return FSK_EINVAL;
  }
  return FSK_OK;
}

// Data mean (FP64 sums into mu[0..63]) and the centred padded copy of the data.
static void nd_center(const float *x, int64_t ld, int64_t n, int d, int DP, int64_t n_pad,
                      double *mu, float *out, hipStream_t s) {
  int64_t blocks = cdiv(n, 4);
  if (blocks > kNdMeanBlocks) blocks = kNdMeanBlocks;
  hipLaunchKernelGGL(mean_nd_kernel, dim3(blocks), dim3(256), 0, s, x, ld, n, d, mu + 64);
  hipLaunchKernelGGL(mean_nd_finish_kernel, dim3(1), dim3(64), 0, s, mu + 64, int(blocks), d, mu);
  hipLaunchKernelGGL(center_nd_kernel, dim3(cdiv(n_pad * DP, 256)), dim3(256), 0, s, x, ld, n,
                     d, DP, n_pad, mu, 1.0 / double(n), out);
}

template <int ND, int MODE>
static void nd_launch(const NdLayout &L, int64_t n_data_pad, int64_t m, float coef,
                      float lap_base, const SplitPlan &p, hipStream_t s) {
  constexpr int QT = MODE == 2 ? nd_score_qt<ND>() : kNdKdeQT;
  const int64_t ntasks = p.qblocks * p.splits;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
  hipLaunchKernelGGL((nd_pair_kernel<ND, QT, MODE>), dim3(grid), dim3(kWG), 0, s, L.xc, L.bn,
                     n_data_pad, MODE == 2 ? L.xc : L.qc, L.qa, m, coef, lap_base, p.split_len,
                     p.qblocks, ntasks, per_xcd, L.part, L.w_part);
}

template <int MODE>
static void nd_dispatch(int ND, const NdLayout &L, int64_t n_data_pad, int64_t m, float coef,
                        float lap_base, const SplitPlan &p, hipStream_t s) {
  switch (ND) {
    case 1: nd_launch<1, MODE>(L, n_data_pad, m, coef, lap_base, p, s); break;
    case 2: nd_launch<2, MODE>(L, n_data_pad, m, coef, lap_base, p, s); break;
    case 3: nd_launch<3, MODE>(L, n_data_pad, m, coef, lap_base, p, s); break;
    default: nd_launch<4, MODE>(L, n_data_pad, m, coef, lap_base, p, s); break;
  }
}
static_assert(kNdMaxSlices == 4, "nd_dispatch covers 1..4 slices");

}  // namespace fsk

using namespace fsk;

extern "C" size_t fsk_nd_workspace_bytes(int op, int64_t n, int64_t m, int d) {
  if (n <= 0 || d < 1 || d > FSK_ND_MAX_DIM) return 0;
  NdLayout L;
  if (op == FSK_OP_KDE_ND_SPLITK) return m > 0 ? nd_layout(nullptr, n, m, d, false, &L) : 0;
  if (op == FSK_OP_SCORE_ND_ORDERED) return nd_layout(nullptr, n, n, d, true, &L);
  return 0;
}

extern "C" int fsk_kde_nd_splitk(const float *data, int64_t ld, int64_t n, const float *query,
                                 int64_t ldq, int64_t m, int d, double bandwidth, unsigned flags,
                                 void *ws, size_t ws_bytes, float *out, void *stream) {
  const char *what = "fsk_kde_nd_splitk";
  int st;
  if ((st = check_nd(what, data, ld, n, d)) != FSK_OK) return st;
  if ((st = check_nd(what, query, ldq, m, d)) != FSK_OK) return st;
  if (!(bandwidth > 0.0) || !isfinite(bandwidth) || !out) {
    set_error("%s: bandwidth must be positive and the output non-null", what);
    return FSK_EINVAL;
  }
  if (flags & ~FSK_FLAG_LAPLACIAN) {
    set_error("%s: flags 0x%x not available (FP32 only, optional Laplacian)", what, flags);
    return FSK_EUNSUPPORTED;
  }
  NdLayout L;
  const size_t need = nd_layout(ws, n, m, d, false, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const int ND = nd_slices(d), DP = kD * ND;
  const double coef = exp2_coef(bandwidth);
  const int64_t n_norm = round_up(n, kNormPad);
  nd_center(data, ld, n, d, DP, n_norm, L.mu, L.xc, s);
  // queries centred on the data mean (mu_sum / n of the data)
  hipLaunchKernelGGL(center_nd_kernel, dim3(cdiv(m * DP, 256)), dim3(256), 0, s, query, ldq, m,
                     d, DP, m, L.mu, 1.0 / double(n), L.qc);
  hipLaunchKernelGGL(norms_nd_kernel, dim3(cdiv(n_norm, 256)), dim3(256), 0, s, L.xc, DP, n,
                     n_norm, 1.0, float(fmin(1e30 / -coef, 1e38)), L.bn);
  hipLaunchKernelGGL(norms_nd_kernel, dim3(cdiv(m, 256)), dim3(256), 0, s, L.qc, DP, m, m, coef,
                     kPadNeg, L.qa);
  const SplitPlan p = nd_plan(n, m, ND, false);
  const bool lapl = (flags & FSK_FLAG_LAPLACIAN) != 0;
  prof_begin(FSK_OP_KDE_ND_SPLITK, s);
  if (lapl)
    nd_dispatch<1>(ND, L, round_up(n, kTile), m, float(coef), float(1.0 + 0.5 * d), p, s);
  else
    nd_dispatch<0>(ND, L, round_up(n, kTile), m, float(coef), 0.f, p, s);
  prof_end(FSK_OP_KDE_ND_SPLITK, s);
  if ((st = launch_status(what)) != FSK_OK) return st;
  return fsk_reduce_partials_f64(L.part, p.splits, m, m, out, stream);
}

extern "C" int fsk_score_nd_ordered(const float *data, int64_t ld, int64_t n, int d,
                                    double bandwidth, unsigned flags, void *ws, size_t ws_bytes,
                                    float *pdf, float *w, int64_t ldw_row, int64_t ldw_col,
                                    void *stream) {
  const char *what = "fsk_score_nd_ordered";
  int st;
  if ((st = check_nd(what, data, ld, n, d)) != FSK_OK) return st;
  if ((st = check_elem_rows(what, n)) != FSK_OK) return st;
  if (!(bandwidth > 0.0) || !isfinite(bandwidth) || !pdf || !w) {
    set_error("%s: bandwidth must be positive and the outputs non-null", what);
    return FSK_EINVAL;
  }
  if (flags != 0) {
    set_error("%s: flags 0x%x not available (FP32 only)", what, flags);
    return FSK_EUNSUPPORTED;
  }
  NdLayout L;
  const size_t need = nd_layout(ws, n, n, d, true, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const int ND = nd_slices(d), DP = kD * ND;
  const double coef = exp2_coef(bandwidth);
  const int64_t n_norm = round_up(n, kNormPad);
  nd_center(data, ld, n, d, DP, n_norm, L.mu, L.xc, s);
  // data side in natural units (padded rows: c P = -1e30), row side times c
  hipLaunchKernelGGL(norms_nd_kernel, dim3(cdiv(n_norm, 256)), dim3(256), 0, s, L.xc, DP, n,
                     n_norm, 1.0, float(fmin(1e30 / -coef, 1e38)), L.bn);
  hipLaunchKernelGGL(norms_nd_kernel, dim3(cdiv(n_norm, 256)), dim3(256), 0, s, L.xc, DP, n,
                     n_norm, coef, kPadNeg, L.qa);
  const SplitPlan p = nd_plan(n, n, ND, true);
  prof_begin(FSK_OP_SCORE_ND_ORDERED, s);
  nd_dispatch<2>(ND, L, round_up(n, kTile), n, float(coef), 0.f, p, s);
  prof_end(FSK_OP_SCORE_ND_ORDERED, s);
  if ((st = launch_status(what)) != FSK_OK) return st;
  hipLaunchKernelGGL(score_nd_finish_kernel, dim3(cdiv(n * d, 256)), dim3(256), 0, s, L.part,
                     L.w_part, p.splits, n, d, DP, L.mu, 1.0 / double(n), pdf, w, ldw_row,
                     ldw_col);
  return launch_status(what);
}

extern "C" int fsk_sdkde_debias_nd_f64(const float *x, int64_t ldx, int64_t n, int d,
                                       const float *pdf, const float *w, int64_t ldw_row,
                                       int64_t ldw_col, double bandwidth, double eps, float *out,
                                       int64_t ldo, void *stream) {
  const char *what = "fsk_sdkde_debias_nd_f64";
  int st;
  if ((st = check_nd(what, x, ldx, n, d)) != FSK_OK) return st;
  if ((st = check_elem_rows(what, n)) != FSK_OK) return st;
  if (!(bandwidth > 0.0) || !isfinite(bandwidth) || !pdf || !w || !out || ldo < d) {
    set_error("%s: bad arguments", what);
    return FSK_EINVAL;
  }
  hipLaunchKernelGGL(debias_nd_f64_kernel, dim3(cdiv(n * d, 256)), dim3(256), 0,
                     as_stream(stream), x, ldx, n, d, pdf, w, ldw_row, ldw_col,
                     1.0 / (bandwidth * bandwidth), 0.5 * bandwidth * bandwidth, eps, out, ldo);
  return launch_status(what);
}
