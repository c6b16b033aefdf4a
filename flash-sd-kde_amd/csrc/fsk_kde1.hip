// fsk_kde1.hip - 1-D Gaussian KDE sums and density-derivative sums (VALU).
//
// Per pair, difference first as the reference's 1-D kernels do
// (kernels/kde_eval_1d_splitk.py:43-46): d = q - x (exact for nearby
// points), t = (d d) K with K = k^2 = log2(e) / (2 h^2) rounded once, and
// exp(-(q-x)^2/(2h^2)) = exp2(-t): a subtraction, two multiplies, one exp2
// and one add. t carries three roundings whatever |x|/h is (a pre-scaled
// x' = k x would carry an absolute error of |k x| ulp into k (q - x), and
// squaring a rounded k doubles its rounding). The Laplacian factor
// (1.5 - ln2 t) multiplies each pair before the
// sum, as the reference's phi * (1 + 1/2 - z^2/2) does
// (flash_sd_kde/reference.py:200-213), so the sum carries no cancellation
// between two separately rounded totals. A data chunk is staged in LDS and
// read as wave-wide broadcasts; each thread carries k1dQPT queries so every
// LDS read feeds several pairs.
//
// Sorted order. The data and the queries are first sorted (hipCUB radix
// sorts, fsk_sort.hip), so a block of consecutive sorted queries spans a
// narrow range [q_lo, q_hi]. A pair with k |q - x| > 12.7 has t > 161 and
// exp2(-t) < 2^-161, which is exactly 0 in FP32 (the smallest denormal is
// 2^-149); so each query block only visits the data chunks that meet
// [q_lo - 12.8/k, q_hi + 12.8/k] (binary search in FP64), and every pair it
// skips would have added exactly +0. The plain and score sums go further:
// a query's largest term is its nearest data point's, so its terms more than
// 2^-T below that one (T = ceil(log2 n) + 32) add at most n 2^-T <= 2^-32 of
// its sum; each query reaches sqrt(d_nn^2 + T/k^2) (kde1_reach_kernel), and a
// block visits the chunks that meet the union of its queries' reaches.
// Results are written at the sorted positions and scattered back to the
// callers' order.
#include "fsk_common.h"
#include "fsk_plan.h"

namespace fsk {

// Reach of a query in units of 1/k: beyond it every kernel term is exactly 0.
constexpr double kReach1d = 12.8;
// Points per FP32 run of the sums before they are added in FP64.
constexpr int kFlush1d = 256;
static_assert(k1dChunk % kFlush1d == 0, "a chunk holds whole runs");

size_t sort_f32_temp_bytes(int64_t n, bool pairs);
int sort_keys_f32(const char *what, const float *in, float *out, int64_t n, void *tmp,
                  size_t tmp_bytes, hipStream_t s);
int sort_pairs_f32_i32(const char *what, const float *kin, float *kout, const int *vin,
                       int *vout, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t s);

// Padded data points (sorted order: past the real ones) sit twice the reach
// beyond the largest data point and query, so their terms are exactly 0 and
// u^2 stays finite (no 0 * inf in the Laplacian and moment sums).
__global__ void pad1_kernel(float *__restrict__ xs, int64_t n, int64_t n_pad,
                            const float *__restrict__ qs, int64_t m, double reach) {
  const int64_t j = n + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n_pad) return;
  const double top = fmax(double(xs[n - 1]), double(qs[m - 1]));
  xs[j] = float(top + 2.0 * reach);
}

__global__ void iota_i32(int *p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = int(i);
}

// Reach of each sorted query for the plain and score sums: terms more than
// 2^-T below the query's largest term (its nearest data point's) are left
// out, T = ceil(log2 n) + 32, so a sum loses at most n 2^-T <= 2^-32 of
// itself: qreach[i] = sqrt(d_nn^2 + T / k^2), at most the exact-zero reach.
__global__ void kde1_reach_kernel(const float *__restrict__ xs, int64_t n,
                                  const float *__restrict__ qs, int64_t m, double t_over_k2,
                                  double zero_reach, float *__restrict__ qreach) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double q = qs[i];
  int64_t a = 0, z = n;  // first j with xs[j] >= q
  while (a < z) {
    const int64_t mid = (a + z) >> 1;
    if (double(xs[mid]) < q) a = mid + 1; else z = mid;
  }
  double d = INFINITY;
  if (a < n) d = double(xs[a]) - q;
  if (a > 0) d = fmin(d, q - double(xs[a - 1]));
  const double r = sqrt(d * d + t_over_k2);
  qreach[i] = float(fmin(r, zero_reach));
}

// Data-chunk range [ranges[2b], ranges[2b+1]) of query block b: the chunks
// that hold the sorted data points within reach of the block's sorted
// queries (empty when there are none). One workgroup per block; with qreach
// each query has its own reach, else every query has `reach`.
__global__ __launch_bounds__(256) void kde1_range_kernel(
    const float *__restrict__ xs, int64_t n, int64_t n_pad, const float *__restrict__ qs,
    int64_t m, int64_t qblocks, double reach, const float *__restrict__ qreach,
    int64_t *__restrict__ ranges) {
  __shared__ double red_lo[256], red_hi[256];
  const int64_t b = blockIdx.x;
  const int64_t i0 = b * k1dRowsPerWG;
  const int64_t i1 = (i0 + k1dRowsPerWG < m ? i0 + k1dRowsPerWG : m) - 1;
  double lo = INFINITY, hi = -INFINITY;
  for (int64_t i = i0 + threadIdx.x; i <= i1; i += 256) {
    const double r = qreach ? double(qreach[i]) : reach;
    lo = fmin(lo, double(qs[i]) - r);
    hi = fmax(hi, double(qs[i]) + r);
  }
  red_lo[threadIdx.x] = lo;
  red_hi[threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
      red_lo[threadIdx.x] = fmin(red_lo[threadIdx.x], red_lo[threadIdx.x + w]);
      red_hi[threadIdx.x] = fmax(red_hi[threadIdx.x], red_hi[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  lo = red_lo[0];
  hi = red_hi[0];
  int64_t a = 0, z = n;  // first j with xs[j] >= lo
  while (a < z) {
    const int64_t mid = (a + z) >> 1;
    if (double(xs[mid]) < lo) a = mid + 1; else z = mid;
  }
  const int64_t jlo = a;
  z = n;  // first j with xs[j] > hi
  while (a < z) {
    const int64_t mid = (a + z) >> 1;
    if (double(xs[mid]) <= hi) a = mid + 1; else z = mid;
  }
  const int64_t jhi = a;
  int64_t c0 = (jlo / k1dChunk) * k1dChunk, c1 = ((jhi + k1dChunk - 1) / k1dChunk) * k1dChunk;
  if (c1 > n_pad) c1 = n_pad;
  if (jhi <= jlo) c0 = c1 = 0;
  ranges[2 * b] = c0;
  ranges[2 * b + 1] = c1;
}

// out[perm[i]] = res[i] (and out2 from res2): sorted positions back to the
// callers' order.
__global__ void kde1_scatter_kernel(const float *__restrict__ res, const float *__restrict__ res2,
                                    const int *__restrict__ perm, int64_t m,
                                    float *__restrict__ out, float *__restrict__ out2) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t o = perm[i];
  out[o] = res[i];
  if (res2) out2[o] = res2[i];
}


// MODE 0: plain sums; MODE 1: sums of phi (1.5 - ln2 u^2);
// MODE 2: pdf sums and derivative sums sum(-z phi) (two outputs);
// MODE 3: sums of phi z^2 = phi 2 ln2 u^2 (the Laplacian correction's moment).
// xs: sorted data (n_pad, padded by pad1_kernel); qs: sorted queries;
// ranges: each query block's data-chunk range (kde1_range_kernel).
template <int MODE, bool ATOMIC>
__global__ __launch_bounds__(256) void kde1_kernel(const float *__restrict__ xs, int64_t n_pad,
                                                   const float *__restrict__ qs, int64_t m,
                                                   const int64_t *__restrict__ ranges,
                                                   int64_t qblocks,
                                                   int64_t ntasks, int64_t per_xcd,
                                                   float ksq, float zscale,
                                                   float *__restrict__ out,
                                                   float *__restrict__ out2, int64_t ld_out) {
  __shared__ f4 chunk[k1dChunk / 4];
  const int64_t task = xcd_task(blockIdx.x, per_xcd);
  if (task >= ntasks) return;
  const int64_t split = task / qblocks;
  const int64_t qb = task - split * qblocks;
  const int tid = threadIdx.x;
  float q[k1dQPT], acc[k1dQPT], accb[k1dQPT], acc2[k1dQPT];
  double a64[k1dQPT], b64[k1dQPT];
#pragma unroll
  for (int k = 0; k < k1dQPT; ++k) {
    int64_t i = qb * k1dRowsPerWG + k * 256 + tid;
    q[k] = qs[i < m ? i : m - 1];
    acc[k] = accb[k] = acc2[k] = 0.f;
    a64[k] = b64[k] = 0.0;
  }
  // this split's share of the block's data range: the range's chunks in
  // `splits` consecutive runs of equal length (the last may be shorter or
  // empty). The XCD-aware task order gives each XCD one split of every query
  // block, so the XCDs get equal work wherever the data is dense.
  const int64_t splits = ntasks / qblocks;
  const int64_t c0 = ranges[2 * qb], c1 = ranges[2 * qb + 1];
  const int64_t per = (c1 - c0 + splits * k1dChunk - 1) / (splits * k1dChunk) * k1dChunk;
  const int64_t j_begin = c0 + split * per;
  const int64_t j_end = j_begin + per < c1 ? j_begin + per : c1;
  for (int64_t j0 = j_begin; j0 < j_end; j0 += k1dChunk) {
    __syncthreads();
    reinterpret_cast<float *>(chunk)[tid] = xs[j0 + tid];
    reinterpret_cast<float *>(chunk)[tid + 256] = xs[j0 + tid + 256];
    reinterpret_cast<float *>(chunk)[tid + 512] = xs[j0 + tid + 512];
    reinterpret_cast<float *>(chunk)[tid + 768] = xs[j0 + tid + 768];
    __syncthreads();
    // FP32 sums over runs of kFlush1d points, two per query (even and odd
    // points: half the chain length at no extra instruction), then FP64.
    // (Adding each four terms pairwise before the running sum, query by
    // query, measured 23 % slower for the score: the eight queries' chains
    // interleave less.)
    for (int jb = 0; jb < k1dChunk / 4; jb += kFlush1d / 4) {
      for (int jj = jb; jj < jb + kFlush1d / 4; ++jj) {
        const f4 xv = chunk[jj];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
#pragma unroll
          for (int k = 0; k < k1dQPT; ++k) {
            const float d = q[k] - xv[s];
            const float t = (d * d) * ksq;
            const float e = fexp2(-t);
            // the even and the odd points of a run go to separate sums
            float &a = (s & 1) ? accb[k] : acc[k];
            if (MODE == 1)
              a = fmaf(e, fmaf(t, -float(kLn2), 1.5f), a);
            else if (MODE == 3)
              a = fmaf(e, t, a);
            else
              a += e;
            if (MODE == 2) acc2[k] = fmaf(e, d, acc2[k]);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < k1dQPT; ++k) {
        a64[k] += double(acc[k]) + double(accb[k]);
        b64[k] += acc2[k];
        acc[k] = accb[k] = acc2[k] = 0.f;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < k1dQPT; ++k) {
    int64_t i = qb * k1dRowsPerWG + k * 256 + tid;
    if (i >= m) continue;
    const double v = MODE == 3 ? 2.0 * kLn2 * a64[k] : a64[k];
    float v2 = 0.f;
    // sum(-z phi) with z = (q - x) / h: acc2 holds sum(phi (q - x)), zscale = 1/h
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
  float *xs, *qs, *res, *res2, *part, *part2, *qreach;
  int *qiota, *qperm;
  int64_t *ranges;
  void *tmp;
  size_t tmp_bytes;
};

// Returns 0 when the sort scratch size cannot be determined (no device).
static size_t kde1_layout(void *ws, int64_t n, int64_t m, bool partials, bool two,
                          Kde1Layout *L) {
  Carve cv(ws);
  SplitPlan p = kde1_plan(n, m);
  const int64_t n_pad = round_up(n, k1dChunk);
  const size_t tk = sort_f32_temp_bytes(n, false), tp = sort_f32_temp_bytes(m, true);
  if (tk == 0 || tp == 0) return 0;
  L->tmp_bytes = tk > tp ? tk : tp;
  L->xs = cv.take<float>(n_pad);
  L->qs = cv.take<float>(m);
  L->qiota = cv.take<int>(m);
  L->qperm = cv.take<int>(m);
  L->ranges = cv.take<int64_t>(2 * p.qblocks);
  L->qreach = cv.take<float>(m);
  L->res = cv.take<float>(m);
  L->res2 = two ? cv.take<float>(m) : nullptr;
  L->part = partials ? cv.take<float>(p.splits * m) : nullptr;
  L->part2 = (partials && two) ? cv.take<float>(p.splits * m) : nullptr;
  L->tmp = cv.take<char>(int64_t(L->tmp_bytes));
  return cv.used;
}

size_t kde1_ws_bytes(int64_t n, int64_t m, bool partials, bool two) {
  if (round_up(n, k1dChunk) > INT32_MAX || m > INT32_MAX) return 0;
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
  if (round_up(n, k1dChunk) > INT32_MAX || m > INT32_MAX) {
    set_error("%s: the 1-D path sorts at most 2^31 - 1 points and queries (n=%lld m=%lld)",
              what, (long long)n, (long long)m);
    return FSK_EUNSUPPORTED;
  }
  {
    // the radix sorts record memset nodes, which do not replay correctly from
    // a HIP graph: refuse capture rather than give wrong results later
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(as_stream(stream), &cs) == hipSuccess &&
        cs != hipStreamCaptureStatusNone) {
      set_error("%s: the 1-D path sorts with library kernels that cannot be captured in a "
                "HIP graph; run it outside capture", what);
      return FSK_EUNSUPPORTED;
    }
  }
  Kde1Layout L;
  size_t need = kde1_layout(ws, n, m, !atomic, mode == 2, &L);
  if (need == 0) {
    set_error("%s: sort scratch size unavailable (no HIP device)", what);
    return FSK_ELAUNCH;
  }
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const double k = sqrt(kLog2e / (2.0 * h * h));
  const int64_t n_pad = round_up(n, k1dChunk);
  int st;
  if ((st = sort_keys_f32(what, data, L.xs, n, L.tmp, L.tmp_bytes, s)) != FSK_OK) return st;
  hipLaunchKernelGGL(iota_i32, dim3(cdiv(m, 256)), dim3(256), 0, s, L.qiota, m);
  if ((st = sort_pairs_f32_i32(what, query, L.qs, L.qiota, L.qperm, m, L.tmp, L.tmp_bytes, s)) !=
      FSK_OK)
    return st;
  if (n_pad > n)
    hipLaunchKernelGGL(pad1_kernel, dim3(cdiv(n_pad - n, 256)), dim3(256), 0, s, L.xs, n, n_pad,
                       L.qs, m, kReach1d / k);
  SplitPlan p = kde1_plan(n, m);
  // plain and score sums: per-query reach from the nearest data point (the
  // Laplacian factor and the moment change sign or grow with distance: they
  // keep the exact-zero reach)
  const bool bounded = mode == 0 || mode == 2;
  if (bounded) {
    const double T = std::ceil(std::log2(double(n))) + 32.0;
    hipLaunchKernelGGL(kde1_reach_kernel, dim3(cdiv(m, 256)), dim3(256), 0, s, L.xs, n, L.qs, m,
                       T / (k * k), kReach1d / k, L.qreach);
  }
  hipLaunchKernelGGL(kde1_range_kernel, dim3(p.qblocks), dim3(256), 0, s, L.xs, n, n_pad, L.qs, m,
                     p.qblocks, kReach1d / k, bounded ? (const float *)L.qreach : nullptr,
                     L.ranges);
  const int64_t ntasks = p.qblocks * p.splits;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
  // the score's derivative sums are sum(phi (q - x)) times -1/h
  const float zscale = float(1.0 / h);
  float *o1 = atomic ? L.res : L.part;
  float *o2 = atomic ? L.res2 : L.part2;
  if (atomic) {
    hipLaunchKernelGGL(fill_f32, dim3(cdiv(m, 256)), dim3(256), 0, s, L.res, m, 0.f);
    if (mode == 2)
      hipLaunchKernelGGL(fill_f32, dim3(cdiv(m, 256)), dim3(256), 0, s, L.res2, m, 0.f);
  }
  const int op = mode == 2 ? FSK_OP_KDE1_SCORE : (atomic ? FSK_OP_KDE1_ATOMIC : FSK_OP_KDE1_SPLITK);
  prof_begin(op, s);
#define FSK_K1(MODE, AT)                                                                      \
  hipLaunchKernelGGL((kde1_kernel<MODE, AT>), dim3(grid), dim3(256), 0, s, L.xs, n_pad, L.qs, m, \
                     L.ranges, p.qblocks, ntasks, per_xcd, float(k * k), zscale, o1, o2, \
                     m)
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
  if ((st = launch_status(what)) != FSK_OK) return st;
  if (!atomic) {
    if ((st = fsk_reduce_partials_f64(L.part, p.splits, m, m, L.res, stream)) != FSK_OK) return st;
    if (mode == 2 && (st = fsk_reduce_partials_f64(L.part2, p.splits, m, m, L.res2, stream)) !=
                         FSK_OK)
      return st;
  }
  hipLaunchKernelGGL(kde1_scatter_kernel, dim3(cdiv(m, 256)), dim3(256), 0, s, L.res,
                     mode == 2 ? L.res2 : nullptr, L.qperm, m, out, out2);
  return launch_status(what);
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
