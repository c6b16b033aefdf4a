// fsk_common.hip - error state, fill/norm prep, FP64 split-K combine, FP64 debias.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "fsk_common.h"
#include "fsk_plan.h"

namespace fsk {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int launch_status(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return FSK_ELAUNCH;
  }
  return FSK_OK;
}

int check_elem_rows(const char *what, int64_t rows) {
  if (rows > kMaxElemRows) {
    set_error("%s: %lld rows, above the supported %lld", what, (long long)rows,
              (long long)kMaxElemRows);
    return FSK_EUNSUPPORTED;
  }
  return FSK_OK;
}

int check_points16(const char *what, const float *p, int64_t ld, int64_t rows) {
  if (rows <= 0) {
    set_error("%s: need at least one row, got %lld", what, (long long)rows);
    return FSK_EINVAL;
  }
  if (!p) {
    set_error("%s: null pointer", what);
    return FSK_EINVAL;
  }
  if (ld < kD || (ld & 3)) {
    set_error("%s: row stride %lld must be >= 16 and a multiple of 4", what, (long long)ld);
    return FSK_EINVAL;
  }
  if (reinterpret_cast<uintptr_t>(p) & 15) {
    set_error("%s: pointer must be 16-byte aligned", what);
    return FSK_EINVAL;
  }
  return FSK_OK;
}

// ---- optional kernel timing: a ring of event pairs per op ----
constexpr int kProfOps = 32;
constexpr int kProfRing = 64;
static bool g_prof = false;
static hipEvent_t g_ev_begin[kProfOps][kProfRing], g_ev_end[kProfOps][kProfRing];
static bool g_ev_made[kProfOps][kProfRing];
static int g_ev_count[kProfOps];

void prof_begin(int op, hipStream_t s) {
  if (!g_prof || op < 0 || op >= kProfOps) return;
  const int k = g_ev_count[op] % kProfRing;
  if (!g_ev_made[op][k]) {
    if (hipEventCreate(&g_ev_begin[op][k]) != hipSuccess) return;
    if (hipEventCreate(&g_ev_end[op][k]) != hipSuccess) return;
    g_ev_made[op][k] = true;
  }
  (void)hipEventRecord(g_ev_begin[op][k], s);
}

void prof_end(int op, hipStream_t s) {
  if (!g_prof || op < 0 || op >= kProfOps) return;
  const int k = g_ev_count[op] % kProfRing;
  if (!g_ev_made[op][k]) return;
  (void)hipEventRecord(g_ev_end[op][k], s);
  ++g_ev_count[op];
}

__global__ void fill_f32(float *p, int64_t n, float v) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void fill_f64(double *p, int64_t n, double v) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void prep_norms16(const float *__restrict__ x, int64_t ld, int64_t n,
                             int64_t n_pad, double coef, float pad, float *__restrict__ out) {
  int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n_pad) return;
  if (j >= n) {
    out[j] = pad;
    return;
  }
  const float *r = x + j * ld;
  double s = 0.0;
#pragma unroll
  for (int f = 0; f < kD; f += 4) {
    f4 v = ld4(r + f);
    s += double(v[0]) * v[0] + double(v[1]) * v[1] + double(v[2]) * v[2] + double(v[3]) * v[3];
  }
  out[j] = float(coef * s);
}

__global__ void center16_kernel(const float *__restrict__ x, int64_t ld, int64_t n, int64_t n_pad,
                                const double *__restrict__ mu_sum, double inv_n,
                                float *__restrict__ xc) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_pad * kD) return;
  const int64_t j = e / kD;
  const int f = int(e % kD);
  xc[e] = j < n ? float(double(x[j * ld + f]) - mu_sum[f] * inv_n) : 0.f;
}

// out[i] = sum_s partials[s*ld + i], FP64, fixed order over s.
__global__ void reduce_partials_f64_kernel(const float *__restrict__ part, int64_t S,
                                           int64_t m, int64_t ld, float *__restrict__ out) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double acc = 0.0;
  for (int64_t s = 0; s < S; ++s) acc += double(part[s * ld + i]);
  out[i] = float(acc);
}

// One thread per (row, feature); partials laid out [S][m][16].
__global__ void reduce_weighted_f64_kernel(const float *__restrict__ part, int64_t S,
                                           int64_t m, int64_t ld, float *__restrict__ out,
                                           int64_t ldo_row, int64_t ldo_col) {
  int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m * kD) return;
  double acc = 0.0;
  for (int64_t s = 0; s < S; ++s) acc += double(part[s * ld + e]);
  int64_t i = e / kD, f = e % kD;
  out[i * ldo_row + f * ldo_col] = float(acc);
}

__global__ void debias_f64_kernel(const float *__restrict__ x, int64_t ldx, int64_t n,
                                  const float *__restrict__ pdf, const float *__restrict__ w,
                                  int64_t ldw_row, int64_t ldw_col, double inv_h2, double delta,
                                  double eps, float *__restrict__ out, int64_t ldo,
                                  float *__restrict__ score_out) {
  int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * kD) return;
  int64_t i = e / kD, f = e % kD;
  double xv = x[i * ldx + f];
  double p = double(pdf[i]) + eps;
  double s = (double(w[i * ldw_row + f * ldw_col]) / p - xv) * inv_h2;
  out[i * ldo + f] = float(xv + delta * s);
  if (score_out) score_out[e] = float(s);
}

}  // namespace fsk

using namespace fsk;

extern "C" const char *fsk_last_error(void) { return g_err; }

extern "C" int fsk_version(void) { return 10000; }

extern "C" int fsk_profile_enable(int on) {
  g_prof = on != 0;
  for (int op = 0; op < kProfOps; ++op) g_ev_count[op] = 0;
  return FSK_OK;
}

extern "C" double fsk_profile_last_ms(int op) {
  if (op < 0 || op >= kProfOps || g_ev_count[op] == 0) return -1.0;
  const int cnt = g_ev_count[op] < kProfRing ? g_ev_count[op] : kProfRing;
  double total = 0.0;
  for (int k = 0; k < cnt; ++k) {
    if (hipEventSynchronize(g_ev_end[op][k]) != hipSuccess) return -1.0;
    float ms = -1.f;
    if (hipEventElapsedTime(&ms, g_ev_begin[op][k], g_ev_end[op][k]) != hipSuccess) return -1.0;
    total += ms;
  }
  return total / cnt;
}

extern "C" size_t fsk_workspace_bytes(int op, int64_t n, int64_t m, unsigned flags) {
  if (n <= 0 || m < 0) return 0;
  switch (op) {
    case FSK_OP_KDE16_SPLITK: return m > 0 ? kde16_ws_bytes(n, m, true) : 0;
    case FSK_OP_KDE16_ATOMIC: return m > 0 ? kde16_ws_bytes(n, m, false) : 0;
    case FSK_OP_KDE16_LOGDENSITY: return m > 0 ? kde16_lse_ws_bytes(n, m, flags) : 0;
    case FSK_OP_KDE1_SPLITK: return m > 0 ? kde1_ws_bytes(n, m, true, false) : 0;
    case FSK_OP_KDE1_ATOMIC: return m > 0 ? kde1_ws_bytes(n, m, false, false) : 0;
    case FSK_OP_KDE1_SCORE: return m > 0 ? kde1_ws_bytes(n, m, true, true) : 0;
    case FSK_OP_SCORE16_SYM: return score16_ws_bytes(n, false);
    case FSK_OP_SCORE16_ROWS:
      if (m <= 0) return 0;
      if ((flags & FSK_PREC_MASK) == FSK_PREC_F32_VALU) return score16_valu_ws_bytes(n, m, false);
      return (flags & FSK_PREC_MASK) == FSK_PREC_F32
                 ? score16_rows_ws_bytes(n, m)
                 : score16_lp_ws_bytes(n, m, flags & FSK_PREC_MASK);
    case FSK_OP_SCORE16_ORDERED:
      if ((flags & FSK_PREC_MASK) == FSK_PREC_F32_VALU) return score16_valu_ws_bytes(n, n, true);
      return (flags & FSK_PREC_MASK) == FSK_PREC_F32
                 ? score16_ws_bytes(n, true)
                 : score16_lp_ws_bytes(n, n, flags & FSK_PREC_MASK);
    case FSK_OP_SPATIAL_CODES16: return spatial_ws_bytes();
    default: return 0;
  }
}

extern "C" int fsk_reduce_partials_f64(const float *partials, int64_t n_splits, int64_t m,
                                       int64_t ld_split, float *out, void *stream) {
  if (!partials || !out || n_splits <= 0 || m <= 0 || ld_split < m) {
    set_error("fsk_reduce_partials_f64: bad arguments (S=%lld m=%lld ld=%lld)",
              (long long)n_splits, (long long)m, (long long)ld_split);
    return FSK_EINVAL;
  }
  hipLaunchKernelGGL(reduce_partials_f64_kernel, dim3(cdiv(m, 256)), dim3(256), 0,
                     as_stream(stream), partials, n_splits, m, ld_split, out);
  return launch_status("fsk_reduce_partials_f64");
}

extern "C" int fsk_reduce_weighted_partials_f64(const float *partials, int64_t n_splits,
                                                int64_t m, int64_t ld_split, float *out,
                                                int64_t ldo_row, int64_t ldo_col,
                                                void *stream) {
  if (!partials || !out || n_splits <= 0 || m <= 0 || ld_split < m * kD) {
    set_error("fsk_reduce_weighted_partials_f64: bad arguments");
    return FSK_EINVAL;
  }
  if (int st = check_elem_rows("fsk_reduce_weighted_partials_f64", m)) return st;
  hipLaunchKernelGGL(reduce_weighted_f64_kernel, dim3(cdiv(m * kD, 256)), dim3(256), 0,
                     as_stream(stream), partials, n_splits, m, ld_split, out, ldo_row,
                     ldo_col);
  return launch_status("fsk_reduce_weighted_partials_f64");
}

extern "C" int fsk_sdkde_debias_f64(const float *x, int64_t ldx, int64_t n, const float *pdf,
                                    const float *w, int64_t ldw_row, int64_t ldw_col,
                                    double bandwidth, double eps, float *out, int64_t ldo,
                                    float *score_out, void *stream) {
  if (!(bandwidth > 0.0) || !isfinite(bandwidth)) {
    set_error("fsk_sdkde_debias_f64: bandwidth must be positive, got %g", bandwidth);
    return FSK_EINVAL;
  }
  if (!x || !pdf || !w || !out || n <= 0 || ldx < kD || ldo < kD) {
    set_error("fsk_sdkde_debias_f64: bad arguments");
    return FSK_EINVAL;
  }
  if (int st = check_elem_rows("fsk_sdkde_debias_f64", n)) return st;
  double inv_h2 = 1.0 / (bandwidth * bandwidth);
  double delta = 0.5 * bandwidth * bandwidth;
  hipLaunchKernelGGL(debias_f64_kernel, dim3(cdiv(n * kD, 256)), dim3(256), 0,
                     as_stream(stream), x, ldx, n, pdf, w, ldw_row, ldw_col, inv_h2, delta,
                     eps, out, ldo, score_out);
  return launch_status("fsk_sdkde_debias_f64");
}
