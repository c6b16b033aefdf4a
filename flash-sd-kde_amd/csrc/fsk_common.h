// fsk_common.h - shared device/host helpers for the gfx950 kernels.
//
// MFMA operand layout used throughout (v_mfma_f32_16x16x4_f32, one block):
//   lane l, g = l / 16, c = l % 16
//   A (16 x 4):  lane holds A[row = c][k = g]
//   B (4 x 16):  lane holds B[k = g][col = c]
//   D (16 x 16): lane holds D[row = 4 g + r][col = c], r = 0..3
// The pair kernels put data points j on the A rows and queries i on the B
// columns, so a lane's four D values belong to ONE query (i = c) and four data
// points (j = 4 g + r): the row sum over data stays inside the lane until the
// final cross-lane reduction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <math.h>

#include "flashsdkde.h"

namespace fsk {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kD = 16;        // feature count of the N-D path
constexpr int kTile = 16;     // MFMA tile edge
constexpr int kWG = 256;      // threads per workgroup (4 wave64)
constexpr int kXCD = 8;       // MI355X: 8 accelerator complex dies
constexpr double kLog2e = 1.4426950408889634;
constexpr double kLn2 = 0.6931471805599453;
// c|x|^2 of padded tail points: exp2 of anything this negative is exactly 0,
// and 0 * (finite) stays finite (no inf - inf or 0 * inf).
constexpr float kPadNeg = -1.0e30f;
// Norm arrays (and 16-bit records) are padded to a multiple of this many rows,
// a multiple of the rows any workgroup owns, so every tile / block read is in
// bounds.
constexpr int64_t kNormPad = 512;

// Elementwise passes over (row, feature) pairs launch one thread each; a
// launch holds at most 2^32 work-items, so the paths that run them (score,
// debias, weighted combine) take at most this many rows.
constexpr int64_t kMaxElemRows = int64_t(1) << 27;

// ---- error plumbing (host) ----
void set_error(const char *fmt, ...);
int launch_status(const char *what);  // FSK_OK or FSK_ELAUNCH after a launch
int check_points16(const char *what, const float *p, int64_t ld, int64_t rows);
int check_elem_rows(const char *what, int64_t rows);  // FSK_EUNSUPPORTED above kMaxElemRows

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return cdiv(a, b) * b; }
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// exp2 scale: exp(-d2 / (2 h^2)) == exp2(c * d2) with c = -log2(e) / (2 h^2)
inline double exp2_coef(double h) { return -kLog2e / (2.0 * h * h); }

// Map a linear block id to a task so that the blocks one XCD runs at the same
// time get consecutive tasks (blocks are dealt round-robin to the 8 XCDs).
__device__ __forceinline__ int64_t xcd_task(int64_t b, int64_t per_xcd) {
  return (b % kXCD) * per_xcd + (b / kXCD);
}
inline int64_t xcd_grid(int64_t tasks, int64_t *per_xcd) {
  *per_xcd = cdiv(tasks, kXCD);
  return *per_xcd * kXCD;
}

__device__ __forceinline__ f4 mfma_f32(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
// exp2(min(x, 0)) in one instruction: exp2 is >= 0 and exceeds 1 exactly when
// x > 0, so clamping its result to [0, 1] (the clamp bit of v_exp_f32, which
// the compiler folds fmed3(., 0, 1) into) equals clamping its argument to 0.
__device__ __forceinline__ float fexp2_le1(float x) {
  return __builtin_amdgcn_fmed3f(__builtin_amdgcn_exp2f(x), 0.f, 1.f);
}

// Symmetric score tasks: window k of block I covers blocks J in
// [I + k jc, I + (k+1) jc) (J >= I). Tasks are numbered window-major (window k
// holds the nb - k jc blocks I with I + k jc < nb), so the workgroups one XCD
// runs together have neighbouring I and the same k: they share J data in L2
// and scatter into staggered blocks. sym_window_offset(k) = first task of
// window k; sym_window_offset(nwin) = number of tasks.
__host__ __device__ inline int64_t sym_window_offset(int64_t k, int64_t nb, int64_t jc) {
  return k * nb - jc * (k * (k - 1) / 2);
}
// Block I and block range [Jlo, Jhi) of a task (binary search over windows).
__device__ __forceinline__ void sym_task(int64_t task, int64_t nb, int64_t jc, int64_t nwin,
                                         int64_t *I, int64_t *Jlo, int64_t *Jhi) {
  int64_t k_lo = 0, k_hi = nwin - 1;  // largest window k with offset(k) <= task
  while (k_lo < k_hi) {
    const int64_t mid = (k_lo + k_hi + 1) >> 1;
    if (sym_window_offset(mid, nb, jc) <= task)
      k_lo = mid;
    else
      k_hi = mid - 1;
  }
  *I = task - sym_window_offset(k_lo, nb, jc);
  *Jlo = *I + k_lo * jc;
  *Jhi = *Jlo + jc < nb ? *Jlo + jc : nb;
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                                               0xF, 0xF, false));
}

// Sums over the 16 lanes of a DPP row of four per-lane values v[0..3], folded
// so that lane c ends with the sum of v[c & 3]: two exchange stages (lane
// bits 0 and 1, each keeping the value the lane is to reduce and sending the
// other to its partner: 2 selects + 1 DPP add per pair) halve the values to
// reduce, then rotations by 4 and 8 lanes finish the sum (11 instructions
// against 16 DPP adds and 3 selects for four separate butterfly row sums).
__device__ __forceinline__ float row16_sum4(const float (&v)[4], int c) {
  const bool b0 = c & 1, b1 = c & 2;
  const float a = (b0 ? v[1] : v[0]) + dpp_f32<0xB1>(b0 ? v[0] : v[1]);  // quad_perm [1,0,3,2]
  const float b = (b0 ? v[3] : v[2]) + dpp_f32<0xB1>(b0 ? v[2] : v[3]);
  float e = (b1 ? b : a) + dpp_f32<0x4E>(b1 ? a : b);  // quad_perm [2,3,0,1]
  e += dpp_f32<0x124>(e);                             // row_ror:4
  e += dpp_f32<0x128>(e);                             // row_ror:8
  return e;
}

__device__ __forceinline__ f4 ld4(const float *p) {
  return *reinterpret_cast<const f4 *>(p);
}

// Sum over the four lane groups (g) of a wave; every lane gets the total of
// lanes {c, c+16, c+32, c+48}.
__device__ __forceinline__ float sum_groups(float v) {
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}
__device__ __forceinline__ double sum_groups(double v) {
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

// Optional event timing around the main kernel of an op (see fsk_profile_enable).
void prof_begin(int op, hipStream_t s);
void prof_end(int op, hipStream_t s);

// Workspace sizes of the individual operations (defined next to each op).
size_t kde16_ws_bytes(int64_t n, int64_t m, bool with_partials);
size_t kde16_lse_ws_bytes(int64_t n, int64_t m, unsigned flags);
size_t spatial_ws_bytes();
size_t score16_ws_bytes(int64_t n, bool ordered);
size_t score16_lp_ws_bytes(int64_t n, int64_t m, unsigned prec);
size_t score16_rows_ws_bytes(int64_t n, int64_t m);
size_t score16_valu_ws_bytes(int64_t n, int64_t m, bool same_rows);
int score16_valu_run(const char *what, const float *rows, int64_t ldr, int64_t m,
                     const float *data, int64_t ld, int64_t n, double h, bool same_rows, void *ws,
                     size_t ws_bytes, float *pdf, float *w, int64_t ldw_row, int64_t ldw_col,
                     void *stream);
int score16_lp_run(const char *what, const float *rows, int64_t ldr, int64_t m,
                   const float *data, int64_t ld, int64_t n, double h, unsigned prec,
                   bool self_rows, bool dense, void *ws, size_t ws_bytes, float *pdf, float *w,
                   int64_t ldw_row, int64_t ldw_col, void *stream);
size_t kde1_ws_bytes(int64_t n, int64_t m, bool partials, bool two);

// Fill kernel used instead of hipMemsetAsync (safe under graph capture).
__global__ void fill_f32(float *p, int64_t n, float v);
__global__ void fill_f64(double *p, int64_t n, double v);

// Per-row scaled squared norms: out[j] = coef * |x_j|^2 (FP64 accumulation,
// one rounding), rows [n, n_pad) get `pad` (kPadNeg for norm terms in base-2
// exponent units) so padded tail points contribute exactly 0.
__global__ void prep_norms16(const float *__restrict__ x, int64_t ld, int64_t n,
                             int64_t n_pad, double coef, float pad, float *__restrict__ out);

// Points centred on the mean mu = mu_sum * inv_n: xc = x - mu (FP64
// difference, FP32 result), 16 floats per row, rows [n, n_pad) zero.
__global__ void center16_kernel(const float *__restrict__ x, int64_t ld, int64_t n, int64_t n_pad,
                                const double *__restrict__ mu_sum, double inv_n,
                                float *__restrict__ xc);

}  // namespace fsk
