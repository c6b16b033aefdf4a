// fsk_lowp.h - 16-bit (BF16 / FP16) operand helpers shared by the
// reduced-precision pair kernels.
//
// Packed records per point, 16-value blocks. BF16X6 uses four blocks
//   data  (A side): [ x1[16] | x2[16] | x3[16] | 1 1 1 bh bm bl 1 1 1 0 ... ]
//   query (B side): [ q1[16] | q2[16] | q3[16] | ah am al 1 1 1 mh mm ml 0 ... ]
// and the other 16-bit forms three ([ x1 | x2 | fold ], 96 bytes: the smaller
// stride measured ~10 % faster for them than the 128-byte one), with
// x = x1 + x2 + x3 the centred point split into 16-bit pieces
// (x1 = rn(x), x2 = rn(x - x1), x3 = rn(x - x1 - x2); exact for BF16 and an
// FP32 x), q' = -2c (q - mu) = q1 + q2 + q3, b = c|x - mu|^2, a = c|q - mu|^2
// and -m (log-sum-exp shift, 0 outside the LSE kernel) each split into three
// 16-bit pieces, so one MFMA over the fold block adds the norm terms (and -m)
// with no per-pair VALU work. A lane of v_mfma_f32_16x16x32_* reads 8
// consecutive K values, so every operand below is one 16-byte load at a
// per-lane offset: the low K half (lane groups 0, 1) and the high K half
// (groups 2, 3) may come from different blocks.
#pragma once

#include "fsk_common.h"

namespace fsk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

template <int PREC>
struct RecLayout {
  static constexpr bool kX6 = PREC == FSK_PREC_BF16X6;
  static constexpr int kSize = kX6 ? 64 : 48;  // 16-bit values per record
  static constexpr int kP1 = 0;                // first piece
  static constexpr int kP2 = 16;               // second piece
  static constexpr int kP3 = 32;               // third piece (BF16X6 only)
  static constexpr int kFold = kX6 ? 48 : 32;  // norm terms and -m
};
inline int64_t rec_size(unsigned prec) { return prec == FSK_PREC_BF16X6 ? 64 : 48; }

// 16-bit type of the cross-term operands of a form.
template <int PREC>
struct CrossType {
  static constexpr int kP16 =
      (PREC == FSK_PREC_F16 || PREC == FSK_PREC_F16X3 || PREC == FSK_PREC_F16X3S) ? FSK_PREC_F16
                                                                                  : FSK_PREC_BF16;
};

__device__ __forceinline__ u16x8 ld_u16x8(const unsigned short *p) {
  return *reinterpret_cast<const u16x8 *>(p);
}

// Offset of lane group g's 8 values for an operand whose low K half is block
// `lo` and high K half is block `hi` (each 16 values).
__device__ __forceinline__ int rec_off(int g, int lo, int hi) {
  return g < 2 ? lo + 8 * g : hi + 8 * (g - 2);
}

// ---- 16-bit helpers (round to nearest even) ----
template <int PREC>
__device__ __forceinline__ unsigned short to16(float v);
template <>
__device__ __forceinline__ unsigned short to16<FSK_PREC_BF16>(float v) {
  __bf16 b = (__bf16)v;
  return __builtin_bit_cast(unsigned short, b);
}
template <>
__device__ __forceinline__ unsigned short to16<FSK_PREC_F16>(float v) {
  _Float16 h = (_Float16)v;
  return __builtin_bit_cast(unsigned short, h);
}
template <int PREC>
__device__ __forceinline__ float from16(unsigned short u);
template <>
__device__ __forceinline__ float from16<FSK_PREC_BF16>(unsigned short u) {
  return float(__builtin_bit_cast(__bf16, u));
}
template <>
__device__ __forceinline__ float from16<FSK_PREC_F16>(unsigned short u) {
  return float(__builtin_bit_cast(_Float16, u));
}
// Split v into three 16-bit pieces (hi, mid, lo) whose sum approximates v.
template <int P16>
__device__ __forceinline__ void split3(double v, unsigned short *p) {
  float r = float(v);
  p[0] = to16<P16>(r);
  double rem = v - double(from16<P16>(p[0]));
  p[1] = to16<P16>(float(rem));
  rem -= double(from16<P16>(p[1]));
  p[2] = to16<P16>(float(rem));
}

// Centred, scaled operands. side 0 = data (A side), 1 = query (B side).
// FP32 forms write float rows (data: x - mu, query: -2c (q - mu)) plus the
// norm term; 16-bit forms write 48-value records. Rows in [rows, rows_pad)
// are padding: zero features and a norm term that drives exp2 to 0.
template <int PREC>
__global__ void prep_lse_kernel(const float *__restrict__ x, int64_t ld, int64_t rows,
                                int64_t rows_pad, const double *__restrict__ mu_sum, double inv_n,
                                double scale, double coef, int side, float *__restrict__ f32_out,
                                float *__restrict__ norm_out,
                                unsigned short *__restrict__ rec_out) {
  int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= rows_pad) return;
  const bool pad = j >= rows;
  double xc[kD];
  double nrm = 0.0;
#pragma unroll
  for (int f = 0; f < kD; ++f) {
    double v = pad ? 0.0 : (double(x[j * ld + f]) - mu_sum[f] * inv_n) * scale;
    xc[f] = v;
    nrm += v * v;
  }
  const double nt = coef * nrm;  // c |x - mu|^2
  const double sc = side == 0 ? 1.0 : -2.0 * coef;
  if (PREC == FSK_PREC_F32 || PREC == FSK_PREC_F32_VALU) {
#pragma unroll
    for (int f = 0; f < kD; ++f) f32_out[j * kD + f] = float(sc * xc[f]);
    norm_out[j] = pad ? kPadNeg : float(nt);
    return;
  }
  constexpr int P16 = CrossType<PREC>::kP16;
  using RL = RecLayout<PREC>;
  unsigned short r[RL::kSize];
#pragma unroll
  for (int f = 0; f < kD; ++f) {
    const float v = float(sc * xc[f]);
    r[RL::kP1 + f] = to16<P16>(v);
    const float r1 = v - from16<P16>(r[RL::kP1 + f]);
    r[RL::kP2 + f] = to16<P16>(r1);
    if (RL::kX6) r[RL::kP3 + f] = to16<P16>(r1 - from16<P16>(r[RL::kP2 + f]));
  }
  const unsigned short one = to16<P16>(1.0f), zero = to16<P16>(0.0f);
  unsigned short *fold = r + RL::kFold;
#pragma unroll
  for (int k = 0; k < kD; ++k) fold[k] = zero;
  unsigned short piece[3];
  if (side == 0) {
    if (pad) {
      piece[0] = to16<P16>(P16 == FSK_PREC_F16 ? -60000.0f : kPadNeg);
      piece[1] = piece[2] = zero;
    } else {
      split3<P16>(nt, piece);
    }
    fold[0] = fold[1] = fold[2] = one;
    fold[3] = piece[0];
    fold[4] = piece[1];
    fold[5] = piece[2];
    fold[6] = fold[7] = fold[8] = one;
  } else {
    split3<P16>(nt, piece);
    fold[0] = piece[0];
    fold[1] = piece[1];
    fold[2] = piece[2];
    fold[3] = fold[4] = fold[5] = one;
    // fold[6 .. 8] = pieces of -m, patched inside the kernel
  }
#pragma unroll
  for (int k = 0; k < RL::kSize; ++k) rec_out[j * RL::kSize + k] = r[k];
}

// Power-of-two prescale of the 16-bit forms: points enter as (x - mu) * s with
// s = 2^round(-log2 h), so the bandwidth becomes h s in [0.71, 1.41] and the
// 16-bit pieces resolve distances relative to h whatever the data's units
// (FP16 subnormals would otherwise swallow the low pieces of small-scale
// data). Exact in binary arithmetic; c is divided by s^2 to match.
inline double lowp_prescale(double h) { return exp2(round(-log2(h))); }

__device__ __forceinline__ f4 mfma16(u16x8 a, u16x8 b, f4 c, int prec) {
  if (prec == FSK_PREC_F16)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h16x8, a),
                                                  __builtin_bit_cast(h16x8, b), c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Column sums of the data rows (FP64, deterministic order) into mu_ws[0..15];
// mean = sum / n. mu_ws holds kMeanWs doubles (the sums, then per-block parts).
constexpr int kMeanBlocks = 256;
constexpr int64_t kMeanWs = int64_t(kD) * (1 + kMeanBlocks);
void launch_mean16(const float *x, int64_t ld, int64_t n, double *mu_ws, hipStream_t s);

// pdf[i] = sum_s pdf_part; w[i,f] = sum_s w_part + mu_f pdf[i] (FP64) for
// partials of points centred on mu = mu_sum * inv_n.
__global__ void score_lp_finish_kernel(const float *__restrict__ pdf_part,
                                       const float *__restrict__ w_part, int64_t S, int64_t n,
                                       const double *__restrict__ mu_sum, double inv_n,
                                       double inv_scale, float *__restrict__ pdf,
                                       float *__restrict__ w, int64_t ldw_row, int64_t ldw_col);

}  // namespace fsk
