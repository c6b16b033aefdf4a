// fsk_lowp.h - 16-bit (BF16 / FP16) operand helpers shared by the
// reduced-precision pair kernels.
//
// Packed 48-value records per point (K = 32 of v_mfma_f32_16x16x32_*):
//   data  (A side): [ xh[16] | xl[16] | 1 1 1 bh bm bl 1 1 1 0 ... ]
//   query (B side): [ qh[16] | ql[16] | ah am al 1 1 1 mh mm ml 0 ... ]
// with x = xh + xl the centred point, q' = -2c (q - mu) = qh + ql, b = c|x - mu|^2,
// a = c|q - mu|^2 and -m (log-sum-exp shift, 0 outside the LSE kernel) each
// split into three 16-bit pieces, so one MFMA over the fold block adds the
// norm terms (and -m) with no per-pair VALU work.
#pragma once

#include "fsk_common.h"

namespace fsk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

constexpr int kRec = 48;            // 16-bit values per packed record

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
                                double coef, int side, float *__restrict__ f32_out,
                                float *__restrict__ norm_out,
                                unsigned short *__restrict__ rec_out) {
  int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= rows_pad) return;
  const bool pad = j >= rows;
  double xc[kD];
  double nrm = 0.0;
#pragma unroll
  for (int f = 0; f < kD; ++f) {
    double v = pad ? 0.0 : double(x[j * ld + f]) - mu_sum[f] * inv_n;
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
  constexpr int P16 = (PREC == FSK_PREC_F16) ? FSK_PREC_F16 : FSK_PREC_BF16;
  unsigned short r[kRec];
#pragma unroll
  for (int f = 0; f < kD; ++f) {
    float v = float(sc * xc[f]);
    r[f] = to16<P16>(v);
    r[kD + f] = to16<P16>(v - from16<P16>(r[f]));
  }
  const unsigned short one = to16<P16>(1.0f), zero = to16<P16>(0.0f);
#pragma unroll
  for (int k = 0; k < kD; ++k) r[2 * kD + k] = zero;
  unsigned short piece[3];
  if (side == 0) {
    if (pad) {
      piece[0] = to16<P16>(P16 == FSK_PREC_F16 ? -60000.0f : kPadNeg);
      piece[1] = piece[2] = zero;
    } else {
      split3<P16>(nt, piece);
    }
    r[32] = r[33] = r[34] = one;
    r[35] = piece[0];
    r[36] = piece[1];
    r[37] = piece[2];
    r[38] = r[39] = r[40] = one;
  } else {
    split3<P16>(nt, piece);
    r[32] = piece[0];
    r[33] = piece[1];
    r[34] = piece[2];
    r[35] = r[36] = r[37] = one;
    // r[38..40] = pieces of -m, patched inside the kernel
  }
#pragma unroll
  for (int k = 0; k < kRec; ++k) rec_out[j * kRec + k] = r[k];
}

__device__ __forceinline__ f4 mfma16(u16x8 a, u16x8 b, f4 c, int prec) {
  if (prec == FSK_PREC_F16)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h16x8, a),
                                                  __builtin_bit_cast(h16x8, b), c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Sum of the data rows (FP64) into sum[16] (pre-zeroed); mean = sum / n.
__global__ void mean16_kernel(const float *__restrict__ x, int64_t ld, int64_t n,
                              double *__restrict__ sum);

}  // namespace fsk
