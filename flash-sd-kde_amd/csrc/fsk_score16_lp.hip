// fsk_score16_lp.hip - empirical-score accumulators with 16-bit MFMA operands
// (BF16 x3 split, single BF16, single FP16), ordered split-K, deterministic.
//
// Cross term: data / query 64-value records (fsk_lowp.h), the same MFMA
// forms as the log-density kernel:
//   BF16 / F16 : [x1 | fold] . [q1 | fold]
//   BF16X3     : [x1 | x2] . [q1 | q1]  +  [x1 | fold] . [q2 | fold]
//   F16X3      : the same with FP16 pieces (phi.X as in BF16X3)
//   BF16X6     : [x1 | x1] . [q1 | q2]  +  [x2 | x1] . [q1 | q3]
//              + [x2 | x3] . [q2 | q1]  +  [x2 | fold] . [q3 | fold]
//                (FP32-class cross term; phi.X as in BF16X3)
// phi = exp2(min(t, 0)) comes out with lane l holding phi(i = l%16, j = 4(l/16)+u),
// u = 0..3, which is exactly four consecutive K slots of the 16x16x32 A operand,
// so phi.X needs no data movement:
//   A = [phi_h(u=0..3) | phi_l(u=0..3)] per lane (phi = phi_h + phi_l,
//   phi_h = phi truncated to BF16, phi_l = rn_bf16(phi - phi_h)),
//   W += A . [Xh | Xh]  +  A . [Xl | 0]      (BF16x3, BF16x6: drops phi_l.Xl)
//   W += [phi | 0] . [Xh | 0]                (single forms)
// X for phi.X comes from a tile-transposed copy: per 16-point tile, per feature
// f and lane group g, [Xh(4g+u, f) u=0..3 | Xl(4g+u, f) u=0..3] (16 bytes).
// Points are centred on the data mean: w' = sum phi (x - mu); the finish
// kernel returns w = w' + mu * pdf in FP64.
#include "fsk_common.h"
#include "fsk_lowp.h"
#include "fsk_plan.h"

namespace fsk {

constexpr int kLpFlush = 32;
static_assert(kNormPad % (64 * kScoreLpQT) == 0 && kNormPad % (64 * kScoreLpQTx6) == 0,
              "records must cover whole row blocks");

template <int P16>
__global__ void prep_xt16_kernel(const float *__restrict__ x, int64_t ld, int64_t n,
                                 int64_t n_pad, const double *__restrict__ mu_sum, double inv_n,
                                 unsigned short *__restrict__ xt) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_pad * kD) return;
  const int64_t j = e / kD;
  const int f = int(e % kD);
  const float v = j < n ? float(double(x[j * ld + f]) - mu_sum[f] * inv_n) : 0.f;
  const unsigned short hi = to16<P16>(v);
  const unsigned short lo = to16<P16>(v - from16<P16>(hi));
  const int64_t T = j / kTile;
  const int g = int((j % kTile) / 4), u = int(j % 4);
  const int64_t base = ((T * kD + f) * 4 + g) * 8;
  xt[base + u] = hi;
  xt[base + 4 + u] = lo;
}

typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef _Float16 h16x2v __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Two floats rounded (nearest even) to 16 bits and packed [a | b << 16]: one
// v_cvt_pk_bf16_f32 for BF16.
template <int P16>
__device__ __forceinline__ unsigned pack2(float a, float b) {
  const f2v v = {a, b};
  if (P16 == FSK_PREC_F16) return __builtin_bit_cast(unsigned, __builtin_convertvector(v, h16x2v));
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2v));
}

// The 16-bit A operand of phi.X from a lane's four phi values (K slots
// 0..3) and, for the split forms, their low parts (K slots 4..7). Split: the
// high part is phi truncated to its upper 16 bits (a BF16 value, exact as
// FP32), the low part phi - high (exact in FP32) rounded to BF16; a byte
// permute packs two high parts, one conversion packs two low parts.
template <int P16, bool SPLIT>
__device__ __forceinline__ u16x8 phi_operand(const float (&e)[4]) {
  u32x4 a;
  if (SPLIT) {
    unsigned b[4];
    float lo[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      b[u] = __builtin_bit_cast(unsigned, e[u]);
      lo[u] = e[u] - __builtin_bit_cast(float, b[u] & 0xffff0000u);
    }
    a[0] = __builtin_amdgcn_perm(b[1], b[0], 0x07060302u);
    a[1] = __builtin_amdgcn_perm(b[3], b[2], 0x07060302u);
    a[2] = pack2<FSK_PREC_BF16>(lo[0], lo[1]);
    a[3] = pack2<FSK_PREC_BF16>(lo[2], lo[3]);
  } else {
    a[0] = pack2<P16>(e[0], e[1]);
    a[1] = pack2<P16>(e[2], e[3]);
    a[2] = a[3] = 0u;
  }
  return __builtin_bit_cast(u16x8, a);
}

template <int PREC, int QT>
__global__ __launch_bounds__(kWG) void score16_lp_kernel(
    const unsigned short *__restrict__ xrec, const unsigned short *__restrict__ qrec,
    const unsigned short *__restrict__ xt, int64_t m, int64_t n_pad16, int64_t split_len,
    int64_t qblocks, int64_t ntasks, int64_t per_xcd, float *__restrict__ pdf_part,
    float *__restrict__ w_part) {
  constexpr int P16 = CrossType<PREC>::kP16;  // cross-term operands
  constexpr bool X6 = (PREC == FSK_PREC_BF16X6);
  constexpr bool X3 = (PREC == FSK_PREC_BF16X3 || PREC == FSK_PREC_F16X3);
  constexpr bool SPLIT_PHI = X3 || X6;  // phi.X with phi = phi_h + phi_l
  // phi.X operands: BF16 for the split forms (phi in [0, 1] keeps full
  // relative precision down to tiny values), else the form's own type
  constexpr int PX = SPLIT_PHI ? FSK_PREC_BF16 : P16;
  using RL = RecLayout<PREC>;
  const int64_t task = xcd_task(blockIdx.x, per_xcd);
  if (task >= ntasks) return;
  const int64_t split = task / qblocks;
  const int64_t qb = task - split * qblocks;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t row0 = qb * (64 * QT) + int64_t(wave) * (16 * QT);

  u16x8 B1[QT], B2[QT], B3[QT], B4[QT];
  f4 w[QT];
  d4 w64[QT];
  float pdf[QT];
  double pdf64[QT];
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    const unsigned short *rq = qrec + (row0 + 16 * R + c) * RL::kSize;  // records padded to whole row blocks
    if (X6) {
      B1[R] = ld_u16x8(rq + rec_off(g, RL::kP1, RL::kP2));
      B2[R] = ld_u16x8(rq + rec_off(g, RL::kP1, RL::kP3));
      B3[R] = ld_u16x8(rq + rec_off(g, RL::kP2, RL::kP1));
      B4[R] = ld_u16x8(rq + rec_off(g, RL::kP3, RL::kFold));
    } else if (X3) {
      B1[R] = ld_u16x8(rq + rec_off(g, RL::kP1, RL::kP1));
      B2[R] = ld_u16x8(rq + rec_off(g, RL::kP2, RL::kFold));
    } else {
      B1[R] = ld_u16x8(rq + rec_off(g, RL::kP1, RL::kFold));
    }
    w[R] = f4{0.f, 0.f, 0.f, 0.f};
    w64[R] = d4{0.0, 0.0, 0.0, 0.0};
    pdf[R] = 0.f;
    pdf64[R] = 0.0;
  }

  const int64_t j_begin = split * split_len;
  int64_t j_end = j_begin + split_len;
  if (j_end > n_pad16) j_end = n_pad16;
  int tc = 0;
  for (int64_t j0 = j_begin; j0 < j_end; j0 += kTile) {
    const unsigned short *rx = xrec + (j0 + c) * RL::kSize;
    u16x8 A1, A2, A3, A4;
    if (X6) {
      A1 = ld_u16x8(rx + rec_off(g, RL::kP1, RL::kP1));
      A2 = ld_u16x8(rx + rec_off(g, RL::kP2, RL::kP1));
      A3 = ld_u16x8(rx + rec_off(g, RL::kP2, RL::kP3));
      A4 = ld_u16x8(rx + rec_off(g, RL::kP2, RL::kFold));
    } else {
      A2 = ld_u16x8(rx + rec_off(g, RL::kP1, RL::kFold));
      if (X3) A1 = ld_u16x8(rx + rec_off(g, RL::kP1, RL::kP2));
    }
    const u16x8 XT = *reinterpret_cast<const u16x8 *>(xt + (((j0 / kTile) * kD + c) * 4 + g) * 8);
    u16x8 BX1, BX2;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      BX1[u] = XT[u];
      BX1[4 + u] = SPLIT_PHI ? XT[u] : (unsigned short)0;
      BX2[u] = XT[4 + u];
      BX2[4 + u] = 0;
    }
#pragma unroll
    for (int R = 0; R < QT; ++R) {
      f4 d;
      if (X6) {
        d = mfma16(A1, B1[R], f4{0.f, 0.f, 0.f, 0.f}, P16);
        d = mfma16(A2, B2[R], d, P16);
        d = mfma16(A3, B3[R], d, P16);
        d = mfma16(A4, B4[R], d, P16);
      } else if (X3) {
        d = mfma16(A1, B1[R], f4{0.f, 0.f, 0.f, 0.f}, P16);
        d = mfma16(A2, B2[R], d, P16);
      } else {
        d = mfma16(A2, B1[R], f4{0.f, 0.f, 0.f, 0.f}, P16);
      }
      float e[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        e[u] = fexp2_le1(d[u]);
        pdf[R] += e[u];
      }
      const u16x8 A = phi_operand<PX, SPLIT_PHI>(e);
      w[R] = mfma16(A, BX1, w[R], PX);
      if (SPLIT_PHI) w[R] = mfma16(A, BX2, w[R], PX);
    }
    if (++tc == kLpFlush) {
      tc = 0;
#pragma unroll
      for (int R = 0; R < QT; ++R) {
        pdf64[R] += pdf[R];
        pdf[R] = 0.f;
#pragma unroll
        for (int r = 0; r < 4; ++r) w64[R][r] += w[R][r];
        w[R] = f4{0.f, 0.f, 0.f, 0.f};
      }
    }
  }
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    double p = sum_groups(pdf64[R] + pdf[R]);
    int64_t i = row0 + 16 * R + c;
    if (g == 0 && i < m) pdf_part[split * m + i] = float(p);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int64_t iw = row0 + 16 * R + 4 * g + r;
      if (iw < m) w_part[(split * m + iw) * kD + c] = float(w64[R][r] + w[R][r]);
    }
  }
}

// pdf[i] = sum_s pdf_part; w[i,f] = sum_s w_part + mu_f pdf[i]  (FP64), n rows.
__global__ void score_lp_finish_kernel(const float *__restrict__ pdf_part,
                                       const float *__restrict__ w_part, int64_t S, int64_t n,
                                       const double *__restrict__ mu_sum, double inv_n,
                                       float *__restrict__ pdf, float *__restrict__ w,
                                       int64_t ldw_row, int64_t ldw_col) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * kD) return;
  const int64_t i = e / kD;
  const int f = int(e % kD);
  double p = 0.0, s = 0.0;
  for (int64_t k = 0; k < S; ++k) {
    p += double(pdf_part[k * n + i]);
    s += double(w_part[(k * n + i) * kD + f]);
  }
  w[i * ldw_row + f * ldw_col] = float(s + mu_sum[f] * inv_n * p);
  if (f == 0) pdf[i] = float(p);
}

struct ScoreLpLayout {
  double *mu;
  unsigned short *xrec, *qrec, *xt;
  float *pdf_part, *w_part;
};

static size_t score_lp_layout(void *ws, int64_t n, int64_t m, unsigned prec, ScoreLpLayout *L) {
  Carve cv(ws);
  SplitPlan p = score16_lp_plan(n, m, prec);
  L->mu = cv.take<double>(kD);
  L->xrec = cv.take<unsigned short>(round_up(n, kNormPad) * rec_size(prec));
  L->qrec = cv.take<unsigned short>(round_up(m, kNormPad) * rec_size(prec));
  L->xt = cv.take<unsigned short>(round_up(n, kTile) * kD * 2);
  L->pdf_part = cv.take<float>(p.splits * m);
  L->w_part = cv.take<float>(p.splits * m * kD);
  return cv.used;
}

size_t score16_lp_ws_bytes(int64_t n, int64_t m, unsigned prec) {
  ScoreLpLayout L;
  return score_lp_layout(nullptr, n, m, prec, &L);
}

// Score accumulators of m rows against the KDE of n data points with 16-bit
// MFMA operands (rows == data for the empirical score of the data set).
int score16_lp_run(const char *what, const float *rows_in, int64_t ldr, int64_t m,
                   const float *data, int64_t ld, int64_t n, double h, unsigned prec, void *ws,
                   size_t ws_bytes, float *pdf, float *w, int64_t ldw_row, int64_t ldw_col,
                   void *stream) {
  ScoreLpLayout L;
  size_t need = score_lp_layout(ws, n, m, prec, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const double coef = exp2_coef(h);
  const int64_t xrows = round_up(n, kNormPad), qrows = round_up(m, kNormPad);
  const int64_t n16 = round_up(n, kTile);
  const double inv_n = 1.0 / double(n);
  hipLaunchKernelGGL(fill_f64, dim3(1), dim3(64), 0, s, L.mu, int64_t(kD), 0.0);
  {
    int64_t blocks = cdiv(n, 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(mean16_kernel, dim3(blocks), dim3(256), 0, s, data, ld, n, L.mu);
  }
  SplitPlan p = score16_lp_plan(n, m, prec);
  const int64_t ntasks = p.qblocks * p.splits;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
#define FSK_LP(P, P16, QT)                                                                       \
  do {                                                                                           \
    hipLaunchKernelGGL(prep_lse_kernel<P>, dim3(cdiv(xrows, 256)), dim3(256), 0, s, data, ld, n, \
                       xrows, L.mu, inv_n, coef, 0, (float *)nullptr, (float *)nullptr, L.xrec);  \
    hipLaunchKernelGGL(prep_lse_kernel<P>, dim3(cdiv(qrows, 256)), dim3(256), 0, s, rows_in,     \
                       ldr, m, qrows, L.mu, inv_n, coef, 1, (float *)nullptr, (float *)nullptr,   \
                       L.qrec);                                                                  \
    hipLaunchKernelGGL(prep_xt16_kernel<P16>, dim3(cdiv(n16 * kD, 256)), dim3(256), 0, s, data,  \
                       ld, n, n16, L.mu, inv_n, L.xt);                                           \
    prof_begin(FSK_OP_SCORE16_ORDERED, s);                                                       \
    hipLaunchKernelGGL((score16_lp_kernel<P, QT>), dim3(grid), dim3(kWG), 0, s, L.xrec, L.qrec,  \
                       L.xt, m, n16, p.split_len, p.qblocks, ntasks, per_xcd, L.pdf_part,        \
                       L.w_part);                                                                \
    prof_end(FSK_OP_SCORE16_ORDERED, s);                                                         \
  } while (0)
  switch (prec) {
    case FSK_PREC_BF16: FSK_LP(FSK_PREC_BF16, FSK_PREC_BF16, kScoreLpQT); break;
    case FSK_PREC_F16: FSK_LP(FSK_PREC_F16, FSK_PREC_F16, kScoreLpQT); break;
    case FSK_PREC_BF16X6: FSK_LP(FSK_PREC_BF16X6, FSK_PREC_BF16, kScoreLpQTx6); break;
    case FSK_PREC_F16X3: FSK_LP(FSK_PREC_F16X3, FSK_PREC_BF16, kScoreLpQT); break;
    default: FSK_LP(FSK_PREC_BF16X3, FSK_PREC_BF16, kScoreLpQT); break;
  }
#undef FSK_LP
  int st = launch_status(what);
  if (st != FSK_OK) return st;
  hipLaunchKernelGGL(score_lp_finish_kernel, dim3(cdiv(m * kD, 256)), dim3(256), 0, s,
                     L.pdf_part, L.w_part, p.splits, m, L.mu, inv_n, pdf, w, ldw_row, ldw_col);
  return launch_status(what);
}

}  // namespace fsk
