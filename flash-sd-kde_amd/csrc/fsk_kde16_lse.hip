// fsk_kde16_lse.hip - 16-D log-density by log-sum-exp, with a choice of
// cross-term precision (the precision form).
//
// Data and queries are first centred on the data mean (distances are
// translation invariant; centring shrinks |x|^2 and with it the cancellation
// in |q|^2 + |x|^2 - 2 q.x). In base 2, with c = -log2(e)/(2h^2),
//     t_ij = c|q_i|^2 + c|x_j|^2 - 2c q_i.x_j  (<= 0 up to rounding)
// and each (query, split) keeps a shift m_i (shared by the four lanes that
// hold query i) and s_i = sum_j exp2(min(t_ij, 0) - m_i). The shift is the
// exact maximum of the split's first tile, raised afterwards whenever a term
// exceeds it (wave-uniform branch); a query whose first-tile maximum is above
// -100 keeps m_i = 0 (no term can overflow, none that matters underflows),
// and a wave whose queries all do runs a loop with no maximum tracking at
// all: one clamped exp and one add per pair. Partials (m, s) are combined in
// FP64: log p_i = ln2 (M + log2 sum_s s_s 2^(m_s - M)) + log_const.
//
// 16-bit forms use v_mfma_f32_16x16x32_{bf16,f16} on 64-value records
// (fsk_lowp.h): x = x1 + x2 + x3, q' = -2c q = q1 + q2 + q3 in 16-bit pieces,
// plus a "fold" block that adds the norm terms and -m_i inside the MFMA, so no
// per-pair VALU add is needed ([lo half | hi half] of K = 32 per operand):
//   BF16 / F16 : one MFMA,    [x1 | fold] . [q1 | fold]
//   BF16X3     : two MFMAs,   [x1 | x2] . [q1 | q1]  +  [x1 | fold] . [q2 | fold]
//                (x1q1 + x2q1 + x1q2 + norms - m; the 2^-16 terms dropped)
//   F16X3      : the same with FP16 pieces (11-bit significands: the dropped
//                x2q2 term is ~2^-22 relative; values are centred and scaled so
//                they stay inside FP16's range)
//   BF16X6     : four MFMAs,  [x1 | x1] . [q1 | q2]  +  [x2 | x1] . [q1 | q3]
//                           + [x2 | x3] . [q2 | q1]  +  [x2 | fold] . [q3 | fold]
//                (all products down to 2^-16 relative plus x2q3; the dropped
//                x3q2 and x3q3 are at or below FP32 rounding: an FP32-class
//                cross term at BF16 MFMA rate)
// -m_i is split into three pieces and patched into the B operand whose high
// half is the fold block.
// FP32 forms: MFMA f32 16x16x4 with the norm terms (and -m) in the
// accumulator initialiser, or plain FP32 FMAs on the VALU.
#include "fsk_common.h"
#include "fsk_far.h"
#include "fsk_plan.h"
#include "fsk_lowp.h"

namespace fsk {

constexpr float kLseSlack = 8.0f;  // raise m only when a term exceeds 2^8 * 2^m
// A split whose first tile already holds a term above 2^-kLseDirect sums
// exp2(t) unshifted (m = 0): no term can overflow (t <= 0) and the sum keeps
// full relative precision down to its largest term.
constexpr float kLseDirect = 100.0f;
constexpr int kLseFlush = 32;

// Column sums of the data in FP64, deterministic: block b writes its 16 sums
// to part[16 b ..], one thread per feature then adds the blocks in order.
__global__ void mean16_kernel(const float *__restrict__ x, int64_t ld, int64_t n,
                              double *__restrict__ part) {
  __shared__ double red[256][kD + 1];
  double acc[kD];
#pragma unroll
  for (int f = 0; f < kD; ++f) acc[f] = 0.0;
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n;
       j += int64_t(gridDim.x) * blockDim.x) {
#pragma unroll
    for (int f = 0; f < kD; f += 4) {
      f4 v = ld4(x + j * ld + f);
      acc[f] += v[0];
      acc[f + 1] += v[1];
      acc[f + 2] += v[2];
      acc[f + 3] += v[3];
    }
  }
#pragma unroll
  for (int f = 0; f < kD; ++f) red[threadIdx.x][f] = acc[f];
  __syncthreads();
  if (threadIdx.x < kD) {
    double s = 0.0;
    for (int t = 0; t < 256; ++t) s += red[t][threadIdx.x];
    part[int64_t(blockIdx.x) * kD + threadIdx.x] = s;
  }
}

__global__ void mean16_finish_kernel(const double *__restrict__ part, int nblocks,
                                     double *__restrict__ sum) {
  if (threadIdx.x >= kD) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += part[b * kD + threadIdx.x];
  sum[threadIdx.x] = s;
}

void launch_mean16(const float *x, int64_t ld, int64_t n, double *mu_ws, hipStream_t s) {
  int64_t blocks = cdiv(n, 256);
  if (blocks > kMeanBlocks) blocks = kMeanBlocks;
  hipLaunchKernelGGL(mean16_kernel, dim3(blocks), dim3(256), 0, s, x, ld, n, mu_ws + kD);
  hipLaunchKernelGGL(mean16_finish_kernel, dim3(1), dim3(64), 0, s, mu_ws + kD, int(blocks), mu_ws);
}

__device__ __forceinline__ float max4(f4 v) { return fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3])); }
__device__ __forceinline__ float gmax(float v) {
  v = fmaxf(v, __shfl_xor(v, 16));
  v = fmaxf(v, __shfl_xor(v, 32));
  return v;
}

// Waves per SIMD the compiler must leave room for: 4 (at most 128 VGPRs), 3
// for BF16X6 (its four query operands per row tile).
template <int PREC>
constexpr int lse_min_waves() { return PREC == FSK_PREC_BF16X6 ? 3 : 4; }

// SKIP instances (16-bit forms) take per-query lower bounds `lower` of
// log2 sum_j exp2(t_ij) (from a pass over a sample of the data) and leave out
// terms below lower - margin, margin = ceil(log2 n) + 28: a query loses at
// most n 2^(lower - margin) <= 2^-28 of its sum. Pairs of data tiles the
// far-block bitmap (fsk_far.h) marks are not read; of the others, a tile
// whose every exponent is below every threshold of the wave's rows adds
// nothing. Waves whose rows all have lower > direct_min (their largest term
// is then above 2^-100) run the unshifted sum; the others run the dense
// loops below.
template <int PREC, int QT, bool SKIP>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(lse_min_waves<PREC>())))
void kde16_lse_mfma_kernel(
    const float *__restrict__ xf, const float *__restrict__ bn,
    const unsigned short *__restrict__ xrec, int64_t n_pad, const float *__restrict__ qf,
    const float *__restrict__ qa, const unsigned short *__restrict__ qrec, int64_t m,
    int64_t split_len, int64_t qblocks, int64_t ntasks, int64_t per_xcd,
    const float *__restrict__ lower, float margin, float direct_min,
    const unsigned long long *__restrict__ far_bits, int64_t far_words,
    float2 *__restrict__ part) {
  constexpr bool F32 = (PREC == FSK_PREC_F32);
  constexpr bool X3 = (PREC == FSK_PREC_BF16X3 || PREC == FSK_PREC_F16X3);
  constexpr bool X6 = (PREC == FSK_PREC_BF16X6);
  constexpr int P16 = CrossType<PREC>::kP16;
  using RL = RecLayout<PREC>;
  static_assert(!SKIP || !F32, "the 16-bit forms leave out terms");
  // SKIP: task = workgroup index, so the round-robin dispatch spreads the
  // row blocks (whose work now depends on their neighbourhood) over the XCDs
  const int64_t task = SKIP ? int64_t(blockIdx.x) : xcd_task(blockIdx.x, per_xcd);
  if (task >= ntasks) return;
  const int64_t split = task / qblocks;
  const int64_t qb = task - split * qblocks;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t row0 = qb * (64 * QT) + int64_t(wave) * (16 * QT);

  f4 qv[QT];
  u16x8 B1[QT], B2[QT], B3[QT], B4[QT];
  float qav[QT], negm[QT], s[QT];
  double s64[QT];
  double mrec[QT];  // exactly the m subtracted inside the MFMA
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    int64_t i = row0 + 16 * R + c;
    if (i > m - 1) i = m - 1;
    if (F32) {
      qv[R] = ld4(qf + i * kD + 4 * g);
      qav[R] = qa[i];
    } else {
      const unsigned short *rq = qrec + i * RL::kSize;
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
    }
    negm[R] = 0.f;
    mrec[R] = 0.0;
    s[R] = 0.f;
    s64[R] = 0.0;
  }

  auto tile = [&](int64_t j0, f4 *d) {
    if (F32) {
      const f4 xa = ld4(xf + (j0 + c) * kD + 4 * g);
      const f4 bb = ld4(bn + j0 + 4 * g);
#pragma unroll
      for (int R = 0; R < QT; ++R) {
        d[R][0] = qav[R] + bb[0];
        d[R][1] = qav[R] + bb[1];
        d[R][2] = qav[R] + bb[2];
        d[R][3] = qav[R] + bb[3];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int R = 0; R < QT; ++R) d[R] = mfma_f32(xa[k], qv[R][k], d[R]);
      }
    } else {
      const unsigned short *rx = xrec + (j0 + c) * RL::kSize;
      if (X6) {
        const u16x8 A1 = ld_u16x8(rx + rec_off(g, RL::kP1, RL::kP1));
        const u16x8 A2 = ld_u16x8(rx + rec_off(g, RL::kP2, RL::kP1));
        const u16x8 A3 = ld_u16x8(rx + rec_off(g, RL::kP2, RL::kP3));
        const u16x8 A4 = ld_u16x8(rx + rec_off(g, RL::kP2, RL::kFold));
#pragma unroll
        for (int R = 0; R < QT; ++R) {
          d[R] = mfma16(A1, B1[R], f4{0.f, 0.f, 0.f, 0.f}, P16);
          d[R] = mfma16(A2, B2[R], d[R], P16);
          d[R] = mfma16(A3, B3[R], d[R], P16);
          d[R] = mfma16(A4, B4[R], d[R], P16);
        }
      } else if (X3) {
        const u16x8 A1 = ld_u16x8(rx + rec_off(g, RL::kP1, RL::kP2));
        const u16x8 A2 = ld_u16x8(rx + rec_off(g, RL::kP1, RL::kFold));
#pragma unroll
        for (int R = 0; R < QT; ++R) {
          d[R] = mfma16(A1, B1[R], f4{0.f, 0.f, 0.f, 0.f}, P16);
          d[R] = mfma16(A2, B2[R], d[R], P16);
        }
      } else {
        const u16x8 A2 = ld_u16x8(rx + rec_off(g, RL::kP1, RL::kFold));
#pragma unroll
        for (int R = 0; R < QT; ++R) d[R] = mfma16(A2, B1[R], f4{0.f, 0.f, 0.f, 0.f}, P16);
      }
    }
  };

  // Install a new maximum for query row tile R (uniform over the 4 lanes of a query).
  auto set_m = [&](int R, float mnew) {
    if (F32) {
      const int64_t i0 = row0 + 16 * R + c;
      const float base = qa[i0 > m - 1 ? m - 1 : i0];
      qav[R] = base - mnew;
      mrec[R] = double(mnew);
      negm[R] = -mnew;
    } else {
      unsigned short pc[3];
      split3<P16>(-double(mnew), pc);
      double mr = -(double(from16<P16>(pc[0])) + double(from16<P16>(pc[1])) +
                    double(from16<P16>(pc[2])));
      // fold slots 6, 7 live in lane group 2 (elements 6, 7), slot 8 in group 3
      // (element 0) of the operand whose high K half is the fold block
      u16x8 &Bf = X6 ? B4[R] : (X3 ? B2[R] : B1[R]);
      if (g == 2) {
        Bf[6] = pc[0];
        Bf[7] = pc[1];
      }
      if (g == 3) Bf[0] = pc[2];
      mrec[R] = mr;
      negm[R] = float(-mr);
    }
  };

  const int64_t j_begin = split * split_len;
  int64_t j_end = j_begin + split_len;
  if (j_end > n_pad) j_end = n_pad;

  f4 d[QT];
  int tc = 0;
  auto flush = [&]() {
#pragma unroll
    for (int R = 0; R < QT; ++R) {
      s64[R] += s[R];
      s[R] = 0.f;
    }
  };
  if constexpr (SKIP) {
    float thr[QT];
    bool direct = true;
#pragma unroll
    for (int R = 0; R < QT; ++R) {
      int64_t i = row0 + 16 * R + c;
      if (i > m - 1) i = m - 1;
      const float lo = lower[i];
      thr[R] = lo - margin;
      direct = direct && lo > direct_min;
    }
    if (__all(direct)) {
      // unshifted sums over the pairs of data tiles the bitmap leaves in,
      // 64 pairs per word (p: pair index, tiles 2p and 2p + 1)
      const unsigned long long *far_row =
          far_bits ? far_bits + (qb * (kWG / 64) + wave) * far_words : nullptr;
      const int64_t p0 = j_begin / (2 * kTile), p1 = (j_end + 2 * kTile - 1) / (2 * kTile);
      for (int64_t w = p0 >> 6; w <= (p1 - 1) >> 6; ++w) {
        unsigned long long near = far_row ? ~sload_u64(far_row + w) : ~0ull;
        if (w == (p0 >> 6)) near &= ~0ull << (p0 & 63);
        if (w == ((p1 - 1) >> 6) && (p1 & 63)) near &= ~0ull >> (64 - (p1 & 63));
        while (near) {
          const int64_t p = (w << 6) + __builtin_ctzll(near);
          near &= near - 1;
#pragma unroll
          for (int tt = 0; tt < 2; ++tt) {
            const int64_t j0 = p * (2 * kTile) + tt * kTile;
            if (j0 < j_begin || j0 >= j_end) continue;
            tile(j0, d);
            bool hot = false;
#pragma unroll
            for (int R = 0; R < QT; ++R) hot = hot || max4(d[R]) > thr[R];
            if (__any(hot)) {
#pragma unroll
              for (int R = 0; R < QT; ++R) {
#pragma unroll
                for (int r = 0; r < 4; ++r) s[R] += fexp2_le1(d[R][r]);
              }
            }
            if (++tc == kLseFlush) {
              tc = 0;
              flush();
            }
          }
        }
      }
#pragma unroll
      for (int R = 0; R < QT; ++R) {
        double v = sum_groups(s64[R] + s[R]);
        int64_t i = row0 + 16 * R + c;
        if (g == 0 && i < m) part[split * m + i] = make_float2(0.f, float(v));
      }
      return;
    }
  }
  // First tile: exact per-query maximum of this split's first 16 points; a
  // query whose maximum is above -kLseDirect keeps m = 0.
  tile(j_begin, d);
  bool direct = true;
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    float mx = fminf(gmax(max4(d[R])), 0.f);
    if (mx > -kLseDirect) mx = 0.f;
    direct = direct && mx == 0.f;
    set_m(R, mx);
    const float dm = float(mrec[R]);
#pragma unroll
    for (int r = 0; r < 4; ++r) s[R] += fexp2(fminf(d[R][r] - dm, negm[R]));
  }
  tc = 1;
  if (__all(direct)) {
    // every query of the wave at m = 0: per pair one clamped exp (exp2 of
    // min(t, 0) as v_exp ... clamp) and one add; no maximum to track
    for (int64_t j0 = j_begin + kTile; j0 < j_end; j0 += kTile) {
      tile(j0, d);
#pragma unroll
      for (int R = 0; R < QT; ++R) {
#pragma unroll
        for (int r = 0; r < 4; ++r) s[R] += fexp2_le1(d[R][r]);
      }
      if (++tc == kLseFlush) {
        tc = 0;
        flush();
      }
    }
  } else {
    // some query of the wave keeps a shift: the running maximum as above
    for (int64_t j0 = j_begin + kTile; j0 < j_end; j0 += kTile) {
      tile(j0, d);
      bool hi = false;
#pragma unroll
      for (int R = 0; R < QT; ++R) hi |= max4(d[R]) > kLseSlack;
      if (__any(hi)) {
#pragma unroll
        for (int R = 0; R < QT; ++R) {
          float up = gmax(max4(d[R]));
          if (up > 0.f) {
            // raise m by `up` (bounded by the clamp t <= 0)
            float mnew = fminf(float(mrec[R]) + up, 0.f);
            float old = float(mrec[R]);
            set_m(R, mnew);
            const float sc = fexp2(float(-(mrec[R] - double(old))));
            s[R] *= sc;
            s64[R] *= double(sc);
            const float dd = float(mrec[R] - double(old));
            d[R] -= dd;
          }
        }
      }
#pragma unroll
      for (int R = 0; R < QT; ++R) {
#pragma unroll
        for (int r = 0; r < 4; ++r) s[R] += fexp2(fminf(d[R][r], negm[R]));
      }
      if (++tc == kLseFlush) {
        tc = 0;
        flush();
      }
    }
  }
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    double v = sum_groups(s64[R] + s[R]);
    int64_t i = row0 + 16 * R + c;
    if (g == 0 && i < m) part[split * m + i] = make_float2(float(mrec[R]), float(v));
  }
}

// FP32 cross term on the VALU: one thread per query, data staged in LDS.
constexpr int kValuChunk = 128;
__global__ __launch_bounds__(256) void kde16_lse_valu_kernel(
    const float *__restrict__ xf, const float *__restrict__ bn, int64_t n_pad,
    const float *__restrict__ qf, const float *__restrict__ qa, int64_t m, int64_t split_len,
    int64_t qblocks, int64_t ntasks, int64_t per_xcd, float2 *__restrict__ part) {
  __shared__ float xs[kValuChunk][kD + 1];
  __shared__ float bs[kValuChunk];
  const int64_t task = xcd_task(blockIdx.x, per_xcd);
  if (task >= ntasks) return;
  const int64_t split = task / qblocks;
  const int64_t qb = task - split * qblocks;
  const int tid = threadIdx.x;
  int64_t i = qb * 256 + tid;
  const int64_t ic = i > m - 1 ? m - 1 : i;
  float q[kD];
#pragma unroll
  for (int f = 0; f < kD; ++f) q[f] = qf[ic * kD + f];
  const float a = qa[ic];
  float mx = -INFINITY, s = 0.f;
  double s64 = 0.0;
  const int64_t j_begin = split * split_len;
  int64_t j_end = j_begin + split_len;
  if (j_end > n_pad) j_end = n_pad;
  for (int64_t j0 = j_begin; j0 < j_end; j0 += kValuChunk) {
    __syncthreads();
    for (int e = tid; e < kValuChunk * kD; e += 256) {
      int64_t j = j0 + e / kD;
      xs[e / kD][e % kD] = j < n_pad ? xf[j * kD + (e % kD)] : 0.f;
    }
    for (int e = tid; e < kValuChunk; e += 256) bs[e] = (j0 + e) < n_pad ? bn[j0 + e] : kPadNeg;
    __syncthreads();
    for (int jj = 0; jj < kValuChunk; jj += 16) {
      float t[16];
      float tm = -INFINITY;
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        float acc = a + bs[jj + u];
#pragma unroll
        for (int f = 0; f < kD; ++f) acc = fmaf(q[f], xs[jj + u][f], acc);
        t[u] = fminf(acc, 0.f);
        tm = fmaxf(tm, t[u]);
      }
      if (tm > mx) {
        float sc = fexp2(mx - tm);  // mx = -inf on the first group -> 0
        s *= sc;
        s64 *= double(sc);
        mx = tm;
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) s += fexp2(t[u] - mx);
    }
    s64 += s;
    s = 0.f;
  }
  if (i < m) part[split * m + i] = make_float2(mx, float(s64));
}

// FP64 combine of (m, s) partials into natural-log densities.
__global__ void lse_combine_kernel(const float2 *__restrict__ part, int64_t S, int64_t m,
                                   double log_const, double *__restrict__ out) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double M = -INFINITY;
  for (int64_t s = 0; s < S; ++s) {
    float2 p = part[s * m + i];
    if (p.y > 0.f && double(p.x) > M) M = p.x;
  }
  double acc = 0.0;
  for (int64_t s = 0; s < S; ++s) {
    float2 p = part[s * m + i];
    if (p.y > 0.f) acc += double(p.y) * exp2(double(p.x) - M);
  }
  out[i] = (acc > 0.0) ? kLn2 * (M + log2(acc)) + log_const : -INFINITY;
}

// log2 of each query's sum from split partials (m_s, s_s): the lower bounds
// of the SKIP instances when the partials come from a sample of the data
// (-inf where no term survived).
__global__ void lse_lower_kernel(const float2 *__restrict__ part, int64_t S, int64_t m,
                                 float *__restrict__ lower) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double M = -INFINITY;
  for (int64_t s = 0; s < S; ++s) {
    float2 p = part[s * m + i];
    if (p.y > 0.f && double(p.x) > M) M = p.x;
  }
  double acc = 0.0;
  for (int64_t s = 0; s < S; ++s) {
    float2 p = part[s * m + i];
    if (p.y > 0.f) acc += double(p.y) * exp2(double(p.x) - M);
  }
  lower[i] = acc > 0.0 ? float(M + log2(acc)) : -INFINITY;
}

// Margin of the SKIP instances below a query's lower bound (log2 units):
// n terms below 2^-(ceil(log2 n) + 28) of the sum add at most 2^-28 of it.
inline double lse_margin(int64_t n) { return std::ceil(std::log2(double(n))) + 28.0; }

// The 16-bit log-densities leave out negligible terms (SKIP instances) from
// this many data points on; the lower bounds come from every k-th point with
// k = n / kLseSample.
constexpr int64_t kLseSkipMinN = 16384;
constexpr int64_t kLseSample = 16384;

struct LseLayout {
  double *mu;
  float *xf, *bn, *qf, *qa;
  unsigned short *xrec, *qrec;
  float2 *part;
  // SKIP: sample records and partials, lower bounds, far-block bitmap
  unsigned short *srec;
  float2 *spart;
  float *lower, *frame, *pair_boxes, *row_boxes;
  unsigned long long *far_bits;
  int64_t stride, ns, npairs;
};

static bool lse_is16(unsigned prec) {
  return prec == FSK_PREC_BF16 || prec == FSK_PREC_F16 || prec == FSK_PREC_BF16X3 ||
         prec == FSK_PREC_BF16X6 || prec == FSK_PREC_F16X3 || prec == FSK_PREC_F16X3S ||
         prec == FSK_PREC_BF16X3S;
}

static size_t lse_layout(void *ws, int64_t n, int64_t m, unsigned flags, LseLayout *L) {
  Carve cv(ws);
  const unsigned prec = flags & FSK_PREC_MASK;
  SplitPlan p = kde16_plan(n, m);
  const int64_t n_pad = round_up(n, kValuChunk);
  L->mu = cv.take<double>(kMeanWs);
  L->xf = L->bn = L->qf = L->qa = nullptr;
  L->xrec = L->qrec = nullptr;
  if (lse_is16(prec)) {
    L->xrec = cv.take<unsigned short>(n_pad * rec_size(prec));
    L->qrec = cv.take<unsigned short>(m * rec_size(prec));
  } else {
    L->xf = cv.take<float>(n_pad * kD);
    L->bn = cv.take<float>(n_pad);
    L->qf = cv.take<float>(m * kD);
    L->qa = cv.take<float>(m);
  }
  L->part = cv.take<float2>(p.splits * m);
  L->srec = nullptr;
  L->spart = nullptr;
  L->lower = L->frame = L->pair_boxes = L->row_boxes = nullptr;
  L->far_bits = nullptr;
  L->stride = L->ns = L->npairs = 0;
  if (lse_is16(prec) && n >= kLseSkipMinN) {
    L->stride = n / kLseSample > 1 ? n / kLseSample : 1;
    L->ns = cdiv(n, L->stride);
    L->srec = cv.take<unsigned short>(round_up(L->ns, kValuChunk) * rec_size(prec));
    L->spart = cv.take<float2>(kde16_plan(L->ns, m).splits * m);
    L->lower = cv.take<float>(m);
    L->npairs = cdiv(n, kFarPairPts);
    if (far_bits_count(p.qblocks, kWG / 64, L->npairs) * 8 <= kFarBitsCap) {
      L->frame = cv.take<float>(kBoxAxes * kD);
      L->pair_boxes = cv.take<float>(L->npairs * kBoxFloats);
      L->row_boxes = cv.take<float>(p.qblocks * (kWG / 64) * kBoxFloats);
      L->far_bits = cv.take<unsigned long long>(far_bits_count(p.qblocks, kWG / 64, L->npairs));
    }
  }
  return cv.used;
}

size_t kde16_lse_ws_bytes(int64_t n, int64_t m, unsigned flags) {
  LseLayout L;
  return lse_layout(nullptr, n, m, flags, &L);
}

}  // namespace fsk

using namespace fsk;

extern "C" int fsk_kde16_logdensity(const float *data, int64_t ld, int64_t n,
                                    const float *query, int64_t ldq, int64_t m,
                                    double bandwidth, unsigned flags, void *ws,
                                    size_t ws_bytes, double *out, void *stream) {
  const char *what = "fsk_kde16_logdensity";
  int st;
  if ((st = check_points16(what, data, ld, n)) != FSK_OK) return st;
  if ((st = check_points16(what, query, ldq, m)) != FSK_OK) return st;
  if (!(bandwidth > 0.0) || !isfinite(bandwidth)) {
    set_error("%s: bandwidth must be positive, got %g", what, bandwidth);
    return FSK_EINVAL;
  }
  if (!out) {
    set_error("%s: null output", what);
    return FSK_EINVAL;
  }
  const unsigned prec = flags & FSK_PREC_MASK;
  if (prec > FSK_PREC_BF16X3S || (flags & ~(FSK_PREC_MASK | FSK_FLAG_DENSE))) {
    set_error("%s: flags 0x%x not available (Laplacian has no log form)", what, flags);
    return FSK_EUNSUPPORTED;
  }
  LseLayout L;
  size_t need = lse_layout(ws, n, m, flags, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const double coef = exp2_coef(bandwidth);
  const int64_t n_pad = round_up(n, kValuChunk);
  launch_mean16(data, ld, n, L.mu, s);
  const double inv_n = 1.0 / double(n);
  const double scl = lse_is16(prec) ? lowp_prescale(bandwidth) : 1.0;
  const double coef_s = coef / (scl * scl);
#define FSK_PREP(P)                                                                             \
  do {                                                                                          \
    hipLaunchKernelGGL(prep_lse_kernel<P>, dim3(cdiv(n_pad, 256)), dim3(256), 0, s, data, ld, n, \
                       n_pad, L.mu, inv_n, scl, coef_s, 0, L.xf, L.bn, L.xrec);                 \
    hipLaunchKernelGGL(prep_lse_kernel<P>, dim3(cdiv(m, 256)), dim3(256), 0, s, query, ldq, m,  \
                       m, L.mu, inv_n, scl, coef_s, 1, L.qf, L.qa, L.qrec);                     \
  } while (0)
  switch (prec) {
    case FSK_PREC_F32: FSK_PREP(FSK_PREC_F32); break;
    case FSK_PREC_F32_VALU: FSK_PREP(FSK_PREC_F32_VALU); break;
    case FSK_PREC_BF16: FSK_PREP(FSK_PREC_BF16); break;
    case FSK_PREC_F16: FSK_PREP(FSK_PREC_F16); break;
    case FSK_PREC_BF16X6: FSK_PREP(FSK_PREC_BF16X6); break;
    case FSK_PREC_F16X3:
    case FSK_PREC_F16X3S: FSK_PREP(FSK_PREC_F16X3); break;
    default: FSK_PREP(FSK_PREC_BF16X3); break;
  }
#undef FSK_PREP
  SplitPlan p = kde16_plan(n, m);
  int64_t per_xcd;
  prof_begin(FSK_OP_KDE16_LOGDENSITY, s);
  if (prec == FSK_PREC_F32_VALU) {
    const int64_t qblocks = cdiv(m, 256);
    const int64_t ntasks = qblocks * p.splits;
    const int64_t grid = xcd_grid(ntasks, &per_xcd);
    hipLaunchKernelGGL(kde16_lse_valu_kernel, dim3(grid), dim3(256), 0, s, L.xf, L.bn,
                       round_up(n, kTile), L.qf, L.qa, m, p.split_len, qblocks, ntasks, per_xcd,
                       L.part);
  } else {
    const int64_t ntasks = p.qblocks * p.splits;
    const int64_t grid = xcd_grid(ntasks, &per_xcd);
    const bool skip = L.lower && !(flags & FSK_FLAG_DENSE);
    if (skip) {
      // lower bounds: the sums over every stride-th point (a subset of the
      // data, so below each full sum), from the dense kernel on its records
      const int64_t ns_pad = round_up(L.ns, kValuChunk);
      SplitPlan ps = kde16_plan(L.ns, m);
      int64_t per_xcd_s;
      const int64_t ntasks_s = ps.qblocks * ps.splits;
      const int64_t grid_s = xcd_grid(ntasks_s, &per_xcd_s);
#define FSK_SAMPLE(P)                                                                            \
  do {                                                                                           \
    hipLaunchKernelGGL(prep_lse_kernel<P>, dim3(cdiv(ns_pad, 256)), dim3(256), 0, s, data,       \
                       ld * L.stride, L.ns, ns_pad, L.mu, inv_n, scl, coef_s, 0, (float *)nullptr, \
                       (float *)nullptr, L.srec);                                                \
    hipLaunchKernelGGL((kde16_lse_mfma_kernel<P, kKdeQT, false>), dim3(grid_s), dim3(kWG), 0, s, \
                       L.xf, L.bn, L.srec, round_up(L.ns, kTile), L.qf, L.qa, L.qrec, m,         \
                       ps.split_len, ps.qblocks, ntasks_s, per_xcd_s, (const float *)nullptr,    \
                       0.f, 0.f, (const unsigned long long *)nullptr, int64_t(0), L.spart);      \
  } while (0)
      switch (prec) {
        case FSK_PREC_BF16: FSK_SAMPLE(FSK_PREC_BF16); break;
        case FSK_PREC_F16: FSK_SAMPLE(FSK_PREC_F16); break;
        case FSK_PREC_BF16X6: FSK_SAMPLE(FSK_PREC_BF16X6); break;
        case FSK_PREC_F16X3:
        case FSK_PREC_F16X3S: FSK_SAMPLE(FSK_PREC_F16X3); break;
        default: FSK_SAMPLE(FSK_PREC_BF16X3); break;
      }
#undef FSK_SAMPLE
      hipLaunchKernelGGL(lse_lower_kernel, dim3(cdiv(m, 256)), dim3(256), 0, s, L.spart,
                         ps.splits, m, L.lower);
      if (L.far_bits)
        launch_far_bits(data, ld, n, query, ldq, m, L.mu, inv_n, -coef, L.npairs, p.qblocks,
                        kWG / 64, 0.f, L.lower, float(lse_margin(n)), L.frame, L.pair_boxes,
                        L.row_boxes, L.far_bits, s);
    }
    const float direct_min = float(-kLseDirect + std::log2(double(L.ns > 0 ? L.ns : 1)));
#define FSK_LSE(P)                                                                                \
  do {                                                                                            \
    if (skip)                                                                                     \
      hipLaunchKernelGGL((kde16_lse_mfma_kernel<P, kKdeQT, true>), dim3(grid), dim3(kWG), 0, s,   \
                         L.xf, L.bn, L.xrec, round_up(n, kTile), L.qf, L.qa, L.qrec, m,           \
                         p.split_len, p.qblocks, ntasks, per_xcd, L.lower, float(lse_margin(n)),  \
                         direct_min, L.far_bits, far_words(L.npairs), L.part);                    \
    else                                                                                          \
      hipLaunchKernelGGL((kde16_lse_mfma_kernel<P, kKdeQT, false>), dim3(grid), dim3(kWG), 0, s,  \
                         L.xf, L.bn, L.xrec, round_up(n, kTile), L.qf, L.qa, L.qrec, m,           \
                         p.split_len, p.qblocks, ntasks, per_xcd, (const float *)nullptr, 0.f,    \
                         0.f, (const unsigned long long *)nullptr, int64_t(0), L.part);           \
  } while (0)
#define FSK_LSE32(P)                                                                             \
  hipLaunchKernelGGL((kde16_lse_mfma_kernel<P, kKdeQT, false>), dim3(grid), dim3(kWG), 0, s,     \
                     L.xf, L.bn, L.xrec, round_up(n, kTile), L.qf, L.qa, L.qrec, m, p.split_len, \
                     p.qblocks, ntasks, per_xcd, (const float *)nullptr, 0.f, 0.f,               \
                     (const unsigned long long *)nullptr, int64_t(0), L.part)
    switch (prec) {
      case FSK_PREC_F32: FSK_LSE32(FSK_PREC_F32); break;
      case FSK_PREC_BF16: FSK_LSE(FSK_PREC_BF16); break;
      case FSK_PREC_F16: FSK_LSE(FSK_PREC_F16); break;
      case FSK_PREC_BF16X6: FSK_LSE(FSK_PREC_BF16X6); break;
      case FSK_PREC_F16X3:
      case FSK_PREC_F16X3S: FSK_LSE(FSK_PREC_F16X3); break;
      default: FSK_LSE(FSK_PREC_BF16X3); break;
    }
#undef FSK_LSE
#undef FSK_LSE32
  }
  prof_end(FSK_OP_KDE16_LOGDENSITY, s);
  if ((st = launch_status(what)) != FSK_OK) return st;
  const double log_const = -log(double(n)) - kD * log(bandwidth) - 0.5 * kD * log(2.0 * M_PI);
  hipLaunchKernelGGL(lse_combine_kernel, dim3(cdiv(m, 256)), dim3(256), 0, s, L.part, p.splits,
                     m, log_const, out);
  return launch_status(what);
}
