// fsk_score16.hip - 16-D empirical-score accumulators on MFMA.
//
//   pdf[i] = sum_j phi_ij,   w[i,:] = sum_j phi_ij x_j,   phi_ij = exp(-|x_i-x_j|^2/(2h^2))
// over all ordered pairs, j = i included (phi_ii = 1), as in
// flash_sd_kde/reference.py:263-284 (empirical_score_nd_numpy).
//
// phi tiles come out of the cross-term MFMA with lane l holding
// phi[i = l%16][j = 4(l/16) + r], r = 0..3. That is exactly the A-operand
// layout of v_mfma_f32_16x16x4_f32 for  W_I += phi . X_J  (k = l/16 <-> j),
// so the phi-weighted data sum needs no data movement: MFMA r consumes the
// r-th phi value of every lane against X_J rows 4(l/16)+r.
//
// Ordered variant: deterministic split-K, FP32 partials per split, FP64
// combine (plus FP64 outer sums inside each split).
// Symmetric variant: each unordered 512x512 block pair once; the column side
// (W_J += phi^T X_I) needs phi transposed, done through per-wave LDS slots;
// column sums are gathered in per-wave LDS regions and scattered with FP64
// global atomics, then converted to the FP32 outputs.
#include <type_traits>

#include "fsk_common.h"
#include "fsk_lowp.h"
#include "fsk_plan.h"
#include "fsk_tile.h"

namespace fsk {

constexpr int kScoreFlush = 32;
static_assert(kNormPad % kScoreRowsPerWG == 0, "row norms must cover whole row blocks");

// Largest spread of the query terms A_i = -c|q_i|^2 (c < 0) inside one
// workgroup for which the factored form below is used, and the largest
// exponent the form lets exp2 see.
constexpr float kFactSpread = 160.f;
constexpr float kFactHalf = 80.f;

// Shift K of the factored form for a workgroup whose A_i lie in [amin, amax]
// (amax - amin <= kFactSpread): the largest d of row i is A_i - K and the
// per-row factor is exp2(K - A_i), both within +-kFactHalf. K is kept as small
// as that allows (0 while amax <= kFactHalf): the accumulator then starts from
// c|x_j|^2 - K with the least magnitude, and the chained FP32 roundings of the
// cross term happen at the smallest partial sums.
__device__ __forceinline__ float fact_shift(float amax) { return fmaxf(0.f, amax - kFactHalf); }

// Rows (queries) and data may differ: rows = data gives the empirical score of
// the data set itself; a row subset gives the score of those rows under the
// whole data set's KDE. qa = c|row|^2 padded to a multiple of 256 rows.
//
// FACT (rows = data only): phi_ij = exp2(t_ij) with t_ij = qa_i + c|x_j|^2 -
// 2c q_i.x_j is evaluated as exp2(d_ij) * exp2(qa_i + K), d_ij = (c|x_j|^2 - K)
// - 2c q_i.x_j: the MFMA accumulator starts from the data-side term (shared by
// every row tile of a data tile), so no per-pair add is needed, and the
// per-row factor multiplies the finished FP64 sums once. K = fact_shift(max
// A_i); with the self pair t_ii = 0 in every row, the largest d of row i is
// A_i - K, within +-kFactHalf, so exp2 neither overflows nor leaves the
// normal range for any pair that matters. A workgroup whose rows spread wider
// takes the direct form (accumulator from c|x_j|^2, qa_i added after the
// cross term).
template <int QT, bool FACT>
__global__ __launch_bounds__(kWG) void score16_ordered_kernel(
    const float *__restrict__ data, int64_t ld, int64_t n, const float *__restrict__ bn,
    int64_t n_pad, const float *__restrict__ rows, int64_t ldr, int64_t m,
    const float *__restrict__ qa, float qscale, int64_t split_len, int64_t qblocks,
    int64_t ntasks, int64_t per_xcd, float *__restrict__ pdf_part, float *__restrict__ w_part) {
  const int64_t task = xcd_task(blockIdx.x, per_xcd);
  if (task >= ntasks) return;
  const int64_t split = task / qblocks;
  const int64_t qb = task - split * qblocks;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t row0 = qb * (64 * QT) + int64_t(wave) * (16 * QT);

  f4 qv[QT], w[QT];
  d4 w64[QT];
  float qav[QT], pdf[QT];
  double pdf64[QT];
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    int64_t i = row0 + 16 * R + c;
    int64_t ic = i > m - 1 ? m - 1 : i;
    qv[R] = ld4(rows + ic * ldr + 4 * g) * qscale;
    qav[R] = qa[i];  // padded to a multiple of 256 rows
    w[R] = f4{0.f, 0.f, 0.f, 0.f};
    w64[R] = d4{0.0, 0.0, 0.0, 0.0};
    pdf[R] = 0.f;
    pdf64[R] = 0.0;
  }

  // workgroup range of A_i = -qa_i over its real rows (FACT only)
  float kq = 0.f;
  bool fact = false;
  if (FACT) {
    __shared__ float red[8];
    float amin = INFINITY, amax = -INFINITY;
#pragma unroll
    for (int R = 0; R < QT; ++R) {
      const float a = R == 0 || row0 + 16 * R + c < m ? -qav[R] : -qav[0];
      amin = fminf(amin, a);
      amax = fmaxf(amax, a);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      amin = fminf(amin, __shfl_xor(amin, o));
      amax = fmaxf(amax, __shfl_xor(amax, o));
    }
    if (lane == 0) {
      red[wave] = amin;
      red[4 + wave] = amax;
    }
    __syncthreads();
    amin = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
    amax = fmaxf(fmaxf(red[4], red[5]), fmaxf(red[6], red[7]));
    fact = amax - amin <= kFactSpread;  // finite and narrow (padded-only rows give -inf)
    kq = fact_shift(amax);
  }

  const int64_t j_begin = split * split_len;
  int64_t j_end = j_begin + split_len;
  if (j_end > n_pad) j_end = n_pad;
  auto run = [&](auto factored) {
    constexpr bool F = decltype(factored)::value;
    float unused[QT];
    int tc = 0;
    for_each_tile<true>(data, ld, n, bn, j_begin, j_end, g, c, [&](const TileData &t) {
      f4 d[QT];
      if constexpr (F) {
        const f4 ci = t.bb - f4{kq, kq, kq, kq};
#pragma unroll
        for (int R = 0; R < QT; ++R) d[R] = mfma_f32(t.xa[0], qv[R][0], ci);
#pragma unroll
        for (int s = 1; s < 4; ++s) {
#pragma unroll
          for (int R = 0; R < QT; ++R) d[R] = mfma_f32(t.xa[s], qv[R][s], d[R]);
        }
#pragma unroll
        for (int R = 0; R < QT; ++R) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float e = fexp2(d[R][r]);
            pdf[R] += e;
            d[R][r] = e;
          }
        }
      } else {
        cross_tile<QT>(qv, t, d);
        kernel_tile<QT, false, false>(d, qav, 1.f, pdf, unused);
      }
      phi_x_tile<QT>(d, t, w);
      if (++tc == kScoreFlush) {
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
    });
  };
  if (FACT && fact)
    run(std::true_type{});
  else
    run(std::false_type{});
  // per-row factor exp2(qa_i + K) of the factored form
  auto row_scale = [&](int64_t i) -> double {
    return FACT && fact ? exp2(double(qa[i]) + double(kq)) : 1.0;
  };
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    double p = sum_groups(pdf64[R] + pdf[R]);
    int64_t i = row0 + 16 * R + c;
    if (g == 0 && i < m) pdf_part[split * m + i] = float(p * row_scale(i));
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int64_t iw = row0 + 16 * R + 4 * g + r;
      if (iw < m) w_part[(split * m + iw) * kD + c] = float((w64[R][r] + w[R][r]) * row_scale(iw));
    }
  }
}

// Symmetric variant: every unordered pair of blocks (I, J), J >= I, of
// B = 4 waves x 16 QT points, once. The workgroup holds block I in registers
// (a wave owns QT row tiles) and streams the data tiles of up to jc blocks J.
//  * Row side (pdf_I, W_I): as the ordered kernel; phi is the A operand of
//    W_I += phi . X_J in the layout the cross-term MFMA leaves it in.
//  * Column side (W_J += phi^T . X_I): each phi tile goes to LDS (four
//    ds_write_b32 per lane) and comes back transposed (one ds_read_b128);
//    rows padded to kSymTS floats keep both accesses free of bank conflicts.
//    The column pdf sums are FMAs in the original layout, reduced over the 16
//    lanes of a DPP row once per data tile (row16_sum4).
//  * Each wave stores its column sums of a data tile in its own LDS region;
//    every kSymFlushTiles tiles the workgroup adds the four regions and sends
//    the totals to the FP64 accumulators with global atomics. Row sums go out
//    once per task.
// The diagonal block J = I runs the row side only, over every ordered pair in
// it. Points are centred on the data mean (zero rows pad the last block). The
// factored form of the ordered kernel applies per workgroup (same spread
// guard): the per-row factor f_i = exp2(qa_i + K) enters the column side
// through its B operand (f_i x_i) and the column pdf FMA, and the row sums
// once at the end.
// Tasks: window k of block I covers J in [I + k jc, I + (k+1) jc), decoded by
// sym_task (fsk_common.h).
template <int QT>
__global__ __launch_bounds__(kWG) void score16_sym_kernel(
    const float *__restrict__ xc, const float *__restrict__ bn, int64_t n, float qscale,
    int64_t nb, int64_t jc, int64_t nwin, int64_t task0, int64_t ntasks, int64_t per_xcd,
    double *__restrict__ pdf64, double *__restrict__ w64) {
  constexpr int B = 4 * 16 * QT;
  constexpr int TS = kSymTS;
  constexpr int NS = kSymSlots;
  constexpr int FT = kSymFlushTiles;
  static_assert(QT % NS == 0, "row tiles are processed in groups of the slot count");
  static_assert((B / kTile) % FT == 0, "a block holds whole flush groups");
  __shared__ float tslot[4][NS][16 * TS];
  __shared__ float colw[4][FT * kTile * TS];  // per wave: column sums W_J of FT tiles
  __shared__ float colp[4][FT * kTile];       // per wave: column pdf sums of FT tiles
  __shared__ float red[8];

  // this launch covers tasks [task0, task0 + gridDim.x) of [0, ntasks)
  const int64_t task = task0 + xcd_task(blockIdx.x, per_xcd);
  if (task >= ntasks) return;
  int64_t I, Jlo, Jhi;
  sym_task(task, nb, jc, nwin, &I, &Jlo, &Jhi);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int g = lane >> 4, c = lane & 15;

  // xc and bn cover nb * B rows (zero points, kPadNeg norms past n)
  const int64_t row0 = I * B + int64_t(wave) * (16 * QT);
  f4 qv[QT], wI[QT], xit[QT];
  float qav[QT], fI[QT], pI[QT];
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    const int64_t i = row0 + 16 * R + c;
    qv[R] = ld4(xc + i * kD + 4 * g) * qscale;
    qav[R] = bn[i];
    wI[R] = f4{0.f, 0.f, 0.f, 0.f};
    pI[R] = 0.f;
  }

  // workgroup range of A_i = -qa_i over its real rows (row I*B is always real)
  float amin = INFINITY, amax = -INFINITY;
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    if (row0 + 16 * R + c < n) {
      amin = fminf(amin, -qav[R]);
      amax = fmaxf(amax, -qav[R]);
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    amin = fminf(amin, __shfl_xor(amin, o));
    amax = fmaxf(amax, __shfl_xor(amax, o));
  }
  if (lane == 0) {
    red[wave] = amin;
    red[4 + wave] = amax;
  }
  __syncthreads();  // also orders the LDS zero fill before the first use
  amin = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
  amax = fmaxf(fmaxf(red[4], red[5]), fmaxf(red[6], red[7]));
  const bool fact = amax - amin <= kFactSpread;
  const float kq = fact_shift(amax);

  // per-row factors (padded rows: f = 0 in the factored form; their points are
  // zero, so x f = 0 in both forms) and the column side's B operand f_i x_i
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    fI[R] = fact ? fexp2(qav[R] + kq) : 1.f;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int64_t it = row0 + 16 * R + 4 * g + s;
      const float f = fact ? fexp2(bn[it] + kq) : 1.f;
      xit[R][s] = xc[it * kD + c] * f;
    }
  }

  auto run = [&](auto factored) {
    constexpr bool F = decltype(factored)::value;
    // cross term of row tile R into d: the factored form starts from the
    // data-side term ci = c|x_j|^2 - K, the direct form from c|x_j|^2 (qa_i is
    // added after the cross term, in kern)
    auto cross = [&](const TileData &t, const f4 &ci, int R, f4 &d) {
      d = mfma_f32(t.xa[0], qv[R][0], F ? ci : t.bb);
    };
    auto kern = [&](float v, int R) { return F ? fexp2(v) : fexp2_le1(v + qav[R]); };
    for (int64_t J = Jlo; J < Jhi; ++J) {
      const int64_t jbase = J * B;
      if (J == I) {
        // diagonal block: row side over every ordered pair inside it
        for (int jt = 0; jt < B / kTile; ++jt) {
          TileData t;
          load_tile<false, true>(xc, kD, nb * B, bn, jbase + jt * kTile, g, c, t);
          const f4 ci = t.bb - f4{kq, kq, kq, kq};
          f4 d[QT];
#pragma unroll
          for (int R = 0; R < QT; ++R) cross(t, ci, R, d[R]);
#pragma unroll
          for (int s = 1; s < 4; ++s) {
#pragma unroll
            for (int R = 0; R < QT; ++R) d[R] = mfma_f32(t.xa[s], qv[R][s], d[R]);
          }
#pragma unroll
          for (int R = 0; R < QT; ++R) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const float e = kern(d[R][r], R);
              pI[R] += e;
              d[R][r] = e;
            }
          }
          phi_x_tile<QT>(d, t, wI);
        }
        continue;
      }
      for (int jt = 0; jt < B / kTile; ++jt) {
        TileData t;
        load_tile<false, true>(xc, kD, nb * B, bn, jbase + jt * kTile, g, c, t);
        const f4 ci = t.bb - f4{kq, kq, kq, kq};
        f4 wJ = f4{0.f, 0.f, 0.f, 0.f};
        float pj[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int hb = 0; hb < QT / NS; ++hb) {
          f4 d[NS];
#pragma unroll
          for (int q = 0; q < NS; ++q) cross(t, ci, hb * NS + q, d[q]);
#pragma unroll
          for (int s = 1; s < 4; ++s) {
#pragma unroll
            for (int q = 0; q < NS; ++q) d[q] = mfma_f32(t.xa[s], qv[hb * NS + q][s], d[q]);
          }
          // lane holds phi(i = c, j = 4g + r): row sums, column pdf FMAs, and
          // the transposed copy T[j][i] in this wave's LDS slot q
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            const int R = hb * NS + q;
            float *sb = tslot[wave][q];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const float e = kern(d[q][r], R);
              pI[R] += e;
              pj[r] = fmaf(e, fI[R], pj[r]);
              d[q][r] = e;
              sb[(4 * g + r) * TS + c] = e;
            }
          }
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int q = 0; q < NS; ++q) wI[hb * NS + q] = mfma_f32(d[q][r], t.xt[r], wI[hb * NS + q]);
          }
          // column side: A = phi(i = 4g + s, j = c), B = f_i x_i of the same rows
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            const f4 et = ld4(&tslot[wave][q][c * TS + 4 * g]);
#pragma unroll
            for (int s = 0; s < 4; ++s) wJ = mfma_f32(et[s], xit[hb * NS + q][s], wJ);
          }
          __builtin_amdgcn_wave_barrier();
        }
        // pj[r]: this lane's part of the column pdf of point 4g + r; lanes
        // c < 4 get the total for point 4g + c. The wave's column sums of this
        // tile go to its own LDS region (plain stores: LDS float atomics
        // measured 40 % slower for the whole kernel)
        const float pv = row16_sum4(pj, c);
        const int lt = jt % FT;
        if (c < 4) colp[wave][lt * kTile + 4 * g + c] = pv;
#pragma unroll
        for (int r = 0; r < 4; ++r) colw[wave][(lt * kTile + 4 * g + r) * TS + c] = wJ[r];
        if (lt == FT - 1) {
          // the four waves' sums of the last FT tiles -> FP64 accumulators
          // (two alternating region sets with one barrier per flush measured
          // no faster)
          __syncthreads();
          const int64_t p0 = jbase + (jt - (FT - 1)) * kTile;
          for (int e = tid; e < FT * kTile * kD; e += kWG) {
            const int pl = e >> 4, f = e & 15;
            const float v = (colw[0][pl * TS + f] + colw[1][pl * TS + f]) +
                            (colw[2][pl * TS + f] + colw[3][pl * TS + f]);
            if (p0 + pl < n) atomicAdd(&w64[(p0 + pl) * kD + f], double(v));
          }
          for (int e = tid; e < FT * kTile; e += kWG) {
            const float v = (colp[0][e] + colp[1][e]) + (colp[2][e] + colp[3][e]);
            if (p0 + e < n) atomicAdd(&pdf64[p0 + e], double(v));
          }
          __syncthreads();
        }
      }
    }
  };
  if (fact)
    run(std::true_type{});
  else
    run(std::false_type{});

  // row sums, times the per-row factor of the factored form (FP64)
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    const double p = sum_groups(double(pI[R]));
    const int64_t i = row0 + 16 * R + c;
    if (g == 0 && i < n) atomicAdd(&pdf64[i], fact ? p * exp2(double(qav[R]) + double(kq)) : p);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t iw = row0 + 16 * R + 4 * g + r;
      if (iw < n) {
        const double sc = fact ? exp2(double(bn[iw]) + double(kq)) : 1.0;
        atomicAdd(&w64[iw * kD + c], double(wI[R][r]) * sc);
      }
    }
  }
}

// pdf = pdf64, w = w64 + mu pdf64 (FP64) for sums over points centred on mu.
__global__ void score_finish_kernel(const double *__restrict__ pdf64,
                                    const double *__restrict__ w64, int64_t n,
                                    const double *__restrict__ mu_sum, double inv_n,
                                    float *__restrict__ pdf, float *__restrict__ w,
                                    int64_t ldw_row, int64_t ldw_col) {
  int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * kD) return;
  int64_t i = e / kD, f = e % kD;
  w[i * ldw_row + f * ldw_col] = float(w64[e] + mu_sum[f] * inv_n * pdf64[i]);
  if (f == 0) pdf[i] = float(pdf64[i]);
}

static_assert(kNormPad % kSymRows == 0, "padded rows must cover whole symmetric blocks");

struct ScoreLayout {
  float *bn;
  float *qa;
  float *pdf_part;
  float *w_part;
  double *pdf64;
  double *w64;
  double *mu;       // ordered / rows: data mean workspace
  float *xc, *rc;   // ordered / rows: centred data and rows (16 floats per point)
};

static size_t score_layout(void *ws, int64_t n, bool ordered, ScoreLayout *L, int64_t m = -1) {
  Carve cv(ws);
  L->bn = cv.take<float>(round_up(n, kNormPad));
  L->qa = nullptr;
  L->pdf_part = L->w_part = nullptr;
  L->pdf64 = L->w64 = nullptr;
  L->rc = nullptr;
  // data mean and centred data (the symmetric kernel reads whole blocks of
  // kSymRows points, so its copy is padded with zero rows to kNormPad)
  L->mu = cv.take<double>(kMeanWs);
  L->xc = cv.take<float>(round_up(n, (m < 0 && !ordered) ? kNormPad : kTile) * kD);
  if (m >= 0) {  // separate rows
    L->qa = cv.take<float>(round_up(m, kNormPad));
    L->rc = cv.take<float>(m * kD);
    SplitPlan p = score16_rows_plan(n, m);
    L->pdf_part = cv.take<float>(p.splits * m);
    L->w_part = cv.take<float>(p.splits * m * kD);
  } else if (ordered) {
    SplitPlan p = score16_ordered_plan(n);
    L->pdf_part = cv.take<float>(p.splits * n);
    L->w_part = cv.take<float>(p.splits * n * kD);
  } else {
    L->pdf64 = cv.take<double>(n);
    L->w64 = cv.take<double>(n * kD);
  }
  return cv.used;
}

size_t score16_ws_bytes(int64_t n, bool ordered) {
  ScoreLayout L;
  return score_layout(nullptr, n, ordered, &L);
}

size_t score16_rows_ws_bytes(int64_t n, int64_t m) {
  ScoreLayout L;
  return score_layout(nullptr, n, true, &L, m);
}

static int score_check(const char *what, const float *data, int64_t ld, int64_t n, double h,
                       unsigned flags, const float *pdf, const float *w, bool lowp_ok,
                       unsigned extra_flags = 0u) {
  int st;
  if ((st = check_points16(what, data, ld, n)) != FSK_OK) return st;
  if ((st = check_elem_rows(what, n)) != FSK_OK) return st;
  if (!(h > 0.0) || !isfinite(h)) {
    set_error("%s: bandwidth must be positive, got %g", what, h);
    return FSK_EINVAL;
  }
  if (!pdf || !w) {
    set_error("%s: null output", what);
    return FSK_EINVAL;
  }
  const unsigned prec = flags & FSK_PREC_MASK;
  const bool lowp = prec == FSK_PREC_BF16 || prec == FSK_PREC_F16 || prec == FSK_PREC_BF16X3 ||
                    prec == FSK_PREC_BF16X6 || prec == FSK_PREC_F16X3 || prec == FSK_PREC_F16X3S ||
                    prec == FSK_PREC_BF16X3S || prec == FSK_PREC_F32_VALU;
  if ((prec != FSK_PREC_F32 && !(lowp && lowp_ok)) ||
      (flags & ~(FSK_PREC_MASK | extra_flags | (lowp_ok ? FSK_FLAG_DENSE : 0u)))) {
    set_error("%s: flags 0x%x not available on this path", what, flags);
    return FSK_EUNSUPPORTED;
  }
  return FSK_OK;
}

}  // namespace fsk

using namespace fsk;

extern "C" int fsk_score16_ordered(const float *data, int64_t ld, int64_t n, double bandwidth,
                                   unsigned flags, void *ws, size_t ws_bytes, float *pdf,
                                   float *w, int64_t ldw_row, int64_t ldw_col, void *stream) {
  const char *what = "fsk_score16_ordered";
  int st = score_check(what, data, ld, n, bandwidth, flags, pdf, w, true);
  if (st != FSK_OK) return st;
  if ((flags & FSK_PREC_MASK) == FSK_PREC_F32_VALU)
    return score16_valu_run(what, data, ld, n, data, ld, n, bandwidth, true, ws, ws_bytes, pdf, w,
                            ldw_row, ldw_col, stream);
  if ((flags & FSK_PREC_MASK) != FSK_PREC_F32)
    return score16_lp_run(what, data, ld, n, data, ld, n, bandwidth, flags & FSK_PREC_MASK, true,
                          (flags & FSK_FLAG_DENSE) != 0, ws, ws_bytes, pdf, w, ldw_row, ldw_col,
                          stream);
  ScoreLayout L;
  size_t need = score_layout(ws, n, true, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const double coef = exp2_coef(bandwidth);
  const int64_t n_norm = round_up(n, kNormPad), n16 = round_up(n, kTile);
  const double inv_n = 1.0 / double(n);
  // centre on the data mean (smaller norm terms, less cancellation in t)
  launch_mean16(data, ld, n, L.mu, s);
  hipLaunchKernelGGL(center16_kernel, dim3(cdiv(n16 * kD, 256)), dim3(256), 0, s, data, ld, n,
                     n16, L.mu, inv_n, L.xc);
  hipLaunchKernelGGL(prep_norms16, dim3(cdiv(n_norm, 256)), dim3(256), 0, s, L.xc, int64_t(kD), n,
                     n_norm, coef, kPadNeg, L.bn);
  SplitPlan p = score16_ordered_plan(n);
  const int64_t ntasks = p.qblocks * p.splits;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
  prof_begin(FSK_OP_SCORE16_ORDERED, s);
  hipLaunchKernelGGL((score16_ordered_kernel<kScoreQT, true>), dim3(grid), dim3(kWG), 0, s, L.xc,
                     int64_t(kD), n, L.bn, n16, L.xc, int64_t(kD), n, L.bn, float(-2.0 * coef),
                     p.split_len, p.qblocks, ntasks, per_xcd, L.pdf_part, L.w_part);
  prof_end(FSK_OP_SCORE16_ORDERED, s);
  if ((st = launch_status(what)) != FSK_OK) return st;
  hipLaunchKernelGGL(score_lp_finish_kernel, dim3(cdiv(n * kD, 256)), dim3(256), 0, s,
                     L.pdf_part, L.w_part, p.splits, n, L.mu, inv_n, 1.0, pdf, w, ldw_row, ldw_col);
  return launch_status(what);
}

extern "C" int fsk_score16_sym_atomic(const float *data, int64_t ld, int64_t n,
                                      double bandwidth, unsigned flags, void *ws,
                                      size_t ws_bytes, float *pdf, float *w, int64_t ldw_row,
                                      int64_t ldw_col, void *stream) {
  const char *what = "fsk_score16_sym_atomic";
  int st = score_check(what, data, ld, n, bandwidth, flags, pdf, w, false);
  if (st != FSK_OK) return st;
  ScoreLayout L;
  size_t need = score_layout(ws, n, false, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const double coef = exp2_coef(bandwidth);
  const int64_t n_norm = round_up(n, kNormPad);
  const double inv_n = 1.0 / double(n);
  // centre on the data mean; rows [n, n_norm) of the copy are zero
  launch_mean16(data, ld, n, L.mu, s);
  hipLaunchKernelGGL(center16_kernel, dim3(cdiv(n_norm * kD, 256)), dim3(256), 0, s, data, ld, n,
                     n_norm, L.mu, inv_n, L.xc);
  hipLaunchKernelGGL(prep_norms16, dim3(cdiv(n_norm, 256)), dim3(256), 0, s, L.xc, int64_t(kD), n,
                     n_norm, coef, kPadNeg, L.bn);
  hipLaunchKernelGGL(fill_f64, dim3(cdiv(n, 256) < 4096 ? cdiv(n, 256) : 4096), dim3(256), 0, s,
                     L.pdf64, n, 0.0);
  hipLaunchKernelGGL(fill_f64, dim3(cdiv(n * kD, 256) < 4096 ? cdiv(n * kD, 256) : 4096),
                     dim3(256), 0, s, L.w64, n * kD, 0.0);
  const int64_t nb = cdiv(n, kSymRows);
  const int64_t nwin = cdiv(nb, kSymJC);
  const int64_t ntasks = sym_window_offset(nwin, nb, kSymJC);
  prof_begin(FSK_OP_SCORE16_SYM, s);
  for (int64_t t0 = 0; t0 < ntasks; t0 += kSymTasksPerLaunch) {
    const int64_t nt = ntasks - t0 < kSymTasksPerLaunch ? ntasks - t0 : kSymTasksPerLaunch;
    int64_t per_xcd;
    const int64_t grid = xcd_grid(nt, &per_xcd);
    hipLaunchKernelGGL((score16_sym_kernel<kSymQT>), dim3(grid), dim3(kWG), 0, s, L.xc, L.bn, n,
                       float(-2.0 * coef), nb, kSymJC, nwin, t0, t0 + nt, per_xcd, L.pdf64,
                       L.w64);
    if ((st = launch_status(what)) != FSK_OK) break;
  }
  prof_end(FSK_OP_SCORE16_SYM, s);
  if (st != FSK_OK) return st;
  hipLaunchKernelGGL(score_finish_kernel, dim3(cdiv(n * kD, 256)), dim3(256), 0, s, L.pdf64,
                     L.w64, n, L.mu, inv_n, pdf, w, ldw_row, ldw_col);
  return launch_status(what);
}

extern "C" int fsk_score16_rows(const float *rows, int64_t ldr, int64_t m, const float *data,
                                int64_t ld, int64_t n, double bandwidth, unsigned flags, void *ws,
                                size_t ws_bytes, float *pdf, float *w, int64_t ldw_row,
                                int64_t ldw_col, void *stream) {
  const char *what = "fsk_score16_rows";
  int st = score_check(what, data, ld, n, bandwidth, flags, pdf, w, true, FSK_FLAG_ROWS_IN_DATA);
  if (st != FSK_OK) return st;
  if ((st = check_points16(what, rows, ldr, m)) != FSK_OK) return st;
  if ((st = check_elem_rows(what, m)) != FSK_OK) return st;
  if ((flags & FSK_PREC_MASK) == FSK_PREC_F32_VALU)
    return score16_valu_run(what, rows, ldr, m, data, ld, n, bandwidth, false, ws, ws_bytes, pdf,
                            w, ldw_row, ldw_col, stream);
  if ((flags & FSK_PREC_MASK) != FSK_PREC_F32)
    return score16_lp_run(what, rows, ldr, m, data, ld, n, bandwidth, flags & FSK_PREC_MASK,
                          (rows == data && ldr == ld && m == n) ||
                              (flags & FSK_FLAG_ROWS_IN_DATA) != 0,
                          (flags & FSK_FLAG_DENSE) != 0, ws,
                          ws_bytes, pdf, w, ldw_row, ldw_col, stream);
  ScoreLayout L;
  size_t need = score_layout(ws, n, true, &L, m);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const double coef = exp2_coef(bandwidth);
  const int64_t n_norm = round_up(n, kNormPad), m_norm = round_up(m, kNormPad);
  const int64_t n16 = round_up(n, kTile);
  const double inv_n = 1.0 / double(n);
  // rows and data centred on the data mean
  launch_mean16(data, ld, n, L.mu, s);
  hipLaunchKernelGGL(center16_kernel, dim3(cdiv(n16 * kD, 256)), dim3(256), 0, s, data, ld, n,
                     n16, L.mu, inv_n, L.xc);
  hipLaunchKernelGGL(center16_kernel, dim3(cdiv(m * kD, 256)), dim3(256), 0, s, rows, ldr, m, m,
                     L.mu, inv_n, L.rc);
  hipLaunchKernelGGL(prep_norms16, dim3(cdiv(n_norm, 256)), dim3(256), 0, s, L.xc, int64_t(kD), n,
                     n_norm, coef, kPadNeg, L.bn);
  hipLaunchKernelGGL(prep_norms16, dim3(cdiv(m_norm, 256)), dim3(256), 0, s, L.rc, int64_t(kD), m,
                     m_norm, coef, kPadNeg, L.qa);
  SplitPlan p = score16_rows_plan(n, m);
  const int64_t ntasks = p.qblocks * p.splits;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
  prof_begin(FSK_OP_SCORE16_ORDERED, s);
  hipLaunchKernelGGL((score16_ordered_kernel<kScoreQT, false>), dim3(grid), dim3(kWG), 0, s, L.xc,
                     int64_t(kD), n, L.bn, n16, L.rc, int64_t(kD), m, L.qa, float(-2.0 * coef),
                     p.split_len, p.qblocks, ntasks, per_xcd, L.pdf_part, L.w_part);
  prof_end(FSK_OP_SCORE16_ORDERED, s);
  if ((st = launch_status(what)) != FSK_OK) return st;
  hipLaunchKernelGGL(score_lp_finish_kernel, dim3(cdiv(m * kD, 256)), dim3(256), 0, s,
                     L.pdf_part, L.w_part, p.splits, m, L.mu, inv_n, 1.0, pdf, w, ldw_row, ldw_col);
  return launch_status(what);
}
