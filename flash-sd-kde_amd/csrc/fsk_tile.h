// fsk_tile.h - the per-tile building blocks of the 16-D pair kernels.
//
// A wave holds QT row tiles of 16 queries (B operand, pre-scaled) and
// streams data in tiles of 16 points (A operand). For one data tile:
//   cross_tile : D[R] = bb_j + 4 chained v_mfma_f32_16x16x4_f32
//                -> lane holds D(i = lane%16, j = 4(lane/16) + r), r = 0..3
//   kernel_tile: t = D * cs + qa_i (NAT) or D + qa_i, phi = exp2(min(t, 0)) in
//                place, row sums into acc[R]
//   phi_x_tile : W[R] += phi . X_J  (4 MFMAs; phi is already the A operand)
// The accumulator starts from the data-side norm term only and the query-side
// term is added once after the cross term: the chained FP32 roundings of the
// MFMA then happen at |bb_j + partial cross term|, which stays well below the
// full |qa_i + bb_j| that a combined initialiser would carry through all four
// steps (on 16-D N(0, I) data this halves the FP32 error of the densities).
// Two unit systems share the code:
//   log2 units: bb = c|x|^2, query operand -2c q, t = D + qa       (score paths)
//   natural   : bb = |x|^2,  query operand -2 q (exact), t = fma(D, c, qa)
//               (the KDE: no rounded query pre-scale, one rounding for c)
// TileSrc addresses one data tile: xa = features 4g..4g+3 of point j0 + c,
// bb = norm terms of points j0 + 4g .. +3, xt[r] = feature c of point j0 + 4g + r.
#pragma once

#include "fsk_common.h"

namespace fsk {

struct TileData {
  f4 xa, bb;
  float xt[4];
};

// Loads for tile j0. CLAMP = true for the one tile that crosses n (rows >= n
// are replaced by row n-1; their norm term is already the pad value, so phi = 0).
template <bool CLAMP, bool WITH_XT>
__device__ __forceinline__ void load_tile(const float *__restrict__ data, int64_t ld, int64_t n,
                                          const float *__restrict__ bn, int64_t j0, int g, int c,
                                          TileData &t) {
  int64_t ja = j0 + c;
  if (CLAMP && ja > n - 1) ja = n - 1;
  t.xa = ld4(data + ja * ld + 4 * g);
  t.bb = ld4(bn + j0 + 4 * g);
  if (WITH_XT) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int64_t jt = j0 + 4 * g + r;
      if (CLAMP && jt > n - 1) jt = n - 1;
      t.xt[r] = data[jt * ld + c];
    }
  }
}

template <int QT>
__device__ __forceinline__ void cross_tile(const f4 (&qv)[QT], const TileData &t, f4 (&d)[QT]) {
#pragma unroll
  for (int R = 0; R < QT; ++R) d[R] = mfma_f32(t.xa[0], qv[R][0], t.bb);
#pragma unroll
  for (int s = 1; s < 4; ++s) {
#pragma unroll
    for (int R = 0; R < QT; ++R) d[R] = mfma_f32(t.xa[s], qv[R][s], d[R]);
  }
}

// t = D * cs + qa (NAT) or D + qa; phi = exp2(min(t, 0)) written back into d;
// acc[R] += row sums. LAPL also accumulates accT[R] += phi * min(t, 0).
template <int QT, bool LAPL, bool NAT>
__device__ __forceinline__ void kernel_tile(f4 (&d)[QT], const float (&qav)[QT], float cs,
                                            float (&acc)[QT], float (&accT)[QT]) {
#pragma unroll
  for (int R = 0; R < QT; ++R) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v = NAT ? fmaf(d[R][r], cs, qav[R]) : d[R][r] + qav[R];
      float e;
      if (LAPL) {
        const float t = fminf(v, 0.f);
        e = fexp2(t);
        accT[R] = fmaf(e, t, accT[R]);
      } else {
        e = fexp2_le1(v);
      }
      acc[R] += e;
      d[R][r] = e;
    }
  }
}

template <int QT>
__device__ __forceinline__ void phi_x_tile(const f4 (&d)[QT], const TileData &t, f4 (&w)[QT]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int R = 0; R < QT; ++R) w[R] = mfma_f32(d[R][r], t.xt[r], w[R]);
  }
}

// Walk the data range [j_begin, j_end) of one split tile by tile: full tiles
// (all 16 points < n) without index clamps, then at most one clamped tail
// tile. body(TileData&) does the compute for one tile. (Issuing the next
// tile's loads before the current tile's compute measured no faster: with
// 2-3 waves per SIMD the MFMA pipe already covers the load latency.)
template <bool WITH_XT, class Body>
__device__ __forceinline__ void for_each_tile(const float *__restrict__ data, int64_t ld, int64_t n,
                                              const float *__restrict__ bn, int64_t j_begin,
                                              int64_t j_end, int g, int c, Body &&body) {
  const int64_t n_full = n & ~int64_t(kTile - 1);
  const int64_t j_mid = j_end < n_full ? j_end : n_full;
  int64_t j0 = j_begin;
  for (; j0 < j_mid; j0 += kTile) {
    TileData t;
    load_tile<false, WITH_XT>(data, ld, n, bn, j0, g, c, t);
    body(t);
  }
  if (j0 < j_end) {
    TileData t;
    load_tile<true, WITH_XT>(data, ld, n, bn, j0, g, c, t);
    body(t);
  }
}

}  // namespace fsk
