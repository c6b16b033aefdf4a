// fsk_far.h - far-block bitmaps for the pair kernels that leave out
// negligible terms (16-bit score forms with FP16 phi, 16-bit log-densities).
//
// A pre-pass bounds every (block of 64 rows, pair of 16-point data tiles)
// from below: the data's five leading principal axes (block power iteration
// on the covariance of a 4096-point sample) span a frame P with orthonormal
// rows, so |q - x| >= |P q - P x| >= D, the gap between the two blocks' boxes
// along those axes, and every exponent of the block (log2 units) is at most
//   -c D^2 + 2^-10 (max a + max b) + 1,
// where a = c |q - mu|^2, b = c |x - mu|^2 and 2^-10 (a + b) bounds a cross
// term's error from 16-bit coordinates (2^-11 relative each). A block whose
// bound is at or below its row block's threshold gets a bit; the kernels
// leave it out without reading its operands. The bound holds for any data:
// the frame only decides how many blocks it proves far.
#pragma once

#include "fsk_common.h"

namespace fsk {

constexpr int kFrameSample = 4096;  // points of the covariance sample
// Five axes: on the 8-component 16-D mixture at n = 2^20 they prove 79 % of
// the (64-row, 32-point) blocks far at the FP16 threshold and 65 % at
// 2^-48, against 74 % and 60 % with three (85 % and 76 % are far).
constexpr int kBoxAxes = 5;
constexpr int kBoxNorm = 2 * kBoxAxes;      // box field: max c|x - mu|^2
constexpr int kBoxThr = 2 * kBoxAxes + 1;   // box field: threshold
constexpr int kBoxFloats = 2 * kBoxAxes + 2;  // lo[axes], hi[axes], norm, threshold
constexpr int kFarRows = 64;        // rows per bitmap row (a wave's rows)
constexpr int kFarPairPts = 32;     // data points per bit
// bitmap sizes above this run without it (the kernels' own tests alone)
constexpr int64_t kFarBitsCap = int64_t(256) << 20;

// Bitmap layout: rows g * waves + k (k < waves) for the waves of workgroup
// g, then one row per workgroup g at groups * waves + g holding the AND of
// its waves' rows; each row has `words` 64-bit words, bit p % 64 of word
// p / 64 for data pair p.
inline int64_t far_words(int64_t npairs) { return (npairs + 63) / 64; }
inline int64_t far_bits_count(int64_t groups, int waves, int64_t npairs) {
  return (groups * waves + groups) * far_words(npairs);
}

// One bitmap word by a scalar load that completes inside the asm block: a
// uniform 8-byte read that stays off the vector-memory counter a kernel's
// data copies are waited on with (a vector load would make the compiler wait
// for every copy in flight before reading it).
// The address must be the same on every lane; it is read from the first.
__device__ __forceinline__ unsigned long long sload_u64(const unsigned long long *p) {
  const unsigned long long a = reinterpret_cast<unsigned long long>(p);
  const unsigned long long u =
      (unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(int(a >> 32)) << 32 |
      (unsigned)__builtin_amdgcn_readfirstlane(int(a & 0xffffffffu));
  unsigned long long v;
  __asm__ volatile("s_load_dwordx2 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=s"(v) : "s"(u) : "memory");
  return v;
}

// Builds the bitmap for `rows` (m of them, stride ldr; groups * waves row
// blocks of kFarRows rows) against the n data points (npairs pairs).
// Thresholds: skip_t for every row block, or, with row_lower (device, m
// floats, log2 units), the smallest row_lower - margin over the block's rows.
// mu_sum / inv_n: the data mean; c = log2(e) / (2 h^2). Scratch: frame
// (kBoxAxes * 16 floats), pair_boxes (npairs * kBoxFloats), row_boxes
// (groups * waves * kBoxFloats).
void launch_far_bits(const float *data, int64_t ld, int64_t n, const float *rows, int64_t ldr,
                     int64_t m, const double *mu_sum, double inv_n, double c, int64_t npairs,
                     int64_t groups, int waves, float skip_t, const float *row_lower,
                     float margin, float *frame, float *pair_boxes, float *row_boxes,
                     unsigned long long *bits, hipStream_t s);

}  // namespace fsk
