// fsk_plan.h - launch plans (tile sizes, split counts, workspace carving).
// Everything here depends only on (n, m, flags), so a plan - and with it the
// FP32 partial boundaries and the FP64 combine order - is identical for every
// call with the same sizes: the split-K paths are bitwise reproducible.
#pragma once

#include <stdint.h>
#include <stddef.h>

#include "flashsdkde.h"

namespace fsk {

// 16-D KDE split-K: a workgroup owns 4 waves x QT16 row tiles of 16 queries.
constexpr int kKdeQT = 4;
constexpr int kKdeRowsPerWG = 4 * 16 * kKdeQT;  // 256 queries
// 16-D score (ordered, FP32): rows per workgroup. Eight row tiles per wave
// reuse each loaded data tile for 128 rows (229 VGPRs, 2 waves per SIMD).
#ifndef FSK_SCORE_QT
#define FSK_SCORE_QT 8
#endif
constexpr int kScoreQT = FSK_SCORE_QT;
constexpr int kScoreRowsPerWG = 4 * 16 * kScoreQT;  // 512 rows
// 16-bit score forms: row tiles per wave and rows per workgroup (BF16X6
// holds four query operands per row tile, so it may use fewer tiles).
#ifndef FSK_LP_QT
#define FSK_LP_QT 4
#endif
constexpr int kScoreLpQT = FSK_LP_QT;
#ifndef FSK_LP_QT_X6
#define FSK_LP_QT_X6 4
#endif
constexpr int kScoreLpQTx6 = FSK_LP_QT_X6;
inline int score_lp_qt(unsigned prec) { return prec == FSK_PREC_BF16X6 ? kScoreLpQTx6 : kScoreLpQT; }
// Waves per workgroup of the 16-bit score kernels: the forms whose phi.X takes
// FP16 phi leave out most tile pairs of spatially ordered data and then do far
// less work per staged pair, so eight waves share each LDS chunk.
constexpr int score_lp_waves(unsigned prec) {
  return prec == FSK_PREC_F16 || prec == FSK_PREC_F16X3S || prec == FSK_PREC_BF16X3S ? 8 : 4;
}
// The 16-bit score kernels take data tiles in pairs: their splits are whole pairs.
constexpr int kLpPairPts = 32;
// Symmetric score: I block = rows of one workgroup (4 waves x kSymQT tiles).
constexpr int kSymQT = 8;
constexpr int kSymRows = 4 * 16 * kSymQT;  // 512 points per block
// J blocks per task. The row sums of a task stay in FP32 over jc blocks
// before their FP64 atomic add, so jc trades task overhead for accuracy: at
// n = 2^20 (8-component mixture) jc = 4, 8, 16 ran the FP32 kernel in 478,
// 465, 459 ms with largest mean-shift errors (w/pdf vs FP64) of 1.13, 1.46
// and 2.07e-5 h (the ordered kernel: 0.92e-5 h). Flushing the row sums every
// 4 blocks inside an 8-block task would keep the smaller error, but the
// flush inside the block loop took the kernel past 256 VGPRs.
#ifndef FSK_SYM_JC
#define FSK_SYM_JC 8
#endif
constexpr int64_t kSymJC = FSK_SYM_JC;
constexpr int kSymTS = 20;         // LDS row stride (floats) of transposed tiles and column sums
constexpr int kSymSlots = 2;       // phi tiles a wave keeps in LDS at a time
constexpr int kSymFlushTiles = 8;  // data tiles between two flushes of the column sums
// Tasks (workgroups) per launch of the symmetric kernel. Their number grows as
// n^2 / 2^22 (jc = 8) and a launch may hold at most 2^32 work-items (2^24
// workgroups of 256), so larger n runs as several launches over consecutive
// task ranges; n = 2^20 (263,168 tasks) is one launch, n = 2^21 two, n = 2^22
// five.
constexpr int64_t kSymTasksPerLaunch = int64_t(1) << 20;
// Target number of (query block, split) tasks per launch: a few per CU slot.
constexpr int64_t kTargetTasks = 8192;
// Cap on split-K partial workspace for the ordered score (bytes).
constexpr int64_t kOrderedPartialCap = int64_t(2) << 30;
// 1-D paths: queries per thread and threads per workgroup.
constexpr int k1dQPT = 8;
constexpr int k1dRowsPerWG = 256 * k1dQPT;
constexpr int k1dChunk = 1024;  // data points staged per LDS round

struct SplitPlan {
  int64_t qblocks;    // query blocks (workgroups along queries)
  int64_t splits;     // data splits
  int64_t split_len;  // data points per split (multiple of the tile)
};

inline int64_t p_cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline SplitPlan make_split_plan(int64_t n, int64_t m, int64_t rows_per_wg,
                                 int64_t tile, int64_t min_split_len,
                                 int64_t max_splits) {
  SplitPlan p;
  p.qblocks = p_cdiv(m, rows_per_wg);
  int64_t s = p_cdiv(kTargetTasks, p.qblocks);
  int64_t s_max_len = p_cdiv(n, min_split_len);
  if (s > s_max_len) s = s_max_len;
  if (s > max_splits) s = max_splits;
  if (s < 1) s = 1;
  int64_t len = p_cdiv(p_cdiv(n, s), tile) * tile;
  p.split_len = len;
  p.splits = p_cdiv(n, len);
  return p;
}

inline SplitPlan kde16_plan(int64_t n, int64_t m) {
  return make_split_plan(n, m, kKdeRowsPerWG, 16, 1024, 1 << 20);
}

inline SplitPlan score16_ordered_plan(int64_t n) {
  int64_t cap = kOrderedPartialCap / (n * 17 * 4 > 0 ? n * 17 * 4 : 1);
  if (cap < 1) cap = 1;
  return make_split_plan(n, n, kScoreRowsPerWG, 16, 2048, cap);
}

// 16-bit score of m rows against n data points (m = n for the data set itself).
inline SplitPlan score16_lp_plan(int64_t n, int64_t m, unsigned prec) {
  int64_t cap = kOrderedPartialCap / (m * 17 * 4 > 0 ? m * 17 * 4 : 1);
  if (cap < 1) cap = 1;
  return make_split_plan(n, m, score_lp_waves(prec) * 16 * score_lp_qt(prec), kLpPairPts, 2048,
                         cap);
}

inline SplitPlan score16_rows_plan(int64_t n, int64_t m) {
  int64_t cap = kOrderedPartialCap / (m * 17 * 4 > 0 ? m * 17 * 4 : 1);
  if (cap < 1) cap = 1;
  return make_split_plan(n, m, kScoreRowsPerWG, 16, 2048, cap);
}

inline SplitPlan kde1_plan(int64_t n, int64_t m) {
  return make_split_plan(n, m, k1dRowsPerWG, k1dChunk, 4096, 1 << 20);
}

// Bump allocator over the caller's workspace (256-byte aligned slices).
struct Carve {
  char *base;
  size_t used;
  explicit Carve(void *b) : base(static_cast<char *>(b)), used(0) {}
  template <class T>
  T *take(int64_t count) {
    size_t off = used;
    used += (size_t(count) * sizeof(T) + 255) & ~size_t(255);
    return base ? reinterpret_cast<T *>(base + off) : nullptr;
  }
};

}  // namespace fsk
