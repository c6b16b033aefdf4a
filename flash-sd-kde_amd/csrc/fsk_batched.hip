// fsk_batched.hip - leave-one-out KDE log-likelihood of many datasets over a
// bandwidth grid, in one persistent launch.
//
// For dataset d (n points) and bandwidth h_b:
//   LL_b = 1/n sum_i log( 1/(n-1) sum_{j != i} (2 pi)^-8 h_b^-16 exp(-|x_i - x_j|^2 / 2h_b^2) )
// Each task = (dataset, block of rows) makes one pass over all points of its
// dataset and accumulates, per row i and bandwidth b,
//   S_ib = sum_{j != i} exp2(c_b (|x_i - x_j|^2 - m_i)),  c_b = -log2(e)/(2 h_b^2),
// with the shift m_i the running minimum of |x_i - x_j|^2 over the points
// seen so far: when a tile brings a smaller distance, the row's sums are
// rescaled by exp2(c_b (m_old - m_new)) <= 1 and m_i drops. At the end m_i is
// the nearest-neighbour distance dmin_i, so
//   log sum_j exp(-d2/2h^2) = c_b dmin_i ln 2 + log S_ib
// exactly (log-sum-exp): every term is <= 1, the nearest neighbour
// contributes 1, and no bandwidth of the grid can underflow. Tiles are visited
// from the row block's own tile onwards (wrapping around), and the points are
// in spatial order, so the minimum is nearly found on the first tiles and
// later rescales are rare (about 1 % of the tiles of a wave).
// One distance feeds all n_bw (<= 32) exponentials; the cross term stays FP32
// MFMA. Each dataset is first centred on its own FP64 mean (distances are
// translation invariant; the norm terms, whose cancellation sets the FP32
// error of |x_i|^2 + |x_j|^2 - 2 x_i.x_j, shrink to the dataset's spread), and
// the MFMA accumulator starts from |x_j|^2 alone, |x_i|^2 being added after
// the cross term.
// Tasks come from an atomic work queue, largest datasets first; per task the
// row sums of log S_ib and of dmin_i are reduced in LDS and added to FP64
// per-(dataset, bandwidth) accumulators; a finish kernel forms LL.
#include <type_traits>

#include "fsk_common.h"
#include "fsk_plan.h"

namespace fsk {

constexpr int kLooMaxSets = 64;
constexpr int kLooMaxBw = 32;
constexpr int kLooFlush = 16;
// kLooFar: the shift m_i of a row before its first real distance. kLooPad:
// the squared distance of padded points, of a row's own point and of idle
// rows, four times farther, so its delta against any shift stays >= 3e30 and
// its terms are exp2(-huge) = 0 (and it never lowers a shift).
constexpr float kLooFar = 1.0e30f;
constexpr float kLooPad = 4.0e30f;

struct LooMeta {
  int n_sets, n_bw;
  int64_t ntasks;
  int64_t off[kLooMaxSets + 1];      // dataset row offsets (original order)
  int64_t poff[kLooMaxSets + 1];     // the same in the padded centred copy (multiples of 16)
  int64_t task_end[kLooMaxSets];     // cumulative task count in processing order
  int order[kLooMaxSets];            // processing order (largest first)
};

// One workgroup per dataset: its FP64 mean (fixed summation order: per
// thread over a strided row set, then a fixed LDS tree), then the centred
// copy xc (16 floats per row, one rounding of the FP64 difference) and the
// squared norms of the centred rows. Each dataset's rows start at poff[d] and
// are padded to a multiple of 16 with zero rows whose norm is kLooPad, so
// every 16-row data tile is read whole and padded points add exp2(-huge) = 0.
__global__ __launch_bounds__(256) void loo_center_kernel(const float *__restrict__ x, int64_t ld,
                                                         LooMeta meta, float *__restrict__ xc,
                                                         float *__restrict__ nrm) {
  __shared__ double red[256][kD + 1];
  __shared__ double mu[kD];
  const int d = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t beg = meta.off[d], end = meta.off[d + 1];
  double acc[kD];
#pragma unroll
  for (int f = 0; f < kD; ++f) acc[f] = 0.0;
  for (int64_t j = beg + tid; j < end; j += 256) {
#pragma unroll
    for (int f = 0; f < kD; f += 4) {
      const f4 v = ld4(x + j * ld + f);
      acc[f] += v[0];
      acc[f + 1] += v[1];
      acc[f + 2] += v[2];
      acc[f + 3] += v[3];
    }
  }
#pragma unroll
  for (int f = 0; f < kD; ++f) red[tid][f] = acc[f];
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int f = 0; f < kD; ++f) red[tid][f] += red[tid + w][f];
    }
    __syncthreads();
  }
  if (tid < kD) mu[tid] = red[0][tid] / double(end - beg);
  __syncthreads();
  const int64_t pbeg = meta.poff[d], pend = meta.poff[d + 1];
  for (int64_t j = beg + tid; j < end; j += 256) {
    const int64_t jp = pbeg + (j - beg);
    double s = 0.0;
#pragma unroll
    for (int f = 0; f < kD; f += 4) {
      const f4 v = ld4(x + j * ld + f);
      f4 o;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        o[e] = float(double(v[e]) - mu[f + e]);
        s += double(o[e]) * double(o[e]);
      }
      *reinterpret_cast<f4 *>(xc + jp * kD + f) = o;
    }
    nrm[jp] = float(s);
  }
  for (int64_t jp = pbeg + (end - beg) + tid; jp < pend; jp += 256) {
#pragma unroll
    for (int f = 0; f < kD; f += 4) *reinterpret_cast<f4 *>(xc + jp * kD + f) = f4{0.f, 0.f, 0.f, 0.f};
    nrm[jp] = kLooPad;
  }
}

// Bandwidth skip groups. For a dataset of n points let T = ceil(log2 n) +
// kLooSkipMargin. A term exp2(c_b delta) with c_b delta <= -T is below 2^-T,
// and S_ib >= 1 (the nearest neighbour's term is exp2(0)), so all such terms
// of a row together change S_ib by less than n 2^-T <= 2^-28 relative: below
// one FP32 rounding of S_ib, and at most 2^-28 on log S_ib. Per data tile a
// group of kLooGroup bandwidths is evaluated only if some lane of the wave has
// a delta below T / |c_b| for the group's widest member (one compare and
// ballot per group); every other group is skipped. At the grid's narrow end
// (0.05 h, c_b 400 times that at h) most tiles skip most groups.
constexpr int kLooGroup = 4;
constexpr int kLooGroups = kLooMaxBw / kLooGroup;
constexpr float kLooSkipMargin = 28.f;

// Register swaps between lane halves (32 lanes) and between neighbouring
// 16-lane rows. v_permlane32_swap a, b swaps the upper 32 lanes of a with the
// lower 32 lanes of b, leaving [a_lo | b_lo] in a and [a_hi | b_hi] in b;
// v_permlane16_swap does the same with the odd 16-lane rows of a and the even
// rows of b. They are written as register-only inline asm: hipcc 7.2 folds a
// combination of the two results of __builtin_amdgcn_permlane*_swap into the
// first result alone (min(a[0], a[1]) becomes a[0]). The s_nop 1 gives the
// two wait states a swap needs after a VALU write of its operands.
__device__ __forceinline__ void swap32(float &a, float &b) {
  asm volatile("s_nop 1\n\tv_permlane32_swap_b32 %0, %1" : "+v"(a), "+v"(b));
}
__device__ __forceinline__ void swap16(float &a, float &b) {
  asm volatile("s_nop 1\n\tv_permlane16_swap_b32 %0, %1" : "+v"(a), "+v"(b));
}

// Sum over the four lane groups of a wave for four values at once: lane
// (g, c) returns v[g] summed over the lanes {c, c+16, c+32, c+48}. After
// swap32(v0, v2) the sum v0 + v2 holds, in the lanes of groups 0 and 1, v0
// plus v0 of the lane 32 above, and in groups 2 and 3, v2 plus v2 of the lane
// 32 below; likewise v1 + v3 after swap32(v1, v3); one swap16 of the two
// partial sums then adds the neighbouring row's, leaving in group g the total
// of v[g]. Three swaps and three adds for four sums.
__device__ __forceinline__ float group_sum_exchange4(float v0, float v1, float v2, float v3) {
  swap32(v0, v2);
  swap32(v1, v3);
  float a0 = v0 + v2, a1 = v1 + v3;
  swap16(a0, a1);
  return a0 + a1;
}

// Minimum over the four lane groups: lane (g, c) returns the minimum of v over
// the lanes {c, c+16, c+32, c+48}: started from a = b = v, the two results
// of each swap hold v and its partner 32 (then 16) lanes away.
__device__ __forceinline__ float group_min4(float v) {
  float a = v, b = v;
  swap32(a, b);
  v = fminf(a, b);
  a = v;
  b = v;
  swap16(a, b);
  return fminf(a, b);
}

// WAVES waves per workgroup, each owning RT row tiles (16 RT rows): a task is
// 16 RT WAVES rows of one dataset against all of its points. RT trades
// registers for data loads and per-tile work shared by more rows; the
// autotuner picks (WAVES, RT, workgroups per CU).
// Per lane (g, c) and row tile: FP32 sums S[b] of its own four pairs
// (row c, data points 4g + r) for all bandwidths, flushed every kLooFlush
// tiles into FP64 sums that are spread over the lane groups: lane (g, c) keeps
// row c's FP64 sums of the bandwidths b = 4q + g.
template <int WAVES, int RT>
__global__ __launch_bounds__(64 * WAVES) void loo_grid_kernel(
    const float *__restrict__ x, int64_t ld, const float *__restrict__ nrm,
    const double *__restrict__ bw, LooMeta meta, int *__restrict__ queue,
    double *__restrict__ sumlog, double *__restrict__ sumdmin,
    unsigned long long *__restrict__ exps_done) {
  constexpr int ROWS = 16 * RT * WAVES;
  constexpr int kQ = kLooMaxBw / 4;
  __shared__ int task_sh;
  __shared__ float coef_sh[kLooMaxBw];
  __shared__ double red_sh[WAVES][kLooMaxBw + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int nbw = meta.n_bw;

  for (;;) {
    __syncthreads();
    if (tid == 0) task_sh = atomicAdd(queue, 1);
    __syncthreads();
    const int64_t t = task_sh;
    if (t >= meta.ntasks) break;
    // decode (scalar): which dataset, which row block
    int k = 0;
    while (t >= meta.task_end[k]) ++k;
    const int d = meta.order[k];
    const int64_t blk = t - (k ? meta.task_end[k - 1] : 0);
    const int64_t n = meta.off[d + 1] - meta.off[d];
    const int64_t pbeg = meta.poff[d], pend = meta.poff[d + 1];
    if (tid < kLooMaxBw) {
      const double h = tid < nbw ? bw[d * kLooMaxBw + tid] : 0.0;
      coef_sh[tid] = tid < nbw ? float(-kLog2e / (2.0 * h * h)) : 0.f;
    }
    __syncthreads();
    // The coefficients stay in vector registers (the same value in every
    // lane): as scalars they overflow the SGPR file next to the loop state.
    float cf[kLooMaxBw];
#pragma unroll
    for (int b = 0; b < kLooMaxBw; ++b) cf[b] = coef_sh[b];
    // skip threshold of each group: delta < T / min|c_b| over its bandwidths
    // with T = ceil(log2 n) + kLooSkipMargin (0, never active, for a group
    // past n_bw)
    const float skip_exp = float(n > 1 ? 64 - __clzll((unsigned long long)(n - 1)) : 0) + kLooSkipMargin;
    float thr[kLooGroups];
#pragma unroll
    for (int q = 0; q < kLooGroups; ++q) {
      float mn = __builtin_inff();
#pragma unroll
      for (int e = 0; e < kLooGroup; ++e)
        if (q * kLooGroup + e < nbw) mn = fminf(mn, -cf[q * kLooGroup + e]);
      thr[q] = skip_exp / mn;
    }

    // this wave's 16 RT rows (padded coordinates; the rows past n are idle)
    const int64_t rbase = pbeg + blk * ROWS + wave * 16 * RT;
    bool row_ok[RT];
    f4 qv[RT];
    float qa[RT], m[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      const int64_t i = rbase + 16 * rt + c;
      row_ok[rt] = i - pbeg < n;
      const int64_t ic = row_ok[rt] ? i : pbeg;
      qv[rt] = ld4(x + ic * ld + 4 * g) * -2.0f;
      // idle rows: distances ~kLooPad against a shift of 0, far above every
      // skip threshold (exps of 0), and never a rescale
      qa[rt] = row_ok[rt] ? nrm[ic] : kLooPad;
      m[rt] = row_ok[rt] ? kLooFar : 0.f;
    }
    // coefficients of the FP64 sums this lane keeps (bandwidths 4q + g)
    float cfl[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) cfl[q] = coef_sh[4 * q + g];

    // S: per lane and row tile, FP32 sums of the lane's own four pairs (row c,
    // data points 4g + r) for all bandwidths, flushed every kLooFlush tiles
    // into the FP64 sums S64 spread over the lane groups.
    float S[RT][kLooMaxBw];
    double S64[RT][kQ];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
#pragma unroll
      for (int b = 0; b < kLooMaxBw; ++b) S[rt][b] = 0.f;
#pragma unroll
      for (int q = 0; q < kQ; ++q) S64[rt][q] = 0.0;
    }
    auto flush = [&]() {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
          const float v = group_sum_exchange4(S[rt][4 * q], S[rt][4 * q + 1], S[rt][4 * q + 2],
                                              S[rt][4 * q + 3]);
          S64[rt][q] += double(v);
#pragma unroll
          for (int e = 0; e < 4; ++e) S[rt][4 * q + e] = 0.f;
        }
    };
    unsigned n_active = 0;  // active (tile, group) pairs of this wave and task
    // One data tile: squared distances, the running row minima (and the
    // rescale of a row's sums when its minimum drops), deltas, the wave's
    // smallest delta, and the exps of the bandwidth groups it leaves active.
    // DIAG: the tile holds the rows of one of the wave's row tiles, whose self
    // pairs are dropped.
    auto tile = [&](int64_t j0, auto diag) {
      const f4 xa = ld4(x + (j0 + c) * ld + 4 * g);
      const f4 bb = ld4(nrm + j0 + 4 * g);
      float delta[RT][4];
      float lo = kLooFar;
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        f4 dd = bb;
#pragma unroll
        for (int s = 0; s < 4; ++s) dd = mfma_f32(xa[s], qv[rt][s], dd);
        dd += qa[rt];
        if constexpr (decltype(diag)::value) {
          if (j0 == rbase + 16 * rt) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (4 * g + r == c) dd[r] = kLooPad;
          }
        }
        const float lmin = fminf(fminf(dd[0], dd[1]), fminf(dd[2], dd[3]));
        // a new minimum in some row of the tile (about 1 % of the tiles):
        // the row's four lanes agree on it, and its sums move to the new shift
        if (__builtin_amdgcn_ballot_w64(lmin < m[rt])) {
          const float tm = group_min4(lmin);
          if (tm < m[rt]) {
            const float dm = m[rt] - tm;
#pragma unroll
            for (int b = 0; b < kLooMaxBw; ++b) S[rt][b] *= fexp2(cf[b] * dm);
#pragma unroll
            for (int q = 0; q < kQ; ++q) S64[rt][q] *= double(fexp2(cfl[q] * dm));
            m[rt] = tm;
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          delta[rt][r] = dd[r] - m[rt];  // >= 0: m is at most every distance seen
          lo = fminf(lo, delta[rt][r]);
        }
      }
#pragma unroll
      for (int q = 0; q < kLooGroups; ++q) {
        if (__builtin_amdgcn_ballot_w64(lo < thr[q])) {
          ++n_active;
#pragma unroll
          for (int e = 0; e < kLooGroup; ++e) {
            const int b = q * kLooGroup + e;
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
              for (int r = 0; r < 4; ++r) S[rt][b] += fexp2(delta[rt][r] * cf[b]);
          }
        }
      }
    };
    // tiles from the wave's own rows onwards, wrapping around
    int64_t j0 = rbase < pend ? rbase : pbeg;
    int tc = 0;
    for (int64_t k = pbeg; k < pend; k += kTile) {
      if (j0 >= rbase && j0 < rbase + 16 * RT)
        tile(j0, std::true_type{});
      else
        tile(j0, std::false_type{});
      j0 += kTile;
      if (j0 == pend) j0 = pbeg;
      if (++tc == kLooFlush) {
        tc = 0;
        flush();
      }
    }
    flush();
    if (exps_done && lane == 0)
      atomicAdd(exps_done, (unsigned long long)n_active * (kLooGroup * 64 * 4 * RT));
    // ---- per-task reduction: sum over rows of log S_ib and of dmin_i ----
    // lane (g, c) holds row c's sums of the bandwidths 4q + g
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const int b = 4 * q + g;
      double v = 0.0;
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) v += (row_ok[rt] && b < nbw) ? log(S64[rt][q]) : 0.0;
      // sum over the 16 rows (lanes c of the group)
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      v += __shfl_xor(v, 4);
      v += __shfl_xor(v, 8);
      if (c == 0) red_sh[wave][b] = v;
    }
    double dm = 0.0;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) dm += row_ok[rt] ? double(m[rt]) : 0.0;
    if (g != 0) dm = 0.0;
    dm += __shfl_xor(dm, 1);
    dm += __shfl_xor(dm, 2);
    dm += __shfl_xor(dm, 4);
    dm += __shfl_xor(dm, 8);
    if (lane == 0) red_sh[wave][kLooMaxBw] = dm;
    __syncthreads();
    if (tid <= kLooMaxBw) {
      double s = 0.0;
      for (int w = 0; w < WAVES; ++w) s += red_sh[w][tid];
      if (tid < nbw) atomicAdd(&sumlog[d * kLooMaxBw + tid], s);
      if (tid == kLooMaxBw) atomicAdd(&sumdmin[d], s);
    }
  }
}

__global__ void loo_finish_kernel(const double *__restrict__ sumlog,
                                  const double *__restrict__ sumdmin,
                                  const double *__restrict__ bw, LooMeta meta,
                                  double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= meta.n_sets * meta.n_bw) return;
  const int d = e / meta.n_bw, b = e % meta.n_bw;
  const double n = double(meta.off[d + 1] - meta.off[d]);
  const double h = bw[d * kLooMaxBw + b];
  out[e] = sumlog[d * kLooMaxBw + b] / n - sumdmin[d] / (2.0 * h * h * n) - log(n - 1.0) -
           0.5 * kD * log(2.0 * M_PI) - kD * log(h);
}

__global__ void loo_pack_bw_kernel(const double *__restrict__ bw_in, int n_sets, int n_bw,
                                   double *__restrict__ bw_out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_sets * kLooMaxBw) return;
  const int d = e / kLooMaxBw, b = e % kLooMaxBw;
  bw_out[e] = b < n_bw ? bw_in[d * n_bw + b] : 1.0;
}

struct LooLayout {
  int *queue;
  float *xc;  // rows centred on their dataset's mean
  float *nrm;
  double *bw, *sumlog, *sumdmin;
};

// Rows of the padded centred copy: every dataset rounded up to whole tiles.
static int64_t loo_padded_rows(const int64_t *offsets, int n_sets) {
  int64_t np = 0;
  for (int d = 0; d < n_sets; ++d) np += round_up(offsets[d + 1] - offsets[d], kTile);
  return np;
}

static size_t loo_layout(void *ws, int64_t N, int n_sets, LooLayout *L) {
  Carve cv(ws);
  L->queue = cv.take<int>(64);
  L->xc = cv.take<float>(N * kD);
  L->nrm = cv.take<float>(N);
  L->bw = cv.take<double>(int64_t(n_sets) * kLooMaxBw);
  L->sumlog = cv.take<double>(int64_t(n_sets) * kLooMaxBw);
  L->sumdmin = cv.take<double>(n_sets);
  return cv.used;
}

}  // namespace fsk

using namespace fsk;

extern "C" size_t fsk_loo_workspace_bytes(const int64_t *offsets, int n_sets, int n_bw) {
  if (!offsets || n_sets <= 0 || n_sets > kLooMaxSets || n_bw <= 0 || n_bw > kLooMaxBw) return 0;
  for (int d = 0; d < n_sets; ++d)
    if (offsets[d + 1] < offsets[d]) return 0;
  LooLayout L;
  return loo_layout(nullptr, loo_padded_rows(offsets, n_sets), n_sets, &L);
}

extern "C" int fsk_loo_bandwidth_grid(const float *data, int64_t ld, const int64_t *offsets,
                                      int n_sets, const double *bandwidths, int n_bw,
                                      int waves_per_wg, int row_tiles, int wg_per_cu, void *ws,
                                      size_t ws_bytes, double *out,
                                      unsigned long long *exps_done, void *stream) {
  const char *what = "fsk_loo_bandwidth_grid";
  if (!offsets || n_sets <= 0 || n_sets > kLooMaxSets || n_bw <= 0 || n_bw > kLooMaxBw ||
      !bandwidths || !out) {
    set_error("%s: need 1..%d datasets, 1..%d bandwidths, non-null pointers", what, kLooMaxSets,
              kLooMaxBw);
    return FSK_EINVAL;
  }
  if (offsets[0] != 0) {
    set_error("%s: offsets[0] must be 0", what);
    return FSK_EINVAL;
  }
  for (int d = 0; d < n_sets; ++d) {
    if (offsets[d + 1] - offsets[d] < 2) {
      set_error("%s: dataset %d has %lld points; leave-one-out needs at least 2", what, d,
                (long long)(offsets[d + 1] - offsets[d]));
      return FSK_EINVAL;
    }
  }
  int st;
  if ((st = check_points16(what, data, ld, offsets[n_sets])) != FSK_OK) return st;
  if (waves_per_wg != 4 && waves_per_wg != 8) {
    set_error("%s: waves_per_wg must be 4 or 8", what);
    return FSK_EINVAL;
  }
  if (row_tiles != 1 && row_tiles != 2) {
    set_error("%s: row_tiles must be 1 or 2", what);
    return FSK_EINVAL;
  }
  if (wg_per_cu < 1 || wg_per_cu > 8) {
    set_error("%s: wg_per_cu must be in 1..8", what);
    return FSK_EINVAL;
  }
  LooLayout L;
  const size_t need = loo_layout(ws, loo_padded_rows(offsets, n_sets), n_sets, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  // plan on the host: largest datasets first (longest tasks start first)
  LooMeta meta;
  meta.n_sets = n_sets;
  meta.n_bw = n_bw;
  for (int d = 0; d <= n_sets; ++d) meta.off[d] = offsets[d];
  meta.poff[0] = 0;
  for (int d = 0; d < n_sets; ++d)
    meta.poff[d + 1] = meta.poff[d] + round_up(offsets[d + 1] - offsets[d], kTile);
  for (int d = 0; d < n_sets; ++d) meta.order[d] = d;
  for (int a = 0; a < n_sets; ++a)
    for (int b = a + 1; b < n_sets; ++b) {
      const int64_t na = offsets[meta.order[a] + 1] - offsets[meta.order[a]];
      const int64_t nb = offsets[meta.order[b] + 1] - offsets[meta.order[b]];
      if (nb > na) {
        const int tmp = meta.order[a];
        meta.order[a] = meta.order[b];
        meta.order[b] = tmp;
      }
    }
  const int rows = 16 * row_tiles * waves_per_wg;
  int64_t acc = 0;
  for (int k = 0; k < n_sets; ++k) {
    const int d = meta.order[k];
    acc += cdiv(offsets[d + 1] - offsets[d], rows);
    meta.task_end[k] = acc;
  }
  for (int k = n_sets; k < kLooMaxSets; ++k) meta.task_end[k] = acc;
  meta.ntasks = acc;

  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(fill_f32, dim3(1), dim3(64), 0, s, reinterpret_cast<float *>(L.queue),
                     int64_t(64), 0.f);
  hipLaunchKernelGGL(fill_f64, dim3(cdiv(int64_t(n_sets) * kLooMaxBw, 256)), dim3(256), 0, s,
                     L.sumlog, int64_t(n_sets) * kLooMaxBw, 0.0);
  hipLaunchKernelGGL(fill_f64, dim3(1), dim3(64), 0, s, L.sumdmin, int64_t(n_sets), 0.0);
  hipLaunchKernelGGL(loo_pack_bw_kernel, dim3(cdiv(int64_t(n_sets) * kLooMaxBw, 256)), dim3(256),
                     0, s, bandwidths, n_sets, n_bw, L.bw);
  hipLaunchKernelGGL(loo_center_kernel, dim3(n_sets), dim3(256), 0, s, data, ld, meta, L.xc,
                     L.nrm);
  static int cus_cached = 0;  // compute units of the current device (queried once)
  if (cus_cached == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cus > 0)
      cus_cached = cus;
    else
      cus_cached = 256;
  }
  const int cus = cus_cached;
  int64_t grid = int64_t(cus) * wg_per_cu;
  if (grid > meta.ntasks) grid = meta.ntasks;
  prof_begin(FSK_OP_LOO_GRID, s);
  const dim3 blk(64 * waves_per_wg);
  if (waves_per_wg == 4 && row_tiles == 1)
    hipLaunchKernelGGL((loo_grid_kernel<4, 1>), dim3(grid), blk, 0, s, L.xc, int64_t(kD), L.nrm,
                       L.bw, meta, L.queue, L.sumlog, L.sumdmin, exps_done);
  else if (waves_per_wg == 4)
    hipLaunchKernelGGL((loo_grid_kernel<4, 2>), dim3(grid), blk, 0, s, L.xc, int64_t(kD), L.nrm,
                       L.bw, meta, L.queue, L.sumlog, L.sumdmin, exps_done);
  else if (row_tiles == 1)
    hipLaunchKernelGGL((loo_grid_kernel<8, 1>), dim3(grid), blk, 0, s, L.xc, int64_t(kD), L.nrm,
                       L.bw, meta, L.queue, L.sumlog, L.sumdmin, exps_done);
  else
    hipLaunchKernelGGL((loo_grid_kernel<8, 2>), dim3(grid), blk, 0, s, L.xc, int64_t(kD), L.nrm,
                       L.bw, meta, L.queue, L.sumlog, L.sumdmin, exps_done);
  prof_end(FSK_OP_LOO_GRID, s);
  if ((st = launch_status(what)) != FSK_OK) return st;
  hipLaunchKernelGGL(loo_finish_kernel, dim3(cdiv(int64_t(n_sets) * n_bw, 256)), dim3(256), 0, s,
                     L.sumlog, L.sumdmin, L.bw, meta, out);
  return launch_status(what);
}
