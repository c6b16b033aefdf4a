// fsk_kde16.hip - 16-D Gaussian KDE sums on MFMA (split-K and atomic forms).
//
// Data and queries are first centred on the data mean (FP64 mean, one FP32
// rounding of each difference): distances are translation invariant and the
// norm terms, whose cancellation sets the FP32 error, shrink. Per pair (i, j)
// of centred points the kernel evaluates
//     D_ij = |x_j|^2 - 2 q_i . x_j                  (MFMA, natural units)
//     t_ij = c D_ij + c |q_i|^2,   c = -log2(e) / (2 h^2)   (one FMA)
//     phi  = exp2(min(t_ij, 0))
// which is exp(-max(|q|^2 + |x|^2 - 2 q.x, 0) / (2 h^2)), the reference's
// clamped norm expansion (kernels/kde_eval_16d_stream_splitk.py:71-75).
// |x_j|^2 (FP64, rounded once) initialises the MFMA accumulator, the cross
// term is four chained v_mfma_f32_16x16x4_f32 against -2 q (an exact
// scaling), and fma / exp2 / add finish the pair on the VALU. Padded tail
// points carry a norm term P with c P = -1e30, so they contribute exactly 0.
#include <mutex>

#include "fsk_common.h"
#include "fsk_lowp.h"
#include "fsk_plan.h"
#include "fsk_tile.h"

namespace fsk {

constexpr int kFlushTiles = 32;  // FP32 inner sums of 4*32 terms, then FP64

// One workgroup = 4 waves; a wave owns QT row tiles of 16 queries and streams
// the data range of its split in tiles of 16 points.
// MODE 0: sums of phi; 1: sums of phi (1 + 8 - |q-x|^2/(2h^2)) (the fused
// Laplacian correction); 2: sums of phi |q-x|^2/h^2 alone (the correction's
// moment, for the two-pass form). With t = log2(phi) = -|q-x|^2 log2(e)/(2h^2),
// |q-x|^2/(2h^2) = -ln2 t, so modes 1 and 2 need the one extra FMA phi t.
template <int QT, int MODE, bool ATOMIC>
__global__ __launch_bounds__(kWG) void kde16_f32_kernel(
    const float *__restrict__ data, int64_t ld, int64_t n, const float *__restrict__ bn,
    int64_t n_pad, const float *__restrict__ query, int64_t ldq, int64_t m,
    const float *__restrict__ qa, float coef, int64_t split_len, int64_t qblocks,
    int64_t ntasks, int64_t per_xcd, float *__restrict__ out, int64_t ld_out) {
  constexpr bool LAPL = MODE != 0;
  const int64_t task = xcd_task(blockIdx.x, per_xcd);
  if (task >= ntasks) return;
  const int64_t split = task / qblocks;
  const int64_t qb = task - split * qblocks;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t row0 = qb * (64 * QT) + int64_t(wave) * (16 * QT);

  f4 qv[QT];
  float qav[QT], acc[QT], accT[QT];
  double acc64[QT], accT64[QT];
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    int64_t i = row0 + 16 * R + c;
    if (i > m - 1) i = m - 1;
    qv[R] = ld4(query + i * ldq + 4 * g) * -2.f;
    qav[R] = qa[i];
    acc[R] = 0.f;
    accT[R] = 0.f;
    acc64[R] = 0.0;
    accT64[R] = 0.0;
  }

  const int64_t j_begin = split * split_len;
  int64_t j_end = j_begin + split_len;
  if (j_end > n_pad) j_end = n_pad;
  int tc = 0;
  for_each_tile<false>(data, ld, n, bn, j_begin, j_end, g, c, [&](const TileData &t) {
    f4 d[QT];
    cross_tile<QT>(qv, t, d);
    kernel_tile<QT, LAPL, true>(d, qav, coef, acc, accT);
    if (++tc == kFlushTiles) {
      tc = 0;
#pragma unroll
      for (int R = 0; R < QT; ++R) {
        acc64[R] += acc[R];
        acc[R] = 0.f;
        if (LAPL) {
          accT64[R] += accT[R];
          accT[R] = 0.f;
        }
      }
    }
  });
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    double v = acc64[R] + acc[R];
    if (MODE == 1) v = 9.0 * v + kLn2 * (accT64[R] + accT[R]);
    if (MODE == 2) v = -2.0 * kLn2 * (accT64[R] + accT[R]);
    v = sum_groups(v);
    int64_t i = row0 + 16 * R + c;
    if (g == 0 && i < m) {
      if (ATOMIC)
        atomicAdd(out + i, float(v));
      else
        out[split * ld_out + i] = float(v);
    }
  }
}

struct Kde16Layout {
  double *mu;   // data mean (column sums) workspace
  float *xc;    // centred data, zero rows up to a multiple of the tile
  float *qc;    // centred queries
  float *bn;
  float *qa;
  float *partials;
};

static size_t kde16_layout(void *ws, int64_t n, int64_t m, bool with_partials,
                           Kde16Layout *L) {
  Carve cv(ws);
  SplitPlan p = kde16_plan(n, m);
  L->mu = cv.take<double>(kMeanWs);
  L->xc = cv.take<float>(round_up(n, kTile) * kD);
  L->qc = cv.take<float>(m * kD);
  L->bn = cv.take<float>(round_up(n, kNormPad));
  L->qa = cv.take<float>(m);
  L->partials = with_partials ? cv.take<float>(p.splits * m) : nullptr;
  return cv.used;
}

size_t kde16_ws_bytes(int64_t n, int64_t m, bool with_partials) {
  Kde16Layout L;
  return kde16_layout(nullptr, n, m, with_partials, &L);
}

static int kde16_run(const char *what, bool atomic, const float *data, int64_t ld, int64_t n,
                     const float *query, int64_t ldq, int64_t m, double h, unsigned flags,
                     void *ws, size_t ws_bytes, float *out, void *stream) {
  int st;
  if ((st = check_points16(what, data, ld, n)) != FSK_OK) return st;
  if ((st = check_points16(what, query, ldq, m)) != FSK_OK) return st;
  if (!(h > 0.0) || !isfinite(h)) {
    set_error("%s: bandwidth must be positive, got %g", what, h);
    return FSK_EINVAL;
  }
  if (!out) {
    set_error("%s: null output", what);
    return FSK_EINVAL;
  }
  const unsigned prec = flags & FSK_PREC_MASK;
  if (prec != FSK_PREC_F32 || (flags & ~(FSK_PREC_MASK | FSK_FLAG_LAPLACIAN | FSK_FLAG_PHI_S))) {
    set_error("%s: flags 0x%x not available on this path", what, flags);
    return FSK_EUNSUPPORTED;
  }
  const bool lapl = (flags & FSK_FLAG_LAPLACIAN) != 0, moment = (flags & FSK_FLAG_PHI_S) != 0;
  if (lapl && moment) {
    set_error("%s: FSK_FLAG_LAPLACIAN and FSK_FLAG_PHI_S exclude each other", what);
    return FSK_EINVAL;
  }
  const int mode = lapl ? 1 : (moment ? 2 : 0);
  Kde16Layout L;
  size_t need = kde16_layout(ws, n, m, !atomic, &L);
  if (!ws || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, got %zu", what, need, ws_bytes);
    return FSK_EWORKSPACE;
  }
  hipStream_t s = as_stream(stream);
  const double coef = exp2_coef(h);
  const int64_t n_pad = round_up(n, kTile);
  const double inv_n = 1.0 / double(n);
  // centre data and queries on the data mean; data norms in natural units
  // (padded rows: c P = -1e30, P kept finite), query norms times c
  const double pad = fmin(1e30 / -coef, 1e38);
  launch_mean16(data, ld, n, L.mu, s);
  hipLaunchKernelGGL(center16_kernel, dim3(cdiv(n_pad * kD, 256)), dim3(256), 0, s, data, ld, n,
                     n_pad, L.mu, inv_n, L.xc);
  hipLaunchKernelGGL(center16_kernel, dim3(cdiv(m * kD, 256)), dim3(256), 0, s, query, ldq, m, m,
                     L.mu, inv_n, L.qc);
  hipLaunchKernelGGL(prep_norms16, dim3(cdiv(round_up(n, kNormPad), 256)), dim3(256), 0, s,
                     L.xc, int64_t(kD), n, round_up(n, kNormPad), 1.0, float(pad), L.bn);
  hipLaunchKernelGGL(prep_norms16, dim3(cdiv(m, 256)), dim3(256), 0, s, L.qc, int64_t(kD), m, m,
                     coef, kPadNeg, L.qa);
  SplitPlan p = kde16_plan(n, m);
  const int64_t ntasks = p.qblocks * p.splits;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
  const float fcoef = float(coef);
  const int op = atomic ? FSK_OP_KDE16_ATOMIC : FSK_OP_KDE16_SPLITK;
  float *dst = atomic ? out : L.partials;
  const int64_t ld_out = atomic ? 0 : m;
  if (atomic) hipLaunchKernelGGL(fill_f32, dim3(cdiv(m, 256)), dim3(256), 0, s, out, m, 0.f);
  prof_begin(op, s);
#define FSK_KDE16(MODE, AT)                                                                     \
  hipLaunchKernelGGL((kde16_f32_kernel<kKdeQT, MODE, AT>), dim3(grid), dim3(kWG), 0, s, L.xc,  \
                     int64_t(kD), n, L.bn, n_pad, L.qc, int64_t(kD), m, L.qa, fcoef,            \
                     p.split_len, p.qblocks, ntasks, per_xcd, dst, ld_out)
  if (atomic) {
    if (mode == 0) FSK_KDE16(0, true);
    else if (mode == 1) FSK_KDE16(1, true);
    else FSK_KDE16(2, true);
  } else {
    if (mode == 0) FSK_KDE16(0, false);
    else if (mode == 1) FSK_KDE16(1, false);
    else FSK_KDE16(2, false);
  }
#undef FSK_KDE16
  if (atomic) {
    prof_end(op, s);
    return launch_status(what);
  }
  prof_end(op, s);
  if ((st = launch_status(what)) != FSK_OK) return st;
  return fsk_reduce_partials_f64(L.partials, p.splits, m, m, out, stream);
}

// ---- allocation-free variant of the split-K KDE ----
// Its workspace is a file-scope device array instead of a caller buffer, for
// callers that cannot allocate (code under graph capture, serving loops).
// Every device the library runs on holds its own copy until the process ends.
// The host never uses the array's name as an address: the per-device address
// comes from hipGetSymbolAddress on the first call on that device.
constexpr size_t kNoAllocScratchBytes = size_t(FSK_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device (looked up once per device).
static int scratch_for_current_device(const char *what, void **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("%s: no usable current device (%s, id %d)", what, hipGetErrorString(e), dev);
    return FSK_ELAUNCH;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("%s: hipGetSymbolAddress(scratch) failed on device %d: %s", what, dev,
                hipGetErrorString(e));
      return FSK_ELAUNCH;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FSK_OK;
}

}  // namespace fsk

using namespace fsk;

extern "C" int fsk_kde16_splitk_noalloc(const float *data, int64_t ld, int64_t n,
                                        const float *query, int64_t ldq, int64_t m,
                                        double bandwidth, unsigned flags, float *out,
                                        void *stream) {
  const char *what = "fsk_kde16_splitk_noalloc";
  int st;
  if ((st = check_points16(what, data, ld, n)) != FSK_OK) return st;
  if ((st = check_points16(what, query, ldq, m)) != FSK_OK) return st;
  const size_t need = kde16_ws_bytes(n, m, true);
  if (need > kNoAllocScratchBytes) {
    set_error("%s: n = %lld points and m = %lld queries need %zu bytes of scratch, above the "
              "%zu-byte static buffer; use fsk_kde16_splitk with a workspace",
              what, (long long)n, (long long)m, need, kNoAllocScratchBytes);
    return FSK_EUNSUPPORTED;
  }
  void *ws = nullptr;
  if ((st = scratch_for_current_device(what, &ws)) != FSK_OK) return st;
  // every byte the kernels read from the scratch is written earlier in the
  // same call (mean, centred copies, norms, all split partials): nothing
  // needs clearing between calls
  return kde16_run(what, false, data, ld, n, query, ldq, m, bandwidth, flags, ws,
                   kNoAllocScratchBytes, out, stream);
}

extern "C" int fsk_kde16_splitk(const float *data, int64_t ld, int64_t n, const float *query,
                                int64_t ldq, int64_t m, double bandwidth, unsigned flags,
                                void *ws, size_t ws_bytes, float *out, void *stream) {
  return kde16_run("fsk_kde16_splitk", false, data, ld, n, query, ldq, m, bandwidth, flags, ws,
                   ws_bytes, out, stream);
}

extern "C" int fsk_kde16_atomic(const float *data, int64_t ld, int64_t n, const float *query,
                                int64_t ldq, int64_t m, double bandwidth, unsigned flags,
                                void *ws, size_t ws_bytes, float *out, void *stream) {
  return kde16_run("fsk_kde16_atomic", true, data, ld, n, query, ldq, m, bandwidth, flags, ws,
                   ws_bytes, out, stream);
}
