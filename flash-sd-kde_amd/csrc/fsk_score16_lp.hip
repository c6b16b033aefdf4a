// fsk_score16_lp.hip - empirical-score accumulators with 16-bit MFMA operands
// (BF16 x3 / x6 splits, FP16 x3 split, single BF16, single FP16), ordered
// split-K, deterministic.
//
// Cross term: data / query records (fsk_lowp.h), the same MFMA forms as the
// log-density kernel:
//   BF16 / F16 : [x1 | fold] . [q1 | fold]
//   BF16X3     : [x1 | x2] . [q1 | q1]  +  [x1 | fold] . [q2 | fold]
//   F16X3      : the same with FP16 pieces (phi.X as in BF16X3)
//   F16X3S     : the F16X3 cross term; phi.X with one FP16 phi (below)
//   BF16X3S    : the BF16X3 cross term; phi.X as F16X3S
//   BF16X6     : [x1 | x1] . [q1 | q2]  +  [x2 | x1] . [q1 | q3]
//              + [x2 | x3] . [q2 | q1]  +  [x2 | fold] . [q3 | fold]
//                (FP32-class cross term; phi.X as in BF16X3)
// phi = exp2(min(t, 0)) comes out with lane l holding phi(i = l%16, j = 4(l/16)+u),
// u = 0..3, which is exactly four consecutive K slots of the 16x16x32 A operand,
// so phi.X needs no data movement:
//   A = [phi_h(u=0..3) | phi_l(u=0..3)] per lane (phi = phi_h + phi_l,
//   phi_h = phi truncated to BF16, phi_l = rn_bf16(phi - phi_h)),
//   W += A . [Xh | Xh]  +  A . [Xl | 0]      (BF16x3, BF16x6, FP16x3: drops phi_l.Xl)
//   W += [phi | 0] . [Xh | Xl]               (single forms; the zero half drops Xl)
//   W += [phi | phi] . [Xh | Xl]             (F16X3S, BF16X3S: phi rounded to
//                                             FP16 once, X in FP16 pieces to
//                                             ~22 bits, one MFMA)
// The row sums pdf_i = sum_j phi_ij are FP32 VALU adds of the unrounded phi
// (an MFMA against a ones operand measured slower here).
//
// Data path. These kernels need several 16-byte operands per lane per 16-point
// data tile (the FP32 kernel needs one), so fetching them per wave through the
// vector-memory path is what limits them. Each data tile is therefore stored
// once, by a prep pass, as ready-made MFMA operand slots (per slot 64 lanes x
// 16 bytes, in lane order), and each workgroup copies chunks of tiles into a
// double-buffered LDS area with direct global-to-LDS loads
// (global_load_lds_dwordx4, one 1 KiB slot per wave instruction); the four
// waves then read their operands with ds_read_b128. Global and L2 traffic per
// workgroup drop four-fold (one copy for four waves) and every access is a
// full, aligned 1 KiB.
//
// Points are centred on the data mean: w' = sum phi (x - mu); the finish
// kernel returns w = w' + mu * pdf in FP64.
#include "fsk_common.h"
#include "fsk_lowp.h"
#include "fsk_plan.h"

namespace fsk {

constexpr int kLpFlush = 32;
static_assert(kNormPad % (64 * kScoreLpQT) == 0 && kNormPad % (64 * kScoreLpQTx6) == 0,
              "records must cover whole row blocks");

typedef __attribute__((address_space(3))) void lds_void;

// Operand slots of one 16-point data tile.
template <int PREC>
struct LpBlock {
  static constexpr bool kX6 = PREC == FSK_PREC_BF16X6;
  static constexpr bool kOnePhi = PREC == FSK_PREC_F16X3S || PREC == FSK_PREC_BF16X3S;
  static constexpr bool kX3 = PREC == FSK_PREC_BF16X3 || PREC == FSK_PREC_F16X3 || kOnePhi;
  static constexpr bool kSplitPhi = kX6 || (kX3 && !kOnePhi);
  static constexpr int kCross = kX6 ? 4 : kX3 ? 2 : 1;  // cross-term A operands
  static constexpr int kX = kSplitPhi ? 2 : 1;          // phi.X B operands
  static constexpr int kSlots = kCross + kX;
  static constexpr int kSlotU16 = 64 * 8;               // 16 bytes per lane
  static constexpr int kTileU16 = kSlots * kSlotU16;
  // tiles per LDS chunk: up to 16 KiB per buffer, two buffers, a whole number
  // of slots per wave
  static constexpr int kChunk0 = 16384 / (kTileU16 * 2);
  static constexpr int kChunk =
      kSlots % 4 == 0 ? kChunk0 : kSlots % 2 == 0 ? (kChunk0 & ~1) : (kChunk0 & ~3);
  static constexpr int kChunkU16 = kChunk * kTileU16;
  static_assert((kChunk * kSlots) % 4 == 0, "four waves copy whole slots");
};

// Record halves (lo, hi) of cross-term A operand `s` of a form.
template <int PREC>
__device__ __forceinline__ int cross_slot_off(int s, int g) {
  using RL = RecLayout<PREC>;
  if (LpBlock<PREC>::kX6) {
    const int lo = s == 0 ? RL::kP1 : RL::kP2;
    const int hi = s == 0 ? RL::kP1 : s == 1 ? RL::kP1 : s == 2 ? RL::kP3 : RL::kFold;
    return rec_off(g, lo, hi);
  }
  if (LpBlock<PREC>::kX3) return rec_off(g, RL::kP1, s == 0 ? RL::kP2 : RL::kFold);
  return rec_off(g, RL::kP1, RL::kFold);
}

// Tile-transposed X for phi.X: per 16-point tile T, feature f and lane group g,
// [Xh(4g+u, f) | Xl(4g+u, f)], u = 0..3, as P16 hi/lo pieces of (x - mu) * scale.
template <int P16>
__global__ void prep_xt16_kernel(const float *__restrict__ x, int64_t ld, int64_t n,
                                 int64_t n_pad, const double *__restrict__ mu_sum, double inv_n,
                                 double scale, unsigned short *__restrict__ xt) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_pad * kD) return;
  const int64_t j = e / kD;
  const int f = int(e % kD);
  const float v = j < n ? float((double(x[j * ld + f]) - mu_sum[f] * inv_n) * scale) : 0.f;
  const unsigned short hi = to16<P16>(v);
  const unsigned short lo = to16<P16>(v - from16<P16>(hi));
  const int64_t T = j / kTile;
  const int g = int((j % kTile) / 4), u = int(j % 4);
  const int64_t base = ((T * kD + f) * 4 + g) * 8;
  xt[base + u] = hi;
  xt[base + 4 + u] = lo;
}

// Operand slots of every data tile from the records and the transposed X;
// tiles [ntiles_data, ntiles_alloc) are the zero slack that whole-chunk copies
// may read past the last tile.
template <int PREC>
__global__ void prep_lp_blocks_kernel(const unsigned short *__restrict__ xrec,
                                      const unsigned short *__restrict__ xt,
                                      int64_t ntiles_data, int64_t ntiles_alloc,
                                      unsigned short *__restrict__ blk) {
  using B = LpBlock<PREC>;
  using RL = RecLayout<PREC>;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // (tile, slot, lane)
  if (e >= ntiles_alloc * B::kSlots * 64) return;
  const int lane = int(e % 64);
  const int slot = int((e / 64) % B::kSlots);
  const int64_t T = e / (64 * B::kSlots);
  const int g = lane >> 4, c = lane & 15;
  u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
  if (T < ntiles_data) {
    if (slot < B::kCross) {
      v = ld_u16x8(xrec + (T * kTile + c) * RL::kSize + cross_slot_off<PREC>(slot, g));
    } else {
      const u16x8 t = ld_u16x8(xt + ((T * kD + c) * 4 + g) * 8);  // [Xh | Xl]
      if (B::kX == 1) {
        v = t;
      } else if (slot == B::kCross) {
        v = u16x8{t[0], t[1], t[2], t[3], t[0], t[1], t[2], t[3]};  // [Xh | Xh]
      } else {
        v = u16x8{t[4], t[5], t[6], t[7], 0, 0, 0, 0};  // [Xl | 0]
      }
    }
  }
  *reinterpret_cast<u16x8 *>(blk + e * 8) = v;
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
template <int P16, bool SPLIT, bool DUP = false>
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
    a[2] = DUP ? a[0] : 0u;  // DUP: [phi | phi] against [Xh | Xl]
    a[3] = DUP ? a[1] : 0u;
  }
  return __builtin_bit_cast(u16x8, a);
}

__device__ __forceinline__ u16x8 lds_u16x8(const unsigned short *p) {
  return *reinterpret_cast<const u16x8 *>(p);
}

template <int PREC, int QT>
__global__ __launch_bounds__(kWG) void score16_lp_kernel(
    const unsigned short *__restrict__ blk, const unsigned short *__restrict__ qrec, int64_t m,
    int64_t ntiles_data, int64_t split_tiles, int64_t qblocks, int64_t ntasks, int64_t per_xcd,
    float *__restrict__ pdf_part, float *__restrict__ w_part) {
  using BL = LpBlock<PREC>;
  constexpr int P16 = CrossType<PREC>::kP16;  // cross-term operands
  constexpr bool X6 = BL::kX6;
  constexpr bool X3 = BL::kX3;
  constexpr bool SPLIT_PHI = BL::kSplitPhi;  // phi.X with phi = phi_h + phi_l
  constexpr bool DUP_PHI = BL::kOnePhi;
  // phi.X operands: BF16 for the split forms (phi in [0, 1] keeps full
  // relative precision down to tiny values), FP16 for the one-phi forms (phi
  // to 2^-12, X in FP16 pieces), else the form's own type
  constexpr int PX = SPLIT_PHI ? FSK_PREC_BF16 : DUP_PHI ? FSK_PREC_F16 : P16;
  constexpr int CH = BL::kChunk;
  constexpr int S1 = BL::kSlotU16;
  using RL = RecLayout<PREC>;
  __shared__ __attribute__((aligned(16))) unsigned short lds[2 * BL::kChunkU16];

  // task, split and the data range are wave-uniform: keep them scalar
  const int64_t task = __builtin_amdgcn_readfirstlane(int(xcd_task(blockIdx.x, per_xcd)));
  if (task >= ntasks) return;
  const int64_t split = __builtin_amdgcn_readfirstlane(int(task / qblocks));
  const int64_t qb = task - split * qblocks;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, c = lane & 15;
  const int64_t row0 = qb * (64 * QT) + int64_t(wave) * (16 * QT);

  u16x8 B1[QT], B2[QT], B3[QT], B4[QT];
  f4 w[QT];
  float pdf[QT];
  d4 w64[QT];
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
    pdf[R] = 0.f;
    w64[R] = d4{0.0, 0.0, 0.0, 0.0};
    pdf64[R] = 0.0;
  }
  // this task's data tiles [t_begin, t_end), copied to LDS CH tiles at a time
  const int64_t t_begin = split * split_tiles;
  int64_t t_end = t_begin + split_tiles;
  if (t_end > ntiles_data) t_end = ntiles_data;
  const int ntile = int(t_end - t_begin);
  const int nchunk = (ntile + CH - 1) / CH;
  const unsigned short *src0 = blk + t_begin * BL::kTileU16;
  // Wave `wave` copies slots wave, wave + 4, ... of a chunk: one 1 KiB slot per
  // instruction, lane l's 16 bytes landing at LDS slot base + 16 l.
  auto stage = [&](int ch, int buf) {
    const unsigned short *src = src0 + int64_t(ch) * BL::kChunkU16 + lane * 8;
#pragma unroll
    for (int q = 0; q < CH * BL::kSlots / 4; ++q) {
      const int k = 4 * q + wave;
      __builtin_amdgcn_global_load_lds(src + k * S1, (lds_void *)(lds + buf * BL::kChunkU16 + k * S1),
                                       16, 0, 0);
    }
  };

  stage(0, 0);
  int tc = 0;
  // (1, 1) in FP16, kept opaque so that the FP16 halves enter the FP32 row
  // sums through one mixed-precision FMA each instead of a convert and an add
  h16x2v one_h = {1.0, 1.0};
  asm volatile("" : "+v"(one_h));
  for (int ch = 0; ch < nchunk; ++ch) {
    // this wave's copies of chunk ch have landed; the barrier makes every
    // wave's copies visible and retires all reads of the buffer staged next
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (ch + 1 < nchunk) stage(ch + 1, (ch + 1) & 1);
    const unsigned short *cb = lds + (ch & 1) * BL::kChunkU16 + lane * 8;
    const int nt = ntile - ch * CH < CH ? ntile - ch * CH : CH;
    for (int t = 0; t < nt; ++t) {
      const unsigned short *tb = cb + t * BL::kTileU16;
      u16x8 A1, A2, A3, A4;
      if (X6) {
        A1 = lds_u16x8(tb);
        A2 = lds_u16x8(tb + S1);
        A3 = lds_u16x8(tb + 2 * S1);
        A4 = lds_u16x8(tb + 3 * S1);
      } else if (X3) {
        A1 = lds_u16x8(tb);
        A2 = lds_u16x8(tb + S1);
      } else {
        A1 = lds_u16x8(tb);
      }
      const u16x8 BX1 = lds_u16x8(tb + BL::kCross * S1);
      const u16x8 BX2 = SPLIT_PHI ? lds_u16x8(tb + (BL::kCross + 1) * S1) : BX1;

      // Row tile by row tile: cross term, phi, its row sums, phi.X.
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
          d = mfma16(A1, B1[R], f4{0.f, 0.f, 0.f, 0.f}, P16);
        }
        // Single FP16 form: the row sums add the FP16 phi that phi.X
        // multiplies, two at a time: one packed FP16 add of the lane's two
        // phi pairs, then both halves into the FP32 sum (v_fma_mix_f32),
        // 3 instructions instead of 4 FP32 adds. (The one-phi forms keep FP32
        // row sums: there the FP16 phi, subnormal below 2^-14, would raise
        // the pdf error of an FP32-class form about six-fold.)
        constexpr bool PAIR_SUM = !SPLIT_PHI && !DUP_PHI && PX == FSK_PREC_F16;
        float e[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          e[u] = fexp2_le1(d[u]);
          if (!PAIR_SUM) pdf[R] += e[u];
        }
        if constexpr (PAIR_SUM) {
          const h16x2v sp = __builtin_bit_cast(h16x2v, pack2<PX>(e[0], e[1])) +
                            __builtin_bit_cast(h16x2v, pack2<PX>(e[2], e[3]));
          pdf[R] = __builtin_fmaf(float(sp.x), float(one_h.x), pdf[R]);
          pdf[R] = __builtin_fmaf(float(sp.y), float(one_h.y), pdf[R]);
        }
        const u16x8 A = phi_operand<PX, SPLIT_PHI, DUP_PHI>(e);
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
  }
#pragma unroll
  for (int R = 0; R < QT; ++R) {
    const double p = sum_groups(pdf64[R] + pdf[R]);
    const int64_t i = row0 + 16 * R + c;
    if (g == 0 && i < m) pdf_part[split * m + i] = float(p);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int64_t iw = row0 + 16 * R + 4 * g + r;
      if (iw < m) w_part[(split * m + iw) * kD + c] = float(w64[R][r] + w[R][r]);
    }
  }
}

// pdf[i] = sum_s pdf_part; w[i,f] = sum_s w_part / scale + mu_f pdf[i]
// (FP64), n rows.
__global__ void score_lp_finish_kernel(const float *__restrict__ pdf_part,
                                       const float *__restrict__ w_part, int64_t S, int64_t n,
                                       const double *__restrict__ mu_sum, double inv_n,
                                       double inv_scale, float *__restrict__ pdf,
                                       float *__restrict__ w, int64_t ldw_row, int64_t ldw_col) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * kD) return;
  const int64_t i = e / kD;
  const int f = int(e % kD);
  double p = 0.0, s = 0.0;
  for (int64_t k = 0; k < S; ++k) {
    p += double(pdf_part[k * n + i]);
    s += double(w_part[(k * n + i) * kD + f]);
  }
  w[i * ldw_row + f * ldw_col] = float(s * inv_scale + mu_sum[f] * inv_n * p);
  if (f == 0) pdf[i] = float(p);
}

static int lp_tile_u16(unsigned prec) {
  switch (prec) {
    case FSK_PREC_BF16X6: return LpBlock<FSK_PREC_BF16X6>::kTileU16;
    case FSK_PREC_BF16X3: return LpBlock<FSK_PREC_BF16X3>::kTileU16;
    case FSK_PREC_F16X3: return LpBlock<FSK_PREC_F16X3>::kTileU16;
    case FSK_PREC_F16X3S: return LpBlock<FSK_PREC_F16X3S>::kTileU16;
    case FSK_PREC_BF16X3S: return LpBlock<FSK_PREC_BF16X3S>::kTileU16;
    default: return LpBlock<FSK_PREC_BF16>::kTileU16;
  }
}
static int lp_chunk(unsigned prec) {
  switch (prec) {
    case FSK_PREC_BF16X6: return LpBlock<FSK_PREC_BF16X6>::kChunk;
    case FSK_PREC_BF16X3: return LpBlock<FSK_PREC_BF16X3>::kChunk;
    case FSK_PREC_F16X3: return LpBlock<FSK_PREC_F16X3>::kChunk;
    case FSK_PREC_F16X3S: return LpBlock<FSK_PREC_F16X3S>::kChunk;
    case FSK_PREC_BF16X3S: return LpBlock<FSK_PREC_BF16X3S>::kChunk;
    default: return LpBlock<FSK_PREC_BF16>::kChunk;
  }
}

struct ScoreLpLayout {
  double *mu;
  unsigned short *xrec, *qrec, *xt, *blk;
  float *pdf_part, *w_part;
  int64_t ntiles_data, ntiles_alloc;
};

static size_t score_lp_layout(void *ws, int64_t n, int64_t m, unsigned prec, ScoreLpLayout *L) {
  Carve cv(ws);
  SplitPlan p = score16_lp_plan(n, m, prec);
  L->ntiles_data = cdiv(n, kTile);
  L->ntiles_alloc = L->ntiles_data + lp_chunk(prec);  // whole-chunk copies may read past the end
  L->mu = cv.take<double>(kMeanWs);
  L->xrec = cv.take<unsigned short>(round_up(n, kNormPad) * rec_size(prec));
  L->qrec = cv.take<unsigned short>(round_up(m, kNormPad) * rec_size(prec));
  L->xt = cv.take<unsigned short>(L->ntiles_data * kTile * kD * 2);
  L->blk = cv.take<unsigned short>(L->ntiles_alloc * lp_tile_u16(prec));
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
  const double scl = lowp_prescale(h);
  const double coef = exp2_coef(h) / (scl * scl);
  const int64_t xrows = round_up(n, kNormPad), qrows = round_up(m, kNormPad);
  const int64_t n16 = L.ntiles_data * kTile;
  const double inv_n = 1.0 / double(n);
  launch_mean16(data, ld, n, L.mu, s);
  SplitPlan p = score16_lp_plan(n, m, prec);
  const int64_t ntasks = p.qblocks * p.splits;
  const int64_t split_tiles = p.split_len / kTile;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
#define FSK_LP(P, PXT, QT)                                                                       \
  do {                                                                                           \
    hipLaunchKernelGGL(prep_lse_kernel<P>, dim3(cdiv(xrows, 256)), dim3(256), 0, s, data, ld, n, \
                       xrows, L.mu, inv_n, scl, coef, 0, (float *)nullptr, (float *)nullptr,      \
                       L.xrec);                                                                  \
    hipLaunchKernelGGL(prep_lse_kernel<P>, dim3(cdiv(qrows, 256)), dim3(256), 0, s, rows_in,     \
                       ldr, m, qrows, L.mu, inv_n, scl, coef, 1, (float *)nullptr,                \
                       (float *)nullptr,                                                         \
                       L.qrec);                                                                  \
    hipLaunchKernelGGL(prep_xt16_kernel<PXT>, dim3(cdiv(n16 * kD, 256)), dim3(256), 0, s, data,  \
                       ld, n, n16, L.mu, inv_n, scl, L.xt);                                      \
    hipLaunchKernelGGL(prep_lp_blocks_kernel<P>,                                                 \
                       dim3(cdiv(L.ntiles_alloc * LpBlock<P>::kSlots * 64, 256)), dim3(256), 0,  \
                       s, L.xrec, L.xt, L.ntiles_data, L.ntiles_alloc, L.blk);                    \
    prof_begin(FSK_OP_SCORE16_ORDERED, s);                                                       \
    hipLaunchKernelGGL((score16_lp_kernel<P, QT>), dim3(grid), dim3(kWG), 0, s, L.blk, L.qrec,   \
                       m, L.ntiles_data, split_tiles, p.qblocks, ntasks, per_xcd, L.pdf_part,     \
                       L.w_part);                                                                \
    prof_end(FSK_OP_SCORE16_ORDERED, s);                                                         \
  } while (0)
  switch (prec) {
    case FSK_PREC_BF16: FSK_LP(FSK_PREC_BF16, FSK_PREC_BF16, kScoreLpQT); break;
    case FSK_PREC_F16: FSK_LP(FSK_PREC_F16, FSK_PREC_F16, kScoreLpQT); break;
    case FSK_PREC_BF16X6: FSK_LP(FSK_PREC_BF16X6, FSK_PREC_BF16, kScoreLpQTx6); break;
    case FSK_PREC_F16X3: FSK_LP(FSK_PREC_F16X3, FSK_PREC_BF16, kScoreLpQT); break;
    case FSK_PREC_F16X3S: FSK_LP(FSK_PREC_F16X3S, FSK_PREC_F16, kScoreLpQT); break;
    case FSK_PREC_BF16X3S: FSK_LP(FSK_PREC_BF16X3S, FSK_PREC_F16, kScoreLpQT); break;
    default: FSK_LP(FSK_PREC_BF16X3, FSK_PREC_BF16, kScoreLpQT); break;
  }
#undef FSK_LP
  int st = launch_status(what);
  if (st != FSK_OK) return st;
  hipLaunchKernelGGL(score_lp_finish_kernel, dim3(cdiv(m * kD, 256)), dim3(256), 0, s,
                     L.pdf_part, L.w_part, p.splits, m, L.mu, inv_n, 1.0 / scl, pdf, w, ldw_row,
                     ldw_col);
  return launch_status(what);
}

}  // namespace fsk
