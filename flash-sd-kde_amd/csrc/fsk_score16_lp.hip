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
// so phi.X needs no data movement. The data tiles go in pairs (a, b), and the
// K slots of one phi.X MFMA take the products of both tiles wherever a form
// would otherwise leave half of them zero:
//   single forms : [phi_a | phi_b] . [Xh_a | Xh_b]                       1 MFMA / pair
//   split forms  : [phi_h,a | phi_l,a] . [Xh_a | Xh_a]
//                + [phi_h,b | phi_l,b] . [Xh_b | Xh_b]
//                + [phi_h,a | phi_h,b] . [Xl_a | Xl_b]                   3 MFMAs / pair
//                  (phi = phi_h + phi_l, phi_h = phi truncated to BF16,
//                   phi_l = rn_bf16(phi - phi_h); drops phi_l . Xl)
//   one-phi forms: [phi | phi] . [Xh | Xl] per tile (F16X3S, BF16X3S: phi
//                  rounded to FP16 once, X in FP16 pieces to ~22 bits)  2 MFMAs / pair
// 16x16x32 MFMAs cost as much issue time as several VALU instructions here
// (the MFMA, exp and VALU issue times add up in these kernels), so the pairing
// takes a quarter of the single forms' MFMA work away.
// Row sums pdf_i = sum_j phi_ij: FP32 VALU adds of the unrounded phi, except
// in the single forms, which add the 16-bit phi that phi.X multiplies, two per
// v_dot2_f32_f16 / v_dot2_f32_bf16 against (1, 1) (FP32 accumulation, products
// exact): half the VALU instructions, and an error far below the 16-bit cross
// term's.
//
// Data path. These kernels need several 16-byte operands per lane per data
// tile (the FP32 kernel needs one), so fetching them per wave through the
// vector-memory path is what limits them. Each pair of data tiles is therefore
// stored once, by a prep pass, as ready-made MFMA operand slots (per slot 64
// lanes x 16 bytes, in lane order), and each workgroup copies chunks of pairs
// into a double-buffered LDS area with direct global-to-LDS loads
// (global_load_lds_dwordx4, one 1 KiB slot per wave instruction); the four
// waves then read their operands with ds_read_b128. Global and L2 traffic per
// workgroup drop four-fold (one copy for four waves) and every access is a
// full, aligned 1 KiB.
//
// Points are centred on the data mean: w' = sum phi (x - mu); the finish
// kernel returns w = w' + mu * pdf in FP64.
//
// Negligible tiles. In the forms whose phi.X operand is phi rounded to FP16
// (F16, F16X3S, BF16X3S), every phi <= 2^-26 becomes exactly 0 (below half
// of FP16's smallest subnormal). The SKIP instances test a wave's cross terms
// against a pair of data tiles (F16) or a tile (the others) with one ballot
// and leave out the exps, conversions, row sums and the phi.X MFMAs when every
// exponent of the wave's row tiles against the tile (pair) is <= skip_t:
//   F16          skip_t = -26: row sums and phi.X both use the FP16 phi, so
//                the results are bitwise those of the dense kernel;
//   F16X3S,      phi.X as above (the centred sums bitwise); the row sums
//   BF16X3S      add the unrounded phi, so skip_t = -(ceil(log2 n) + 28)
//                bounds what a row loses by n 2^skip_t <= 2^-28 of its self
//                term (rows = data only; other rows run dense).
// In data put in spatial order, most tile pairs of a large set are far apart
// and skip.
//
// Far pairs before any MFMA: a bitmap from box bounds along the data's
// principal axes (fsk_far.h) marks the (wave's 64 rows, tile pair) blocks
// whose every exponent is provably <= skip_t; the kernel leaves them out
// without reading their operands.
#include "fsk_common.h"
#include "fsk_far.h"
#include "fsk_lowp.h"
#include "fsk_plan.h"

namespace fsk {

static_assert(kNormPad % (128 * kScoreLpQT) == 0 && kNormPad % (64 * kScoreLpQTx6) == 0,
              "records must cover whole row blocks");
static_assert(kNormPad % kLpPairPts == 0, "records must cover whole tile pairs");

typedef __attribute__((address_space(3))) void lds_void;

// Operand slots of one pair of 16-point data tiles (a = 2P, b = 2P + 1):
// the cross-term A operands of tile a, then of tile b, then the phi.X B
// operands of the pair.
template <int PREC>
struct LpPair {
  static constexpr bool kX6 = PREC == FSK_PREC_BF16X6;
  static constexpr bool kOnePhi = PREC == FSK_PREC_F16X3S || PREC == FSK_PREC_BF16X3S;
  static constexpr bool kF16Phi = kOnePhi || PREC == FSK_PREC_F16;  // phi.X takes FP16 phi
  static constexpr bool kX3 = PREC == FSK_PREC_BF16X3 || PREC == FSK_PREC_F16X3 || kOnePhi;
  static constexpr bool kSplitPhi = kX6 || (kX3 && !kOnePhi);
  static constexpr int kCross = kX6 ? 4 : kX3 ? 2 : 1;  // cross-term A operands per tile
  // split: [Xh_a | Xh_a], [Xh_b | Xh_b], [Xl_a | Xl_b]; one-phi: [Xh_a | Xl_a],
  // [Xh_b | Xl_b]; single: [Xh_a | Xh_b]
  static constexpr int kX = kSplitPhi ? 3 : kOnePhi ? 2 : 1;
  static constexpr int kSlots = 2 * kCross + kX;
  static constexpr int kSlotU16 = 64 * 8;  // 16 bytes per lane
  static constexpr int kPairU16 = kSlots * kSlotU16;
  // pairs per LDS chunk: up to 16 KiB per buffer
  static constexpr int kChunk = 8192 / kPairU16 > 0 ? 8192 / kPairU16 : 1;
  static constexpr int kChunkU16 = kChunk * kPairU16;
  static constexpr int kChunkSlots = kChunk * kSlots;
  // Waves per workgroup sharing each chunk, and LDS buffers (chunks staged
  // kBufs - 1 ahead). The forms that leave out negligible tiles spend a few
  // hundred cycles per staged pair where they skip, against the chunk's
  // L2 / MALL latency of a microsecond or more: eight waves share a chunk
  // and it is fetched two chunks ahead.
  static constexpr int kWaves = score_lp_waves(PREC);
  static constexpr int kBufs = kF16Phi ? 3 : 2;
};

// Record halves (lo, hi) of cross-term A operand `s` of a form.
template <int PREC>
__device__ __forceinline__ int cross_slot_off(int s, int g) {
  using RL = RecLayout<PREC>;
  if (LpPair<PREC>::kX6) {
    const int lo = s == 0 ? RL::kP1 : RL::kP2;
    const int hi = s == 0 ? RL::kP1 : s == 1 ? RL::kP1 : s == 2 ? RL::kP3 : RL::kFold;
    return rec_off(g, lo, hi);
  }
  if (LpPair<PREC>::kX3) return rec_off(g, RL::kP1, s == 0 ? RL::kP2 : RL::kFold);
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

// Operand slots of every tile pair from the records and the transposed X;
// pairs [npairs_data, npairs_alloc) are the zero slack that whole-chunk copies
// may read past the last pair. The records and xt cover 32 npairs_data points
// (rows past n are padding: zero X, a norm term that drives phi to 0).
template <int PREC>
__global__ void prep_lp_pairs_kernel(const unsigned short *__restrict__ xrec,
                                     const unsigned short *__restrict__ xt, int64_t npairs_data,
                                     int64_t npairs_alloc, unsigned short *__restrict__ blk) {
  using B = LpPair<PREC>;
  using RL = RecLayout<PREC>;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // (pair, slot, lane)
  if (e >= npairs_alloc * B::kSlots * 64) return;
  const int lane = int(e % 64);
  const int slot = int((e / 64) % B::kSlots);
  const int64_t P = e / (64 * B::kSlots);
  const int g = lane >> 4, c = lane & 15;
  u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
  if (P < npairs_data) {
    if (slot < 2 * B::kCross) {
      const int64_t T = 2 * P + slot / B::kCross;
      v = ld_u16x8(xrec + (T * kTile + c) * RL::kSize + cross_slot_off<PREC>(slot % B::kCross, g));
    } else {
      const int xs = slot - 2 * B::kCross;
      // [Xh | Xl] of feature c, lane group g, for tiles a and b
      const u16x8 ta = ld_u16x8(xt + (((2 * P) * kD + c) * 4 + g) * 8);
      const u16x8 tb = ld_u16x8(xt + (((2 * P + 1) * kD + c) * 4 + g) * 8);
      if (B::kSplitPhi) {
        if (xs == 0) v = u16x8{ta[0], ta[1], ta[2], ta[3], ta[0], ta[1], ta[2], ta[3]};
        else if (xs == 1) v = u16x8{tb[0], tb[1], tb[2], tb[3], tb[0], tb[1], tb[2], tb[3]};
        else v = u16x8{ta[4], ta[5], ta[6], ta[7], tb[4], tb[5], tb[6], tb[7]};
      } else if (B::kOnePhi) {
        v = xs == 0 ? ta : tb;
      } else {
        v = u16x8{ta[0], ta[1], ta[2], ta[3], tb[0], tb[1], tb[2], tb[3]};
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
// v_cvt_pk_bf16_f32 / v_cvt_pk_f16_f32.
template <int P16>
__device__ __forceinline__ unsigned pack2(float a, float b) {
  const f2v v = {a, b};
  if (P16 == FSK_PREC_F16) return __builtin_bit_cast(unsigned, __builtin_convertvector(v, h16x2v));
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2v));
}

// BF16 split of a lane's four phi values: hi = the two packed words of phi
// truncated to its upper 16 bits (a BF16 value, exact as FP32: one byte
// permute per two values), lo = the two packed words of rn_bf16(phi - hi)
// (the difference is exact in FP32).
__device__ __forceinline__ void split_phi(const float (&e)[4], unsigned (&hi)[2],
                                          unsigned (&lo)[2]) {
  unsigned b[4];
  float l[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    b[u] = __builtin_bit_cast(unsigned, e[u]);
    l[u] = e[u] - __builtin_bit_cast(float, b[u] & 0xffff0000u);
  }
  hi[0] = __builtin_amdgcn_perm(b[1], b[0], 0x07060302u);
  hi[1] = __builtin_amdgcn_perm(b[3], b[2], 0x07060302u);
  lo[0] = pack2<FSK_PREC_BF16>(l[0], l[1]);
  lo[1] = pack2<FSK_PREC_BF16>(l[2], l[3]);
}

// acc + (the two 16-bit values packed in p): one v_dot2_f32_f16 / _bf16
// against (1, 1); the products are exact, the sum FP32.
template <int P16>
__device__ __forceinline__ float dot2_ones(unsigned p, float acc) {
  if (P16 == FSK_PREC_F16) {
    const h16x2v one = {1.0, 1.0};
    return __builtin_amdgcn_fdot2(__builtin_bit_cast(h16x2v, p), one, acc, false);
  }
  const bf16x2v one = {1.0, 1.0};
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2v, p), one, acc, false);
}

__device__ __forceinline__ u16x8 words(unsigned a, unsigned b, unsigned c, unsigned d) {
  return __builtin_bit_cast(u16x8, u32x4{a, b, c, d});
}

__device__ __forceinline__ u16x8 lds_u16x8(const unsigned short *p) {
  return *reinterpret_cast<const u16x8 *>(p);
}

// Cross term of one data tile (A operands from LDS) against one row tile.
template <int PREC>
__device__ __forceinline__ f4 lp_cross(const u16x8 (&A)[LpPair<PREC>::kCross], u16x8 b1,
                                       u16x8 b2, u16x8 b3, u16x8 b4) {
  constexpr int P16 = CrossType<PREC>::kP16;
  f4 d = mfma16(A[0], b1, f4{0.f, 0.f, 0.f, 0.f}, P16);
  if (LpPair<PREC>::kCross >= 2) d = mfma16(A[1 % LpPair<PREC>::kCross], b2, d, P16);
  if (LpPair<PREC>::kCross == 4) {
    d = mfma16(A[2 % LpPair<PREC>::kCross], b3, d, P16);
    d = mfma16(A[3 % LpPair<PREC>::kCross], b4, d, P16);
  }
  return d;
}

// Waves per SIMD the compiler must leave room for: 4 (at most 128 VGPRs) for
// every form but BF16X6, whose four query operands per row tile need more and
// which runs best at 3 (at most 168 VGPRs, a few spilled outside the loop:
// n = 2^20 at 2 / 3 / 4 waves took 327 / 302 / 401 ms).
template <int PREC>
constexpr int lp_min_waves() { return PREC == FSK_PREC_BF16X6 ? 3 : 4; }

// s_waitcnt vmcnt(n) for n = 0..3, expcnt and lgkmcnt not waited on (gfx9
// encoding: vmcnt in bits 3:0, expcnt 6:4, lgkmcnt 11:8).
__device__ __forceinline__ void wait_vmcnt(int n) {
  if (n <= 0) __builtin_amdgcn_s_waitcnt(0xF70);
  else if (n == 1) __builtin_amdgcn_s_waitcnt(0xF71);
  else if (n == 2) __builtin_amdgcn_s_waitcnt(0xF72);
  else __builtin_amdgcn_s_waitcnt(0xF73);
}

// Largest of a lane's cross-term exponents.
__device__ __forceinline__ float max4(f4 d) {
  return fmaxf(fmaxf(d[0], d[1]), fmaxf(d[2], d[3]));
}

template <int PREC, int QT, bool SKIP>
__global__ __launch_bounds__(64 * LpPair<PREC>::kWaves)
__attribute__((amdgpu_waves_per_eu(lp_min_waves<PREC>())))
void score16_lp_kernel(
    const unsigned short *__restrict__ blk, const unsigned short *__restrict__ qrec, int64_t m,
    int64_t npairs_data, int64_t split_pairs, int64_t qblocks, int64_t ntasks, int64_t per_xcd,
    float skip_t, const unsigned long long *__restrict__ far_bits, int64_t far_words,
    float *__restrict__ pdf_part, float *__restrict__ w_part) {
  using BL = LpPair<PREC>;
  static_assert(!SKIP || BL::kF16Phi, "tiles are skipped only where phi.X takes FP16 phi");
  constexpr bool X6 = BL::kX6;
  constexpr bool X3 = BL::kX3;
  constexpr bool SPLIT_PHI = BL::kSplitPhi;  // phi.X with phi = phi_h + phi_l
  constexpr bool ONE_PHI = BL::kOnePhi;
  // phi.X operands: BF16 for the split forms (phi in [0, 1] keeps full
  // relative precision down to tiny values), FP16 for the one-phi forms (phi
  // to 2^-12, X in FP16 pieces), else the form's own type
  constexpr int PX = SPLIT_PHI ? FSK_PREC_BF16 : ONE_PHI ? FSK_PREC_F16 : CrossType<PREC>::kP16;
  constexpr bool SINGLE = !SPLIT_PHI && !ONE_PHI;
  // tile pairs between FP64 flushes of the FP32 sums: the single forms' own
  // error (16-bit cross term and phi, ~1e-2 relative) dwarfs FP32 rounding
  // over 2048 terms
  constexpr int FLUSH = SINGLE ? 64 : 16;
  constexpr int KC = BL::kCross;
  constexpr int CH = BL::kChunk;
  constexpr int S1 = BL::kSlotU16;
  constexpr int W = BL::kWaves;
  constexpr int NB = BL::kBufs;
  static_assert(BL::kChunkSlots <= 4 * W, "at most 3 newer loads per wave are waited past");
  using RL = RecLayout<PREC>;
  __shared__ __attribute__((aligned(16))) unsigned short lds[NB * BL::kChunkU16];

  // task, split and the data range are wave-uniform: keep them scalar
  // Dense instances: each XCD takes a contiguous range of tasks (row blocks
  // of one data split share its L2). SKIP instances: task = workgroup index,
  // so the round-robin dispatch gives each XCD every eighth row block; the
  // work of a row block depends on how many tile pairs lie near it, and
  // contiguous ranges of spatially ordered rows left some XCDs with most of
  // the near pairs.
  const int64_t task = __builtin_amdgcn_readfirstlane(
      int(SKIP ? int64_t(blockIdx.x) : xcd_task(blockIdx.x, per_xcd)));
  if (task >= ntasks) return;
  const int64_t split = __builtin_amdgcn_readfirstlane(int(task / qblocks));
  const int64_t qb = task - split * qblocks;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, c = lane & 15;
  const int64_t row0 = qb * (16 * W * QT) + int64_t(wave) * (16 * QT);

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
      B3[R] = B4[R] = B2[R];
    } else {
      B1[R] = ld_u16x8(rq + rec_off(g, RL::kP1, RL::kFold));
      B2[R] = B3[R] = B4[R] = B1[R];
    }
    w[R] = f4{0.f, 0.f, 0.f, 0.f};
    pdf[R] = 0.f;
    w64[R] = d4{0.0, 0.0, 0.0, 0.0};
    pdf64[R] = 0.0;
  }
  // this task's tile pairs [p_begin, p_end), copied to LDS CH pairs at a time
  const int64_t p_begin = split * split_pairs;
  int64_t p_end = p_begin + split_pairs;
  if (p_end > npairs_data) p_end = npairs_data;
  const int npair = int(p_end - p_begin);
  const int nchunk = (npair + CH - 1) / CH;
  const unsigned short *src0 = blk + p_begin * BL::kPairU16;
  // Wave `wave` copies slots wave, wave + W, ... of a chunk: one 1 KiB slot per
  // instruction, lane l's 16 bytes landing at LDS slot base + 16 l.
  const int my_loads = (BL::kChunkSlots - wave + W - 1) / W;  // per chunk, this wave
  auto stage = [&](int ch, int buf) {
    const unsigned short *src = src0 + int64_t(ch) * BL::kChunkU16 + lane * 8;
#pragma unroll
    for (int q = 0; q < (BL::kChunkSlots + W - 1) / W; ++q) {
      const int k = W * q + wave;
      if (k < BL::kChunkSlots)
        __builtin_amdgcn_global_load_lds(src + k * S1,
                                         (lds_void *)(lds + buf * BL::kChunkU16 + k * S1), 16, 0, 0);
    }
  };

  // this wave's row of the far-block bitmap and the workgroup's (far for all
  // of its waves), read 64 pairs at a time
  const unsigned long long *far_row =
      SKIP && far_bits ? far_bits + (qb * W + wave) * far_words : nullptr;
  const unsigned long long *wg_row =
      SKIP && far_bits ? far_bits + (qblocks * W + qb) * far_words : nullptr;
  int64_t widx = -1, gidx = -1;
  unsigned long long fword = 0, gword = 0;
  // Bits of pairs [P, P + cnt) (cnt <= 64) of a bitmap row, bit t for pair
  // P + t; idx / word cache the row's last word read (rows are walked in
  // increasing order).
  auto bits_at = [&](const unsigned long long *row, int64_t P, int cnt, int64_t &idx,
                     unsigned long long &word) -> unsigned long long {
    const int64_t w0 = P >> 6;
    const int sh = int(P & 63);
    if (w0 != idx) {
      idx = w0;
      word = sload_u64(row + w0);
    }
    unsigned long long b = word >> sh;
    if (sh + cnt > 64) {
      idx = w0 + 1;
      word = sload_u64(row + w0 + 1);
      b |= word << (64 - sh);
    }
    return cnt == 64 ? b : b & ((1ull << cnt) - 1);
  };
  auto flush = [&]() {
#pragma unroll
    for (int R = 0; R < QT; ++R) {
      pdf64[R] += pdf[R];
      pdf[R] = 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) w64[R][r] += w[R][r];
      w[R] = f4{0.f, 0.f, 0.f, 0.f};
    }
  };
  // the computation of one tile pair, read from LDS at pb
  auto pair_body = [&](const unsigned short *pb) {
    if constexpr (SINGLE) {
      // single forms: both tiles at once, one phi.X MFMA [phi_a | phi_b]
      u16x8 Aa[KC], Ab[KC];
      Aa[0] = lds_u16x8(pb);
      Ab[0] = lds_u16x8(pb + S1);
      const u16x8 BX = lds_u16x8(pb + 2 * S1);
      if constexpr (SKIP) {
        // all cross terms, then one ballot for the wave's row tiles: in
        // spatially ordered data a tile pair is negligible for nearly all
        // of a wave's 64 consecutive rows or for none
        f4 da[QT], db[QT];
#pragma unroll
        for (int R = 0; R < QT; ++R) {
          da[R] = lp_cross<PREC>(Aa, B1[R], B2[R], B3[R], B4[R]);
          db[R] = lp_cross<PREC>(Ab, B1[R], B2[R], B3[R], B4[R]);
        }
        float mx = fmaxf(max4(da[0]), max4(db[0]));
#pragma unroll
        for (int R = 1; R < QT; ++R) mx = fmaxf(mx, fmaxf(max4(da[R]), max4(db[R])));
        if (__builtin_amdgcn_ballot_w64(mx > skip_t)) {
#pragma unroll
          for (int R = 0; R < QT; ++R) {
            float ea[4], eb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              ea[u] = fexp2_le1(da[R][u]);
              eb[u] = fexp2_le1(db[R][u]);
            }
            const unsigned a0 = pack2<PX>(ea[0], ea[1]), a1 = pack2<PX>(ea[2], ea[3]);
            const unsigned b0 = pack2<PX>(eb[0], eb[1]), b1 = pack2<PX>(eb[2], eb[3]);
            pdf[R] = dot2_ones<PX>(a0, pdf[R]);
            pdf[R] = dot2_ones<PX>(a1, pdf[R]);
            pdf[R] = dot2_ones<PX>(b0, pdf[R]);
            pdf[R] = dot2_ones<PX>(b1, pdf[R]);
            w[R] = mfma16(words(a0, a1, b0, b1), BX, w[R], PX);
          }
        }
      } else {
#pragma unroll
        for (int R = 0; R < QT; ++R) {
          const f4 da = lp_cross<PREC>(Aa, B1[R], B2[R], B3[R], B4[R]);
          const f4 db = lp_cross<PREC>(Ab, B1[R], B2[R], B3[R], B4[R]);
          float ea[4], eb[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            ea[u] = fexp2_le1(da[u]);
            eb[u] = fexp2_le1(db[u]);
          }
          const unsigned a0 = pack2<PX>(ea[0], ea[1]), a1 = pack2<PX>(ea[2], ea[3]);
          const unsigned b0 = pack2<PX>(eb[0], eb[1]), b1 = pack2<PX>(eb[2], eb[3]);
          pdf[R] = dot2_ones<PX>(a0, pdf[R]);
          pdf[R] = dot2_ones<PX>(a1, pdf[R]);
          pdf[R] = dot2_ones<PX>(b0, pdf[R]);
          pdf[R] = dot2_ones<PX>(b1, pdf[R]);
          w[R] = mfma16(words(a0, a1, b0, b1), BX, w[R], PX);
        }
        // Issue order of one pair (3 LDS reads, 12 MFMAs, 32 exps, 32 other
        // VALU ops per wave): the compiler's own order clusters the MFMAs and
        // the exps; interleaving 3 exps and 2 VALU ops per MFMA keeps the
        // matrix core busy under the transcendental work (FP16 score at
        // n = 2^20: 125 -> 119 ms; other mixes measured 119-123 ms).
        __builtin_amdgcn_sched_group_barrier(0x100, 3, 0);  // LDS reads
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);  // MFMA
#pragma unroll
        for (int gi = 0; gi < 10; ++gi) {
          __builtin_amdgcn_sched_group_barrier(0x400, 3, 0);  // transcendental
          __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // VALU
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        }
        __builtin_amdgcn_sched_group_barrier(0x002, 64, 0);
      }
    } else {
      // split and one-phi forms: tile a, then tile b (one tile's operands
      // live at a time); the split forms keep phi_h of tile a per row tile
      // for the shared [phi_h,a | phi_h,b] . [Xl_a | Xl_b] product
      unsigned ha[QT][2];
#pragma unroll
      for (int tt = 0; tt < 2; ++tt) {
        const unsigned short *tb = pb + tt * KC * S1;
        u16x8 A[KC];
#pragma unroll
        for (int s = 0; s < KC; ++s) A[s] = lds_u16x8(tb + s * S1);
        const u16x8 BX = lds_u16x8(pb + (2 * KC + tt) * S1);
        const u16x8 BXL = SPLIT_PHI && tt == 1 ? lds_u16x8(pb + (2 * KC + 2) * S1) : BX;
        if constexpr (SKIP) {
          // one ballot per tile for the wave's row tiles, as above
          f4 d[QT];
#pragma unroll
          for (int R = 0; R < QT; ++R) d[R] = lp_cross<PREC>(A, B1[R], B2[R], B3[R], B4[R]);
          float mx = max4(d[0]);
#pragma unroll
          for (int R = 1; R < QT; ++R) mx = fmaxf(mx, max4(d[R]));
          if (__builtin_amdgcn_ballot_w64(mx > skip_t)) {
#pragma unroll
            for (int R = 0; R < QT; ++R) {
              float e[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                e[u] = fexp2_le1(d[R][u]);
                pdf[R] += e[u];
              }
              const unsigned p0 = pack2<PX>(e[0], e[1]), p1 = pack2<PX>(e[2], e[3]);
              w[R] = mfma16(words(p0, p1, p0, p1), BX, w[R], PX);
            }
          }
          continue;
        }
#pragma unroll
        for (int R = 0; R < QT; ++R) {
          const f4 d = lp_cross<PREC>(A, B1[R], B2[R], B3[R], B4[R]);
          float e[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            e[u] = fexp2_le1(d[u]);
            pdf[R] += e[u];
          }
          if constexpr (SPLIT_PHI) {
            unsigned hi[2], lo[2];
            split_phi(e, hi, lo);
            w[R] = mfma16(words(hi[0], hi[1], lo[0], lo[1]), BX, w[R], PX);
            if (tt == 0) {
              ha[R][0] = hi[0];
              ha[R][1] = hi[1];
            } else {
              w[R] = mfma16(words(ha[R][0], ha[R][1], hi[0], hi[1]), BXL, w[R], PX);
            }
          } else {
            const unsigned p0 = pack2<PX>(e[0], e[1]), p1 = pack2<PX>(e[2], e[3]);
            w[R] = mfma16(words(p0, p1, p0, p1), BX, w[R], PX);
          }
        }
      }
    }
  };
  int tc = 0;
  // every pair of chunk ch, from LDS buffer buf (dense instances)
  auto run_chunk = [&](int ch, int buf) {
    const unsigned short *cb = lds + buf * BL::kChunkU16 + lane * 8;
    const int np = npair - ch * CH < CH ? npair - ch * CH : CH;
    for (int t = 0; t < np; ++t) {
      pair_body(cb + t * BL::kPairU16);
      if (++tc == FLUSH) {
        tc = 0;
        flush();
      }
    }
  };
  // the pairs of chunk ch this wave needs (SKIP instances): the bitmap's far
  // pairs add exact zeros to phi.X and at most n 2^skip_t to a row sum, so
  // the walk visits only the set bits of the near mask. The FP32 sums are
  // flushed before the first pair past each flush point, which keeps them
  // bitwise those of the dense order (nothing is added in between).
  int64_t next_flush = FLUSH;
  auto run_chunk_near = [&](int ch, int buf) {
    const unsigned short *cb = lds + buf * BL::kChunkU16 + lane * 8;
    const int np = npair - ch * CH < CH ? npair - ch * CH : CH;
    const int64_t q0 = int64_t(ch) * CH;  // pair index within the task
    unsigned long long near = np == 64 ? ~0ull : (1ull << np) - 1;
    if (far_row) near &= ~bits_at(far_row, p_begin + q0, np, widx, fword);
    while (near) {
      const int t = __builtin_ctzll(near);
      near &= near - 1;
      const int64_t q = q0 + t;
      if (q >= next_flush) {
        flush();
        next_flush = (q / FLUSH + 1) * FLUSH;
      }
      pair_body(cb + t * BL::kPairU16);
    }
  };
  if constexpr (SKIP) {
    // Copies run NB - 1 needed chunks ahead: a stretch of chunks nobody needs
    // costs a scan of their bits, not a copy latency. ring[b] is the chunk
    // buffer b holds (nchunk: none).
    int scan = 0;
    auto next_needed = [&]() -> int {
      if (!wg_row) return scan < nchunk ? scan++ : nchunk;
      // A chunk is needed unless every pair of it is far for every wave of
      // the workgroup; every wave reads the same words, so all agree, and
      // chunks nobody needs are neither copied nor waited for. The first
      // pair at or after chunk `scan` that some wave needs is found a 64-pair
      // word at a time.
      int64_t P = p_begin + int64_t(scan) * CH;
      while (P < p_end) {
        const int64_t w = P >> 6;
        if (w != gidx) {
          gidx = w;
          gword = sload_u64(wg_row + w);
        }
        const unsigned long long near = ~gword >> (P & 63);
        if (near) {
          P += __builtin_ctzll(near);
          break;
        }
        P = (w + 1) << 6;
      }
      if (P >= p_end) {
        scan = nchunk;
        return nchunk;
      }
      const int c = int((P - p_begin) / CH);
      scan = c + 1;
      return c;
    };
    int ring[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) ring[b] = nchunk;
#pragma unroll
    for (int b = 0; b < NB - 1; ++b) {
      ring[b] = next_needed();
      if (ring[b] < nchunk) stage(ring[b], b);
    }
    for (int k = 0;; ++k) {
      const int ch = ring[k % NB];
      if (ch >= nchunk) break;
      // This wave's copies of chunk ch have landed once at most those of the
      // chunks staged after it are outstanding (loads retire in order); the
      // barrier makes every wave's copies visible, and every wave has finished
      // reading the buffer staged next. A plain s_barrier: a fenced one would
      // also wait for the chunks in flight.
      int newer = 0;
#pragma unroll
      for (int j = 1; j <= NB - 2; ++j)
        if (ring[(k + j) % NB] < nchunk) newer += my_loads;
      wait_vmcnt(newer);
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();
      __asm__ volatile("" ::: "memory");
      {
        const int bn = (k + NB - 1) % NB;
        ring[bn] = next_needed();
        if (ring[bn] < nchunk) stage(ring[bn], bn);
      }
      run_chunk_near(ch, k % NB);
    }
  } else {
    // every chunk in order, copied NB - 1 ahead
    for (int c = 0; c < NB - 1 && c < nchunk; ++c) stage(c, c);
    for (int ch = 0; ch < nchunk; ++ch) {
      // this wave's copies of chunk ch have landed once at most those of the
      // chunks staged after it are outstanding; the barrier makes every
      // wave's copies visible and retires all reads of the buffer staged next
      const int newer = nchunk - 1 - ch < NB - 2 ? nchunk - 1 - ch : NB - 2;
      wait_vmcnt(newer * my_loads);
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();
      __asm__ volatile("" ::: "memory");
      if (ch + NB - 1 < nchunk) stage(ch + NB - 1, (ch + NB - 1) % NB);
      run_chunk(ch, ch % NB);
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

static int lp_pair_u16(unsigned prec) {
  switch (prec) {
    case FSK_PREC_BF16X6: return LpPair<FSK_PREC_BF16X6>::kPairU16;
    case FSK_PREC_BF16X3: return LpPair<FSK_PREC_BF16X3>::kPairU16;
    case FSK_PREC_F16X3: return LpPair<FSK_PREC_F16X3>::kPairU16;
    case FSK_PREC_F16X3S: return LpPair<FSK_PREC_F16X3S>::kPairU16;
    case FSK_PREC_BF16X3S: return LpPair<FSK_PREC_BF16X3S>::kPairU16;
    default: return LpPair<FSK_PREC_BF16>::kPairU16;
  }
}
static int lp_chunk(unsigned prec) {
  switch (prec) {
    case FSK_PREC_BF16X6: return LpPair<FSK_PREC_BF16X6>::kChunk;
    case FSK_PREC_BF16X3: return LpPair<FSK_PREC_BF16X3>::kChunk;
    case FSK_PREC_F16X3: return LpPair<FSK_PREC_F16X3>::kChunk;
    case FSK_PREC_F16X3S: return LpPair<FSK_PREC_F16X3S>::kChunk;
    case FSK_PREC_BF16X3S: return LpPair<FSK_PREC_BF16X3S>::kChunk;
    default: return LpPair<FSK_PREC_BF16>::kChunk;
  }
}

struct ScoreLpLayout {
  double *mu;
  unsigned short *xrec, *qrec, *xt, *blk;
  float *pdf_part, *w_part;
  int64_t npairs_data, npairs_alloc;
  // far-block bitmap (forms with FP16 phi, when within kFarBitsCap)
  float *frame, *pair_boxes, *row_boxes;
  unsigned long long *far_bits;
  int64_t row_blocks, far_words;
};

static size_t score_lp_layout(void *ws, int64_t n, int64_t m, unsigned prec, ScoreLpLayout *L) {
  Carve cv(ws);
  SplitPlan p = score16_lp_plan(n, m, prec);
  L->npairs_data = cdiv(n, kLpPairPts);
  L->npairs_alloc = L->npairs_data + lp_chunk(prec);  // whole-chunk copies may read past the end
  L->mu = cv.take<double>(kMeanWs);
  L->xrec = cv.take<unsigned short>(round_up(n, kNormPad) * rec_size(prec));
  L->qrec = cv.take<unsigned short>(round_up(m, kNormPad) * rec_size(prec));
  L->xt = cv.take<unsigned short>(L->npairs_data * kLpPairPts * kD * 2);
  L->blk = cv.take<unsigned short>(L->npairs_alloc * lp_pair_u16(prec));
  L->pdf_part = cv.take<float>(p.splits * m);
  L->w_part = cv.take<float>(p.splits * m * kD);
  L->row_blocks = p.qblocks * score_lp_waves(prec);
  L->far_words = cdiv(L->npairs_data, 64);
  static_assert(kFarRows == 16 * kScoreLpQT, "a bitmap row holds one wave's rows");
  const bool bits = (prec == FSK_PREC_F16 || prec == FSK_PREC_F16X3S || prec == FSK_PREC_BF16X3S) &&
                    far_bits_count(p.qblocks, score_lp_waves(prec), L->npairs_data) * 8 <=
                        kFarBitsCap;
  L->frame = L->pair_boxes = L->row_boxes = nullptr;
  L->far_bits = nullptr;
  if (bits) {
    L->frame = cv.take<float>(kBoxAxes * kD);
    L->pair_boxes = cv.take<float>(L->npairs_data * kBoxFloats);
    L->row_boxes = cv.take<float>(L->row_blocks * kBoxFloats);
    // the waves' rows, then one row per workgroup (the AND of its waves')
    L->far_bits = cv.take<unsigned long long>(
        far_bits_count(p.qblocks, score_lp_waves(prec), L->npairs_data));
  } else {
    L->row_blocks = L->far_words = 0;
  }
  return cv.used;
}

size_t score16_lp_ws_bytes(int64_t n, int64_t m, unsigned prec) {
  ScoreLpLayout L;
  return score_lp_layout(nullptr, n, m, prec, &L);
}

// The score kernel of form P, with the tile test where the form has one and
// `skip` asks for it.
template <int P, int QT>
static void launch_score_lp(bool skip, float skip_t, const unsigned long long *far_bits,
                            int64_t grid, hipStream_t s, const ScoreLpLayout &L, int64_t m,
                            int64_t split_pairs, int64_t qblocks, int64_t ntasks,
                            int64_t per_xcd) {
  if constexpr (LpPair<P>::kF16Phi) {
    if (skip) {
      hipLaunchKernelGGL((score16_lp_kernel<P, QT, true>), dim3(grid), dim3(64 * LpPair<P>::kWaves),
                         0, s, L.blk, L.qrec, m, L.npairs_data, split_pairs, qblocks, ntasks,
                         per_xcd, skip_t, far_bits, L.far_words, L.pdf_part, L.w_part);
      return;
    }
  }
  hipLaunchKernelGGL((score16_lp_kernel<P, QT, false>), dim3(grid), dim3(64 * LpPair<P>::kWaves),
                     0, s, L.blk, L.qrec, m, L.npairs_data, split_pairs, qblocks, ntasks, per_xcd,
                     0.f, (const unsigned long long *)nullptr, int64_t(0), L.pdf_part, L.w_part);
}

// Score accumulators of m rows against the KDE of n data points with 16-bit
// MFMA operands. self_rows: the rows are the data (the empirical score of the
// data set), so every row holds its self term; dense: evaluate every tile.
int score16_lp_run(const char *what, const float *rows_in, int64_t ldr, int64_t m,
                   const float *data, int64_t ld, int64_t n, double h, unsigned prec,
                   bool self_rows, bool dense, void *ws, size_t ws_bytes, float *pdf, float *w,
                   int64_t ldw_row, int64_t ldw_col, void *stream) {
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
  const int64_t n32 = L.npairs_data * kLpPairPts;
  const double inv_n = 1.0 / double(n);
  launch_mean16(data, ld, n, L.mu, s);
  SplitPlan p = score16_lp_plan(n, m, prec);
  const int64_t ntasks = p.qblocks * p.splits;
  const int64_t split_pairs = p.split_len / kLpPairPts;
  int64_t per_xcd;
  const int64_t grid = xcd_grid(ntasks, &per_xcd);
  // tile test threshold (exponents in log2 units): FP16 phi is exactly 0 at
  // or below 2^-26; the unrounded row sums of the one-phi forms lose at most
  // n 2^skip_t <= 2^-28 of a row's self term
  bool skip = false;
  float skip_t = 0.f;
  if (!dense && prec == FSK_PREC_F16) {
    skip = true;
    skip_t = -26.f;
  } else if (!dense && self_rows && (prec == FSK_PREC_F16X3S || prec == FSK_PREC_BF16X3S)) {
    skip = true;
    skip_t = -float(std::ceil(std::log2(double(n))) + 28.0);
  }
  // far-block bitmap: frame of the data, boxes of tile pairs and of the
  // waves' row blocks, one bit per (row block, pair)
  const unsigned long long *far_bits = nullptr;
  if (skip && L.far_bits && n > 0 && m > 0) {
    const double cabs = -exp2_coef(h);
    launch_far_bits(data, ld, n, rows_in, ldr, m, L.mu, inv_n, cabs, L.npairs_data, p.qblocks,
                    score_lp_waves(prec), skip_t, nullptr, 0.f, L.frame, L.pair_boxes,
                    L.row_boxes, L.far_bits, s);
    far_bits = L.far_bits;
  }
#define FSK_LP(P, PXT, QT)                                                                       \
  do {                                                                                           \
    hipLaunchKernelGGL(prep_lse_kernel<P>, dim3(cdiv(xrows, 256)), dim3(256), 0, s, data, ld, n, \
                       xrows, L.mu, inv_n, scl, coef, 0, (float *)nullptr, (float *)nullptr,      \
                       L.xrec);                                                                  \
    hipLaunchKernelGGL(prep_lse_kernel<P>, dim3(cdiv(qrows, 256)), dim3(256), 0, s, rows_in,     \
                       ldr, m, qrows, L.mu, inv_n, scl, coef, 1, (float *)nullptr,                \
                       (float *)nullptr,                                                         \
                       L.qrec);                                                                  \
    hipLaunchKernelGGL(prep_xt16_kernel<PXT>, dim3(cdiv(n32 * kD, 256)), dim3(256), 0, s, data,  \
                       ld, n, n32, L.mu, inv_n, scl, L.xt);                                      \
    hipLaunchKernelGGL(prep_lp_pairs_kernel<P>,                                                  \
                       dim3(cdiv(L.npairs_alloc * LpPair<P>::kSlots * 64, 256)), dim3(256), 0,   \
                       s, L.xrec, L.xt, L.npairs_data, L.npairs_alloc, L.blk);                    \
    prof_begin(FSK_OP_SCORE16_ORDERED, s);                                                       \
    launch_score_lp<P, QT>(skip, skip_t, far_bits, grid, s, L, m, split_pairs, p.qblocks, ntasks, \
                           per_xcd);                                                             \
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
