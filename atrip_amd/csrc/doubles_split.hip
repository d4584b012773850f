// Doubles contractions W(ijk) in split FP32 (the mixed-precision form).
//
// Same regrouped class GEMMs and tuple windows as the FP64 kernel
// (doubles_f64.hip).  Every operand x is held as hi = (float)x plus the
// residual lo = bf16(x - hi); a product is formed as
//     hi_a hi_b                      on v_mfma_f32_16x16x4_f32
//   + (hi_a lo_b + lo_a hi_b)        on one v_mfma_f32_16x16x32_bf16
// (the FP32 form drops the bracket: one FP32 product per product)
// with both MFMAs accumulating into the same FP32 accumulator.  For a K chunk
// of 16, lane group q (lanes 16q..16q+15) holds kappa = k0 + 4q + s, s < 4,
// of its row/column for the FP32 MFMAs; the bf16 instruction's 8 K slots of
// the same lane group take [bf16(hi_a), lo_a] against [lo_b, bf16(hi_b)] at
// those four kappa, so its K = 32 covers the 16 hi*lo and the 16 lo*hi
// products of the chunk and bf16(hi) is made in registers from the hi
// already loaded (only hi and lo are stored: 6 bytes per element).  The
// correction terms carry the part of each product below FP32's 24 bits
// (their own bf16 rounding is 2^-8 of an already 2^-24-small term).  The FP32 result of each
// class GEMM is stored as FP32 (it carries no more bits) and widened to FP64
// by the energy kernel; singles, the energy and all sums stay FP64
// (energy.hip).
#include "doubles_common.hpp"

namespace atrip_amd {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// bf16 operands of the correction MFMA from a lane's four hi values h and
// four lo values l (same kappa): A takes [bf16(h), l], B takes [l, bf16(h)].
__device__ inline bf8 corr_a(f4 h, bf4 l) {
  const bf4 b = __builtin_convertvector(h, bf4);
  return __builtin_shufflevector(b, l, 0, 1, 2, 3, 4, 5, 6, 7);
}
__device__ inline bf8 corr_b(f4 h, bf4 l) {
  const bf4 b = __builtin_convertvector(h, bf4);
  return __builtin_shufflevector(l, b, 0, 1, 2, 3, 4, 5, 6, 7);
}

struct SplitParams {
  // per operand family: hi (float) and lo (bf16)
  const float *T3h, *H3h, *V3h, *T2Nh, *T2TNh;
  const __bf16 *T3l, *H3l, *V3l, *T2Nl, *T2TNl;
  const int32_t *windows;
  const uint64_t *aprefix;
  float *G;
  int64_t No, Nv, Kv, Ko, N2, W3, slots;
  uint64_t first, last;
  int nb, mtiles, ntiles;
  int64_t total, per_xcd;
};

template <int FM, int FN, bool CORR>
__global__ __launch_bounds__(256) void doubles_split_kernel(SplitParams P) {
  constexpr int MT = 32 * FM, NT = 32 * FN;
  const int64_t bid = blockIdx.x;
  const int64_t logical = (bid % 8) * P.per_xcd + bid / 8;  // XCD-aware
  if (logical >= P.total) return;
  int64_t rem = logical;
  const int nt = (int)(rem % P.ntiles); rem /= P.ntiles;
  const int mt = (int)(rem % P.mtiles); rem /= P.mtiles;
  const int nb = P.nb;
  const int pair = (int)(rem % (nb * nb)); rem /= (nb * nb);
  const int cl = (int)(rem % 3); rem /= 3;
  const int64_t w = rem;
  const int64_t al = P.windows[3 * w], be = P.windows[3 * w + 1],
                ga = P.windows[3 * w + 2];
  const int o1 = pair / nb, o2 = pair % nb;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kk = lane >> 4, l16 = lane & 15;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t No = P.No, N2 = P.N2, Nv = P.Nv;

  int32_t ca[FN], cb[FN], cc[FN], cslot[FN], crr[FN];  // 32-bit: live all loop
  bool cvalid[FN];
  bool any = false;
#pragma unroll
  for (int g = 0; g < FN; g++) {
    const int64_t n = (int64_t)nt * NT + wn * (NT / 2) + g * 16 + l16;
    const int t = (int)(n / No);
    crr[g] = (int32_t)(n % No);
    int64_t a, b, c, slot;
    class_tuple(cl, al, be, ga, nb, o1, o2, t, a, b, c, slot);
    bool ok = t < nb && tuple_ok(Nv, a, b, c);
    if (ok && (P.first > 0 || P.last != ~0ULL)) {
      const uint64_t idx = tuple_index(P.aprefix, Nv, a, b, c);
      ok = idx >= P.first && idx < P.last;
    }
    cvalid[g] = ok;
    any |= ok;
    if (!ok) { a = 0; b = 0; c = 0; crr[g] = 0; }
    ca[g] = (int32_t)a; cb[g] = (int32_t)b; cc[g] = (int32_t)c;
    cslot[g] = (int32_t)slot;
  }
  if (!__syncthreads_or(any)) return;

  int64_t x1, x2;
  class_pair(cl, al, be, ga, nb, o1, o2, x1, x2);
  if (x1 >= Nv) x1 = Nv - 1;
  if (x2 >= Nv) x2 = Nv - 1;
  int32_t rowS[FM], rowW[FM];  // 32-bit: live all loop
#pragma unroll
  for (int f = 0; f < FM; f++) {
    int64_t m = (int64_t)mt * MT + wm * (MT / 2) + f * 16 + l16;
    if (m >= N2) m = N2 - 1;
    const int64_t p = m % No, q = m / No;
    rowS[f] = (int32_t)m;
    rowW[f] = (int32_t)(q + No * p);
  }

  const int64_t Kv = P.Kv, Ko = P.Ko;
  const float *aph[FM], *bph[FN];
  const __bf16 *apc[FM], *bpc[FN];
#define SET_PTRS(SG)                                                          \
  {                                                                           \
    const int64_t K_ = (SG) < 2 ? Kv : Ko;                                    \
    const int64_t xa_ = ((SG) == 0 || (SG) == 3) ? x1 : x2;                   \
    const int64_t abase_ = xa_ * N2 * K_;                                     \
    const float *Ah_ = ((SG) < 2 ? P.T3h : P.H3h) + abase_;                   \
    const __bf16 *Al_ = ((SG) < 2 ? P.T3l : P.H3l) + abase_;                  \
    _Pragma("unroll") for (int f = 0; f < FM; f++) {                          \
      const int64_t r_ = (((SG)&1) ? rowW[f] : rowS[f]) * K_;                 \
      aph[f] = Ah_ + r_ + 4 * kk;                                             \
      apc[f] = Al_ + r_ + 4 * kk;                                             \
    }                                                                         \
    _Pragma("unroll") for (int g = 0; g < FN; g++) {                          \
      int64_t u_, v_;                                                         \
      const int wh_ = b_pair(cl, (SG), ca[g], cb[g], cc[g], u_, v_);          \
      const int64_t c_ = ((u_ + Nv * v_) * No + crr[g]) * K_;                 \
      const float *Bh_ = wh_ == 0 ? P.V3h : wh_ == 1 ? P.T2Nh : P.T2TNh;      \
      const __bf16 *Bl_ = wh_ == 0 ? P.V3l : wh_ == 1 ? P.T2Nl : P.T2TNl;     \
      bph[g] = Bh_ + c_ + 4 * kk;                                             \
      bpc[g] = Bl_ + c_ + 4 * kk;                                             \
    }                                                                         \
  }
#define LOAD(AH, AL, BH, BL, KOFF)                                           \
  {                                                                          \
    _Pragma("unroll") for (int f = 0; f < FM; f++) {                         \
      AH[f] = *(const f4 *)(aph[f] + (KOFF));                                \
      if constexpr (CORR) AL[f] = *(const bf4 *)(apc[f] + (KOFF));          \
    }                                                                        \
    _Pragma("unroll") for (int g = 0; g < FN; g++) {                         \
      BH[g] = *(const f4 *)(bph[g] + (KOFF));                                \
      if constexpr (CORR) BL[g] = *(const bf4 *)(bpc[g] + (KOFF));          \
    }                                                                        \
    __builtin_amdgcn_sched_barrier(0);                                       \
  }
#define COMPUTE(AH, AL, BH, BL)                                              \
  {                                                                          \
    _Pragma("unroll") for (int s = 0; s < 4; s++)                            \
        _Pragma("unroll") for (int f = 0; f < FM; f++)                       \
            _Pragma("unroll") for (int g = 0; g < FN; g++) acc[f][g] =       \
                __builtin_amdgcn_mfma_f32_16x16x4f32(AH[f][s], BH[g][s],     \
                                                     acc[f][g], 0, 0, 0);    \
    if constexpr (CORR) {                                                    \
      bf8 ca_[FM], cb_[FN];                                                  \
      _Pragma("unroll") for (int f = 0; f < FM; f++) ca_[f] =                \
          corr_a(AH[f], AL[f]);                                              \
      _Pragma("unroll") for (int g = 0; g < FN; g++) cb_[g] =                \
          corr_b(BH[g], BL[g]);                                              \
      _Pragma("unroll") for (int f = 0; f < FM; f++)                         \
          _Pragma("unroll") for (int g = 0; g < FN; g++) acc[f][g] =         \
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(ca_[f], cb_[g],        \
                                                      acc[f][g], 0, 0, 0);   \
    }                                                                        \
  }

  f4 acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; f++)
#pragma unroll
    for (int g = 0; g < FN; g++) acc[f][g] = f4{0.f, 0.f, 0.f, 0.f};

  // three operand sets in rotation: two chunks of loads stay in flight
  // behind the MFMAs of the current one (FP32 MFMAs take half the cycles of
  // FP64 ones, so one chunk of compute alone does not cover the latency)
  f4 ah0[FM], bh0[FN], ah1[FM], bh1[FN], ah2[FM], bh2[FN];
  bf4 ac0[FM], bc0[FN], ac1[FM], bc1[FN], ac2[FM], bc2[FN];
#define CLAMP(K_) ((K_) < last ? (K_) : last)
#pragma unroll 1
  for (int sg = 0; sg < 4; sg++) {
    SET_PTRS(sg);
    const int64_t last = (sg < 2 ? Kv : Ko) - KCHUNK;
    LOAD(ah0, ac0, bh0, bc0, 0);
    LOAD(ah1, ac1, bh1, bc1, CLAMP(KCHUNK));
    for (int64_t k0 = 0;; k0 += 3 * KCHUNK) {
      LOAD(ah2, ac2, bh2, bc2, CLAMP(k0 + 2 * KCHUNK));
      COMPUTE(ah0, ac0, bh0, bc0);
      if (k0 + KCHUNK > last) break;
      LOAD(ah0, ac0, bh0, bc0, CLAMP(k0 + 3 * KCHUNK));
      COMPUTE(ah1, ac1, bh1, bc1);
      if (k0 + 2 * KCHUNK > last) break;
      LOAD(ah1, ac1, bh1, bc1, CLAMP(k0 + 4 * KCHUNK));
      COMPUTE(ah2, ac2, bh2, bc2);
      if (k0 + 3 * KCHUNK > last) break;
    }
  }
#undef CLAMP
#undef SET_PTRS
#undef LOAD
#undef COMPUTE

  // store: for the 16x16 FP32-output MFMAs, D[i][j] sits in lane
  // j + 16*(i/4), register i%4 (measured on gfx950)
  float *Gc = P.G + ((int64_t)cl * P.slots + w * (int64_t)nb * nb * nb) * P.W3;
#pragma unroll
  for (int g = 0; g < FN; g++) {
    if (!cvalid[g]) continue;
    float *Gs = Gc + cslot[g] * P.W3;
    const int64_t rr = crr[g];
#pragma unroll
    for (int f = 0; f < FM; f++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t m =
            (int64_t)mt * MT + wm * (MT / 2) + f * 16 + 4 * kk + r;
        if (m >= N2) continue;
        Gs[m * No + rr] = acc[f][g][r];  // GEMM order: 64-byte rows
      }
    }
  }
}

// x -> hi = (float)x, l = bf16(x - hi); bf16 by round to nearest even on
// the float bits.
__global__ __launch_bounds__(256) void split_kernel(const double *x, float *h,
                                                    uint16_t *l, int64_t n) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    split_store(x[q], h, l, q);
  }
}

}  // namespace

hipError_t launch_split(const double *x, DevSystem::Split &out, int64_t n,
                        hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536L * 16) blocks = 65536L * 16;
  hipLaunchKernelGGL(split_kernel, dim3((unsigned)blocks), dim3(256), 0, st, x,
                     out.h, out.l, n);
  return hipGetLastError();
}

hipError_t launch_doubles_split(const DevSystem &s, const BatchArgs &b,
                                bool use_j, float *G, bool corrections,
                                int tile, hipStream_t st) {
  SplitParams P;
  const DevSystem::Split &V = use_j ? s.sV3J : s.sV3;
  const DevSystem::Split &H = use_j ? s.sH3J : s.sH3;
  P.T3h = s.sT3.h; P.H3h = H.h; P.V3h = V.h; P.T2Nh = s.sT2N.h;
  P.T2TNh = s.sT2TN.h;
  P.T3l = (const __bf16 *)s.sT3.l; P.H3l = (const __bf16 *)H.l;
  P.V3l = (const __bf16 *)V.l; P.T2Nl = (const __bf16 *)s.sT2N.l;
  P.T2TNl = (const __bf16 *)s.sT2TN.l;
  P.windows = b.windows;
  P.aprefix = s.aprefix;
  P.G = G;
  P.No = s.No;
  P.Nv = s.Nv;
  P.Kv = s.Kv;
  P.Ko = s.Ko;
  P.N2 = s.No * s.No;
  P.W3 = s.No * s.No * s.No;
  P.slots = b.slots;
  P.first = b.first;
  P.last = b.last;
  P.nb = b.nb;
  tile = pick_tile(tile, s.No, b.nb);
  const int FM = tile_fm(tile), FN = tile_fn(tile);
  const int64_t MT = 32 * FM, NT = 32 * FN;
  P.mtiles = (int)((P.N2 + MT - 1) / MT);
  P.ntiles = (int)(((int64_t)b.nb * s.No + NT - 1) / NT);
  P.total = (int64_t)b.nw * 3 * b.nb * b.nb * P.mtiles * P.ntiles;
  P.per_xcd = (P.total + 7) / 8;
  const int64_t grid = P.per_xcd * 8;
  if (grid <= 0) return hipSuccess;
#define LAUNCH(A, B)                                                         \
  if (FM == A && FN == B) {                                                  \
    if (corrections)                                                         \
      hipLaunchKernelGGL((doubles_split_kernel<A, B, true>),                 \
                         dim3((unsigned)grid), dim3(256), 0, st, P);         \
    else                                                                     \
      hipLaunchKernelGGL((doubles_split_kernel<A, B, false>),                \
                         dim3((unsigned)grid), dim3(256), 0, st, P);         \
  }
  LAUNCH(2, 2) LAUNCH(4, 2) LAUNCH(2, 4) LAUNCH(4, 4) LAUNCH(3, 3)
  LAUNCH(5, 5) LAUNCH(4, 5) LAUNCH(5, 4)
#undef LAUNCH
  return hipGetLastError();
}

}  // namespace atrip_amd
