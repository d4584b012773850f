// Doubles contractions W(ijk) in split FP32 (the mixed-precision form).
//
// Same regrouped class GEMMs and tuple windows as the FP64 kernel
// (doubles_f64.hip).  Every operand x is held as hi = (float)x plus the
// residual lo = bf16(x - hi); a product is formed as
//     hi_a hi_b                      on v_mfma_f32_32x32x2_f32
//   + hi_a lo_b + lo_a hi_b          on two v_mfma_f32_32x32x16_bf16
// (the FP32 form drops the corrections: one FP32 product per product), all
// accumulating into the same FP32 accumulator.  Fragments are 32 rows
// (columns); for a K chunk of 16, lane l feeds row/column l % 32 at the 8
// consecutive kappa k0 + 8 (l / 32) + m: FP32 MFMA m (m < 8) takes kappa
// m of each half, and the bf16 MFMAs take all 8 (bf16(hi) made in
// registers from the hi already loaded, lo loaded): bf16(hi_a) . lo_b and
// lo_a . bf16(hi_b).  Per 32 x 32 block and chunk: 8 x 64 + 2 x 32 cycles,
// against 16 x 64 in FP64.  The corrections carry the part of each product
// below FP32's 24 bits (their own bf16 rounding is 2^-8 of an already
// 2^-24-small term).  The FP32 result of each class GEMM is stored as FP32
// (it carries no more bits) and widened to FP64 by the class-sum kernel;
// singles, the energy and all sums stay FP64 (energy.hip).
#include "doubles_common.hpp"

namespace atrip_amd {
namespace {

typedef float f8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// A lane's 8 hi values and 8 lo values of a chunk.
struct SParts {
  f8 h;
  bf8 l;
};

struct SplitParams {
  // per operand family: hi (float) and lo (bf16), row-major [row][K]
  const float *T3h, *H3h, *V3h, *T2Nh, *T2TNh;
  const __bf16 *T3l, *H3l, *V3l, *T2Nl, *T2TNl;
  const int32_t *windows;
  const uint64_t *aprefix;
  float *G;
  int64_t No, Nv, Kv, Ko, N2, W3, NoP, slots;
  uint64_t first, last;
  int nb, mtiles, ntiles;
  int64_t total, per_xcd;
};

template <int FM, int FN, bool CORR>
__global__ __launch_bounds__(256) void doubles_split_kernel(SplitParams P) {
  constexpr int MT = 64 * FM, NT = 64 * FN;
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
  const int kh = lane >> 5, l32 = lane & 31;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t No = P.No, N2 = P.N2, Nv = P.Nv;

  int32_t ca[FN], cb[FN], cc[FN], cslot[FN], crr[FN];  // 32-bit: live all loop
  bool cvalid[FN];
  bool any = false;
#pragma unroll
  for (int g = 0; g < FN; g++) {
    const int64_t n = (int64_t)nt * NT + wn * (NT / 2) + g * 32 + l32;
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
    int64_t m = (int64_t)mt * MT + wm * (MT / 2) + f * 32 + l32;
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
      aph[f] = Ah_ + r_ + 8 * kh;                                             \
      apc[f] = Al_ + r_ + 8 * kh;                                             \
    }                                                                         \
    _Pragma("unroll") for (int g = 0; g < FN; g++) {                          \
      int64_t u_, v_;                                                         \
      const int wh_ = b_pair(cl, (SG), ca[g], cb[g], cc[g], u_, v_);          \
      const int64_t c_ = ((u_ + Nv * v_) * No + crr[g]) * K_;                 \
      const float *Bh_ = wh_ == 0 ? P.V3h : wh_ == 1 ? P.T2Nh : P.T2TNh;      \
      const __bf16 *Bl_ = wh_ == 0 ? P.V3l : wh_ == 1 ? P.T2Nl : P.T2TNl;     \
      bph[g] = Bh_ + c_ + 8 * kh;                                             \
      bpc[g] = Bl_ + c_ + 8 * kh;                                             \
    }                                                                         \
  }
#define LOAD(A, B, KOFF)                                                     \
  {                                                                          \
    _Pragma("unroll") for (int f = 0; f < FM; f++) {                         \
      A[f].h = *(const f8 *)(aph[f] + (KOFF));                               \
      if constexpr (CORR) A[f].l = *(const bf8 *)(apc[f] + (KOFF));         \
    }                                                                        \
    _Pragma("unroll") for (int g = 0; g < FN; g++) {                         \
      B[g].h = *(const f8 *)(bph[g] + (KOFF));                               \
      if constexpr (CORR) B[g].l = *(const bf8 *)(bpc[g] + (KOFF));         \
    }                                                                        \
    __builtin_amdgcn_sched_barrier(0);                                       \
  }
  // FP32 MFMA m over all blocks before m + 1 (consecutive MFMAs update
  // different accumulators), then the two correction MFMAs
#define COMPUTE(A, B)                                                        \
  {                                                                          \
    _Pragma("unroll") for (int m = 0; m < 8; m++)                            \
        _Pragma("unroll") for (int f = 0; f < FM; f++)                       \
            _Pragma("unroll") for (int g = 0; g < FN; g++) acc[f][g] =       \
                __builtin_amdgcn_mfma_f32_32x32x2f32(A[f].h[m], B[g].h[m],   \
                                                     acc[f][g], 0, 0, 0);    \
    if constexpr (CORR) {                                                    \
      bf8 ha_[FM], hb_[FN];                                                  \
      _Pragma("unroll") for (int f = 0; f < FM; f++) ha_[f] =                \
          __builtin_convertvector(A[f].h, bf8);                              \
      _Pragma("unroll") for (int g = 0; g < FN; g++) hb_[g] =                \
          __builtin_convertvector(B[g].h, bf8);                              \
      _Pragma("unroll") for (int f = 0; f < FM; f++)                         \
          _Pragma("unroll") for (int g = 0; g < FN; g++) acc[f][g] =         \
              __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha_[f], B[g].l,        \
                                                      acc[f][g], 0, 0, 0);   \
      _Pragma("unroll") for (int f = 0; f < FM; f++)                         \
          _Pragma("unroll") for (int g = 0; g < FN; g++) acc[f][g] =         \
              __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[f].l, hb_[g],        \
                                                      acc[f][g], 0, 0, 0);   \
    }                                                                        \
  }

  f16v acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; f++)
#pragma unroll
    for (int g = 0; g < FN; g++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[f][g][r] = 0.f;

  // two operand sets in ping-pong: the loads of chunk c + 1 are in flight
  // behind the MFMAs of chunk c.  One pipeline runs through the four
  // segments: (sg, c) is the next chunk to load, and past a segment's end
  // the pointers move to the next segment's chunk 0 in place, so every
  // segment but the first starts with its first chunk in flight.  After the
  // last chunk the prefetch re-reads it (a load never used).
  static_assert(6 * (FM + FN) <= 63, "two chunks of loads must fit vmcnt");
  SParts a0[FM], b0[FN], a1[FM], b1[FN];
  int sg = 0;
  int64_t nch = Kv / KCHUNK, c = 0;
  SET_PTRS(0);
#define ADVANCE()                                                            \
  {                                                                          \
    if (++c == nch && ++sg < 4) {                                            \
      SET_PTRS(sg);                                                          \
      nch = (sg < 2 ? Kv : Ko) / KCHUNK;                                     \
      c = 0;                                                                 \
    }                                                                        \
  }
#define NEXT_OFF() (KCHUNK * (sg < 4 ? c : nch - 1))
  LOAD(a0, b0, 0);
  ADVANCE();
#pragma unroll 1
  for (;;) {
    const bool v1 = sg < 4;
    LOAD(a1, b1, NEXT_OFF());
    if (v1) ADVANCE();
    COMPUTE(a0, b0);
    if (!v1) break;
    const bool v2 = sg < 4;
    LOAD(a0, b0, NEXT_OFF());
    if (v2) ADVANCE();
    COMPUTE(a1, b1);
    if (!v2) break;
  }
#undef ADVANCE
#undef NEXT_OFF
#undef SET_PTRS
#undef LOAD
#undef COMPUTE

  // 32x32 FP32-output MFMAs: D[i][j] in lane j + 32 ((i / 4) % 2), register
  // 4 (i / 8) + i % 4
  float *Gc = P.G + ((int64_t)cl * P.slots + w * (int64_t)nb * nb * nb) * P.W3;
#pragma unroll
  for (int g = 0; g < FN; g++) {
    if (!cvalid[g]) continue;
    float *Gs = Gc + cslot[g] * P.W3;
    const int64_t rr = crr[g];
#pragma unroll
    for (int f = 0; f < FM; f++) {
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int64_t m = (int64_t)mt * MT + wm * (MT / 2) + f * 32 +
                          8 * (r >> 2) + 4 * kh + (r & 3);
        if (m >= N2) continue;
        Gs[m * P.NoP + rr] = acc[f][g][r];  // GEMM order: contiguous rows
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
  P.NoP = g_pitch(s.No, true);
  P.W3 = g_elems(s.No, true);
  P.slots = b.slots;
  P.first = b.first;
  P.last = b.last;
  P.nb = b.nb;
  tile = pick_t32_tile(tile, s.No, b.nb);
  const int FM = t32_fm(tile), FN = t32_fn(tile);
  const int64_t MT = 64 * FM, NT = 64 * FN;
  P.mtiles = (int)((P.N2 + MT - 1) / MT);
  P.ntiles = (int)(((int64_t)b.nb * s.No + NT - 1) / NT);
  P.total = (int64_t)b.nw * 3 * b.nb * b.nb * P.mtiles * P.ntiles;
  P.per_xcd = (P.total + 7) / 8;
  const int64_t grid = P.per_xcd * 8;
  if (grid <= 0) return hipSuccess;
  if (grid > MAX_GRID) return hipErrorInvalidConfiguration;
#define LAUNCH(A, B)                                                         \
  if (FM == A && FN == B) {                                                  \
    if (corrections)                                                         \
      hipLaunchKernelGGL((doubles_split_kernel<A, B, true>),                 \
                         dim3((unsigned)grid), dim3(256), 0, st, P);         \
    else                                                                     \
      hipLaunchKernelGGL((doubles_split_kernel<A, B, false>),                \
                         dim3((unsigned)grid), dim3(256), 0, st, P);         \
  }
  LAUNCH(2, 2) LAUNCH(2, 3) LAUNCH(3, 2) LAUNCH(2, 4) LAUNCH(3, 3)
  LAUNCH(2, 5) LAUNCH(3, 5) LAUNCH(4, 4) LAUNCH(3, 4) LAUNCH(4, 3)
#undef LAUNCH
  return hipGetLastError();
}

}  // namespace atrip_amd
