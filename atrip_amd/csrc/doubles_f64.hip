// Doubles contractions W(ijk) of the (T) hot path on FP64 MFMA (gfx950).
//
// The twelve GEMM terms of doubles_contribution (reference
// src/atrip/Equations.cxx:620-680, loop form :686-726) are regrouped by the
// W index that the single-occupied-index operand carries:
//   class w ("ij|k"): P0, P1, H5, H2   A-pair (a,b), columns over c
//   class v ("ik|j"): P3, P5, H1, H4   A-pair (a,c), columns over b
//   class u ("jk|i"): P4, P2, H0, H3   A-pair (b,c), columns over a
// Within a class the four terms share one output layout, so they are one GEMM
//   C[(p,q), (t,r)] = sum_kappa A[(p,q), kappa] B[kappa, (t,r)]
// with kappa running over four segments (E, E, L, L) and the hole operands
// stored pre-negated.  The A operand depends only on the class's index pair,
// so one A tile serves every tuple of the window that shares the pair (the
// cross-tuple reuse that per-tuple GEMMs lack), and the t index runs over
// the third tuple index inside the window.  Each class writes its own buffer
// in the W layout W[i + No j + No^2 k]; the energy kernel sums the three.
//
// MFMA v_mfma_f64_16x16x4f64, operands straight from global memory (L2/MALL
// resident) into registers.  A K chunk of 16 is split over the lane groups
// q = lane/16 as kappa = k0 + 2q + {0, 1} (k-steps 0, 1) and
// k0 + 8 + 2q + {0, 1} (k-steps 2, 3): each of a lane's two 16-byte loads
// then sits next to those of the other lane groups, so one load instruction
// covers 64 contiguous bytes of each of its 16 rows (two whole 32-byte
// sectors) and the L1 serves half the sector requests of a 32-byte-per-lane
// split.
#include "doubles_common.hpp"

namespace atrip_amd {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

// A lane's four doubles of a chunk: 2q, 2q+1 and 8+2q, 9+2q past p.
__device__ __forceinline__ d4 load_chunk(const double *p) {
  const d2 lo = *(const d2 *)p, hi = *(const d2 *)(p + 8);
  return __builtin_shufflevector(lo, hi, 0, 1, 2, 3);
}

struct DoublesParams {
  const double *T3, *V3, *H3, *T2N, *T2TN;
  const int32_t *windows;
  const uint64_t *aprefix;
  double *G;
  int64_t No, Nv, Kv, Ko, N2, W3, NoP, slots;
  uint64_t first, last;
  int nb, mtiles, ntiles;
  int64_t total, per_xcd;
};

// B operand base of segment sg for a column tuple.
__device__ inline const double *b_base(const DoublesParams &P, int cl, int sg,
                                       int64_t a, int64_t b, int64_t c) {
  int64_t u, v;
  const int which = b_pair(cl, sg, a, b, c, u, v);
  const double *Bt = which == 0 ? P.V3 : which == 1 ? P.T2N : P.T2TN;
  return Bt + (u + P.Nv * v) * P.No * (sg < 2 ? P.Kv : P.Ko);
}

template <int FM, int FN>
__global__ __launch_bounds__(256) void doubles_f64_kernel(DoublesParams P) {
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

  // columns (t, rr) of this lane's B fragments
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
    if (!ok) { a = 0; b = 0; c = 0; crr[g] = 0; }  // safe load address
    ca[g] = (int32_t)a; cb[g] = (int32_t)b; cc[g] = (int32_t)c;
    cslot[g] = (int32_t)slot;
  }
  if (!__syncthreads_or(any)) return;

  // A-pair of this class: w (a,b), v (a,c), u (b,c)
  int64_t x1, x2;
  class_pair(cl, al, be, ga, nb, o1, o2, x1, x2);
  if (x1 >= Nv) x1 = Nv - 1;
  if (x2 >= Nv) x2 = Nv - 1;

  // rows: straight (p + No q) and swapped (q + No p) row of the A slab
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
  // segments: 0 T3[x1] straight, 1 T3[x2] swapped, 2 H3[x2] straight,
  //           3 H3[x1] swapped
#define SET_PTRS(SG)                                                        \
  {                                                                         \
    const int64_t K_ = (SG) < 2 ? Kv : Ko;                                  \
    const double *Ab_ = (SG) < 2 ? P.T3 + ((SG) == 0 ? x1 : x2) * N2 * Kv   \
                                 : P.H3 + ((SG) == 2 ? x2 : x1) * N2 * Ko;  \
    _Pragma("unroll") for (int f = 0; f < FM; f++) ap[f] =                  \
        Ab_ + (((SG)&1) ? rowW[f] : rowS[f]) * K_ + 2 * kk;                 \
    _Pragma("unroll") for (int g = 0; g < FN; g++) bp[g] =                  \
        b_base(P, cl, (SG), ca[g], cb[g], cc[g]) + crr[g] * K_ + 2 * kk;    \
  }
  d4 acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; f++)
#pragma unroll
    for (int g = 0; g < FN; g++) acc[f][g] = d4{0.0, 0.0, 0.0, 0.0};

  // Software pipeline: two operand sets in ping-pong, so the loads of one
  // chunk are in flight while the MFMAs of the other run (no register
  // copies between the sets, hence no wait on the in-flight loads).
  const double *ap[FM], *bp[FN];
#define LOAD(AV, BV, KOFF)                                              \
  {                                                                     \
    _Pragma("unroll") for (int f = 0; f < FM; f++) AV[f] =              \
        load_chunk(ap[f] + (KOFF));                                     \
    _Pragma("unroll") for (int g = 0; g < FN; g++) BV[g] =              \
        load_chunk(bp[g] + (KOFF));                                     \
    __builtin_amdgcn_sched_barrier(0);                                  \
  }
  // k-steps 0, 1 hold kappa k0 .. k0 + 7 and steps 2, 3 kappa k0 + 8 ..
  // k0 + 15: a last chunk with at most 8 occupied kappa (the rest zero
  // padding, which adds exact zeros) skips steps 2, 3
#define STEPS(AV, BV, S0, S1)                                           \
  _Pragma("unroll") for (int s = S0; s < S1; s++)                       \
      _Pragma("unroll") for (int f = 0; f < FM; f++)                    \
          _Pragma("unroll") for (int g = 0; g < FN; g++) acc[f][g] =    \
              __builtin_amdgcn_mfma_f64_16x16x4f64(AV[f][s], BV[g][s],  \
                                                   acc[f][g], 0, 0, 0);
#define COMPUTE(AV, BV, HALF)                                           \
  {                                                                     \
    STEPS(AV, BV, 0, 2)                                                 \
    if (!(HALF)) { STEPS(AV, BV, 2, 4) }                                \
  }
  d4 a0[FM], b0[FN], a1[FM], b1[FN];
  // One pipeline through the four segments: (sg, c) is the next chunk to
  // load, and past a segment's end the pointers move to the next segment's
  // chunk 0 in place, so every segment but the first starts with its first
  // chunk in flight (3 % faster at No=40 than a pipeline restarted per
  // segment).  The prefetch is issued unconditionally (after the last chunk
  // it re-reads it, a load never used) so every path into a COMPUTE has the
  // same loads outstanding.  h0 / h1: the chunk in set 0 / 1 is a segment's
  // last and holds at most 8 occupied kappa.
  int sg = 0;
  int64_t nch = Kv / KCHUNK, c = 0;
  bool t8 = P.Nv - (Kv - KCHUNK) <= 8;  // tail rule of segment sg
  SET_PTRS(0);
#define ADVANCE()                                                            \
  {                                                                          \
    if (++c == nch && ++sg < 4) {                                            \
      SET_PTRS(sg);                                                          \
      const int64_t K_ = sg < 2 ? Kv : Ko;                                   \
      nch = K_ / KCHUNK;                                                     \
      t8 = (sg < 2 ? P.Nv : No) - (K_ - KCHUNK) <= 8;                        \
      c = 0;                                                                 \
    }                                                                        \
  }
#define NEXT_OFF() (KCHUNK * (sg < 4 ? c : nch - 1))
#define HALF_NEXT() (sg < 4 && t8 && c == nch - 1)
  bool h0 = HALF_NEXT(), h1 = false;
  LOAD(a0, b0, 0);
  ADVANCE();
#pragma unroll 1
  for (;;) {
    const bool v1 = sg < 4;
    h1 = HALF_NEXT();
    LOAD(a1, b1, NEXT_OFF());
    if (v1) ADVANCE();
    COMPUTE(a0, b0, h0);
    if (!v1) break;
    const bool v2 = sg < 4;
    h0 = HALF_NEXT();
    LOAD(a0, b0, NEXT_OFF());
    if (v2) ADVANCE();
    COMPUTE(a1, b1, h1);
    if (!v2) break;
  }
#undef ADVANCE
#undef NEXT_OFF
#undef HALF_NEXT
#undef LOAD
#undef STEPS
#undef COMPUTE

#undef SET_PTRS
  // store: for v_mfma_f64_16x16x4f64, D[i][j] sits in lane j + 16*(i%4),
  // register i/4 (measured on gfx950)
  double *Gc = P.G + ((int64_t)cl * P.slots + w * (int64_t)nb * nb * nb) * P.W3;
#pragma unroll
  for (int g = 0; g < FN; g++) {
    if (!cvalid[g]) continue;
    double *Gs = Gc + cslot[g] * P.W3;
    const int64_t rr = crr[g];
#pragma unroll
    for (int f = 0; f < FM; f++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t m =
            (int64_t)mt * MT + wm * (MT / 2) + f * 16 + 4 * r + kk;
        if (m >= N2) continue;
        Gs[m * P.NoP + rr] = acc[f][g][r];  // GEMM order: contiguous rows
      }
    }
  }
}

}  // namespace

hipError_t launch_doubles_f64(const DevSystem &s, const BatchArgs &b,
                              bool use_j, double *G, int tile,
                              hipStream_t st) {
  DoublesParams P;
  P.T3 = s.T3;
  P.V3 = use_j ? s.V3J : s.V3;
  P.H3 = use_j ? s.H3J : s.H3;
  P.T2N = s.T2N;
  P.T2TN = s.T2TN;
  P.windows = b.windows;
  P.aprefix = s.aprefix;
  P.G = G;
  P.No = s.No;
  P.Nv = s.Nv;
  P.Kv = s.Kv;
  P.Ko = s.Ko;
  P.N2 = s.No * s.No;
  P.NoP = g_pitch(s.No, false);
  P.W3 = g_elems(s.No, false);
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
  if (grid > MAX_GRID) return hipErrorInvalidConfiguration;
#define LAUNCH(A, B)                                                       \
  if (FM == A && FN == B)                                                  \
    hipLaunchKernelGGL((doubles_f64_kernel<A, B>), dim3((unsigned)grid),   \
                       dim3(256), 0, st, P);
  LAUNCH(2, 2) LAUNCH(4, 2) LAUNCH(2, 4) LAUNCH(4, 4) LAUNCH(3, 3)
  LAUNCH(5, 5) LAUNCH(4, 5) LAUNCH(5, 4)
#undef LAUNCH
  return hipGetLastError();
}

}  // namespace atrip_amd
