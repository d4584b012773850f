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
// resident) into registers.  A K chunk of 16 is split as
// kappa = k0 + 4*(lane/16) + s, s = 0..3, so every lane reads 4 consecutive
// doubles (32 B) of one row/column per chunk and feeds 4 MFMA k-steps.
#include "internal.hpp"

namespace atrip_amd {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

struct DoublesParams {
  const double *T3, *V3, *H3, *T2N, *T2TN;
  const int32_t *windows;
  const uint64_t *aprefix;
  double *G;
  int64_t No, Nv, Kv, Ko, N2, N3, slots;
  uint64_t first, last;
  int nb, mtiles, ntiles;
  int64_t total, per_xcd;
};

// Column tuple of class `cl` for A-pair offsets (o1,o2) and third offset t.
__device__ inline void class_tuple(int cl, int64_t al, int64_t be, int64_t ga,
                                   int nb, int o1, int o2, int t, int64_t &a,
                                   int64_t &b, int64_t &c, int64_t &slot) {
  if (cl == 0) {
    a = al * nb + o1; b = be * nb + o2; c = ga * nb + t;
    slot = ((int64_t)o1 * nb + o2) * nb + t;
  } else if (cl == 1) {
    a = al * nb + o1; b = be * nb + t; c = ga * nb + o2;
    slot = ((int64_t)o1 * nb + t) * nb + o2;
  } else {
    a = al * nb + t; b = be * nb + o1; c = ga * nb + o2;
    slot = ((int64_t)t * nb + o1) * nb + o2;
  }
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

  // column information
  int64_t colB[FN];   // tuple pair indices are resolved per segment below
  int64_t ca[FN], cb[FN], cc[FN], cslot[FN], crr[FN];
  bool cvalid[FN];
  bool any = false;
#pragma unroll
  for (int g = 0; g < FN; g++) {
    const int64_t n = (int64_t)nt * NT + wn * (NT / 2) + g * 16 + l16;
    const int t = (int)(n / No);
    crr[g] = n % No;
    int64_t a, b, c, slot;
    class_tuple(cl, al, be, ga, nb, o1, o2, t, a, b, c, slot);
    bool ok = t < nb && tuple_ok(Nv, a, b, c);
    if (ok && (P.first > 0 || P.last != ~0ULL)) {
      const uint64_t idx = tuple_index(P.aprefix, Nv, a, b, c);
      ok = idx >= P.first && idx < P.last;
    }
    cvalid[g] = ok;
    any |= ok;
    // clamp to a valid address for the loads of invalid columns
    if (!ok) { a = 0; b = 0; c = 0; crr[g] = 0; }
    ca[g] = a; cb[g] = b; cc[g] = c;
    cslot[g] = slot;
    colB[g] = 0;
  }
  if (!__syncthreads_or(any)) return;

  // A-pair indices of this class (valid whenever a column is valid)
  int64_t x1, x2;  // class w: (a,b)  v: (a,c)  u: (b,c)
  if (cl == 0) { x1 = al * nb + o1; x2 = be * nb + o2; }
  else if (cl == 1) { x1 = al * nb + o1; x2 = ga * nb + o2; }
  else { x1 = be * nb + o1; x2 = ga * nb + o2; }
  if (x1 >= Nv) x1 = Nv - 1;
  if (x2 >= Nv) x2 = Nv - 1;

  // rows
  int64_t rowS[FM], rowW[FM];
#pragma unroll
  for (int f = 0; f < FM; f++) {
    int64_t m = (int64_t)mt * MT + wm * (MT / 2) + f * 16 + l16;
    if (m >= N2) m = N2 - 1;
    const int64_t p = m % No, q = m / No;
    rowS[f] = m;
    rowW[f] = q + No * p;
  }

  d4 acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; f++)
#pragma unroll
    for (int g = 0; g < FN; g++) acc[f][g] = d4{0.0, 0.0, 0.0, 0.0};

  const int64_t Kv = P.Kv, Ko = P.Ko;
  for (int sg = 0; sg < 4; sg++) {
    // segment table (see file comment)
    const double *Abase;
    bool swap;
    int64_t K;
    if (sg < 2) {
      K = Kv;
      const int64_t xa = (cl == 0) ? (sg == 0 ? x1 : x2)
                       : (cl == 1) ? (sg == 0 ? x1 : x2)
                                   : (sg == 0 ? x1 : x2);
      Abase = P.T3 + xa * N2 * Kv;
      swap = sg == 1;
    } else {
      K = Ko;
      // w: H5 (b, straight), H2 (a, swapped); v: H1 (c), H4 (a, swapped);
      // u: H0 (c), H3 (b, swapped)
      int64_t xa;
      if (cl == 0) xa = (sg == 2) ? x2 : x1;
      else if (cl == 1) xa = (sg == 2) ? x2 : x1;
      else xa = (sg == 2) ? x2 : x1;
      Abase = P.H3 + xa * N2 * Ko;
      swap = sg == 3;
    }
    const double *ap[FM];
#pragma unroll
    for (int f = 0; f < FM; f++)
      ap[f] = Abase + (swap ? rowW[f] : rowS[f]) * K + 4 * kk;
    const double *bp[FN];
#pragma unroll
    for (int g = 0; g < FN; g++) {
      const int64_t a = ca[g], b = cb[g], c = cc[g];
      int64_t u, v;  // B pair (u,v) -> index u + Nv v
      const double *Bt;
      if (cl == 0) {
        if (sg == 0) { u = b; v = c; Bt = P.V3; }
        else if (sg == 1) { u = a; v = c; Bt = P.V3; }
        else if (sg == 2) { u = a; v = c; Bt = P.T2N; }
        else { u = b; v = c; Bt = P.T2N; }
      } else if (cl == 1) {
        if (sg == 0) { u = c; v = b; Bt = P.V3; }
        else if (sg == 1) { u = a; v = b; Bt = P.V3; }
        else if (sg == 2) { u = a; v = b; Bt = P.T2N; }
        else { u = b; v = c; Bt = P.T2TN; }
      } else {
        if (sg == 0) { u = c; v = a; Bt = P.V3; }
        else if (sg == 1) { u = b; v = a; Bt = P.V3; }
        else if (sg == 2) { u = a; v = b; Bt = P.T2TN; }
        else { u = a; v = c; Bt = P.T2TN; }
      }
      bp[g] = Bt + ((u + Nv * v) * No + crr[g]) * K + 4 * kk;
    }
    for (int64_t k0 = 0; k0 < K; k0 += KCHUNK) {
      d4 av[FM], bv[FN];
#pragma unroll
      for (int f = 0; f < FM; f++) av[f] = *(const d4 *)(ap[f] + k0);
#pragma unroll
      for (int g = 0; g < FN; g++) bv[g] = *(const d4 *)(bp[g] + k0);
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int f = 0; f < FM; f++)
#pragma unroll
          for (int g = 0; g < FN; g++)
            acc[f][g] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                av[f][s], bv[g][s], acc[f][g], 0, 0, 0);
    }
  }
  (void)colB;

  // store: for v_mfma_f64_16x16x4f64, D[i][j] sits in lane j + 16*(i%4),
  // register i/4 (measured on gfx950)
  double *Gc = P.G + ((int64_t)cl * P.slots + w * (int64_t)nb * nb * nb) * P.N3;
#pragma unroll
  for (int g = 0; g < FN; g++) {
    if (!cvalid[g]) continue;
    double *Gs = Gc + cslot[g] * P.N3;
    const int64_t rr = crr[g];
#pragma unroll
    for (int f = 0; f < FM; f++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t m =
            (int64_t)mt * MT + wm * (MT / 2) + f * 16 + 4 * r + kk;
        if (m >= N2) continue;
        const int64_t p = m % No, q = m / No;
        int64_t addr;
        if (cl == 0) addr = m + N2 * rr;            // (i,j,k) = (p,q,rr)
        else if (cl == 1) addr = p + No * rr + N2 * q;  // (p,rr,q)
        else addr = rr + No * m;                    // (rr,p,q)
        Gs[addr] = acc[f][g][r];
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
  P.N3 = P.N2 * s.No;
  P.slots = b.slots;
  P.first = b.first;
  P.last = b.last;
  P.nb = b.nb;
  int FM = 2, FN = 2;
  if (tile == 1) { FM = 4; FN = 2; }
  if (tile == 2) { FM = 2; FN = 4; }
  const int64_t MT = 32 * FM, NT = 32 * FN;
  P.mtiles = (int)((P.N2 + MT - 1) / MT);
  P.ntiles = (int)(((int64_t)b.nb * s.No + NT - 1) / NT);
  P.total = (int64_t)b.nw * 3 * b.nb * b.nb * P.mtiles * P.ntiles;
  P.per_xcd = (P.total + 7) / 8;
  const int64_t grid = P.per_xcd * 8;
  if (grid <= 0) return hipSuccess;
  if (FM == 4)
    hipLaunchKernelGGL((doubles_f64_kernel<4, 2>), dim3((unsigned)grid),
                       dim3(256), 0, st, P);
  else if (FN == 4)
    hipLaunchKernelGGL((doubles_f64_kernel<2, 4>), dim3((unsigned)grid),
                       dim3(256), 0, st, P);
  else
    hipLaunchKernelGGL((doubles_f64_kernel<2, 2>), dim3((unsigned)grid),
                       dim3(256), 0, st, P);
  return hipGetLastError();
}

}  // namespace atrip_amd
