// Batched form of the (T) hot path: the tuples of many small resident
// systems of uneven No/Nv in one persistent launch.
//
// A work item is one window (block triple of edge nb) of one system.  A
// workgroup takes work items from a device counter in the dispatch order the
// host scheduler chose, and does everything for its window itself:
//   1. the doubles contractions (reference src/atrip/Equations.cxx:455-728,
//      regrouped into the three classes w, v, u of the FP64 doubles kernel),
//      class by class.  Each of the four waves takes wave tiles of
//      (16*FM) x (16*FN) outputs (FM, FN per system, from the tuner), runs
//      the FP64 MFMA K loop with operands straight from global memory and
//      adds the tile into W of the window's tuples in the workgroup's own
//      scratch (class w stores, v and u add after a barrier, so W sums the
//      classes in the fixed order (w + v) + u);
//   2. singles and energy of each tuple of the window (reference
//      src/atrip/Equations.cxx:101-238 and :387-426, selection as
//      src/atrip/Atrip.cxx:636-670), one wave per tuple over the i>=j>=k
//      points, reduced by a fixed butterfly;
//   3. a compensated sum of the window's tuple energies in slot order.
// Nothing depends on which workgroup runs which window or on the other
// systems of the launch, so a system's energy is the same alone and in any
// mix.  No workgroup waits on another; the loop ends when the counter passes
// the number of work items.
#include "doubles_common.hpp"

namespace atrip_amd {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// correction-MFMA operands from four hi and four lo values of one kappa
// range: A [bf16(hi), lo], B [lo, bf16(hi)] (as the plain split kernel)
__device__ inline bf8 corr_a(f4 h, bf4 l) {
  const bf4 b = __builtin_convertvector(h, bf4);
  return __builtin_shufflevector(b, l, 0, 1, 2, 3, 4, 5, 6, 7);
}
__device__ inline bf8 corr_b(f4 h, bf4 l) {
  const bf4 b = __builtin_convertvector(h, bf4);
  return __builtin_shufflevector(l, b, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// B operand base of segment sg for column tuple (a,b,c) of class cl.
__device__ __forceinline__ const double *bbase(const BSys &S, int cl, int sg, int64_t a,
                                      int64_t b, int64_t c) {
  int64_t u, v;
  const int which = b_pair(cl, sg, a, b, c, u, v);
  const double *Bt = which == 0 ? S.V3 : which == 1 ? S.T2N : S.T2TN;
  return Bt + (u + S.Nv * v) * S.No * (sg < 2 ? S.Kv : S.Ko);
}

// Adds one wave tile into W (order i + No j + No^2 k).  Class u's column
// index is i, so its fragment rows are stored as they are (16 consecutive
// i per row, 128 bytes).  For classes w and v the fragment is first
// transposed through the wave's 16 x 17 LDS area so that a store
// instruction covers 16 consecutive rows m (consecutive i) of 4 columns.
// Per column group the element offsets come first and every old value
// (classes v and u add) is loaded before any store, the next group's
// before this group's stores, so the adds wait on memory about once per
// tile instead of once per element (a load after a store to a possibly
// aliasing address cannot be moved above it).  Every element
// of W receives one contribution per class, so within a class the tiles'
// elements are disjoint.
// OUT = 8: FP64 MFMA layout D[i][j] in lane j + 16 (i % 4), register i / 4;
// OUT = 4: FP32 MFMA layout D[i][j] in lane j + 16 (i / 4), register i % 4.
template <int FM, int FN, int OUT, typename V4>
__device__ __forceinline__ void store_tile(const V4 (&acc)[FM][FN], int cl,
                                           int mt, int nt, int64_t No,
                                           const bool *cvalid,
                                           const int32_t *cslot,
                                           const int32_t *crr, double *Gwin,
                                           bool add, int lane, double *xp) {
  const int kk = lane >> 4, l16 = lane & 15;
  const int no = (int)No, n2 = no * no, w3 = n2 * no;
  int off[2][FM][4];     // element offsets in the window's W, -1 if none
  double old[2][FM][4];  // their old values (classes v and u)
  // offsets and old values of column group g into set h
#define GROUP_LOAD(g, h)                                                       \
  {                                                                            \
    if (cl == 2) {                                                             \
      _Pragma("unroll") for (int f = 0; f < FM; f++)                           \
          _Pragma("unroll") for (int r = 0; r < 4; r++) {                      \
        const int row = OUT == 8 ? 4 * r + kk : 4 * kk + r;                    \
        const int m = mt * 16 * FM + f * 16 + row;                             \
        const int p = m % no, q = m / no;                                      \
        off[h][f][r] = cvalid[g] && m < n2                                     \
                           ? cslot[g] * w3 + crr[g] + no * (p + no * q)        \
                           : -1;                                               \
      }                                                                        \
    } else {                                                                   \
      _Pragma("unroll") for (int r = 0; r < 4; r++) {                          \
        const int col = 4 * r + kk;                                            \
        const int slot = __shfl(cslot[g], col, 64);                            \
        const int rr = __shfl(crr[g], col, 64);                                \
        const bool ok = __shfl((int)cvalid[g], col, 64) != 0;                  \
        _Pragma("unroll") for (int f = 0; f < FM; f++) {                       \
          const int m = mt * 16 * FM + f * 16 + l16;                           \
          const int p = m % no, q = m / no;                                    \
          const int o =                                                        \
              cl == 0 ? p + no * (q + no * rr) : p + no * (rr + no * q);       \
          off[h][f][r] = ok && m < n2 ? slot * w3 + o : -1;                    \
        }                                                                      \
      }                                                                        \
    }                                                                          \
    if (add)                                                                   \
      _Pragma("unroll") for (int f = 0; f < FM; f++)                           \
          _Pragma("unroll") for (int r = 0; r < 4; r++) old[h][f][r] =         \
              Gwin[off[h][f][r] >= 0 ? off[h][f][r] : 0];                      \
  }
  GROUP_LOAD(0, 0)
#pragma unroll
  for (int g = 0; g < FN; g++) {
    const int h = g & 1;
    // the next group's old values are requested before this group's stores
    if (g + 1 < FN) GROUP_LOAD(g + 1, h ^ 1)
#pragma unroll
    for (int f = 0; f < FM; f++) {
      double v[4];
      if (cl == 2) {
#pragma unroll
        for (int r = 0; r < 4; r++) v[r] = (double)acc[f][g][r];
      } else {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int row = OUT == 8 ? 4 * r + kk : 4 * kk + r;
          xp[row * 17 + l16] = (double)acc[f][g][r];
        }
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): LDS writes done
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < 4; r++) v[r] = xp[l16 * 17 + 4 * r + kk];
        __builtin_amdgcn_s_waitcnt(0xC07F);  // reads done before the next tile
        __builtin_amdgcn_wave_barrier();
      }
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (off[h][f][r] >= 0)
          Gwin[off[h][f][r]] = add ? old[h][f][r] + v[r] : v[r];
    }
  }
#undef GROUP_LOAD
}

// One wave tile: rows mt*16*FM .. of the No^2 (p,q) rows, columns
// nt*16*FN .. of the nb*No (t, r) columns of class cl at A-pair (o1,o2).
// W of window slot s sits at Gwin + s*No^3 in the order i + No j + No^2 k.
template <int FM, int FN>
__device__ __forceinline__ void wave_tile(const BSys &S, int cl, int al,
                                          int be, int ga, int o1, int o2,
                                          int mt, int nt, double *Gwin,
                                          bool add, int lane, double *xp) {
  const int nb = S.nb;
  const int64_t No = S.No, Nv = S.Nv, N2 = No * No, Kv = S.Kv, Ko = S.Ko;
  const int kk = lane >> 4, l16 = lane & 15;
  int32_t ca[FN], cb[FN], cc[FN], cslot[FN], crr[FN];
  bool cvalid[FN];
  bool any = false;
#pragma unroll
  for (int g = 0; g < FN; g++) {
    const int64_t n = (int64_t)nt * 16 * FN + g * 16 + l16;
    const int t = (int)(n / No);
    crr[g] = (int32_t)(n % No);
    int64_t a, b, c, slot;
    class_tuple(cl, al, be, ga, nb, o1, o2, t, a, b, c, slot);
    const bool ok = t < nb && tuple_ok(Nv, a, b, c);
    cvalid[g] = ok;
    any |= ok;
    if (!ok) { a = 0; b = 0; c = 0; crr[g] = 0; slot = 0; }
    ca[g] = (int32_t)a; cb[g] = (int32_t)b; cc[g] = (int32_t)c;
    cslot[g] = (int32_t)slot;
  }
  if (!__any(any)) return;

  int64_t x1, x2;
  class_pair(cl, al, be, ga, nb, o1, o2, x1, x2);
  if (x1 >= Nv) x1 = Nv - 1;
  if (x2 >= Nv) x2 = Nv - 1;
  int32_t rowS[FM], rowW[FM];
#pragma unroll
  for (int f = 0; f < FM; f++) {
    int64_t m = (int64_t)mt * 16 * FM + f * 16 + l16;
    if (m >= N2) m = N2 - 1;
    rowS[f] = (int32_t)m;
    rowW[f] = (int32_t)(m / No + No * (m % No));
  }

  d4 acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; f++)
#pragma unroll
    for (int g = 0; g < FN; g++) acc[f][g] = d4{0.0, 0.0, 0.0, 0.0};
  const double *ap[FM], *bp[FN];
  d4 a0[FM], b0[FN], a1[FM], b1[FN];
#define LOAD(AV, BV, KOFF)                                              \
  {                                                                     \
    _Pragma("unroll") for (int f = 0; f < FM; f++) AV[f] =              \
        *(const d4 *)(ap[f] + (KOFF));                                  \
    _Pragma("unroll") for (int g = 0; g < FN; g++) BV[g] =              \
        *(const d4 *)(bp[g] + (KOFF));                                  \
    __builtin_amdgcn_sched_barrier(0);                                  \
  }
#define COMPUTE(AV, BV)                                                   \
  {                                                                       \
    _Pragma("unroll") for (int s = 0; s < 4; s++)                         \
        _Pragma("unroll") for (int f = 0; f < FM; f++)                    \
            _Pragma("unroll") for (int g = 0; g < FN; g++) acc[f][g] =    \
                __builtin_amdgcn_mfma_f64_16x16x4f64(AV[f][s], BV[g][s],  \
                                                     acc[f][g], 0, 0, 0); \
  }
  // segments: 0 T3[x1] straight, 1 T3[x2] swapped, 2 H3[x2] straight,
  //           3 H3[x1] swapped; K runs are multiples of 16.  Each segment
  // has its own two-stage pipeline: one pipeline through all four keeps the
  // next segment's pointers live as well and costs the kernel its second
  // wave per SIMD.
#pragma unroll 1
  for (int sg = 0; sg < 4; sg++) {
    const int64_t K_ = sg < 2 ? Kv : Ko;
    const double *Ab = sg < 2 ? S.T3 + (sg == 0 ? x1 : x2) * N2 * Kv
                              : S.H3 + (sg == 2 ? x2 : x1) * N2 * Ko;
#pragma unroll
    for (int f = 0; f < FM; f++)
      ap[f] = Ab + (int64_t)((sg & 1) ? rowW[f] : rowS[f]) * K_ + 4 * kk;
#pragma unroll
    for (int g = 0; g < FN; g++)
      bp[g] = bbase(S, cl, sg, ca[g], cb[g], cc[g]) + (int64_t)crr[g] * K_ +
              4 * kk;
    const int64_t last = K_ - KCHUNK;
    LOAD(a0, b0, 0);
    for (int64_t k0 = 0;; k0 += 2 * KCHUNK) {
      LOAD(a1, b1, k0 + KCHUNK < last ? k0 + KCHUNK : last);
      COMPUTE(a0, b0);
      if (k0 + KCHUNK > last) break;
      LOAD(a0, b0, k0 + 2 * KCHUNK < last ? k0 + 2 * KCHUNK : last);
      COMPUTE(a1, b1);
      if (k0 + 2 * KCHUNK > last) break;
    }
  }
#undef LOAD
#undef COMPUTE
  store_tile<FM, FN, 8>(acc, cl, mt, nt, No, cvalid, cslot, crr, Gwin, add,
                        lane, xp);
}

// The same wave tile with FP32 products (mixed forms): hi*hi on
// v_mfma_f32_16x16x4f32 and, when CORR, hi*lo + lo*hi in one
// v_mfma_f32_16x16x32_bf16 per 16-deep chunk (each lane group pairs
// [bf16(hi_a), lo_a] with [lo_b, bf16(hi_b)] over its four kappa), FP32
// accumulation; the tile is widened to FP64 when it is added into W.
template <int FM, int FN, bool CORR>
__device__ __forceinline__ void wave_tile_split(const BSys &S, const BSplit &X,
                                                int cl, int al, int be, int ga,
                                                int o1, int o2, int mt, int nt,
                                                double *Gwin, bool add,
                                                int lane, double *xp) {
  const int nb = S.nb;
  const int64_t No = S.No, Nv = S.Nv, N2 = No * No, Kv = S.Kv, Ko = S.Ko;
  const int kk = lane >> 4, l16 = lane & 15;
  int32_t ca[FN], cb[FN], cc[FN], cslot[FN], crr[FN];
  bool cvalid[FN];
  bool any = false;
#pragma unroll
  for (int g = 0; g < FN; g++) {
    const int64_t n = (int64_t)nt * 16 * FN + g * 16 + l16;
    const int t = (int)(n / No);
    crr[g] = (int32_t)(n % No);
    int64_t a, b, c, slot;
    class_tuple(cl, al, be, ga, nb, o1, o2, t, a, b, c, slot);
    const bool ok = t < nb && tuple_ok(Nv, a, b, c);
    cvalid[g] = ok;
    any |= ok;
    if (!ok) { a = 0; b = 0; c = 0; crr[g] = 0; slot = 0; }
    ca[g] = (int32_t)a; cb[g] = (int32_t)b; cc[g] = (int32_t)c;
    cslot[g] = (int32_t)slot;
  }
  if (!__any(any)) return;

  int64_t x1, x2;
  class_pair(cl, al, be, ga, nb, o1, o2, x1, x2);
  if (x1 >= Nv) x1 = Nv - 1;
  if (x2 >= Nv) x2 = Nv - 1;
  int32_t rowS[FM], rowW[FM];
#pragma unroll
  for (int f = 0; f < FM; f++) {
    int64_t m = (int64_t)mt * 16 * FM + f * 16 + l16;
    if (m >= N2) m = N2 - 1;
    rowS[f] = (int32_t)m;
    rowW[f] = (int32_t)(m / No + No * (m % No));
  }
  f4 acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; f++)
#pragma unroll
    for (int g = 0; g < FN; g++) acc[f][g] = f4{0.f, 0.f, 0.f, 0.f};
  const float *aph[FM], *bph[FN];
  const __bf16 *apc[FM], *bpc[FN];
  f4 ah0[FM], bh0[FN], ah1[FM], bh1[FN];
  bf4 ac0[FM], bc0[FN], ac1[FM], bc1[FN];
#define LOAD(AH, AC, BH, BC, KOFF)                                          \
  {                                                                         \
    _Pragma("unroll") for (int f = 0; f < FM; f++) {                        \
      AH[f] = *(const f4 *)(aph[f] + (KOFF));                               \
      if constexpr (CORR) AC[f] = *(const bf4 *)(apc[f] + (KOFF));         \
    }                                                                       \
    _Pragma("unroll") for (int g = 0; g < FN; g++) {                        \
      BH[g] = *(const f4 *)(bph[g] + (KOFF));                               \
      if constexpr (CORR) BC[g] = *(const bf4 *)(bpc[g] + (KOFF));         \
    }                                                                       \
    __builtin_amdgcn_sched_barrier(0);                                      \
  }
#define COMPUTE(AH, AC, BH, BC)                                             \
  {                                                                         \
    _Pragma("unroll") for (int s = 0; s < 4; s++)                           \
        _Pragma("unroll") for (int f = 0; f < FM; f++)                      \
            _Pragma("unroll") for (int g = 0; g < FN; g++) acc[f][g] =      \
                __builtin_amdgcn_mfma_f32_16x16x4f32(AH[f][s], BH[g][s],    \
                                                     acc[f][g], 0, 0, 0);   \
    if constexpr (CORR) {                                                   \
      bf8 ca_[FM], cb_[FN];                                                 \
      _Pragma("unroll") for (int f = 0; f < FM; f++) ca_[f] =               \
          corr_a(AH[f], AC[f]);                                             \
      _Pragma("unroll") for (int g = 0; g < FN; g++) cb_[g] =               \
          corr_b(BH[g], BC[g]);                                             \
      _Pragma("unroll") for (int f = 0; f < FM; f++)                        \
          _Pragma("unroll") for (int g = 0; g < FN; g++) acc[f][g] =        \
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(ca_[f], cb_[g],       \
                                                      acc[f][g], 0, 0, 0);  \
    }                                                                       \
  }
  // per-segment pipelines: one pipeline across the segments (as in
  // wave_tile) keeps more pointers live and costs this tile its second wave
  // per SIMD
#pragma unroll 1
  for (int sg = 0; sg < 4; sg++) {
    const int64_t K_ = sg < 2 ? Kv : Ko;
    const int64_t xa = (sg == 0 || sg == 3) ? x1 : x2;
    const int64_t abase = xa * N2 * K_;
    const float *Ah = (sg < 2 ? X.T3h : X.H3h) + abase;
    const __bf16 *Ac = (const __bf16 *)(sg < 2 ? X.T3l : X.H3l) + abase;
#pragma unroll
    for (int f = 0; f < FM; f++) {
      const int64_t r = (int64_t)((sg & 1) ? rowW[f] : rowS[f]) * K_;
      aph[f] = Ah + r + 4 * kk;
      apc[f] = Ac + r + 4 * kk;
    }
#pragma unroll
    for (int g = 0; g < FN; g++) {
      int64_t u, v;
      const int wh = b_pair(cl, sg, ca[g], cb[g], cc[g], u, v);
      const int64_t c = ((u + Nv * v) * No + crr[g]) * K_;
      const float *Bh = wh == 0 ? X.V3h : wh == 1 ? X.T2Nh : X.T2TNh;
      const uint16_t *Bc = wh == 0 ? X.V3l : wh == 1 ? X.T2Nl : X.T2TNl;
      bph[g] = Bh + c + 4 * kk;
      bpc[g] = (const __bf16 *)Bc + c + 4 * kk;
    }
    const int64_t last = K_ - KCHUNK;
    LOAD(ah0, ac0, bh0, bc0, 0);
    for (int64_t k0 = 0;; k0 += 2 * KCHUNK) {
      LOAD(ah1, ac1, bh1, bc1, k0 + KCHUNK < last ? k0 + KCHUNK : last);
      COMPUTE(ah0, ac0, bh0, bc0);
      if (k0 + KCHUNK > last) break;
      LOAD(ah0, ac0, bh0, bc0, k0 + 2 * KCHUNK < last ? k0 + 2 * KCHUNK : last);
      COMPUTE(ah1, ac1, bh1, bc1);
      if (k0 + 2 * KCHUNK > last) break;
    }
  }
#undef LOAD
#undef COMPUTE

  store_tile<FM, FN, 4>(acc, cl, mt, nt, No, cvalid, cslot, crr, Gwin, add,
                        lane, xp);
}

// Energy of tuple (a,b,c) from its W (order i + No j + No^2 k), by one
// wave; tA/tB/tC hold Tai at (a,x), (b,x), (c,x).
__device__ __forceinline__ double tuple_energy(const BSys &S, const double *W, int64_t a,
                               int64_t b, int64_t c, const double *tA,
                               const double *tB, const double *tC, int lane) {
  const int64_t No = S.No, Nv = S.Nv, N2 = No * No;
  int distinct = 0;
  if (a == b) distinct++;
  if (b == c) distinct--;
  const double epsabc = (S.ea[a] + S.ea[b]) + S.ea[c];
  const double *Vbc = S.VHH + (b + Nv * c) * N2;
  const double *Vac = S.VHH + (a + Nv * c) * N2;
  const double *Vab = S.VHH + (a + Nv * b) * N2;
  const int no = (int)No;
  auto Wv = [&](int x, int y, int z) { return W[x + no * (y + no * z)]; };
  // singles at a permutation (x,y,z): t_a[x] V^bc_yz + t_b[y] V^ac_xz
  // + t_c[z] V^ab_xy
  auto Z = [&](double wv, int x, int y, int z) {
    double zz = wv + tA[x] * Vbc[y + no * z];
    zz = zz + tB[y] * Vac[x + no * z];
    zz = zz + tC[z] * Vab[x + no * y];
    return zz;
  };
  double acc = 0.0;
#pragma unroll 1
  for (int64_t pt = lane; pt < S.npoints; pt += 64) {
    const uint32_t pk = S.points[pt];
    const int i = pk & 1023, j = (pk >> 10) & 1023, k = pk >> 20;
    const double facjk = j == k ? 0.5 : 1.0;
    const double facij = i == j ? 0.5 : 1.0;
    const double eijk = (S.ei[i] + S.ei[j]) + S.ei[k];
    const double denominator = epsabc - eijk;
    if (distinct == 0) {
      const double A = Wv(i, j, k), B = Wv(i, k, j), C = Wv(j, i, k),
                   D = Wv(j, k, i), E = Wv(k, i, j), F = Wv(k, j, i);
      const double U = Z(A, i, j, k), V = Z(B, i, k, j), W_ = Z(C, j, i, k),
                   X = Z(D, j, k, i), Y = Z(E, k, i, j), Zz = Z(F, k, j, i);
      const double UXY = U + (X + Y), VWZ = V + (W_ + Zz);
      const double first =
          A * U + (B * V + (C * W_ + (D * X + (E * Y + F * Zz))));
      const double second = (UXY - 2.0 * VWZ) * (A + (D + E));
      const double third = (VWZ - 2.0 * UXY) * (B + (C + F));
      const double value = 3.0 * first + (second + third);
      acc += ((2.0 * value) / denominator) * (facjk * facij);
    } else {
      const double A = Wv(i, j, k), B = Wv(j, k, i), C = Wv(k, i, j);
      const double U = Z(A, i, j, k), V = Z(B, j, k, i), W_ = Z(C, k, i, j);
      const double UVW = U + (V + W_);
      const double value =
          3.0 * ((A * U + B * V) + C * W_) - (A + (B + C)) * UVW;
      acc += ((2.0 * value) / denominator) * (facjk * facij);
    }
  }
  return wave_sum(acc);
}

constexpr int BMAX_NO = 64;  // singles tables in LDS

// Load a descriptor the whole workgroup shares and make every 32-bit word
// wave-uniform, so it lives in scalar registers.
template <typename T>
__device__ __forceinline__ T uniform_load(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "descriptor size");
  T v = *p;
  uint32_t w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; i++)
    w[i] = __builtin_amdgcn_readfirstlane(w[i]);
  __builtin_memcpy(&v, w, sizeof(T));
  return v;
}

template <int PREC>
__global__ __launch_bounds__(256, 2) void batched_kernel(BatchedLaunch P) {
  __shared__ int task_sh;
  __shared__ double etup[512];
  __shared__ double tsing[4][3][BMAX_NO];
  __shared__ double xpose[4][16 * 17];  // per-wave tile transposition
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double *Gwin = P.scratch + (int64_t)blockIdx.x * P.scratch_stride;
#pragma unroll 1
  for (;;) {
    if (tid == 0) task_sh = (int)atomicAdd(P.counter, 1u);
    __syncthreads();
    // uniform: the work item's descriptors then load into scalar registers
    const int q = __builtin_amdgcn_readfirstlane(task_sh);
    __syncthreads();
    if (q >= P.ntasks) break;
    const int tix = uniform_load(P.order + q);
    const BTask T = uniform_load(P.tasks + tix);
    const BSys S = uniform_load(P.sys + T.sys);
    BSplit X{};
    if constexpr (PREC != ATRIP_FP64) X = uniform_load(P.split + T.sys);
    const int nb = S.nb, nb3 = nb * nb * nb;
    const int64_t No = S.No, N2 = No * No;
    const int fm = bfm(S.tile), fn = bfn(S.tile);
    const int mtiles = (int)((N2 + 16 * fm - 1) / (16 * fm));
    const int ntiles = (int)((nb * No + 16 * fn - 1) / (16 * fn));
    const int nsub = nb * nb * mtiles * ntiles;

#pragma unroll 1
    for (int cl = 0; cl < 3; cl++) {
#pragma unroll 1
      for (int u = wave; u < nsub; u += 4) {
        const int nt = u % ntiles, mt = (u / ntiles) % mtiles;
        const int pair = u / (ntiles * mtiles);
        const int o1 = pair / nb, o2 = pair % nb;
        switch (S.tile) {
#define TILE_CASE(V)                                                       \
  case V:                                                                  \
    if constexpr (PREC == ATRIP_FP64)                                      \
      wave_tile<bfm(V), bfn(V)>(S, cl, T.al, T.be, T.ga, o1, o2, mt, nt,   \
                                Gwin, cl > 0, lane, xpose[wave]);          \
    else                                                                   \
      wave_tile_split<bfm(V), bfn(V), PREC == ATRIP_SPLIT_FP32>(           \
          S, X, cl, T.al, T.be, T.ga, o1, o2, mt, nt, Gwin, cl > 0, lane,  \
          xpose[wave]);                                                    \
    break;
          TILE_CASE(0)
          TILE_CASE(1)
          TILE_CASE(2)
          TILE_CASE(3)
          TILE_CASE(4)
          TILE_CASE(5)
          TILE_CASE(6)
          TILE_CASE(7)
          TILE_CASE(8)
          TILE_CASE(9)
#undef TILE_CASE
          default:
            break;
        }
      }
      __syncthreads();  // class done before the next adds to W
    }

    // energies, one tuple per wave, each wave through its own slots: the
    // singles table tsing[wave] is the wave's own, so its LDS writes need
    // only to complete before the same wave reads them
#pragma unroll 1
    for (int slot = wave; slot < nb3; slot += 4) {
      const int64_t a = (int64_t)T.al * nb + slot / (nb * nb);
      const int64_t b = (int64_t)T.be * nb + (slot / nb) % nb;
      const int64_t c = (int64_t)T.ga * nb + slot % nb;
      const bool ok = tuple_ok(S.Nv, a, b, c);
      double e = 0.0;
      if (ok) {
        if (lane < No) {
          tsing[wave][0][lane] = S.Tai[a + S.Nv * lane];
          tsing[wave][1][lane] = S.Tai[b + S.Nv * lane];
          tsing[wave][2][lane] = S.Tai[c + S.Nv * lane];
        }
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): LDS writes done
        __builtin_amdgcn_wave_barrier();
        e = tuple_energy(S, Gwin + (int64_t)slot * N2 * No, a, b, c,
                         tsing[wave][0], tsing[wave][1], tsing[wave][2],
                         lane);
        __builtin_amdgcn_wave_barrier();  // reads done before the next writes
      }
      if (lane == 0) etup[slot] = e;
    }
    __syncthreads();  // every slot's energy is in etup
    if (P.tuple_e)
      for (int s = tid; s < nb3; s += 256) P.tuple_e[T.tuple_off + s] = etup[s];
    if (tid == 0) {
      // Neumaier over the slots in order
      double sum = 0.0, comp = 0.0;
      for (int s = 0; s < nb3; s++) {
        const double x = etup[s];
        const double t = sum + x;
        comp += (fabs(sum) >= fabs(x)) ? (sum - t) + x : (x - t) + sum;
        sum = t;
      }
      P.task_e[2 * tix] = sum;
      P.task_e[2 * tix + 1] = comp;
    }
    // the next window overwrites W and etup only after the barriers at the
    // top of the loop
  }
}

}  // namespace

hipError_t launch_batched(const BatchedLaunch &L, hipStream_t st) {
  if (L.ntasks <= 0 || L.grid <= 0) return hipSuccess;
  if (L.precision == ATRIP_FP64)
    hipLaunchKernelGGL(batched_kernel<ATRIP_FP64>, dim3(L.grid), dim3(256), 0,
                       st, L);
  else if (L.precision == ATRIP_SPLIT_FP32)
    hipLaunchKernelGGL(batched_kernel<ATRIP_SPLIT_FP32>, dim3(L.grid),
                       dim3(256), 0, st, L);
  else if (L.precision == ATRIP_FP32)
    hipLaunchKernelGGL(batched_kernel<ATRIP_FP32>, dim3(L.grid), dim3(256), 0,
                       st, L);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

int batched_blocks_per_cu(int precision) {
  int n = 0;
  hipError_t e;
  if (precision == ATRIP_SPLIT_FP32)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, batched_kernel<ATRIP_SPLIT_FP32>, 256, 0);
  else if (precision == ATRIP_FP32)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, batched_kernel<ATRIP_FP32>, 256, 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, batched_kernel<ATRIP_FP64>, 256, 0);
  if (e != hipSuccess) return 1;
  return n > 0 ? n : 1;
}

}  // namespace atrip_amd
