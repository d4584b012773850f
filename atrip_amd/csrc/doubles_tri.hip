// Doubles contractions W(ijk) from three-part bf16 operands (the bf16x3
// mixed-precision form).
//
// Same regrouped class GEMMs and tuple windows as the FP64 kernel
// (doubles_f64.hip).  Every operand x is held as three bf16 parts,
//     b0 = bf16(x),  b1 = bf16(x - b0),  b2 = bf16(x - b0 - b1),
// which carry 24 significant bits, as many as an FP32 value (each part is
// at most 2^-9 of the one before).  A product is formed from the six part
// products down to the 2^-18 level,
//     a0 b0 + a0 b1 + a1 b0 + a1 b1 + a0 b2 + a2 b0,
// on v_mfma_f32_32x32x16_bf16 with FP32 accumulation; the three dropped
// terms (a1 b2, a2 b1, a2 b2) are at most 2^-27 of the product, below an
// FP32 rounding (2^-24).  Each MFMA takes one part of a 32-row A fragment
// and one part of a 32-column B fragment over a whole 16-deep K chunk (lane
// l feeds row/column l % 32 at kappa k0 + 8 (l / 32) + 0..7), so a chunk is
// six MFMAs per fragment pair (6 x 32 cycles per 32 x 32 block, against
// 16 x 64 cycles in FP64) and the operands are the loaded registers as
// they are.  Operands are in the tiled layout (TriGeom): each part of a
// 32 x 16 fragment chunk is 1 KiB contiguous in lane order, one 16-byte load
// per lane; the swapped A segments read the row-swapped copies, so every A
// fragment is 32 consecutive rows.  Class buffers are FP32 and are widened
// to FP64 by the class-sum kernel; singles, the energy and all sums stay
// FP64 (energy.hip).
#include "doubles_common.hpp"

namespace atrip_amd {
namespace {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// A lane's three parts at its eight kappa of a chunk.
struct Parts {
  bf8 p0, p1, p2;
};

__device__ __forceinline__ Parts load_parts(const __bf16 *p) {
  Parts r;
  r.p0 = *(const bf8 *)p;
  r.p1 = *(const bf8 *)(p + 512);
  r.p2 = *(const bf8 *)(p + 1024);
  return r;
}

struct TriParams {
  // tiled parts; T3s / H3s: the row-swapped copies
  const __bf16 *T3, *T3s, *H3, *H3s, *V3, *T2N, *T2TN;
  const int32_t *windows;
  const uint64_t *aprefix;
  float *G;
  int64_t No, Nv, Kv, Ko, N2, N2p, W3, NoP, slots;
  uint64_t first, last;
  int nb, mtiles, ntiles;
  int64_t total, per_xcd;
};

template <int FM, int FN>
__global__ __launch_bounds__(256) void doubles_tri_kernel(TriParams P) {
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

  int32_t ca[FN], cb[FN], cc[FN], cslot[FN], crr[FN];
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
  // the swapped segments read the same rows of the row-swapped copies.
  // Rows past the padded slice (never stored) are folded back into its
  // last tile at the same position mod 32, so every lane of an A fragment
  // sits at its own lane slot of the tile and reads in bounds.
  const int64_t Kv = P.Kv, Ko = P.Ko, N2p = P.N2p;
  int32_t rowS[FM];
#pragma unroll
  for (int f = 0; f < FM; f++) {
    int64_t m = (int64_t)mt * MT + wm * (MT / 2) + f * 32 + l32;
    if (m >= N2p) m = N2p - 32 + (m & 31);
    rowS[f] = (int32_t)m;
  }

  // a lane's parts of chunk k0 start at its slot of the tile of k0: the
  // tile of k0 = 0 plus 96 k0 (1536 parts per 16-deep chunk); A lanes sit
  // at slot lane (aligned rows), B lanes at the slot of their row
  const __bf16 *ap[FM], *bp[FN];
#define SET_PTRS(SG)                                                          \
  {                                                                           \
    const int64_t K_ = (SG) < 2 ? Kv : Ko;                                    \
    const int64_t xa_ = ((SG) == 0 || (SG) == 3) ? x1 : x2;                   \
    const __bf16 *A_ = (SG) == 0   ? P.T3                                     \
                       : (SG) == 1 ? P.T3s                                    \
                       : (SG) == 2 ? P.H3                                     \
                                   : P.H3s;                                   \
    _Pragma("unroll") for (int f = 0; f < FM; f++) ap[f] =                    \
        A_ + tri_tile(xa_ * N2p + rowS[f], 0, K_) + 8 * lane;                 \
    _Pragma("unroll") for (int g = 0; g < FN; g++) {                          \
      int64_t u_, v_;                                                         \
      const int wh_ = b_pair(cl, (SG), ca[g], cb[g], cc[g], u_, v_);          \
      const __bf16 *B_ = wh_ == 0 ? P.V3 : wh_ == 1 ? P.T2N : P.T2TN;         \
      const int64_t R_ = (u_ + Nv * v_) * No + crr[g];                        \
      bp[g] = B_ + tri_tile(R_, 0, K_) + 8 * tri_lane(R_, 8 * kh);            \
    }                                                                         \
  }
#define LOAD(A, B, KOFF)                                                     \
  {                                                                          \
    _Pragma("unroll") for (int f = 0; f < FM; f++) A[f] =                    \
        load_parts(ap[f] + 96 * (KOFF));                                     \
    _Pragma("unroll") for (int g = 0; g < FN; g++) B[g] =                    \
        load_parts(bp[g] + 96 * (KOFF));                                     \
    __builtin_amdgcn_sched_barrier(0);                                       \
  }
  // a0 b0, a0 b1, a1 b0, a1 b1, a0 b2, a2 b0: each term over all blocks
  // before the next, so consecutive MFMAs update different accumulators
#define TERM(A, B, AP, BP)                                                   \
  _Pragma("unroll") for (int f = 0; f < FM; f++)                             \
      _Pragma("unroll") for (int g = 0; g < FN; g++) acc[f][g] =             \
          __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[f].AP, B[g].BP,          \
                                                  acc[f][g], 0, 0, 0);
#define COMPUTE(A, B)                                                        \
  {                                                                          \
    TERM(A, B, p0, p0) TERM(A, B, p0, p1) TERM(A, B, p1, p0)                 \
    TERM(A, B, p1, p1) TERM(A, B, p0, p2) TERM(A, B, p2, p0)                 \
  }

  f16v acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; f++)
#pragma unroll
    for (int g = 0; g < FN; g++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[f][g][r] = 0.f;

  // two operand sets in ping-pong: the loads of chunk c + 1 are in flight
  // behind the MFMAs of chunk c (a chunk is 6 FM FN MFMAs of 32 cycles, at
  // least 24 x 32 cycles; a third set would cost the registers the largest
  // tiles need and measured slower).  One pipeline runs through the four
  // segments: (sg, c) is the next chunk to load, and past a segment's end
  // the pointers move to the next segment's chunk 0 in place, so every
  // segment but the first starts with its first chunk in flight (2.7 %
  // faster at No=60 than a pipeline restarted per segment).  After the last
  // chunk the prefetch re-reads it (a load never used).
  static_assert(6 * (FM + FN) <= 63, "two chunks of loads must fit vmcnt");
  Parts a0[FM], b0[FN], a1[FM], b1[FN];
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
#undef TERM
#undef COMPUTE

  // 32x32 FP32-output MFMAs: D[i][j] in lane j + 32 ((i / 4) % 2), register
  // 4 (i / 8) + i % 4 (measured on gfx950)
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

// FP64 operand -> its tiled bf16x3 copy (and row-swapped copy).
__global__ __launch_bounds__(256) void tri_kernel(const double *x, TriOp out,
                                                  TriGeom g, int64_t No,
                                                  int64_t n) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    tri_store(x[q], out, g, No, q);
}

}  // namespace

hipError_t launch_tri(const double *x, const TriOp &out, const TriGeom &g,
                      int64_t No, hipStream_t st) {
  const int64_t n = g.no * g.nr * g.K;
  if (n <= 0) return hipSuccess;
  if (g.K % 16 || (g.swapped && g.nr != No * No)) return hipErrorInvalidValue;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536L * 16) blocks = 65536L * 16;
  hipLaunchKernelGGL(tri_kernel, dim3((unsigned)blocks), dim3(256), 0, st, x,
                     out, g, No, n);
  return hipGetLastError();
}

hipError_t launch_doubles_tri(const DevSystem &s, const BatchArgs &b,
                              bool use_j, float *G, int tile, hipStream_t st) {
  TriParams P;
  const TriOp &H = use_j ? s.tH3J : s.tH3;
  P.T3 = (const __bf16 *)s.tT3.p;
  P.T3s = (const __bf16 *)s.tT3.sw;
  P.H3 = (const __bf16 *)H.p;
  P.H3s = (const __bf16 *)H.sw;
  P.V3 = (const __bf16 *)(use_j ? s.tV3J.p : s.tV3.p);
  P.T2N = (const __bf16 *)s.tT2N.p;
  P.T2TN = (const __bf16 *)s.tT2TN.p;
  P.windows = b.windows;
  P.aprefix = s.aprefix;
  P.G = G;
  P.No = s.No;
  P.Nv = s.Nv;
  P.Kv = s.Kv;
  P.Ko = s.Ko;
  P.N2 = s.No * s.No;
  P.N2p = pad32(P.N2);
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
  if (FM == A && FN == B)                                                    \
    hipLaunchKernelGGL((doubles_tri_kernel<A, B>), dim3((unsigned)grid),     \
                       dim3(256), 0, st, P);
  LAUNCH(2, 2) LAUNCH(2, 3) LAUNCH(3, 2) LAUNCH(2, 4) LAUNCH(3, 3)
  LAUNCH(2, 5) LAUNCH(3, 5) LAUNCH(4, 4) LAUNCH(3, 4) LAUNCH(4, 3)
#undef LAUNCH
  return hipGetLastError();
}

}  // namespace atrip_amd
