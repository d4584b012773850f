// Doubles contractions W(ijk) from three-part bf16 operands (the bf16x3
// mixed-precision form).
//
// Same regrouped class GEMMs and tuple windows as the FP64 kernel
// (doubles_f64.hip).  Every operand x is held as three bf16 parts,
//     b0 = bf16(x),  b1 = bf16(x - b0),  b2 = bf16(x - b0 - b1),
// which carry 24 significant bits, as many as an FP32 value (each part is
// at most 2^-9 of the one before).  A product is formed from the six part
// products down to the 2^-18 level,
//     a0 b0 + (a0 b1 + a1 b0) + (a1 b1 + a0 b2 + a2 b0),
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation; the three dropped
// terms (a1 b2, a2 b1, a2 b2) are at most 2^-27 of the product, below an
// FP32 rounding (2^-24).  For a K chunk of 16, lane group q (lanes 16q..16q+15)
// holds kappa = k0 + 4q + s, s < 4, of its row/column; each MFMA's 8 K
// slots of a lane group take two parts at those four kappa, so its K = 32
// covers two part products of the chunk:
//     [a0, a1] . [b0, b1]  ->  a0 b0 + a1 b1
//     [a1, a0] . [b0, b1]  ->  a1 b0 + a0 b1
//     [a0, a2] . [b2, b0]  ->  a0 b2 + a2 b0
// Three bf16 MFMAs (16 cycles each) per 16-deep chunk against four FP64
// MFMAs (64 cycles each).  Operands are in the tiled layout (TriGeom): parts
// 0/1 of a 16 x 16 fragment chunk are 1 KiB and part 2 is 512 bytes, both
// contiguous in lane order, so a fragment chunk is two fully coalesced
// loads per lane (16 + 8 bytes); the swapped A segments read the
// row-swapped copies, so every A fragment is 16 consecutive rows too.  Class buffers are FP32 and are
// widened to FP64 by the class-sum kernel; singles, the energy and all sums
// stay FP64 (energy.hip).
#include "doubles_common.hpp"

namespace atrip_amd {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// A lane's three parts at its four kappa of a chunk: [b0, b1] and b2.
struct Parts {
  bf8 p01;
  bf4 p2;
};

__device__ __forceinline__ Parts load_parts(const __bf16 *p, int d2) {
  Parts r;
  r.p01 = *(const bf8 *)p;
  r.p2 = *(const bf4 *)(p + d2);
  return r;
}

__device__ __forceinline__ bf8 swap01(bf8 x) {
  return __builtin_shufflevector(x, x, 4, 5, 6, 7, 0, 1, 2, 3);
}
// [x0, y] and [y, x0] for x = [x0, x1]
__device__ __forceinline__ bf8 with2_a(bf8 x, bf4 y) {
  const bf8 yy = __builtin_shufflevector(y, y, 0, 1, 2, 3, 0, 1, 2, 3);
  return __builtin_shufflevector(x, yy, 0, 1, 2, 3, 8, 9, 10, 11);
}
__device__ __forceinline__ bf8 with2_b(bf8 x, bf4 y) {
  const bf8 yy = __builtin_shufflevector(y, y, 0, 1, 2, 3, 0, 1, 2, 3);
  return __builtin_shufflevector(yy, x, 0, 1, 2, 3, 8, 9, 10, 11);
}

struct TriParams {
  // tiled parts; T3s / H3s: the row-swapped copies
  const __bf16 *T3, *T3s, *H3, *H3s, *V3, *T2N, *T2TN;
  const int32_t *windows;
  const uint64_t *aprefix;
  float *G;
  int64_t No, Nv, Kv, Ko, N2, N2p, W3, slots;
  uint64_t first, last;
  int nb, mtiles, ntiles;
  int64_t total, per_xcd;
};

template <int FM, int FN>
__global__ __launch_bounds__(256) void doubles_tri_kernel(TriParams P) {
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

  int32_t ca[FN], cb[FN], cc[FN], cslot[FN], crr[FN];
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
  // the swapped segments read the same rows of the row-swapped copies.
  // Rows past the padded slice (never stored) are folded back into its
  // last tile at the same position mod 16, so every lane of an A fragment
  // sits at its own lane of the tile (L = lane) and reads in bounds.
  const int64_t Kv = P.Kv, Ko = P.Ko, N2p = P.N2p;
  int32_t rowS[FM];
#pragma unroll
  for (int f = 0; f < FM; f++) {
    int64_t m = (int64_t)mt * MT + wm * (MT / 2) + f * 16 + l16;
    if (m >= N2p) m = N2p - 16 + (m & 15);
    rowS[f] = (int32_t)m;
  }

  // a lane's parts of chunk k0 start at tri_pos(R, 4 kk, K) + 48 k0 (the
  // next tile of the same rows); part 2 is d2 parts past parts 0/1 (the
  // lane constant 512 - 4 lane for A, per column for B)
  const int dA = 512 - 4 * lane;
  const __bf16 *ap[FM], *bp[FN];
  int32_t dB[FN];
#define SET_PTRS(SG)                                                          \
  {                                                                           \
    const int64_t K_ = (SG) < 2 ? Kv : Ko;                                    \
    const int64_t xa_ = ((SG) == 0 || (SG) == 3) ? x1 : x2;                   \
    const __bf16 *A_ = (SG) == 0   ? P.T3                                     \
                       : (SG) == 1 ? P.T3s                                    \
                       : (SG) == 2 ? P.H3                                     \
                                   : P.H3s;                                   \
    _Pragma("unroll") for (int f = 0; f < FM; f++) ap[f] =                    \
        A_ + tri_pos(xa_ * N2p + rowS[f], 4 * kk, K_).p01;                    \
    _Pragma("unroll") for (int g = 0; g < FN; g++) {                          \
      int64_t u_, v_;                                                         \
      const int wh_ = b_pair(cl, (SG), ca[g], cb[g], cc[g], u_, v_);          \
      const __bf16 *B_ = wh_ == 0 ? P.V3 : wh_ == 1 ? P.T2N : P.T2TN;         \
      const TriPos q_ = tri_pos((u_ + Nv * v_) * No + crr[g], 4 * kk, K_);    \
      bp[g] = B_ + q_.p01;                                                    \
      dB[g] = (int32_t)(q_.p2 - q_.p01);                                      \
    }                                                                         \
  }
#define LOAD(A, B, KOFF)                                                     \
  {                                                                          \
    _Pragma("unroll") for (int f = 0; f < FM; f++) A[f] =                    \
        load_parts(ap[f] + 48 * (KOFF), dA);                                 \
    _Pragma("unroll") for (int g = 0; g < FN; g++) B[g] =                    \
        load_parts(bp[g] + 48 * (KOFF), dB[g]);                              \
    __builtin_amdgcn_sched_barrier(0);                                       \
  }
  // [a0,a1].[b0,b1] + [a1,a0].[b0,b1] + [a0,a2].[b2,b0]
#define COMPUTE(A, B)                                                        \
  {                                                                          \
    bf8 b20_[FN];                                                            \
    _Pragma("unroll") for (int g = 0; g < FN; g++) b20_[g] =                 \
        with2_b(B[g].p01, B[g].p2);                                          \
    _Pragma("unroll") for (int f = 0; f < FM; f++) {                         \
      const bf8 a10_ = swap01(A[f].p01), a02_ = with2_a(A[f].p01, A[f].p2);  \
      _Pragma("unroll") for (int g = 0; g < FN; g++) {                       \
        acc[f][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(                 \
            A[f].p01, B[g].p01, acc[f][g], 0, 0, 0);                         \
        acc[f][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(                 \
            a10_, B[g].p01, acc[f][g], 0, 0, 0);                             \
        acc[f][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(                 \
            a02_, b20_[g], acc[f][g], 0, 0, 0);                              \
      }                                                                      \
    }                                                                        \
  }

  f4 acc[FM][FN];
#pragma unroll
  for (int f = 0; f < FM; f++)
#pragma unroll
    for (int g = 0; g < FN; g++) acc[f][g] = f4{0.f, 0.f, 0.f, 0.f};

  // three operand sets in rotation: two chunks of loads stay in flight
  // behind the MFMAs of the current one.  A chunk is 2 (FM + FN) loads, so
  // three chunks (60 at 5 x 5) fit the 63 of the vector-memory counter and
  // the wait before a chunk's MFMAs covers that chunk's loads only.
  static_assert(6 * (FM + FN) <= 63, "three chunks of loads must fit vmcnt");
  Parts a0[FM], b0[FN], a1[FM], b1[FN], a2[FM], b2[FN];
#define CLAMP(K_) ((K_) < last ? (K_) : last)
#pragma unroll 1
  for (int sg = 0; sg < 4; sg++) {
    SET_PTRS(sg);
    const int64_t last = (sg < 2 ? Kv : Ko) - KCHUNK;
    LOAD(a0, b0, 0);
    LOAD(a1, b1, CLAMP(KCHUNK));
    for (int64_t k0 = 0;; k0 += 3 * KCHUNK) {
      LOAD(a2, b2, CLAMP(k0 + 2 * KCHUNK));
      COMPUTE(a0, b0);
      if (k0 + KCHUNK > last) break;
      LOAD(a0, b0, CLAMP(k0 + 3 * KCHUNK));
      COMPUTE(a1, b1);
      if (k0 + 2 * KCHUNK > last) break;
      LOAD(a1, b1, CLAMP(k0 + 4 * KCHUNK));
      COMPUTE(a2, b2);
      if (k0 + 3 * KCHUNK > last) break;
    }
  }
#undef CLAMP
#undef SET_PTRS
#undef LOAD
#undef COMPUTE

  // 16x16 FP32-output MFMAs: D[i][j] in lane j + 16*(i/4), register i%4
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
  P.N2p = pad16(P.N2);
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
  if (FM == A && FN == B)                                                    \
    hipLaunchKernelGGL((doubles_tri_kernel<A, B>), dim3((unsigned)grid),     \
                       dim3(256), 0, st, P);
  LAUNCH(2, 2) LAUNCH(4, 2) LAUNCH(2, 4) LAUNCH(4, 4) LAUNCH(3, 3)
  LAUNCH(5, 5) LAUNCH(4, 5) LAUNCH(5, 4)
#undef LAUNCH
  return hipGetLastError();
}

}  // namespace atrip_amd
