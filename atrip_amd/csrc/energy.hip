// Fused singles + energy epilogue of the (T) hot path.
//
// wsum_kernel first sums the three class buffers the doubles kernel wrote
// (each in its GEMM output order, see internal.hpp; FP64, or FP32 for the
// single-precision-product forms) into W in FP64, (w + v) + u, re-laid out
// as 8x8x8 cubes (w_cube).  It works on 16x16x16 bricks staged in LDS, so
// every class is read along its own fastest index in full 128-byte lines and
// every cube is written as one 4 KiB run.
//
// energy_kernel: one workgroup per tuple slot of a window batch.  Z = W +
// singles is formed in registers (reference src/atrip/Equations.cxx:387-426, with Z copied from W
// as in src/atrip/Atrip.cxx:899-906), and the energy of the tuple follows
// get_energy_distinct / get_energy_same (src/atrip/Equations.cxx:101-238)
// selected as src/atrip/Atrip.cxx:636-670 does.  The i>=j>=k points are
// spread over the workgroup; each thread sums its points in a fixed order
// and a fixed shuffle/LDS tree combines the threads, so the per-tuple energy
// does not depend on which other tuples share the launch.
#include <type_traits>

#include "internal.hpp"

namespace atrip_amd {
namespace {

struct EnergyParams {
  const double *W, *WJ;  // W (and the (cT) pass's W) as cubes, [slot][W3]
  const double *Tai, *VHH, *ei, *ea;
  const int32_t *windows;
  const uint64_t *aprefix;
  double *e_slot;
  int64_t No, Nv, N2, W3, nbk, slots;  // nbk: cubes per index, W3 = (8 nbk)^3
  uint64_t first, last;
  int nb;
  double tai_sign;
};

__device__ inline double wave_sum(double v) {
  // fixed xor butterfly over the 64 lanes
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Class sums of one 16x16x16 brick of W for every slot of the batch:
// out = sum(A) [+ sum(B)], sum(X) = (Xw + Xv) + Xu, written as cubes.
template <typename GT>
__global__ __launch_bounds__(256) void wsum_kernel(const GT *A, const GT *B,
                                                   double *out, int64_t No,
                                                   int64_t NoP, int64_t W3g,
                                                   int64_t slots, int nbk,
                                                   int nbr) {
  // LDS brick [z][y][x] with pitch 17 (conflict-free transposed writes)
  __shared__ double br[16 * 17 * 17];
  const int64_t bid = blockIdx.x;
  const int64_t slot = bid / ((int64_t)nbr * nbr * nbr);
  const int brick = (int)(bid % ((int64_t)nbr * nbr * nbr));
  const int bx = brick % nbr, by = (brick / nbr) % nbr, bz = brick / (nbr * nbr);
  const int tid = threadIdx.x, no = (int)No;
  const int x0 = 16 * bx, y0 = 16 * by, z0 = 16 * bz;
  auto at = [](int x, int y, int z) { return x + 17 * (y + 17 * z); };
  double *o = out + slot * (int64_t)(512 * nbk * nbk * nbk);
  for (int set = 0; set < (B ? 2 : 1); set++) {
    const GT *S = set == 0 ? A : B;
    const GT *Sw = S + slot * W3g, *Sv = S + (slots + slot) * W3g,
             *Su = S + (2 * slots + slot) * W3g;
    GT v[16], t[16];  // widened to FP64 only when added into the brick
    // class w: (x + No y) No + z, z fastest
#pragma unroll
    for (int it = 0; it < 16; it++) {
      const int e = tid + 256 * it;
      const int z = z0 + (e & 15), y = y0 + ((e >> 4) & 15), x = x0 + (e >> 8);
      const bool in = x < no && y < no && z < no;
      v[it] = Sw[in ? (x + no * y) * NoP + z : 0];  // always a valid address
      if (!in) v[it] = (GT)0;
    }
    // class v: (x + No z) No + y, y fastest
#pragma unroll
    for (int it = 0; it < 16; it++) {
      const int e = tid + 256 * it;
      const int y = y0 + (e & 15), x = x0 + ((e >> 4) & 15), z = z0 + (e >> 8);
      const bool in = x < no && y < no && z < no;
      t[it] = Sv[in ? (x + no * z) * NoP + y : 0];
      if (!in) t[it] = (GT)0;
    }
    __syncthreads();  // the previous set's cube writes are done with br
#pragma unroll
    for (int it = 0; it < 16; it++) {
      const int e = tid + 256 * it;
      br[at(e >> 8, (e >> 4) & 15, e & 15)] = (double)v[it];
    }
    // class u: (y + No z) No + x, x fastest
#pragma unroll
    for (int it = 0; it < 16; it++) {
      const int e = tid + 256 * it;
      const int x = x0 + (e & 15), y = y0 + ((e >> 4) & 15), z = z0 + (e >> 8);
      const bool in = x < no && y < no && z < no;
      v[it] = Su[in ? (y + no * z) * NoP + x : 0];
      if (!in) v[it] = (GT)0;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 16; it++) {
      const int e = tid + 256 * it;
      br[at((e >> 4) & 15, e & 15, e >> 8)] += (double)t[it];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 16; it++) {
      const int e = tid + 256 * it;
      br[at(e & 15, (e >> 4) & 15, e >> 8)] += (double)v[it];
    }
    __syncthreads();
    // the brick's cubes, 512 contiguous doubles each; the second set adds
    // to what the same thread wrote for the first
#pragma unroll
    for (int it = 0; it < 16; it++) {
      const int e = tid + 256 * it;
      const int cube = e >> 9, q = e & 511;
      const int cx = cube & 1, cy = (cube >> 1) & 1, cz = cube >> 2;
      const int gx = 2 * bx + cx, gy = 2 * by + cy, gz = 2 * bz + cz;
      if (gx >= nbk || gy >= nbk || gz >= nbk) continue;
      const double val = br[at(8 * cx + (q & 7), 8 * cy + ((q >> 3) & 7), 8 * cz + (q >> 6))];
      double *dst = o + ((int64_t)((gz * nbk + gy) * nbk + gx) << 9) + q;
      *dst = set == 0 ? val : *dst + val;
    }
  }
}

// Energy of one tuple per workgroup.  The i>=j>=k points are visited per
// block triple (B1>=B2>=B3) of edge 8: the (up to) six distinct cubes that
// hold W at the six index permutations of those points are copied into LDS
// with coalesced 4 KiB reads, then every point reads its twelve W/Z values
// from LDS.
// The orbit of block triple (B1, B2, B3) under the six index permutations
// (ijk ikj jik jki kij kji): base offsets cb[s] of its nc distinct cubes and
// the slot al[p] of the cube of each permutation.
__device__ __forceinline__ void orbit(int B1, int B2, int B3, int nbk,
                                      int (&cb)[6], int &nc, int (&al)[6]) {
  const int cx[6] = {B1, B1, B2, B2, B3, B3};
  const int cy[6] = {B2, B3, B1, B3, B1, B2};
  const int cz[6] = {B3, B2, B3, B1, B2, B1};
  nc = 0;
#pragma unroll
  for (int p = 0; p < 6; p++) {
    int q0 = p;
#pragma unroll
    for (int q = 0; q < p; q++)
      if (q0 == p && cx[q] == cx[p] && cy[q] == cy[p] && cz[q] == cz[p]) q0 = q;
    if (q0 == p) {
      al[p] = nc;
      cb[nc < 6 ? nc : 5] = ((cz[p] * nbk + cy[p]) * nbk + cx[p]) << 9;
      nc++;
    } else {
      al[p] = al[q0];
    }
  }
}

// Energy terms of the i>=j>=k points of block triple (B1, B2, B3) (edge 8)
// of one tuple, two points per thread, added to e_acc (and, with the (cT)
// W, to ect_acc).  The six permuted W cubes of the triple's orbit sit in
// LDS (cube[alias[p]] holds W at permutation p, element x + 8 y + CZ z), as
// do the singles tables tA/tB/tC (Tai at (a,x), (b,x), (c,x), slot 8 * position
// + local index), the eps_i table sE and the V^bc, V^ac, V^ab 8x8 block pairs
// sV.  Z = W + singles (reference src/atrip/Equations.cxx:387-426) and the
// get_energy_distinct / get_energy_same formulas (:101-238) in the
// reference's expression order.
// sV holds the 8x8 V blocks of the triple, [slice][block pair][ly + 8 lz]:
// all nine ordered position pairs (V6 false) or only the six with two
// different positions (V6 true), the only ones a point reads.
template <bool HJ, int CZ, int CSZ, bool V6>
__device__ __forceinline__ void triple_points(
    int B1, int B2, int B3, int no, int distinct, double epsabc,
    const int (&alias)[6], const double (*cube)[CSZ], const double *cubeJ,
    const double *tA, const double *tB, const double *tC, const double *sE,
    const double *sV, int tid, double &e_acc, double &ect_acc) {
  auto vb = [](int v, int r, int s2, int q) {
    return V6 ? (v * 6 + 2 * r + s2 - (s2 > r ? 1 : 0)) * 64 + q
              : ((v * 3 + r) * 3 + s2) * 64 + q;
  };
  // both points of a thread in one body, so their LDS reads and FP64
  // chains interleave
#pragma unroll
  for (int pt = tid; pt < 512; pt += 256) {
    const int li = pt & 7, lj = (pt >> 3) & 7, lk = pt >> 6;
    const int i = B1 * 8 + li, j = B2 * 8 + lj, k = B3 * 8 + lk;
    if (i >= no || j >= no || k >= no || i < j || j < k) continue;
    const double ei = sE[li], ej = sE[8 + lj], ek = sE[16 + lk];
    const double facjk = j == k ? 0.5 : 1.0;
    const double facij = i == j ? 0.5 : 1.0;
    const double eijk = (ei + ej) + ek;
    const double denominator = epsabc - eijk;
    // local offsets of the permutations inside their cubes
    const int o_ijk = li + 8 * lj + CZ * lk, o_ikj = li + 8 * lk + CZ * lj,
              o_jik = lj + 8 * li + CZ * lk, o_jki = lj + 8 * lk + CZ * li,
              o_kij = lk + 8 * li + CZ * lj, o_kji = lk + 8 * lj + CZ * li;
    // singles at a permutation (x,y,z): t_a[x] V^bc_yz + t_b[y] V^ac_xz
    // + t_c[z] V^ab_xy; slot s = 8 * position + local index (position 0:
    // i, 1: j, 2: k)
    auto Z = [&](double wv, int sx, int sy, int sz) {
      const int px = sx >> 3, py = sy >> 3, pz = sz >> 3;
      const int lx = sx & 7, ly = sy & 7, lz = sz & 7;
      double zz = wv + tA[sx] * sV[vb(0, py, pz, ly + 8 * lz)];
      zz = zz + tB[sy] * sV[vb(1, px, pz, lx + 8 * lz)];
      zz = zz + tC[sz] * sV[vb(2, px, py, lx + 8 * ly)];
      return zz;
    };
    auto J = [&](int p, int o) { return cubeJ[alias[p] * CSZ + o]; };
    const int si = li, sj = 8 + lj, sk = 16 + lk;
    if (distinct == 0) {
      const double A = cube[alias[0]][o_ijk], B = cube[alias[1]][o_ikj],
                   C = cube[alias[2]][o_jik], D = cube[alias[3]][o_jki],
                   E = cube[alias[4]][o_kij], F = cube[alias[5]][o_kji];
      const double U = Z(A, si, sj, sk), V = Z(B, si, sk, sj),
                   W = Z(C, sj, si, sk), X = Z(D, sj, sk, si),
                   Y = Z(E, sk, si, sj), Zz = Z(F, sk, sj, si);
      const double UXY = U + (X + Y), VWZ = V + (W + Zz);
      {
        const double first =
            A * U + (B * V + (C * W + (D * X + (E * Y + F * Zz))));
        const double second = (UXY - 2.0 * VWZ) * (A + (D + E));
        const double third = (VWZ - 2.0 * UXY) * (B + (C + F));
        const double value = 3.0 * first + (second + third);
        e_acc += ((2.0 * value) / denominator) * (facjk * facij);
      }
      if constexpr (HJ) {
        const double A2 = J(0, o_ijk), B2 = J(1, o_ikj), C2 = J(2, o_jik),
                     D2 = J(3, o_jki), E2 = J(4, o_kij), F2 = J(5, o_kji);
        const double first =
            A2 * U + (B2 * V + (C2 * W + (D2 * X + (E2 * Y + F2 * Zz))));
        const double second = (UXY - 2.0 * VWZ) * (A2 + (D2 + E2));
        const double third = (VWZ - 2.0 * UXY) * (B2 + (C2 + F2));
        const double value = 3.0 * first + (second + third);
        ect_acc += ((2.0 * value) / denominator) * (facjk * facij);
      }
    } else {
      const double A = cube[alias[0]][o_ijk], B = cube[alias[3]][o_jki],
                   C = cube[alias[4]][o_kij];
      const double U = Z(A, si, sj, sk), V = Z(B, sj, sk, si),
                   W = Z(C, sk, si, sj);
      const double UVW = U + (V + W);
      {
        const double value =
            3.0 * ((A * U + B * V) + C * W) - (A + (B + C)) * UVW;
        e_acc += ((2.0 * value) / denominator) * (facjk * facij);
      }
      if constexpr (HJ) {
        const double A2 = J(0, o_ijk), B2 = J(3, o_jki), C2 = J(4, o_kij);
        const double value =
            3.0 * ((A2 * U + B2 * V) + C2 * W) - (A2 + (B2 + C2)) * UVW;
        ect_acc += ((2.0 * value) / denominator) * (facjk * facij);
      }
    }
  }
}

template <bool HJ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
void energy_kernel(EnergyParams P) {
  const int64_t gslot = blockIdx.x;
  const int nb = P.nb;
  const int64_t nb3 = (int64_t)nb * nb * nb;
  const int64_t w = gslot / nb3, s = gslot % nb3;
  const int64_t ta = s / (nb * nb), tb = (s / nb) % nb, tc = s % nb;
  const int64_t a = P.windows[3 * w] * (int64_t)nb + ta;
  const int64_t b = P.windows[3 * w + 1] * (int64_t)nb + tb;
  const int64_t c = P.windows[3 * w + 2] * (int64_t)nb + tc;
  const int64_t No = P.No, Nv = P.Nv, N2 = P.N2;
  bool ok = tuple_ok(Nv, a, b, c);
  if (ok && (P.first > 0 || P.last != ~0ULL)) {
    const uint64_t idx = tuple_index(P.aprefix, Nv, a, b, c);
    ok = idx >= P.first && idx < P.last;
  }
  if (!ok) {
    if (threadIdx.x == 0) {
      P.e_slot[2 * gslot] = 0.0;
      P.e_slot[2 * gslot + 1] = 0.0;
    }
    return;
  }
  constexpr bool have_j = HJ;

  // cube element x + 8 y + 64 z at x + 8 y + CZ z: the six permutation
  // reads of a wave (lanes over two of x, y, z) are then at most 2-way bank
  // conflicts instead of 8-way
  constexpr int CZ = 69, CSZ = 7 * CZ + 64;
  __shared__ double cube[6][CSZ];
  __shared__ double cubeJ[HJ ? 6 : 1][HJ ? CSZ : 1];
  __shared__ double tA[64], tB[64], tC[64];  // No <= 64 handled in tiles below
  __shared__ double sE[24];                  // eps_i at the triple's blocks
  // V^bc, V^ac, V^ab on the 8x8 block pairs (r, s) of the current triple
  // (positions r != s of blocks B1, B2, B3): sV[slice][r][s][ly + 8 lz]
  __shared__ double sV[3][3][3][64];
  __shared__ double red[2][4];

  const double *Vbc = P.VHH + (b + Nv * c) * N2;
  const double *Vac = P.VHH + (a + Nv * c) * N2;
  const double *Vab = P.VHH + (a + Nv * b) * N2;
  const double *Ws = P.W + gslot * P.W3;
  const double *Js = have_j ? P.WJ + gslot * P.W3 : Ws;

  int distinct = 0;
  if (a == b) distinct++;
  if (b == c) distinct--;
  const double epsabc = (P.ea[a] + P.ea[b]) + P.ea[c];
  const int nbk = (int)P.nbk;
  const int no = (int)No;
  const int tid = threadIdx.x;

  double e_acc = 0.0, ect_acc = 0.0;
  // Block triples B1 >= B2 >= B3 in nested order (B1 fastest).  The loads
  // of the next triple (cubes, V blocks, t and eps_i tables) are issued into
  // registers before the points of the current one are computed and written
  // to LDS after them, so each triple's memory latency hides behind the
  // previous triple's arithmetic.  Loads are unconditional (the last triple
  // fetches itself again) so none waits behind a branch, and the point loop
  // reads LDS only.
  double tt[12];             // t at (a,x), (b,x), (c,x) and eps_i, x in B1,B2,B3
  double vv[7];              // V^bc, V^ac, V^ab block pairs
  double vw[12];             // cubes of the orbit (2 elements per cube)
  double vwJ[HJ ? 12 : 1];
  int ncur = 0, nnext = 0;
#define FETCH(F1, F2, F3, NC)                                                  \
  {                                                                            \
    const int t_ = tid & 7;                                                    \
    const int xi_ = (F1) * 8 + t_, xj_ = (F2) * 8 + t_, xk_ = (F3) * 8 + t_;   \
    const int ci_ = xi_ < no ? xi_ : 0, cj_ = xj_ < no ? xj_ : 0,              \
              ck_ = xk_ < no ? xk_ : 0;                                        \
    tt[0] = P.Tai[a + Nv * ci_]; tt[1] = P.Tai[a + Nv * cj_];                  \
    tt[2] = P.Tai[a + Nv * ck_]; tt[3] = P.Tai[b + Nv * ci_];                  \
    tt[4] = P.Tai[b + Nv * cj_]; tt[5] = P.Tai[b + Nv * ck_];                  \
    tt[6] = P.Tai[c + Nv * ci_]; tt[7] = P.Tai[c + Nv * cj_];                  \
    tt[8] = P.Tai[c + Nv * ck_]; tt[9] = P.ei[ci_];                            \
    tt[10] = P.ei[cj_]; tt[11] = P.ei[ck_];                                    \
    const int Bp_[3] = {(F1), (F2), (F3)};                                     \
    _Pragma("unroll") for (int it = 0; it < 7; it++) {                         \
      const int e = tid + 256 * it;                                            \
      const int v = e / 576, rs = (e / 64) % 9, l = e & 63;                    \
      const int r = rs / 3, sp = rs % 3;                                       \
      const int y = Bp_[r] * 8 + (l & 7), z = Bp_[sp] * 8 + (l >> 3);          \
      const double *V = v == 0 ? Vbc : v == 1 ? Vac : Vab;                     \
      const bool ok_ = e < 3 * 9 * 64 && r != sp && y < No && z < No;          \
      const double val_ = V[ok_ ? y + No * z : 0];                             \
      vv[it] = ok_ ? val_ : 0.0;                                               \
    }                                                                          \
    int cb_[6], al_[6];                                                        \
    orbit((F1), (F2), (F3), nbk, cb_, NC, al_);                                \
    _Pragma("unroll") for (int it = 0; it < 12; it++) {                        \
      const int p_ = it >> 1, q_ = tid + 256 * (it & 1);                       \
      vw[it] = Ws[cb_[p_ < (NC) ? p_ : 0] + q_];                               \
      if constexpr (HJ) vwJ[it] = Js[cb_[p_ < (NC) ? p_ : 0] + q_];            \
    }                                                                          \
  }
#define COMMIT(C1, C2, C3, NC)                                                 \
  {                                                                            \
    if (tid < 8) {                                                             \
      const int ii_ = (C1) * 8 + tid, jj_ = (C2) * 8 + tid,                    \
                kk_ = (C3) * 8 + tid;                                          \
      tA[tid] = ii_ < No ? P.tai_sign * tt[0] : 0.0;                           \
      tA[8 + tid] = jj_ < No ? P.tai_sign * tt[1] : 0.0;                       \
      tA[16 + tid] = kk_ < No ? P.tai_sign * tt[2] : 0.0;                      \
      tB[tid] = ii_ < No ? P.tai_sign * tt[3] : 0.0;                           \
      tB[8 + tid] = jj_ < No ? P.tai_sign * tt[4] : 0.0;                       \
      tB[16 + tid] = kk_ < No ? P.tai_sign * tt[5] : 0.0;                      \
      tC[tid] = ii_ < No ? P.tai_sign * tt[6] : 0.0;                           \
      tC[8 + tid] = jj_ < No ? P.tai_sign * tt[7] : 0.0;                       \
      tC[16 + tid] = kk_ < No ? P.tai_sign * tt[8] : 0.0;                      \
      sE[tid] = tt[9];                                                         \
      sE[8 + tid] = tt[10];                                                    \
      sE[16 + tid] = tt[11];                                                   \
    }                                                                          \
    _Pragma("unroll") for (int it = 0; it < 7; it++) {                         \
      const int e = tid + 256 * it;                                            \
      if (e < 3 * 9 * 64) (&sV[0][0][0][0])[e] = vv[it];                       \
    }                                                                          \
    _Pragma("unroll") for (int it = 0; it < 12; it++) {                        \
      const int p_ = it >> 1, q_ = tid + 256 * (it & 1);                       \
      if (p_ < (NC)) {                                                         \
        const int at_ = (q_ & 63) + CZ * (q_ >> 6);                            \
        cube[p_][at_] = vw[it];                                                \
        if constexpr (HJ) cubeJ[p_][at_] = vwJ[it];                            \
      }                                                                        \
    }                                                                          \
  }
  int B1 = 0, B2 = 0, B3 = 0;
  FETCH(0, 0, 0, ncur);
#pragma unroll 1
  for (;;) {
    __syncthreads();  // the previous triple's points are done with the LDS
    COMMIT(B1, B2, B3, ncur);
    __syncthreads();
    int N1 = B1 + 1, N2b = B2, N3 = B3;
    if (N1 == nbk) {
      N2b++;
      if (N2b == nbk) {
        N3++;
        N2b = N3;
      }
      N1 = N2b;
    }
    const bool more = N3 < nbk;
    if (!more) {
      N1 = B1;
      N2b = B2;
      N3 = B3;
    }
    FETCH(N1, N2b, N3, nnext);
    int alias[6];
    {
      int cb_[6], nc_;
      orbit(B1, B2, B3, nbk, cb_, nc_, alias);
    }
    triple_points<HJ, CZ, CSZ, false>(B1, B2, B3, no, distinct, epsabc,
                                      alias, cube, &cubeJ[0][0], tA, tB, tC,
                                      sE, &sV[0][0][0][0], tid, e_acc,
                                      ect_acc);
    if (!more) break;
    B1 = N1;
    B2 = N2b;
    B3 = N3;
    ncur = nnext;
  }
#undef FETCH
#undef COMMIT

  e_acc = wave_sum(e_acc);
  ect_acc = wave_sum(ect_acc);
  const int wave = tid >> 6, lane = tid & 63;
  if (lane == 0) {
    red[0][wave] = e_acc;
    red[1][wave] = ect_acc;
  }
  __syncthreads();
  if (tid == 0) {
    const double e = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    const double ect = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    P.e_slot[2 * gslot] = e;
    P.e_slot[2 * gslot + 1] = have_j ? ect : e;
  }
}

// Fused class sums + singles + energy: one workgroup per (tuple slot, block
// triple B1 >= B2 >= B3 of edge 8).  The workgroup gathers the (up to) six
// permuted W cubes of the triple's orbit straight from the three class
// buffers the doubles kernel wrote (each read along its own fastest index
// in 64-byte row pieces) and sums them in LDS in the fixed order
// (w + v) + u, the order of wsum_kernel, so W is the same bits as in the
// cube path; then the triple's points follow triple_points, and the
// workgroup's energy (a fixed shuffle/LDS tree) goes to its own entry of
// e_part.  triple_sum_kernel adds a tuple's triples in order.  Nothing
// about W leaves the chip: no cube is written to memory and read back.
struct FusedParams {
  const void *G, *GJ;  // class buffers [3][slots][No^2 rows][NoP] (FP64 or FP32)
  const double *Tai, *VHH, *ei, *ea;
  const int32_t *windows;
  const uint64_t *aprefix;
  double *e_part;  // [slot][triple][2]
  int64_t No, Nv, N2, NoP, W3g, slots, ntri;
  uint64_t first, last;
  int nb, nbk;
  double tai_sign;
};

// Triple t of the nested order B3 outer, B2 = B3.., B1 = B2.. innermost
// (the order the per-tuple kernel walks).
__device__ __forceinline__ void triple_of(int t, int nbk, int &B1, int &B2,
                                          int &B3) {
  int r = t;
  B3 = 0;
  for (;;) {
    const int m = nbk - B3, cnt = m * (m + 1) / 2;
    if (r < cnt) break;
    r -= cnt;
    B3++;
  }
  B2 = B3;
  for (;;) {
    const int cnt = nbk - B2;
    if (r < cnt) break;
    r -= cnt;
    B2++;
  }
  B1 = B2 + r;
}

// The loads, class phases and points of one block triple's orbit, whose
// NC distinct cubes (6, 3 or 1) are qc[0 .. NC - 1]: one compile-time
// instance per NC, so every load of an instance is issued before any is
// waited on and none is spent on a repeated cube.
template <typename GT, bool HJ, int NC, int CZ, int CSZ>
__device__ __forceinline__ void fused_orbit(
    const FusedParams &P, int64_t gid, int64_t gslot, int tid, int no,
    int64_t a, int64_t b, int64_t c, int B1, int B2, int B3, int distinct,
    double epsabc, const int (&qc)[6], const int (&alias)[6],
    double (*cube)[CSZ], double (*cubeJ)[HJ ? CSZ : 1], double *tA,
    double *tB, double *tC, double *sE, double *sV, double (*red)[4]) {
  const int64_t No = P.No, Nv = P.Nv, N2 = P.N2;
  // element e = tid + 256 h of a cube, per class: row e >> 3 of the class's
  // two slower indices, f = e & 7 along its fastest index
  const GT *G = (const GT *)P.G;
  const GT *GJ = (const GT *)P.GJ;
  const int64_t W3g = P.W3g, slots = P.slots;
  const int nop = (int)P.NoP;  // 32-bit: offsets use the 32x32->64 multiply-add
  const int64_t bw = gslot * W3g, bv = (slots + gslot) * W3g,
                bu = (2 * slots + gslot) * W3g;
  // element (q, h) of class cls for this thread: its offset inside the
  // slot's class buffer (0 outside the occupied range, with in false) and
  // its LDS position.  Branch-free, and 32-bit: the class base is
  // wave-uniform, so a load is a scalar base plus a 32-bit lane offset
  // (No^2 times the row pitch stays far below 2^32).
  auto elem = [&](int q, int h, int cls, int &pos, bool &in) -> uint32_t {
    const int e = tid + 256 * h, r = e >> 3, f = e & 7;
    const int r0 = r & 7, r1 = r >> 3;
    const int x = cls == 2 ? f : r0;
    const int y = cls == 0 ? r1 : cls == 1 ? f : r0;
    const int z = cls == 0 ? f : r1;
    pos = x + 8 * y + CZ * z;
    const int i = 8 * (qc[q] & 255) + x, j = 8 * ((qc[q] >> 8) & 255) + y,
              k = 8 * (qc[q] >> 16) + z;
    in = (i < no) & (j < no) & (k < no);
    const uint32_t o =
        cls == 0 ? (uint32_t)((i + no * j) * nop + k)      // w: k fastest
      : cls == 1 ? (uint32_t)((i + no * k) * nop + j)      // v: j fastest
                 : (uint32_t)((j + no * k) * nop + i);     // u: i fastest
    return in ? o : 0u;  // always a valid address
  };
  auto gather = [&](const GT *S, int q, int h, int cls) -> double {
    int pos;
    bool in;
    const uint32_t o = elem(q, h, cls, pos, in);
    const GT *Sc = S + (cls == 0 ? bw : cls == 1 ? bv : bu);
    const GT v = Sc[o];
    return in ? (double)v : 0.0;
  };
  // The loads, class phases and points for the orbit's nc distinct cubes
  // (6, 3 or 1), one compile-time instance each, selected by a uniform
  // branch: every load of an instance is issued before any is waited on
  // and none is spent on a repeated cube.
  // Every global load of the workgroup is issued before any is waited on
  // (a load behind a branch is waited on before the branch joins, which
  // serialised the first version's loads): W and the (cT) W of the orbit's
  // cubes, the V^bc, V^ac, V^ab blocks of the triple, and (wave 0) t at
  // (a,x), (b,x), (c,x) and eps_i for x in the blocks B1, B2, B3.
  double vals[HJ ? 2 : 1][3][NC][2];
#pragma unroll
  for (int set = 0; set < (HJ ? 2 : 1); set++)
#pragma unroll
    for (int cls = 0; cls < 3; cls++)
#pragma unroll
      for (int q = 0; q < NC; q++)
#pragma unroll
        for (int h = 0; h < 2; h++)
          vals[set][cls][q][h] = gather(set ? GJ : G, q, h, cls);
  double vv[5];
  {
    const double *Vbc = P.VHH + (b + Nv * c) * N2;
    const double *Vac = P.VHH + (a + Nv * c) * N2;
    const double *Vab = P.VHH + (a + Nv * b) * N2;
#pragma unroll
    for (int it = 0; it < 5; it++) {
      // entry e: slice v, block pair ps = 2 r + (s - [s > r]) of positions
      // r != s, element l = ly + 8 lz
      const int e = tid + 256 * it;
      const int v = e / 384, ps = (e / 64) % 6, l = e & 63;
      const int r = ps >> 1, s2 = (ps & 1) < r ? (ps & 1) : (ps & 1) + 1;
      // (selects, not an array: a run-time-indexed array lives in
      // scratch, and its loads count in vmcnt behind the gathers)
      const int Br = r == 0 ? B1 : r == 1 ? B2 : B3;
      const int Bs = s2 == 0 ? B1 : s2 == 1 ? B2 : B3;
      const int y = Br * 8 + (l & 7), z = Bs * 8 + (l >> 3);
      const double *V = v == 0 ? Vbc : v == 1 ? Vac : Vab;
      const bool ok_ = (e < 3 * 6 * 64) & (y < No) & (z < No);
      const double val = V[ok_ ? y + No * z : 0];
      vv[it] = ok_ ? val : 0.0;
    }
  }
  if (tid < 64) {  // wave 0
    const int l = tid < 24 ? tid : 0;
    const int blk = l >> 3 == 0 ? B1 : l >> 3 == 1 ? B2 : B3;
    const int x = blk * 8 + (l & 7);
    const bool in = tid < 24 && x < no;
    const int cx = x < no ? x : 0;
    const double ta = P.Tai[a + Nv * cx], tb = P.Tai[b + Nv * cx],
                 tc = P.Tai[c + Nv * cx], ex = P.ei[cx];
    tA[tid] = in ? P.tai_sign * ta : 0.0;
    tB[tid] = in ? P.tai_sign * tb : 0.0;
    tC[tid] = in ? P.tai_sign * tc : 0.0;
    sE[tid] = ex;
  }
#pragma unroll
  for (int it = 0; it < 5; it++) sV[tid + 256 * it] = vv[it];
  // the three class phases into LDS in the order w, v, u
#pragma unroll
  for (int cls = 0; cls < 3; cls++) {
#pragma unroll
    for (int q = 0; q < NC; q++)
#pragma unroll
      for (int h = 0; h < 2; h++) {
        int at;
        bool in_;
        (void)elem(q, h, cls, at, in_);
        if (cls == 0) {
          cube[q][at] = vals[0][0][q][h];
          if constexpr (HJ) cubeJ[q][at] = vals[HJ ? 1 : 0][0][q][h];
        } else {
          cube[q][at] += vals[0][cls][q][h];
          if constexpr (HJ) cubeJ[q][at] += vals[HJ ? 1 : 0][cls][q][h];
        }
      }
    __syncthreads();
  }

  double e_acc = 0.0, ect_acc = 0.0;
  triple_points<HJ, CZ, CSZ, true>(B1, B2, B3, no, distinct, epsabc, alias,
                                   cube, &cubeJ[0][0], tA, tB, tC, sE, sV,
                                   tid, e_acc, ect_acc);
  e_acc = wave_sum(e_acc);
  ect_acc = wave_sum(ect_acc);
  const int wave = tid >> 6, lane = tid & 63;
  if (lane == 0) {
    red[0][wave] = e_acc;
    red[1][wave] = ect_acc;
  }
  __syncthreads();
  if (tid == 0) {
    const double e = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    const double ect = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    P.e_part[2 * gid] = e;
    P.e_part[2 * gid + 1] = HJ ? ect : e;
  }
}

template <typename GT, bool HJ>
__global__ __launch_bounds__(256) void energy_fused_kernel(FusedParams P) {
  // workgroups go to the 8 XCDs round-robin: XCD x runs the triples of
  // slots x, x + 8, x + 16, ... one slot after the other, so the triples of
  // a slot (whose row pieces share 128-byte lines) meet in one L2, and the
  // slots stay spread over the XCDs
  // (all of the decode in 32-bit unsigned arithmetic: a 64-bit division is
  // ~130 scalar instructions, and the scalar unit, shared by the CU's
  // waves, was the kernel's bound)
  const uint32_t bid = blockIdx.x, r = bid >> 3, ntri = (uint32_t)P.ntri;
  const uint32_t rq = r / ntri;
  const uint32_t gs = 8 * rq + (bid & 7);
  const int tri = (int)(r - rq * ntri);
  if ((int64_t)gs >= P.slots) return;
  const int64_t gslot = gs;
  const int64_t gid = gslot * P.ntri + tri;
  const int nb = P.nb;
  const uint32_t nb2 = (uint32_t)(nb * nb), nb3 = nb2 * (uint32_t)nb;
  const uint32_t w32 = gs / nb3, s32 = gs - w32 * nb3;
  const uint32_t ta32 = s32 / nb2, r2 = s32 - ta32 * nb2;
  const uint32_t tb32 = r2 / (uint32_t)nb, tc32 = r2 - tb32 * (uint32_t)nb;
  const int64_t w = w32, ta = ta32, tb = tb32, tc = tc32;
  const int64_t a = P.windows[3 * w] * (int64_t)nb + ta;
  const int64_t b = P.windows[3 * w + 1] * (int64_t)nb + tb;
  const int64_t c = P.windows[3 * w + 2] * (int64_t)nb + tc;
  const int64_t No = P.No, Nv = P.Nv;
  const int tid = threadIdx.x;
  bool ok = tuple_ok(Nv, a, b, c);
  if (ok && (P.first > 0 || P.last != ~0ULL)) {
    const uint64_t idx = tuple_index(P.aprefix, Nv, a, b, c);
    ok = idx >= P.first && idx < P.last;
  }
  if (!ok) {
    if (tid == 0) {
      P.e_part[2 * gid] = 0.0;
      P.e_part[2 * gid + 1] = 0.0;
    }
    return;
  }
  constexpr int CZ = 69, CSZ = 7 * CZ + 64;
  __shared__ double cube[6][CSZ];
  __shared__ double cubeJ[HJ ? 6 : 1][HJ ? CSZ : 1];
  // singles tables (entries 0..23 used; wave 0 writes all 64) and the V
  // blocks of the six block pairs of different positions (1152 used; every
  // thread writes its five, 1280)
  __shared__ double tA[64], tB[64], tC[64], sE[64];
  __shared__ double sV[5 * 256];
  __shared__ double red[2][4];

  int distinct = 0;
  if (a == b) distinct++;
  if (b == c) distinct--;
  const double epsabc = (P.ea[a] + P.ea[b]) + P.ea[c];
  const int no = (int)No;
  int B1, B2, B3;
  triple_of(tri, P.nbk, B1, B2, B3);
  // The orbit's distinct cubes.  Permutation p (ijk ikj jik jki kij kji)
  // holds its W in the cube of block coordinates code[p] (i | j << 8 |
  // k << 16); the distinct codes, in order of first appearance, are cubes
  // 0 .. nc - 1 (qc), and alias[p] is the cube of permutation p.  Every
  // array index below is a compile-time constant, so all of it stays in
  // registers.
  const int code[6] = {B1 | B2 << 8 | B3 << 16, B1 | B3 << 8 | B2 << 16,
                       B2 | B1 << 8 | B3 << 16, B2 | B3 << 8 | B1 << 16,
                       B3 | B1 << 8 | B2 << 16, B3 | B2 << 8 | B1 << 16};
  int alias[6], qc[6], nc = 0;
#pragma unroll
  for (int q = 0; q < 6; q++) qc[q] = 0;
#pragma unroll
  for (int p = 0; p < 6; p++) {
    int first = p;  // earliest permutation with the same cube
#pragma unroll
    for (int q = p - 1; q >= 0; q--)
      if (code[q] == code[p]) first = q;
    int al = nc;
#pragma unroll
    for (int q = 0; q < p; q++)
      if (q == first) al = alias[q];
    alias[p] = al;
    if (first == p) {
#pragma unroll
      for (int q = 0; q < 6; q++)
        if (q == nc) qc[q] = code[p];
      nc++;
    }
  }

  if (nc == 6)
    fused_orbit<GT, HJ, 6, CZ, CSZ>(P, gid, gslot, tid, no, a, b, c, B1, B2,
                                    B3, distinct, epsabc, qc, alias, cube,
                                    cubeJ, tA, tB, tC, sE, sV, red);
  else if (nc == 3)
    fused_orbit<GT, HJ, 3, CZ, CSZ>(P, gid, gslot, tid, no, a, b, c, B1, B2,
                                    B3, distinct, epsabc, qc, alias, cube,
                                    cubeJ, tA, tB, tC, sE, sV, red);
  else
    fused_orbit<GT, HJ, 1, CZ, CSZ>(P, gid, gslot, tid, no, a, b, c, B1, B2,
                                    B3, distinct, epsabc, qc, alias, cube,
                                    cubeJ, tA, tB, tC, sE, sV, red);
}

// Energy of each tuple slot: its triples' partials added in triple order.
__global__ __launch_bounds__(256) void triple_sum_kernel(const double *e_part,
                                                         int64_t slots,
                                                         int64_t ntri,
                                                         double *e_slot) {
  const int64_t slot = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (slot >= slots) return;
  const double *p = e_part + 2 * slot * ntri;
  double e = 0.0, ect = 0.0;
  for (int64_t t = 0; t < ntri; t++) {
    e += p[2 * t];
    ect += p[2 * t + 1];
  }
  e_slot[2 * slot] = e;
  e_slot[2 * slot + 1] = ect;
}

// Per window of a batch, a compensated (Neumaier) sum of its slots' energies
// in slot order: (sum, comp) to out[2 w], out[2 w + 1].
__global__ __launch_bounds__(256) void window_sum_kernel(const double *e_slot,
                                                         int nw, int64_t nb3,
                                                         double *out) {
  const int w = blockIdx.x * 256 + threadIdx.x;
  if (w >= nw) return;
  double sum = 0.0, comp = 0.0;
  for (int64_t q = 0; q < nb3; q++) {
    const double x = e_slot[2 * (w * nb3 + q)];
    const double t = sum + x;
    comp += (fabs(sum) >= fabs(x)) ? (sum - t) + x : (x - t) + sum;
    sum = t;
  }
  out[2 * w] = sum;
  out[2 * w + 1] = comp;
}

// Compensated sum of the slot energies of one batch: 256 threads each sum a
// fixed strided subset (Neumaier), then a fixed pairwise tree of (sum, comp).
__global__ __launch_bounds__(256) void batch_reduce_kernel(
    const double *e_slot, int64_t slots, double *partials, int index) {
  __shared__ double sh[4][256];
  double s0 = 0.0, c0 = 0.0, s1 = 0.0, c1 = 0.0;
  auto add = [](double &s, double &c, double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  };
  for (int64_t q = threadIdx.x; q < slots; q += 256) {
    add(s0, c0, e_slot[2 * q]);
    add(s1, c1, e_slot[2 * q + 1]);
  }
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = c0;
  sh[2][threadIdx.x] = s1;
  sh[3][threadIdx.x] = c1;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      double s = sh[0][threadIdx.x], c = sh[1][threadIdx.x] + sh[1][threadIdx.x + h];
      add(s, c, sh[0][threadIdx.x + h]);
      sh[0][threadIdx.x] = s;
      sh[1][threadIdx.x] = c;
      s = sh[2][threadIdx.x];
      c = sh[3][threadIdx.x] + sh[3][threadIdx.x + h];
      add(s, c, sh[2][threadIdx.x + h]);
      sh[2][threadIdx.x] = s;
      sh[3][threadIdx.x] = c;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[4 * index + 0] = sh[0][0];
    partials[4 * index + 1] = sh[1][0];
    partials[4 * index + 2] = sh[2][0];
    partials[4 * index + 3] = sh[3][0];
  }
}

}  // namespace

hipError_t launch_wsum(const DevSystem &s, const BatchArgs &b, const void *A,
                       const void *B, bool g32, double *out, hipStream_t st) {
  const int nbk = (int)w_blocks(s.No), nbr = (nbk + 1) / 2;
  const int64_t grid = b.slots * (int64_t)nbr * nbr * nbr;
  if (grid <= 0) return hipSuccess;
  if (grid > MAX_GRID) return hipErrorInvalidConfiguration;
  const int64_t NoP = g_pitch(s.No, g32), W3g = g_elems(s.No, g32);
  if (g32)
    hipLaunchKernelGGL(wsum_kernel<float>, dim3((unsigned)grid), dim3(256), 0,
                       st, (const float *)A, (const float *)B, out, s.No, NoP, W3g,
                       b.slots, nbk, nbr);
  else
    hipLaunchKernelGGL(wsum_kernel<double>, dim3((unsigned)grid), dim3(256),
                       0, st, (const double *)A, (const double *)B, out, s.No,
                       NoP, W3g, b.slots, nbk, nbr);
  return hipGetLastError();
}

hipError_t launch_energy(const DevSystem &s, const BatchArgs &b,
                         const void *G, const void *GJ, bool g32,
                         double tai_sign, double *e_slot, double *scratch,
                         bool fused, hipStream_t st) {
  if (fused) {
    const int64_t nbk = w_blocks(s.No), ntri = n_triples(nbk);
    FusedParams P;
    P.G = G;
    P.GJ = GJ;
    P.Tai = s.Tai;
    P.VHH = s.VHH;
    P.ei = s.ei;
    P.ea = s.ea;
    P.windows = b.windows;
    P.aprefix = s.aprefix;
    P.e_part = scratch;
    P.No = s.No;
    P.Nv = s.Nv;
    P.N2 = s.No * s.No;
    P.NoP = g_pitch(s.No, g32);
    P.W3g = g_elems(s.No, g32);
    P.slots = b.slots;
    P.ntri = ntri;
    P.first = b.first;
    P.last = b.last;
    P.nb = b.nb;
    P.nbk = (int)nbk;
    P.tai_sign = tai_sign;
    const int64_t grid = (b.slots + 7) / 8 * 8 * ntri;
    if (grid <= 0) return hipSuccess;
    // (the kernel decodes its block id and class-buffer offsets in 32 bits)
    if (grid > MAX_GRID || g_elems(s.No, g32) >= (int64_t(1) << 31))
      return hipErrorInvalidConfiguration;
    if (g32 && GJ)
      hipLaunchKernelGGL((energy_fused_kernel<float, true>), dim3((unsigned)grid),
                         dim3(256), 0, st, P);
    else if (g32)
      hipLaunchKernelGGL((energy_fused_kernel<float, false>), dim3((unsigned)grid),
                         dim3(256), 0, st, P);
    else if (GJ)
      hipLaunchKernelGGL((energy_fused_kernel<double, true>), dim3((unsigned)grid),
                         dim3(256), 0, st, P);
    else
      hipLaunchKernelGGL((energy_fused_kernel<double, false>), dim3((unsigned)grid),
                         dim3(256), 0, st, P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(triple_sum_kernel, dim3((unsigned)((b.slots + 255) / 256)),
                       dim3(256), 0, st, (const double *)scratch, b.slots, ntri,
                       e_slot);
    return hipGetLastError();
  }
  const int64_t W3c = w_cube(s.No);
  double *W = scratch, *WJ = GJ ? scratch + b.slots * W3c : nullptr;
  hipError_t e = launch_wsum(s, b, G, nullptr, g32, W, st);
  if (e != hipSuccess) return e;
  if (GJ && (e = launch_wsum(s, b, GJ, nullptr, g32, WJ, st)) != hipSuccess)
    return e;
  EnergyParams P;
  P.W = W;
  P.WJ = WJ;
  P.Tai = s.Tai;
  P.VHH = s.VHH;
  P.ei = s.ei;
  P.ea = s.ea;
  P.windows = b.windows;
  P.aprefix = s.aprefix;
  P.e_slot = e_slot;
  P.No = s.No;
  P.Nv = s.Nv;
  P.N2 = s.No * s.No;
  P.nbk = w_blocks(s.No);
  P.W3 = W3c;
  P.slots = b.slots;
  P.first = b.first;
  P.last = b.last;
  P.nb = b.nb;
  P.tai_sign = tai_sign;
  const dim3 grid((unsigned)b.slots), block(256);
  if (GJ)
    hipLaunchKernelGGL(energy_kernel<true>, grid, block, 0, st, P);
  else
    hipLaunchKernelGGL(energy_kernel<false>, grid, block, 0, st, P);
  return hipGetLastError();
}

hipError_t launch_window_sums(const double *e_slot, int nw, int64_t nb3,
                              double *out, hipStream_t st) {
  if (nw <= 0) return hipSuccess;
  hipLaunchKernelGGL(window_sum_kernel, dim3((unsigned)((nw + 255) / 256)),
                     dim3(256), 0, st, e_slot, nw, nb3, out);
  return hipGetLastError();
}

hipError_t launch_batch_reduce(const DevSystem &, const BatchArgs &b,
                               const double *e_slot, double *partials,
                               int index, hipStream_t st) {
  hipLaunchKernelGGL(batch_reduce_kernel, dim3(1), dim3(256), 0, st, e_slot,
                     b.slots, partials, index);
  return hipGetLastError();
}

}  // namespace atrip_amd
