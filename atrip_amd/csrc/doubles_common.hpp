// Work decomposition shared by the doubles kernels (FP64 and split FP32):
// the class/segment table of the regrouped doubles contractions.
#pragma once

#include "internal.hpp"

namespace atrip_amd {

// Column tuple of class `cl` for A-pair offsets (o1,o2) and third offset t;
// slot is the tuple's position ((ta*nb + tb)*nb + tc) inside its window.
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

// B operand of segment sg of class cl for column tuple (a,b,c): returns the
// array (0: V(xy) slabs, 1: -T(xy)[L,n], 2: -T(xy)[n,L]) and its pair (u,v).
//   class w: P0 V(bc), P1 V(ac), H5 -TAC[L,k], H2 -TBC[L,k]
//   class v: P3 V(cb), P5 V(ab), H1 -TAB[L,j], H4 -TBC[j,L]
//   class u: P4 V(ca), P2 V(ba), H0 -TAB[i,L], H3 -TAC[i,L]
__device__ inline int b_pair(int cl, int sg, int64_t a, int64_t b, int64_t c,
                             int64_t &u, int64_t &v) {
  if (cl == 0) {
    if (sg == 0) { u = b; v = c; return 0; }
    if (sg == 1) { u = a; v = c; return 0; }
    if (sg == 2) { u = a; v = c; return 1; }
    u = b; v = c; return 1;
  }
  if (cl == 1) {
    if (sg == 0) { u = c; v = b; return 0; }
    if (sg == 1) { u = a; v = b; return 0; }
    if (sg == 2) { u = a; v = b; return 1; }
    u = b; v = c; return 2;
  }
  if (sg == 0) { u = c; v = a; return 0; }
  if (sg == 1) { u = b; v = a; return 0; }
  if (sg == 2) { u = a; v = b; return 2; }
  u = a; v = c; return 2;
}

// A-pair (x1, x2) of class cl: w (a,b), v (a,c), u (b,c).  The A operand of
// segment 0 is T(x1) straight, 1 T(x2) swapped, 2 Vhhh(x2) straight,
// 3 Vhhh(x1) swapped.
__device__ inline void class_pair(int cl, int64_t al, int64_t be, int64_t ga,
                                  int nb, int o1, int o2, int64_t &x1,
                                  int64_t &x2) {
  if (cl == 0) { x1 = al * nb + o1; x2 = be * nb + o2; }
  else if (cl == 1) { x1 = al * nb + o1; x2 = ga * nb + o2; }
  else { x1 = be * nb + o1; x2 = ga * nb + o2; }
}

// W position (i,j,k) of GEMM element (row (p,q), column index rr).
__device__ inline void class_out(int cl, int64_t p, int64_t q, int64_t rr,
                                 int64_t &i, int64_t &j, int64_t &k) {
  if (cl == 0) { i = p; j = q; k = rr; }       // "ij|k"
  else if (cl == 1) { i = p; j = rr; k = q; }  // "ik|j"
  else { i = rr; j = p; k = q; }               // "jk|i"
}

// Tile shapes (wave tile FM x FN fragments of 16x16, 2x2 waves) and the
// fraction of peak each reaches per useful tile element on MI355X in FP64:
// 5x5 at No=40, Nv=400; 4x2, 4x4, 3x3, 4x5 and 5x4 relative to it from
// whole-system runs at No=26/Nv=308, 31/270, 38/204 and 20/183, each divided
// by its padded-grid fraction (5x4 measured 4 % above 5x5 there, entered
// equal to it, since no shape measured both at full use); 2x2 and 2x4 from
// an earlier kernel at No=40.
constexpr int N_TILES = 8;
__host__ __device__ inline int tile_fm(int t) {
  const int v[N_TILES] = {2, 4, 2, 4, 3, 5, 4, 5};
  return v[t];
}
__host__ __device__ inline int tile_fn(int t) {
  const int v[N_TILES] = {2, 2, 4, 4, 3, 5, 5, 4};
  return v[t];
}
inline double tile_rate(int t) {
  const double v[N_TILES] = {0.44, 0.575, 0.51, 0.627, 0.61, 0.69, 0.677, 0.69};
  return v[t];
}

// Automatic tile: best measured rate times the useful fraction of the padded
// M = No^2 by N = nb*No tile grid.
inline int pick_tile(int tile, int64_t No, int nb) {
  if (tile >= 0 && tile < N_TILES) return tile;
  double best = -1.0;
  int pick = 0;
  const double M = (double)No * No, N = (double)nb * No;
  for (int t = 0; t < N_TILES; t++) {
    const double mt = 32.0 * tile_fm(t), nt = 32.0 * tile_fn(t);
    const double use = M / (__builtin_ceil(M / mt) * mt) *
                       N / (__builtin_ceil(N / nt) * nt);
    const double score = tile_rate(t) * use;
    if (score > best) {
      best = score;
      pick = t;
    }
  }
  return pick;
}

// Automatic window edge of the plain path: the doubles GEMMs of a window
// have N = nb No columns, so the edge sets how much of the last column tile
// is padding.  Each candidate edge is scored by its best tile's modelled
// rate (pick_tile's model) times the share of valid tuples in the windows;
// an edge other than 8 is taken only when it scores 15 % above 8 (the
// model's own error).  Measured in FP64 on MI355X: No=26, Nv=308 runs at
// 40.6 TFLOP/s with edge 6 against 33.2 with edge 8; No=40 and No=60 keep 8.
inline int auto_window(int64_t No, int64_t Nv) {
  auto score = [&](int nb) {
    const int t = pick_tile(-1, No, nb);
    const double mt = 32.0 * tile_fm(t), nt = 32.0 * tile_fn(t);
    const double M = (double)No * No, N = (double)nb * No;
    const double use = M / (__builtin_ceil(M / mt) * mt) *
                       N / (__builtin_ceil(N / nt) * nt);
    const double nblk = __builtin_ceil((double)Nv / nb);
    const double cells = nblk * (nblk + 1) * (nblk + 2) / 6 * nb * nb * nb;
    const double tuples = (double)Nv * (Nv + 1) * (Nv + 2) / 6;
    return tile_rate(t) * use * tuples / cells;
  };
  const double base = score(8);
  int pick = 8;
  double best = 1.15 * base;
  for (int nb : {6, 12}) {
    const double s = score(nb);
    if (s > best) {
      best = s;
      pick = nb;
    }
  }
  return pick;
}

// Wave tiles of FM x FN fragments of 32 x 32 (2 x 2 waves per workgroup)
// of the 32x32-MFMA kernels (bf16x3, split FP32), variant t of the "tile"
// option.
constexpr int N_T32_TILES = 10;
__host__ __device__ inline int t32_fm(int t) {
  const int v[N_T32_TILES] = {2, 2, 3, 2, 3, 2, 3, 4, 3, 4};
  return v[t];
}
__host__ __device__ inline int t32_fn(int t) {
  const int v[N_T32_TILES] = {2, 3, 2, 4, 3, 5, 5, 4, 4, 3};
  return v[t];
}
// Rate of each variant relative to 4x4, per useful tile element: variants
// 4-9 measured in bf16x3 on MI355X (mean of No=40, Nv=400 and No=60,
// Nv=600, each divided by its padded-grid fraction; 2x5 re-measured against
// 3x5 at No=40 with the one-pipeline K loop, 1.5 % faster there); variants
// 0-3 from the operand bytes per MFMA cycle, FM FN / (FM + FN), scaled by the
// ratio that measured variants 4 and 5 show over the same formula.
inline double t32_rate(int t) {
  const double v[N_T32_TILES] = {0.62, 0.75, 0.75, 0.83, 0.985,
                                 0.966, 0.988, 1.0, 0.989, 0.952};
  return v[t];
}
// Automatic variant: rate times the useful fraction of the padded
// M = No^2 by N = nb No tile grid.
inline int pick_t32_tile(int tile, int64_t No, int nb) {
  if (tile >= 0 && tile < N_T32_TILES) return tile;
  double best = -1.0;
  int pick = 0;
  const double M = (double)No * No, N = (double)nb * No;
  for (int t = 0; t < N_T32_TILES; t++) {
    const double mt = 64.0 * t32_fm(t), nt = 64.0 * t32_fn(t);
    const double use = M / (__builtin_ceil(M / mt) * mt) *
                       N / (__builtin_ceil(N / nt) * nt);
    const double score = t32_rate(t) * use;
    if (score > best) {
      best = score;
      pick = t;
    }
  }
  return pick;
}

}  // namespace atrip_amd
