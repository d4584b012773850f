// Internal declarations shared by the HIP translation units of libatrip_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/atrip_hip.h"

namespace atrip_amd {

// Largest pinned host buffer of a reader upload (bytes).
constexpr int64_t PINNED_MAX = (int64_t)256 << 20;

// Contraction-index padding of the device layouts: every K run of the
// doubles GEMMs is a multiple of the 16-element K chunk of the kernels.
constexpr int KCHUNK = 16;
__host__ __device__ inline int64_t pad16(int64_t n) { return (n + 15) / 16 * 16; }

// bf16x3 operands.  An FP64 operand [o][r < nr][k < K] is held as three
// bf16 parts per element (x ~ b0 + b1 + b2) in tiles of 32 rows x 16 k,
// R = o * nrp + r (nrp: nr, or nr padded to 32 so that every o starts a
// tile row).  Tile (R / 32, k / 16) sits at ((R / 32) (K / 16) + k / 16)
// * 1536 parts; inside it part p of (R, k) sits at 512 p + 8 L + k % 8 with
// L = R % 32 + 32 ((k % 16) / 8), the lane that feeds (R, k) to a 32x32x16
// MFMA, so each part of a 32 x 16 fragment chunk is 1 KiB contiguous in
// lane order (one 16-byte load per lane).  The A operands (nr = No^2 rows
// (p, q)) also keep a row-swapped copy whose row p + No q holds row
// q + No p.
__host__ __device__ inline int64_t pad32(int64_t n) { return (n + 31) / 32 * 32; }
struct TriGeom {
  int64_t no, nr, nrp, K;
  bool swapped;  // keeps a row-swapped copy (A operands)
};
struct TriOp {
  uint16_t *p = nullptr, *sw = nullptr;
};
__host__ __device__ inline int64_t tri_parts(const TriGeom &g) {
  return 3 * pad32(g.no * g.nrp) * g.K;
}
// Start of tile (R / 32, k / 16) and the lane slot of (R, k).
__host__ __device__ inline int64_t tri_tile(int64_t R, int64_t k, int64_t K) {
  return ((R >> 5) * (K >> 4) + (k >> 4)) * 1536;
}
__host__ __device__ inline int tri_lane(int64_t R, int64_t k) {
  return (int)((R & 31) + 32 * ((k & 15) >> 3));
}

// Layout of the per-tuple class buffers written by the doubles kernels and
// read by the energy kernels: each class keeps its GEMM output order,
// element (row m, column index r) at m * P + r with the row pitch
// P = g_pitch(No), so a 16-column MFMA output fragment row is one contiguous
// store and every row starts on a 64-byte boundary.  In W indices:
//   class w: m = i + No j, r = k   ->  (i + No j) P + k
//   class v: m = i + No k, r = j   ->  (i + No k) P + j
//   class u: m = j + No k, r = i   ->  (j + No k) P + i
// Row pitch: No rounded up to 8 elements for FP64 buffers, 16 for FP32 (the
// FP32-product forms), so the 64-byte row pieces the epilogue gathers never
// straddle a line; the pad columns are never written or read.
__host__ __device__ inline int64_t g_pitch(int64_t No, bool g32) {
  const int64_t q = g32 ? 16 : 8;
  return (No + q - 1) / q * q;
}
// Largest one-dimensional grid a launch takes (workgroups); launchers
// refuse larger ones instead of truncating the count.
constexpr int64_t MAX_GRID = 0xFFFFFFFFLL;
// Elements of one tuple's class buffer.
__host__ __device__ inline int64_t g_elems(int64_t No, bool g32) {
  return No * No * g_pitch(No, g32);
}
// The energy kernels stage 8x8x8 blocks of W in LDS (edge WCUBE).
constexpr int WCUBE = 8;
__host__ __device__ inline int64_t w_blocks(int64_t No) {
  return (No + WCUBE - 1) / WCUBE;
}
// Doubles of one tuple's W in the cube layout the energy kernels read: the
// occupied range padded to whole cubes, cube (bx, by, bz) at
// ((bz nbk + by) nbk + bx) * 512, element x + 8 y + 64 z inside.
__host__ __device__ inline int64_t w_cube(int64_t No) {
  const int64_t n = w_blocks(No) * WCUBE;
  return n * n * n;
}
// Block triples B1 >= B2 >= B3 over nbk blocks per index.
__host__ __device__ inline int64_t n_triples(int64_t nbk) {
  return nbk * (nbk + 1) * (nbk + 2) / 6;
}

// Device-resident system.  Pair-indexed arrays use the CTF pair order
// p(x,y) = x + Nv*y.  Layouts (all row-major with the last index fastest):
//   T3  [x][p + No q][E < Kv]    = Tpphh[x,E,p,q]          (T(x) slices)
//   V3  [p(x,y)][r][E < Kv]      = Vppph[x,y,E,r]          (V(xy) slices)
//   H3  [x][p + No q][L < Ko]    = Vhhhp[p,q,L,x]          (Vhhh(x) slices)
//   T2N [p(x,y)][n][L < Ko]      = -Tpphh[x,y,L,n]         (-T(xy)[L,n])
//   T2TN[p(x,y)][n][L < Ko]      = -Tpphh[x,y,n,L]         (-T(xy)[n,L])
//   VHH [p(x,y)][i + No j]       = Vpphh[x,y,i,j]          (V(xy)ij)
//   Tai [a + Nv i], ei[i], ea[a] as in CTF.
// Padding entries (E >= Nv, L >= No) are zero.
struct DevSystem {
  int64_t No = 0, Nv = 0, Kv = 0, Ko = 0;
  double *T3 = nullptr, *V3 = nullptr, *H3 = nullptr, *T2N = nullptr,
         *T2TN = nullptr, *VHH = nullptr, *Tai = nullptr, *ei = nullptr,
         *ea = nullptr;
  double *V3J = nullptr, *H3J = nullptr;  // (cT) integrals, may be null
  bool has_j = false;                     // (cT) integrals present
  // Complex systems: the arrays above hold the real parts, these the
  // imaginary parts (same layouts), and the *n arrays the negated imaginary
  // parts of T3 and the hole integrals (operands of the real-GEMM passes).
  bool cplx = false;
  double *T3i = nullptr, *V3i = nullptr, *H3i = nullptr, *T2Ni = nullptr,
         *T2TNi = nullptr, *VHHi = nullptr, *Taii = nullptr, *eii = nullptr,
         *eai = nullptr, *V3Ji = nullptr, *H3Ji = nullptr;
  double *T3in = nullptr, *H3in = nullptr, *H3Jin = nullptr;
  // Split-FP32 copies of the GEMM operands, made on the first mixed run:
  // h = (float)x and l = bf16(x - h), same layouts as above (bf16(h), the
  // other factor of the correction products, is formed in registers).
  struct Split {
    float *h = nullptr;
    uint16_t *l = nullptr;
  };
  Split sT3, sV3, sH3, sT2N, sT2TN, sV3J, sH3J;
  bool split_ready = false;
  // bf16x3 copies in the tiled layout (TriGeom): T3 and H3 rows padded to
  // 32 per slice with row-swapped copies, the pair-indexed arrays unpadded.
  TriOp tT3, tV3, tH3, tT2N, tT2TN, tV3J, tH3J;
  bool tri_ready = false;
  // Operand store: 0 FP64 (split or bf16x3 copies made on the first mixed
  // run of that form), 1 split FP32 only (h + l, 6 bytes per element), 2
  // FP32 (h) only, 3 bf16x3 only (6 bytes per element).  Modes 1-3 keep no
  // FP64 copy of the doubles operands (T3, V3, H3, T2N, T2TN and the J
  // arrays).
  int store = 0;
  uint32_t *points = nullptr;             // packed (i>=j>=k) triples
  int64_t npoints = 0;
  uint64_t *aprefix = nullptr;            // tuples with first index < a
  uint64_t bytes = 0;
  std::vector<void *> allocations;
};

// bf16 of a float, round to nearest even (finite inputs).
__device__ inline uint16_t to_bf16(float v) {
  uint32_t u = __float_as_uint(v);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Split-FP32 element q of x: h = (float)x, l = bf16(x - h); l may be null
// (FP32-only store).
__device__ inline void split_store(double x, float *h, uint16_t *l,
                                   int64_t q) {
  const float hi = (float)x;
  h[q] = hi;
  if (l) l[q] = to_bf16((float)(x - (double)hi));
}

// bf16x3 parts of x into element (R, k) of a tiled copy: b0 = bf16(x),
// b1 = bf16(x - b0), b2 = bf16(x - b0 - b1), each residual formed in FP64,
// so b0 + b1 + b2 is x within 2^-27 |x|.
__device__ inline void tri_put(double x, uint16_t *out, int64_t R, int64_t k,
                               int64_t K) {
  const uint16_t p0 = to_bf16((float)x);
  const double r1 = x - (double)__uint_as_float((uint32_t)p0 << 16);
  const uint16_t p1 = to_bf16((float)r1);
  const double r2 = r1 - (double)__uint_as_float((uint32_t)p1 << 16);
  const int64_t at = tri_tile(R, k, K) + 8 * tri_lane(R, k) + (k & 7);
  out[at] = p0;
  out[at + 512] = p1;
  out[at + 1024] = to_bf16((float)r2);
}

// Element q of FP64 operand [o][r][k] into its tiled copy (and the
// row-swapped one when the geometry keeps it; No = sqrt(nr) there).
__device__ inline void tri_store(double x, const TriOp &t, const TriGeom &g,
                                 int64_t No, int64_t q) {
  const int64_t k = q % g.K, row = q / g.K, r = row % g.nr, o = row / g.nr;
  tri_put(x, t.p, o * g.nrp + r, k, g.K);
  if (g.swapped) {
    const int64_t rs = r / No + No * (r % No);
    tri_put(x, t.sw, o * g.nrp + rs, k, g.K);
  }
}

// Tuple windows: block triples (alpha <= beta <= gamma) of edge nb.
struct WindowPlan {
  int nb = 0;
  int nblk = 0;
  std::vector<int32_t> windows;  // 3 per window, lexicographic
  // per window: smallest and largest canonical index and number of tuples
  std::vector<uint64_t> wmin, wmax, wcount;
};

struct KernelStat {
  uint64_t launches = 0;
  double ms = 0.0;
  double flop = 0.0;
};

struct Batch;

// Tuned plan of the batched form for one system shape.
struct BatchTune {
  int nb = 1, tile = 0;
  double sec_window = 0.0;  // measured (or modelled) time of a full window
  bool measured = false;
};

}  // namespace atrip_amd

struct atrip_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::map<int, atrip_amd::DevSystem> systems;
  std::map<int, atrip_amd::WindowPlan> plans;  // keyed by system
  int next_sys = 1;
  // options
  int window = 0;  // window edge; 0: automatic per system (auto_window)
  int batch = 4;
  int timing = 0;
  int tile = -1;  // -1: chosen per system shape
  int operand_store = 0;  // store mode of the next uploads (DevSystem::store)
  int epilogue = 0;  // 0: fused class sums + energy, 1: W cubes first
  int serve_plain_no = 16;  // served requests with No >= this: plain path
  int64_t stage_kb = 4 << 20;  // device staging of host uploads (KiB)
  // scratch, grown on demand
  double *G = nullptr;       // class buffers [ncls][slots][g_elems(No)]
  size_t G_bytes = 0;
  double *wcube = nullptr;   // W as cubes for the energy kernels
  size_t wcube_bytes = 0;
  double *e_slot = nullptr;  // per-slot energies (2 per slot: e, e_ct)
  size_t e_slot_bytes = 0;
  double *partials = nullptr;  // per-batch (sum, comp) x 2
  size_t partials_bytes = 0;
  void *sel = nullptr;  // a run's window list (and per-window sums)
  size_t sel_bytes = 0;
  // batched form: tuned plans keyed by (No, padded Nv, precision),
  // workgroup scratch
  std::map<std::array<int64_t, 3>, atrip_amd::BatchTune> btune;
  double *bscratch = nullptr;
  size_t bscratch_bytes = 0;
  int64_t batch_scratch_mb = 8192;
  int batch_grid = 0;   // 0: occupancy x CUs
  int batch_order = 0;  // 0: longest window first, 1: canonical
  int batch_nb = 0;     // forced window edge of the batched form (0: tuned)
  int batch_tile = -1;  // forced tile variant of the batched form (-1: tuned)
  std::map<std::string, atrip_amd::KernelStat> stats;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending_events;
  std::vector<std::string> pending_names;
  std::vector<double> pending_flop;
};

namespace atrip_amd {

// ---- launch wrappers (one translation unit each) -----------------------

// Batch of consecutive windows handed to the hot kernels.
struct BatchArgs {
  const int32_t *windows;  // device, 3 per window of this batch
  int nw;                  // windows in the batch
  int nb;                  // window edge
  uint64_t first, last;    // canonical tuple range [first, last)
  int64_t slots;           // nw * nb^3
};

// Doubles contractions in FP64 MFMA: class buffers G[3][slots][g_elems(No)].
hipError_t launch_doubles_f64(const DevSystem &s, const BatchArgs &b,
                              bool use_j, double *G, int tile,
                              hipStream_t st);

// Doubles contractions in split FP32.
// Doubles contractions with FP32 products (class buffers in FP32):
// corrections = true adds the hi*lo + lo*hi terms (split FP32).
hipError_t launch_doubles_split(const DevSystem &s, const BatchArgs &b,
                                bool use_j, float *G, bool corrections,
                                int tile, hipStream_t st);

// Doubles contractions in bf16x3 (six part products on bf16 MFMA, FP32
// accumulation, class buffers in FP32).
hipError_t launch_doubles_tri(const DevSystem &s, const BatchArgs &b,
                              bool use_j, float *G, int tile, hipStream_t st);

// Class sums: out[slot] = sum(A) (+ sum(B)) as cubes (w_cube per slot), for
// class-buffer sets A, B (FP64, or FP32 when g32).
hipError_t launch_wsum(const DevSystem &s, const BatchArgs &b, const void *A,
                       const void *B, bool g32, double *out, hipStream_t st);

// Class sums and singles + energy epilogue; writes e_slot[2*slot + {0,1}].
// fused: one kernel gathers each block triple's W straight from the class
// buffers (scratch: slots * n_triples * 2 doubles of per-triple partials);
// otherwise the class-sum kernel writes W as cubes first (scratch: (GJ ? 2
// : 1) * slots * w_cube(No) doubles).  Same W bits either way.
hipError_t launch_energy(const DevSystem &s, const BatchArgs &b,
                         const void *G, const void *GJ, bool g32,
                         double tai_sign, double *e_slot, double *scratch,
                         bool fused, hipStream_t st);
// Doubles of energy scratch per tuple slot (see launch_energy).
inline int64_t energy_scratch_per_slot(int64_t No, bool use_j, bool cplx,
                                       bool fused) {
  if (fused && !cplx) return 2 * n_triples(w_blocks(No));
  return (use_j ? 2 : 1) * (cplx ? 2 : 1) * w_cube(No);
}

// Singles + energy of complex systems from the four class-buffer sets of
// the real passes (Wr = G0 + G1, Wi = G2 + G3); GJ null without (cT).
// wcube: scratch of (GJ ? 4 : 2) * slots * w_cube(No) doubles.
hipError_t launch_energy_z(const DevSystem &s, const BatchArgs &b,
                           double *const G[4], double *const GJ[4],
                           double tai_sign, double *e_slot, double *wcube,
                           hipStream_t st);

// Per window of a batch (nw windows of nb3 slots), the Neumaier sum of its
// slot energies in slot order: out[2 w] = sum, out[2 w + 1] = compensation.
hipError_t launch_window_sums(const double *e_slot, int nw, int64_t nb3,
                              double *out, hipStream_t st);

// Deterministic compensated sum of the valid slot energies of a batch into
// partials[4*index + {sum, comp, sum_ct, comp_ct}].
hipError_t launch_batch_reduce(const DevSystem &s, const BatchArgs &b,
                               const double *e_slot, double *partials,
                               int index, hipStream_t st);

// Upload helpers.
hipError_t launch_transpose(const double *src, double *dst, int64_t n_outer,
                            int64_t n_a, int64_t n_b, int64_t src_o,
                            int64_t src_a, int64_t src_b, int64_t dst_o,
                            int64_t dst_a, int64_t dst_b, double scale,
                            hipStream_t st);
// Generate a synthetic system in place; tensor_flag is or-ed into every
// tensor id (ATRIP_SYN_IMAG fills the imaginary parts of a complex system,
// given a DevSystem whose arrays are its imaginary planes).
// Every amplitude and integral value is multiplied by `scale` (orbital
// energies are not).
hipError_t launch_synth(DevSystem &s, uint64_t seed, bool with_j,
                        hipStream_t st, int tensor_flag = 0,
                        double scale = 1.0);
// dst[q] = scale * src[q], q < n.
hipError_t launch_scale(const double *src, double *dst, int64_t n,
                        double scale, hipStream_t st);
// Fill a Split copy of n doubles.
hipError_t launch_split(const double *x, DevSystem::Split &out, int64_t n,
                        hipStream_t st);
// Fill the bf16x3 copy (and row-swapped copy) of FP64 operand x of
// geometry g (No: occupied indices, for the swap).
hipError_t launch_tri(const double *x, const TriOp &out, const TriGeom &g,
                      int64_t No, hipStream_t st);
// Geometry of a system's bf16x3 operands: A (T3, H3 and H3J) and the
// pair-indexed B arrays (V3, T2N, T2TN, V3J), for contraction length K.
inline TriGeom tri_geom_a(int64_t No, int64_t Nv, int64_t K) {
  return TriGeom{Nv, No * No, pad32(No * No), K, true};
}
inline TriGeom tri_geom_b(int64_t No, int64_t Nv, int64_t K) {
  return TriGeom{Nv * Nv, No, No, K, false};
}
inline uint64_t tri_bytes(const TriGeom &g) {
  return (uint64_t)tri_parts(g) * 2 * (g.swapped ? 2 : 1);
}

// ---- batched form: many small resident systems in one persistent launch --

// One system of a batched launch, with the window edge and tile variant
// chosen for it.  Tile variant v is the wave tile of (16*bfm(v)) x
// (16*bfn(v)) outputs.
struct BSys {
  const double *T3, *V3, *H3, *T2N, *T2TN, *VHH, *Tai, *ei, *ea;
  const uint32_t *points;
  int64_t No, Nv, Kv, Ko, npoints;
  int32_t nb, tile;
};
constexpr int N_BTILES = 10;
constexpr __host__ __device__ inline int bfm(int v) {
  constexpr int t[N_BTILES] = {1, 1, 2, 2, 1, 4, 2, 4, 3, 3};
  return t[v];
}
constexpr __host__ __device__ inline int bfn(int v) {
  constexpr int t[N_BTILES] = {1, 2, 1, 2, 4, 1, 4, 2, 2, 3};
  return t[v];
}

// Split-FP32 operands of a system in a batched launch (h = FP32 values,
// l = bf16(x - h); l unused by the FP32-only form).
struct BSplit {
  const float *T3h, *V3h, *H3h, *T2Nh, *T2TNh;
  const uint16_t *T3l, *V3l, *H3l, *T2Nl, *T2TNl;
};

// One work item: a window (block triple of edge nb) of one system.
struct BTask {
  int32_t sys, al, be, ga;
  int64_t tuple_off;  // first entry of the task's slots in the per-tuple output
};

struct BatchedLaunch {
  int precision;            // ATRIP_FP64, ATRIP_SPLIT_FP32 or ATRIP_FP32
  const BSys *sys;          // device
  const BSplit *split;      // device, one per system (mixed forms only)
  const BTask *tasks;       // device, canonical order (system-major)
  const int32_t *order;     // device, dispatch order (a permutation)
  int ntasks;
  unsigned *counter;        // device, zero before the launch
  double *scratch;          // grid * scratch_stride doubles
  int64_t scratch_stride;
  double *task_e;           // 2 per task: (sum, compensation)
  double *tuple_e;          // optional per-slot energies
  int grid;
};

// Persistent launch: every workgroup takes windows from the dispatch order
// until none is left; doubles, singles and energy of a window stay in the
// workgroup.
hipError_t launch_batched(const BatchedLaunch &L, hipStream_t st);
// Workgroups per CU the persistent kernel of a precision can hold.
int batched_blocks_per_cu(int precision);
// Canonical (lexicographic) index of tuple (a,b,c), a<=b<=c, not all equal.
__host__ __device__ inline uint64_t tuple_index(const uint64_t *aprefix,
                                                uint64_t Nv, uint64_t a,
                                                uint64_t b, uint64_t c) {
  // tuples (a, b', c') with a <= b' < b: sum_{b'=a}^{b-1} (Nv - b')
  const uint64_t nb = b - a;
  const uint64_t before_b = nb * Nv - (nb * (a + b - 1)) / 2;
  return aprefix[a] + before_b - (b > a ? 1 : 0) + (c - b) - (a == b ? 1 : 0);
}

__host__ __device__ inline bool tuple_ok(uint64_t Nv, uint64_t a, uint64_t b,
                                         uint64_t c) {
  return a <= b && b <= c && c < Nv && !(a == b && b == c);
}

}  // namespace atrip_amd
