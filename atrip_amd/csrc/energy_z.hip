// Singles + energy epilogue of complex (T) (atrip::Atrip::run<Complex>).
//
// The doubles of a complex system run as four real class-GEMM passes on
// real/imaginary operand planes (see the run loop), so W = Wr + i Wi arrives
// as four sets of class buffers: Wr = sum(G0) + sum(G1), Wi = sum(G2) +
// sum(G3), summed into cubes by launch_wsum.  Per tuple (one workgroup), for each block
// triple of edge 8 the cubes holding W at the six index permutations are
// summed into LDS as (re, im), Z = W + singles is formed with complex
// products (reference src/atrip/Equations.cxx:387-426), and the energy
// follows get_energy_distinct / get_energy_same (:101-238) with F = Complex:
// A..F = conj(W) (maybe_conjugate_scalar), U..Z = Z, complex denominators,
// real part of the tuple's sum (:176-178).  Fixed per-thread order and a
// fixed reduction tree, as the real epilogue.
#include "internal.hpp"

namespace atrip_amd {
namespace {

struct cplx {
  double r, i;
};
__device__ inline cplx operator+(cplx a, cplx b) { return {a.r + b.r, a.i + b.i}; }
__device__ inline cplx operator-(cplx a, cplx b) { return {a.r - b.r, a.i - b.i}; }
__device__ inline cplx operator*(cplx a, cplx b) {
  return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__device__ inline cplx operator*(double s, cplx a) { return {s * a.r, s * a.i}; }
__device__ inline cplx conj(cplx a) { return {a.r, -a.i}; }
// Real part of a / b (Smith's scaling, exact for real b).
__device__ inline double re_div(cplx a, cplx b) {
  if (b.i == 0.0) return a.r / b.r;
  if (fabs(b.r) >= fabs(b.i)) {
    const double t = b.i / b.r, d = b.r + b.i * t;
    return (a.r + a.i * t) / d;
  }
  const double t = b.r / b.i, d = b.r * t + b.i;
  return (a.r * t + a.i) / d;
}

struct EnergyZParams {
  const double *Wr, *Wi, *WJr, *WJi;  // cubes, [slot][W3]
  const double *Tair, *Taii, *VHHr, *VHHi, *eir, *eii, *ear;
  const int32_t *windows;
  const uint64_t *aprefix;
  double *e_slot;
  int64_t No, Nv, N2, W3, nbk, slots;  // nbk: cubes per index, W3 = (8 nbk)^3
  uint64_t first, last;
  int nb;
  double tai_sign;
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <bool HJ>
__global__ __launch_bounds__(256) void energy_z_kernel(EnergyZParams P) {
  const int64_t gslot = blockIdx.x;
  const int nb = P.nb;
  const int64_t nb3 = (int64_t)nb * nb * nb;
  const int64_t w = gslot / nb3, s = gslot % nb3;
  const int64_t a = P.windows[3 * w] * (int64_t)nb + s / (nb * nb);
  const int64_t b = P.windows[3 * w + 1] * (int64_t)nb + (s / nb) % nb;
  const int64_t c = P.windows[3 * w + 2] * (int64_t)nb + s % nb;
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
  __shared__ double cr[6][512], ci[6][512];
  __shared__ double jr[HJ ? 6 : 1][HJ ? 512 : 1], ji[HJ ? 6 : 1][HJ ? 512 : 1];
  __shared__ double red[2][4];

  const int64_t pbc = (b + Nv * c) * N2, pac = (a + Nv * c) * N2,
                pab = (a + Nv * b) * N2;
  int distinct = 0;
  if (a == b) distinct++;
  if (b == c) distinct--;
  // the reference keeps only the real part of the virtual-orbital sum
  // (std::real, src/atrip/Atrip.cxx:643-644); the occupied energies of the
  // denominator stay complex
  const cplx epsabc = cplx{(P.ear[a] + P.ear[b]) + P.ear[c], 0.0};
  const int nbk = (int)P.nbk, no = (int)No, tid = threadIdx.x;
  auto tai = [&](int64_t x, int occ) {
    return cplx{P.tai_sign * P.Tair[x + Nv * occ], P.tai_sign * P.Taii[x + Nv * occ]};
  };
  auto vhh = [&](int64_t base, int p, int q) {
    return cplx{P.VHHr[base + p + no * q], P.VHHi[base + p + no * q]};
  };
  const int64_t off = gslot * P.W3;

  double e_acc = 0.0, ect_acc = 0.0;
#pragma unroll 1
  for (int B3 = 0; B3 < nbk; B3++)
#pragma unroll 1
    for (int B2 = B3; B2 < nbk; B2++)
#pragma unroll 1
      for (int B1 = B2; B1 < nbk; B1++) {
        // cubes of the permutations ijk ikj jik jki kij kji
        const int cx[6] = {B1, B1, B2, B2, B3, B3};
        const int cy[6] = {B2, B3, B1, B3, B1, B2};
        const int cz[6] = {B3, B2, B3, B1, B2, B1};
        __syncthreads();
        for (int p = 0; p < 6; p++) {
          const int64_t base = off + (((int64_t)(cz[p] * nbk + cy[p]) * nbk + cx[p]) << 9);
          for (int q = tid; q < 512; q += 256) {
            cr[p][q] = P.Wr[base + q];
            ci[p][q] = P.Wi[base + q];
            if constexpr (HJ) {
              jr[p][q] = P.WJr[base + q];
              ji[p][q] = P.WJi[base + q];
            }
          }
        }
        __syncthreads();
#pragma unroll 1
        for (int pt = tid; pt < 512; pt += 256) {
          const int li = pt & 7, lj = (pt >> 3) & 7, lk = pt >> 6;
          const int i = B1 * 8 + li, j = B2 * 8 + lj, k = B3 * 8 + lk;
          if (i >= no || j >= no || k >= no || i < j || j < k) continue;
          const double facjk = j == k ? 0.5 : 1.0, facij = i == j ? 0.5 : 1.0;
          const cplx den = epsabc - ((cplx{P.eir[i], P.eii[i]} + cplx{P.eir[j], P.eii[j]}) +
                                     cplx{P.eir[k], P.eii[k]});
          const int o[6] = {li + 8 * lj + 64 * lk, li + 8 * lk + 64 * lj,
                            lj + 8 * li + 64 * lk, lj + 8 * lk + 64 * li,
                            lk + 8 * li + 64 * lj, lk + 8 * lj + 64 * li};
          const int px[6] = {i, i, j, j, k, k}, py[6] = {j, k, i, k, i, j},
                    pz[6] = {k, j, k, i, j, i};
          cplx Wp[6], Zp[6];
#pragma unroll
          for (int p = 0; p < 6; p++) {
            Wp[p] = cplx{cr[p][o[p]], ci[p][o[p]]};
            // Z = W + t_a[x] V^bc_yz + t_b[y] V^ac_xz + t_c[z] V^ab_xy
            cplx z = Wp[p] + tai(a, px[p]) * vhh(pbc, py[p], pz[p]);
            z = z + tai(b, py[p]) * vhh(pac, px[p], pz[p]);
            z = z + tai(c, pz[p]) * vhh(pab, px[p], py[p]);
            Zp[p] = z;
          }
          auto point = [&](const cplx *Wq) {
            cplx value;
            if (distinct == 0) {
              const cplx A = conj(Wq[0]), B = conj(Wq[1]), C = conj(Wq[2]),
                         D = conj(Wq[3]), E = conj(Wq[4]), F = conj(Wq[5]);
              const cplx U = Zp[0], V = Zp[1], W = Zp[2], X = Zp[3], Y = Zp[4],
                         Z = Zp[5];
              const cplx UXY = U + (X + Y), VWZ = V + (W + Z);
              const cplx first = A * U + (B * V + (C * W + (D * X + (E * Y + F * Z))));
              value = 3.0 * first + ((UXY - 2.0 * VWZ) * (A + (D + E)) +
                                     (VWZ - 2.0 * UXY) * (B + (C + F)));
            } else {
              const cplx A = conj(Wq[0]), B = conj(Wq[3]), C = conj(Wq[4]);
              const cplx U = Zp[0], V = Zp[3], W = Zp[4];
              value = 3.0 * ((A * U + B * V) + C * W) - (A + (B + C)) * (U + (V + W));
            }
            return re_div(2.0 * value, den) * (facjk * facij);
          };
          e_acc += point(Wp);
          if constexpr (HJ) {
            cplx Jp[6];
#pragma unroll
            for (int p = 0; p < 6; p++) Jp[p] = cplx{jr[p][o[p]], ji[p][o[p]]};
            ect_acc += point(Jp);
          }
        }
      }
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
    P.e_slot[2 * gslot + 1] = HJ ? ect : e;
  }
}

}  // namespace

hipError_t launch_energy_z(const DevSystem &s, const BatchArgs &b,
                           double *const G[4], double *const GJ[4],
                           double tai_sign, double *e_slot, double *wcube,
                           hipStream_t st) {
  const int64_t W3c = w_cube(s.No), n = b.slots * W3c;
  EnergyZParams P;
  P.Wr = wcube;
  P.Wi = wcube + n;
  P.WJr = GJ ? wcube + 2 * n : nullptr;
  P.WJi = GJ ? wcube + 3 * n : nullptr;
  hipError_t e;
  if ((e = launch_wsum(s, b, G[0], G[1], false, (double *)P.Wr, st)) != hipSuccess ||
      (e = launch_wsum(s, b, G[2], G[3], false, (double *)P.Wi, st)) != hipSuccess)
    return e;
  if (GJ &&
      ((e = launch_wsum(s, b, GJ[0], GJ[1], false, (double *)P.WJr, st)) != hipSuccess ||
       (e = launch_wsum(s, b, GJ[2], GJ[3], false, (double *)P.WJi, st)) != hipSuccess))
    return e;
  P.Tair = s.Tai;
  P.Taii = s.Taii;
  P.VHHr = s.VHH;
  P.VHHi = s.VHHi;
  P.eir = s.ei;
  P.eii = s.eii;
  P.ear = s.ea;
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
    hipLaunchKernelGGL(energy_z_kernel<true>, grid, block, 0, st, P);
  else
    hipLaunchKernelGGL(energy_z_kernel<false>, grid, block, 0, st, P);
  return hipGetLastError();
}

}  // namespace atrip_amd
