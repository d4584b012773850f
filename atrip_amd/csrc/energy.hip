// Fused singles + energy epilogue of the (T) hot path.
//
// One workgroup per tuple slot of a window batch.  W is the sum of the three
// class buffers written by the doubles kernel, Z = W + singles is formed in
// registers (reference src/atrip/Equations.cxx:387-426, with Z copied from W
// as in src/atrip/Atrip.cxx:899-906), and the energy of the tuple follows
// get_energy_distinct / get_energy_same (src/atrip/Equations.cxx:101-238)
// selected as src/atrip/Atrip.cxx:636-670 does.  The i>=j>=k points are
// spread over the workgroup; each thread sums its points in a fixed order
// and a fixed shuffle/LDS tree combines the threads, so the per-tuple energy
// does not depend on which other tuples share the launch.
#include "internal.hpp"

namespace atrip_amd {
namespace {

struct EnergyParams {
  const double *G, *GJ;
  const double *Tai, *VHH, *ei, *ea;
  const uint32_t *points;
  const int32_t *windows;
  const uint64_t *aprefix;
  double *e_slot;
  int64_t No, Nv, N2, N3, npoints, slots;
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

__global__ __launch_bounds__(256) void energy_kernel(EnergyParams P) {
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

  extern __shared__ double sm[];
  double *tA = sm, *tB = sm + No, *tC = sm + 2 * No;
  for (int64_t q = threadIdx.x; q < No; q += blockDim.x) {
    tA[q] = P.tai_sign * P.Tai[a + Nv * q];
    tB[q] = P.tai_sign * P.Tai[b + Nv * q];
    tC[q] = P.tai_sign * P.Tai[c + Nv * q];
  }
  // V(xy)ij slabs of the singles term stay in global memory (L1/L2 hits)
  const double *Vbc = P.VHH + (b + Nv * c) * N2;
  const double *Vac = P.VHH + (a + Nv * c) * N2;
  const double *Vab = P.VHH + (a + Nv * b) * N2;
  __syncthreads();

  const double *Gw = P.G + (0 * P.slots + gslot) * P.N3;
  const double *Gv = P.G + (1 * P.slots + gslot) * P.N3;
  const double *Gu = P.G + (2 * P.slots + gslot) * P.N3;
  const bool have_j = P.GJ != nullptr;
  const double *Jw = have_j ? P.GJ + (0 * P.slots + gslot) * P.N3 : Gw;
  const double *Jv = have_j ? P.GJ + (1 * P.slots + gslot) * P.N3 : Gv;
  const double *Ju = have_j ? P.GJ + (2 * P.slots + gslot) * P.N3 : Gu;

  int distinct = 0;
  if (a == b) distinct++;
  if (b == c) distinct--;
  const double epsabc = (P.ea[a] + P.ea[b]) + P.ea[c];

  auto Wat = [&](int64_t x, int64_t y, int64_t z) {
    const int64_t q = x + No * y + N2 * z;
    return (Gw[q] + Gv[q]) + Gu[q];
  };
  auto WJat = [&](int64_t x, int64_t y, int64_t z) {
    const int64_t q = x + No * y + N2 * z;
    return (Jw[q] + Jv[q]) + Ju[q];
  };
  auto Zat = [&](double wv, int64_t x, int64_t y, int64_t z) {
    double zz = wv + tA[x] * Vbc[y + No * z];
    zz = zz + tB[y] * Vac[x + No * z];
    zz = zz + tC[z] * Vab[x + No * y];
    return zz;
  };

  double e_acc = 0.0, ect_acc = 0.0;
  for (int64_t pt = threadIdx.x; pt < P.npoints; pt += blockDim.x) {
    const uint32_t packed = P.points[pt];
    const int64_t i = packed & 1023u, j = (packed >> 10) & 1023u,
                  k = (packed >> 20) & 1023u;
    const double ei = P.ei[i], ej = P.ei[j], ek = P.ei[k];
    const double facjk = j == k ? 0.5 : 1.0;
    const double facij = i == j ? 0.5 : 1.0;
    const double eijk = (ei + ej) + ek;
    const double denominator = epsabc - eijk;
    if (distinct == 0) {
      const double A = Wat(i, j, k), B = Wat(i, k, j), C = Wat(j, i, k),
                   D = Wat(j, k, i), E = Wat(k, i, j), F = Wat(k, j, i);
      const double U = Zat(A, i, j, k), V = Zat(B, i, k, j),
                   W = Zat(C, j, i, k), X = Zat(D, j, k, i),
                   Y = Zat(E, k, i, j), Z = Zat(F, k, j, i);
      const double UXY = U + (X + Y), VWZ = V + (W + Z);
      {
        const double first =
            A * U + (B * V + (C * W + (D * X + (E * Y + F * Z))));
        const double second = (UXY - 2.0 * VWZ) * (A + (D + E));
        const double third = (VWZ - 2.0 * UXY) * (B + (C + F));
        const double value = 3.0 * first + (second + third);
        e_acc += ((2.0 * value) / denominator) * (facjk * facij);
      }
      if (have_j) {
        const double A2 = WJat(i, j, k), B2 = WJat(i, k, j),
                     C2 = WJat(j, i, k), D2 = WJat(j, k, i),
                     E2 = WJat(k, i, j), F2 = WJat(k, j, i);
        const double first =
            A2 * U + (B2 * V + (C2 * W + (D2 * X + (E2 * Y + F2 * Z))));
        const double second = (UXY - 2.0 * VWZ) * (A2 + (D2 + E2));
        const double third = (VWZ - 2.0 * UXY) * (B2 + (C2 + F2));
        const double value = 3.0 * first + (second + third);
        ect_acc += ((2.0 * value) / denominator) * (facjk * facij);
      }
    } else {
      const double A = Wat(i, j, k), B = Wat(j, k, i), C = Wat(k, i, j);
      const double U = Zat(A, i, j, k), V = Zat(B, j, k, i),
                   W = Zat(C, k, i, j);
      const double UVW = U + (V + W);
      {
        const double value = 3.0 * ((A * U + B * V) + C * W) - (A + (B + C)) * UVW;
        e_acc += ((2.0 * value) / denominator) * (facjk * facij);
      }
      if (have_j) {
        const double A2 = WJat(i, j, k), B2 = WJat(j, k, i),
                     C2 = WJat(k, i, j);
        const double value =
            3.0 * ((A2 * U + B2 * V) + C2 * W) - (A2 + (B2 + C2)) * UVW;
        ect_acc += ((2.0 * value) / denominator) * (facjk * facij);
      }
    }
  }

  __shared__ double red[2][4];
  e_acc = wave_sum(e_acc);
  ect_acc = wave_sum(ect_acc);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    red[0][wave] = e_acc;
    red[1][wave] = ect_acc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double e = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    const double ect = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    P.e_slot[2 * gslot] = e;
    P.e_slot[2 * gslot + 1] = have_j ? ect : e;
  }
}

// Neumaier sum of the slot energies of one batch, in slot order.
__global__ void batch_reduce_kernel(const double *e_slot, int64_t slots,
                                    double *partials, int index) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s0 = 0.0, c0 = 0.0, s1 = 0.0, c1 = 0.0;
  for (int64_t q = 0; q < slots; q++) {
    const double x0 = e_slot[2 * q], x1 = e_slot[2 * q + 1];
    double t = s0 + x0;
    c0 += (fabs(s0) >= fabs(x0)) ? (s0 - t) + x0 : (x0 - t) + s0;
    s0 = t;
    t = s1 + x1;
    c1 += (fabs(s1) >= fabs(x1)) ? (s1 - t) + x1 : (x1 - t) + s1;
    s1 = t;
  }
  partials[4 * index + 0] = s0;
  partials[4 * index + 1] = c0;
  partials[4 * index + 2] = s1;
  partials[4 * index + 3] = c1;
}

}  // namespace

hipError_t launch_energy(const DevSystem &s, const BatchArgs &b,
                         const double *G, const double *GJ, double tai_sign,
                         double *e_slot, hipStream_t st) {
  EnergyParams P;
  P.G = G;
  P.GJ = GJ;
  P.Tai = s.Tai;
  P.VHH = s.VHH;
  P.ei = s.ei;
  P.ea = s.ea;
  P.points = s.points;
  P.windows = b.windows;
  P.aprefix = s.aprefix;
  P.e_slot = e_slot;
  P.No = s.No;
  P.Nv = s.Nv;
  P.N2 = s.No * s.No;
  P.N3 = P.N2 * s.No;
  P.npoints = s.npoints;
  P.slots = b.slots;
  P.first = b.first;
  P.last = b.last;
  P.nb = b.nb;
  P.tai_sign = tai_sign;
  const size_t shm = sizeof(double) * (3 * s.No);
  hipLaunchKernelGGL(energy_kernel, dim3((unsigned)b.slots), dim3(256), shm,
                     st, P);
  return hipGetLastError();
}

hipError_t launch_batch_reduce(const DevSystem &, const BatchArgs &b,
                               const double *e_slot, double *partials,
                               int index, hipStream_t st) {
  hipLaunchKernelGGL(batch_reduce_kernel, dim3(1), dim3(64), 0, st, e_slot,
                     b.slots, partials, index);
  return hipGetLastError();
}

}  // namespace atrip_amd
