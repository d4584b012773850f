// Device re-layout of CTF tensors and on-device synthetic generation.
//
// The reference cuts slices out of the CTF tensors with Tensor::slice into
// contiguous buffers once per run (src/atrip/Unions.hpp:22-75) and then
// ships them per tuple.  Here every tensor is re-laid out once, on the
// device, into the layouts documented on DevSystem (internal.hpp), and stays
// resident.  The same layouts can also be filled straight from the
// counter-based generator, which avoids host generation and PCIe traffic for
// large synthetic systems.
#include "../../include/atrip_synth.h"
#include "internal.hpp"

namespace atrip_amd {
namespace {

constexpr int TT = 32;  // transpose tile edge

// dst[o*dst_o + a*dst_a + b*dst_b] = scale * src[o*src_o + a*src_a + b*src_b]
// b is the contiguous source index, a the contiguous destination index.
__global__ __launch_bounds__(256) void transpose_kernel(
    const double *__restrict__ src, double *__restrict__ dst, int64_t n_a,
    int64_t n_b, int64_t tiles_a, int64_t src_o, int64_t src_a, int64_t src_b,
    int64_t dst_o, int64_t dst_a, int64_t dst_b, double scale) {
  __shared__ double tile[TT][TT + 1];
  const int64_t o = blockIdx.y;
  const int64_t ta = blockIdx.x % tiles_a, tb = blockIdx.x / tiles_a;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const double *s = src + o * src_o;
  double *d = dst + o * dst_o;
#pragma unroll
  for (int r = 0; r < TT; r += 8) {
    const int64_t a = ta * TT + ty + r, b = tb * TT + tx;
    if (a < n_a && b < n_b) tile[ty + r][tx] = s[a * src_a + b * src_b];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < TT; r += 8) {
    const int64_t b = tb * TT + ty + r, a = ta * TT + tx;
    if (a < n_a && b < n_b) d[a * dst_a + b * dst_b] = scale * tile[tx][ty + r];
  }
}

enum SynthTarget { S_T3, S_V3, S_H3, S_T2N, S_T2TN, S_VHH, S_PLAIN };

// Fill one device array from the generator; element q of the destination is
// mapped back to the CTF linear index of the tensor it holds.
__global__ __launch_bounds__(256) void synth_kernel(double *dst, float *sh,
                                                    uint16_t *sl, TriOp tri,
                                                    TriGeom tg, int64_t n,
                                                    int target, int tensor,
                                                    uint64_t stream, int64_t No,
                                                    int64_t Nv, int64_t Kv,
                                                    int64_t Ko, double scale) {
  const int64_t N2 = No * No, Nv2 = Nv * Nv;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    int64_t idx = -1;
    double sign = 1.0;
    switch (target) {
    case S_T3: {  // [x][row][E<Kv] = Tpphh[x,E,row]
      const int64_t E = q % Kv, r = q / Kv, row = r % N2, x = r / N2;
      if (E < Nv) idx = x + Nv * E + Nv2 * row;
      break;
    }
    case S_V3: {  // [pair][r][E<Kv] = Vppph[pair,E,r]
      const int64_t E = q % Kv, t = q / Kv, r = t % No, pair = t / No;
      if (E < Nv) idx = pair + Nv2 * E + Nv2 * Nv * r;
      break;
    }
    case S_H3: {  // [x][row][L<Ko] = Vhhhp[row,L,x]
      const int64_t L = q % Ko, t = q / Ko, row = t % N2, x = t / N2;
      if (L < No) idx = row + N2 * L + N2 * No * x;
      break;
    }
    case S_T2N: {  // [pair][n][L<Ko] = -Tpphh[pair,L,n]
      const int64_t L = q % Ko, t = q / Ko, nn = t % No, pair = t / No;
      if (L < No) idx = pair + Nv2 * L + Nv2 * No * nn;
      sign = -1.0;
      break;
    }
    case S_T2TN: {  // [pair][n][L<Ko] = -Tpphh[pair,n,L]
      const int64_t L = q % Ko, t = q / Ko, nn = t % No, pair = t / No;
      if (L < No) idx = pair + Nv2 * nn + Nv2 * No * L;
      sign = -1.0;
      break;
    }
    case S_VHH: {  // [pair][ij] = Vpphh[pair,ij]
      const int64_t ij = q % N2, pair = q / N2;
      idx = pair + Nv2 * ij;
      break;
    }
    default: idx = q;
    }
    double v = 0.0;
    if (idx >= 0)
      v = sign * atrip_synth_value(stream, tensor, (uint64_t)No, (uint64_t)Nv,
                                   (uint64_t)idx) * scale;
    if (dst) dst[q] = v;
    if (sh) split_store(v, sh, sl, q);
    if (tri.p) tri_store(v, tri, tg, No, q);
  }
}

// Fill n elements of a target layout into dst (FP64), the split arrays o
// (o.h null: none) and/or the bf16x3 copy tri of geometry tg (tri.p null:
// none).
hipError_t synth_fill(double *dst, const DevSystem::Split &o, const TriOp &tri,
                      const TriGeom &tg, int64_t n, SynthTarget target,
                      int tensor, uint64_t seed, const DevSystem &s,
                      double scale, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const uint64_t stream = atrip_synth_stream(seed, tensor);
  // amplitudes and integrals carry the scale, orbital energies do not
  const int base = tensor & 15;
  if (base == ATRIP_SYN_EI || base == ATRIP_SYN_EA) scale = 1.0;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536L * 16) blocks = 65536L * 16;
  hipLaunchKernelGGL(synth_kernel, dim3((unsigned)blocks), dim3(256), 0, st,
                     dst, o.h, o.l, tri, tg, n, (int)target, tensor, stream, s.No,
                     s.Nv, s.Kv, s.Ko, scale);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_transpose(const double *src, double *dst, int64_t n_outer,
                            int64_t n_a, int64_t n_b, int64_t src_o,
                            int64_t src_a, int64_t src_b, int64_t dst_o,
                            int64_t dst_a, int64_t dst_b, double scale,
                            hipStream_t st) {
  if (n_outer <= 0 || n_a <= 0 || n_b <= 0) return hipSuccess;
  const int64_t tiles_a = (n_a + TT - 1) / TT, tiles_b = (n_b + TT - 1) / TT;
  // blockIdx.y carries the outer index (<= 65535 per launch)
  for (int64_t o0 = 0; o0 < n_outer; o0 += 65535) {
    const int64_t no = n_outer - o0 < 65535 ? n_outer - o0 : 65535;
    hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)(tiles_a * tiles_b),
                                              (unsigned)no),
                       dim3(256), 0, st, src + o0 * src_o, dst + o0 * dst_o,
                       n_a, n_b, tiles_a, src_o, src_a, src_b, dst_o, dst_a,
                       dst_b, scale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_synth(DevSystem &s, uint64_t seed, bool with_j,
                        hipStream_t st, int tensor_flag, double scale) {
  const int64_t No = s.No, Nv = s.Nv, N2 = No * No, Nv2 = Nv * Nv;
  hipError_t e;
  const DevSystem::Split none;
  const TriOp tnone;
  const TriGeom ga = tri_geom_a(No, Nv, s.Kv), gb = tri_geom_b(No, Nv, s.Kv),
                gh = tri_geom_a(No, Nv, s.Ko), gt = tri_geom_b(No, Nv, s.Ko);
  // doubles operands: FP64 array, or straight into the split arrays or the
  // bf16x3 copy when the system keeps no FP64 copy
#define SYN(ptr, sp, tp, g, tgt, tensor)                                    \
  if ((e = synth_fill(s.store == 0 ? ptr : nullptr,                         \
                      s.store == 1 || s.store == 2 ? sp : none,             \
                      s.store == 3 ? tp : tnone, g, (g).no * (g).nr * (g).K, \
                      tgt, (tensor) | tensor_flag, seed, s, scale, st)) !=  \
      hipSuccess)                                                           \
    return e;
#define SYN64(ptr, n, tgt, tensor)                                         \
  if ((e = synth_fill(ptr, none, tnone, ga, n, tgt, (tensor) | tensor_flag, \
                      seed, s, scale, st)) != hipSuccess)                   \
    return e;
  SYN(s.T3, s.sT3, s.tT3, ga, S_T3, ATRIP_SYN_TABIJ);
  SYN(s.V3, s.sV3, s.tV3, gb, S_V3, ATRIP_SYN_VABCI);
  SYN(s.H3, s.sH3, s.tH3, gh, S_H3, ATRIP_SYN_VIJKA);
  SYN(s.T2N, s.sT2N, s.tT2N, gt, S_T2N, ATRIP_SYN_TABIJ);
  SYN(s.T2TN, s.sT2TN, s.tT2TN, gt, S_T2TN, ATRIP_SYN_TABIJ);
  SYN64(s.VHH, Nv2 * N2, S_VHH, ATRIP_SYN_VABIJ);
  SYN64(s.Tai, Nv * No, S_PLAIN, ATRIP_SYN_TAI);
  SYN64(s.ei, No, S_PLAIN, ATRIP_SYN_EI);
  SYN64(s.ea, Nv, S_PLAIN, ATRIP_SYN_EA);
  if (with_j) {
    SYN(s.V3J, s.sV3J, s.tV3J, gb, S_V3, ATRIP_SYN_JABCI);
    SYN(s.H3J, s.sH3J, s.tH3J, gh, S_H3, ATRIP_SYN_JIJKA);
  }
#undef SYN
#undef SYN64
  return hipSuccess;
}

__global__ __launch_bounds__(256) void scale_kernel(const double *src,
                                                    double *dst, int64_t n,
                                                    double scale) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    dst[q] = scale * src[q];
}

hipError_t launch_scale(const double *src, double *dst, int64_t n,
                        double scale, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536L * 16) blocks = 65536L * 16;
  hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(256), 0, st,
                     src, dst, n, scale);
  return hipGetLastError();
}

}  // namespace atrip_amd
