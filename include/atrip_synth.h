/*
 * Counter-based synthetic tensor generator for (T) inputs.
 *
 * Every element is a pure function of (seed, tensor id, column-major CTF
 * linear index), so the host (C or C++) and the GPU (HIP device code)
 * produce bit-identical tensors without sharing any state.  The arithmetic
 * is restricted to IEEE-754 operations that are correctly rounded on both
 * sides (+ - * /, fma) plus a platform sqrt that is corrected to the
 * correctly rounded result; log and cos are evaluated here from their
 * series so no libm/ocml difference can leak in.  Contraction of a*b+c into
 * fma is switched off for this header (explicit fma() is used where wanted).
 *
 * Methods (published):
 *   - SplitMix64 (Steele, Lea, Flood, OOPSLA 2014): the n-th output of a
 *     stream with state s is mix(s + (n+1)*0x9E3779B97F4A7C15).
 *   - Box-Muller transform (1958): z = sqrt(-2 ln u1) cos(2 pi u2).
 *   - ln(m) = 2 atanh((m-1)/(m+1)) series; cos/sin by Taylor series after
 *     reduction to |theta| <= pi/4.
 *
 * Distributions (one per tensor id, standard deviation = second parameter):
 *   ei ~ U[-1.5,-0.3], ea ~ U[0.1,3.0], Tai ~ N(0,1e-3), Tabij ~ N(0,1e-2)
 *   with Tabij = Tbaji, Vabij ~ N(0,5e-2) with Vabij = Vbaji,
 *   Vijka, Vabci, Jijka, Jabci ~ N(0,5e-2).
 * Symmetrisation draws the element whose (a,i) pair index a+Nv*i is not
 * larger than (b,j)'s and mirrors it, so marginals stay unchanged.
 * Complex systems: the imaginary part of tensor t is tensor t |
 * ATRIP_SYN_IMAG, with t's distribution and symmetrisation from its own
 * stream; orbital energies stay real (imaginary part 0), as in the
 * reference bench's complex mode (bench/main.cxx:224-232).
 */
#ifndef ATRIP_SYNTH_H
#define ATRIP_SYNTH_H

#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define ATRIP_SYNTH_FN __host__ __device__ static inline
#else
#define ATRIP_SYNTH_FN static inline
#endif

#if defined(__clang__)
#define ATRIP_SYNTH_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define ATRIP_SYNTH_NOCONTRACT
#endif

enum atrip_synth_tensor {
  ATRIP_SYN_EI = 0,
  ATRIP_SYN_EA = 1,
  ATRIP_SYN_TAI = 2,
  ATRIP_SYN_TABIJ = 3,
  ATRIP_SYN_VABIJ = 4,
  ATRIP_SYN_VIJKA = 5,
  ATRIP_SYN_VABCI = 6,
  ATRIP_SYN_JIJKA = 7,
  ATRIP_SYN_JABCI = 8,
  ATRIP_SYN_IMAG = 16 /* flag: imaginary part of the tensor */
};

ATRIP_SYNTH_FN uint64_t atrip_synth_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* State of the stream that feeds one tensor. */
ATRIP_SYNTH_FN uint64_t atrip_synth_stream(uint64_t seed, int tensor) {
  return atrip_synth_mix64(atrip_synth_mix64(seed + 0x9E3779B97F4A7C15ULL)
                           ^ ((uint64_t)(tensor + 1) * 0xD1B54A32D192ED03ULL));
}

/* n-th SplitMix64 output of stream state s. */
ATRIP_SYNTH_FN uint64_t atrip_synth_raw(uint64_t s, uint64_t n) {
  return atrip_synth_mix64(s + (n + 1ULL) * 0x9E3779B97F4A7C15ULL);
}

ATRIP_SYNTH_FN double atrip_synth_bits2d(uint64_t b) {
  double d;
  __builtin_memcpy(&d, &b, sizeof d);
  return d;
}

ATRIP_SYNTH_FN uint64_t atrip_synth_d2bits(double d) {
  uint64_t b;
  __builtin_memcpy(&b, &d, sizeof b);
  return b;
}

/* Uniform in the open interval (0,1): 53 random bits, centred in the cell. */
ATRIP_SYNTH_FN double atrip_synth_u01(uint64_t s, uint64_t n) {
  ATRIP_SYNTH_NOCONTRACT
  const uint64_t r = atrip_synth_raw(s, n) >> 11;
  return ((double)r + 0.5) * (1.0 / 9007199254740992.0);
}

/* Natural log for x in (0, 1], from ln m = 2 atanh((m-1)/(m+1)). */
ATRIP_SYNTH_FN double atrip_synth_log(double x) {
  ATRIP_SYNTH_NOCONTRACT
  const uint64_t bits = atrip_synth_d2bits(x);
  int e = (int)((bits >> 52) & 0x7FF) - 1023;
  double m = atrip_synth_bits2d((bits & 0x000FFFFFFFFFFFFFULL)
                                | 0x3FF0000000000000ULL); /* [1,2) */
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e = e + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 1.0 / 23.0;
  p = fma(p, z, 1.0 / 21.0);
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  p = fma(p, z, 1.0);
  return fma((double)e, 0.6931471805599453, 2.0 * s * p);
}

/* cos(2 pi u) for u in [0, 1). */
ATRIP_SYNTH_FN double atrip_synth_cos2pi(double u) {
  ATRIP_SYNTH_NOCONTRACT
  /* y in [-0.5, 0.5); cos(2 pi u) = -cos(2 pi y) */
  const double y = u - 0.5;
  const double qd = floor(fma(4.0, y, 0.5));
  const int q = (int)qd;
  const double f = fma(-0.25, qd, y); /* [-1/8, 1/8] */
  const double t = f * 6.283185307179586;
  const double z = t * t;
  double c = 1.0 / 6402373705728000.0; /* 1/18! */
  c = fma(c, -z, 1.0 / 20922789888000.0);
  c = fma(c, -z, 1.0 / 87178291200.0);
  c = fma(c, -z, 1.0 / 479001600.0);
  c = fma(c, -z, 1.0 / 3628800.0);
  c = fma(c, -z, 1.0 / 40320.0);
  c = fma(c, -z, 1.0 / 720.0);
  c = fma(c, -z, 1.0 / 24.0);
  c = fma(c, -z, 1.0 / 2.0);
  c = fma(c, -z, 1.0);
  double sn = 1.0 / 121645100408832000.0; /* 1/19! */
  sn = fma(sn, -z, 1.0 / 355687428096000.0);
  sn = fma(sn, -z, 1.0 / 1307674368000.0);
  sn = fma(sn, -z, 1.0 / 6227020800.0);
  sn = fma(sn, -z, 1.0 / 39916800.0);
  sn = fma(sn, -z, 1.0 / 362880.0);
  sn = fma(sn, -z, 1.0 / 5040.0);
  sn = fma(sn, -z, 1.0 / 120.0);
  sn = fma(sn, -z, 1.0 / 6.0);
  sn = fma(sn, -z, 1.0);
  sn = sn * t;
  /* cos(2 pi y) = cos(t + q pi/2) */
  const int qm = q & 3;
  const double cy = qm == 0 ? c : qm == 1 ? -sn : qm == 2 ? -c : sn;
  return -cy;
}

/* Correctly rounded sqrt: platform result then one Tuckerman correction. */
ATRIP_SYNTH_FN double atrip_synth_sqrt(double x) {
  ATRIP_SYNTH_NOCONTRACT
  double y = sqrt(x);
  const double up = atrip_synth_bits2d(atrip_synth_d2bits(y) + 1ULL);
  const double dn = atrip_synth_bits2d(atrip_synth_d2bits(y) - 1ULL);
  const double r = fma(-y, y, x);
  if (r > y * (up - y))
    y = up;
  else if (r < -(y * (y - dn)))
    y = dn;
  return y;
}

/* Standard normal variate number n of stream s. */
ATRIP_SYNTH_FN double atrip_synth_normal(uint64_t s, uint64_t n) {
  ATRIP_SYNTH_NOCONTRACT
  const double u1 = atrip_synth_u01(s, 2ULL * n);
  const double u2 = atrip_synth_u01(s, 2ULL * n + 1ULL);
  const double r = atrip_synth_sqrt(-2.0 * atrip_synth_log(u1));
  return r * atrip_synth_cos2pi(u2);
}

/*
 * Value of element `idx` (column-major CTF linear index) of tensor `tensor`
 * for a system with No occupied and Nv virtual orbitals.  `s` is
 * atrip_synth_stream(seed, tensor).
 */
ATRIP_SYNTH_FN double atrip_synth_value(uint64_t s, int tensor, uint64_t No,
                                        uint64_t Nv, uint64_t idx) {
  ATRIP_SYNTH_NOCONTRACT
  const int base = tensor & 15;
  if ((tensor & ATRIP_SYN_IMAG) && base <= ATRIP_SYN_EA) return 0.0;
  switch (base) {
  case ATRIP_SYN_EI: return fma(1.2, atrip_synth_u01(s, idx), -1.5);
  case ATRIP_SYN_EA: return fma(2.9, atrip_synth_u01(s, idx), 0.1);
  case ATRIP_SYN_TAI: return 1e-3 * atrip_synth_normal(s, idx);
  case ATRIP_SYN_TABIJ:
  case ATRIP_SYN_VABIJ: {
    /* lens {Nv, Nv, No, No}: idx = a + Nv b + Nv^2 i + Nv^2 No j */
    const uint64_t NvNv = Nv * Nv;
    const uint64_t a = idx % Nv, b = (idx / Nv) % Nv;
    const uint64_t i = (idx / NvNv) % No, j = idx / (NvNv * No);
    uint64_t key = idx;
    if (a + Nv * i > b + Nv * j) key = b + Nv * a + NvNv * j + NvNv * No * i;
    const double sd = base == ATRIP_SYN_TABIJ ? 1e-2 : 5e-2;
    return sd * atrip_synth_normal(s, key);
  }
  default: return 5e-2 * atrip_synth_normal(s, idx);
  }
}

#endif /* ATRIP_SYNTH_H */
