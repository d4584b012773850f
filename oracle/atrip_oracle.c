/*
 * CPU oracle for atrip's (T) hot path -- TEST INFRASTRUCTURE ONLY.
 * See atrip_oracle.h for the contract and the reference lines each function
 * restates.  Built with -ffp-contract=off so that expression order (not the
 * compiler) decides rounding.
 */
#define _GNU_SOURCE
#include "atrip_oracle.h"

#include <dlfcn.h>
#include <complex.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "../include/atrip_synth.h"

/* ------------------------------------------------------------------ */
/* Tuples                                                              */
/* ------------------------------------------------------------------ */

uint64_t oracle_n_tuples(uint64_t Nv) {
  /* src/atrip/Tuples.cxx:93 */
  return Nv * (Nv + 1) * (Nv + 2) / 6 - Nv;
}

void oracle_all_tuples(uint64_t Nv, uint64_t *abc) {
  /* src/atrip/Tuples.cxx:121-134: lexicographic, a=b=c skipped */
  uint64_t u = 0;
  for (uint64_t a = 0; a < Nv; a++)
    for (uint64_t b = a; b < Nv; b++)
      for (uint64_t c = b; c < Nv; c++) {
        if (a == b && b == c) continue;
        abc[3 * u + 0] = a;
        abc[3 * u + 1] = b;
        abc[3 * u + 2] = c;
        u++;
      }
}

/* Number of tuples whose first index is <= T (cumulative a-block size),
 * src/atrip/DatabaseCommunicator.cxx:51-61.  The grouping of the integer
 * divisions is kept: only the sum of the last two parts is even. */
static int64_t a_block_cumulative(int64_t T, int64_t nv) {
  const int64_t n1 = nv - 1, t1 = T + 1;
  return t1 * n1 * nv + n1 * t1 - (n1 * (T * t1) / 2) - (t1 * (n1 * nv) / 2)
       + (((T * t1 * (1 + 2 * T)) / 6) - 3 * ((T * t1) / 2)) / 2;
}

/* Number of tuples (a, b', c) with a <= b' <= T inside the a-block,
 * src/atrip/DatabaseCommunicator.cxx:63-65. */
static int64_t b_block_cumulative(int64_t a, int64_t T, int64_t nv) {
  return nv * ((T - a) + 1) - (T * (T + 1) - a * (a - 1)) / 2 - 1;
}

void oracle_nth_tuple(uint64_t it, uint64_t Nv, uint64_t out[3]) {
  /* src/atrip/DatabaseCommunicator.cxx:67-106, recomputed per call */
  const int64_t nv = (int64_t)Nv, target = (int64_t)it;
  int64_t a = -1, block_a = 0;
  for (int64_t t = 0; t < nv; t++) {
    const int64_t s = a_block_cumulative(t, nv);
    ++a;
    if (s > target) break;
    block_a = s;
  }
  int64_t b = a - 1, block_b = block_a;
  for (int64_t t = a; t < nv; t++) {
    const int64_t s = b_block_cumulative(a, t, nv);
    ++b;
    if (s + block_a > target) break;
    block_b = s + block_a;
  }
  const int64_t c = b + target - block_b + (a == b);
  out[0] = (uint64_t)a;
  out[1] = (uint64_t)b;
  out[2] = (uint64_t)c;
}

/* ------------------------------------------------------------------ */
/* BLAS                                                                */
/* ------------------------------------------------------------------ */

typedef void (*dgemm_fn)(const char *, const char *, const int *, const int *,
                         const int *, const double *, const double *,
                         const int *, const double *, const int *,
                         const double *, double *, const int *);
static dgemm_fn g_dgemm = NULL;
static const char *g_blas_name = NULL;

const char *oracle_blas_init(const char *path) {
  static const char *candidates[] = {
      "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so", "libopenblas.so.0",
      "libopenblas.so", "libblas.so.3", NULL};
  if (g_dgemm) return g_blas_name;
  const char *single[] = {path, NULL};
  const char **list = path ? single : candidates;
  for (int i = 0; list[i]; i++) {
    void *h = dlopen(list[i], RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    dgemm_fn f = (dgemm_fn)dlsym(h, "dgemm_");
    if (!f) {
      dlclose(h);
      continue;
    }
    /* one BLAS thread per caller thread */
    void (*mkl_threads)(int) = (void (*)(int))dlsym(h, "MKL_Set_Num_Threads");
    if (mkl_threads) mkl_threads(1);
    void (*ob_threads)(int) =
        (void (*)(int))dlsym(h, "openblas_set_num_threads");
    if (ob_threads) ob_threads(1);
    g_dgemm = f;
    g_blas_name = list[i];
    return g_blas_name;
  }
  return NULL;
}

const char *oracle_blas_name(void) { return g_blas_name; }

void oracle_dgemm(char transa, char transb, int m, int n, int k, double alpha,
                  const double *A, int lda, const double *B, int ldb,
                  double beta, double *C, int ldc) {
  if (g_dgemm) {
    const char ta[2] = {transa, 0}, tb[2] = {transb, 0};
    g_dgemm(ta, tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
    return;
  }
  const int ta = (transa == 'T' || transa == 't');
  const int tb = (transb == 'T' || transb == 't');
  for (int j = 0; j < n; j++) {
    double *cj = C + (size_t)j * ldc;
    for (int i = 0; i < m; i++) cj[i] = beta == 0.0 ? 0.0 : beta * cj[i];
    for (int l = 0; l < k; l++) {
      const double bl = tb ? B[j + (size_t)l * ldb] : B[l + (size_t)j * ldb];
      const double s = alpha * bl;
      if (!ta) {
        const double *al = A + (size_t)l * lda;
        for (int i = 0; i < m; i++) cj[i] += al[i] * s;
      } else {
        for (int i = 0; i < m; i++) cj[i] += A[l + (size_t)i * lda] * s;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* Doubles                                                             */
/* ------------------------------------------------------------------ */

typedef struct {
  int mistake;
} oracle_opts;

static void doubles_loop_impl(size_t No, size_t Nv, const double *VABph,
                              const double *VACph, const double *VBCph,
                              const double *VBAph, const double *VCAph,
                              const double *VCBph, const double *VhhhA,
                              const double *VhhhB, const double *VhhhC,
                              const double *TAphh, const double *TBphh,
                              const double *TCphh, const double *TABhh,
                              const double *TAChh, const double *TBChh,
                              double *Tijk, int mistake) {
  /* src/atrip/Equations.cxx:686-726 */
  const size_t NoNo = No * No, NoNv = No * Nv;
  for (size_t k = 0; k < No; k++)
    for (size_t j = 0; j < No; j++)
      for (size_t i = 0; i < No; i++) {
        const size_t ijk = i + j * No + k * NoNo;
        double w = 0.0;
        for (size_t L = 0; L < No; L++) {
          const double tab_h1 = mistake == ORACLE_MISTAKE_H1_TRANSPOSE
                                    ? TABhh[j + L * No]
                                    : TABhh[L + j * No];
          w -= tab_h1 * VhhhC[i + k * No + L * NoNo];                 /* H1 */
          if (mistake == ORACLE_MISTAKE_H0_SIGN)
            w += TABhh[i + L * No] * VhhhC[j + k * No + L * NoNo];
          else
            w -= TABhh[i + L * No] * VhhhC[j + k * No + L * NoNo];    /* H0 */
          w -= TAChh[L + k * No] * VhhhB[i + j * No + L * NoNo];      /* H5 */
          w -= TAChh[i + L * No] * VhhhB[k + j * No + L * NoNo];      /* H3 */
          w -= TBChh[L + k * No] * VhhhA[j + i * No + L * NoNo];      /* H2 */
          w -= TBChh[j + L * No] * VhhhA[k + i * No + L * NoNo];      /* H4 */
        }
        for (size_t E = 0; E < Nv; E++) {
          w += TAphh[E + i * Nv + j * NoNv] * VBCph[E + k * Nv];      /* P0 */
          if (mistake == ORACLE_MISTAKE_P3_PERM)
            w += TAphh[E + i * Nv + j * NoNv] * VCBph[E + k * Nv];
          else
            w += TAphh[E + i * Nv + k * NoNv] * VCBph[E + j * Nv];    /* P3 */
          w += TCphh[E + k * Nv + i * NoNv] * VABph[E + j * Nv];      /* P5 */
          w += TCphh[E + k * Nv + j * NoNv] * VBAph[E + i * Nv];      /* P2 */
          w += TBphh[E + j * Nv + i * NoNv] * VACph[E + k * Nv];      /* P1 */
          if (mistake != ORACLE_MISTAKE_DROP_P4)
            w += TBphh[E + j * Nv + k * NoNv] * VCAph[E + i * Nv];    /* P4 */
        }
        Tijk[ijk] = w;
      }
}

void oracle_doubles_loop(size_t No, size_t Nv, const double *VABph,
                         const double *VACph, const double *VBCph,
                         const double *VBAph, const double *VCAph,
                         const double *VCBph, const double *VhhhA,
                         const double *VhhhB, const double *VhhhC,
                         const double *TAphh, const double *TBphh,
                         const double *TCphh, const double *TABhh,
                         const double *TAChh, const double *TBChh,
                         double *Tijk) {
  doubles_loop_impl(No, Nv, VABph, VACph, VBCph, VBAph, VCAph, VCBph, VhhhA,
                    VhhhB, VhhhC, TAphh, TBphh, TCphh, TABhh, TAChh, TBChh,
                    Tijk, ORACLE_OK);
}

/* to[i + j N + k N^2] += from[perm(i,j,k)], src/atrip/Equations.cxx:26-89 */
enum { R_IJK, R_IKJ, R_JIK, R_JKI, R_KIJ, R_KJI };
static void reorder_add(int r, size_t N, double *to, const double *from) {
  const size_t NN = N * N;
  size_t idx = 0;
  for (size_t k = 0; k < N; k++)
    for (size_t j = 0; j < N; j++)
      for (size_t i = 0; i < N; i++, idx++) {
        size_t s;
        switch (r) {
        case R_IJK: s = i + j * N + k * NN; break;
        case R_IKJ: s = i + k * N + j * NN; break;
        case R_JIK: s = j + i * N + k * NN; break;
        case R_JKI: s = j + k * N + i * NN; break;
        case R_KIJ: s = k + i * N + j * NN; break;
        default: s = k + j * N + i * NN; break;
        }
        to[idx] += from[s];
      }
}

void oracle_doubles_gemm(size_t No, size_t Nv, const double *VABph,
                         const double *VACph, const double *VBCph,
                         const double *VBAph, const double *VCAph,
                         const double *VCBph, const double *VhhhA,
                         const double *VhhhB, const double *VhhhC,
                         const double *TAphh, const double *TBphh,
                         const double *TCphh, const double *TABhh,
                         const double *TAChh, const double *TBChh,
                         double *Tijk, double *tbuf, double *vbuf) {
  /* src/atrip/Equations.cxx:585-680 (DGEMM branch, CPU macros :538-578) */
  const size_t N3 = No * No * No;
  const int NoNo = (int)(No * No), no = (int)No, nv = (int)Nv;
  memset(tbuf, 0, N3 * sizeof(double));
  memset(vbuf, 0, N3 * sizeof(double));
  memset(Tijk, 0, N3 * sizeof(double));
#define HOLES(V, T, TR, R)                                                     \
  do {                                                                         \
    memcpy(vbuf, V, N3 * sizeof(double));                                      \
    oracle_dgemm('N', TR, NoNo, no, no, -1.0, vbuf, NoNo, T, no, 0.0, tbuf,    \
                 NoNo);                                                        \
    reorder_add(R, No, Tijk, tbuf);                                            \
  } while (0)
#define PARTICLES(TX, VX, R)                                                   \
  do {                                                                         \
    oracle_dgemm('T', 'N', NoNo, no, nv, 1.0, TX, nv, VX, nv, 0.0, tbuf,       \
                 NoNo);                                                        \
    reorder_add(R, No, Tijk, tbuf);                                            \
  } while (0)
  HOLES(VhhhC, TABhh, 'N', R_IKJ); /* H1 */
  HOLES(VhhhC, TABhh, 'T', R_JKI); /* H0 */
  HOLES(VhhhB, TAChh, 'N', R_IJK); /* H5 */
  HOLES(VhhhB, TAChh, 'T', R_KJI); /* H3 */
  HOLES(VhhhA, TBChh, 'N', R_JIK); /* H2 */
  HOLES(VhhhA, TBChh, 'T', R_KIJ); /* H4 */
  PARTICLES(TAphh, VBCph, R_IJK);  /* P0 */
  PARTICLES(TAphh, VCBph, R_IKJ);  /* P3 */
  PARTICLES(TCphh, VABph, R_KIJ);  /* P5 */
  PARTICLES(TCphh, VBAph, R_KJI);  /* P2 */
  PARTICLES(TBphh, VACph, R_JIK);  /* P1 */
  PARTICLES(TBphh, VCAph, R_JKI);  /* P4 */
#undef HOLES
#undef PARTICLES
}

/* ------------------------------------------------------------------ */
/* Singles and energies                                                */
/* ------------------------------------------------------------------ */

static void singles_impl(size_t No, size_t Nv, size_t a, size_t b, size_t c,
                         const double *Tph, const double *VABij,
                         const double *VACij, const double *VBCij,
                         double *Zijk, int mistake) {
  /* src/atrip/Equations.cxx:387-426 */
  const size_t NoNo = No * No;
  for (size_t k = 0; k < No; k++)
    for (size_t i = 0; i < No; i++)
      for (size_t j = 0; j < No; j++) {
        const size_t ijk = i + j * No + k * NoNo;
        Zijk[ijk] += Tph[a + i * Nv] * VBCij[j + k * No];
        if (mistake == ORACLE_MISTAKE_SINGLES_IDX)
          Zijk[ijk] += Tph[b + j * Nv] * VACij[k + i * No];
        else
          Zijk[ijk] += Tph[b + j * Nv] * VACij[i + k * No];
        Zijk[ijk] += Tph[c + k * Nv] * VABij[i + j * No];
      }
}

void oracle_singles(size_t No, size_t Nv, size_t a, size_t b, size_t c,
                    const double *Tph, const double *VABij,
                    const double *VACij, const double *VBCij, double *Zijk) {
  singles_impl(No, Nv, a, b, c, Tph, VABij, VACij, VBCij, Zijk, ORACLE_OK);
}

#define ORACLE_MIN(x, y) ((x) < (y) ? (x) : (y))

static double energy_distinct_impl(double epsabc, size_t No,
                                   const double *epsi, const double *Tijk,
                                   const double *Zijk, int mistake) {
  /* src/atrip/Equations.cxx:101-180, same blocking and expression order */
  const size_t bs = 16;
  double energy = 0.0;
  for (size_t kk = 0; kk < No; kk += bs) {
    const size_t kend = ORACLE_MIN(No, kk + bs);
    for (size_t jj = kk; jj < No; jj += bs) {
      const size_t jend = ORACLE_MIN(No, jj + bs);
      for (size_t ii = jj; ii < No; ii += bs) {
        const size_t iend = ORACLE_MIN(No, ii + bs);
        for (size_t k = kk; k < kend; k++) {
          const double ek = epsi[k];
          const size_t jstart = jj > k ? jj : k;
          for (size_t j = jstart; j < jend; j++) {
            const double ej = epsi[j];
            const double facjk = j == k ? 0.5 : 1.0;
            const size_t istart = ii > j ? ii : j;
            for (size_t i = istart; i < iend; i++) {
              const double ei = epsi[i];
              const double facij =
                  (i == j && mistake != ORACLE_MISTAKE_FACTOR) ? 0.5 : 1.0;
              const double eijk = (ei + ej) + ek;
              const double denominator = epsabc - eijk;
              const double U = Zijk[i + No * j + No * No * k];
              const double V = Zijk[i + No * k + No * No * j];
              const double W = Zijk[j + No * i + No * No * k];
              const double X = Zijk[j + No * k + No * No * i];
              const double Y = Zijk[k + No * i + No * No * j];
              const double Z = Zijk[k + No * j + No * No * i];
              const double A = Tijk[i + No * j + No * No * k];
              const double B = Tijk[i + No * k + No * No * j];
              const double C = Tijk[j + No * i + No * No * k];
              const double D = Tijk[j + No * k + No * No * i];
              const double E = Tijk[k + No * i + No * No * j];
              const double F = Tijk[k + No * j + No * No * i];
              const double AU = A * U, BV = B * V, CW = C * W, DX = D * X,
                           EY = E * Y, FZ = F * Z;
              const double UXY = U + (X + Y), VWZ = V + (W + Z);
              const double ADE = A + (D + E), BCF = B + (C + F);
              const double first = AU + (BV + (CW + (DX + (EY + FZ))));
              const double second = (UXY - 2.0 * VWZ) * ADE;
              const double third = (VWZ - 2.0 * UXY) * BCF;
              const double value = 3.0 * first + (second + third);
              const double loop_energy =
                  ((2.0 * value) / denominator) * (facjk * facij);
              energy += loop_energy;
            }
          }
        }
      }
    }
  }
  return energy;
}

static double energy_same_impl(double epsabc, size_t No, const double *epsi,
                               const double *Tijk, const double *Zijk,
                               int mistake) {
  /* src/atrip/Equations.cxx:182-238 */
  const size_t bs = 16;
  double energy = 0.0;
  for (size_t kk = 0; kk < No; kk += bs) {
    const size_t kend = ORACLE_MIN(kk + bs, No);
    for (size_t jj = kk; jj < No; jj += bs) {
      const size_t jend = ORACLE_MIN(jj + bs, No);
      for (size_t ii = jj; ii < No; ii += bs) {
        const size_t iend = ORACLE_MIN(ii + bs, No);
        for (size_t k = kk; k < kend; k++) {
          const double ek = epsi[k];
          const size_t jstart = jj > k ? jj : k;
          for (size_t j = jstart; j < jend; j++) {
            const double facjk = j == k ? 0.5 : 1.0;
            const double ej = epsi[j];
            const size_t istart = ii > j ? ii : j;
            for (size_t i = istart; i < iend; i++) {
              const double ei = epsi[i];
              const double facij =
                  (i == j && mistake != ORACLE_MISTAKE_FACTOR) ? 0.5 : 1.0;
              const double eijk = (ei + ej) + ek;
              const double denominator = epsabc - eijk;
              const double U = Zijk[i + No * j + No * No * k];
              const double V = Zijk[j + No * k + No * No * i];
              const double W = Zijk[k + No * i + No * No * j];
              const double A = Tijk[i + No * j + No * No * k];
              const double B = Tijk[j + No * k + No * No * i];
              const double C = Tijk[k + No * i + No * No * j];
              const double ABC = A + (B + C), UVW = U + (V + W);
              const double AU = A * U, BV = B * V, CW = C * W;
              const double sum3 = (AU + BV) + CW;
              const double value = 3.0 * sum3 - ABC * UVW;
              const double loop_energy =
                  ((2.0 * value) / denominator) * (facjk * facij);
              energy += loop_energy;
            }
          }
        }
      }
    }
  }
  return energy;
}

double oracle_energy_distinct(double epsabc, size_t No, const double *epsi,
                              const double *Tijk, const double *Zijk) {
  return energy_distinct_impl(epsabc, No, epsi, Tijk, Zijk, ORACLE_OK);
}

double oracle_energy_same(double epsabc, size_t No, const double *epsi,
                          const double *Tijk, const double *Zijk) {
  return energy_same_impl(epsabc, No, epsi, Tijk, Zijk, ORACLE_OK);
}

/* ------------------------------------------------------------------ */
/* Slices and the tuple loop                                           */
/* ------------------------------------------------------------------ */

size_t oracle_slices_size(size_t No, size_t Nv) {
  /* 6 V(xy)ph + 3 T(xy)hh + 3 V(xy)hh */
  return 6 * Nv * No + 6 * No * No;
}

typedef struct {
  const double *V[6]; /* AB, AC, BC, BA, CA, CB (as (E,i)) */
  const double *Vh[3];/* A, B, C (No^3) */
  const double *T[3]; /* A, B, C (E,i,j) */
  const double *T2[3];/* AB, AC, BC (i,j) */
  const double *V2[3];/* AB, AC, BC (i,j) */
} tuple_slices;

/* V(xy)ph[E + Nv i] = Vabci[x + Nv y + Nv^2 E + Nv^3 i]  (Unions.hpp:155-197)*/
static void cut_vph(const double *Vabci, size_t No, size_t Nv, size_t x,
                    size_t y, double *out) {
  const size_t NvNv = Nv * Nv, Nv3 = NvNv * Nv;
  for (size_t i = 0; i < No; i++)
    for (size_t E = 0; E < Nv; E++)
      out[E + Nv * i] = Vabci[x + Nv * y + NvNv * E + Nv3 * i];
}

/* X(xy)hh[i + No j] = Xpphh[x + Nv y + Nv^2 i + Nv^2 No j] (Unions.hpp:200-278)*/
static void cut_hh(const double *X, size_t No, size_t Nv, size_t x, size_t y,
                   double *out) {
  const size_t NvNv = Nv * Nv;
  for (size_t j = 0; j < No; j++)
    for (size_t i = 0; i < No; i++)
      out[i + No * j] = X[x + Nv * y + NvNv * i + NvNv * No * j];
}

typedef struct {
  const oracle_input *in;
  int ijkabc, use_gemm, mistake;
  const double *Tslices; /* Nv blocks of Nv*No^2: T(x)[E + Nv i + Nv No j] */
  const double *Tai;     /* possibly negated copy */
  uint64_t lo, hi;       /* tuple index range for this worker */
  double *per_tuple, *per_tuple_ct; /* indexed by (it - first) */
  uint64_t first;
  int status;
} worker_args;

/* x86-64-v3 (AVX2 + FMA) clone where the host has it; same results, since
 * contraction stays off and fma() is correctly rounded either way. */
#if defined(__x86_64__) && defined(__GNUC__) && !defined(__clang__)
#define ORACLE_CLONES __attribute__((target_clones("arch=haswell", "default")))
#else
#define ORACLE_CLONES
#endif

/* Per-tuple arithmetic shared by the whole-tensor and the synthetic-slice
 * loops: doubles (src/atrip/Atrip.cxx:855-885), Z = copy of W plus singles
 * (:895-922), the energy selector (:636-670) and, with J slices, the (cT)
 * pass (:927-963).  work holds 5 No^3 doubles. */
ORACLE_CLONES static void tuple_compute(size_t No, size_t Nv, size_t a,
                                        size_t b, size_t c,
                                        const tuple_slices *s,
                                        const double *const *VJ,
                                        const double *const *VhJ,
                                        const double *Tai, const double *ei,
                                        const double *ea, int use_gemm,
                                        int mistake, double *work, double *e_out,
                                        double *ect_out) {
  const size_t N3 = No * No * No;
  double *Tijk = work, *Zijk = work + N3, *tbuf = work + 2 * N3,
         *vbuf = work + 3 * N3, *TJ = work + 4 * N3;
  if (use_gemm && mistake == ORACLE_OK)
    oracle_doubles_gemm(No, Nv, s->V[0], s->V[1], s->V[2], s->V[3], s->V[4],
                        s->V[5], s->Vh[0], s->Vh[1], s->Vh[2], s->T[0],
                        s->T[1], s->T[2], s->T2[0], s->T2[1], s->T2[2], Tijk,
                        tbuf, vbuf);
  else
    doubles_loop_impl(No, Nv, s->V[0], s->V[1], s->V[2], s->V[3], s->V[4],
                      s->V[5], s->Vh[0], s->Vh[1], s->Vh[2], s->T[0], s->T[1],
                      s->T[2], s->T2[0], s->T2[1], s->T2[2], Tijk, mistake);
  memcpy(Zijk, Tijk, N3 * sizeof(double));
  singles_impl(No, Nv, a, b, c, Tai, s->V2[0], s->V2[1], s->V2[2], Zijk,
               mistake);
  int distinct = 0;
  if (a == b) distinct++;
  if (b == c) distinct--;
  if (mistake == ORACLE_MISTAKE_SAME_AS_DIST) distinct = 0;
  const double epsabc = (ea[a] + ea[b]) + ea[c];
  const double e =
      distinct == 0 ? energy_distinct_impl(epsabc, No, ei, Tijk, Zijk, mistake)
                    : energy_same_impl(epsabc, No, ei, Tijk, Zijk, mistake);
  double ect = e;
  if (VJ) {
    if (use_gemm && mistake == ORACLE_OK)
      oracle_doubles_gemm(No, Nv, VJ[0], VJ[1], VJ[2], VJ[3], VJ[4], VJ[5],
                          VhJ[0], VhJ[1], VhJ[2], s->T[0], s->T[1], s->T[2],
                          s->T2[0], s->T2[1], s->T2[2], TJ, tbuf, vbuf);
    else
      doubles_loop_impl(No, Nv, VJ[0], VJ[1], VJ[2], VJ[3], VJ[4], VJ[5],
                        VhJ[0], VhJ[1], VhJ[2], s->T[0], s->T[1], s->T[2],
                        s->T2[0], s->T2[1], s->T2[2], TJ, mistake);
    ect = distinct == 0
              ? energy_distinct_impl(epsabc, No, ei, TJ, Zijk, mistake)
              : energy_same_impl(epsabc, No, ei, TJ, Zijk, mistake);
  }
  *e_out = e;
  *ect_out = ect;
}

static void *worker(void *p) {
  worker_args *w = (worker_args *)p;
  const oracle_input *in = w->in;
  const size_t No = in->No, Nv = in->Nv, N3 = No * No * No, NvNo = Nv * No;
  const size_t NoNo = No * No;
  double *buf = (double *)malloc(
      sizeof(double) * (5 * N3 + 2 * oracle_slices_size(No, Nv)));
  if (!buf) {
    w->status = -1;
    return NULL;
  }
  double *sl = buf + 5 * N3, *slJ = sl + oracle_slices_size(No, Nv);
  const int have_J = in->Jijka && in->Jabci;
  for (uint64_t it = w->lo; it < w->hi; it++) {
    uint64_t abc[3];
    oracle_nth_tuple(it, Nv, abc);
    const size_t a = abc[0], b = abc[1], c = abc[2];
    const size_t xy[6][2] = {{a, b}, {a, c}, {b, c}, {b, a}, {c, a}, {c, b}};
    tuple_slices s;
    for (int q = 0; q < 6; q++) {
      double *dst = sl + q * NvNo;
      cut_vph(in->Vabci, No, Nv, xy[q][0], xy[q][1], dst);
      s.V[q] = dst;
    }
    for (int q = 0; q < 3; q++) {
      s.Vh[q] = in->Vijka + N3 * abc[q];
      s.T[q] = w->Tslices + abc[q] * (Nv * NoNo);
      double *t2 = sl + 6 * NvNo + q * NoNo;
      double *v2 = sl + 6 * NvNo + 3 * NoNo + q * NoNo;
      cut_hh(in->Tabij, No, Nv, xy[q][0], xy[q][1], t2);
      cut_hh(in->Vabij, No, Nv, xy[q][0], xy[q][1], v2);
      s.T2[q] = t2;
      s.V2[q] = v2;
    }
    const double *VJ[6], *VhJ[3];
    if (have_J) {
      for (int q = 0; q < 6; q++) {
        double *dst = slJ + q * NvNo;
        cut_vph(in->Jabci, No, Nv, xy[q][0], xy[q][1], dst);
        VJ[q] = dst;
      }
      for (int q = 0; q < 3; q++) VhJ[q] = in->Jijka + N3 * abc[q];
    }
    tuple_compute(No, Nv, a, b, c, &s, have_J ? VJ : NULL, VhJ, w->Tai, in->ei,
                  in->ea, w->use_gemm, w->mistake, buf,
                  &w->per_tuple[it - w->first], &w->per_tuple_ct[it - w->first]);
  }
  free(buf);
  w->status = 0;
  return NULL;
}

int oracle_run(const oracle_input *in, int ijkabc, uint64_t first,
               uint64_t count, int use_gemm, int nthreads, int mistake,
               double *energy, double *ct_energy, double *per_tuple,
               double *per_tuple_ct) {
  const size_t No = in->No, Nv = in->Nv, NoNo = No * No;
  const uint64_t n = oracle_n_tuples(Nv);
  if (first > n) first = n;
  if (count > n - first) count = n - first;
  if (nthreads < 1) nthreads = 1;
  /* T(x) slices for every x, src/atrip/Unions.hpp:78-113 */
  double *Tsl = (double *)malloc(sizeof(double) * Nv * Nv * NoNo);
  double *Tai = (double *)malloc(sizeof(double) * Nv * No);
  double *pt = per_tuple, *ptc = per_tuple_ct;
  if (!pt) pt = (double *)malloc(sizeof(double) * (count ? count : 1));
  if (!ptc) ptc = (double *)malloc(sizeof(double) * (count ? count : 1));
  if (!Tsl || !Tai || !pt || !ptc) return -1;
  for (size_t x = 0; x < Nv; x++)
    for (size_t r = 0; r < Nv * NoNo; r++)
      Tsl[x * Nv * NoNo + r] = in->Tabij[x + Nv * r];
  for (size_t q = 0; q < Nv * No; q++)
    Tai[q] = ijkabc ? -in->Tai[q] : in->Tai[q];

  worker_args *args = (worker_args *)calloc(nthreads, sizeof(worker_args));
  pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
  const uint64_t chunk = (count + nthreads - 1) / nthreads;
  int status = 0;
  for (int t = 0; t < nthreads; t++) {
    worker_args *w = &args[t];
    w->in = in;
    w->ijkabc = ijkabc;
    w->use_gemm = use_gemm;
    w->mistake = mistake;
    w->Tslices = Tsl;
    w->Tai = Tai;
    w->first = first;
    w->lo = first + chunk * t;
    w->hi = w->lo + chunk;
    if (w->lo > first + count) w->lo = first + count;
    if (w->hi > first + count) w->hi = first + count;
    w->per_tuple = pt;
    w->per_tuple_ct = ptc;
    if (nthreads == 1)
      worker(w);
    else
      pthread_create(&th[t], NULL, worker, w);
  }
  if (nthreads > 1)
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  for (int t = 0; t < nthreads; t++) status |= args[t].status;

  /* src/atrip/Atrip.cxx:925 accumulates in tuple order; :1108-1111 sign */
  double e = 0.0, ect = 0.0;
  for (uint64_t q = 0; q < count; q++) {
    e += pt[q];
    ect += ptc[q];
  }
  if (!ijkabc) {
    e = -e;
    ect = -ect;
  }
  *energy = e;
  *ct_energy = ect;
  if (pt != per_tuple) free(pt);
  if (ptc != per_tuple_ct) free(ptc);
  free(args);
  free(th);
  free(Tsl);
  free(Tai);
  return status;
}

/* ------------------------------------------------------------------ */
/* Complex (T)                                                         */
/* ------------------------------------------------------------------ */

typedef double _Complex zdouble;

/* src/atrip/Equations.cxx:620-680 with F = Complex: holes use conj(Vhhh)
 * (MAYBE_CONJ), particles plain products. */
static void doubles_z(size_t No, size_t Nv, const zdouble *V[6],
                      const zdouble *Vh[3], const zdouble *T[3],
                      const zdouble *T2[3], zdouble *Tijk) {
  const size_t NoNo = No * No, NoNv = No * Nv;
  const zdouble *VABph = V[0], *VACph = V[1], *VBCph = V[2], *VBAph = V[3],
                *VCAph = V[4], *VCBph = V[5];
  const zdouble *VhhhA = Vh[0], *VhhhB = Vh[1], *VhhhC = Vh[2];
  const zdouble *TAphh = T[0], *TBphh = T[1], *TCphh = T[2];
  const zdouble *TABhh = T2[0], *TAChh = T2[1], *TBChh = T2[2];
  for (size_t k = 0; k < No; k++)
    for (size_t j = 0; j < No; j++)
      for (size_t i = 0; i < No; i++) {
        zdouble w = 0.0;
        for (size_t L = 0; L < No; L++) {
          w -= TABhh[L + j * No] * conj(VhhhC[i + k * No + L * NoNo]); /* H1 */
          w -= TABhh[i + L * No] * conj(VhhhC[j + k * No + L * NoNo]); /* H0 */
          w -= TAChh[L + k * No] * conj(VhhhB[i + j * No + L * NoNo]); /* H5 */
          w -= TAChh[i + L * No] * conj(VhhhB[k + j * No + L * NoNo]); /* H3 */
          w -= TBChh[L + k * No] * conj(VhhhA[j + i * No + L * NoNo]); /* H2 */
          w -= TBChh[j + L * No] * conj(VhhhA[k + i * No + L * NoNo]); /* H4 */
        }
        for (size_t E = 0; E < Nv; E++) {
          w += TAphh[E + i * Nv + j * NoNv] * VBCph[E + k * Nv]; /* P0 */
          w += TAphh[E + i * Nv + k * NoNv] * VCBph[E + j * Nv]; /* P3 */
          w += TCphh[E + k * Nv + i * NoNv] * VABph[E + j * Nv]; /* P5 */
          w += TCphh[E + k * Nv + j * NoNv] * VBAph[E + i * Nv]; /* P2 */
          w += TBphh[E + j * Nv + i * NoNv] * VACph[E + k * Nv]; /* P1 */
          w += TBphh[E + j * Nv + k * NoNv] * VCAph[E + i * Nv]; /* P4 */
        }
        Tijk[i + j * No + k * NoNo] = w;
      }
}

/* src/atrip/Equations.cxx:101-238 with F = Complex, without the 16-blocking
 * (the sum order within a tuple is free to differ at the tolerance). */
static double energy_z(int distinct, zdouble epsabc, size_t No,
                       const zdouble *epsi, const zdouble *Tijk,
                       const zdouble *Zijk) {
  zdouble energy = 0.0;
  for (size_t k = 0; k < No; k++)
    for (size_t j = k; j < No; j++)
      for (size_t i = j; i < No; i++) {
        const zdouble facjk = j == k ? 0.5 : 1.0, facij = i == j ? 0.5 : 1.0;
        const zdouble denominator = epsabc - ((epsi[i] + epsi[j]) + epsi[k]);
        zdouble value;
        if (distinct == 0) {
          const zdouble U = Zijk[i + No * j + No * No * k],
                        V = Zijk[i + No * k + No * No * j],
                        W = Zijk[j + No * i + No * No * k],
                        X = Zijk[j + No * k + No * No * i],
                        Y = Zijk[k + No * i + No * No * j],
                        Z = Zijk[k + No * j + No * No * i];
          const zdouble A = conj(Tijk[i + No * j + No * No * k]),
                        B = conj(Tijk[i + No * k + No * No * j]),
                        C = conj(Tijk[j + No * i + No * No * k]),
                        D = conj(Tijk[j + No * k + No * No * i]),
                        E = conj(Tijk[k + No * i + No * No * j]),
                        F = conj(Tijk[k + No * j + No * No * i]);
          const zdouble UXY = U + (X + Y), VWZ = V + (W + Z);
          const zdouble first = A * U + (B * V + (C * W + (D * X + (E * Y + F * Z))));
          value = 3.0 * first + ((UXY - 2.0 * VWZ) * (A + (D + E)) +
                                 (VWZ - 2.0 * UXY) * (B + (C + F)));
        } else {
          const zdouble U = Zijk[i + No * j + No * No * k],
                        V = Zijk[j + No * k + No * No * i],
                        W = Zijk[k + No * i + No * No * j];
          const zdouble A = conj(Tijk[i + No * j + No * No * k]),
                        B = conj(Tijk[j + No * k + No * No * i]),
                        C = conj(Tijk[k + No * i + No * No * j]);
          value = 3.0 * ((A * U + B * V) + C * W) - (A + (B + C)) * (U + (V + W));
        }
        energy += ((2.0 * value) / denominator) * (facjk * facij);
      }
  return creal(energy);
}

typedef struct {
  const oracle_input_z *in;
  const zdouble *Tai;
  uint64_t lo, hi, first;
  double *pt, *ptc;
  int status;
} worker_z_args;

static void *worker_z(void *p) {
  worker_z_args *w = (worker_z_args *)p;
  const oracle_input_z *in = w->in;
  const size_t No = in->No, Nv = in->Nv, NoNo = No * No, N3 = NoNo * No;
  const size_t Nv2 = Nv * Nv, Nv3 = Nv2 * Nv;
  const zdouble *Tabij = (const zdouble *)in->Tabij,
                *Vabij = (const zdouble *)in->Vabij,
                *Vijka = (const zdouble *)in->Vijka,
                *Vabci = (const zdouble *)in->Vabci,
                *Jijka = (const zdouble *)in->Jijka,
                *Jabci = (const zdouble *)in->Jabci,
                *ei = (const zdouble *)in->ei, *ea = (const zdouble *)in->ea;
  const int have_J = Jijka && Jabci;
  const size_t nbuf = 4 * N3 + 12 * Nv * No + 3 * Nv * NoNo + 6 * NoNo;
  zdouble *buf = (zdouble *)malloc(sizeof(zdouble) * nbuf);
  if (!buf) {
    w->status = -1;
    return NULL;
  }
  zdouble *Tijk = buf, *Zijk = buf + N3, *TJ = buf + 2 * N3;
  zdouble *vs = buf + 4 * N3, *vsJ = vs + 6 * Nv * No;
  zdouble *ts = vsJ + 6 * Nv * No, *hh = ts + 3 * Nv * NoNo;
  for (uint64_t it = w->lo; it < w->hi; it++) {
    uint64_t abc[3];
    oracle_nth_tuple(it, Nv, abc);
    const size_t a = abc[0], b = abc[1], c = abc[2];
    const size_t xy[6][2] = {{a, b}, {a, c}, {b, c}, {b, a}, {c, a}, {c, b}};
    const zdouble *V[6], *VJ[6], *Vh[3], *VhJ[3], *T[3], *T2[3], *V2[3];
    for (int q = 0; q < 6; q++) {
      zdouble *d = vs + q * Nv * No, *dJ = vsJ + q * Nv * No;
      for (size_t i = 0; i < No; i++)
        for (size_t E = 0; E < Nv; E++) {
          const size_t src = xy[q][0] + Nv * xy[q][1] + Nv2 * E + Nv3 * i;
          d[E + Nv * i] = Vabci[src];
          if (have_J) dJ[E + Nv * i] = Jabci[src];
        }
      V[q] = d;
      VJ[q] = dJ;
    }
    for (int q = 0; q < 3; q++) {
      Vh[q] = Vijka + N3 * abc[q];
      VhJ[q] = have_J ? Jijka + N3 * abc[q] : NULL;
      zdouble *t = ts + q * Nv * NoNo;
      for (size_t r = 0; r < Nv * NoNo; r++) t[r] = Tabij[abc[q] + Nv * r];
      T[q] = t;
      zdouble *t2 = hh + q * NoNo, *v2 = hh + 3 * NoNo + q * NoNo;
      for (size_t j = 0; j < No; j++)
        for (size_t i = 0; i < No; i++) {
          const size_t src = xy[q][0] + Nv * xy[q][1] + Nv2 * i + Nv2 * No * j;
          t2[i + No * j] = Tabij[src];
          v2[i + No * j] = Vabij[src];
        }
      T2[q] = t2;
      V2[q] = v2;
    }
    doubles_z(No, Nv, V, Vh, T, T2, Tijk);
    memcpy(Zijk, Tijk, N3 * sizeof(zdouble));
    for (size_t k = 0; k < No; k++)
      for (size_t i = 0; i < No; i++)
        for (size_t j = 0; j < No; j++) {
          const size_t ijk = i + j * No + k * NoNo;
          Zijk[ijk] += w->Tai[a + i * Nv] * V2[2][j + k * No];
          Zijk[ijk] += w->Tai[b + j * Nv] * V2[1][i + k * No];
          Zijk[ijk] += w->Tai[c + k * Nv] * V2[0][i + j * No];
        }
    int distinct = 0;
    if (a == b) distinct++;
    if (b == c) distinct--;
    /* real part only, src/atrip/Atrip.cxx:643-644 (std::real of the sum) */
    const zdouble epsabc = creal((ea[a] + ea[b]) + ea[c]);
    const double e = energy_z(distinct, epsabc, No, ei, Tijk, Zijk);
    double ect = e;
    if (have_J) {
      doubles_z(No, Nv, VJ, VhJ, T, T2, TJ);
      ect = energy_z(distinct, epsabc, No, ei, TJ, Zijk);
    }
    w->pt[it - w->first] = e;
    w->ptc[it - w->first] = ect;
  }
  free(buf);
  w->status = 0;
  return NULL;
}

int oracle_run_z(const oracle_input_z *in, int ijkabc, uint64_t first,
                 uint64_t count, int nthreads, double *energy,
                 double *ct_energy, double *per_tuple, double *per_tuple_ct) {
  const size_t No = in->No, Nv = in->Nv;
  const uint64_t n = oracle_n_tuples(Nv);
  if (first > n) first = n;
  if (count > n - first) count = n - first;
  if (nthreads < 1) nthreads = 1;
  zdouble *Tai = (zdouble *)malloc(sizeof(zdouble) * Nv * No);
  double *pt = per_tuple, *ptc = per_tuple_ct;
  if (!pt) pt = (double *)malloc(sizeof(double) * (count ? count : 1));
  if (!ptc) ptc = (double *)malloc(sizeof(double) * (count ? count : 1));
  if (!Tai || !pt || !ptc) return -1;
  for (size_t q = 0; q < Nv * No; q++)
    Tai[q] = ijkabc ? -((const zdouble *)in->Tai)[q] : ((const zdouble *)in->Tai)[q];
  worker_z_args *args = (worker_z_args *)calloc(nthreads, sizeof(worker_z_args));
  pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
  const uint64_t chunk = (count + nthreads - 1) / nthreads;
  int status = 0;
  for (int t = 0; t < nthreads; t++) {
    worker_z_args *w = &args[t];
    w->in = in;
    w->Tai = Tai;
    w->first = first;
    w->lo = first + chunk * t;
    w->hi = w->lo + chunk;
    if (w->lo > first + count) w->lo = first + count;
    if (w->hi > first + count) w->hi = first + count;
    w->pt = pt;
    w->ptc = ptc;
    if (nthreads == 1)
      worker_z(w);
    else
      pthread_create(&th[t], NULL, worker_z, w);
  }
  if (nthreads > 1)
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  for (int t = 0; t < nthreads; t++) status |= args[t].status;
  double e = 0.0, ect = 0.0;
  for (uint64_t q = 0; q < count; q++) {
    e += pt[q];
    ect += ptc[q];
  }
  if (!ijkabc) {
    e = -e;
    ect = -ect;
  }
  *energy = e;
  *ct_energy = ect;
  if (pt != per_tuple) free(pt);
  if (ptc != per_tuple_ct) free(ptc);
  free(args);
  free(th);
  free(Tai);
  return status;
}

typedef struct {
  uint64_t stream;
  int tensor;
  size_t No, Nv, lo, hi;
  double *out;
} fill_args;

ORACLE_CLONES static void *fill_worker(void *p) {
  fill_args *f = (fill_args *)p;
  for (size_t q = f->lo; q < f->hi; q++)
    f->out[q] = atrip_synth_value(f->stream, f->tensor, f->No, f->Nv, q);
  return NULL;
}

void oracle_fill_synthetic(uint64_t seed, int tensor, size_t No, size_t Nv,
                           double *out, size_t n, int nthreads) {
  const uint64_t s = atrip_synth_stream(seed, tensor);
  if (nthreads < 1) nthreads = 1;
  if ((size_t)nthreads > n / 4096 + 1) nthreads = (int)(n / 4096 + 1);
  fill_args *args = (fill_args *)calloc(nthreads, sizeof(fill_args));
  pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
  const size_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; t++) {
    fill_args *f = &args[t];
    f->stream = s;
    f->tensor = tensor;
    f->No = No;
    f->Nv = Nv;
    f->out = out;
    f->lo = chunk * t < n ? chunk * t : n;
    f->hi = chunk * (t + 1) < n ? chunk * (t + 1) : n;
    if (nthreads == 1)
      fill_worker(f);
    else
      pthread_create(&th[t], NULL, fill_worker, f);
  }
  if (nthreads > 1)
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(args);
  free(th);
}

/* ------------------------------------------------------------------ */
/* The (T) loop on slices drawn from the generator                      */
/* ------------------------------------------------------------------ */

/* Slices of one index x (T(x) and Vhhh(x), plus Jhhh(x) with J), kept for
 * the next tuples: consecutive tuples share a and mostly b.  A tuple needs
 * at most 9 distinct ones, and the least recently used is replaced, so the
 * slices of the current tuple are never evicted while it is set up. */
#define SYN_CACHE 9
typedef struct {
  long x[SYN_CACHE];
  int kind[SYN_CACHE];
  double *data[SYN_CACHE];
  uint64_t used[SYN_CACHE], clock;
} syn_cache;

typedef struct {
  uint64_t seed;
  size_t No, Nv;
  int with_j, use_gemm;
  double scale;
  const double *Tai, *ei, *ea;
  uint64_t lo, hi, first;
  double *per_tuple, *per_tuple_ct;
  int status;
} syn_args;

/* kind 0: T(x)[E + Nv i + Nv No j] = Tabij[x, E, i, j] (Unions.hpp:78-113);
 * kind 1: Vhhh(x)[q] = Vijka[q + No^3 x]; kind 2: the same of Jijka. */
static const double *syn_x_slice(const syn_args *w, syn_cache *c, int kind,
                                 size_t x) {
  const size_t No = w->No, Nv = w->Nv, N3 = No * No * No;
  int slot = 0;
  for (int q = 0; q < SYN_CACHE; q++) {
    if (c->data[q] && c->kind[q] == kind && c->x[q] == (long)x) {
      c->used[q] = ++c->clock;
      return c->data[q];
    }
    if (c->used[q] < c->used[slot]) slot = q;
  }
  const size_t n = kind == 0 ? Nv * No * No : N3;
  free(c->data[slot]);
  double *d = (double *)malloc(sizeof(double) * n);
  c->data[slot] = d;
  c->kind[slot] = kind;
  c->x[slot] = (long)x;
  c->used[slot] = ++c->clock;
  if (!d) return NULL;
  if (kind == 0) {
    const uint64_t st = atrip_synth_stream(w->seed, ATRIP_SYN_TABIJ);
    for (size_t r = 0; r < n; r++)  /* r = E + Nv (i + No j) */
      d[r] = atrip_synth_value(st, ATRIP_SYN_TABIJ, No, Nv, x + Nv * r) *
             w->scale;
  } else {
    const int t = kind == 1 ? ATRIP_SYN_VIJKA : ATRIP_SYN_JIJKA;
    const uint64_t st = atrip_synth_stream(w->seed, t);
    for (size_t q = 0; q < N3; q++)
      d[q] = atrip_synth_value(st, t, No, Nv, q + N3 * x) * w->scale;
  }
  return d;
}

static void *syn_worker(void *p) {
  syn_args *w = (syn_args *)p;
  const size_t No = w->No, Nv = w->Nv, N3 = No * No * No, NvNo = Nv * No;
  const size_t NoNo = No * No, NvNv = Nv * Nv;
  syn_cache cache;
  memset(&cache, 0, sizeof cache);
  double *buf = (double *)malloc(
      sizeof(double) * (5 * N3 + 2 * oracle_slices_size(No, Nv)));
  w->status = -1;
  if (!buf) return NULL;
  double *sl = buf + 5 * N3, *slJ = sl + oracle_slices_size(No, Nv);
  const uint64_t sV = atrip_synth_stream(w->seed, ATRIP_SYN_VABCI);
  const uint64_t sJ = atrip_synth_stream(w->seed, ATRIP_SYN_JABCI);
  const uint64_t sT = atrip_synth_stream(w->seed, ATRIP_SYN_TABIJ);
  const uint64_t sW = atrip_synth_stream(w->seed, ATRIP_SYN_VABIJ);
  int ok = 1;
  for (uint64_t it = w->lo; it < w->hi && ok; it++) {
    uint64_t abc[3];
    oracle_nth_tuple(it, Nv, abc);
    const size_t a = abc[0], b = abc[1], c = abc[2];
    const size_t xy[6][2] = {{a, b}, {a, c}, {b, c}, {b, a}, {c, a}, {c, b}};
    tuple_slices s;
    const double *VJ[6], *VhJ[3];
    /* V(xy)ph[E + Nv i] = Vabci[x + Nv y + Nv^2 E + Nv^3 i] (cut_vph) */
    for (int q = 0; q < 6; q++) {
      double *dst = sl + q * NvNo, *dJ = slJ + q * NvNo;
      for (size_t i = 0; i < No; i++)
        for (size_t E = 0; E < Nv; E++) {
          const uint64_t idx = xy[q][0] + Nv * xy[q][1] + NvNv * E + NvNv * Nv * i;
          dst[E + Nv * i] =
              atrip_synth_value(sV, ATRIP_SYN_VABCI, No, Nv, idx) * w->scale;
          if (w->with_j)
            dJ[E + Nv * i] =
                atrip_synth_value(sJ, ATRIP_SYN_JABCI, No, Nv, idx) * w->scale;
        }
      s.V[q] = dst;
      VJ[q] = dJ;
    }
    /* X(xy)hh[i + No j] = Xpphh[x + Nv y + Nv^2 i + Nv^2 No j] (cut_hh) */
    for (int q = 0; q < 3; q++) {
      double *t2 = sl + 6 * NvNo + q * NoNo;
      double *v2 = sl + 6 * NvNo + 3 * NoNo + q * NoNo;
      for (size_t j = 0; j < No; j++)
        for (size_t i = 0; i < No; i++) {
          const uint64_t idx = xy[q][0] + Nv * xy[q][1] + NvNv * i + NvNv * No * j;
          t2[i + No * j] =
              atrip_synth_value(sT, ATRIP_SYN_TABIJ, No, Nv, idx) * w->scale;
          v2[i + No * j] =
              atrip_synth_value(sW, ATRIP_SYN_VABIJ, No, Nv, idx) * w->scale;
        }
      s.T2[q] = t2;
      s.V2[q] = v2;
    }
    for (int q = 0; q < 3; q++) {
      s.T[q] = syn_x_slice(w, &cache, 0, abc[q]);
      s.Vh[q] = syn_x_slice(w, &cache, 1, abc[q]);
      VhJ[q] = w->with_j ? syn_x_slice(w, &cache, 2, abc[q]) : NULL;
      if (!s.T[q] || !s.Vh[q] || (w->with_j && !VhJ[q])) ok = 0;
    }
    if (ok)
      tuple_compute(No, Nv, a, b, c, &s, w->with_j ? VJ : NULL, VhJ, w->Tai,
                    w->ei, w->ea, w->use_gemm, ORACLE_OK, buf,
                    &w->per_tuple[it - w->first],
                    &w->per_tuple_ct[it - w->first]);
  }
  for (int q = 0; q < SYN_CACHE; q++) free(cache.data[q]);
  free(buf);
  w->status = ok ? 0 : -1;
  return NULL;
}

int oracle_run_synthetic(uint64_t seed, size_t No, size_t Nv, int with_j,
                         double scale, int ijkabc, uint64_t first,
                         uint64_t count, int use_gemm, int nthreads,
                         double *per_tuple, double *per_tuple_ct) {
  const uint64_t n = oracle_n_tuples(Nv);
  if (first > n) first = n;
  if (count > n - first) count = n - first;
  if (nthreads < 1) nthreads = 1;
  double *Tai = (double *)malloc(sizeof(double) * Nv * No);
  double *ei = (double *)malloc(sizeof(double) * No);
  double *ea = (double *)malloc(sizeof(double) * Nv);
  syn_args *args = (syn_args *)calloc(nthreads, sizeof(syn_args));
  pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
  if (!Tai || !ei || !ea || !args || !th) return -1;
  const uint64_t sTai = atrip_synth_stream(seed, ATRIP_SYN_TAI);
  for (size_t q = 0; q < Nv * No; q++) {
    const double v = atrip_synth_value(sTai, ATRIP_SYN_TAI, No, Nv, q) * scale;
    Tai[q] = ijkabc ? -v : v;  /* src/atrip/Atrip.cxx:183-187 */
  }
  oracle_fill_synthetic(seed, ATRIP_SYN_EI, No, Nv, ei, No, 1);
  oracle_fill_synthetic(seed, ATRIP_SYN_EA, No, Nv, ea, Nv, 1);
  const uint64_t chunk = (count + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; t++) {
    syn_args *w = &args[t];
    w->seed = seed;
    w->No = No;
    w->Nv = Nv;
    w->with_j = with_j;
    w->use_gemm = use_gemm;
    w->scale = scale;
    w->Tai = Tai;
    w->ei = ei;
    w->ea = ea;
    w->first = first;
    w->lo = first + chunk * t;
    w->hi = w->lo + chunk;
    if (w->lo > first + count) w->lo = first + count;
    if (w->hi > first + count) w->hi = first + count;
    w->per_tuple = per_tuple;
    w->per_tuple_ct = per_tuple_ct;
    if (nthreads == 1)
      syn_worker(w);
    else
      pthread_create(&th[t], NULL, syn_worker, w);
  }
  int status = 0;
  if (nthreads > 1)
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  for (int t = 0; t < nthreads; t++) status |= args[t].status;
  free(args);
  free(th);
  free(Tai);
  free(ei);
  free(ea);
  return status;
}
