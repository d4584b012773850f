/*
 * CPU oracle for atrip's perturbative-triples (T) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the
 * reference's single-rank CPU algorithm so that the GPU implementation can be
 * checked against it.  Only the tests, the smoke check and the CPU-baseline
 * leg of the benchmark may load it, and only as the checker or as the
 * reported CPU baseline; the product path never calls it.
 *
 * Every function names the reference lines it follows (paths relative to the
 * atrip source tree).  How each function is pinned (closed forms, invariants,
 * brute force) is written next to its declaration.
 *
 * Layout conventions are the reference's CTF ones: column-major, first index
 * fastest, lens as in bench/main.cxx:198-200:
 *   Tpphh[a,b,i,j], Vpphh[a,b,i,j]: {Nv,Nv,No,No}
 *   Vhhhp[i,j,k,a]: {No,No,No,Nv}
 *   Vppph[a,b,c,i]: {Nv,Nv,Nv,No}
 *   Tph[a,i]: {Nv,No};  ei: {No};  ea: {Nv}
 */
#ifndef ATRIP_ORACLE_H
#define ATRIP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of tuples a<=b<=c without a=b=c: src/atrip/Tuples.cxx:93.
 * Pinned: equals Nv(Nv+1)(Nv+2)/6 - Nv and the brute-force count. */
uint64_t oracle_n_tuples(uint64_t Nv);

/* Lexicographic enumeration, src/atrip/Tuples.cxx:121-134.
 * Pinned: brute force over all (a,b,c) in the tests. */
void oracle_all_tuples(uint64_t Nv, uint64_t *abc /* 3*n */);

/* Closed-form decode of the it-th tuple, src/atrip/DatabaseCommunicator.cxx:
 * 51-106 (a_block_sum_atrip, b_block_sum_atrip, nth_atrip), without the
 * static cache.  Pinned: exhaustive comparison with the enumeration. */
void oracle_nth_tuple(uint64_t it, uint64_t Nv, uint64_t out[3]);

/* Doubles contribution W(ijk), direct-loop form, src/atrip/Equations.cxx:
 * 686-726.  Slices as produced by src/atrip/Unions.hpp (see
 * oracle_gather_slices).  Pinned: simultaneous-permutation symmetry of W,
 * separable rank-one closed forms per term, agreement with the GEMM form. */
void oracle_doubles_loop(size_t No, size_t Nv, const double *VABph,
                         const double *VACph, const double *VBCph,
                         const double *VBAph, const double *VCAph,
                         const double *VCBph, const double *VhhhA,
                         const double *VhhhB, const double *VhhhC,
                         const double *TAphh, const double *TBphh,
                         const double *TCphh, const double *TABhh,
                         const double *TAChh, const double *TBChh,
                         double *Tijk);

/* Doubles contribution, DGEMM form (12 dgemm + permuted accumulate),
 * src/atrip/Equations.cxx:26-89 (reorder), :494-578 (call shapes),
 * :585-680 (term order).  tbuf and vbuf are No^3 scratch each. */
void oracle_doubles_gemm(size_t No, size_t Nv, const double *VABph,
                         const double *VACph, const double *VBCph,
                         const double *VBAph, const double *VCAph,
                         const double *VCBph, const double *VhhhA,
                         const double *VhhhB, const double *VhhhC,
                         const double *TAphh, const double *TBphh,
                         const double *TCphh, const double *TABhh,
                         const double *TAChh, const double *TBChh,
                         double *Tijk, double *tbuf, double *vbuf);

/* Singles contribution Z += t V, src/atrip/Equations.cxx:387-426.
 * Pinned: closed form (Tai = 0 leaves Z = W; linear in Tai). */
void oracle_singles(size_t No, size_t Nv, size_t a, size_t b, size_t c,
                    const double *Tph, const double *VABij,
                    const double *VACij, const double *VBCij, double *Zijk);

/* Energy of a tuple with a<b<c, src/atrip/Equations.cxx:101-180.
 * Pinned: equals the unrestricted sum over all (i,j,k) of
 * 2 W_ijk (4Z_ijk + Z_jki + Z_kij - 2Z_ikj - 2Z_jik - 2Z_kji) / D_ijk. */
double oracle_energy_distinct(double epsabc, size_t No, const double *epsi,
                              const double *Tijk, const double *Zijk);

/* Energy of a tuple with exactly two equal indices,
 * src/atrip/Equations.cxx:182-238.  Pinned through the full-space identity
 * in the tests (the restricted sum equals one sixth of the unrestricted
 * sum over all (a,b,c) for inputs with Tabij = Tbaji, Vabij = Vbaji). */
double oracle_energy_same(double epsabc, size_t No, const double *epsi,
                          const double *Tijk, const double *Zijk);

/* Whole-tensor inputs in CTF layout (non-owning). Jijka/Jabci may be NULL. */
typedef struct {
  size_t No, Nv;
  const double *ei, *ea, *Tai, *Tabij, *Vabij, *Vijka, *Vabci;
  const double *Jijka, *Jabci;
} oracle_input;

/* The 15 doubles slices plus 3 singles slices of tuple (a,b,c), as
 * src/atrip/Unions.hpp:78-278 cuts them and src/atrip/Atrip.cxx:857-883,
 * :919-921 hands them out.  `buf` must hold oracle_slices_size(No,Nv)
 * doubles.  When use_J is set, Vabci/Vijka are replaced by Jabci/Jijka
 * (src/atrip/Atrip.cxx:928-963). */
size_t oracle_slices_size(size_t No, size_t Nv);

/* Planted mistakes for sensitivity tests of the pins (0 = none). */
enum oracle_mistake {
  ORACLE_OK = 0,
  ORACLE_MISTAKE_P3_PERM = 1,     /* P3 accumulated with IJK instead of IKJ */
  ORACLE_MISTAKE_H0_SIGN = 2,     /* H0 added instead of subtracted */
  ORACLE_MISTAKE_DROP_P4 = 3,     /* P4 term missing */
  ORACLE_MISTAKE_SAME_AS_DIST = 4,/* distinct formula used for a=b or b=c */
  ORACLE_MISTAKE_FACTOR = 5,      /* f_ij = 1 on i == j */
  ORACLE_MISTAKE_H1_TRANSPOSE = 6,/* TAB read transposed in H1 */
  ORACLE_MISTAKE_SINGLES_IDX = 7  /* t_bj V^ac_ik read as t_bj V^ac_ki */
};

/*
 * Full single-rank (T) loop, src/atrip/Atrip.cxx:623-684 (energy selector),
 * :686-963 (tuple loop body incl. the (cT) pass), :1094-1111 (sum and sign),
 * :179-187 (ijkabc negates Tai).  Processes the tuples with global index in
 * [first, first+count) of the lexicographic order.  Writes the physical
 * energies (sign applied) to *energy and *ct_energy.  Without J tensors the
 * (cT) accumulator receives the same per-tuple energy, as in the reference.
 * per_tuple / per_tuple_ct (may be NULL) receive the raw per-tuple values
 * before the sign.  use_gemm selects oracle_doubles_gemm over the loop form.
 * nthreads > 1 shards the tuples over threads (sum order is the tuple
 * order either way).  Returns 0 on success.
 */
int oracle_run(const oracle_input *in, int ijkabc, uint64_t first,
               uint64_t count, int use_gemm, int nthreads, int mistake,
               double *energy, double *ct_energy, double *per_tuple,
               double *per_tuple_ct);

/*
 * Complex (T) (atrip::Atrip::run<Complex>, src/atrip/Atrip.cxx:1136), loop
 * form.  Tensors are interleaved complex (re, im) in CTF layout, i.e. the
 * memory of std::complex<double> arrays.  Follows the reference's GEMM
 * path: the hole terms use the conjugated Vhhh slices (MAYBE_CONJ before
 * each hole DGEMM pair, src/atrip/Equations.cxx:620-656), the particle terms
 * and the singles use plain products (zgemm "T"/"N", :484-489, and
 * :387-426), and the energy conjugates W (maybe_conjugate_scalar on Tijk,
 * :135-146, :212-217), divides by the complex denominator and keeps the
 * real part of the tuple's sum (:176-178).  Same selector, tuple order and
 * sign as oracle_run; per_tuple values are the real parts.
 */
typedef struct {
  size_t No, Nv;
  const double *ei, *ea, *Tai, *Tabij, *Vabij, *Vijka, *Vabci;
  const double *Jijka, *Jabci;
} oracle_input_z;

int oracle_run_z(const oracle_input_z *in, int ijkabc, uint64_t first,
                 uint64_t count, int nthreads, double *energy,
                 double *ct_energy, double *per_tuple, double *per_tuple_ct);

/*
 * The same loop as oracle_run (DGEMM or loop form, same per-tuple
 * arithmetic), on the seeded synthetic system of include/atrip_synth.h with
 * every amplitude and integral times `scale`, but building each tuple's
 * slices straight from the generator instead of cutting them out of whole
 * tensors: for shapes whose tensors do not fit host memory (No=60, Nv=600
 * holds 125 GB).  Writes the raw per-tuple energies (before the sign) of
 * tuples [first, first+count) to per_tuple / per_tuple_ct.  Pinned by
 * equality with oracle_run on the whole synthetic tensors (tests).
 */
int oracle_run_synthetic(uint64_t seed, size_t No, size_t Nv, int with_j,
                         double scale, int ijkabc, uint64_t first,
                         uint64_t count, int use_gemm, int nthreads,
                         double *per_tuple, double *per_tuple_ct);

/* Fill a whole tensor in CTF layout from the counter-based generator
 * (include/atrip_synth.h).  n = number of elements of that tensor; the
 * elements are split over nthreads threads. */
void oracle_fill_synthetic(uint64_t seed, int tensor, size_t No, size_t Nv,
                           double *out, size_t n, int nthreads);

/* Host BLAS used by the GEMM form: loads dgemm_ from `path` (or searches
 * a few standard names when NULL).  Returns the library name, or NULL when
 * none is found (then a plain C GEMM is used).  Single-threaded BLAS is
 * requested where the library supports it. */
const char *oracle_blas_init(const char *path);
const char *oracle_blas_name(void);

/* Column-major GEMM used by the GEMM form (dgemm_ semantics). */
void oracle_dgemm(char transa, char transb, int m, int n, int k, double alpha,
                  const double *A, int lda, const double *B, int ldb,
                  double beta, double *C, int ldc);

#ifdef __cplusplus
}
#endif

#endif /* ATRIP_ORACLE_H */
