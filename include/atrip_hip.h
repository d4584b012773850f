/*
 * atrip_hip.h -- C-ABI of the MI355X-native (T) hot path.
 *
 * This is the drop-in boundary: plain C types only (no C++, no torch, no HIP
 * types).  The C++ layer (include/atrip/Atrip.hpp) implements the
 * reference's atrip::Atrip::run(Input) -> Output on top of it, and any FFI
 * (ctypes, cgo, JNI) can bind it directly.
 *
 * Every call returns 0 on success and a non-zero status otherwise; the text
 * of the failure is then available from atrip_hip_last_error(ctx).  The C++
 * layer turns a failure into `throw std::string(...)`, the reference's own
 * error convention (src/atrip/Acc.hpp:95-107, src/atrip/HIP.hpp:31-42).
 *
 * A context is bound to one device and is not thread-safe: use one host
 * thread per context (the serving loop owns its own context).
 */
#ifndef ATRIP_HIP_H
#define ATRIP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct atrip_hip_ctx atrip_hip_ctx;

/*
 * One (T) system in the reference's CTF layout: column-major, first index
 * fastest, lens as in bench/main.cxx:198-200 of the reference:
 *   ei {No}, ea {Nv}, Tai {Nv,No}, Tabij {Nv,Nv,No,No}, Vabij {Nv,Nv,No,No},
 *   Vijka {No,No,No,Nv}, Vabci {Nv,Nv,Nv,No}, Jijka {No,No,No,Nv},
 *   Jabci {Nv,Nv,Nv,No}.
 * Host pointers; Jijka/Jabci may be NULL (no (cT) pass).  Replaces the
 * tensor fields of atrip::Atrip::Input<double>
 * (src/atrip/Atrip.hpp:68-107) and the slicing of src/atrip/Unions.hpp.
 */
typedef struct {
  int64_t No, Nv;
  const double *ei, *ea, *Tai, *Tabij, *Vabij, *Vijka, *Vabci;
  const double *Jijka, *Jabci;
} atrip_sys_desc;

/* Arithmetic of the doubles contractions. */
enum atrip_precision {
  ATRIP_FP64 = 0,        /* FP64 MFMA, the plain path */
  ATRIP_SPLIT_FP32 = 1,  /* split FP32: hi*hi on FP32 MFMA + hi*lo + lo*hi
                            corrections, FP32 accumulation, FP64 energy */
  ATRIP_FP32 = 2         /* hi*hi only on FP32 MFMA, FP64 energy */
};

/* Create a context on HIP device `device`.  Replaces the device selection
 * and BLAS-handle creation of src/atrip/Atrip.cxx:78-152. */
int atrip_hip_create(int device, atrip_hip_ctx **out);

/* Free every device buffer of the context and the context itself. */
int atrip_hip_destroy(atrip_hip_ctx *ctx);

/* Text of the last failure on this context ("" when none). */
const char *atrip_hip_last_error(atrip_hip_ctx *ctx);

/* Copy a system to the device once, re-laid out for the kernels, and keep
 * it resident.  Replaces the per-run CTF slicing into SliceUnion sources
 * (src/atrip/Unions.hpp:22-75, src/atrip/SliceUnion.cxx:305-332) and the
 * per-tuple slice transfers (src/atrip/Slice.cxx:157-199).  *sys receives a
 * handle used by the run calls. */
int atrip_hip_upload(atrip_hip_ctx *ctx, const atrip_sys_desc *desc,
                     int *sys);

/* Generate the seeded synthetic system (include/atrip_synth.h) directly in
 * device memory; bit-identical to filling host tensors with the same
 * generator and uploading them.  No counterpart in the reference (its bench
 * fills CTF tensors with fill_random, bench/main.cxx:53-89). */
int atrip_hip_upload_synthetic(atrip_hip_ctx *ctx, int64_t No, int64_t Nv,
                               uint64_t seed, int with_j, int *sys);

/* Release a resident system's device memory. */
int atrip_hip_release(atrip_hip_ctx *ctx, int sys);

/*
 * (T) energy over the tuples with global index in [first, first+count) of
 * the reference's lexicographic a<=b<=c order (src/atrip/Tuples.cxx:89-134);
 * count = UINT64_MAX means "to the end".  Writes the physical energies (sign
 * of src/atrip/Atrip.cxx:1108-1111 applied unless ijkabc, with Tai negated
 * when ijkabc as in :183-187) to *energy and *ct_energy.  Without J tensors
 * *ct_energy equals *energy, as in the reference's loop (:925, :961).
 * Replaces the body of atrip::Atrip::run (src/atrip/Atrip.cxx:686-1111):
 * doubles_contribution, singles_contribution, get_energy_distinct/same.
 */
int atrip_hip_run(atrip_hip_ctx *ctx, int sys, int precision, int ijkabc,
                  uint64_t first, uint64_t count, double *energy,
                  double *ct_energy);

/* Per-tuple energies (raw, before the sign) for the tuples
 * [first, first+count) in lexicographic order; e_ct may be NULL.  The
 * reference keeps these only in a never-filled debug map
 * (src/atrip/Atrip.cxx:575). */
int atrip_hip_tuple_energies(atrip_hip_ctx *ctx, int sys, int precision,
                             int ijkabc, uint64_t first, uint64_t count,
                             double *e, double *e_ct);

/* Tuning and instrumentation knobs (no counterpart in the reference):
 *   "window"  block edge nb of the tuple windows (1..32)
 *   "batch"   windows per launch
 *   "timing"  1: time every hot-kernel launch with HIP events
 *   "tile"    doubles-kernel tile variant 0..7, -1 (default) = automatic */
int atrip_hip_set_option(atrip_hip_ctx *ctx, const char *key, int64_t value);
int atrip_hip_get_option(atrip_hip_ctx *ctx, const char *key, int64_t *value);

/* Accumulated HIP-event time of kernel `name` ("doubles", "doubles_mixed",
 * "energy") since the last reset: launches and total milliseconds, plus the
 * algorithmic doubles FLOP those launches covered (12 No^3 (No+Nv) per
 * tuple, src/atrip/Atrip.cxx:578-580). */
int atrip_hip_kernel_stats(atrip_hip_ctx *ctx, const char *name,
                           uint64_t *launches, double *total_ms,
                           double *flop);
int atrip_hip_reset_stats(atrip_hip_ctx *ctx);

/* Sizes of a resident system: No, Nv, number of tuples, device bytes. */
int atrip_hip_sys_info(atrip_hip_ctx *ctx, int sys, int64_t *No, int64_t *Nv,
                       uint64_t *n_tuples, uint64_t *device_bytes);

/* Number of HIP devices visible (0 when none); never fails. */
int atrip_hip_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* ATRIP_HIP_H */
