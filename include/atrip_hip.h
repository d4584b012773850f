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
  ATRIP_FP32 = 2,        /* hi*hi only on FP32 MFMA, FP64 energy */
  ATRIP_BF16X3 = 3       /* every operand as three bf16 parts (24 bits):
                            six part products on bf16 MFMA, FP32
                            accumulation, FP64 energy */
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

/* The synthetic system with every amplitude and integral (Tai, Tabij,
 * Vabij, Vijka, Vabci and the J tensors) multiplied by `scale`; orbital
 * energies unchanged.  E(T) then scales by scale^4 (each tuple energy is a
 * product of a doubles or singles term, both bilinear in the amplitudes and
 * integrals, with W).  Used to state absolute energy errors at molecular
 * magnitudes on the synthetic shapes.  No counterpart in the reference. */
int atrip_hip_upload_synthetic_scaled(atrip_hip_ctx *ctx, int64_t No,
                                      int64_t Nv, uint64_t seed, int with_j,
                                      double scale, int *sys);

/* Tensor ids of the streaming upload (the order of atrip_sys_desc). */
enum atrip_tensor {
  ATRIP_TENSOR_EI = 0,
  ATRIP_TENSOR_EA = 1,
  ATRIP_TENSOR_TAI = 2,
  ATRIP_TENSOR_TABIJ = 3,
  ATRIP_TENSOR_VABIJ = 4,
  ATRIP_TENSOR_VIJKA = 5,
  ATRIP_TENSOR_VABCI = 6,
  ATRIP_TENSOR_JIJKA = 7,
  ATRIP_TENSOR_JABCI = 8
};

/* Reader of the streaming upload: copy the `count` values with CTF linear
 * (column-major, first index fastest) indices [first, first + count) of
 * tensor `tensor` (enum atrip_tensor) into `out`; for complex systems each
 * value is an interleaved (re, im) pair of doubles.  Return 0 on success;
 * anything else aborts the upload with an error. */
typedef int (*atrip_read_fn)(void *user, int tensor, int64_t first,
                             int64_t count, double *out);

/* Flags of atrip_hip_upload_stream. */
enum {
  ATRIP_STREAM_WITH_J = 1,   /* Jijka and Jabci are read ((cT) pass) */
  ATRIP_STREAM_COMPLEX = 2   /* complex values, as atrip_hip_upload_complex */
};

/* Upload through a reader, so that the caller never holds a whole tensor in
 * host memory: each large tensor is requested in slabs of its slowest CTF
 * index (Tabij and Vabij per j, Vijka and Jijka per a, Vabci and Jabci per
 * i), through a pinned host buffer of at most min(stage_kb, 256 MiB) bytes,
 * and re-laid out on the device chunk by chunk; ei, ea and Tai are read
 * whole.  Replaces the CTF slicing of the reference
 * (Tensor::slice per SliceUnion, src/atrip/Unions.hpp:22-75) and its
 * read_all of the small tensors (src/atrip/Atrip.cxx:176-181).  Same
 * resident system, bit for bit, as atrip_hip_upload(_complex) of the same
 * values; honours operand_store like the other uploads. */
int atrip_hip_upload_stream(atrip_hip_ctx *ctx, int64_t No, int64_t Nv,
                            int flags, atrip_read_fn read, void *user,
                            int *sys);

/* Complex systems (atrip::Atrip::run<Complex>, src/atrip/Atrip.cxx:1136):
 * the desc pointers address interleaved (re, im) pairs, i.e. the memory of
 * std::complex<double> CTF tensors.  Runs of a complex system use FP64
 * (precision ATRIP_FP64); its energies follow the reference's complex GEMM
 * path (conjugated Vhhh in the hole terms, conj(W) in the energy, real part
 * of each tuple's sum).  The synthetic variant draws the imaginary parts
 * from the generator's ATRIP_SYN_IMAG streams (orbital energies real). */
int atrip_hip_upload_complex(atrip_hip_ctx *ctx, const atrip_sys_desc *desc,
                             int *sys);
int atrip_hip_upload_synthetic_complex(atrip_hip_ctx *ctx, int64_t No,
                                       int64_t Nv, uint64_t seed, int with_j,
                                       int *sys);

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

/* Bytes of the fixed device region of atrip_hip_run_noalloc (per GPU). */
#define ATRIP_NOALLOC_SCRATCH_BYTES (256ull * 1024ull * 1024ull)

/* atrip_hip_run without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and sign as atrip_hip_run; per-tuple energies are
 * bitwise those of atrip_hip_run, and E(T) is the same sum when the same
 * windows per launch fit (else it agrees to rounding).  All device scratch
 * (class buffers, W cubes, slot energies, batch partials, window list)
 * comes from one fixed region of ATRIP_NOALLOC_SCRATCH_BYTES per GPU, held
 * from the first call on that GPU until the process ends; the windows per
 * launch shrink until a launch fits.  Fails with a clear error, before any
 * launch, when one window of the context's "window" edge needs more
 * (FP64 without (cT): No=40 fits edge 5, No=60 edge 3; the message names
 * the largest edge that fits), and for a mixed form whose operand copies are not
 * resident yet (upload with that operand_store, or run the form once
 * through atrip_hip_run).  No timing events are recorded.  Calls share the
 * GPU's region: calls that overlap on one GPU (two threads, or contexts
 * on different streams of one device) race and give wrong results. */
int atrip_hip_run_noalloc(atrip_hip_ctx *ctx, int sys, int precision,
                          int ijkabc, uint64_t first, uint64_t count,
                          double *energy, double *ct_energy);

/* Plan that atrip_hip_run takes for a shape, on the host only (no device
 * needed): *nb the window edge (`window` 0 = automatic, as the context
 * option) and *fm x *fn the doubles kernel's wave tile in MFMA fragments
 * (16x16 FP64 fragments for ATRIP_FP64, 32x32 for the other forms), the
 * automatic pick of a context whose "tile" option is -1.  No counterpart in
 * the reference, which has no tiles; for callers that size or schedule runs
 * from the plan.  Returns nonzero for No or Nv below 1, a window outside
 * 0..32 or an unknown precision. */
int atrip_hip_plan_model(int64_t No, int64_t Nv, int precision, int window,
                         int *nb, int *fm, int *fn);

/* (T) energy of a systematic sample of the tuple space.  The window plan
 * cuts the virtual index into blocks of edge "window"; its windows are the
 * block triples (alpha <= beta <= gamma) holding a tuple, and a row is all
 * windows of one block pair (alpha, beta).  The sample is the rows offset,
 * offset + stride, offset + 2 stride, ... in lexicographic order, every
 * valid tuple of each; *tuples receives their number.  Rows keep the
 * operand reuse of a contiguous run, and the sample spans the whole
 * first-index range: for measuring throughput and precision-form
 * differences on shapes whose full tuple space takes too long.  stride 1
 * is the whole space.  No counterpart in the reference. */
int atrip_hip_run_windows(atrip_hip_ctx *ctx, int sys, int precision,
                          int ijkabc, uint64_t stride, uint64_t offset,
                          double *energy, double *ct_energy,
                          uint64_t *tuples);

/* Per-tuple energies (raw, before the sign) for the tuples
 * [first, first+count) in lexicographic order; e_ct may be NULL.  The
 * reference keeps these only in a never-filled debug map
 * (src/atrip/Atrip.cxx:575). */
int atrip_hip_tuple_energies(atrip_hip_ctx *ctx, int sys, int precision,
                             int ijkabc, uint64_t first, uint64_t count,
                             double *e, double *e_ct);

/* Tuning and instrumentation knobs (no counterpart in the reference):
 *   "window"  block edge nb of the tuple windows (1..32); 0 (default):
 *             chosen per system shape (8 unless a modelled tile fit of
 *             edge 6 or 12 is 15 % better)
 *   "batch"   windows per launch
 *   "timing"  1: time every hot-kernel launch with HIP events
 *   "tile"    doubles-kernel tile variant: 0..7 for the FP64 kernel, 0..9
 *             for the FP32-product kernels (a variant a kernel lacks
 *             means automatic for it), -1 (default) = automatic
 *   "epilogue"  0 (default): fused class sums + energy per block triple;
 *               1: class sums written as W cubes first (complex systems
 *               always take this form)
 *   "serve_plain_no"  servers run requests with No at or above it through
 *               the plain class-GEMM path and the rest in the persistent
 *               batched launch (default 16; a bf16x3 server runs every
 *               request on the plain path)
 *   "stage_kb"  device staging buffer of host uploads, KiB (default 4 GiB;
 *               at least one slab of the slowest CTF index is staged per
 *               step whatever the budget), also the bound of the pinned
 *               host buffer of atrip_hip_upload_stream (at most 256 MiB)
 *   "operand_store"  how the next uploads keep the doubles operands:
 *                    0 FP64 (split or bf16x3 copies made by the first
 *                    mixed run of that form), 1 split FP32 only, 2 FP32
 *                    only, 3 bf16x3 only (no FP64 copy: for systems whose
 *                    FP64 and mixed copies do not fit together)
 *   "batch_scratch_mb"  device scratch bound of the batched form (MiB)
 *   "batch_grid"        batched persistent grid (0 = occupancy x CUs)
 *   "batch_order"       batched dispatch: 0 longest window first,
 *                       1 canonical order
 *   "batch_nb", "batch_tile"  force the batched window edge (0 = tuned) and
 *                       tile variant (-1 = tuned) */
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

/*
 * Batched form: the full tuple spaces of nsys resident systems (No <= 64) in
 * one persistent launch, in `precision` (ATRIP_FP64, ATRIP_SPLIT_FP32 or
 * ATRIP_FP32; the mixed forms widen each FP32 tile into an FP64 W;
 * ATRIP_BF16X3 is refused); energies[s] receives E(T) of
 * systems[s] with the sign of src/atrip/Atrip.cxx:1108-1111.  No
 * counterpart in the reference, which runs one system per atrip::Atrip::run
 * (src/atrip/Atrip.cxx:686-1111); each system's result is that of
 * atrip_hip_run on it (same formulas; summation order of the windows
 * differs), and does not depend on the other systems of the batch.
 * Window edge and tile variant of each system come from the tuner (below);
 * shapes not tuned yet use the cost model.
 */
int atrip_hip_batch_run(atrip_hip_ctx *ctx, const int *systems, int nsys,
                        int precision, double *energies);

/*
 * Autotuner of the batched form.  mode 0: cost model only; mode 1: for each
 * distinct shape (No, Nv rounded up to 16) among the systems, time the six
 * (window edge x tile variant) candidates the cost model ranks best, with
 * the persistent kernel itself on one window per workgroup of that shape,
 * and keep the fastest per tuple.  The choices are kept in the context and
 * used by atrip_hip_batch_run and the serving loop.
 */
int atrip_hip_batch_tune(atrip_hip_ctx *ctx, const int *systems, int nsys,
                         int precision, int mode);

/* Plan used for a resident system: window edge, tile variant (wave tile of
 * 16*fm x 16*fn outputs), whether it was measured, and the estimated
 * seconds per full window. */
int atrip_hip_batch_plan(atrip_hip_ctx *ctx, int sys, int precision, int *nb,
                         int *fm, int *fn, int *measured, double *sec_window);

/* Cost-model choice for a shape, on the host only (no device needed). */
int atrip_hip_batch_model(int64_t No, int64_t Nv, int precision,
                          int64_t scratch_mb, int grid, int *nb, int *fm,
                          int *fn);

/*
 * Serving form: a request loop with continuous batching over one context.
 * No counterpart in the reference (one Atrip::run per process).  Requests
 * wait in a FIFO queue and are admitted while a slot is free and their
 * resident arrays fit hbm_budget_bytes; each atrip_hip_serve_step admits
 * what fits and runs one persistent launch over windows of every active
 * request (shortest remaining work first by default), about iter_ms long,
 * then retires finished requests and frees their memory.  All requests of
 * a server compute in its `precision` (ATRIP_FP64, ATRIP_SPLIT_FP32 or
 * ATRIP_FP32); the systems keep only that operand form resident
 * (operand_store 0, 1 or 2).  A request's energy does not
 * depend on the mix it was served in.  The server uses its
 * context's stream and must be driven from the context's thread; the
 * context must outlive it.  Errors are reported on the context.
 */
typedef struct atrip_hip_server atrip_hip_server;

int atrip_hip_serve_create(atrip_hip_ctx *ctx, uint64_t hbm_budget_bytes,
                           int max_active, double iter_ms, int precision,
                           atrip_hip_server **out);
int atrip_hip_serve_destroy(atrip_hip_server *srv);
/* 0: shortest remaining work first (default), 1: equal share per request */
int atrip_hip_serve_set_policy(atrip_hip_server *srv, int policy);
/* Milliseconds since the server was created (the clock of all times). */
double atrip_hip_serve_clock_ms(atrip_hip_server *srv);
/* Enqueue a seeded synthetic system (generated on the device when admitted)
 * or host tensors (read when admitted: they must stay valid until then).
 * arrival_ms < 0 stamps the request with the current clock. */
int atrip_hip_serve_submit_synthetic(atrip_hip_server *srv, int64_t No,
                                     int64_t Nv, uint64_t seed,
                                     double arrival_ms, int64_t *request);
int atrip_hip_serve_submit(atrip_hip_server *srv, const atrip_sys_desc *desc,
                           double arrival_ms, int64_t *request);
/* Admission, one launch, retirement; *finished = requests ended by it. */
int atrip_hip_serve_step(atrip_hip_server *srv, int *finished);
/* State (0 queued, 1 active, 2 done, 3 failed), E(T) when done, and the
 * arrival, admission and completion times in ms.  Fails for a failed
 * request, with its reason. */
int atrip_hip_serve_result(atrip_hip_server *srv, int64_t request, int *state,
                           double *energy, double *t_submit_ms,
                           double *t_admit_ms, double *t_done_ms);
int atrip_hip_serve_stats(atrip_hip_server *srv, int *queued, int *active,
                          uint64_t *resident_bytes, uint64_t *peak_bytes,
                          uint64_t *tuples_done, uint64_t *launches,
                          double *busy_ms);

#ifdef __cplusplus
}
#endif

#endif /* ATRIP_HIP_H */
