// Host-side helpers shared by the C-ABI translation units: the error
// convention (every entry point returns 0 or a non-zero status and keeps the
// message for atrip_hip_last_error) and the resident-system lookup.
#pragma once

#include <hip/hip_runtime.h>

#include <exception>
#include <string>
#include <vector>

#include "internal.hpp"

namespace atrip_amd {

struct Fail {
  std::string msg;
};

inline void check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw Fail{std::string(what) + ": " + hipGetErrorString(e)};
}

template <typename F>
int guarded(atrip_hip_ctx *ctx, F &&f) {
  if (!ctx) return 1;
  try {
    ctx->err.clear();
    f();
    return 0;
  } catch (const Fail &x) {
    ctx->err = x.msg;
  } catch (const std::exception &x) {
    ctx->err = x.what();
  }
  if (ctx->err.empty()) ctx->err = "unknown failure";
  // HIP keeps the error of a failed runtime call (an allocation that found
  // no memory, say) until it is read; read it here so that the next launch
  // check on this thread does not report it again.
  (void)hipGetLastError();
  return 2;
}

// Resident system of a handle; throws Fail for an unknown handle.
DevSystem &get_sys(atrip_hip_ctx *ctx, int sys);

// Split-FP32 copies of a system's doubles operands (made once).
void ensure_split(atrip_hip_ctx *ctx, DevSystem &s);
// bf16x3 copies of a system's doubles operands (made once).
void ensure_tri(atrip_hip_ctx *ctx, DevSystem &s);
// Throws Fail unless `precision` is known and the system's operand store
// holds what it reads: FP64 needs store 0, split FP32 0 or 1, FP32 0-2,
// bf16x3 0 or 3; complex systems run in FP64 only.
void check_precision(const DevSystem &s, int precision);

// Device bytes of the resident arrays of a system uploaded with operand
// store mode `store` (what admission charges).
uint64_t system_bytes(int64_t No, int64_t Nv, bool with_j, int store);

// atrip_hip_run over [first, first + count) with every piece of device
// scratch (class buffers, W cubes, slot energies, batch partials, window
// list) carved out of [base, base + bytes): no device allocation, no timing
// events.  The windows per launch shrink until a launch fits; a run whose
// single window does not fit, or a mixed form whose operand copies are not
// resident yet, throws Fail.  Writes the signed energies like atrip_hip_run.
void run_fixed(atrip_hip_ctx *ctx, int sys, int precision, int ijkabc,
               uint64_t first, uint64_t count, void *base, size_t bytes,
               double *energy, double *ct_energy);

// Window edge of the plain path for system s: the "window" option, or
// when it is 0 the automatic edge of the system's shape.
int window_edge(const atrip_hip_ctx *ctx, const DevSystem &s);
// Number of windows of the plain-path plan of edge nb of a system (the
// lexicographic block triples holding a tuple).
uint64_t plan_windows(atrip_hip_ctx *ctx, int sys, int nb);
// Plain-path run (doubles kernel, energy epilogue) over windows
// [wbegin, wend) of the plan of edge nb of a system, every tuple of each,
// in launches of the context's "batch" windows.  out receives per window,
// in plan order, the (sum, comp) of a compensated sum of its raw tuple
// energies in slot order, so a window's pair does not depend on how windows
// were grouped into launches or calls.
void run_plan_windows(atrip_hip_ctx *ctx, int sys, int precision, int nb,
                      uint64_t wbegin, uint64_t wend,
                      std::vector<double> &out);

// Number of tuples a <= b <= c < Nv, not all equal.
inline uint64_t n_tuples(uint64_t Nv) {
  return Nv * (Nv + 1) * (Nv + 2) / 6 - Nv;
}

}  // namespace atrip_amd
