// Allocation-free run: atrip_hip_run with all of its device scratch in one
// fixed region of this code object (`scratch`), for callers that cannot
// call hipMalloc (work under graph capture, tight serving loops, harnesses
// that forbid allocation).  No counterpart in the reference.
//
// Every GPU holds its own copy of `scratch` from the moment this code
// object is loaded on it until the process ends.  Its device address is
// looked up once per device with hipGetSymbolAddress, with that device
// current, and only that address is handed to kernels and copies (in host
// code the array's name is a host-side stand-in, not a device address).
// All calls on one GPU share the region and it keeps the last call's
// contents; nothing in it needs to start at zero (every kernel writes what
// it later reads), and calls that overlap on one GPU (two host threads, or
// two contexts on different streams) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "host.hpp"

using namespace atrip_amd;

namespace {

static __device__ __align__(256) unsigned char scratch[ATRIP_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex addr_mutex;
void *addr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device `dev`.
void *scratch_on(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES)
    throw Fail{"device index " + std::to_string(dev) +
               " beyond the allocation-free table"};
  std::lock_guard<std::mutex> lock(addr_mutex);
  if (!addr[dev]) {
    void *p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
          "hipGetSymbolAddress(scratch)");
    if (!p) throw Fail{"hipGetSymbolAddress(scratch) returned null"};
    addr[dev] = p;
  }
  return addr[dev];
}

}  // namespace

extern "C" int atrip_hip_run_noalloc(atrip_hip_ctx *ctx, int sys,
                                     int precision, int ijkabc,
                                     uint64_t first, uint64_t count,
                                     double *energy, double *ct_energy) {
  return guarded(ctx, [&] {
    if (!energy) throw Fail{"null energy pointer"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    void *base = scratch_on(ctx->device);
    run_fixed(ctx, sys, precision, ijkabc, first, count, base,
              (size_t)ATRIP_NOALLOC_SCRATCH_BYTES, energy, ct_energy);
  });
}
