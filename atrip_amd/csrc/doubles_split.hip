// Split-FP32 doubles contractions (placeholder until the mixed form lands).
#include "internal.hpp"

namespace atrip_amd {

hipError_t launch_doubles_split(const DevSystem &, const BatchArgs &, bool,
                                double *, hipStream_t) {
  return hipErrorNotSupported;
}

}  // namespace atrip_amd
