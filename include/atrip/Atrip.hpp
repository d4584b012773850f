// atrip::Atrip -- the reference's (T) interface, served by the MI355X-native
// HIP path through the C-ABI of include/atrip_hip.h.
//
// Source-compatible with the reference header src/atrip/Atrip.hpp:41-140:
// the same Input<F> fields and with_* builders (:68-107), the same attribute
// names and defaults (:113-130), the same Output (:134-137) and the same
// run<F> entry point (:138-139).  Existing callers recompile unchanged; MPI
// is not needed (single process, one GPU), so init() accepts any
// communicator value and ignores it.
//
// Differences a caller can observe are listed next to the members they
// concern.  Errors are thrown as std::string, the reference's convention
// (src/atrip/Acc.hpp:95-107).
#pragma once

#include <complex>
#include <cstddef>
#include <functional>
#include <string>

#if defined(ATRIP_USE_CYCLOPS)
#include <ctf.hpp>
#else
#include <atrip/ctf_shim.hpp>
#endif

#define ATRIP_ADD_ATTRIBUTE(_type, _name, _default) \
  _type _name = _default;                           \
  Input &with_##_name(_type i) {                    \
    _name = i;                                      \
    return *this;                                   \
  }

namespace atrip {

using Complex = std::complex<double>;

// Progress callback, src/atrip/Debug.hpp:95-109 of the reference.  Called at
// the end of every processed segment of tuples.
struct IterationDescription;
using IterationDescriptor = std::function<void(IterationDescription const &)>;
struct IterationDescription {
  static IterationDescriptor descriptor;
  size_t current_iteration;
  size_t total_iterations;
  double current_elapsed_time;
};
void register_iteration_descriptor(IterationDescriptor);

struct Atrip {
  static size_t rank;  // always 0: one process drives one GPU
  static size_t np;    // always 1

  // Reference: static void init(MPI_Comm) (src/atrip/Atrip.hpp:65).
  template <typename Comm>
  static void init(Comm) {
    init();
  }
  static void init();

  template <typename F = double>
  struct Input {
    CTF::Tensor<F> *ei = nullptr, *ea = nullptr, *Tph = nullptr,
                   *Tpphh = nullptr, *Vpphh = nullptr, *Vhhhp = nullptr,
                   *Vppph = nullptr, *Jppph = nullptr, *Jhhhp = nullptr;
    Input &with_epsilon_i(CTF::Tensor<F> *t) {
      ei = t;
      return *this;
    }
    Input &with_epsilon_a(CTF::Tensor<F> *t) {
      ea = t;
      return *this;
    }
    Input &with_Tai(CTF::Tensor<F> *t) {
      Tph = t;
      return *this;
    }
    Input &with_Tabij(CTF::Tensor<F> *t) {
      Tpphh = t;
      return *this;
    }
    Input &with_Vabij(CTF::Tensor<F> *t) {
      Vpphh = t;
      return *this;
    }
    Input &with_Vijka(CTF::Tensor<F> *t) {
      Vhhhp = t;
      return *this;
    }
    Input &with_Vabci(CTF::Tensor<F> *t) {
      Vppph = t;
      return *this;
    }
    Input &with_Jijka(CTF::Tensor<F> *t) {
      Jhhhp = t;
      return *this;
    }
    Input &with_Jabci(CTF::Tensor<F> *t) {
      Jppph = t;
      return *this;
    }
    // Accepted for source compatibility; one GPU always runs the whole
    // tuple space in window order.
    enum TuplesDistribution { NAIVE, GROUP_AND_SORT };

    ATRIP_ADD_ATTRIBUTE(bool, delete_Vppph, false)
    ATRIP_ADD_ATTRIBUTE(bool, rank_round_robin, false)  // no effect (1 rank)
    ATRIP_ADD_ATTRIBUTE(bool, chrono, false)
    ATRIP_ADD_ATTRIBUTE(bool, barrier, false)           // no effect (1 rank)
    ATRIP_ADD_ATTRIBUTE(bool, blocking, false)          // no effect
    ATRIP_ADD_ATTRIBUTE(size_t, max_iterations, 0)
    ATRIP_ADD_ATTRIBUTE(int, iteration_mod, -1)
    ATRIP_ADD_ATTRIBUTE(int, percentage_mod, -1)
    ATRIP_ADD_ATTRIBUTE(TuplesDistribution, tuples_distribution, NAIVE)
    ATRIP_ADD_ATTRIBUTE(std::string, checkpoint_path, "atrip-checkpoint.yaml")
    ATRIP_ADD_ATTRIBUTE(bool, read_checkpoint_if_exists, true)
    ATRIP_ADD_ATTRIBUTE(bool, writeCheckpoint, true)
    ATRIP_ADD_ATTRIBUTE(float, checkpoint_at_percentage, 10)
    ATRIP_ADD_ATTRIBUTE(size_t, checkpoint_at_every_iteration, 0)
    ATRIP_ADD_ATTRIBUTE(bool, ijkabc, 0)
    ATRIP_ADD_ATTRIBUTE(size_t, ooo_threads, 0)  // no effect (no such kernel)
    ATRIP_ADD_ATTRIBUTE(size_t, ooo_blocks, 0)   // no effect
    // Additions of this build (no counterpart in the reference):
    ATRIP_ADD_ATTRIBUTE(int, device, 0)          // HIP device index
    ATRIP_ADD_ATTRIBUTE(bool, mixed_precision, false)  // mixed doubles:
    // "bf16x3" (three bf16 parts per operand, six part products on bf16
    // MFMA: the fastest and most accurate), "split_fp32" (hi*hi on FP32
    // MFMA plus hi*lo + lo*hi on bf16 MFMA) or "fp32" (hi*hi only); FP32
    // accumulation and FP64 energy in every form
    ATRIP_ADD_ATTRIBUTE(std::string, mixed_form, "bf16x3")
    // Upload staging (KiB): device chunk of each large tensor and bound of
    // the pinned host buffer its slabs are read through (at most 256 MiB)
    ATRIP_ADD_ATTRIBUTE(int64_t, stage_kb, (int64_t)4 << 20)
  };

  struct Output {
    double energy;
    double ct_energy;
  };

  template <typename F = double>
  static Output run(Input<F> const &in);
};

}  // namespace atrip
