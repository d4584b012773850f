// atrip::Atrip::run on the MI355X C-ABI.  Host side of the drop-in: reads the
// CTF tensors (reference src/atrip/Atrip.cxx:72-73, :176-181), uploads them
// once, honours the checkpoint and max_iterations semantics of the
// reference's main loop (:587-621, :686-1057) and applies its final sign
// (:1094-1111).  The per-tuple work runs on the GPU through atrip_hip_run.
#include <atrip/Atrip.hpp>
#include <atrip/Checkpoint.hpp>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <type_traits>
#include <vector>

#include "../../include/atrip_hip.h"

namespace atrip {

size_t Atrip::rank = 0;
size_t Atrip::np = 1;
IterationDescriptor IterationDescription::descriptor;

void register_iteration_descriptor(IterationDescriptor d) {
  IterationDescription::descriptor = d;
}

void Atrip::init() {
  rank = 0;
  np = 1;
}

namespace {

struct Ctx {
  atrip_hip_ctx *h = nullptr;
  explicit Ctx(int device) {
    if (atrip_hip_create(device, &h) != 0)
      throw std::string("atrip: cannot open HIP device ") +
          std::to_string(device);
  }
  ~Ctx() { atrip_hip_destroy(h); }
  void check(int rc, const char *what) const {
    if (rc != 0)
      throw std::string("atrip: ") + what + ": " + atrip_hip_last_error(h);
  }
};

template <typename F>
void check_size(CTF::Tensor<F> *t, const char *name, size_t expect) {
  if (!t) throw std::string("atrip: missing input tensor ") + name;
  size_t n = 1;
  for (int d = 0; d < t->order; d++) n *= (size_t)t->lens[d];
  if (n != expect)
    throw std::string("atrip: tensor ") + name + " has " + std::to_string(n) +
        " elements, expected " + std::to_string(expect);
}

// Elements [first, first + count) of a tensor in column-major order.  The
// shim serves the contiguous range directly; with Cyclops the range is
// asked for by global index, CTF's Tensor::read(npair, global_idx, data).
template <typename F>
void read_range(CTF::Tensor<F> *t, int64_t first, int64_t count, F *out) {
#if defined(ATRIP_USE_CYCLOPS)
  std::vector<int64_t> idx((size_t)count);
  for (int64_t q = 0; q < count; q++) idx[q] = first + q;
  t->read(count, idx.data(), out);
#else
  t->read_range(first, count, out);
#endif
}

// The input tensors in atrip_tensor order, read slab by slab by the
// library (atrip_hip_upload_stream): the host never holds a copy of a
// large tensor, where the reference slices each one per SliceUnion
// (src/atrip/Unions.hpp:22-75).
template <typename F>
struct TensorReader {
  CTF::Tensor<F> *t[9] = {};
  std::string error;
  static int read(void *user, int tensor, int64_t first, int64_t count,
                  double *out) {
    auto *r = static_cast<TensorReader *>(user);
    try {
      if (tensor < 0 || tensor > 8 || !r->t[tensor])
        throw std::string("no tensor with id ") + std::to_string(tensor);
      read_range(r->t[tensor], first, count, reinterpret_cast<F *>(out));
      return 0;
    } catch (const std::string &e) {
      r->error = e;
    } catch (const std::exception &e) {
      r->error = e.what();
    }
    return 1;
  }
};

template <typename F>
Atrip::Output run_impl(Atrip::Input<F> const &in, size_t np);

}  // namespace

template <>
Atrip::Output Atrip::run<double>(Atrip::Input<double> const &in) {
  return run_impl<double>(in, np);
}

// Reference: explicit instantiation for Complex (src/atrip/Atrip.cxx:1136).
template <>
Atrip::Output Atrip::run<Complex>(Atrip::Input<Complex> const &in) {
  if (in.mixed_precision)
    throw std::string("atrip: the complex (T) path runs in FP64 only");
  return run_impl<Complex>(in, np);
}

namespace {

template <typename F>
Atrip::Output run_impl(Atrip::Input<F> const &in, size_t np) {
  if (!in.ei || !in.ea) throw std::string("atrip: epsilon tensors missing");
  const size_t No = (size_t)in.ei->lens[0];
  const size_t Nv = (size_t)in.ea->lens[0];
  std::cout << "Atrip: No: " << No << "\n"
            << "Atrip: Nv: " << Nv << "\n"
            << "Atrip: np: " << np << "\n";

  const bool with_j = in.Jhhhp && in.Jppph;
  TensorReader<F> reader;
  CTF::Tensor<F> *ts[9] = {in.ei,    in.ea,    in.Tph,   in.Tpphh, in.Vpphh,
                           in.Vhhhp, in.Vppph, in.Jhhhp, in.Jppph};
  const char *names[9] = {"ei",    "ea",    "Tai",   "Tabij", "Vabij",
                          "Vijka", "Vabci", "Jijka", "Jabci"};
  const size_t sizes[9] = {No,           Nv,           Nv * No,
                           Nv * Nv * No * No, Nv * Nv * No * No,
                           No * No * No * Nv, Nv * Nv * Nv * No,
                           No * No * No * Nv, Nv * Nv * Nv * No};
  for (int q = 0; q < (with_j ? 9 : 7); q++) {
    check_size(ts[q], names[q], sizes[q]);
    reader.t[q] = ts[q];
  }

  Ctx ctx(in.device);
  ctx.check(atrip_hip_set_option(ctx.h, "stage_kb", in.stage_kb), "stage_kb");
  int sys = 0;
  const int flags = (with_j ? ATRIP_STREAM_WITH_J : 0) |
                    (std::is_same<F, Complex>::value ? ATRIP_STREAM_COMPLEX : 0);
  if (atrip_hip_upload_stream(ctx.h, (int64_t)No, (int64_t)Nv, flags,
                              &TensorReader<F>::read, &reader, &sys) != 0) {
    std::string why = atrip_hip_last_error(ctx.h);
    if (!reader.error.empty()) why += " (" + reader.error + ")";
    throw std::string("atrip: upload: ") + why;
  }
  if (in.delete_Vppph) delete in.Vppph;  // reference: src/atrip/Atrip.cxx:310

  const uint64_t n = (uint64_t)Nv * (Nv + 1) * (Nv + 2) / 6 - Nv;
  int precision = ATRIP_FP64;
  if (in.mixed_precision) {
    if (in.mixed_form == "split_fp32") precision = ATRIP_SPLIT_FP32;
    else if (in.mixed_form == "fp32") precision = ATRIP_FP32;
    else if (in.mixed_form == "bf16x3") precision = ATRIP_BF16X3;
    else
      throw std::string("atrip: unknown mixed_form '" + in.mixed_form +
                        "' (split_fp32, fp32 or bf16x3)");
  }

  // checkpoint resume (reference :587-621): the stored energy is physical
  uint64_t first = 0;
  double energy_base = 0.0;  // raw-convention energy carried in
  if (in.read_checkpoint_if_exists) {
    std::ifstream probe(in.checkpoint_path);
    if (probe.is_open()) {
      std::cout << "Atrip: Reading checkpoint from " << in.checkpoint_path
                << "\n";
      const Checkpoint c = read_checkpoint(probe);
      first = std::min<uint64_t>(c.iteration, n);
      energy_base = -c.energy;
      std::cout << "Atrip: energy from checkpoint " << energy_base << "\n"
                << "Atrip: iteration from checkpoint " << first << "\n";
    }
  }
  // max_iterations: the reference processes tuple i, then stops when
  // i >= max_iterations (:1052-1056), i.e. indices first..max_iterations.
  uint64_t last = n;
  if (in.max_iterations != 0 && first < n)
    last = std::min<uint64_t>(n, std::max<uint64_t>(first + 1,
                                                    in.max_iterations + 1));

  // Checkpoint interval as the reference computes it (:588-591).  The
  // reference writes only when that interval is 0 (its interval branch is
  // disabled by `&& false`, :713), then before every tuple; here the writes
  // happen at segment ends, one segment being 1 % of the range.  With a
  // nonzero interval nothing is written, and the interval only paces the
  // progress callback.
  const uint64_t mod =
      in.checkpoint_at_every_iteration != 0
          ? in.checkpoint_at_every_iteration
          : (uint64_t)((double)n * in.checkpoint_at_percentage / 100.0);
  const bool write_ck = mod == 0 && in.writeCheckpoint;
  // (at most ~1000 segments per run)
  uint64_t seg = mod != 0 ? std::max<uint64_t>(mod, (last - first) / 1000)
                          : std::max<uint64_t>(1, (last - first) / 100);
  if (seg == 0) seg = 1;

  // Progress lines at the reference's reporting points (:402-404, :734):
  // iteration numbers (tuple index + 1) that are multiples of iteration_mod,
  // or of percentage_mod % of the run when that is set, and the 1 % point.
  // An iteration_mod of -1 (the default) converts to SIZE_MAX as there, so
  // only the 1 % point reports.  Here the line comes at the end of the
  // segment that reaches such a point and names the segment's last iteration.
  const uint64_t report_mod =
      in.percentage_mod > 0
          ? (uint64_t)((double)n * in.percentage_mod / 100.0)
          : (uint64_t)(size_t)in.iteration_mod;
  const uint64_t one_percent = (uint64_t)((double)n * 0.01);
  // the reference counts complex products as four real ones (bench/main.cxx:119)
  const double gflop_tuple = 12.0 * (double)No * No * No * (double)(No + Nv) /
                             1e9 * (std::is_same<F, Complex>::value ? 4.0 : 1.0);

  const auto t0 = std::chrono::steady_clock::now();
  double raw = 0.0, raw_ct = 0.0;  // sums before the final sign
  for (uint64_t s0 = first; s0 < last; s0 += seg) {
    const uint64_t cnt = std::min(seg, last - s0);
    double e = 0.0, ect = 0.0;
    // ijkabc=0 returns -sum; ijkabc=1 returns +sum with Tai negated
    ctx.check(atrip_hip_run(ctx.h, sys, precision, in.ijkabc ? 1 : 0, s0, cnt,
                            &e, &ect),
              "run");
    raw += in.ijkabc ? e : -e;
    raw_ct += in.ijkabc ? ect : -ect;
    const double elapsed =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
            .count();
    if (write_ck && s0 + cnt < last) {
      Checkpoint c;
      c.no = No;
      c.nv = Nv;
      c.nranks = 1;
      c.nnodes = 1;
      c.energy = -(energy_base + raw);
      c.iteration = s0 + cnt;
      c.rank_round_robin = in.rank_round_robin;
      write_checkpoint(c, in.checkpoint_path);
    }
    if (IterationDescription::descriptor)
      IterationDescription::descriptor({s0 + cnt, n, elapsed});
    const uint64_t it0 = s0 + 1, it1 = s0 + cnt;  // iterations of the segment
    if (report_mod == 0 || it1 / report_mod > (it0 - 1) / report_mod ||
        (one_percent >= it0 && one_percent <= it1)) {
      std::cout << "Atrip: iteration " << it1 << " [" << 100 * it1 / n
                << "%] ("
                << (elapsed > 0 ? gflop_tuple * (double)(it1 - first) / elapsed
                                : -1.0)
                << "GF)" << std::endl;
      if (in.chrono) std::cout << "  iterations :: " << elapsed << std::endl;
    }
  }
  const double elapsed =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
          .count();

  Atrip::Output out{energy_base + raw, raw_ct};
  if (!in.ijkabc) {
    out.energy = -out.energy;
    out.ct_energy = -out.ct_energy;
  }
  std::cout << "Atrip: Energy: " << std::setprecision(15) << std::setw(23)
            << out.energy << std::endl;
  std::cout << "Atrip: Energy (cT): " << std::setprecision(15) << std::setw(23)
            << out.ct_energy << std::endl;
  if (elapsed > 0)
    std::cout << "atrip:flops(iterations): "
              << gflop_tuple * (double)(last - first) / elapsed << "\n";
  return out;
}

}  // namespace

}  // namespace atrip
