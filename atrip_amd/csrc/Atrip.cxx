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
std::vector<F> read_tensor(CTF::Tensor<F> *t, const char *name,
                           size_t expect) {
  if (!t) throw std::string("atrip: missing input tensor ") + name;
  size_t n = 1;
  for (int d = 0; d < t->order; d++) n *= (size_t)t->lens[d];
  if (n != expect)
    throw std::string("atrip: tensor ") + name + " has " + std::to_string(n) +
        " elements, expected " + std::to_string(expect);
  std::vector<F> v(n);
  t->read_all(v.data());
  return v;
}

// Host memory of a tensor as the C-ABI sees it: doubles, interleaved
// (re, im) for complex tensors (std::complex<double> is two doubles).
template <typename F>
const double *raw(const std::vector<F> &v) {
  return reinterpret_cast<const double *>(v.data());
}

template <typename F>
int upload(atrip_hip_ctx *h, const atrip_sys_desc *d, int *sys);
template <>
int upload<double>(atrip_hip_ctx *h, const atrip_sys_desc *d, int *sys) {
  return atrip_hip_upload(h, d, sys);
}
template <>
int upload<Complex>(atrip_hip_ctx *h, const atrip_sys_desc *d, int *sys) {
  return atrip_hip_upload_complex(h, d, sys);
}

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

  // host copies in CTF layout (the reference's read_all, :179-181)
  std::vector<F> ei = read_tensor(in.ei, "ei", No);
  std::vector<F> ea = read_tensor(in.ea, "ea", Nv);
  std::vector<F> tai = read_tensor(in.Tph, "Tai", Nv * No);
  std::vector<F> tabij = read_tensor(in.Tpphh, "Tabij", Nv * Nv * No * No);
  std::vector<F> vabij = read_tensor(in.Vpphh, "Vabij", Nv * Nv * No * No);
  std::vector<F> vijka = read_tensor(in.Vhhhp, "Vijka", No * No * No * Nv);
  std::vector<F> vabci = read_tensor(in.Vppph, "Vabci", Nv * Nv * Nv * No);
  std::vector<F> jijka, jabci;
  const bool with_j = in.Jhhhp && in.Jppph;
  if (with_j) {
    jijka = read_tensor(in.Jhhhp, "Jijka", No * No * No * Nv);
    jabci = read_tensor(in.Jppph, "Jabci", Nv * Nv * Nv * No);
  }

  Ctx ctx(in.device);
  atrip_sys_desc d{};
  d.No = (int64_t)No;
  d.Nv = (int64_t)Nv;
  d.ei = raw(ei);
  d.ea = raw(ea);
  d.Tai = raw(tai);
  d.Tabij = raw(tabij);
  d.Vabij = raw(vabij);
  d.Vijka = raw(vijka);
  d.Vabci = raw(vabci);
  d.Jijka = with_j ? raw(jijka) : nullptr;
  d.Jabci = with_j ? raw(jabci) : nullptr;
  int sys = 0;
  ctx.check(upload<F>(ctx.h, &d, &sys), "upload");
  // host copies are not needed once resident
  std::vector<F>().swap(vabci);
  std::vector<F>().swap(jabci);
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
  // the reference counts complex products as four real ones (bench/main.cxx:119)
  const double gflop_tuple = 12.0 * (double)No * No * No * (double)(No + Nv) /
                             1e9 * (std::is_same<F, Complex>::value ? 4.0 : 1.0);
  if (elapsed > 0)
    std::cout << "atrip:flops(iterations): "
              << gflop_tuple * (double)(last - first) / elapsed << "\n";
  return out;
}

}  // namespace

}  // namespace atrip
