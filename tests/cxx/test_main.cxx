// C++ tests of the drop-in interface, written in the style of the reference's
// test/main.cxx (TESTCASE blocks with assert).  Without arguments only the
// host-side cases run; with --gpu the atrip::Atrip::run cases run on HIP
// device 0 and are checked against the CPU oracle (oracle/_build/liboracle.so).
#undef NDEBUG
#include <atrip.hpp>
#include <atrip_synth.h>

#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../oracle/atrip_oracle.h"

#define TESTCASE(_name, ...)                                      \
  {                                                               \
    std::cout << "\x1b[35m-> \x1b[0m" << _name << std::endl;      \
    __VA_ARGS__                                                   \
  }

namespace {

struct System {
  int no, nv;
  std::vector<CTF::Tensor<double> *> t;  // ei ea Tai Tabij Vabij Vijka Vabci Jijka Jabci
  std::vector<std::vector<double>> host;
  System(int no_, int nv_, uint64_t seed, bool with_j) : no(no_), nv(nv_) {
    const std::vector<std::vector<int>> lens = {
        {no}, {nv}, {nv, no}, {nv, nv, no, no}, {nv, nv, no, no},
        {no, no, no, nv}, {nv, nv, nv, no}, {no, no, no, nv}, {nv, nv, nv, no}};
    for (int k = 0; k < (with_j ? 9 : 7); k++) {
      auto *x = new CTF::Tensor<double>((int)lens[k].size(), lens[k].data());
      std::vector<double> v((size_t)x->size());
      oracle_fill_synthetic(seed, k, no, nv, v.data(), v.size(), 4);
      x->write_all(v.data());
      t.push_back(x);
      host.push_back(v);
    }
    while (t.size() < 9) t.push_back(nullptr);
  }
  atrip::Atrip::Input<double> input() const {
    return atrip::Atrip::Input<double>()
        .with_epsilon_i(t[0]).with_epsilon_a(t[1]).with_Tai(t[2])
        .with_Tabij(t[3]).with_Vabij(t[4]).with_Vijka(t[5]).with_Vabci(t[6])
        .with_Jijka(t[7]).with_Jabci(t[8])
        .with_read_checkpoint_if_exists(false).with_writeCheckpoint(false);
  }
  oracle_input oin() const {
    oracle_input o{};
    o.No = no;
    o.Nv = nv;
    o.ei = host[0].data(); o.ea = host[1].data(); o.Tai = host[2].data();
    o.Tabij = host[3].data(); o.Vabij = host[4].data();
    o.Vijka = host[5].data(); o.Vabci = host[6].data();
    o.Jijka = host.size() > 7 ? host[7].data() : nullptr;
    o.Jabci = host.size() > 8 ? host[8].data() : nullptr;
    return o;
  }
  std::pair<double, double> oracle(uint64_t first, uint64_t count,
                                   int ijkabc = 0) const {
    oracle_input o = oin();
    double e = 0, ect = 0;
    assert(oracle_run(&o, ijkabc, first, count, 0, 4, 0, &e, &ect, nullptr,
                      nullptr) == 0);
    return {e, ect};
  }
};

// The same system with complex tensors: real parts as System, imaginary
// parts from the generator's ATRIP_SYN_IMAG streams.
struct ComplexSystem {
  int no, nv;
  std::vector<CTF::Tensor<atrip::Complex> *> t;
  std::vector<std::vector<atrip::Complex>> host;
  ComplexSystem(int no_, int nv_, uint64_t seed, bool with_j) : no(no_), nv(nv_) {
    const std::vector<std::vector<int>> lens = {
        {no}, {nv}, {nv, no}, {nv, nv, no, no}, {nv, nv, no, no},
        {no, no, no, nv}, {nv, nv, nv, no}, {no, no, no, nv}, {nv, nv, nv, no}};
    for (int k = 0; k < (with_j ? 9 : 7); k++) {
      auto *x = new CTF::Tensor<atrip::Complex>((int)lens[k].size(), lens[k].data());
      const size_t n = (size_t)x->size();
      std::vector<double> re(n), im(n);
      oracle_fill_synthetic(seed, k, no, nv, re.data(), n, 4);
      oracle_fill_synthetic(seed, k | ATRIP_SYN_IMAG, no, nv, im.data(), n, 4);
      std::vector<atrip::Complex> v(n);
      for (size_t q = 0; q < n; q++) v[q] = atrip::Complex(re[q], im[q]);
      x->write_all(v.data());
      t.push_back(x);
      host.push_back(v);
    }
    while (t.size() < 9) t.push_back(nullptr);
  }
  atrip::Atrip::Input<atrip::Complex> input() const {
    return atrip::Atrip::Input<atrip::Complex>()
        .with_epsilon_i(t[0]).with_epsilon_a(t[1]).with_Tai(t[2])
        .with_Tabij(t[3]).with_Vabij(t[4]).with_Vijka(t[5]).with_Vabci(t[6])
        .with_Jijka(t[7]).with_Jabci(t[8])
        .with_read_checkpoint_if_exists(false).with_writeCheckpoint(false);
  }
  std::pair<double, double> oracle() const {
    oracle_input_z o{};
    o.No = no;
    o.Nv = nv;
    auto p = [&](size_t k) {
      return k < host.size() ? reinterpret_cast<const double *>(host[k].data())
                             : nullptr;
    };
    o.ei = p(0); o.ea = p(1); o.Tai = p(2); o.Tabij = p(3); o.Vabij = p(4);
    o.Vijka = p(5); o.Vabci = p(6); o.Jijka = p(7); o.Jabci = p(8);
    double e = 0, ect = 0;
    assert(oracle_run_z(&o, 0, 0, oracle_n_tuples(nv), 4, &e, &ect, nullptr,
                        nullptr) == 0);
    return {e, ect};
  }
};

bool close(double a, double b, double rel = 1e-10) {
  return std::fabs(a - b) <= rel * std::fabs(b);
}

}  // namespace

int main(int argc, char **argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;

#define _CMP_CHECK(what)                                   \
  do {                                                     \
    std::cout << "\t Checking " << #what << std::endl;     \
    assert(in.what == what);                               \
    assert(out.what == what);                              \
  } while (0)

  TESTCASE("Testing checkpoint reader and writers",
           const std::string out_checkpoint = "/tmp/atrip-test-checkpoint.yaml";
           const double energy = -1.493926352289995443;
           const size_t no = 154, nv = 1500, nranks = 48 * 10, nnodes = 10;
           const size_t iteration = 546;
           for (bool rank_round_robin : {true, false}) {
             atrip::Checkpoint out = {no,     nv,        nranks, nnodes,
                                      energy, iteration, rank_round_robin},
                               in;
             atrip::write_checkpoint(out, out_checkpoint);
             in = atrip::read_checkpoint(out_checkpoint);
             _CMP_CHECK(no);
             _CMP_CHECK(nv);
             _CMP_CHECK(nranks);
             _CMP_CHECK(nnodes);
             _CMP_CHECK(iteration);
             _CMP_CHECK(rank_round_robin);
             _CMP_CHECK(energy);
           })
#undef _CMP_CHECK

  TESTCASE("Input builder keeps the reference defaults",
           atrip::Atrip::Input<double> in;
           assert(in.max_iterations == 0);
           assert(in.checkpoint_path == "atrip-checkpoint.yaml");
           assert(in.read_checkpoint_if_exists);
           assert(in.checkpoint_at_percentage == 10);
           assert(!in.ijkabc);
           assert(in.with_max_iterations(7).max_iterations == 7);
           // additions of this build: FP64 unless asked, bf16x3 when mixed
           assert(!in.mixed_precision);
           assert(in.mixed_form == "bf16x3");)

  if (!gpu) {
    std::cout << "host cases passed\n";
    return 0;
  }
  atrip::Atrip::init(0);

  TESTCASE("run<double> matches the oracle (No=6, Nv=13)",
           System s(6, 13, 21, false);
           const uint64_t n = oracle_n_tuples(13);
           auto ref = s.oracle(0, n);
           auto out = atrip::Atrip::run(s.input());
           assert(close(out.energy, ref.first));
           assert(out.ct_energy == out.energy);)

  TESTCASE("large tensors are read in slabs through a bounded buffer",
           // 4 KiB of staging: 256 doubles per pinned half, so every large
           // tensor arrives in many ranged reads of at most 256 values and
           // the energies are those of the whole-tensor upload
           System s(6, 13, 21, true);
           const uint64_t n = oracle_n_tuples(13);
           auto ref = s.oracle(0, n);
           auto out = atrip::Atrip::run(s.input().with_stage_kb(4));
           assert(close(out.energy, ref.first));
           assert(close(out.ct_energy, ref.second));
           for (int k = 3; k < 9; k++) {
             assert(s.t[k]->range_reads > 1);
             assert(s.t[k]->max_range_read <= 256);
           }
           auto whole = atrip::Atrip::run(s.input());
           assert(whole.energy == out.energy && whole.ct_energy == out.ct_energy);)

  TESTCASE("complex tensors are read in slabs too",
           ComplexSystem s(4, 9, 29, true);
           auto ref = s.oracle();
           auto out = atrip::Atrip::run(s.input().with_stage_kb(4));
           assert(close(out.energy, ref.first));
           assert(close(out.ct_energy, ref.second));
           for (int k = 3; k < 9; k++) {
             assert(s.t[k]->range_reads > 1);
             assert(s.t[k]->max_range_read <= 128);
           })

  TESTCASE("max_iterations stops after tuple max_iterations",
           System s(5, 11, 22, false);
           auto ref = s.oracle(0, 101);  // tuples 0..100
           auto out = atrip::Atrip::run(s.input().with_max_iterations(100));
           assert(close(out.energy, ref.first));)

  TESTCASE("checkpoint resume continues from Iteration with Energy",
           System s(5, 11, 23, false);
           const uint64_t n = oracle_n_tuples(11);
           auto full = s.oracle(0, n);
           auto head = s.oracle(0, 57);
           atrip::Checkpoint c{5, 11, 1, 1, head.first, 57, false};
           const std::string path = "/tmp/atrip-test-resume.yaml";
           atrip::write_checkpoint(c, path);
           auto out = atrip::Atrip::run(s.input()
                                            .with_checkpoint_path(path)
                                            .with_read_checkpoint_if_exists(true));
           assert(close(out.energy, full.first));
           std::remove(path.c_str());)

  TESTCASE("interval-0 checkpoints are written and resumable",
           System s(4, 9, 24, false);
           const uint64_t n = oracle_n_tuples(9);
           auto full = s.oracle(0, n);
           const std::string path = "/tmp/atrip-test-write.yaml";
           std::remove(path.c_str());
           auto out = atrip::Atrip::run(s.input()
                                            .with_checkpoint_path(path)
                                            .with_checkpoint_at_percentage(0)
                                            .with_writeCheckpoint(true));
           assert(close(out.energy, full.first));
           atrip::Checkpoint c = atrip::read_checkpoint(path);
           assert(c.iteration > 0 && c.iteration < n);
           auto head = s.oracle(0, c.iteration);
           assert(close(c.energy, head.first, 1e-9));
           std::remove(path.c_str());)

  TESTCASE("ijkabc negates Tai and keeps the sign",
           System s(5, 10, 25, false);
           const uint64_t n = oracle_n_tuples(10);
           auto ref = s.oracle(0, n, 1);
           auto out = atrip::Atrip::run(s.input().with_ijkabc(true));
           assert(close(out.energy, ref.first));)

  TESTCASE("(cT) with J tensors",
           System s(4, 9, 26, true);
           const uint64_t n = oracle_n_tuples(9);
           auto ref = s.oracle(0, n);
           auto out = atrip::Atrip::run(s.input());
           assert(close(out.energy, ref.first));
           assert(close(out.ct_energy, ref.second));)

  TESTCASE("run<Complex> matches the complex oracle, with (cT)",
           ComplexSystem s(5, 9, 27, true);
           auto ref = s.oracle();
           auto out = atrip::Atrip::run(s.input());
           assert(close(out.energy, ref.first));
           assert(close(out.ct_energy, ref.second));)

  TESTCASE("errors are std::string",
           bool thrown = false;
           try {
             atrip::Atrip::Input<double> in;
             atrip::Atrip::run(in);
           } catch (const std::string &e) {
             thrown = true;
           }
           assert(thrown);
           thrown = false;
           try {
             atrip::Atrip::Input<atrip::Complex> in;
             atrip::Atrip::run(in);
           } catch (const std::string &e) {
             thrown = true;
           }
           assert(thrown);
           thrown = false;
           try {
             ComplexSystem s(3, 5, 28, false);
             atrip::Atrip::run(s.input().with_mixed_precision(true));
           } catch (const std::string &e) {
             thrown = true;
           }
           assert(thrown);)

  std::cout << "all cases passed\n";
  return 0;
}
