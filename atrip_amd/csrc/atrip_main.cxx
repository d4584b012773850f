// `atrip` command-line driver: the counterpart of the reference's bench
// driver (bench/main.cxx:403-502) for this build.  It fills CTF-shaped
// tensors with the seeded synthetic generator (include/atrip_synth.h) and
// calls atrip::Atrip::run, printing the reference's Energy lines.
#include <atrip.hpp>
#include <atrip_synth.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

namespace {

void usage() {
  std::cerr
      << "usage: atrip --no N --nv N [--seed S] [--max-iterations K]\n"
         "             [--checkpoint-path P] [--nocheckpoint]\n"
         "             [--checkpoint-it K] [--checkpoint-% P] [--ijkabc]\n"
         "             [--cT] [--mixed] [--device D] [--chrono]\n";
}

CTF::Tensor<double> *make(int order, std::vector<int> lens, int tensor,
                          uint64_t seed, int no, int nv) {
  auto *t = new CTF::Tensor<double>(order, lens.data());
  const uint64_t s = atrip_synth_stream(seed, tensor);
  std::vector<double> v((size_t)t->size());
  for (size_t q = 0; q < v.size(); q++)
    v[q] = atrip_synth_value(s, tensor, (uint64_t)no, (uint64_t)nv, q);
  t->write_all(v.data());
  return t;
}

}  // namespace

int main(int argc, char **argv) {
  int no = 0, nv = 0, device = 0;
  uint64_t seed = 1;
  size_t max_it = 0, ck_it = 0;
  float ck_pct = 10;
  bool nockpt = false, ijkabc = false, ct = false, mixed = false,
       chrono = false;
  std::string ck_path = "atrip-checkpoint.yaml";
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto next = [&]() -> const char * {
      if (i + 1 >= argc) {
        usage();
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--no") no = std::atoi(next());
    else if (a == "--nv") nv = std::atoi(next());
    else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
    else if (a == "--max-iterations") max_it = std::strtoull(next(), nullptr, 10);
    else if (a == "--checkpoint-path") ck_path = next();
    else if (a == "--checkpoint-it") ck_it = std::strtoull(next(), nullptr, 10);
    else if (a == "--checkpoint-%") ck_pct = std::atof(next());
    else if (a == "--nocheckpoint") nockpt = true;
    else if (a == "--ijkabc") ijkabc = true;
    else if (a == "--cT") ct = true;
    else if (a == "--mixed") mixed = true;
    else if (a == "--device") device = std::atoi(next());
    else if (a == "--chrono") chrono = true;
    else {
      usage();
      return 2;
    }
  }
  if (no <= 0 || nv <= 0) {
    usage();
    return 2;
  }
  using atrip::Atrip;
  auto *ei = make(1, {no}, ATRIP_SYN_EI, seed, no, nv);
  auto *ea = make(1, {nv}, ATRIP_SYN_EA, seed, no, nv);
  auto *tai = make(2, {nv, no}, ATRIP_SYN_TAI, seed, no, nv);
  auto *tabij = make(4, {nv, nv, no, no}, ATRIP_SYN_TABIJ, seed, no, nv);
  auto *vabij = make(4, {nv, nv, no, no}, ATRIP_SYN_VABIJ, seed, no, nv);
  auto *vijka = make(4, {no, no, no, nv}, ATRIP_SYN_VIJKA, seed, no, nv);
  auto *vabci = make(4, {nv, nv, nv, no}, ATRIP_SYN_VABCI, seed, no, nv);
  CTF::Tensor<double> *jijka = nullptr, *jabci = nullptr;
  if (ct) {
    jijka = make(4, {no, no, no, nv}, ATRIP_SYN_JIJKA, seed, no, nv);
    jabci = make(4, {nv, nv, nv, no}, ATRIP_SYN_JABCI, seed, no, nv);
  }
  Atrip::init(0);
  auto in = Atrip::Input<double>()
                .with_epsilon_i(ei)
                .with_epsilon_a(ea)
                .with_Tai(tai)
                .with_Tabij(tabij)
                .with_Vabij(vabij)
                .with_Vijka(vijka)
                .with_Vabci(vabci)
                .with_Jijka(jijka)
                .with_Jabci(jabci)
                .with_max_iterations(max_it)
                .with_checkpoint_path(ck_path)
                .with_checkpoint_at_every_iteration(ck_it)
                .with_checkpoint_at_percentage(ck_pct)
                .with_read_checkpoint_if_exists(!nockpt)
                .with_writeCheckpoint(!nockpt)
                .with_ijkabc(ijkabc)
                .with_chrono(chrono)
                .with_device(device)
                .with_mixed_precision(mixed)
                .with_delete_Vppph(true);
  try {
    const auto out = Atrip::run(in);
    std::printf("Energy: %.15g\nEnergy (cT): %.15g\n", out.energy,
                out.ct_energy);
  } catch (const std::string &e) {
    std::cerr << "ERROR: " << e << "\n";
    return 1;
  } catch (const char *e) {
    std::cerr << "ERROR: " << e << "\n";
    return 1;
  }
  return 0;
}
