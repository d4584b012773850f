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
         "             [--cT] [--mixed] [--mixed-form split_fp32|fp32|bf16x3]\n"
         "             [--complex] [--device D] [--chrono]\n";
}

template <typename F>
void fill(F &x, uint64_t sr, uint64_t si, int tensor, int no, int nv,
          uint64_t q);
template <>
void fill<double>(double &x, uint64_t sr, uint64_t, int tensor, int no,
                  int nv, uint64_t q) {
  x = atrip_synth_value(sr, tensor, (uint64_t)no, (uint64_t)nv, q);
}
// complex: imaginary part from the ATRIP_SYN_IMAG stream of the tensor
template <>
void fill<atrip::Complex>(atrip::Complex &x, uint64_t sr, uint64_t si,
                          int tensor, int no, int nv, uint64_t q) {
  x = atrip::Complex(
      atrip_synth_value(sr, tensor, (uint64_t)no, (uint64_t)nv, q),
      atrip_synth_value(si, tensor | ATRIP_SYN_IMAG, (uint64_t)no,
                        (uint64_t)nv, q));
}

template <typename F>
CTF::Tensor<F> *make(int order, std::vector<int> lens, int tensor,
                     uint64_t seed, int no, int nv) {
  auto *t = new CTF::Tensor<F>(order, lens.data());
  const uint64_t sr = atrip_synth_stream(seed, tensor);
  const uint64_t si = atrip_synth_stream(seed, tensor | ATRIP_SYN_IMAG);
  std::vector<F> v((size_t)t->size());
  for (size_t q = 0; q < v.size(); q++) fill<F>(v[q], sr, si, tensor, no, nv, q);
  t->write_all(v.data());
  return t;
}

struct Options {
  int no = 0, nv = 0, device = 0;
  uint64_t seed = 1;
  size_t max_it = 0, ck_it = 0;
  float ck_pct = 10;
  bool nockpt = false, ijkabc = false, ct = false, mixed = false,
       chrono = false;
  std::string ck_path = "atrip-checkpoint.yaml";
  std::string mixed_form = "split_fp32";
};

template <typename F>
int run(const Options &o) {
  using atrip::Atrip;
  const int no = o.no, nv = o.nv;
  const uint64_t seed = o.seed;
  auto *ei = make<F>(1, {no}, ATRIP_SYN_EI, seed, no, nv);
  auto *ea = make<F>(1, {nv}, ATRIP_SYN_EA, seed, no, nv);
  auto *tai = make<F>(2, {nv, no}, ATRIP_SYN_TAI, seed, no, nv);
  auto *tabij = make<F>(4, {nv, nv, no, no}, ATRIP_SYN_TABIJ, seed, no, nv);
  auto *vabij = make<F>(4, {nv, nv, no, no}, ATRIP_SYN_VABIJ, seed, no, nv);
  auto *vijka = make<F>(4, {no, no, no, nv}, ATRIP_SYN_VIJKA, seed, no, nv);
  auto *vabci = make<F>(4, {nv, nv, nv, no}, ATRIP_SYN_VABCI, seed, no, nv);
  CTF::Tensor<F> *jijka = nullptr, *jabci = nullptr;
  if (o.ct) {
    jijka = make<F>(4, {no, no, no, nv}, ATRIP_SYN_JIJKA, seed, no, nv);
    jabci = make<F>(4, {nv, nv, nv, no}, ATRIP_SYN_JABCI, seed, no, nv);
  }
  Atrip::init(0);
  auto in = typename Atrip::Input<F>()
                .with_epsilon_i(ei)
                .with_epsilon_a(ea)
                .with_Tai(tai)
                .with_Tabij(tabij)
                .with_Vabij(vabij)
                .with_Vijka(vijka)
                .with_Vabci(vabci)
                .with_Jijka(jijka)
                .with_Jabci(jabci)
                .with_max_iterations(o.max_it)
                .with_checkpoint_path(o.ck_path)
                .with_checkpoint_at_every_iteration(o.ck_it)
                .with_checkpoint_at_percentage(o.ck_pct)
                .with_read_checkpoint_if_exists(!o.nockpt)
                .with_writeCheckpoint(!o.nockpt)
                .with_ijkabc(o.ijkabc)
                .with_chrono(o.chrono)
                .with_device(o.device)
                .with_mixed_precision(o.mixed)
                .with_mixed_form(o.mixed_form)
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

}  // namespace

int main(int argc, char **argv) {
  Options o;
  bool cplx = false;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto next = [&]() -> const char * {
      if (i + 1 >= argc) {
        usage();
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--no") o.no = std::atoi(next());
    else if (a == "--nv") o.nv = std::atoi(next());
    else if (a == "--seed") o.seed = std::strtoull(next(), nullptr, 10);
    else if (a == "--max-iterations") o.max_it = std::strtoull(next(), nullptr, 10);
    else if (a == "--checkpoint-path") o.ck_path = next();
    else if (a == "--checkpoint-it") o.ck_it = std::strtoull(next(), nullptr, 10);
    else if (a == "--checkpoint-%") o.ck_pct = std::atof(next());
    else if (a == "--nocheckpoint") o.nockpt = true;
    else if (a == "--ijkabc") o.ijkabc = true;
    else if (a == "--cT") o.ct = true;
    else if (a == "--mixed") o.mixed = true;
    else if (a == "--mixed-form") {
      o.mixed = true;
      o.mixed_form = next();
    }
    else if (a == "--complex") cplx = true;
    else if (a == "--device") o.device = std::atoi(next());
    else if (a == "--chrono") o.chrono = true;
    else {
      usage();
      return 2;
    }
  }
  if (o.no <= 0 || o.nv <= 0) {
    usage();
    return 2;
  }
  return cplx ? run<atrip::Complex>(o) : run<double>(o);
}
