// `atrip` command-line driver: the counterpart of the reference's bench
// driver (bench/main.cxx:403-502) for this build.  Each tensor is read from a
// dense binary file when its path option names one, else filled by the seeded
// synthetic generator (include/atrip_synth.h); then atrip::Atrip::run runs
// and the reference's Energy lines are printed.
#include <atrip.hpp>
#include <atrip_synth.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

void usage() {
  std::cerr
      << "usage: atrip --no N --nv N [--dist group|naive] [--seed S]\n"
         "             [--max-iterations K] [--mod K] [-% P]\n"
         "             [--checkpoint-path P] [--nocheckpoint]\n"
         "             [--checkpoint-it K] [--checkpoint-% P] [--ijkabc]\n"
         "             [--cT] [--mixed] [--mixed-form bf16x3|split_fp32|fp32]\n"
         "             [--complex] [--device D] [--nochrono] [--keep-vppph]\n"
         "             [--rank-round-robin] [--barrier] [--blocking]\n"
         "             [--ei P] [--ea P] [--Tph P] [--Tpphh P] [--Vpphh P]\n"
         "             [--Vhhhp P] [--Vppph P] [--Jppph P] [--Jhhhp P]\n";
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

bool file_exists(const std::string &path) {
  return std::ifstream(path).good();
}

// out[p + h*q] = in[q + g*p] for p < h, q < g: the index halves of a tensor
// swapped, as the driver's ijkabc assignments Tph["ia"] = iTph["ai"] and
// Tpphh["ijab"] = iTpphh["abij"] (bench/main.cxx:300-306) do.
template <typename F>
void swap_halves(const CTF::Tensor<F> &in, CTF::Tensor<F> &out, int64_t h,
                 int64_t g) {
  std::vector<F> a((size_t)in.size()), b(a.size());
  in.read_all(a.data());
  for (int64_t q = 0; q < g; q++)
    for (int64_t p = 0; p < h; p++) b[(size_t)(p + h * q)] = a[(size_t)(q + g * p)];
  out.write_all(b.data());
}

// The reference driver's read_or_fill (bench/main.cxx:53-89): a tensor of
// `lens` read from the dense column-major file `path` when it exists;
// otherwise, with a warning if a path was given, the seeded synthetic fill.
// With `flip` (ijkabc amplitudes) the file holds the tensor with its index
// halves swapped (lens {no,nv} / {no,no,nv,nv}) and is transposed back.
template <typename F>
CTF::Tensor<F> *read_or_make(const char *name, int order, std::vector<int> lens,
                             int tensor, uint64_t seed, int no, int nv,
                             const std::string &path, bool flip = false) {
  if (path.empty() || !file_exists(path)) {
    if (!path.empty())
      std::cout << "WARNING: file " << path << " provided but not found!\n"
                << "Synthetic initialization for tensor " << name << "\n";
    return make<F>(order, lens, tensor, seed, no, nv);
  }
  auto *t = new CTF::Tensor<F>(order, lens.data());
  if (!flip) {
    t->read_dense_from_file(path.c_str());
    return t;
  }
  const int half = order / 2;
  std::vector<int> fl(lens.begin() + half, lens.end());
  fl.insert(fl.end(), lens.begin(), lens.begin() + half);
  CTF::Tensor<F> f(order, fl.data());
  f.read_dense_from_file(path.c_str());
  int64_t h = 1, g = 1;
  for (int d = 0; d < half; d++) h *= lens[d];
  for (int d = half; d < order; d++) g *= lens[d];
  swap_halves(f, *t, h, g);
  return t;
}

// Orbital energies: the driver reads them as real tensors and converts them
// to the field (bench/main.cxx:207-234).
template <typename F>
CTF::Tensor<F> *read_or_make_eps(const char *name, int n, int tensor,
                                 uint64_t seed, int no, int nv,
                                 const std::string &path) {
  if (path.empty() || !file_exists(path))
    return read_or_make<F>(name, 1, {n}, tensor, seed, no, nv, path);
  CTF::Tensor<double> r(1, &n);
  r.read_dense_from_file(path.c_str());
  std::vector<double> v((size_t)n);
  r.read_all(v.data());
  std::vector<F> w(v.begin(), v.end());
  auto *t = new CTF::Tensor<F>(1, &n);
  t->write_all(w.data());
  return t;
}

// Jabci / Jijka: the driver reads these from their files with no fallback
// (bench/main.cxx:315-343); with no path given they are synthetic here.
template <typename F>
CTF::Tensor<F> *read_or_make_j(int order, std::vector<int> lens, int tensor,
                               uint64_t seed, int no, int nv,
                               const std::string &path) {
  if (path.empty()) return make<F>(order, lens, tensor, seed, no, nv);
  auto *t = new CTF::Tensor<F>(order, lens.data());
  t->read_dense_from_file(path.c_str());
  return t;
}

// Defaults of the reference driver's options (bench/main.cxx:410-470).
struct Options {
  int no = 0, nv = 0, device = 0, it_mod = -1, pct_mod = 10;
  uint64_t seed = 1;
  size_t max_it = 0, ck_it = 0;
  float ck_pct = 0;
  bool nockpt = false, ijkabc = false, ct = false, mixed = false,
       chrono = true, keep_vppph = false, rank_round_robin = false,
       barrier = false, blocking = false;
  bool dist_naive = false;  // --dist naive, else group (GROUP_AND_SORT)
  std::string ck_path = "checkpoint.yaml";
  std::string mixed_form = "bf16x3";
  std::string ei, ea, tph, tpphh, vpphh, vhhhp, vppph, jppph, jhhhp;
};

template <typename F>
int run(const Options &o) try {
  using atrip::Atrip;
  const int no = o.no, nv = o.nv;
  const uint64_t seed = o.seed;
  const bool fl = o.ijkabc;
  auto *ei = read_or_make_eps<F>("ei", no, ATRIP_SYN_EI, seed, no, nv, o.ei);
  auto *ea = read_or_make_eps<F>("ea", nv, ATRIP_SYN_EA, seed, no, nv, o.ea);
  auto *tai = read_or_make<F>("tph", 2, {nv, no}, ATRIP_SYN_TAI, seed, no, nv,
                              o.tph, fl);
  auto *tabij = read_or_make<F>("tpphh", 4, {nv, nv, no, no}, ATRIP_SYN_TABIJ,
                                seed, no, nv, o.tpphh, fl);
  auto *vabij = read_or_make<F>("Vpphh", 4, {nv, nv, no, no}, ATRIP_SYN_VABIJ,
                                seed, no, nv, o.vpphh, fl);
  auto *vijka = read_or_make<F>("Vhhhp", 4, {no, no, no, nv}, ATRIP_SYN_VIJKA,
                                seed, no, nv, o.vhhhp);
  auto *vabci = read_or_make<F>("Vppph", 4, {nv, nv, nv, no}, ATRIP_SYN_VABCI,
                                seed, no, nv, o.vppph);
  CTF::Tensor<F> *jijka = nullptr, *jabci = nullptr;
  if (o.ct || (!o.jppph.empty() && !o.jhhhp.empty())) {
    jijka = read_or_make_j<F>(4, {no, no, no, nv}, ATRIP_SYN_JIJKA, seed, no,
                              nv, o.jhhhp);
    jabci = read_or_make_j<F>(4, {nv, nv, nv, no}, ATRIP_SYN_JABCI, seed, no,
                              nv, o.jppph);
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
                .with_iteration_mod(o.it_mod)
                .with_percentage_mod(o.pct_mod)
                .with_tuples_distribution(
                    o.dist_naive ? Atrip::Input<F>::NAIVE
                                 : Atrip::Input<F>::GROUP_AND_SORT)
                .with_rank_round_robin(o.rank_round_robin)
                .with_barrier(o.barrier)
                .with_blocking(o.blocking)
                .with_device(o.device)
                .with_mixed_precision(o.mixed)
                .with_mixed_form(o.mixed_form)
                .with_delete_Vppph(!o.keep_vppph);
  const auto out = Atrip::run(in);
  std::printf("Energy: %.15g\nEnergy (cT): %.15g\n", out.energy,
              out.ct_energy);
  return 0;
} catch (const std::string &e) {
  std::cerr << "ERROR: " << e << "\n";
  return 1;
} catch (const char *e) {
  std::cerr << "ERROR: " << e << "\n";
  return 1;
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
    else if (a == "--nochrono") o.chrono = false;
    else if (a == "--keep-vppph") o.keep_vppph = true;
    else if (a == "--rank-round-robin") o.rank_round_robin = true;
    else if (a == "--barrier") o.barrier = true;
    else if (a == "--blocking") o.blocking = true;
    else if (a == "--mod") o.it_mod = std::atoi(next());
    else if (a == "-%") o.pct_mod = std::atoi(next());
    else if (a == "--dist") {
      const std::string d = next();
      if (d == "group") o.dist_naive = false;
      else if (d == "naive") o.dist_naive = true;
      else {
        std::cerr << "ERROR: --dist: '" << d << "' is not group or naive\n";
        return 2;
      }
    }
    else if (a == "--ei") o.ei = next();
    else if (a == "--ea") o.ea = next();
    else if (a == "--Tph") o.tph = next();
    else if (a == "--Tpphh") o.tpphh = next();
    else if (a == "--Vpphh") o.vpphh = next();
    else if (a == "--Vhhhp") o.vhhhp = next();
    else if (a == "--Vppph") o.vppph = next();
    else if (a == "--Jppph") o.jppph = next();
    else if (a == "--Jhhhp") o.jhhhp = next();
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
