// Minimal stand-in for the part of Cyclops CTF that the (T) interface uses,
// for builds without Cyclops.  With Cyclops, define ATRIP_USE_CYCLOPS and the
// real <ctf.hpp> is used instead.  atrip::Atrip::run touches `order`,
// `lens` (as in the reference, src/atrip/Atrip.cxx:72-73), `read_all` for
// the small tensors (:179-181) and ranged reads of the large ones: here
// `read_range`, with Cyclops its Tensor::read(npair, global_idx, data).
//
// Storage is a dense, column-major (first index fastest) host array: the
// layout CTF's read_all returns.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

namespace CTF {

// Placeholder for CTF::World: single process, no communicator.
struct World {
  World() = default;
  World(int, char **) {}
  int rank = 0, np = 1;
};

#ifndef NS
#define NS 0  // "no symmetry" tag of CTF constructors
#endif

template <typename F = double>
class Tensor {
 public:
  int order = 0;
  int *lens = nullptr;
  F *data = nullptr;

  Tensor() = default;
  Tensor(int order_, const int *lens_, const int * /*syms*/ = nullptr,
         World & /*world*/ = default_world())
      : order(order_), lens_store_(lens_, lens_ + order_) {
    lens = lens_store_.data();
    values_.assign(size(), F{});
    data = values_.data();
  }
  Tensor(const Tensor &o)
      : order(o.order), lens_store_(o.lens_store_), values_(o.values_) {
    lens = lens_store_.data();
    data = values_.data();
  }
  Tensor &operator=(const Tensor &o) {
    order = o.order;
    lens_store_ = o.lens_store_;
    values_ = o.values_;
    lens = lens_store_.data();
    data = values_.data();
    return *this;
  }

  int64_t size() const {
    int64_t n = 1;
    for (int d = 0; d < order; d++) n *= lens_store_[d];
    return n;
  }

  // Every element, column-major, into a caller buffer of size() elements.
  void read_all(F *out, bool /*unpack*/ = true) const {
    std::memcpy(out, values_.data(), sizeof(F) * values_.size());
  }
  // Elements with global (column-major) indices global_idx[0..npair) into
  // data, the form of CTF's Tensor::read(npair, global_idx, data).
  void read(int64_t npair, const int64_t *global_idx, F *data) const {
    for (int64_t q = 0; q < npair; q++) data[q] = values_[global_idx[q]];
  }
  // Elements [first, first + count) in column-major order: the ranged read
  // the slab-wise upload asks for (a contiguous global index range).
  void read_range(int64_t first, int64_t count, F *out) const {
    if (first < 0 || count < 0 || first + count > size())
      throw std::string("read_range out of bounds");
    std::memcpy(out, values_.data() + first, sizeof(F) * count);
    ++range_reads;
    if (count > max_range_read) max_range_read = count;
  }
  // Instrumentation of read_range: calls and the largest count asked for.
  mutable int64_t range_reads = 0, max_range_read = 0;
  // Fill from a column-major buffer (counterpart of CTF's write).
  void write_all(const F *in) {
    std::memcpy(values_.data(), in, sizeof(F) * values_.size());
  }

  // Uniform fill in [a, b), as CTF's fill_random (bench/main.cxx:85 of the
  // reference); seeded so a run is reproducible.
  void fill_random(F a, F b, uint64_t seed = 0) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (auto &v : values_) v = a + (b - a) * F(u(g));
  }

  // Dense binary file of size() values, the layout CTF's
  // read_dense_from_file expects (column-major doubles).
  void read_dense_from_file(const char *path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::string("cannot open ") + path;
    in.read(reinterpret_cast<char *>(values_.data()),
            sizeof(F) * values_.size());
    if (!in) throw std::string("short read from ") + path;
  }

 private:
  static World &default_world() {
    static World w;
    return w;
  }
  std::vector<int> lens_store_;
  std::vector<F> values_;
};

}  // namespace CTF
