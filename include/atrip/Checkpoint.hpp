// Checkpoint file of a (T) run, in the reference's format
// (src/atrip/Checkpoint.hpp:27-79): one "Key: value" line per field, the
// energy printed with 19 significant digits, booleans as true/false.  The
// stored energy is the physical one (sign applied), Iteration is the number
// of tuples already done, i.e. the index of the next tuple.
#pragma once

#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <string>

namespace atrip {

struct Checkpoint {
  size_t no = 0, nv = 0;
  size_t nranks = 0;
  size_t nnodes = 0;
  double energy = 0.0;
  size_t iteration = 0;
  bool rank_round_robin = false;
};

inline void write_checkpoint(Checkpoint const &c, std::string const &path) {
  std::ofstream out(path);
  if (!out) throw std::string("cannot write checkpoint ") + path;
  out << "No: " << c.no << "\n"
      << "Nv: " << c.nv << "\n"
      << "Nranks: " << c.nranks << "\n"
      << "Nnodes: " << c.nnodes << "\n"
      << "Energy: " << std::setprecision(19) << c.energy << "\n"
      << "Iteration: " << c.iteration << "\n"
      << "RankRoundRobin: " << (c.rank_round_robin ? "true" : "false")
      << "\n";
}

inline Checkpoint read_checkpoint(std::istream &in) {
  Checkpoint c;
  auto strip = [](std::string s) {
    const char *ws = " \t\r";
    const size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return std::string();
    return s.substr(b, s.find_last_not_of(ws) - b + 1);
  };
  std::string line;
  while (std::getline(in, line)) {
    const size_t colon = line.find(':');
    if (colon == std::string::npos) continue;
    const std::string key = strip(line.substr(0, colon));
    const std::string val = strip(line.substr(colon + 1));
    if (key == "No") c.no = std::strtoull(val.c_str(), nullptr, 10);
    else if (key == "Nv") c.nv = std::strtoull(val.c_str(), nullptr, 10);
    else if (key == "Nranks") c.nranks = std::strtoull(val.c_str(), nullptr, 10);
    else if (key == "Nnodes") c.nnodes = std::strtoull(val.c_str(), nullptr, 10);
    else if (key == "Energy") c.energy = std::strtod(val.c_str(), nullptr);
    else if (key == "Iteration")
      c.iteration = std::strtoull(val.c_str(), nullptr, 10);
    else if (key == "RankRoundRobin") c.rank_round_robin = !val.empty() && val[0] == 't';
  }
  return c;
}

inline Checkpoint read_checkpoint(std::string const &path) {
  std::ifstream in(path);
  if (!in) throw std::string("cannot read checkpoint ") + path;
  return read_checkpoint(in);
}

}  // namespace atrip
