// C-ABI of the MI355X (T) hot path: contexts, resident systems, the tuple
// window scheduler and the run loop.  See include/atrip_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "doubles_common.hpp"
#include "host.hpp"
#include "../../include/atrip_synth.h"

using namespace atrip_amd;

DevSystem &atrip_amd::get_sys(atrip_hip_ctx *ctx, int sys) {
  auto it = ctx->systems.find(sys);
  if (it == ctx->systems.end()) throw Fail{"unknown system handle"};
  return it->second;
}

namespace {
void *dev_bytes(DevSystem &s, size_t bytes, const char *what);
// Zeroed tiled bf16x3 copy (and row-swapped copy) of geometry g.
void alloc_tri(DevSystem &s, TriOp &o, const TriGeom &g, const char *what) {
  const size_t bytes = (size_t)tri_parts(g) * 2;
  o.p = (uint16_t *)dev_bytes(s, bytes, what);
  check(hipMemset(o.p, 0, bytes), what);
  if (g.swapped) {
    o.sw = (uint16_t *)dev_bytes(s, bytes, what);
    check(hipMemset(o.sw, 0, bytes), what);
  }
}
}

// Split-FP32 copies of the GEMM operands, made once per system on the first
// mixed-precision run.
void atrip_amd::ensure_split(atrip_hip_ctx *ctx, DevSystem &s) {
  if (s.split_ready) return;
  const int64_t No = s.No, Nv = s.Nv, N2 = No * No, Nv2 = Nv * Nv;
  auto make = [&](const double *x, DevSystem::Split &o, int64_t n) {
    if (!x) return;
    o.h = (float *)dev_bytes(s, (size_t)n * sizeof(float), "alloc split");
    o.l = (uint16_t *)dev_bytes(s, (size_t)n * 2, "alloc split");
    check(launch_split(x, o, n, ctx->stream), "split conversion");
  };
  make(s.T3, s.sT3, Nv * N2 * s.Kv);
  make(s.V3, s.sV3, Nv2 * No * s.Kv);
  make(s.H3, s.sH3, Nv * N2 * s.Ko);
  make(s.T2N, s.sT2N, Nv2 * No * s.Ko);
  make(s.T2TN, s.sT2TN, Nv2 * No * s.Ko);
  make(s.V3J, s.sV3J, Nv2 * No * s.Kv);
  make(s.H3J, s.sH3J, Nv * N2 * s.Ko);
  check(hipStreamSynchronize(ctx->stream), "split conversion");
  s.split_ready = true;
}

void atrip_amd::check_precision(const DevSystem &s, int precision) {
  if (precision != ATRIP_FP64 && precision != ATRIP_SPLIT_FP32 &&
      precision != ATRIP_FP32 && precision != ATRIP_BF16X3)
    throw Fail{"unknown precision " + std::to_string(precision)};
  if (s.cplx && precision != ATRIP_FP64)
    throw Fail{"complex systems run in FP64 only"};
  if (precision == ATRIP_FP64 && s.store != 0)
    throw Fail{"system was uploaded without FP64 operands (operand_store)"};
  if (precision == ATRIP_SPLIT_FP32 && (s.store == 2 || s.store == 3))
    throw Fail{"system was uploaded without split operands (operand_store " +
               std::to_string(s.store) + ")"};
  if (precision == ATRIP_FP32 && s.store == 3)
    throw Fail{"system was uploaded with bf16x3 operands only (operand_store 3)"};
  if (precision == ATRIP_BF16X3 && s.store != 0 && s.store != 3)
    throw Fail{"system was uploaded without bf16x3 operands (operand_store " +
               std::to_string(s.store) + ")"};
}

// bf16x3 copies of the GEMM operands, made once per system on the first
// bf16x3 run.
void atrip_amd::ensure_tri(atrip_hip_ctx *ctx, DevSystem &s) {
  if (s.tri_ready) return;
  const int64_t No = s.No, Nv = s.Nv;
  const TriGeom ga = tri_geom_a(No, Nv, s.Kv), gb = tri_geom_b(No, Nv, s.Kv),
                gh = tri_geom_a(No, Nv, s.Ko), gt = tri_geom_b(No, Nv, s.Ko);
  auto make = [&](const double *x, TriOp &o, const TriGeom &g) {
    if (!x) return;
    alloc_tri(s, o, g, "alloc bf16x3");
    check(launch_tri(x, o, g, No, ctx->stream), "bf16x3 conversion");
  };
  make(s.T3, s.tT3, ga);
  make(s.V3, s.tV3, gb);
  make(s.H3, s.tH3, gh);
  make(s.T2N, s.tT2N, gt);
  make(s.T2TN, s.tT2TN, gt);
  make(s.V3J, s.tV3J, gb);
  make(s.H3J, s.tH3J, gh);
  check(hipStreamSynchronize(ctx->stream), "bf16x3 conversion");
  s.tri_ready = true;
}

namespace {

double *dev_alloc(DevSystem &s, int64_t n, const char *what) {
  void *p = nullptr;
  const size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(double);
  check(hipMalloc(&p, bytes), what);
  s.allocations.push_back(p);
  s.bytes += bytes;
  check(hipMemset(p, 0, bytes), what);
  return (double *)p;
}

void *dev_bytes(DevSystem &s, size_t bytes, const char *what) {
  void *p = nullptr;
  check(hipMalloc(&p, std::max<size_t>(bytes, 1)), what);
  s.allocations.push_back(p);
  s.bytes += bytes;
  return p;
}

void free_system(DevSystem &s) {
  for (void *p : s.allocations) (void)hipFree(p);
  s.allocations.clear();
}

}  // namespace

// Device bytes alloc_system takes for a system (the arrays below, the
// packed point table and the canonical-index prefix).
uint64_t atrip_amd::system_bytes(int64_t No, int64_t Nv, bool with_j,
                                 int store) {
  const uint64_t Kv = pad16(Nv), Ko = pad16(No), N2 = No * No, Nv2 = Nv * Nv;
  const auto d = [](uint64_t n) { return std::max<uint64_t>(n, 1) * 8; };
  // doubles operands: 8 bytes per element as FP64, 6 as split (4 + 2), 4
  // as FP32 only; bf16x3 copies by their tiled geometry
  const auto o = [&](uint64_t n) {
    return store == 2 ? n * 4 : store == 1 ? n * 6 : d(n);
  };
  uint64_t b = d(Nv2 * N2) + d(Nv * No) + d(No) + d(Nv);
  if (store == 3) {
    b += tri_bytes(tri_geom_a(No, Nv, Kv)) + tri_bytes(tri_geom_b(No, Nv, Kv)) +
         tri_bytes(tri_geom_a(No, Nv, Ko)) + 2 * tri_bytes(tri_geom_b(No, Nv, Ko));
    if (with_j)
      b += tri_bytes(tri_geom_b(No, Nv, Kv)) + tri_bytes(tri_geom_a(No, Nv, Ko));
  } else {
    b += o(Nv * N2 * Kv) + o(Nv2 * No * Kv) + o(Nv * N2 * Ko) +
         2 * o(Nv2 * No * Ko);
    if (with_j) b += o(Nv2 * No * Kv) + o(Nv * N2 * Ko);
  }
  b += (uint64_t)No * (No + 1) * (No + 2) / 6 * sizeof(uint32_t);
  b += (Nv + 1) * sizeof(uint64_t);
  return b;
}

namespace {

void alloc_system(atrip_hip_ctx *ctx, DevSystem &s, int64_t No, int64_t Nv,
                  bool with_j, bool cplx = false) {
  if (No < 1 || Nv < 1) throw Fail{"No and Nv must be positive"};
  if (cplx && ctx->operand_store != 0)
    throw Fail{"complex systems keep FP64 operands (operand_store 0)"};
  if (No > 1023) throw Fail{"No > 1023 is not supported"};
  s.No = No;
  s.Nv = Nv;
  s.Kv = pad16(Nv);
  s.Ko = pad16(No);
  const int64_t N2 = No * No, Nv2 = Nv * Nv;
  s.store = ctx->operand_store;
  // doubles operands: FP64, or only their split / FP32 / bf16x3 form
  auto operand = [&](double *&x, DevSystem::Split &sp, TriOp &tp,
                     const TriGeom &g, const char *what) {
    const int64_t n = g.no * g.nr * g.K;
    if (s.store == 0) {
      x = dev_alloc(s, n, what);
      return;
    }
    if (s.store == 3) {
      alloc_tri(s, tp, g, what);
      return;
    }
    sp.h = (float *)dev_bytes(s, (size_t)n * sizeof(float), what);
    if (s.store == 1) {
      sp.l = (uint16_t *)dev_bytes(s, (size_t)n * 2, what);
    }
  };
  const TriGeom ga = tri_geom_a(No, Nv, s.Kv), gb = tri_geom_b(No, Nv, s.Kv),
                gh = tri_geom_a(No, Nv, s.Ko), gt = tri_geom_b(No, Nv, s.Ko);
  operand(s.T3, s.sT3, s.tT3, ga, "alloc T3");
  operand(s.V3, s.sV3, s.tV3, gb, "alloc V3");
  operand(s.H3, s.sH3, s.tH3, gh, "alloc H3");
  operand(s.T2N, s.sT2N, s.tT2N, gt, "alloc T2N");
  operand(s.T2TN, s.sT2TN, s.tT2TN, gt, "alloc T2TN");
  s.split_ready = s.store == 1 || s.store == 2;
  s.tri_ready = s.store == 3;
  s.VHH = dev_alloc(s, Nv2 * N2, "alloc VHH");
  s.Tai = dev_alloc(s, Nv * No, "alloc Tai");
  s.ei = dev_alloc(s, No, "alloc ei");
  s.ea = dev_alloc(s, Nv, "alloc ea");
  s.has_j = with_j;
  if (with_j) {
    operand(s.V3J, s.sV3J, s.tV3J, gb, "alloc V3J");
    operand(s.H3J, s.sH3J, s.tH3J, gh, "alloc H3J");
  }
  s.cplx = cplx;
  if (cplx) {
    s.T3i = dev_alloc(s, Nv * N2 * s.Kv, "alloc T3 (im)");
    s.V3i = dev_alloc(s, Nv2 * No * s.Kv, "alloc V3 (im)");
    s.H3i = dev_alloc(s, Nv * N2 * s.Ko, "alloc H3 (im)");
    s.T2Ni = dev_alloc(s, Nv2 * No * s.Ko, "alloc T2N (im)");
    s.T2TNi = dev_alloc(s, Nv2 * No * s.Ko, "alloc T2TN (im)");
    s.VHHi = dev_alloc(s, Nv2 * N2, "alloc VHH (im)");
    s.Taii = dev_alloc(s, Nv * No, "alloc Tai (im)");
    s.eii = dev_alloc(s, No, "alloc ei (im)");
    s.eai = dev_alloc(s, Nv, "alloc ea (im)");
    s.T3in = dev_alloc(s, Nv * N2 * s.Kv, "alloc -T3 (im)");
    s.H3in = dev_alloc(s, Nv * N2 * s.Ko, "alloc -H3 (im)");
    if (with_j) {
      s.V3Ji = dev_alloc(s, Nv2 * No * s.Kv, "alloc V3J (im)");
      s.H3Ji = dev_alloc(s, Nv * N2 * s.Ko, "alloc H3J (im)");
      s.H3Jin = dev_alloc(s, Nv * N2 * s.Ko, "alloc -H3J (im)");
    }
  }
  // i >= j >= k points of the energy loops, packed 10 bits per index
  std::vector<uint32_t> pts;
  for (int64_t k = 0; k < No; k++)
    for (int64_t j = k; j < No; j++)
      for (int64_t i = j; i < No; i++)
        pts.push_back((uint32_t)(i | (j << 10) | (k << 20)));
  s.npoints = (int64_t)pts.size();
  {
    void *p = nullptr;
    check(hipMalloc(&p, pts.size() * sizeof(uint32_t)), "alloc points");
    s.allocations.push_back(p);
    s.bytes += pts.size() * sizeof(uint32_t);
    check(hipMemcpy(p, pts.data(), pts.size() * sizeof(uint32_t),
                    hipMemcpyHostToDevice),
          "upload points");
    s.points = (uint32_t *)p;
  }
  // tuples with first index < a, for the canonical index
  std::vector<uint64_t> pre(Nv + 1, 0);
  for (int64_t a = 0; a < Nv; a++) {
    const uint64_t m = (uint64_t)(Nv - a);
    pre[a + 1] = pre[a] + m * (m + 1) / 2 - 1;
  }
  {
    void *p = nullptr;
    check(hipMalloc(&p, pre.size() * sizeof(uint64_t)), "alloc prefix");
    s.allocations.push_back(p);
    s.bytes += pre.size() * sizeof(uint64_t);
    check(hipMemcpy(p, pre.data(), pre.size() * sizeof(uint64_t),
                    hipMemcpyHostToDevice),
          "upload prefix");
    s.aprefix = (uint64_t *)p;
  }
  (void)ctx;
}

// Where the values of a system's input tensors come from.  Values are in
// CTF linear (column-major) order; a value is `comp` doubles (1 real, 2 an
// interleaved (re, im) pair).  fetch() copies values [first, first + count)
// of tensor t (enum atrip_tensor) to device memory on stream st; host()
// reads them into host memory (the small tensors).
struct Source {
  int comp = 1;
  virtual ~Source() = default;
  virtual void fetch(int t, int64_t first, int64_t count, double *dst,
                     hipStream_t st) = 0;
  virtual void host(int t, int64_t first, int64_t count, double *out) = 0;
};

// Tensors in caller memory (the pointers of an atrip_sys_desc).
struct DescSource : Source {
  const double *p[9];
  DescSource(const atrip_sys_desc *d, int comp_) {
    comp = comp_;
    const double *q[9] = {d->ei,    d->ea,    d->Tai,   d->Tabij, d->Vabij,
                          d->Vijka, d->Vabci, d->Jijka, d->Jabci};
    std::copy(q, q + 9, p);
  }
  void fetch(int t, int64_t first, int64_t count, double *dst,
             hipStream_t st) override {
    check(hipMemcpyAsync(dst, p[t] + first * comp,
                         (size_t)(count * comp) * sizeof(double),
                         hipMemcpyHostToDevice, st),
          "upload chunk");
  }
  void host(int t, int64_t first, int64_t count, double *out) override {
    std::memcpy(out, p[t] + first * comp, (size_t)(count * comp) * sizeof(double));
  }
};

// Tensors read through the caller's reader, piece by piece into a pinned
// host buffer of bounded size (two halves: one is read into while the copy
// of the other is in flight), so the host holds no tensor whole.
struct ReaderSource : Source {
  atrip_read_fn fn;
  void *user;
  double *pinned = nullptr;
  int64_t half = 0;  // doubles per half
  hipEvent_t done[2] = {nullptr, nullptr};
  bool recorded[2] = {false, false};
  ReaderSource(atrip_read_fn f, void *u, int comp_, int64_t cap_bytes)
      : fn(f), user(u) {
    comp = comp_;
    half = std::max<int64_t>(comp, cap_bytes / (2 * (int64_t)sizeof(double)));
    half = half / comp * comp;
    try {
      check(hipHostMalloc((void **)&pinned,
                          (size_t)(2 * half) * sizeof(double),
                          hipHostMallocDefault),
            "alloc pinned staging");
      for (auto &e : done)
        check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    } catch (...) {
      release();  // no destructor runs for a constructor that throws
      throw;
    }
  }
  ~ReaderSource() override { release(); }
  void release() {
    for (int h = 0; h < 2; h++)
      if (recorded[h]) (void)hipEventSynchronize(done[h]);  // copies done
    for (auto &e : done)
      if (e) (void)hipEventDestroy(e), e = nullptr;
    if (pinned) (void)hipHostFree(pinned), pinned = nullptr;
  }
  void read(int t, int64_t first, int64_t count, double *out) {
    if (fn(user, t, first, count, out) != 0)
      throw Fail{"tensor reader failed (tensor " + std::to_string(t) +
                 ", values " + std::to_string(first) + "+" +
                 std::to_string(count) + ")"};
  }
  void fetch(int t, int64_t first, int64_t count, double *dst,
             hipStream_t st) override {
    const int64_t per = half / comp;  // values per piece
    int h = 0;
    for (int64_t off = 0; off < count; off += per, h ^= 1) {
      const int64_t n = std::min(per, count - off);
      double *buf = pinned + h * half;
      if (recorded[h]) check(hipEventSynchronize(done[h]), "pinned staging");
      read(t, first + off, n, buf);
      check(hipMemcpyAsync(dst + off * comp, buf,
                           (size_t)(n * comp) * sizeof(double),
                           hipMemcpyHostToDevice, st),
            "upload chunk");
      check(hipEventRecord(done[h], st), "event");
      recorded[h] = true;
    }
  }
  void host(int t, int64_t first, int64_t count, double *out) override {
    read(t, first, count, out);
  }
};

// Stage tensor t through a bounded device buffer (the context's stage_kb),
// whole slabs of `slab` values of its slowest CTF index per step (n_slow
// slabs; at least one slab per step), and hand each staged chunk to
// `place(dev_chunk, first_slow, count_slow)`.
template <typename P>
void staged(atrip_hip_ctx *ctx, Source &src, int t, int64_t slab,
            int64_t n_slow, P &&place) {
  const int64_t budget = std::max<int64_t>(
      1, ctx->stage_kb * 1024 / (int64_t)(sizeof(double) * src.comp));
  int64_t per = std::max<int64_t>(1, budget / std::max<int64_t>(slab, 1));
  per = std::min(per, n_slow);
  void *stage = nullptr;
  check(hipMalloc(&stage, (size_t)(per * slab * src.comp) * sizeof(double)),
        "alloc staging");
  try {
    for (int64_t s0 = 0; s0 < n_slow; s0 += per) {
      const int64_t cnt = std::min(per, n_slow - s0);
      src.fetch(t, s0 * slab, cnt * slab, (double *)stage, ctx->stream);
      place((const double *)stage, s0, cnt);
      check(hipStreamSynchronize(ctx->stream), "relayout");
    }
  } catch (...) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(stage);
    throw;
  }
  check(hipFree(stage), "free staging");
}

// A system that keeps only the split / FP32 / bf16x3 form of its doubles
// operands gets each FP64 operand in a temporary device array, converted
// and freed as soon as it is filled.
struct TempOperands {
  DevSystem &s;
  hipStream_t st;
  std::vector<std::pair<double **, DevSystem::Split *>> items;
  std::vector<std::pair<TriOp *, TriGeom>> tris;
  std::vector<int64_t> sizes;
  void add(double *&x, DevSystem::Split &sp, TriOp &tp, const TriGeom &g) {
    const int64_t n = g.no * g.nr * g.K;
    if (s.store == 0) return;
    void *p = nullptr;
    check(hipMalloc(&p, (size_t)std::max<int64_t>(n, 1) * sizeof(double)),
          "alloc temporary operand");
    check(hipMemsetAsync(p, 0, (size_t)std::max<int64_t>(n, 1) * sizeof(double),
                         st),
          "clear temporary operand");
    x = (double *)p;
    items.push_back({&x, &sp});
    tris.push_back({&tp, g});
    sizes.push_back(n);
  }
  void convert() {
    for (size_t q = 0; q < items.size(); q++) {
      if (s.store == 3)
        check(launch_tri(*items[q].first, *tris[q].first, tris[q].second,
                         s.No, st),
              "bf16x3 conversion");
      else
        check(launch_split(*items[q].first, *items[q].second, sizes[q], st),
              "split conversion");
      check(hipStreamSynchronize(st), "split conversion");
      (void)hipFree(*items[q].first);
      *items[q].first = nullptr;
    }
    items.clear();
    tris.clear();
    sizes.clear();
  }
  ~TempOperands() {
    for (auto &it : items) {
      (void)hipFree(*it.first);
      *it.first = nullptr;
    }
  }
};

// Re-lays out every input tensor of `src` into the device arrays of s (and
// of im, the imaginary planes, for a complex source: src.comp == 2).  Each
// large tensor is streamed in slabs of its slowest CTF index through the
// staging buffer; every transpose reads one part of the interleaved values
// (stride comp).
void upload_tensors(atrip_hip_ctx *ctx, DevSystem &s, DevSystem *im,
                    Source &src) {
  const int64_t No = s.No, Nv = s.Nv, N2 = No * No, Nv2 = Nv * Nv;
  const int64_t Kv = s.Kv, Ko = s.Ko;
  const int m = src.comp;
  if ((m == 2) != (im != nullptr)) throw Fail{"internal: complex parts"};
  DevSystem *dst[2] = {&s, im};
  hipStream_t st = ctx->stream;
  TempOperands tmp{s, st, {}, {}};
  tmp.add(s.T3, s.sT3, s.tT3, tri_geom_a(No, Nv, Kv));
  tmp.add(s.T2N, s.sT2N, s.tT2N, tri_geom_b(No, Nv, Ko));
  tmp.add(s.T2TN, s.sT2TN, s.tT2TN, tri_geom_b(No, Nv, Ko));
  // Tabij {Nv,Nv,No,No}: slowest index j
  staged(ctx, src, ATRIP_TENSOR_TABIJ, Nv2 * No, No,
         [&](const double *c, int64_t j0, int64_t cj) {
           for (int part = 0; part < m; part++) {
             DevSystem &d = *dst[part];
             // T3[x][p + No q][E] = T[x,E,p,q], rows of this chunk: q = j
             check(launch_transpose(c + part, d.T3 + j0 * No * Kv, No * cj, Nv,
                                    Nv, m * Nv2, m * Nv, m, Kv, 1, N2 * Kv,
                                    1.0, st),
                   "relayout T3");
             // T2N[pair][n][L] = -T[pair,L,n], n = j
             check(launch_transpose(c + part, d.T2N + j0 * Ko, cj, No, Nv2,
                                    m * Nv2 * No, m * Nv2, m, Ko, 1, No * Ko,
                                    -1.0, st),
                   "relayout T2N");
             // T2TN[pair][n][L] = -T[pair,n,L], L = j
             check(launch_transpose(c + part, d.T2TN + j0, cj, No, Nv2,
                                    m * Nv2 * No, m * Nv2, m, 1, Ko, No * Ko,
                                    -1.0, st),
                   "relayout T2TN");
           }
         });
  // Vabij {Nv,Nv,No,No}: VHH[pair][i + No j]
  staged(ctx, src, ATRIP_TENSOR_VABIJ, Nv2 * No, No,
         [&](const double *c, int64_t j0, int64_t cj) {
           for (int part = 0; part < m; part++)
             check(launch_transpose(c + part, dst[part]->VHH + j0 * No, 1,
                                    No * cj, Nv2, 0, m * Nv2, m, 0, 1, N2, 1.0,
                                    st),
                   "relayout VHH");
         });
  // Vijka {No,No,No,Nv}: H3[x][row][L] = V[row,L,x]
  auto hhh = [&](int t, double *DevSystem::*arr) {
    staged(ctx, src, t, N2 * No, Nv,
           [&](const double *c, int64_t x0, int64_t cx) {
             for (int part = 0; part < m; part++)
               check(launch_transpose(c + part, dst[part]->*arr + x0 * N2 * Ko,
                                      cx, No, N2, m * N2 * No, m * N2, m,
                                      N2 * Ko, 1, Ko, 1.0, st),
                     "relayout H3");
           });
  };
  // Vabci {Nv,Nv,Nv,No}: V3[pair][r][E] = V[pair,E,r]
  auto ppph = [&](int t, double *DevSystem::*arr) {
    staged(ctx, src, t, Nv2 * Nv, No,
           [&](const double *c, int64_t r0, int64_t cr) {
             for (int part = 0; part < m; part++)
               check(launch_transpose(c + part, dst[part]->*arr + r0 * Kv, cr,
                                      Nv, Nv2, m * Nv2 * Nv, m * Nv2, m, Kv, 1,
                                      No * Kv, 1.0, st),
                     "relayout V3");
           });
  };
  tmp.convert();
  tmp.add(s.H3, s.sH3, s.tH3, tri_geom_a(No, Nv, Ko));
  hhh(ATRIP_TENSOR_VIJKA, &DevSystem::H3);
  tmp.convert();
  tmp.add(s.V3, s.sV3, s.tV3, tri_geom_b(No, Nv, Kv));
  ppph(ATRIP_TENSOR_VABCI, &DevSystem::V3);
  tmp.convert();
  if (s.has_j) {
    tmp.add(s.H3J, s.sH3J, s.tH3J, tri_geom_a(No, Nv, Ko));
    hhh(ATRIP_TENSOR_JIJKA, &DevSystem::H3J);
    tmp.convert();
    tmp.add(s.V3J, s.sV3J, s.tV3J, tri_geom_b(No, Nv, Kv));
    ppph(ATRIP_TENSOR_JABCI, &DevSystem::V3J);
    tmp.convert();
  }
  // the small tensors whole (the reference's read_all, Atrip.cxx:176-181)
  auto small = [&](int t, int64_t n, double *DevSystem::*arr, const char *what) {
    std::vector<double> buf((size_t)(n * m)), part((size_t)n);
    src.host(t, 0, n, buf.data());
    for (int q = 0; q < m; q++) {
      for (int64_t e = 0; e < n; e++) part[e] = buf[e * m + q];
      check(hipMemcpy(dst[q]->*arr, part.data(), (size_t)n * sizeof(double),
                      hipMemcpyHostToDevice),
            what);
    }
  };
  small(ATRIP_TENSOR_TAI, Nv * No, &DevSystem::Tai, "upload Tai");
  small(ATRIP_TENSOR_EI, No, &DevSystem::ei, "upload ei");
  small(ATRIP_TENSOR_EA, Nv, &DevSystem::ea, "upload ea");
}

// The imaginary planes of a complex system in the slots of the real ones
// (for filling them with the real-system code).
DevSystem imag_view(const DevSystem &s) {
  DevSystem v;
  v.No = s.No; v.Nv = s.Nv; v.Kv = s.Kv; v.Ko = s.Ko;
  v.has_j = s.has_j;
  v.T3 = s.T3i; v.V3 = s.V3i; v.H3 = s.H3i; v.T2N = s.T2Ni;
  v.T2TN = s.T2TNi; v.VHH = s.VHHi; v.Tai = s.Taii; v.ei = s.eii;
  v.ea = s.eai; v.V3J = s.V3Ji; v.H3J = s.H3Ji;
  v.aprefix = s.aprefix;
  return v;
}

// Negated imaginary planes of T3 and the hole integrals.
void complex_negations(atrip_hip_ctx *ctx, DevSystem &s) {
  const int64_t N2 = s.No * s.No;
  check(launch_scale(s.T3i, s.T3in, s.Nv * N2 * s.Kv, -1.0, ctx->stream),
        "negate T3 (im)");
  check(launch_scale(s.H3i, s.H3in, s.Nv * N2 * s.Ko, -1.0, ctx->stream),
        "negate H3 (im)");
  if (s.has_j)
    check(launch_scale(s.H3Ji, s.H3Jin, s.Nv * N2 * s.Ko, -1.0, ctx->stream),
          "negate H3J (im)");
  check(hipStreamSynchronize(ctx->stream), "negations");
}

// The four real class-GEMM passes of a complex system: W = Wr + i Wi with
//   Wr = [T3r V3r + H3r N_r] + [(-T3i) V3i + H3i N_i]
//   Wi = [T3r V3i + H3r N_i] + [T3i V3r + (-H3i) N_r]
// where N = -T2 (the pre-negated hole operands) and the hole terms carry
// conj(Vhhh) = H3r - i H3i, as the reference's complex GEMM path does
// (MAYBE_CONJ, src/atrip/Equations.cxx:620-656).
void complex_passes(const DevSystem &s, DevSystem v[4]) {
  for (int q = 0; q < 4; q++) {
    v[q].No = s.No; v[q].Nv = s.Nv; v[q].Kv = s.Kv; v[q].Ko = s.Ko;
    v[q].aprefix = s.aprefix;
    v[q].has_j = s.has_j;
  }
  v[0].T3 = s.T3;   v[0].V3 = s.V3;  v[0].H3 = s.H3;   v[0].T2N = s.T2N;
  v[0].T2TN = s.T2TN;  v[0].V3J = s.V3J;  v[0].H3J = s.H3J;
  v[1].T3 = s.T3in; v[1].V3 = s.V3i; v[1].H3 = s.H3i;  v[1].T2N = s.T2Ni;
  v[1].T2TN = s.T2TNi; v[1].V3J = s.V3Ji; v[1].H3J = s.H3Ji;
  v[2].T3 = s.T3;   v[2].V3 = s.V3i; v[2].H3 = s.H3;   v[2].T2N = s.T2Ni;
  v[2].T2TN = s.T2TNi; v[2].V3J = s.V3Ji; v[2].H3J = s.H3J;
  v[3].T3 = s.T3i;  v[3].V3 = s.V3;  v[3].H3 = s.H3in; v[3].T2N = s.T2N;
  v[3].T2TN = s.T2TN;  v[3].V3J = s.V3J;  v[3].H3J = s.H3Jin;
}

// Host-side canonical index, identical to the device one.
struct HostIndex {
  std::vector<uint64_t> pre;
  uint64_t Nv;
  explicit HostIndex(uint64_t nv) : pre(nv + 1, 0), Nv(nv) {
    for (uint64_t a = 0; a < nv; a++) {
      const uint64_t m = nv - a;
      pre[a + 1] = pre[a] + m * (m + 1) / 2 - 1;
    }
  }
  uint64_t operator()(uint64_t a, uint64_t b, uint64_t c) const {
    return tuple_index(pre.data(), Nv, a, b, c);
  }
};

// Window plan for a system at the context's window edge: the block triples
// in lexicographic order with their canonical index span and tuple count.
WindowPlan &get_plan(atrip_hip_ctx *ctx, int sys, const DevSystem &s,
                     int nb_want = 0) {
  if (nb_want <= 0) nb_want = window_edge(ctx, s);
  auto it = ctx->plans.find(sys);
  if (it != ctx->plans.end() && it->second.nb == nb_want) return it->second;
  if (it != ctx->plans.end()) ctx->plans.erase(it);
  WindowPlan p;
  p.nb = nb_want;
  p.nblk = (int)((s.Nv + p.nb - 1) / p.nb);
  const HostIndex hidx((uint64_t)s.Nv);
  const int nb = p.nb;
  for (int al = 0; al < p.nblk; al++)
    for (int be = al; be < p.nblk; be++)
      for (int ga = be; ga < p.nblk; ga++) {
        uint64_t lo = ~0ULL, hi = 0, cnt = 0;
        for (int ta = 0; ta < nb; ta++)
          for (int tb = 0; tb < nb; tb++)
            for (int tc = 0; tc < nb; tc++) {
              const uint64_t a = (uint64_t)al * nb + ta,
                             b = (uint64_t)be * nb + tb,
                             c = (uint64_t)ga * nb + tc;
              if (!tuple_ok(s.Nv, a, b, c)) continue;
              const uint64_t q = hidx(a, b, c);
              lo = std::min(lo, q);
              hi = std::max(hi, q);
              cnt++;
            }
        if (!cnt) continue;
        p.windows.insert(p.windows.end(), {al, be, ga});
        p.wmin.push_back(lo);
        p.wmax.push_back(hi);
        p.wcount.push_back(cnt);
      }
  return ctx->plans.emplace(sys, std::move(p)).first->second;
}

void ensure(void *&ptr, size_t &have, size_t need, const char *what) {
  if (have >= need && ptr) return;
  if (ptr) check(hipFree(ptr), what);
  ptr = nullptr;
  have = 0;
  check(hipMalloc(&ptr, need), what);
  have = need;
}

struct RunResult {
  double e = 0, ect = 0;
  uint64_t tuples = 0;  // tuples computed
};

// Destroys the queued timing events (read at the end of a run, or left by
// a run that failed); each event can end one pair and start the next, so
// they are deduplicated first.
void drop_pending_events(atrip_hip_ctx *ctx) {
  std::vector<hipEvent_t> uniq;
  for (auto &pr : ctx->pending_events) {
    uniq.push_back(pr.first);
    uniq.push_back(pr.second);
  }
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  for (auto ev : uniq) (void)hipEventDestroy(ev);
  ctx->pending_events.clear();
  ctx->pending_names.clear();
  ctx->pending_flop.clear();
}

// Runs every batch of windows intersecting [first, last).  When `sink` is
// given it receives the per-slot energies of each batch.
// A fixed device region a run may carve its scratch from (no allocation).
struct FixedScratch {
  void *base;
  size_t bytes;
};

// Which windows of the plan a run covers besides the tuple range: every
// stride-th row of windows (block pairs) from row offset, and plan windows
// [wbegin, wend) by their index in the plan.
struct WindowSel {
  uint64_t stride = 1, offset = 0;
  uint64_t wbegin = 0, wend = ~0ULL;
  int nb = 0;  // window edge of the plan; 0: the context's (window_edge)
};

template <typename Sink>
RunResult run_range(atrip_hip_ctx *ctx, int sys, int precision, int ijkabc,
                    uint64_t first, uint64_t last, Sink &&sink,
                    const WindowSel &wsel = WindowSel{},
                    const FixedScratch *fixed = nullptr) {
  drop_pending_events(ctx);
  DevSystem &s = get_sys(ctx, sys);
  WindowPlan &plan = get_plan(ctx, sys, s, wsel.nb);
  const uint64_t n = n_tuples(s.Nv);
  last = std::min(last, n);
  RunResult res;
  check_precision(s, precision);
  const bool g32 = precision != ATRIP_FP64;
  if (fixed) {
    // the operand copies of a mixed form are device allocations: a run
    // without allocation needs them resident already
    if (precision == ATRIP_BF16X3 && !s.tri_ready)
      throw Fail{"bf16x3 operand copies are not resident: run the system "
                 "once in bf16x3 through atrip_hip_run, or upload it with "
                 "operand_store 3"};
    if ((precision == ATRIP_SPLIT_FP32 || precision == ATRIP_FP32) &&
        !s.split_ready)
      throw Fail{"split FP32 operand copies are not resident: run the system "
                 "once in a split form through atrip_hip_run, or upload it "
                 "with operand_store 1 or 2"};
  } else if (precision == ATRIP_BF16X3) {
    ensure_tri(ctx, s);
  } else if (g32) {
    ensure_split(ctx, s);
  }
  if (first >= last) return res;
  const int nb = plan.nb;
  const bool full = first == 0 && last == n;
  HostIndex hidx((uint64_t)s.Nv);
  // select windows and count their tuples in range
  std::vector<int32_t> sel;
  std::vector<uint64_t> counts;
  const size_t nwin = plan.windows.size() / 3;
  int64_t pair = -1, prev_al = -1, prev_be = -1;  // (alpha, beta) row index
  for (size_t w = 0; w < nwin; w++) {
    const int64_t al = plan.windows[3 * w], be = plan.windows[3 * w + 1],
                  ga = plan.windows[3 * w + 2];
    if (al != prev_al || be != prev_be) {
      pair++;
      prev_al = al;
      prev_be = be;
    }
    if (w < wsel.wbegin || w >= wsel.wend) continue;
    if (wsel.stride > 1 && (uint64_t)pair % wsel.stride != wsel.offset % wsel.stride)
      continue;
    uint64_t cnt = 0;
    if (full || (plan.wmin[w] >= first && plan.wmax[w] < last)) {
      cnt = plan.wcount[w];
    } else if (plan.wmax[w] < first || plan.wmin[w] >= last) {
      continue;
    } else {
      for (int ta = 0; ta < nb; ta++)
        for (int tb = 0; tb < nb; tb++)
          for (int tc = 0; tc < nb; tc++) {
            const uint64_t a = al * nb + ta, b = be * nb + tb,
                           c = ga * nb + tc;
            if (!tuple_ok(s.Nv, a, b, c)) continue;
            const uint64_t q = hidx(a, b, c);
            if (q >= first && q < last) cnt++;
          }
    }
    if (!cnt) continue;
    sel.insert(sel.end(), {(int32_t)al, (int32_t)be, (int32_t)ga});
    counts.push_back(cnt);
    res.tuples += cnt;
  }
  const int nsel = (int)counts.size();
  if (!nsel) return res;
  // Device scratch of the run: class buffers G (and those of the (cT)
  // pass; four sets for complex systems), per-slot energies, the energy
  // kernels' scratch (per-triple partials, or W cubes for complex systems
  // and the cube epilogue), batch partials and the window list.  Taken from the context's buffers (grown
  // on demand), or carved out of a fixed region without any allocation
  // (then the windows per launch shrink until the run fits).
  const int64_t nb3 = (int64_t)nb * nb * nb;
  const int64_t W3 = g_elems(s.No, g32);  // one tuple's class buffer
  const bool use_j = s.has_j;
  const int nsets = s.cplx ? 4 : 1;
  const size_t gelem = g32 ? sizeof(float) : sizeof(double);
  const bool fused = ctx->epilogue == 0;
  struct Sizes {
    int64_t slots_max = 0;
    size_t gsz = 0, g = 0, e_slot = 0, wcube = 0, partials = 0, windows = 0;
    int nbatch = 0;
  };
  auto sizes = [&](int bw_) {
    Sizes z;
    z.slots_max = (int64_t)std::min(bw_, nsel) * nb3;
    z.gsz = (size_t)(3 * z.slots_max * W3) * gelem;
    z.g = (size_t)((use_j ? 2 : 1) * nsets) * z.gsz;
    z.e_slot = (size_t)(2 * z.slots_max) * sizeof(double);
    z.wcube = (size_t)(z.slots_max * energy_scratch_per_slot(
                                         s.No, use_j, s.cplx, fused)) *
              sizeof(double);
    z.nbatch = (nsel + bw_ - 1) / bw_;
    z.partials = (size_t)(4 * z.nbatch) * sizeof(double);
    z.windows = sel.size() * sizeof(int32_t);
    return z;
  };
  auto a256 = [](size_t x) { return (x + 255) / 256 * 256; };
  auto total = [&](const Sizes &z) {
    return a256(z.g) + a256(z.e_slot) + a256(z.wcube) + a256(z.partials) +
           a256(z.windows);
  };

  int bw = std::max(1, ctx->batch);
  void *Gbuf = nullptr, *e_slot = nullptr, *wcube = nullptr,
       *partials = nullptr, *d_sel = nullptr;
  Sizes z;
  // per-window sums of a window-sum sink, for every selected window
  constexpr bool kWinSums = std::decay_t<Sink>::window_sums();
  const size_t wsz = kWinSums ? (size_t)(2 * nsel) * sizeof(double) : 0;
  double *wsum_dev = nullptr;
  if (fixed && kWinSums)
    throw Fail{"per-window sums are not an allocation-free run"};
  if (fixed) {
    while (bw > 1 && total(sizes(bw)) > fixed->bytes) bw--;
    z = sizes(bw);
    if (total(z) > fixed->bytes) {
      // largest window edge whose single window fits (class buffers, cubes
      // and slot energies scale with edge^3)
      const size_t per_slot = (z.g + z.wcube + z.e_slot) / (size_t)nb3;
      const size_t rest = a256(z.partials) + a256(z.windows) + 3 * 256;
      int fit = 0;
      for (int e = nb - 1; e >= 1 && !fit; e--)
        if ((size_t)e * e * e * per_slot + rest <= fixed->bytes) fit = e;
      throw Fail{"run needs " + std::to_string(total(z)) +
                 " bytes of scratch for one window of edge " +
                 std::to_string(nb) + " at No=" + std::to_string(s.No) +
                 "; the allocation-free entry point holds " +
                 std::to_string(fixed->bytes) +
                 (fit ? " (window edge " + std::to_string(fit) + " fits)"
                      : " (no window edge fits this No)")};
    }
    char *p = (char *)fixed->base;
    Gbuf = p;
    p += a256(z.g);
    e_slot = p;
    p += a256(z.e_slot);
    wcube = p;
    p += a256(z.wcube);
    partials = p;
    p += a256(z.partials);
    d_sel = p;
    check(hipMemcpyAsync(d_sel, sel.data(), z.windows, hipMemcpyHostToDevice,
                         ctx->stream),
          "upload windows");
  } else {
    z = sizes(bw);
    // the window list lives in a context buffer grown on demand (no
    // allocation and release per call); the copy is ordered on the stream
    // before the kernels that read it, and `sel` outlives it (the run
    // waits for the stream before returning)
    ensure(ctx->sel, ctx->sel_bytes, a256(z.windows) + wsz, "alloc windows");
    d_sel = ctx->sel;
    if (kWinSums) wsum_dev = (double *)((char *)d_sel + a256(z.windows));
    check(hipMemcpyAsync(d_sel, sel.data(), z.windows, hipMemcpyHostToDevice,
                         ctx->stream),
          "upload windows");
    ensure((void *&)ctx->G, ctx->G_bytes, z.g, "alloc G");
    ensure((void *&)ctx->e_slot, ctx->e_slot_bytes, z.e_slot, "alloc e_slot");
    ensure((void *&)ctx->wcube, ctx->wcube_bytes, z.wcube,
           "alloc energy scratch");
    ensure((void *&)ctx->partials, ctx->partials_bytes, z.partials,
           "alloc partials");
    Gbuf = ctx->G;
    e_slot = ctx->e_slot;
    wcube = ctx->wcube;
    partials = ctx->partials;
  }
  const size_t gsz = z.gsz;
  const int nbatch = z.nbatch;
  DevSystem zpass[4];
  if (s.cplx) complex_passes(s, zpass);
  // timing events are objects the runtime creates: none without allocation
  const bool timing = ctx->timing && !fixed;
  // complex products are four real ones (the reference bench counts 4x)
  const double flop_tuple = 12.0 * (double)s.No * s.No * s.No * (s.No + s.Nv) *
                            (s.cplx ? 4.0 : 1.0);
  const char *dname = precision == ATRIP_FP64        ? "doubles"
                      : precision == ATRIP_SPLIT_FP32 ? "doubles_mixed"
                      : precision == ATRIP_BF16X3     ? "doubles_bf16x3"
                                                      : "doubles_fp32";
  hipStream_t st = ctx->stream;
  double *Gz[4], *GJz[4];
  for (int q = 0; q < 4; q++) {
    Gz[q] = (double *)((char *)Gbuf + q * gsz);
    GJz[q] = (double *)((char *)Gbuf + (4 + q) * gsz);
  }
  void *G = Gbuf;
  void *GJ = use_j ? (void *)((char *)Gbuf + gsz) : nullptr;
  std::vector<double> host_slots;

  for (int bi = 0; bi < nbatch; bi++) {
    BatchArgs b;
    const int w0 = bi * bw;
    b.nw = std::min(bw, nsel - w0);
    b.windows = (const int32_t *)d_sel + 3 * w0;
    b.nb = nb;
    b.first = full ? 0 : first;
    b.last = full ? ~0ULL : last;
    b.slots = (int64_t)b.nw * nb3;
    uint64_t cnt = 0;
    for (int w = 0; w < b.nw; w++) cnt += counts[w0 + w];
    // the batch's timing events; destroyed here unless handed to the
    // context's pending list (a launch that throws leaves none behind)
    struct BatchEvents {
      hipEvent_t e[3] = {nullptr, nullptr, nullptr};
      bool handed = false;
      ~BatchEvents() {
        if (!handed)
          for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
      }
    } ev;
    if (timing) {
      for (hipEvent_t &x : ev.e) check(hipEventCreate(&x), "event");
      check(hipEventRecord(ev.e[0], st), "event");
    }
    for (int pass = 0; pass < (use_j ? 2 : 1) && s.cplx; pass++)
      for (int q = 0; q < 4; q++)
        check(launch_doubles_f64(zpass[q], b, pass == 1,
                                 pass == 0 ? Gz[q] : GJz[q], ctx->tile, st),
              "doubles kernel (complex pass)");
    for (int pass = 0; pass < (use_j ? 2 : 1) && !s.cplx; pass++) {
      void *Gp = pass == 0 ? G : GJ;
      if (precision == ATRIP_FP64)
        check(launch_doubles_f64(s, b, pass == 1, (double *)Gp, ctx->tile, st),
              "doubles kernel");
      else if (precision == ATRIP_BF16X3)
        check(launch_doubles_tri(s, b, pass == 1, (float *)Gp, ctx->tile, st),
              "doubles (bf16x3) kernel");
      else
        check(launch_doubles_split(s, b, pass == 1, (float *)Gp,
                                   precision == ATRIP_SPLIT_FP32, ctx->tile,
                                   st),
              "doubles (FP32 products) kernel");
    }
    if (timing) check(hipEventRecord(ev.e[1], st), "event");
    if (s.cplx)
      check(launch_energy_z(s, b, Gz, use_j ? GJz : nullptr,
                            ijkabc ? -1.0 : 1.0, (double *)e_slot,
                            (double *)wcube, st),
            "energy kernel (complex)");
    else
      check(launch_energy(s, b, G, GJ, g32, ijkabc ? -1.0 : 1.0,
                          (double *)e_slot, (double *)wcube, fused, st),
            "energy kernel");
    if (timing) {
      check(hipEventRecord(ev.e[2], st), "event");
      ctx->pending_events.push_back({ev.e[0], ev.e[1]});
      ctx->pending_names.push_back(dname);
      ctx->pending_flop.push_back(flop_tuple * cnt * (use_j ? 2 : 1));
      ctx->pending_events.push_back({ev.e[1], ev.e[2]});
      ctx->pending_names.push_back("energy");
      ctx->pending_flop.push_back(0.0);
      ev.handed = true;
    }
    check(launch_batch_reduce(s, b, (double *)e_slot, (double *)partials, bi,
                              st),
          "reduce kernel");
    if constexpr (kWinSums) {
      check(launch_window_sums((const double *)e_slot, b.nw, nb3,
                               wsum_dev + 2 * w0, st),
            "window sums kernel");
    } else if (sink.active()) {
      host_slots.resize(2 * b.slots);
      check(hipMemcpyAsync(host_slots.data(), (double *)e_slot,
                           host_slots.size() * sizeof(double),
                           hipMemcpyDeviceToHost, st),
            "download slots");
      check(hipStreamSynchronize(st), "slots");
      sink(sel.data() + 3 * w0, b.nw, nb, host_slots);
    }
  }
  std::vector<double> part(4 * nbatch);
  check(hipMemcpyAsync(part.data(), (double *)partials, part.size() * sizeof(double),
                       hipMemcpyDeviceToHost, st),
        "download partials");
  std::vector<double> wsums(kWinSums ? 2 * (size_t)nsel : 0);
  if (kWinSums)
    check(hipMemcpyAsync(wsums.data(), wsum_dev, wsz, hipMemcpyDeviceToHost,
                         st),
          "download window sums");
  check(hipStreamSynchronize(st), "run");
  if constexpr (kWinSums) sink.windows(wsums);
  // Neumaier over the batch partials in batch order
  double s0 = 0, c0 = 0, s1 = 0, c1 = 0;
  auto add = [](double &sum, double &comp, double x) {
    const double t = sum + x;
    comp += (std::fabs(sum) >= std::fabs(x)) ? (sum - t) + x : (x - t) + sum;
    sum = t;
  };
  for (int bi = 0; bi < nbatch; bi++) {
    add(s0, c0, part[4 * bi]);
    add(s0, c0, part[4 * bi + 1]);
    add(s1, c1, part[4 * bi + 2]);
    add(s1, c1, part[4 * bi + 3]);
  }
  res.e = s0 + c0;
  res.ect = s1 + c1;
  // collect event timings
  for (size_t q = 0; q < ctx->pending_events.size(); q++) {
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, ctx->pending_events[q].first,
                              ctx->pending_events[q].second),
          "elapsed");
    KernelStat &k = ctx->stats[ctx->pending_names[q]];
    k.launches++;
    k.ms += ms;
    k.flop += ctx->pending_flop[q];
  }
  drop_pending_events(ctx);  // read: release them
  return res;
}

struct NoSink {
  static constexpr bool window_sums() { return false; }
  bool active() const { return false; }
  void operator()(const int32_t *, int, int, const std::vector<double> &) {}
};

}  // namespace

int atrip_amd::window_edge(const atrip_hip_ctx *ctx, const DevSystem &s) {
  return ctx->window > 0 ? ctx->window : auto_window(s.No, s.Nv);
}

uint64_t atrip_amd::plan_windows(atrip_hip_ctx *ctx, int sys, int nb) {
  DevSystem &s = get_sys(ctx, sys);
  return get_plan(ctx, sys, s, nb).windows.size() / 3;
}

void atrip_amd::run_plan_windows(atrip_hip_ctx *ctx, int sys, int precision,
                                 int nb, uint64_t wbegin, uint64_t wend,
                                 std::vector<double> &out) {
  out.clear();
  // per-window (sum, comp) pairs formed on the device, downloaded once
  struct WinSink {
    static constexpr bool window_sums() { return true; }
    std::vector<double> *out;
    bool active() const { return true; }
    void windows(const std::vector<double> &pairs) {
      out->insert(out->end(), pairs.begin(), pairs.end());
    }
  } sink{&out};
  WindowSel ws;
  ws.wbegin = wbegin;
  ws.wend = wend;
  ws.nb = nb;
  run_range(ctx, sys, precision, 0, 0, ~0ULL, sink, ws);
}

void atrip_amd::run_fixed(atrip_hip_ctx *ctx, int sys, int precision,
                          int ijkabc, uint64_t first, uint64_t count,
                          void *base, size_t bytes, double *energy,
                          double *ct_energy) {
  const uint64_t last =
      count == ~0ULL || first + count < first ? ~0ULL : first + count;
  const FixedScratch fx{base, bytes};
  RunResult r = run_range(ctx, sys, precision, ijkabc, first, last, NoSink{},
                          WindowSel{}, &fx);
  *energy = ijkabc ? r.e : -r.e;
  if (ct_energy) *ct_energy = ijkabc ? r.ect : -r.ect;
}

extern "C" {

int atrip_hip_plan_model(int64_t No, int64_t Nv, int precision, int window,
                         int *nb, int *fm, int *fn) {
  if (No < 1 || Nv < 1 || window < 0 || window > 32) return 1;
  if (precision < ATRIP_FP64 || precision > ATRIP_BF16X3) return 1;
  // the same choices as window_edge and the doubles launchers make
  const int e = window > 0 ? window : auto_window(No, Nv);
  int m, n;
  if (precision == ATRIP_FP64) {
    const int t = pick_tile(-1, No, e);
    m = tile_fm(t);
    n = tile_fn(t);
  } else {
    const int t = pick_t32_tile(-1, No, e);
    m = t32_fm(t);
    n = t32_fn(t);
  }
  if (nb) *nb = e;
  if (fm) *fm = m;
  if (fn) *fn = n;
  return 0;
}

int atrip_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int atrip_hip_create(int device, atrip_hip_ctx **out) {
  if (!out) return 1;
  *out = nullptr;
  auto *ctx = new atrip_hip_ctx();
  ctx->device = device;
  int rc = guarded(ctx, [&] {
    int n = 0;
    check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n)
      throw Fail{"device index " + std::to_string(device) + " out of range (" +
                 std::to_string(n) + " devices)"};
    check(hipSetDevice(device), "hipSetDevice");
    check(hipStreamCreate(&ctx->stream), "hipStreamCreate");
  });
  if (rc) {
    static thread_local std::string last;
    last = ctx->err;
    delete ctx;
    return rc;
  }
  *out = ctx;
  return 0;
}

int atrip_hip_destroy(atrip_hip_ctx *ctx) {
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->systems) free_system(kv.second);
  if (ctx->G) (void)hipFree(ctx->G);
  if (ctx->e_slot) (void)hipFree(ctx->e_slot);
  if (ctx->partials) (void)hipFree(ctx->partials);
  if (ctx->sel) (void)hipFree(ctx->sel);
  if (ctx->bscratch) (void)hipFree(ctx->bscratch);
  if (ctx->wcube) (void)hipFree(ctx->wcube);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

const char *atrip_hip_last_error(atrip_hip_ctx *ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int atrip_hip_upload(atrip_hip_ctx *ctx, const atrip_sys_desc *desc,
                     int *sys) {
  return guarded(ctx, [&] {
    if (!desc || !sys) throw Fail{"null argument"};
    if (!desc->ei || !desc->ea || !desc->Tai || !desc->Tabij ||
        !desc->Vabij || !desc->Vijka || !desc->Vabci)
      throw Fail{"missing input tensor"};
    if ((desc->Jijka == nullptr) != (desc->Jabci == nullptr))
      throw Fail{"Jijka and Jabci must be given together"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    DevSystem s;
    try {
      alloc_system(ctx, s, desc->No, desc->Nv, desc->Jijka != nullptr);
      DescSource src(desc, 1);
      upload_tensors(ctx, s, nullptr, src);
    } catch (...) {
      free_system(s);
      throw;
    }
    const int id = ctx->next_sys++;
    ctx->systems.emplace(id, std::move(s));
    *sys = id;
  });
}

int atrip_hip_upload_synthetic(atrip_hip_ctx *ctx, int64_t No, int64_t Nv,
                               uint64_t seed, int with_j, int *sys) {
  return atrip_hip_upload_synthetic_scaled(ctx, No, Nv, seed, with_j, 1.0,
                                           sys);
}

int atrip_hip_upload_synthetic_scaled(atrip_hip_ctx *ctx, int64_t No,
                                      int64_t Nv, uint64_t seed, int with_j,
                                      double scale, int *sys) {
  return guarded(ctx, [&] {
    if (!sys) throw Fail{"null argument"};
    if (!std::isfinite(scale)) throw Fail{"scale must be finite"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    DevSystem s;
    try {
      alloc_system(ctx, s, No, Nv, with_j != 0);
      check(launch_synth(s, seed, with_j != 0, ctx->stream, 0, scale),
            "synthetic fill");
      check(hipStreamSynchronize(ctx->stream), "synthetic fill");
    } catch (...) {
      free_system(s);
      throw;
    }
    const int id = ctx->next_sys++;
    ctx->systems.emplace(id, std::move(s));
    *sys = id;
  });
}

int atrip_hip_upload_complex(atrip_hip_ctx *ctx, const atrip_sys_desc *desc,
                             int *sys) {
  return guarded(ctx, [&] {
    if (!desc || !sys) throw Fail{"null argument"};
    if (!desc->ei || !desc->ea || !desc->Tai || !desc->Tabij ||
        !desc->Vabij || !desc->Vijka || !desc->Vabci)
      throw Fail{"missing input tensor"};
    if ((desc->Jijka == nullptr) != (desc->Jabci == nullptr))
      throw Fail{"Jijka and Jabci must be given together"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    DevSystem s;
    try {
      alloc_system(ctx, s, desc->No, desc->Nv, desc->Jijka != nullptr, true);
      // real and imaginary planes from the interleaved values in one pass
      DescSource src(desc, 2);
      DevSystem im = imag_view(s);
      upload_tensors(ctx, s, &im, src);
      complex_negations(ctx, s);
    } catch (...) {
      free_system(s);
      throw;
    }
    const int id = ctx->next_sys++;
    ctx->systems.emplace(id, std::move(s));
    *sys = id;
  });
}

int atrip_hip_upload_stream(atrip_hip_ctx *ctx, int64_t No, int64_t Nv,
                            int flags, atrip_read_fn read, void *user,
                            int *sys) {
  return guarded(ctx, [&] {
    if (!read || !sys) throw Fail{"null argument"};
    if (flags & ~(ATRIP_STREAM_WITH_J | ATRIP_STREAM_COMPLEX))
      throw Fail{"unknown upload flags"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    const bool with_j = flags & ATRIP_STREAM_WITH_J;
    const bool cplx = flags & ATRIP_STREAM_COMPLEX;
    DevSystem s;
    try {
      alloc_system(ctx, s, No, Nv, with_j, cplx);
      ReaderSource src(read, user, cplx ? 2 : 1,
                       std::min<int64_t>(ctx->stage_kb * 1024, PINNED_MAX));
      if (cplx) {
        DevSystem im = imag_view(s);
        upload_tensors(ctx, s, &im, src);
        complex_negations(ctx, s);
      } else {
        upload_tensors(ctx, s, nullptr, src);
      }
    } catch (...) {
      (void)hipStreamSynchronize(ctx->stream);
      free_system(s);
      throw;
    }
    const int id = ctx->next_sys++;
    ctx->systems.emplace(id, std::move(s));
    *sys = id;
  });
}

int atrip_hip_upload_synthetic_complex(atrip_hip_ctx *ctx, int64_t No,
                                       int64_t Nv, uint64_t seed, int with_j,
                                       int *sys) {
  return guarded(ctx, [&] {
    if (!sys) throw Fail{"null argument"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    DevSystem s;
    try {
      alloc_system(ctx, s, No, Nv, with_j != 0, true);
      check(launch_synth(s, seed, with_j != 0, ctx->stream), "synthetic fill");
      DevSystem v = imag_view(s);
      check(launch_synth(v, seed, with_j != 0, ctx->stream, ATRIP_SYN_IMAG),
            "synthetic fill (im)");
      check(hipStreamSynchronize(ctx->stream), "synthetic fill");
      complex_negations(ctx, s);
    } catch (...) {
      free_system(s);
      throw;
    }
    const int id = ctx->next_sys++;
    ctx->systems.emplace(id, std::move(s));
    *sys = id;
  });
}

int atrip_hip_release(atrip_hip_ctx *ctx, int sys) {
  return guarded(ctx, [&] {
    DevSystem &s = get_sys(ctx, sys);
    check(hipStreamSynchronize(ctx->stream), "sync");
    free_system(s);
    ctx->systems.erase(sys);
    ctx->plans.erase(sys);
  });
}

int atrip_hip_run(atrip_hip_ctx *ctx, int sys, int precision, int ijkabc,
                  uint64_t first, uint64_t count, double *energy,
                  double *ct_energy) {
  return guarded(ctx, [&] {
    if (!energy) throw Fail{"null energy pointer"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    const uint64_t last =
        count == ~0ULL || first + count < first ? ~0ULL : first + count;
    RunResult r = run_range(ctx, sys, precision, ijkabc, first, last, NoSink{});
    // src/atrip/Atrip.cxx:1108-1111: E = -sum unless ijkabc
    *energy = ijkabc ? r.e : -r.e;
    if (ct_energy) *ct_energy = ijkabc ? r.ect : -r.ect;
  });
}

int atrip_hip_run_windows(atrip_hip_ctx *ctx, int sys, int precision,
                          int ijkabc, uint64_t stride, uint64_t offset,
                          double *energy, double *ct_energy,
                          uint64_t *tuples) {
  return guarded(ctx, [&] {
    if (!energy) throw Fail{"null energy pointer"};
    if (stride == 0) throw Fail{"stride must be positive"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    WindowSel ws;
    ws.stride = stride;
    ws.offset = offset;
    RunResult r = run_range(ctx, sys, precision, ijkabc, 0, ~0ULL, NoSink{}, ws);
    *energy = ijkabc ? r.e : -r.e;
    if (ct_energy) *ct_energy = ijkabc ? r.ect : -r.ect;
    if (tuples) *tuples = r.tuples;
  });
}

int atrip_hip_tuple_energies(atrip_hip_ctx *ctx, int sys, int precision,
                             int ijkabc, uint64_t first, uint64_t count,
                             double *e, double *e_ct) {
  return guarded(ctx, [&] {
    if (!e) throw Fail{"null output pointer"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    DevSystem &s = get_sys(ctx, sys);
    const uint64_t n = n_tuples(s.Nv);
    if (first > n) first = n;
    if (count > n - first) count = n - first;
    const uint64_t last = first + count;
    HostIndex hidx((uint64_t)s.Nv);
    struct Sink {
      static constexpr bool window_sums() { return false; }
      bool active() const { return true; }
      const HostIndex *h;
      uint64_t first, last, Nv;
      double *e, *ect;
      void operator()(const int32_t *win, int nw, int nb,
                      const std::vector<double> &slots) {
        const int64_t nb3 = (int64_t)nb * nb * nb;
        for (int w = 0; w < nw; w++)
          for (int64_t q = 0; q < nb3; q++) {
            const uint64_t a = (uint64_t)win[3 * w] * nb + q / (nb * nb);
            const uint64_t b = (uint64_t)win[3 * w + 1] * nb + (q / nb) % nb;
            const uint64_t c = (uint64_t)win[3 * w + 2] * nb + q % nb;
            if (!tuple_ok(Nv, a, b, c)) continue;
            const uint64_t idx = (*h)(a, b, c);
            if (idx < first || idx >= last) continue;
            e[idx - first] = slots[2 * (w * nb3 + q)];
            if (ect) ect[idx - first] = slots[2 * (w * nb3 + q) + 1];
          }
      }
    } sink{&hidx, first, last, (uint64_t)s.Nv, e, e_ct};
    run_range(ctx, sys, precision, ijkabc, first, last, sink);
  });
}

int atrip_hip_set_option(atrip_hip_ctx *ctx, const char *key, int64_t value) {
  return guarded(ctx, [&] {
    const std::string k = key ? key : "";
    if (k == "window") {
      if (value < 0 || value > 32)
        throw Fail{"window must be in 0..32 (0: automatic)"};
      ctx->window = (int)value;
    } else if (k == "batch") {
      if (value < 1 || value > 4096) throw Fail{"batch must be in 1..4096"};
      ctx->batch = (int)value;
    } else if (k == "timing") {
      ctx->timing = value != 0;
    } else if (k == "tile") {
      if (value < -1 || value > 9) throw Fail{"tile must be in -1..9"};
      ctx->tile = (int)value;
    } else if (k == "operand_store") {
      if (value < 0 || value > 3)
        throw Fail{"operand_store must be 0, 1, 2 or 3"};
      ctx->operand_store = (int)value;
    } else if (k == "serve_plain_no") {
      if (value < 1 || value > (1 << 20))
        throw Fail{"serve_plain_no must be in 1..1048576"};
      ctx->serve_plain_no = (int)value;
    } else if (k == "epilogue") {
      if (value < 0 || value > 1) throw Fail{"epilogue must be 0 or 1"};
      ctx->epilogue = (int)value;
    } else if (k == "stage_kb") {
      if (value < 1 || value > ((int64_t)1 << 30))
        throw Fail{"stage_kb must be in 1..2^30"};
      ctx->stage_kb = value;
    } else if (k == "batch_scratch_mb") {
      if (value < 16 || value > (1 << 18))
        throw Fail{"batch_scratch_mb must be in 16..262144"};
      ctx->batch_scratch_mb = value;
    } else if (k == "batch_grid") {
      if (value < 0 || value > (1 << 20))
        throw Fail{"batch_grid must be in 0..1048576"};
      ctx->batch_grid = (int)value;
    } else if (k == "batch_order") {
      if (value < 0 || value > 1) throw Fail{"batch_order must be 0 or 1"};
      ctx->batch_order = (int)value;
    } else if (k == "batch_nb") {
      if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8)
        throw Fail{"batch_nb must be 0, 1, 2, 4 or 8"};
      ctx->batch_nb = (int)value;
    } else if (k == "batch_tile") {
      if (value < -1 || value >= N_BTILES)
        throw Fail{"batch_tile must be in -1.." + std::to_string(N_BTILES - 1)};
      ctx->batch_tile = (int)value;
    } else {
      throw Fail{"unknown option '" + k + "'"};
    }
  });
}

int atrip_hip_get_option(atrip_hip_ctx *ctx, const char *key, int64_t *value) {
  return guarded(ctx, [&] {
    const std::string k = key ? key : "";
    if (!value) throw Fail{"null value pointer"};
    if (k == "window") *value = ctx->window;
    else if (k == "batch") *value = ctx->batch;
    else if (k == "timing") *value = ctx->timing;
    else if (k == "tile") *value = ctx->tile;
    else if (k == "operand_store") *value = ctx->operand_store;
    else if (k == "stage_kb") *value = ctx->stage_kb;
    else if (k == "epilogue") *value = ctx->epilogue;
    else if (k == "serve_plain_no") *value = ctx->serve_plain_no;
    else if (k == "batch_scratch_mb") *value = ctx->batch_scratch_mb;
    else if (k == "batch_grid") *value = ctx->batch_grid;
    else if (k == "batch_order") *value = ctx->batch_order;
    else if (k == "batch_nb") *value = ctx->batch_nb;
    else if (k == "batch_tile") *value = ctx->batch_tile;
    else throw Fail{"unknown option '" + k + "'"};
  });
}

int atrip_hip_kernel_stats(atrip_hip_ctx *ctx, const char *name,
                           uint64_t *launches, double *total_ms,
                           double *flop) {
  return guarded(ctx, [&] {
    const KernelStat k = ctx->stats[name ? name : ""];
    if (launches) *launches = k.launches;
    if (total_ms) *total_ms = k.ms;
    if (flop) *flop = k.flop;
  });
}

int atrip_hip_reset_stats(atrip_hip_ctx *ctx) {
  return guarded(ctx, [&] { ctx->stats.clear(); });
}

int atrip_hip_sys_info(atrip_hip_ctx *ctx, int sys, int64_t *No, int64_t *Nv,
                       uint64_t *n_tup, uint64_t *device_bytes) {
  return guarded(ctx, [&] {
    DevSystem &s = get_sys(ctx, sys);
    if (No) *No = s.No;
    if (Nv) *Nv = s.Nv;
    if (n_tup) *n_tup = n_tuples(s.Nv);
    if (device_bytes) *device_bytes = s.bytes;
  });
}

}  // extern "C"
