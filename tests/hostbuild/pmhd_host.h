// Host (CPU) build of the product's kernel bodies and C ABI -- TEST
// INFRASTRUCTURE ONLY.  It compiles pythonmhd_amd/csrc/*.h with g++ so the
// FP64 arithmetic of the HIP kernels can be checked against the oracle in a
// container without a GPU.  The product package never loads this library.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "pmhd_adjoint.h"

PMHD_NS_BEGIN

template <class Ph, class T>
void host_each(const Ctx<T>& c) {
  int64_t n[3];
  c.extent(Ph::kind, Ph::dir, n);
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t j = 0; j < n[0]; ++j)
    for (int64_t i = 0; i < n[1]; ++i)
      for (int64_t k = 0; k < n[2]; ++k) Ph()(c, I3{j, i, k});
}

inline void max_bits(uint64_t m[3], const double o[3]) {
  for (int d = 0; d < 3; ++d) {
    uint64_t b;
    std::memcpy(&b, &o[d], 8);
    if (b > m[d]) m[d] = b;
  }
}

// a phase that also yields speeds, maximised over its elements (serially, so the
// phase's writes need no other ordering)
template <class Ph, class T>
void host_each_max(const Ctx<T>& c, uint64_t m[3]) {
  int64_t n[3];
  c.extent(Ph::kind, Ph::dir, n);
  for (int64_t j = 0; j < n[0]; ++j)
    for (int64_t i = 0; i < n[1]; ++i)
      for (int64_t k = 0; k < n[2]; ++k) {
        double o[3];
        Ph()(c, I3{j, i, k}, o);
        max_bits(m, o);
      }
}

template <class T>
void host_cfl(const Ctx<T>& c, double out[3]) {
  uint64_t m[3] = {0, 0, 0};
  host_each_max<CflSpeeds>(c, m);
  for (int d = 0; d < 3; ++d) std::memcpy(&out[d], &m[d], 8);
}

struct HostExec {
  uint64_t last[3] = {0, 0, 0};
  template <class Ph, class T>
  void run(const Ctx<T>& c) {
    host_each<Ph>(c);
  }
  // (a GPU executor overlaps these launches with the following ones)
  template <class Ph, class T>
  void run_side(const Ctx<T>& c) {
    run<Ph>(c);
  }
  void join_side() {}
  template <class Ph, class T>
  void run_max(const Ctx<T>& c) {
    for (int d = 0; d < 3; ++d) last[d] = 0;
    if constexpr (Ph::id == K_UPDATE) {
      if (c.fcnt) fold_floor_counts(c.fcnt);
    }
    host_each_max<Ph>(c, last);
  }
  void max_partials(double out[3]) {
    for (int d = 0; d < 3; ++d) std::memcpy(&out[d], &last[d], 8);
  }
  template <class T>
  void cfl(const Ctx<T>& c, double out[3]) {
    host_cfl(c, out);
  }
  template <class F>
  void reduce(int, int64_t n, F f, double sums[9], double maxes[2]) {
    for (int q = 0; q < 9; ++q) sums[q] = 0;
    maxes[0] = maxes[1] = 0;
    for (int64_t i = 0; i < n; ++i) {
      double a[9], b[2];
      f(i, a, b);
      for (int q = 0; q < 9; ++q) sums[q] += a[q];
      maxes[0] = npmax(maxes[0], b[0]);
      maxes[1] = npmax(maxes[1], b[1]);
    }
  }
};

// the batch executor of the host build: the same phases, box after box
struct HostBatchExec {
  BatchLayout* lay = nullptr;
  bool tuning = false;
  int tune_reps = 1;

  template <class T>
  Ctx<T>* ctxs() const {
    return static_cast<Ctx<T>*>(lay->dctx);
  }
  // walk the launch exactly as the HIP batch kernel does (blocks striding over the
  // tile list, threads of a tile through locate_tile) and check that every element
  // of every active box is visited once
  // tuning on the host walks the phase with every candidate tile (checking each
  // walk) and keeps a candidate that varies with the launch id, so later launches
  // exercise many shapes
  template <class Ph, class T>
  void run_side(const Ctx<T>& c0) {
    run<Ph>(c0);
  }
  void join_side() {}
  template <class Ph, class T>
  void run(const Ctx<T>& c0) {
    if (tuning) {
      const std::vector<TileCfg> cands = tile_candidates(*lay, Ph::id);
      for (const TileCfg& t : cands)
        if (lay->apply_tile(Ph::id, t)) walk<Ph, T>();
      (void)lay->apply_tile(Ph::id, cands[size_t(Ph::id) % cands.size()]);
      lay->tuned_ms[Ph::id] = 0.f;
    }
    (void)c0;
    walk<Ph, T>();
  }

  template <class Ph, class T>
  void run_max(const Ctx<T>& c0) {
    if (tuning) {
      const std::vector<TileCfg> cands = tile_candidates(*lay, Ph::id);
      for (const TileCfg& t : cands)
        if (lay->apply_tile(Ph::id, t)) walk<Ph, T, true>();
      (void)lay->apply_tile(Ph::id, cands[size_t(Ph::id) % cands.size()]);
      lay->tuned_ms[Ph::id] = 0.f;
    }
    (void)c0;
    if constexpr (Ph::id == K_UPDATE) {
      for (int b = 0; b < lay->nbox; ++b) {
        const Ctx<T>& c = ctxs<T>()[b];
        if (c.active && c.fcnt) fold_floor_counts(c.fcnt);
      }
    }
    walk<Ph, T, true>();
  }

  template <class Ph, class T, bool kMax = false>
  void walk() {
    const TileCfg t = lay->tiles[Ph::id];
    const int total = lay->total[Ph::id];
    const int grid = t.grid > 0 && t.grid < total ? t.grid : total;
    const int* prefix = lay->dprefix + Ph::id * (lay->nbox + 1);
    std::vector<std::vector<unsigned char>> seen(static_cast<size_t>(lay->nbox));
    for (int b = 0; b < lay->nbox; ++b) {
      int64_t n[3];
      ctxs<T>()[b].extent(Ph::kind, Ph::dir, n);
      seen[size_t(b)].assign(size_t(n[0] * n[1] * n[2]), 0);
    }
    for (int blk = 0; blk < grid; ++blk)
      for (int vb = blk; vb < total; vb += grid)
        for (int tid = 0; tid < t.bk * t.bi; ++tid) {
          const TileRef r = locate_tile(ctxs<T>(), prefix, lay->nbox, vb, Ph::kind, Ph::dir, t.bk, t.bi, tid);
          const Ctx<T>& c = ctxs<T>()[r.box];
          if (!c.active || !r.in) continue;
          call<Ph, kMax>(c, I3{r.j, r.i, r.k}, lay->dred + 3 * r.box);
          int64_t n[3];
          c.extent(Ph::kind, Ph::dir, n);
          seen[size_t(r.box)][size_t((int64_t(r.j) * n[1] + r.i) * n[2] + r.k)] += 1;
        }
    for (int b = 0; b < lay->nbox; ++b) {
      if (!ctxs<T>()[b].active) continue;
      for (unsigned char v : seen[size_t(b)])
        if (v != 1) throw std::runtime_error(std::string("tile walk of ") + kernel_name(Ph::id) + " not exactly once");
    }
  }
  template <class Ph, bool kMax, class T>
  static void call(const Ctx<T>& c, I3 p, unsigned long long* red) {
    if constexpr (kMax) {
      double o[3];
      Ph()(c, p, o);
      uint64_t m[3] = {red[0], red[1], red[2]};
      max_bits(m, o);
      for (int d = 0; d < 3; ++d) red[d] = m[d];
    } else {
      (void)red;
      Ph()(c, p);
    }
  }

  template <class T>
  void cfl(const Ctx<T>&) {
    for (int b = 0; b < lay->nbox; ++b) {
      const Ctx<T>& c = ctxs<T>()[b];
      if (!c.active) continue;
      double m[3];
      host_cfl(c, m);
      for (int d = 0; d < 3; ++d) {
        uint64_t v;
        std::memcpy(&v, &m[d], 8);
        if (v > lay->dred[3 * b + d]) lay->dred[3 * b + d] = v;
      }
    }
  }
  template <class T>
  void begin(double cfl, double t_end, int64_t max_cycles, int commit) {
    for (int b = 0; b < lay->nbox; ++b) {
      double m[3];
      for (int d = 0; d < 3; ++d) {
        uint64_t v = lay->dred[3 * b + d];
        std::memcpy(&m[d], &v, 8);
        lay->dred[3 * b + d] = 0;
      }
      box_begin(ctxs<T>()[b], lay->dclk[b], m, cfl, t_end, max_cycles, commit);
    }
  }
  template <class T>
  void end() {
    for (int b = 0; b < lay->nbox; ++b) box_end(ctxs<T>()[b], lay->dclk[b]);
  }
  template <class T>
  void activate_all() {
    for (int b = 0; b < lay->nbox; ++b) ctxs<T>()[b].active = 1;
  }
};

template <class T>
int riemann_run(int d, int solver, int64_t n, const double* wl, const double* wr, const double* ul,
                const double* ur, const double* bn, double gamma, int efix, double eps, double* F,
                std::string&) {
  for (int64_t i = 0; i < n; ++i) riemann_face<T>(d, solver, n, i, wl, wr, ul, ur, bn, T(gamma), efix != 0, T(eps), F);
  return PMHD_OK;
}

struct HostBackend {
  HostExec ex;
  HostBatchExec bex;
  HostExec& exec() { return ex; }
  HostBatchExec& batch_exec(BatchLayout* lay) {
    bex.lay = lay;
    return bex;
  }
  int compute_units() const { return 1; }
  void* alloc(size_t bytes) { return std::calloc(bytes, 1); }
  void release(void* p) { std::free(p); }
  void h2d(void* d, const void* s, size_t b) { std::memcpy(d, s, b); }
  void d2h(void* d, const void* s, size_t b) { std::memcpy(d, s, b); }
  void zero(void* d, size_t b) { std::memset(d, 0, b); }
  void d2d(void* d, const void* s, size_t b) { std::memcpy(d, s, b); }
  // "device" float64 arrays of the device entry points are host memory here
  template <class T>
  void dev_in(T* dst, const double* src, int64_t n) {
    for (int64_t i = 0; i < n; ++i) dst[i] = T(src[i]);
  }
  template <class T>
  void dev_out(double* dst, const T* src, int64_t n) {
    for (int64_t i = 0; i < n; ++i) dst[i] = double(src[i]);
  }
  template <class T>
  void dev_split(T* hi, T* lo, const double* src, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
      hi[i] = T(src[i]);
      lo[i] = T(src[i] - double(hi[i]));
    }
  }
  template <class T>
  void dev_join(double* dst, const T* hi, const T* lo, int64_t n) {
    for (int64_t i = 0; i < n; ++i) dst[i] = double(hi[i]) + double(lo[i]);
  }
  void set_stream(void*) {}
  bool overflow = false;
  template <class Ph, class T>
  void adjoint_phase(const Ctx<T>& c, const Ctx<double>& a) {
    int64_t n[3];
    c.extent(Ph::kind, Ph::dir, n);
    AdState st;
    for (int64_t j = 0; j < n[0]; ++j)
      for (int64_t i = 0; i < n[1]; ++i)
        for (int64_t k = 0; k < n[2]; ++k)
          if (!adjoint_element<Ph>(c, a, I3{j, i, k}, st, [](double* p, double v) {
                if (v != 0.0) *p += v;
              }))
            overflow = true;
  }
  bool adjoint_overflow() const { return overflow; }
  // one thread: the contributions already add up in a fixed order
  void set_deterministic(bool) {}
  void set_arena(double*, int64_t) {}
  void adjoint_select(int, size_t) {}
  void adjoint_end() {}
  void sync() {}
  const char* error() { return nullptr; }
  void set_profiling(bool) {}
  int profile(int, char*, double*, int64_t*, int64_t*) { return 0; }
};

PMHD_NS_END

