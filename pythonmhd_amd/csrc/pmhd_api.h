// Implementation of the C ABI declared in include/pmhd.h, written once against a
// backend B that provides device memory, copies and an executor:
//   void* B::alloc(size_t); void B::release(void*);
//   void B::h2d(void*, const void*, size_t); void B::d2h(void*, const void*, size_t);
//   void B::sync(); const char* B::error();   (nullptr when no runtime error)
//   B::Exec& B::exec();
#pragma once
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "pmhd.h"
#include "pmhd_step.h"

struct pmhd_handle {
  virtual ~pmhd_handle() {}
  std::string err;
  pmhd_grid g;
  pmhd_params p;
  virtual int set_state(const double*, const double*, const double*, const double*, const double*) = 0;
  virtual int get_state(int, double*, double*, double*, double*, double*) = 0;
  virtual int compute_dt(double*) = 0;
  virtual int step(double) = 0;
  virtual int synchronize() = 0;
  virtual int diagnostics(double* totals, double* max_divb, double* max_b, int64_t* nonfinite) = 0;
  virtual int set_profiling(int on) = 0;
  virtual int get_profile(int maxn, char* names, double* ms, int64_t* launches, int64_t* elems,
                          int32_t* n) = 0;
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

namespace pmhd {

inline int64_t nfaces_of(const pmhd_grid& g, int d) {
  return (g.ny + (d == 1)) * (g.nx + (d == 0)) * (g.nz + (d == 2));
}

template <class T, class B>
struct Handle : public pmhd_handle {
  B be;
  Ctx<T> c;
  T* U[2] = {nullptr, nullptr};
  T* W[2] = {nullptr, nullptr};
  T* FB[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  std::vector<void*> owned;
  int cur = 0;
  int64_t cycles_since_set = 0;

  T* dalloc(int64_t n) {
    if (n <= 0) n = 1;
    void* ptr = be.alloc(size_t(n) * sizeof(T));
    if (!ptr) throw std::bad_alloc();
    owned.push_back(ptr);
    return static_cast<T*>(ptr);
  }

  ~Handle() override {
    for (void* ptr : owned) be.release(ptr);
  }

  int init(const pmhd_grid& g_, const pmhd_params& p_) {
    g = g_;
    p = p_;
    std::memset(&c, 0, sizeof(c));
    c.ndim = g.ndim;
    c.ny = g.ndim >= 2 ? g.ny : 1;
    c.nx = g.nx;
    c.nz = g.ndim == 3 ? g.nz : 1;
    g.ny = c.ny;
    g.nz = c.nz;
    c.bc[0] = g.bcx;
    c.bc[1] = g.bcy;
    c.bc[2] = g.bcz;
    c.bc_zrec = g.bcx;
    c.order = p.reconstruct_order;
    c.efix = p.entropy_fix;
    c.gamma = T(p.gamma);
    c.eps = T(p.epsilon);
    c.min_dens = T(p.min_dens);
    c.min_pres = T(p.min_pres);
    c.min_energy = T(p.min_energy);
    const double w[3] = {g.dx, g.dy, g.dz};
    for (int d = 0; d < 3; ++d) {
      c.width[d] = T(w[d]);
      c.wd[d] = w[d];
      c.invd[d] = T(1.0 / w[d]);
    }
    try {
      const int64_t N = c.ncell();
      int64_t maxf = 0, maxs = 0;
      for (int d = 0; d < g.ndim; ++d) {
        maxf = std::max<int64_t>(maxf, c.nface(d));
        maxs = std::max<int64_t>(maxs, c.nslope(d));
      }
      for (int s = 0; s < 2; ++s) {
        U[s] = dalloc(8 * N);
        W[s] = dalloc(8 * N);
        for (int d = 0; d < 3; ++d) FB[s][d] = dalloc(d < g.ndim && g.ndim >= 2 ? c.nface(d) : 1);
      }
      c.WL = dalloc(7 * maxf);
      c.WR = dalloc(7 * maxf);
      c.S = p.reconstruct_order > 0 ? dalloc(7 * maxs) : dalloc(1);
      for (int d = 0; d < 3; ++d) {
        const bool used = d < g.ndim;
        c.UL[d] = dalloc(used ? 7 * c.nface(d) : 1);
        c.UR[d] = dalloc(used ? 7 * c.nface(d) : 1);
        c.F[d] = dalloc(used ? 7 * c.nface(d) : 1);
        c.Fc[d] = dalloc(used && g.ndim >= 2 ? 7 * c.nface(d) : 1);
        c.nfb[d] = dalloc(used && g.ndim >= 2 ? c.nface(d) : 1);
        c.E[d] = dalloc(g.ndim == 3 || (g.ndim == 2 && d == 2) ? c.nedge() : 1);
      }
      c.Ec = dalloc(g.ndim == 3 ? 3 * N : N);
    } catch (const std::bad_alloc&) {
      return fail(PMHD_ENOMEM, "device allocation failed");
    }
    bind();
    const char* e = be.error();
    if (e) return fail(PMHD_EHIP, e);
    return PMHD_OK;
  }

  void bind() {
    const int nx = 1 - cur;
    c.U = U[cur];
    c.W = W[cur];
    c.U2 = U[nx];
    c.W2 = W[nx];
    for (int d = 0; d < 3; ++d) {
      c.fb[d] = FB[cur][d];
      c.fb2[d] = FB[nx][d];
    }
  }

  void upload(T* dst, const double* src, int64_t n) {
    if (sizeof(T) == sizeof(double)) {
      be.h2d(dst, src, size_t(n) * sizeof(double));
    } else {
      std::vector<T> tmp(static_cast<size_t>(n));
      for (int64_t i = 0; i < n; ++i) tmp[size_t(i)] = T(src[i]);
      be.h2d(dst, tmp.data(), size_t(n) * sizeof(T));
    }
  }
  void download(double* dst, const T* src, int64_t n) {
    if (sizeof(T) == sizeof(double)) {
      be.d2h(dst, src, size_t(n) * sizeof(double));
    } else {
      std::vector<T> tmp(static_cast<size_t>(n));
      be.d2h(tmp.data(), src, size_t(n) * sizeof(T));
      for (int64_t i = 0; i < n; ++i) dst[i] = double(tmp[size_t(i)]);
    }
  }

  int set_state(const double* prim, const double* cons, const double* fbx, const double* fby,
                const double* fbz) override {
    if (!prim || !cons) return fail(PMHD_EINVAL, "prim and cons are required");
    const double* fin[3] = {fbx, fby, fbz};
    for (int d = 0; d < g.ndim && g.ndim >= 2; ++d)
      if (!fin[d]) return fail(PMHD_EINVAL, "missing face array");
    const int64_t N = c.ncell();
    cur = 0;
    cycles_since_set = 0;
    bind();
    upload(U[0], cons, 8 * N);
    upload(W[0], prim, 8 * N);
    if (g.ndim >= 2)
      for (int d = 0; d < g.ndim; ++d) upload(FB[0][d], fin[d], c.nface(d));
    if (g.ndim == 1) c.bx1d = T(prim[4 * N]);
    be.sync();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int get_state(int which, double* prim, double* cons, double* fbx, double* fby,
                double* fbz) override {
    int s = cur;
    if (which == 1 && cycles_since_set > 0) s = 1 - cur;
    const int64_t N = c.ncell();
    be.sync();
    if (cons) download(cons, U[s], 8 * N);
    if (prim) download(prim, W[s], 8 * N);
    double* fo[3] = {fbx, fby, fbz};
    if (g.ndim >= 2)
      for (int d = 0; d < g.ndim; ++d)
        if (fo[d]) download(fo[d], FB[s][d], c.nface(d));
    be.sync();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int compute_dt(double* dt) override {
    double m[3] = {0, 0, 0};
    cfl_maxima(be.exec(), c, m);
    const char* e = be.error();
    if (e) return fail(PMHD_EHIP, e);
    // Source/TimestepCalculations_HelperModule.py:222-225, 285-289, 354-361
    const double cfl = p.cfl;
    if (g.ndim == 1) {
      *dt = cfl / (m[0] / g.dx);
    } else if (g.ndim == 2) {
      *dt = cfl / npmax(m[0] / g.dx, m[1] / g.dy);
    } else {
      *dt = cfl / npmax(m[0] / g.dx, npmax(m[1] / g.dy, m[2] / g.dz));
    }
    return PMHD_OK;
  }

  int step(double dt) override {
    const double w[3] = {g.dx, g.dy, g.dz};
    c.dt = T(dt);
    c.hdt = T(0.5 * dt);
    for (int d = 0; d < 3; ++d) {
      const double dtd = dt / w[d];
      const double h = 0.5 * dtd;
      c.dtd[d] = T(dtd);
      c.hdtd[d] = T(h);
      c.emfq[d] = T(h * 0.5);
    }
    {
      // Source/CTU_HelperModule.py:455: dtdy = longdouble(dt)/longdouble(dy)
      const long double hl = 0.5L * (static_cast<long double>(dt) / static_cast<long double>(g.dy));
      c.ldh[0] = static_cast<double>(hl);
      c.ldh[1] = static_cast<double>(hl - static_cast<long double>(c.ldh[0]));
    }
    cycle(be.exec(), c);
    cur = 1 - cur;
    bind();
    ++cycles_since_set;
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int synchronize() override {
    be.sync();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int diagnostics(double* totals, double* max_divb, double* max_b, int64_t* nonfinite) override {
    double sums[9], maxes[2];
    pmhd::diagnostics(be.exec(), c, sums, maxes);
    const char* e = be.error();
    if (e) return fail(PMHD_EHIP, e);
    if (totals)
      for (int q = 0; q < 8; ++q) totals[q] = sums[q];
    if (nonfinite) *nonfinite = int64_t(sums[8]);
    if (max_divb) *max_divb = g.ndim >= 2 && maxes[1] > 0 ? maxes[0] / maxes[1] : 0.0;
    if (max_b) *max_b = maxes[1];
    return PMHD_OK;
  }

  int set_profiling(int on) override {
    be.set_profiling(on != 0);
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int get_profile(int maxn, char* names, double* ms, int64_t* launches, int64_t* elems,
                  int32_t* n) override {
    *n = be.profile(maxn, names, ms, launches, elems);
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }
};

template <class B>
int create_handle(const pmhd_grid* g, const pmhd_params* p, pmhd_handle** out, std::string& err) {
  if (!g || !p || !out) {
    err = "null argument";
    return PMHD_EINVAL;
  }
  *out = nullptr;
  if (g->ndim < 1 || g->ndim > 3) { err = "ndim must be 1, 2 or 3"; return PMHD_EINVAL; }
  if (g->nx < 1 || (g->ndim >= 2 && g->ny < 1) || (g->ndim == 3 && g->nz < 1)) {
    err = "number of cells must be positive"; return PMHD_EINVAL;
  }
  if (!(g->dx > 0) || (g->ndim >= 2 && !(g->dy > 0)) || (g->ndim == 3 && !(g->dz > 0))) {
    err = "cell widths must be positive"; return PMHD_EINVAL;
  }
  const int bcs[3] = {g->bcx, g->bcy, g->bcz};
  for (int d = 0; d < g->ndim; ++d)
    if (bcs[d] != 0 && bcs[d] != 1) { err = "boundary condition must be 0 (outflow) or 1 (periodic)"; return PMHD_EINVAL; }
  if (p->reconstruct_order < 0 || p->reconstruct_order > 2) {
    err = "reconstruct_order must be 0, 1 or 2"; return PMHD_EINVAL;
  }
  if (p->precision != 64 && p->precision != 32) { err = "precision must be 64 or 32"; return PMHD_EINVAL; }
  pmhd_grid gg = *g;
  if (gg.ndim < 3) { gg.nz = 1; gg.bcz = gg.bcx; gg.dz = 1.0; }
  if (gg.ndim < 2) { gg.ny = 1; gg.bcy = gg.bcx; gg.dy = 1.0; }
  pmhd_handle* h = nullptr;
  int rc;
  if (p->precision == 64) {
    auto* hh = new (std::nothrow) Handle<double, B>();
    if (!hh) { err = "out of host memory"; return PMHD_ENOMEM; }
    rc = hh->init(gg, *p);
    h = hh;
  } else {
    auto* hh = new (std::nothrow) Handle<float, B>();
    if (!hh) { err = "out of host memory"; return PMHD_ENOMEM; }
    rc = hh->init(gg, *p);
    h = hh;
  }
  if (rc != PMHD_OK) {
    err = h->err;
    delete h;
    return rc;
  }
  *out = h;
  return PMHD_OK;
}

inline int advance(pmhd_handle* h, double t0, double t_end, int64_t cycle0, int64_t max_cycles,
                   double* t_out, int64_t* cycle_out, double* t2_out) {
  double t = t0, t2 = t0;
  int64_t cyc = cycle0;
  while (t < t_end && cyc < max_cycles) {
    t2 = t;
    double dt;
    int rc = h->compute_dt(&dt);
    if (rc) return rc;
    if (h->g.ndim == 3) {
      if (t_end - t < dt) dt = t_end - t;   // Source/EvolveFunctions_MHD.py:851-852
    } else {
      if (t + dt > t_end) dt = t_end - t;   // Source/EvolveFunctions_MHD.py:259-260, 472-473
    }
    rc = h->step(dt);
    if (rc) return rc;
    t += dt;
    ++cyc;
  }
  if (t_out) *t_out = t;
  if (cycle_out) *cycle_out = cyc;
  if (t2_out) *t2_out = t2;
  return PMHD_OK;
}

}  // namespace pmhd

// C ABI bodies; the including translation unit defines PMHD_BACKEND.
#define PMHD_DEFINE_C_ABI(BACKEND)                                                          \
  static std::string g_pmhd_create_error;                                                   \
  extern "C" int pmhd_create(const pmhd_grid* g, const pmhd_params* p, pmhd_handle** out) { \
    try {                                                                                   \
      return pmhd::create_handle<BACKEND>(g, p, out, g_pmhd_create_error);                  \
    } catch (...) {                                                                         \
      g_pmhd_create_error = "unexpected exception in pmhd_create";                          \
      return PMHD_EINVAL;                                                                   \
    }                                                                                       \
  }                                                                                         \
  extern "C" int pmhd_destroy(pmhd_handle* h) {                                             \
    delete h;                                                                               \
    return PMHD_OK;                                                                         \
  }                                                                                         \
  extern "C" int pmhd_set_state(pmhd_handle* h, const double* prim, const double* cons,     \
                                const double* fbx, const double* fby, const double* fbz) {  \
    if (!h) return PMHD_EINVAL;                                                             \
    try { return h->set_state(prim, cons, fbx, fby, fbz); }                                 \
    catch (...) { return h->fail(PMHD_EINVAL, "exception in pmhd_set_state"); }             \
  }                                                                                         \
  extern "C" int pmhd_get_state(pmhd_handle* h, int32_t which, double* prim, double* cons,  \
                                double* fbx, double* fby, double* fbz) {                    \
    if (!h) return PMHD_EINVAL;                                                             \
    try { return h->get_state(which, prim, cons, fbx, fby, fbz); }                          \
    catch (...) { return h->fail(PMHD_EINVAL, "exception in pmhd_get_state"); }             \
  }                                                                                         \
  extern "C" int pmhd_compute_dt(pmhd_handle* h, double* dt) {                              \
    if (!h || !dt) return PMHD_EINVAL;                                                      \
    return h->compute_dt(dt);                                                               \
  }                                                                                         \
  extern "C" int pmhd_step(pmhd_handle* h, double dt) {                                     \
    if (!h) return PMHD_EINVAL;                                                             \
    return h->step(dt);                                                                     \
  }                                                                                         \
  extern "C" int pmhd_advance(pmhd_handle* h, double t0, double t_end, int64_t cycle0,      \
                              int64_t max_cycles, double* t_out, int64_t* cycle_out,        \
                              double* t2) {                                                 \
    if (!h) return PMHD_EINVAL;                                                             \
    return pmhd::advance(h, t0, t_end, cycle0, max_cycles, t_out, cycle_out, t2);           \
  }                                                                                         \
  extern "C" int pmhd_synchronize(pmhd_handle* h) {                                         \
    if (!h) return PMHD_EINVAL;                                                             \
    return h->synchronize();                                                                \
  }                                                                                         \
  extern "C" int pmhd_diagnostics(pmhd_handle* h, double* totals, double* max_divb,          \
                                  double* max_b, int64_t* nonfinite) {                      \
    if (!h) return PMHD_EINVAL;                                                             \
    return h->diagnostics(totals, max_divb, max_b, nonfinite);                              \
  }                                                                                         \
  extern "C" int pmhd_set_profiling(pmhd_handle* h, int32_t enable) {                       \
    if (!h) return PMHD_EINVAL;                                                             \
    return h->set_profiling(enable);                                                        \
  }                                                                                         \
  extern "C" int pmhd_get_profile(pmhd_handle* h, int32_t max_entries, char* names,         \
                                  double* total_ms, int64_t* launches, int64_t* elements,   \
                                  int32_t* n_entries) {                                     \
    if (!h || !names || !total_ms || !launches || !elements || !n_entries) return PMHD_EINVAL; \
    return h->get_profile(max_entries, names, total_ms, launches, elements, n_entries);     \
  }                                                                                         \
  extern "C" const char* pmhd_last_error(const pmhd_handle* h) {                            \
    return h ? h->err.c_str() : g_pmhd_create_error.c_str();                                \
  }
