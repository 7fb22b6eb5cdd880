// Implementation of the C ABI declared in include/pmhd.h, written once against a
// backend B that provides device memory, copies and executors:
//   void* B::alloc(size_t); void B::release(void*); void B::zero(void*, size_t);
//   void B::h2d(void*, const void*, size_t); void B::d2h(void*, const void*, size_t);
//   void B::sync(); const char* B::error();   (nullptr when no runtime error)
//   B::exec()                 -- single-grid executor
//   B::batch_exec(BatchLayout*) -- executor over all boxes of a batch
#pragma once
#include <algorithm>
#include <cstring>
#include <functional>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "pmhd.h"
#include "pmhd_hydro.h"

struct pmhd_handle {
  virtual ~pmhd_handle() {}
  std::string err;
  pmhd_grid g;
  pmhd_params p;
  virtual int init(const pmhd_grid&, const pmhd_params&) = 0;
  virtual int set_state(const double*, const double*, const double*, const double*, const double*) = 0;
  virtual int get_state(int, double*, double*, double*, double*, double*) = 0;
  virtual int compute_dt(double*) = 0;
  virtual int step(double) = 0;
  virtual int synchronize() = 0;
  virtual int diagnostics(double* totals, double* max_divb, double* max_b, int64_t* nonfinite) = 0;
  virtual int set_profiling(int on) = 0;
  virtual int get_profile(int maxn, char* names, double* ms, int64_t* launches, int64_t* elems,
                          int32_t* n) = 0;
  virtual int set_sweep(const int32_t seg[3]) = 0;
  virtual int get_stats(pmhd_stats* out) = 0;
  virtual int set_state_device(const double*, const double*, const double*, const double*, const double*) = 0;
  virtual int get_state_device(int, double*, double*, double*, double*, double*) = 0;
  virtual int set_stream(void* stream) = 0;
  // the evolve loop with the clock on the device (no host round trip per cycle)
  virtual int advance_clocked(double t0, double t_end, int64_t cycle0, int64_t max_cycles, double* t_out,
                              int64_t* cycle_out, double* t2_out) = 0;
  int clock = 0;   // PMHD_CLOCK_AUTO, PMHD_CLOCK_HOST or PMHD_CLOCK_DEVICE
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

// The message pmhd_last_error(NULL) returns after a failed call without a handle
// (pmhd_cfl_dt_noalloc): a string literal, recorded without allocating.  Defined
// with the C ABI.
void pmhd_note_null_error(const char* literal);

// Constructors of the handles of one precision (FP64 or FP32 bulk), each defined
// in the translation unit compiled for that precision; the result still needs
// init().  nullptr when out of host memory.
pmhd_handle* pmhd_new_handle_f64();
pmhd_handle* pmhd_new_handle_f32();
// the FP32 step built with non-temporal result stores, for grids whose launch
// outputs are far larger than the caches (pmhd_create picks it by size)
pmhd_handle* pmhd_new_handle_f32nt();
struct pmhd_batch;
pmhd_batch* pmhd_new_batch_f64();
pmhd_batch* pmhd_new_batch_f32();
// the Riemann flux of n faces in the precision's arithmetic (pmhd_riemann_flux)
int pmhd_riemann_f64(int d, int solver, int64_t n, const double* wl, const double* wr, const double* ul,
                     const double* ur, const double* bn, double gamma, int efix, double eps, double* F,
                     std::string& err);
int pmhd_riemann_f32(int d, int solver, int64_t n, const double* wl, const double* wr, const double* ul,
                     const double* ur, const double* bn, double gamma, int efix, double eps, double* F,
                     std::string& err);

struct pmhd_batch {
  virtual ~pmhd_batch() {}
  std::string err;
  int nbox = 0;
  virtual int init(int n, const pmhd_grid* grids, const pmhd_params& p) = 0;
  virtual int set_state(int box, const double*, const double*, const double*, const double*, const double*) = 0;
  virtual int get_state(int box, int which, double*, double*, double*, double*, double*) = 0;
  virtual int advance(double t0, double t_end, int64_t cycle0, int64_t max_cycles, double* t_out,
                      int64_t* cycle_out, double* t2_out) = 0;
  virtual int autotune(int reps) = 0;
  virtual int get_tiles(int maxn, char* names, int32_t* bk, int32_t* bi, int32_t* grid, double* ms,
                        int32_t* n) = 0;
  virtual int synchronize() = 0;
  virtual int set_profiling(int on) = 0;
  virtual int get_profile(int maxn, char* names, double* ms, int64_t* launches, int64_t* elems,
                          int32_t* n) = 0;
  virtual int get_stats(int box, pmhd_stats* out) = 0;
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

PMHD_NS_BEGIN

// floor counters (device words, FLOOR_WORDS of them) -> the ABI's statistics
inline void floor_stats(const unsigned long long* f, pmhd_stats* out) {
  out->floored_total = 0;
  for (int s = 0; s < FLOOR_SITES; ++s)
    for (int v = 0; v < 2; ++v) {
      out->floored[s][v] = int64_t(f[FLOOR_TOT + 2 * s + v]);
      out->floored_max[s][v] = int64_t(f[FLOOR_MAX + 2 * s + v]);
      out->floored_total += out->floored[s][v];
    }
}

// ---------------------------------------------------------------------------
// batch launch tables: per launch id, every box's index box and the tile shape
// (bk threads along z by bi along x) and grid; prefix[id][b] = first tile of box b
// ---------------------------------------------------------------------------
struct TileCfg {
  int bk, bi;
  int grid;   // blocks of the launch; 0 = one block per tile
};

struct BatchLayout {
  int nbox = 0;
  int ndim = 0;
  int cus = 256;
  int64_t max_n2 = 1;                 // longest contiguous extent over launches and boxes
  void* dctx = nullptr;               // device Ctx<T>[nbox]
  BoxClock* dclk = nullptr;           // device clocks
  unsigned long long* dred = nullptr; // device CFL maxima, 3 per box
  int* dprefix = nullptr;             // device [K_COUNT][nbox + 1]
  std::vector<int64_t> ext;           // [K_COUNT][nbox][3]
  std::vector<int> hprefix;           // host copy of the prefix tables
  TileCfg tiles[K_COUNT];
  int total[K_COUNT] = {};
  int64_t elems[K_COUNT] = {};
  float tuned_ms[K_COUNT];
  int line_block = 64;                // threads per workgroup of the line sweeps
  std::function<void(int)> upload;    // copies hprefix[id] to the device

  BatchLayout() {
    for (int id = 0; id < K_COUNT; ++id) {
      tiles[id] = TileCfg{1, 256, 0};
      tuned_ms[id] = -1.f;
    }
  }

  // use tile t for launch id; false (and no change) when its tile count would
  // not fit the 32-bit tile index of the batch kernels
  bool apply_tile(int id, TileCfg t) {
    int64_t acc = 0, el = 0;
    for (int b = 0; b < nbox; ++b) {
      const int64_t* e = ext.data() + (size_t(id) * size_t(nbox) + size_t(b)) * 3;
      acc += e[0] * ((e[1] + t.bi - 1) / t.bi) * ((e[2] + t.bk - 1) / t.bk);
      el += e[0] * e[1] * e[2];
    }
    if (acc > int64_t(std::numeric_limits<int>::max())) return false;
    tiles[id] = t;
    int* pre = hprefix.data() + size_t(id) * size_t(nbox + 1);
    acc = 0;
    for (int b = 0; b < nbox; ++b) {
      pre[b] = int(acc);
      const int64_t* e = ext.data() + (size_t(id) * size_t(nbox) + size_t(b)) * 3;
      acc += e[0] * ((e[1] + t.bi - 1) / t.bi) * ((e[2] + t.bk - 1) / t.bk);
    }
    pre[nbox] = int(acc);
    total[id] = int(acc);
    elems[id] = el;
    if (upload) upload(id);
    return true;
  }
};

// block of 256 threads, as wide along z as the boxes allow
inline TileCfg default_tile(const BatchLayout& lay) {
  int bk = 1;
  while (bk < 64 && bk < lay.max_n2) bk <<= 1;
  return TileCfg{bk, 256 / bk, 0};
}

// the shapes the tile tuner tries: 64-256 threads, one block per tile or a
// persistent grid of a few blocks per compute unit
inline std::vector<TileCfg> tile_candidates(const BatchLayout& lay, int) {
  std::vector<std::pair<int, int>> shapes;
  if (lay.max_n2 <= 1) {
    shapes = {{1, 64}, {1, 128}, {1, 256}};
  } else {
    shapes = {{64, 1}, {64, 2}, {64, 4}, {32, 4}, {32, 8}, {16, 8}, {16, 16}};
  }
  std::vector<TileCfg> out;
  for (const auto& s : shapes) {
    out.push_back(TileCfg{s.first, s.second, 0});
    out.push_back(TileCfg{s.first, s.second, 8 * lay.cus});
  }
  return out;
}

// ---------------------------------------------------------------------------
// grid set-up shared by single grids and batch boxes
// ---------------------------------------------------------------------------
template <class T>
void init_ctx(Ctx<T>& c, const pmhd_grid& g, const pmhd_params& p) {
  std::memset(&c, 0, sizeof(c));
  c.ndim = g.ndim;
  c.active = 1;
  c.physics = p.physics;
  c.ny = g.ndim >= 2 ? g.ny : 1;
  c.nx = g.nx;
  c.nz = g.ndim == 3 ? g.nz : 1;
  c.bc[0] = g.bcx;
  c.bc[1] = g.bcy;
  c.bc[2] = g.bcz;
  c.fixq = p.fix_quirks;
  c.riemann = p.riemann;
  // the reference's z reconstruction uses the x boundary condition
  c.bc_zrec = (p.fix_quirks & PMHD_FIX_Z_BC) ? g.bcz : g.bcx;
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
}

// Faces per thread of the fused predictor sweep along direction d (Ctx::seg).  In
// FP32 across the contiguous axis: the largest power of two, at most 16, that still
// leaves about 2048 threads per compute unit for nfaces faces, and at least 2 (a run
// of one face reconstructs two cells); long runs amortise the cell reconstructed
// before each run.  Along the contiguous axis (z in 3D, x in 2D), where the threads
// of a run launch would read addresses a run apart: one workgroup per line (-1) for
// lines of at least 64 cells, else runs of 2.  In FP64 the fused kernels need more
// registers than they save memory traffic, so the separate launches are used (0).
inline bool line_axis(int ndim, int d) { return (ndim == 3 && d == 2) || (ndim == 2 && d == 0); }
inline int sweep_segment(int ndim, int d, int64_t nfaces, int64_t n_along, int cus, bool fp64) {
  if (fp64) return 0;
  if (line_axis(ndim, d)) return n_along >= 64 ? -1 : 2;
  const int64_t target = int64_t(cus > 0 ? cus : 256) * 2048;
  int s = 2;
  while (s < 16 && nfaces / (2 * s) >= target) s *= 2;
  return s;
}

// the two state buffers of a grid (the update writes the other one)
template <class T>
struct StateBufs {
  T* U[2] = {nullptr, nullptr};
  T* W[2] = {nullptr, nullptr};
  T* FB[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  T* FBLO[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};   // split faces only
};

// cell variables and interface-state variables of a physics
inline int cell_vars(int physics) { return physics == 1 ? 5 : 8; }
inline int face_vars(int physics) { return physics == 1 ? 5 : 7; }

// every device buffer of one grid; alloc(n) returns room for n values of T
template <class T, class A>
void alloc_grid(Ctx<T>& c, StateBufs<T>& s, int order, A&& alloc) {
  const int nd = c.ndim;
  const int64_t N = c.ncell();
  const int64_t cv = cell_vars(c.physics), fv = face_vars(c.physics);
  const bool mhd = c.physics == 0;
  int64_t maxf = 0, maxs = 0;
  for (int d = 0; d < nd; ++d) {
    maxf = std::max<int64_t>(maxf, c.nface(d));
    maxs = std::max<int64_t>(maxs, c.nslope(d));
  }
  for (int b = 0; b < 2; ++b) {
    s.U[b] = alloc(cv * N);
    s.W[b] = alloc(cv * N);
    for (int d = 0; d < 3; ++d) {
      const int64_t nf = mhd && d < nd && nd >= 2 ? c.nface(d) : 1;
      s.FB[b][d] = alloc(nf);
      s.FBLO[b][d] = alloc(split_faces<T>() ? nf : 1);
    }
  }
  c.WL = alloc(fv * maxf);
  c.WR = alloc(fv * maxf);
  c.S = order > 0 ? alloc(fv * maxs) : alloc(1);
  for (int d = 0; d < 3; ++d) {
    const bool used = d < nd;
    c.UL[d] = alloc(used ? fv * c.nface(d) : 1);
    c.UR[d] = alloc(used ? fv * c.nface(d) : 1);
    c.F[d] = alloc(used ? fv * c.nface(d) : 1);
    c.Fc[d] = alloc(used && nd >= 2 ? fv * c.nface(d) : 1);
    c.nfb[d] = alloc(mhd && used && nd >= 2 ? c.nface(d) : 1);
    c.nfblo[d] = alloc(split_faces<T>() && mhd && used && nd >= 2 ? c.nface(d) : 1);
    c.E[d] = alloc(mhd && (nd == 3 || (nd == 2 && d == 2)) ? c.nedge() : 1);
  }
  c.Ec = alloc(mhd ? (nd == 3 ? 3 * N : N) : 1);
  c.fcnt = reinterpret_cast<unsigned long long*>(alloc(int64_t(FLOOR_WORDS * sizeof(unsigned long long) / sizeof(T))));
  c.zend = alloc(mhd && nd >= 2 ? 14 * (c.ncell() / c.n_along(nd == 3 ? 2 : 0)) : 1);
}

// point the context at buffer set `cur` (the update writes 1 - cur)
template <class T>
void bind_bufs(Ctx<T>& c, const StateBufs<T>& s, int cur) {
  const int nx = 1 - cur;
  c.U = s.U[cur];
  c.W = s.W[cur];
  c.U2 = s.U[nx];
  c.W2 = s.W[nx];
  for (int d = 0; d < 3; ++d) {
    c.fb[d] = s.FB[cur][d];
    c.fb2[d] = s.FB[nx][d];
    c.fblo[d] = s.FBLO[cur][d];
    c.fb2lo[d] = s.FBLO[nx][d];
  }
}

// host float64 <-> device T
template <class T, class B>
void upload_vals(B& be, T* dst, const double* src, int64_t n) {
  if (sizeof(T) == sizeof(double)) {
    be.h2d(dst, src, size_t(n) * sizeof(double));
  } else {
    std::vector<T> tmp(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) tmp[size_t(i)] = T(src[i]);
    be.h2d(dst, tmp.data(), size_t(n) * sizeof(T));
  }
}
template <class T, class B>
void download_vals(B& be, double* dst, const T* src, int64_t n) {
  if (sizeof(T) == sizeof(double)) {
    be.d2h(dst, src, size_t(n) * sizeof(double));
  } else {
    std::vector<T> tmp(static_cast<size_t>(n));
    be.d2h(tmp.data(), src, size_t(n) * sizeof(T));
    for (int64_t i = 0; i < n; ++i) dst[i] = double(tmp[size_t(i)]);
  }
}

// load / store the reference-layout state of one grid into buffer set `s`
template <class T, class B>
int load_state(B& be, Ctx<T>& c, const pmhd_grid& g, const double* prim, const double* cons,
               const double* const fin[3], T* U, T* W, T* const FB[3], T* const FBLO[3], std::string& err) {
  if (!prim || !cons) {
    err = "prim and cons are required";
    return PMHD_EINVAL;
  }
  const bool mhd = c.physics == 0;
  for (int d = 0; mhd && d < g.ndim && g.ndim >= 2; ++d)
    if (!fin[d]) {
      err = "missing face array";
      return PMHD_EINVAL;
    }
  const int64_t N = c.ncell();
  upload_vals(be, U, cons, cell_vars(c.physics) * N);
  upload_vals(be, W, prim, cell_vars(c.physics) * N);
  if (mhd && g.ndim >= 2)
    for (int d = 0; d < g.ndim; ++d) {
      if (split_faces<T>()) {
        // high part: the float rounding; low part: the rounding of the remainder
        const int64_t nf = c.nface(d);
        std::vector<double> lo(static_cast<size_t>(nf));
        for (int64_t i = 0; i < nf; ++i) lo[size_t(i)] = fin[d][i] - double(T(fin[d][i]));
        upload_vals(be, FB[d], fin[d], nf);
        upload_vals(be, FBLO[d], lo.data(), nf);
      } else {
        upload_vals(be, FB[d], fin[d], c.nface(d));
      }
    }
  if (mhd && g.ndim == 1) c.bx1d = T(prim[4 * N]);
  return PMHD_OK;
}

// the same from / to device float64 arrays of the reference layout (conversions on
// the device, on the backend's stream)
template <class T, class B>
int load_state_dev(B& be, Ctx<T>& c, const pmhd_grid& g, const double* prim, const double* cons,
                   const double* const fin[3], T* U, T* W, T* const FB[3], T* const FBLO[3], std::string& err) {
  if (!prim || !cons) {
    err = "prim and cons are required";
    return PMHD_EINVAL;
  }
  const bool mhd = c.physics == 0;
  for (int d = 0; mhd && d < g.ndim && g.ndim >= 2; ++d)
    if (!fin[d]) {
      err = "missing face array";
      return PMHD_EINVAL;
    }
  const int64_t N = c.ncell();
  be.dev_in(U, cons, cell_vars(c.physics) * N);
  be.dev_in(W, prim, cell_vars(c.physics) * N);
  if (mhd && g.ndim >= 2)
    for (int d = 0; d < g.ndim; ++d) {
      if (split_faces<T>()) be.dev_split(FB[d], FBLO[d], fin[d], c.nface(d));
      else be.dev_in(FB[d], fin[d], c.nface(d));
    }
  if (mhd && g.ndim == 1) {
    double bx = 0.0;
    be.d2h(&bx, prim + 4 * N, sizeof(double));
    c.bx1d = T(bx);
  }
  return PMHD_OK;
}

template <class T, class B>
void store_state_dev(B& be, const Ctx<T>& c, const pmhd_grid& g, const T* U, const T* W, const T* const FB[3],
                     const T* const FBLO[3], double* prim, double* cons, double* const fo[3]) {
  const int64_t N = c.ncell();
  if (cons) be.dev_out(cons, U, cell_vars(c.physics) * N);
  if (prim) be.dev_out(prim, W, cell_vars(c.physics) * N);
  if (c.physics == 0 && g.ndim >= 2)
    for (int d = 0; d < g.ndim; ++d)
      if (fo[d]) {
        if (split_faces<T>()) be.dev_join(fo[d], FB[d], FBLO[d], c.nface(d));
        else be.dev_out(fo[d], FB[d], c.nface(d));
      }
}

template <class T, class B>
void store_state(B& be, const Ctx<T>& c, const pmhd_grid& g, const T* U, const T* W, const T* const FB[3],
                 const T* const FBLO[3], double* prim, double* cons, double* const fo[3]) {
  const int64_t N = c.ncell();
  if (cons) download_vals(be, cons, U, cell_vars(c.physics) * N);
  if (prim) download_vals(be, prim, W, cell_vars(c.physics) * N);
  if (c.physics == 0 && g.ndim >= 2)
    for (int d = 0; d < g.ndim; ++d)
      if (fo[d]) {
        download_vals(be, fo[d], FB[d], c.nface(d));
        if (split_faces<T>()) {
          std::vector<double> lo(static_cast<size_t>(c.nface(d)));
          download_vals(be, lo.data(), FBLO[d], c.nface(d));
          for (int64_t i = 0; i < c.nface(d); ++i) fo[d][i] += lo[size_t(i)];
        }
      }
}

// ---------------------------------------------------------------------------
// device-clocked runs: the launch tables of a batch of boxes (also used for one
// grid whose clock stays on the device) and the loop that advances them
// ---------------------------------------------------------------------------
// Allocate (once: a layout built again for the same boxes keeps its memory) the
// device contexts, clocks, CFL maxima and tile tables of the boxes hctx (bound to
// their buffers), fill the tables with the default tiles and copy the contexts to
// the device.  dalloc(bytes) returns device memory or throws std::bad_alloc.
// false when a launch would exceed 2^31 tiles.
template <class T, class B, class A>
bool init_layout(B& be, BatchLayout& lay, std::vector<Ctx<T>>& hctx, A&& dalloc) {
  const int n = int(hctx.size());
  if (!lay.dctx || lay.nbox != n) {
    lay.dctx = dalloc(sizeof(Ctx<T>) * size_t(n));
    lay.dclk = static_cast<BoxClock*>(dalloc(sizeof(BoxClock) * size_t(n)));
    lay.dred = static_cast<unsigned long long*>(dalloc(3 * sizeof(unsigned long long) * size_t(n)));
    lay.dprefix = static_cast<int*>(dalloc(sizeof(int) * size_t(K_COUNT) * size_t(n + 1)));
  }
  lay.nbox = n;
  be.zero(lay.dred, 3 * sizeof(unsigned long long) * size_t(n));
  lay.ndim = hctx[0].ndim;
  lay.cus = be.compute_units();
  lay.ext.assign(size_t(K_COUNT) * size_t(n) * 3, 0);
  lay.max_n2 = 1;
  for (int id = 0; id < K_COUNT; ++id) {
    int kind, dir;
    phase_shape(id, kind, dir);
    for (int b = 0; b < n; ++b) {
      int64_t e[3];
      hctx[size_t(b)].extent(kind, dir, e);
      for (int q = 0; q < 3; ++q) lay.ext[(size_t(id) * size_t(n) + size_t(b)) * 3 + size_t(q)] = e[q];
      // launches along a direction the boxes do not have never run
      if (dir < lay.ndim) lay.max_n2 = std::max<int64_t>(lay.max_n2, e[2]);
    }
  }
  lay.hprefix.assign(size_t(K_COUNT) * size_t(n + 1), 0);
  BatchLayout* L = &lay;
  B* bp = &be;
  lay.upload = [L, bp](int id) {
    bp->h2d(L->dprefix + size_t(id) * size_t(L->nbox + 1), L->hprefix.data() + size_t(id) * size_t(L->nbox + 1),
            sizeof(int) * size_t(L->nbox + 1));
  };
  for (int id = 0; id < K_COUNT; ++id) {
    int kind, dir;
    phase_shape(id, kind, dir);
    // line launches: one tile per line
    const TileCfg t = kind == EXT_LINE ? TileCfg{1, 1, 0} : default_tile(lay);
    if (!lay.apply_tile(id, t)) return false;
  }
  be.h2d(lay.dctx, hctx.data(), sizeof(Ctx<T>) * size_t(n));
  return true;
}

// Cycles of the boxes of lay with their clocks on the device: each cycle's begin
// launch takes every box's CFL maxima, decides whether the box still runs (the
// reference's loop condition), clips its dt and forms its scalars, and the end
// launch swaps its buffers and advances its clock.  The host reads the clocks back
// every 16 cycles only to stop early.  The clocks hclk are uploaded first and read
// back at the end; c0 supplies the launch-invariant decisions.  ec_first: the
// boxes' cell E of the current state is already stored (the first cycle skips it).
template <class T, class B>
void clocked_loop(B& be, BatchLayout& lay, Ctx<T>& c0, std::vector<BoxClock>& hclk, double cfl, double t_end,
                  int64_t max_cycles, int64_t iters, bool ec_first) {
  be.h2d(lay.dclk, hclk.data(), sizeof(BoxClock) * size_t(lay.nbox));
  auto& bx = be.batch_exec(&lay);
  bx.template activate_all<T>();
  const int ec = c0.ec_valid;
  auto one_cycle = [&] {
    bx.template begin<T>(cfl, t_end, max_cycles, 1);
    run_cycle(bx, c0);
    bx.template end<T>();
  };
  for (int64_t it = 0; it < iters; ++it) {
    // the first cycle forms the cell E of the set state unless it is stored; later
    // ones take the one each box's update stored
    c0.ec_valid = it > 0 || ec_first;
    // later cycles take their speeds from the previous update's epilogue
    if (it == 0) bx.cfl(c0);
    one_cycle();
    if ((it + 1) % 16 == 0 && it + 1 < iters) {
      be.d2h(hclk.data(), lay.dclk, sizeof(BoxClock) * size_t(lay.nbox));
      bool any = false;
      for (const BoxClock& k : hclk) any = any || (k.t < t_end && k.cyc < max_cycles);
      if (!any) break;
    }
    if (be.error()) break;
  }
  c0.ec_valid = ec;
  be.d2h(hclk.data(), lay.dclk, sizeof(BoxClock) * size_t(lay.nbox));
}

// ---------------------------------------------------------------------------
// one grid
// ---------------------------------------------------------------------------
template <class T, class B>
struct Handle : public pmhd_handle {
  B be;
  Ctx<T> c;
  StateBufs<T> sb;
  std::vector<void*> owned;
  int cur = 0;
  int64_t cycles_since_set = 0;
  bool speeds_ready = false;   // the last update left the new state's CFL maxima
  bool ec_ready = false;       // the last update left the new state's cell E
  // the grid as a one-box batch for the device-clocked loop (built on first use)
  BatchLayout clay;
  bool clay_ready = false;

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

  int init(const pmhd_grid& g_, const pmhd_params& p_) override {
    g = g_;
    p = p_;
    init_ctx(c, g, p);
    try {
      alloc_grid(c, sb, p.reconstruct_order, [&](int64_t n) { return dalloc(n); });
    } catch (const std::bad_alloc&) {
      return fail(PMHD_ENOMEM, "device allocation failed");
    }
    bind_bufs(c, sb, cur);
    be.zero(c.fcnt, FLOOR_WORDS * sizeof(unsigned long long));
    for (int d = 0; d < g.ndim; ++d)
      c.seg[d] = sweep_segment(g.ndim, d, c.nface(d), c.n_along(d), be.compute_units(), sizeof(T) == sizeof(double));
    c.ec_next = p.physics == 0 && g.ndim >= 2;
    const char* e = be.error();
    if (e) return fail(PMHD_EHIP, e);
    return PMHD_OK;
  }

  int set_sweep(const int32_t seg[3]) override {
    for (int d = 0; d < 3; ++d) {
      const int lo = d >= g.ndim || line_axis(g.ndim, d) ? -1 : 0;
      if (seg[d] < lo || seg[d] > 4096)
        return fail(PMHD_EINVAL, "sweep run length must be in [0, 4096], or -1 (one workgroup per line) along "
                                 "the contiguous axis (z in 3D, x in 2D)");
    }
    for (int d = 0; d < 3; ++d) c.seg[d] = d < g.ndim ? seg[d] : 0;
    clay_ready = false;
    return PMHD_OK;
  }

  int set_state(const double* prim, const double* cons, const double* fbx, const double* fby,
                const double* fbz) override {
    const double* fin[3] = {fbx, fby, fbz};
    cur = 0;
    cycles_since_set = 0;
    speeds_ready = false;
    ec_ready = false;
    bind_bufs(c, sb, cur);
    int rc = load_state(be, c, g, prim, cons, fin, sb.U[0], sb.W[0], sb.FB[0], sb.FBLO[0], err);
    if (rc) return rc;
    be.zero(c.fcnt, FLOOR_WORDS * sizeof(unsigned long long));
    be.sync();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int get_state(int which, double* prim, double* cons, double* fbx, double* fby, double* fbz) override {
    int s = cur;
    if (which == 1 && cycles_since_set > 0) s = 1 - cur;
    double* fo[3] = {fbx, fby, fbz};
    be.sync();
    store_state(be, c, g, sb.U[s], sb.W[s], sb.FB[s], sb.FBLO[s], prim, cons, fo);
    be.sync();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int set_state_device(const double* prim, const double* cons, const double* fbx, const double* fby,
                       const double* fbz) override {
    const double* fin[3] = {fbx, fby, fbz};
    cur = 0;
    cycles_since_set = 0;
    speeds_ready = false;
    ec_ready = false;
    bind_bufs(c, sb, cur);
    int rc = load_state_dev(be, c, g, prim, cons, fin, sb.U[0], sb.W[0], sb.FB[0], sb.FBLO[0], err);
    if (rc) return rc;
    be.zero(c.fcnt, FLOOR_WORDS * sizeof(unsigned long long));
    be.sync();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int get_state_device(int which, double* prim, double* cons, double* fbx, double* fby, double* fbz) override {
    int s = cur;
    if (which == 1 && cycles_since_set > 0) s = 1 - cur;
    double* fo[3] = {fbx, fby, fbz};
    store_state_dev(be, c, g, sb.U[s], sb.W[s], sb.FB[s], sb.FBLO[s], prim, cons, fo);
    be.sync();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int set_stream(void* stream) override {
    be.set_stream(stream);
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  // The evolve loop with the clock on the device: the grid runs as a batch of one
  // box over its own buffers (the batch kernels read the context from device
  // memory, where each cycle's begin launch forms dt and its scalars), so cycles
  // are queued without a host round trip.  The results are those of the host loop
  // (bit for bit in FP64).
  int advance_clocked(double t0, double t_end, int64_t cycle0, int64_t max_cycles, double* t_out,
                      int64_t* cycle_out, double* t2_out) override {
    bind_bufs(c, sb, cur);
    c.active = 1;
    if (!clay_ready) {
      // (again after pmhd_set_sweep: the run lengths set the launches' index boxes)
      std::vector<Ctx<T>> one(1, c);
      try {
        if (!init_layout(be, clay, one, [&](size_t bytes) { return static_cast<void*>(dalloc(int64_t(
                                                                  (bytes + sizeof(T) - 1) / sizeof(T)))); }))
          return fail(PMHD_EINVAL, "grid too large for the device-clocked loop (more than 2^31 tiles)");
      } catch (const std::bad_alloc&) {
        return fail(PMHD_ENOMEM, "device allocation failed");
      }
      // line sweeps: the single grid's workgroup (the smallest of 64, 128, 256
      // threads that covers the line, else 256 in chunks), not the batch's 64
      const int ax = g.ndim == 3 ? 2 : 0;
      int lb = 64;
      while (lb < 256 && int64_t(lb) < c.n_along(ax)) lb <<= 1;
      clay.line_block = lb;
      clay_ready = true;
    } else {
      be.h2d(clay.dctx, &c, sizeof(Ctx<T>));
    }
    std::vector<BoxClock> k(1, BoxClock{t0, t0, 0.0, cycle0, 0, 0});
    clocked_loop(be, clay, c, k, p.cfl, t_end, max_cycles, max_cycles - cycle0, ec_ready);
    const char* e = be.error();
    if (e) return fail(PMHD_EHIP, e);
    const int64_t done = k[0].done;
    if (done > 0) {
      cur = int((cur + done) & 1);
      bind_bufs(c, sb, cur);
      cycles_since_set += done;
      // the speeds went to the batch's maxima, not to this grid's partials
      speeds_ready = false;
      ec_ready = c.ec_next != 0;
    }
    if (t_out) *t_out = k[0].t;
    if (cycle_out) *cycle_out = k[0].cyc;
    if (t2_out) *t2_out = k[0].t2;
    if (k[0].nonfinite)
      return fail(PMHD_ENONFINITE, "the state became non-finite (NaN timestep); the results are those of the "
                                   "reference's loop, which ends with a NaN time");
    return PMHD_OK;
  }

  int compute_dt(double* dt) override {
    double m[3] = {0, 0, 0};
    if (speeds_ready) be.exec().max_partials(m);
    else be.exec().cfl(c, m);
    const char* e = be.error();
    if (e) return fail(PMHD_EHIP, e);
    *dt = dt_of(c, p.cfl, m);
    return PMHD_OK;
  }

  int step(double dt) override {
    set_step_scalars(c, dt);
    {
      // Source/CTU_HelperModule.py:455: dtdy = longdouble(dt)/longdouble(dy)
      const long double hl = 0.5L * (static_cast<long double>(dt) / static_cast<long double>(g.dy));
      c.ldh[0] = static_cast<double>(hl);
      c.ldh[1] = static_cast<double>(hl - static_cast<long double>(c.ldh[0]));
    }
    c.ec_valid = ec_ready;
    run_cycle(be.exec(), c);
    cur = 1 - cur;
    bind_bufs(c, sb, cur);
    ++cycles_since_set;
    speeds_ready = true;
    ec_ready = c.ec_next != 0;
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
    if (max_divb) *max_divb = c.physics == 0 && g.ndim >= 2 && maxes[1] > 0 ? maxes[0] / maxes[1] : 0.0;
    if (max_b) *max_b = maxes[1];
    return PMHD_OK;
  }

  int set_profiling(int on) override {
    be.set_profiling(on != 0);
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int get_profile(int maxn, char* names, double* ms, int64_t* launches, int64_t* elems, int32_t* n) override {
    *n = be.profile(maxn, names, ms, launches, elems);
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int get_stats(pmhd_stats* out) override {
    std::memset(out, 0, sizeof(*out));
    unsigned long long f[FLOOR_WORDS];
    be.sync();
    be.d2h(f, c.fcnt, sizeof(f));
    floor_stats(f, out);
    double sums[9], maxes[2];
    pmhd::diagnostics(be.exec(), c, sums, maxes);
    out->nonfinite = int64_t(sums[8]);
    out->cycles = cycles_since_set;
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }
};

// ---------------------------------------------------------------------------
// a batch of independent boxes (no counterpart in the reference): every phase is
// one launch over the tiles of all boxes, each box keeps its own timestep and
// clock on the device, so a run of many cycles needs no host round trip
// ---------------------------------------------------------------------------
template <class T, class B>
struct BatchHandle : public pmhd_batch {
  B be;
  pmhd_params p;
  std::vector<pmhd_grid> grids;
  std::vector<Ctx<T>> hctx;           // host mirror of the device contexts
  std::vector<StateBufs<T>> bufs;
  std::vector<BoxClock> hclk;
  BatchLayout lay;
  void* arena = nullptr;
  std::vector<void*> owned;

  ~BatchHandle() override {
    for (void* ptr : owned) be.release(ptr);
  }

  void* dalloc_bytes(size_t bytes) {
    void* ptr = be.alloc(bytes > 0 ? bytes : 1);
    if (!ptr) throw std::bad_alloc();
    owned.push_back(ptr);
    return ptr;
  }

  int init(int n, const pmhd_grid* gs, const pmhd_params& p_) override {
    nbox = n;
    p = p_;
    grids.assign(gs, gs + n);
    hctx.resize(size_t(n));
    bufs.resize(size_t(n));
    hclk.assign(size_t(n), BoxClock{0, 0, 0, 0, 0, 0});
    for (int b = 0; b < n; ++b) init_ctx(hctx[size_t(b)], grids[size_t(b)], p);
    try {
      // one arena for all boxes: a sizing pass, then the carving pass
      size_t total = 0;
      const size_t align = 256;
      for (int pass = 0; pass < 2; ++pass) {
        char* base = static_cast<char*>(arena);
        size_t off = 0;
        for (int b = 0; b < n; ++b) {
          alloc_grid(hctx[size_t(b)], bufs[size_t(b)], p.reconstruct_order, [&](int64_t cnt) {
            const size_t bytes = size_t(std::max<int64_t>(cnt, 1)) * sizeof(T);
            T* ptr = pass == 0 ? nullptr : reinterpret_cast<T*>(base + off);
            off += (bytes + align - 1) / align * align;
            return ptr;
          });
        }
        if (pass == 0) {
          total = off;
          arena = dalloc_bytes(total);
        }
      }
    } catch (const std::bad_alloc&) {
      return fail(PMHD_ENOMEM, "device allocation failed");
    }
    // fused predictor sweeps with one plan per direction for all boxes: line
    // launches along the contiguous axis (one workgroup per line of a box), runs
    // across it
    int seg[3] = {0, 0, 0};
    {
      int64_t nf[3] = {0, 0, 0};
      for (int b = 0; b < n; ++b)
        for (int d = 0; d < grids[0].ndim; ++d) nf[d] += hctx[size_t(b)].nface(d);
      for (int d = 0; d < grids[0].ndim; ++d) {
        seg[d] = sweep_segment(grids[0].ndim, d, nf[d], int64_t(0), be.compute_units(), sizeof(T) == sizeof(double));
        if (seg[d] > 0 && line_axis(grids[0].ndim, d)) seg[d] = -1;
      }
    }
    for (int b = 0; b < n; ++b) {
      bind_bufs(hctx[size_t(b)], bufs[size_t(b)], 0);
      hctx[size_t(b)].ec_next = p.physics == 0 && grids[0].ndim >= 2;
      for (int d = 0; d < 3; ++d) hctx[size_t(b)].seg[d] = seg[d];
    }
    for (int b = 0; b < n; ++b) be.zero(hctx[size_t(b)].fcnt, FLOOR_WORDS * sizeof(unsigned long long));
    try {
      if (!init_layout(be, lay, hctx, [&](size_t bytes) { return dalloc_bytes(bytes); }))
        return fail(PMHD_EINVAL, "batch too large: more than 2^31 tiles");
    } catch (const std::bad_alloc&) {
      return fail(PMHD_ENOMEM, "device allocation failed");
    }
    const char* e = be.error();
    if (e) return fail(PMHD_EHIP, e);
    return PMHD_OK;
  }

  void push_ctx() { be.h2d(lay.dctx, hctx.data(), sizeof(Ctx<T>) * size_t(nbox)); }
  void pull_ctx() { be.d2h(hctx.data(), lay.dctx, sizeof(Ctx<T>) * size_t(nbox)); }

  int check_box(int box) {
    if (box < 0 || box >= nbox) return fail(PMHD_EINVAL, "box index out of range");
    return PMHD_OK;
  }

  int set_state(int box, const double* prim, const double* cons, const double* fbx, const double* fby,
                const double* fbz) override {
    if (check_box(box)) return PMHD_EINVAL;
    Ctx<T>& c = hctx[size_t(box)];
    const double* fin[3] = {fbx, fby, fbz};
    bind_bufs(c, bufs[size_t(box)], 0);
    int rc = load_state(be, c, grids[size_t(box)], prim, cons, fin, bufs[size_t(box)].U[0], bufs[size_t(box)].W[0],
                        bufs[size_t(box)].FB[0], bufs[size_t(box)].FBLO[0], err);
    if (rc) return rc;
    hclk[size_t(box)].done = 0;
    be.zero(c.fcnt, FLOOR_WORDS * sizeof(unsigned long long));
    be.h2d(static_cast<Ctx<T>*>(lay.dctx) + box, &c, sizeof(Ctx<T>));
    be.sync();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int get_state(int box, int which, double* prim, double* cons, double* fbx, double* fby, double* fbz) override {
    if (check_box(box)) return PMHD_EINVAL;
    const Ctx<T>& c = hctx[size_t(box)];
    const bool second = which == 1 && hclk[size_t(box)].done > 0;
    const T* fb[3] = {second ? c.fb2[0] : c.fb[0], second ? c.fb2[1] : c.fb[1], second ? c.fb2[2] : c.fb[2]};
    const T* fblo[3] = {second ? c.fb2lo[0] : c.fblo[0], second ? c.fb2lo[1] : c.fblo[1],
                        second ? c.fb2lo[2] : c.fblo[2]};
    double* fo[3] = {fbx, fby, fbz};
    be.sync();
    store_state(be, c, grids[size_t(box)], second ? c.U2 : c.U, second ? c.W2 : c.W, fb, fblo, prim, cons, fo);
    be.sync();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int advance(double t0, double t_end, int64_t cycle0, int64_t max_cycles, double* t_out, int64_t* cycle_out,
              double* t2_out) override {
    for (int b = 0; b < nbox; ++b) {
      BoxClock& k = hclk[size_t(b)];
      k.t = t0;
      k.t2 = t0;
      k.dt = 0;
      k.cyc = cycle0;
      k.nonfinite = 0;
    }
    clocked_loop(be, lay, hctx[0], hclk, p.cfl, t_end, max_cycles, max_cycles - cycle0, false);
    pull_ctx();
    for (int b = 0; b < nbox; ++b) {
      if (t_out) t_out[b] = hclk[size_t(b)].t;
      if (cycle_out) cycle_out[b] = hclk[size_t(b)].cyc;
      if (t2_out) t2_out[b] = hclk[size_t(b)].t2;
    }
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  // time every candidate tile of every phase on this batch (one cycle's worth of
  // launches per candidate, writing only work and next-state buffers) and keep
  // the fastest; the current state is unchanged, the second-last is dropped
  int autotune(int reps) override {
    auto& bx = be.batch_exec(&lay);
    bx.template activate_all<T>();
    bx.tuning = true;
    bx.tune_reps = reps > 0 ? reps : 3;
    bx.cfl(hctx[0]);
    bx.template begin<T>(p.cfl, std::numeric_limits<double>::infinity(), std::numeric_limits<int64_t>::max(), 0);
    run_cycle(bx, hctx[0]);
    bx.tuning = false;
    be.zero(lay.dred, 3 * sizeof(unsigned long long) * size_t(nbox));
    // the tuning launches floored and counted: start the counts afresh
    for (int b = 0; b < nbox; ++b) be.zero(hctx[size_t(b)].fcnt, FLOOR_WORDS * sizeof(unsigned long long));
    be.sync();
    for (BoxClock& k : hclk) k.done = 0;
    pull_ctx();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int get_tiles(int maxn, char* names, int32_t* bk, int32_t* bi, int32_t* grid, double* ms, int32_t* n) override {
    int k = 0;
    for (int id = 0; id < K_COUNT && k < maxn; ++id) {
      if (lay.tuned_ms[id] < 0.f) continue;
      std::strncpy(names + 32 * k, kernel_name(id), 31);
      names[32 * k + 31] = 0;
      bk[k] = lay.tiles[id].bk;
      bi[k] = lay.tiles[id].bi;
      grid[k] = lay.tiles[id].grid;
      ms[k] = double(lay.tuned_ms[id]);
      ++k;
    }
    *n = k;
    return PMHD_OK;
  }

  int synchronize() override {
    be.sync();
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }
  int set_profiling(int on) override {
    be.set_profiling(on != 0);
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }
  int get_profile(int maxn, char* names, double* ms, int64_t* launches, int64_t* elems, int32_t* n) override {
    *n = be.profile(maxn, names, ms, launches, elems);
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  int get_stats(int box, pmhd_stats* out) override {
    if (check_box(box)) return PMHD_EINVAL;
    std::memset(out, 0, sizeof(*out));
    unsigned long long f[FLOOR_WORDS];
    be.sync();
    be.d2h(f, hctx[size_t(box)].fcnt, sizeof(f));
    floor_stats(f, out);
    double sums[9], maxes[2];
    pmhd::diagnostics(be.exec(), hctx[size_t(box)], sums, maxes);
    out->nonfinite = int64_t(sums[8]);
    out->cycles = hclk[size_t(box)].done;
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }
};

// the first problem of a grid, as a string literal (nullptr: none)
inline const char* grid_problem(const pmhd_grid* g) {
  if (g->ndim < 1 || g->ndim > 3) return "ndim must be 1, 2 or 3";
  if (g->nx < 1 || (g->ndim >= 2 && g->ny < 1) || (g->ndim == 3 && g->nz < 1)) return "number of cells must be positive";
  if (g->nx > (1 << 24) || g->ny > (1 << 24) || g->nz > (1 << 24))
    return "number of cells along an axis must not exceed 2^24";
  {
    // every variable plane of a state, work or edge array stays below 4 GiB: the
    // kernels form byte offsets within a plane in 32 bits
    const int64_t ext = (g->nx + 2) * (g->ndim >= 2 ? g->ny + 2 : 1) * (g->ndim == 3 ? g->nz + 2 : 1);
    if (ext > (int64_t(1) << 29)) return "grid too large: (nx+2)(ny+2)(nz+2) must not exceed 2^29 cells";
  }
  if (!(g->dx > 0) || (g->ndim >= 2 && !(g->dy > 0)) || (g->ndim == 3 && !(g->dz > 0)))
    return "cell widths must be positive";
  const int bcs[3] = {g->bcx, g->bcy, g->bcz};
  for (int d = 0; d < g->ndim; ++d)
    if (bcs[d] != 0 && bcs[d] != 1) return "boundary condition must be 0 (outflow) or 1 (periodic)";
  return nullptr;
}

inline int check_grid(const pmhd_grid* g, std::string& err) {
  const char* m = grid_problem(g);
  if (m) {
    err = m;
    return PMHD_EINVAL;
  }
  return PMHD_OK;
}

// every launch of a single grid fits the HIP grid limits (blocks along y and z
// below 65536) with the block shape the single-grid executor uses: `threads`
// threads (256; the adjoint's phase launches use 64), the smallest power of two
// >= n2 (at most 64) of them along the contiguous axis
inline int check_launch_extents(const pmhd_grid* g, std::string& err, int threads = 256) {
  Ctx<double> c;
  std::memset(&c, 0, sizeof(c));
  c.ndim = g->ndim;
  c.ny = g->ndim >= 2 ? g->ny : 1;
  c.nx = g->nx;
  c.nz = g->ndim == 3 ? g->nz : 1;
  for (int kind = 0; kind < 4; ++kind)
    for (int d = 0; d < 3; ++d) {
      int64_t n[3];
      c.extent(kind, d, n);
      int64_t bx = 1;
      while (bx < 64 && bx < n[2]) bx <<= 1;
      const int64_t by = threads / bx > 0 ? threads / bx : 1;
      if (n[0] > 65535 || (n[1] + by - 1) / by > 65535) {
        err = "grid too large for one launch (cells along y, or along x per block row)";
        return PMHD_EINVAL;
      }
    }
  return PMHD_OK;
}

inline const char* params_problem(const pmhd_params* p) {
  if (p->reconstruct_order < 0 || p->reconstruct_order > 2) return "reconstruct_order must be 0, 1 or 2";
  if (p->precision != 64 && p->precision != 32) return "precision must be 64 or 32";
  if (p->physics != 0 && p->physics != 1) return "physics must be 0 (MHD) or 1 (hydrodynamics)";
  if (p->riemann < 0 || p->riemann > 2) return "riemann must be 0 (Roe), 1 (HLLE) or 2 (Roe with HLLE fallback)";
  if (p->physics == 1 && p->riemann != 0) return "the hydrodynamic step has the Roe solver only (riemann 0)";
  if (p->fix_quirks & ~7) return "fix_quirks has unknown bits (PMHD_FIX_Z_BC | PMHD_FIX_OUTFLOW_GHOSTS | "
                                 "PMHD_FIX_DTDY_80BIT)";
  return nullptr;
}

inline int check_params(const pmhd_params* p, std::string& err) {
  const char* m = params_problem(p);
  if (m) {
    err = m;
    return PMHD_EINVAL;
  }
  return PMHD_OK;
}

inline pmhd_grid normal_grid(const pmhd_grid& g) {
  pmhd_grid gg = g;
  if (gg.ndim < 3) { gg.nz = 1; gg.bcz = gg.bcx; gg.dz = 1.0; }
  if (gg.ndim < 2) { gg.ny = 1; gg.bcy = gg.bcx; gg.dy = 1.0; }
  return gg;
}

// FP32 grids whose interface fluxes along x alone (7 or 5 words per face) exceed
// the 256 MiB last-level cache: every launch's results stream through HBM anyway,
// so they are stored past the caches (the non-temporal unit), which leaves the L2
// to the inputs' neighbour re-reads; smaller grids keep ordinary stores, whose
// values the next launch may still find cached
inline bool streaming_grid(const pmhd_grid& g, const pmhd_params& p) {
  const int64_t ny = g.ndim >= 2 ? g.ny : 1, nz = g.ndim == 3 ? g.nz : 1;
  const int64_t faces = ny * (int64_t(g.nx) + 1) * nz;
  return int64_t(face_vars(p.physics)) * faces * int64_t(sizeof(float)) > (int64_t(256) << 20);
}

inline int create_handle(const pmhd_grid* g, const pmhd_params* p, pmhd_handle** out, std::string& err) {
  if (!g || !p || !out) {
    err = "null argument";
    return PMHD_EINVAL;
  }
  *out = nullptr;
  int rc = check_grid(g, err);
  if (rc) return rc;
  rc = check_launch_extents(g, err);
  if (rc) return rc;
  rc = check_params(p, err);
  if (rc) return rc;
  const pmhd_grid gg = normal_grid(*g);
  pmhd_handle* h = p->precision == 64 ? pmhd_new_handle_f64()
                   : streaming_grid(*g, *p)   ? pmhd_new_handle_f32nt()
                                             : pmhd_new_handle_f32();
  if (!h) { err = "out of host memory"; return PMHD_ENOMEM; }
  rc = h->init(gg, *p);
  if (rc != PMHD_OK) {
    err = h->err;
    delete h;
    return rc;
  }
  *out = h;
  return PMHD_OK;
}

inline int create_batch(int nbox, const pmhd_grid* gs, const pmhd_params* p, pmhd_batch** out, std::string& err) {
  if (!gs || !p || !out) {
    err = "null argument";
    return PMHD_EINVAL;
  }
  *out = nullptr;
  if (nbox < 1) { err = "a batch needs at least one box"; return PMHD_EINVAL; }
  int rc = check_params(p, err);
  if (rc) return rc;
  std::vector<pmhd_grid> gg(static_cast<size_t>(nbox));
  for (int b = 0; b < nbox; ++b) {
    rc = check_grid(&gs[b], err);
    if (rc) { err = "box " + std::to_string(b) + ": " + err; return rc; }
    if (gs[b].ndim != gs[0].ndim) { err = "all boxes of a batch must have the same ndim"; return PMHD_EINVAL; }
    gg[size_t(b)] = normal_grid(gs[b]);
  }
  pmhd_batch* h = p->precision == 64 ? pmhd_new_batch_f64() : pmhd_new_batch_f32();
  if (!h) { err = "out of host memory"; return PMHD_ENOMEM; }
  rc = h->init(nbox, gg.data(), *p);
  if (rc != PMHD_OK) {
    err = h->err;
    delete h;
    return rc;
  }
  *out = h;
  return PMHD_OK;
}

// grids of at most this many cells run the evolve loop with the device clock by
// default: their cycles are short enough that a host round trip per cycle shows;
// larger grids run faster in the single-grid kernels than as a one-box batch
constexpr int64_t kClockDeviceCells = int64_t(1) << 16;

inline int advance(pmhd_handle* h, double t0, double t_end, int64_t cycle0, int64_t max_cycles,
                   double* t_out, int64_t* cycle_out, double* t2_out) {
  const int64_t cells = int64_t(h->g.nx) * (h->g.ndim >= 2 ? h->g.ny : 1) * (h->g.ndim == 3 ? h->g.nz : 1);
  const bool dev = h->clock == PMHD_CLOCK_DEVICE || (h->clock == PMHD_CLOCK_AUTO && cells <= kClockDeviceCells);
  if (dev) return h->advance_clocked(t0, t_end, cycle0, max_cycles, t_out, cycle_out, t2_out);
  double t = t0, t2 = t0;
  int64_t cyc = cycle0;
  bool finite = true;
  while (t < t_end && cyc < max_cycles) {
    t2 = t;
    double dt;
    int rc = h->compute_dt(&dt);
    if (rc) return rc;
    // a non-finite state gives a NaN timestep; the reference still runs that cycle
    // (its clock becomes NaN and its loop ends), and so does this loop
    if (!(dt - dt == 0.0)) finite = false;
    dt = clip_dt(h->g.ndim, h->p.physics, t, t_end, dt);
    rc = h->step(dt);
    if (rc) return rc;
    t += dt;
    ++cyc;
  }
  if (t_out) *t_out = t;
  if (cycle_out) *cycle_out = cyc;
  if (t2_out) *t2_out = t2;
  if (!finite) return h->fail(PMHD_ENONFINITE, "the state became non-finite (NaN timestep); the results are those "
                                               "of the reference's loop, which ends with a NaN time");
  return PMHD_OK;
}

PMHD_NS_END

// The constructors of one precision, in the translation unit compiled for it.
#define PMHD_DEFINE_FACTORIES(SUFFIX, T, BACKEND)                                           \
  pmhd_handle* pmhd_new_handle_##SUFFIX() { return new (std::nothrow) pmhd::Handle<T, BACKEND>(); } \
  pmhd_batch* pmhd_new_batch_##SUFFIX() { return new (std::nothrow) pmhd::BatchHandle<T, BACKEND>(); } \
  int pmhd_riemann_##SUFFIX(int d, int solver, int64_t n, const double* wl, const double* wr,          \
                            const double* ul, const double* ur, const double* bn, double gamma,        \
                            int efix, double eps, double* F, std::string& err) {                      \
    return pmhd::riemann_run<T>(d, solver, n, wl, wr, ul, ur, bn, gamma, efix, eps, F, err);          \
  }

// C ABI bodies (one translation unit; the handles come from the constructors above).
#define PMHD_DEFINE_C_ABI()                                                                 \
  static std::string g_pmhd_create_error;                                                   \
  static const char* g_pmhd_null_error = nullptr;                                           \
  void pmhd_note_null_error(const char* literal) { g_pmhd_null_error = literal; }           \
  extern "C" int pmhd_create(const pmhd_grid* g, const pmhd_params* p, pmhd_handle** out) { \
    g_pmhd_null_error = nullptr;                                                            \
    try {                                                                                   \
      return pmhd::create_handle(g, p, out, g_pmhd_create_error);                           \
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
  extern "C" int pmhd_set_state_device(pmhd_handle* h, const double* prim, const double* cons, \
                                       const double* fbx, const double* fby, const double* fbz) { \
    if (!h) return PMHD_EINVAL;                                                             \
    try { return h->set_state_device(prim, cons, fbx, fby, fbz); }                          \
    catch (...) { return h->fail(PMHD_EINVAL, "exception in pmhd_set_state_device"); }      \
  }                                                                                         \
  extern "C" int pmhd_get_state_device(pmhd_handle* h, int32_t which, double* prim,         \
                                       double* cons, double* fbx, double* fby, double* fbz) { \
    if (!h) return PMHD_EINVAL;                                                             \
    try { return h->get_state_device(which, prim, cons, fbx, fby, fbz); }                   \
    catch (...) { return h->fail(PMHD_EINVAL, "exception in pmhd_get_state_device"); }      \
  }                                                                                         \
  extern "C" int pmhd_set_stream(pmhd_handle* h, void* stream) {                            \
    if (!h) return PMHD_EINVAL;                                                             \
    return h->set_stream(stream);                                                           \
  }                                                                                         \
  extern "C" int pmhd_compute_dt(pmhd_handle* h, double* dt) {                              \
    if (!h || !dt) return PMHD_EINVAL;                                                      \
    return h->compute_dt(dt);                                                               \
  }                                                                                         \
  extern "C" int pmhd_step(pmhd_handle* h, double dt) {                                     \
    if (!h) return PMHD_EINVAL;                                                             \
    return h->step(dt);                                                                     \
  }                                                                                         \
  extern "C" int pmhd_set_sweep(pmhd_handle* h, int32_t seg_x, int32_t seg_y, int32_t seg_z) { \
    if (!h) return PMHD_EINVAL;                                                             \
    const int32_t seg[3] = {seg_x, seg_y, seg_z};                                           \
    try { return h->set_sweep(seg); }                                                       \
    catch (...) { return h->fail(PMHD_EINVAL, "exception in pmhd_set_sweep"); }             \
  }                                                                                         \
  extern "C" int pmhd_set_clock(pmhd_handle* h, int32_t mode) {                             \
    if (!h) return PMHD_EINVAL;                                                             \
    if (mode != PMHD_CLOCK_AUTO && mode != PMHD_CLOCK_HOST && mode != PMHD_CLOCK_DEVICE)    \
      return h->fail(PMHD_EINVAL, "clock mode must be PMHD_CLOCK_AUTO, _HOST or _DEVICE");  \
    h->clock = mode;                                                                        \
    return PMHD_OK;                                                                         \
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
  extern "C" int pmhd_riemann_flux(int32_t precision, int32_t direction, int32_t riemann,  \
                                   int64_t n, const double* wl, const double* wr,           \
                                   const double* ul, const double* ur, const double* bn,    \
                                   double gamma, int32_t entropy_fix, double epsilon,       \
                                   double* flux) {                                          \
    g_pmhd_null_error = nullptr;                                                            \
    if ((precision != 64 && precision != 32) || direction < 0 || direction > 2 ||           \
        riemann < 0 || riemann > 2 || n < 0 || (n > 0 && (!wl || !wr || !ul || !ur || !bn || !flux))) { \
      g_pmhd_create_error = "pmhd_riemann_flux: precision 64/32, direction 0-2, riemann 0-2, n >= 0 " \
                            "and all arrays are required";                                  \
      return PMHD_EINVAL;                                                                   \
    }                                                                                       \
    try {                                                                                   \
      return precision == 64                                                                \
                 ? pmhd_riemann_f64(direction, riemann, n, wl, wr, ul, ur, bn, gamma,       \
                                    entropy_fix, epsilon, flux, g_pmhd_create_error)        \
                 : pmhd_riemann_f32(direction, riemann, n, wl, wr, ul, ur, bn, gamma,       \
                                    entropy_fix, epsilon, flux, g_pmhd_create_error);       \
    } catch (...) {                                                                         \
      g_pmhd_create_error = "unexpected exception in pmhd_riemann_flux";                    \
      return PMHD_EINVAL;                                                                   \
    }                                                                                       \
  }                                                                                         \
  extern "C" int pmhd_get_stats(pmhd_handle* h, pmhd_stats* out) {                          \
    if (!h || !out) return PMHD_EINVAL;                                                     \
    try { return h->get_stats(out); }                                                       \
    catch (...) { return h->fail(PMHD_EINVAL, "exception in pmhd_get_stats"); }             \
  }                                                                                         \
  extern "C" const char* pmhd_last_error(const pmhd_handle* h) {                            \
    if (h) return h->err.c_str();                                                           \
    return g_pmhd_null_error ? g_pmhd_null_error : g_pmhd_create_error.c_str();             \
  }                                                                                         \
  extern "C" int pmhd_batch_create(int32_t nbox, const pmhd_grid* grids,                    \
                                   const pmhd_params* p, pmhd_batch** out) {                \
    g_pmhd_null_error = nullptr;                                                            \
    try {                                                                                   \
      return pmhd::create_batch(nbox, grids, p, out, g_pmhd_create_error);                  \
    } catch (...) {                                                                         \
      g_pmhd_create_error = "unexpected exception in pmhd_batch_create";                    \
      return PMHD_EINVAL;                                                                   \
    }                                                                                       \
  }                                                                                         \
  extern "C" int pmhd_batch_destroy(pmhd_batch* b) {                                        \
    delete b;                                                                               \
    return PMHD_OK;                                                                         \
  }                                                                                         \
  extern "C" int pmhd_batch_set_state(pmhd_batch* b, int32_t box, const double* prim,       \
                                      const double* cons, const double* fbx,                \
                                      const double* fby, const double* fbz) {               \
    if (!b) return PMHD_EINVAL;                                                             \
    try { return b->set_state(box, prim, cons, fbx, fby, fbz); }                            \
    catch (...) { return b->fail(PMHD_EINVAL, "exception in pmhd_batch_set_state"); }       \
  }                                                                                         \
  extern "C" int pmhd_batch_get_state(pmhd_batch* b, int32_t box, int32_t which,            \
                                      double* prim, double* cons, double* fbx,              \
                                      double* fby, double* fbz) {                           \
    if (!b) return PMHD_EINVAL;                                                             \
    try { return b->get_state(box, which, prim, cons, fbx, fby, fbz); }                     \
    catch (...) { return b->fail(PMHD_EINVAL, "exception in pmhd_batch_get_state"); }       \
  }                                                                                         \
  extern "C" int pmhd_batch_advance(pmhd_batch* b, double t0, double t_end, int64_t cycle0, \
                                    int64_t max_cycles, double* t_out, int64_t* cycle_out,  \
                                    double* t2_out) {                                       \
    if (!b) return PMHD_EINVAL;                                                             \
    try { return b->advance(t0, t_end, cycle0, max_cycles, t_out, cycle_out, t2_out); }     \
    catch (...) { return b->fail(PMHD_EINVAL, "exception in pmhd_batch_advance"); }         \
  }                                                                                         \
  extern "C" int pmhd_batch_autotune(pmhd_batch* b, int32_t reps) {                         \
    if (!b) return PMHD_EINVAL;                                                             \
    try { return b->autotune(reps); }                                                       \
    catch (...) { return b->fail(PMHD_EINVAL, "exception in pmhd_batch_autotune"); }        \
  }                                                                                         \
  extern "C" int pmhd_batch_get_tiles(pmhd_batch* b, int32_t max_entries, char* names,      \
                                      int32_t* bk, int32_t* bi, int32_t* grid, double* ms,  \
                                      int32_t* n_entries) {                                 \
    if (!b || !names || !bk || !bi || !grid || !ms || !n_entries) return PMHD_EINVAL;       \
    return b->get_tiles(max_entries, names, bk, bi, grid, ms, n_entries);                   \
  }                                                                                         \
  extern "C" int pmhd_batch_synchronize(pmhd_batch* b) {                                    \
    if (!b) return PMHD_EINVAL;                                                             \
    return b->synchronize();                                                                \
  }                                                                                         \
  extern "C" int pmhd_batch_set_profiling(pmhd_batch* b, int32_t enable) {                  \
    if (!b) return PMHD_EINVAL;                                                             \
    return b->set_profiling(enable);                                                        \
  }                                                                                         \
  extern "C" int pmhd_batch_get_profile(pmhd_batch* b, int32_t max_entries, char* names,    \
                                        double* total_ms, int64_t* launches,                \
                                        int64_t* elements, int32_t* n_entries) {            \
    if (!b || !names || !total_ms || !launches || !elements || !n_entries) return PMHD_EINVAL; \
    return b->get_profile(max_entries, names, total_ms, launches, elements, n_entries);     \
  }                                                                                         \
  extern "C" int pmhd_batch_get_stats(pmhd_batch* b, int32_t box, pmhd_stats* out) {       \
    if (!b || !out) return PMHD_EINVAL;                                                     \
    try { return b->get_stats(box, out); }                                                  \
    catch (...) { return b->fail(PMHD_EINVAL, "exception in pmhd_batch_get_stats"); }       \
  }                                                                                         \
  extern "C" const char* pmhd_batch_last_error(const pmhd_batch* b) {                       \
    if (b) return b->err.c_str();                                                           \
    return g_pmhd_null_error ? g_pmhd_null_error : g_pmhd_create_error.c_str();             \
  }
