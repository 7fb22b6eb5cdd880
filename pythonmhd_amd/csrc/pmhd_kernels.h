// Kernel bodies of one CTU+CT cycle: one function per launch, called once per
// element (cell, face or edge point).  The same bodies run under the HIP
// executor (one thread per element) and under the host test executor.
//
// Layouts are the reference's (Source/SimulationGrid.py:47-169,
// Source/CT_HelperModule.py:29-103): cells (v, ny, nx, nz) with z contiguous,
// faces (ny, nx+1, nz) / (ny+1, nx, nz) / (ny, nx, nz+1), edges (ny+1, nx+1, nz+1).
// 2D uses nz = 1, 1D uses ny = nz = 1.  Interface states and fluxes are stored in
// the reference's directional 7-vector order (Source/EvolveFunctions_MHD.py:775-807).
#pragma once
#include <type_traits>

#include "pmhd.h"
#include "pmhd_math.h"

PMHD_NS_BEGIN

enum { OUTFLOW = 0, PERIODIC = 1 };
// element kinds a launch runs over
// EXT_SEG: runs of consecutive faces along the launch direction (one run per thread);
// EXT_LINE: whole lines of faces along the launch direction (one per element);
// EXT_ENDS: the two end faces of every line along the launch direction
enum ExtKind { EXT_CELL = 0, EXT_FACE = 1, EXT_SLOPE = 2, EXT_EDGE = 3, EXT_SEG = 4, EXT_LINE = 5, EXT_ENDS = 6 };

// launch identifiers (for per-kernel timing); "+ d" for the per-direction ones
enum KernelId {
  K_SLOPES = 0, K_EDGES = 3, K_COPY = 6, K_PREDICT = 9, K_CELL_E = 12, K_EMF = 13, K_EMF2 = 14,
  K_CT_HALF = 17, K_HALF_E = 20, K_CORRECT = 21, K_CT_FULL = 24, K_UPDATE = 27, K_CFL = 28,
  K_DIAG = 29, K_SWEEP = 30, K_SWEEP_LINE_Z = 33, K_SWEEP_ENDS_Z = 34, K_ADJOINT = 35, K_SWEEP_LINE_X = 36,
  K_SWEEP_ENDS_X = 37, K_COUNT = 38
};
inline const char* kernel_name(int id) {
  static const char* names[K_COUNT] = {
      "slopes_x", "slopes_y", "slopes_z", "edges_x", "edges_y", "edges_z", "copy_x", "copy_y", "copy_z",
      "predict_x", "predict_y", "predict_z", "cell_e", "emf", "emf_corr", "unused_15", "unused_16",
      "ct_half", "unused_18", "unused_19", "half_e", "ctu_correct_x", "ctu_correct_y", "ctu_correct_z",
      "ct_full", "unused_25", "unused_26", "update", "cfl", "diag", "sweep_x", "sweep_y", "sweep_z",
      "sweep_z", "sweep_ends_z", "adjoint", "sweep_x", "sweep_ends_x"};
  return (id >= 0 && id < K_COUNT) ? names[id] : "?";
}
// element kind (ExtKind) and direction of each launch id
inline void phase_shape(int id, int& kind, int& dir) {
  dir = 0;
  kind = EXT_CELL;
  if (id < K_EDGES) { kind = EXT_SLOPE; dir = id - K_SLOPES; }
  else if (id < K_COPY) { dir = id - K_EDGES; }
  else if (id < K_PREDICT) { kind = EXT_FACE; dir = id - K_COPY; }
  else if (id < K_CELL_E) { kind = EXT_FACE; dir = id - K_PREDICT; }
  else if (id == K_EMF || id == K_EMF2) { kind = EXT_EDGE; }
  else if (id == K_CT_HALF || id == K_CT_FULL) { kind = EXT_EDGE; }
  else if (id >= K_CORRECT && id < K_CT_FULL) { kind = EXT_FACE; dir = id - K_CORRECT; }
  else if (id >= K_SWEEP && id < K_SWEEP + 3) { kind = EXT_SEG; dir = id - K_SWEEP; }
  else if (id == K_SWEEP_LINE_Z) { kind = EXT_LINE; dir = 2; }
  else if (id == K_SWEEP_ENDS_Z) { kind = EXT_ENDS; dir = 2; }
  else if (id == K_SWEEP_LINE_X) { kind = EXT_LINE; dir = 0; }
  else if (id == K_SWEEP_ENDS_X) { kind = EXT_ENDS; dir = 0; }
}

// pointer types of a context's arrays (the adjoint's dual contexts use accessors)
template <class T>
struct Mem {
  using cp = const T*;
  using p = T*;
};
template <class T>
using CPtr = typename Mem<T>::cp;
template <class T>
using Ptr = typename Mem<T>::p;

// The FP32-bulk step keeps its face-centred field as two float arrays, the high
// and low parts of a double value (its translation unit defines PMHD_FACES_SPLIT):
// every consumer reads the high part, which is the float rounding of the double
// value, and the constrained-transport update works on the double value, so div B
// stays at double rounding while the consumers move float words.
template <class T>
PMHD_FI constexpr bool split_faces() {
  return false;
}
#ifdef PMHD_FACES_SPLIT
template <>
PMHD_FI constexpr bool split_faces<float>() {
  return true;
}
#endif

template <class T>
struct Ctx {
  int ndim;
  int active;        // 0: a batch box that has finished (its launches are skipped)
  int physics;       // 0 ideal MHD (8 cell variables + face B), 1 hydrodynamics (5 variables)
  int64_t ny, nx, nz;
  int bc[3];         // boundary condition per direction (x, y, z)
  int bc_zrec;       // boundary condition the z reconstruction/corrections use
  int order;         // 0 copy, 1 PLM, 2 PPM
  int efix;
  int ec_next;       // the update also stores the new state's cell E (the next cycle's CellE)
  int ec_valid;      // Ec already holds the current state's cell E (the cycle skips CellE)
  int riemann;       // RIEMANN_ROE (the reference's), RIEMANN_HLLE or RIEMANN_ROE_HLLE
  int fixq;          // PMHD_FIX_* bits: reference behaviours corrected instead of reproduced
  int seg[3];        // faces per thread of the fused predictor sweep along x, y, z (0: separate
                     // slope / edge / predictor launches; -1 along z: one workgroup per line)
  T gamma, eps, min_dens, min_pres, min_energy;
  // per-cycle scalars (all formed on the host in double)
  T dt, hdt;
  T dtd[3];          // dt/dx, dt/dy, dt/dz
  T hdtd[3];         // 0.5*(dt/dx), ...
  T invd[3];         // 1/dx, ...
  T width[3];        // dx, dy, dz
  double wd[3];      // dx, dy, dz in double (diagnostics)
  T emfq[3];         // 0.5*(0.5*(dt/dx)), ...
  T bx1d;            // uniform B_x of the 1D step
  double ldh[2];     // 0.5*dt/dy as the reference's 80-bit value, hi + lo (FP64 only)
  // state
  CPtr<T> U;
  CPtr<T> W;
  Ptr<T> U2;
  Ptr<T> W2;
  CPtr<T> fb[3];
  Ptr<T> fb2[3];
  Ptr<T> nfb[3];
  // low parts of the face field when split_faces<T>() (else unused)
  CPtr<T> fblo[3];
  Ptr<T> fb2lo[3];
  Ptr<T> nfblo[3];
  // work
  Ptr<T> WL;
  Ptr<T> WR;
  Ptr<T> S;
  Ptr<T> UL[3];
  Ptr<T> UR[3];
  Ptr<T> F[3];           // predictor fluxes
  Ptr<T> Fc[3];          // corrector fluxes
  Ptr<T> Ec;
  Ptr<T> zend;       // per line along the contiguous axis (z in 3D, x in 2D): left edge state
                     // of the first cell (7) and right edge state of the last cell (7),
                     // planes [v][line], for the end faces
  Ptr<T> E[3];
  // floor counters (FLOOR_WORDS words; null: not counted): the values floored in
  // the running cycle per call site and variable, their totals since the state was
  // set, and the most one call site floored in one cycle
  unsigned long long* fcnt;
  // element counts
  PMHD_FI int64_t ncell() const { return ny * nx * nz; }
  PMHD_FI int64_t n_along(int d) const { return d == 0 ? nx : (d == 1 ? ny : nz); }
  PMHD_FI int64_t fny(int d) const { return ny + (d == 1); }
  PMHD_FI int64_t fnx(int d) const { return nx + (d == 0); }
  PMHD_FI int64_t fnz(int d) const { return nz + (d == 2); }
  PMHD_FI int64_t nface(int d) const { return fny(d) * fnx(d) * fnz(d); }
  PMHD_FI int64_t cid(int64_t j, int64_t i, int64_t k) const { return (j * nx + i) * nz + k; }
  PMHD_FI int64_t fid(int d, int64_t j, int64_t i, int64_t k) const {
    return (j * fnx(d) + i) * fnz(d) + k;
  }
  PMHD_FI int64_t eny() const { return ny + 1; }
  PMHD_FI int64_t enx() const { return nx + 1; }
  PMHD_FI int64_t enz() const { return ndim == 3 ? nz + 1 : 1; }
  PMHD_FI int64_t nedge() const { return eny() * enx() * enz(); }
  PMHD_FI int64_t eid(int64_t j, int64_t i, int64_t k) const { return (j * enx() + i) * enz() + k; }
  // extended-cell array for slopes: one ghost each side along d
  PMHD_FI int64_t sny(int d) const { return ny + 2 * (d == 1); }
  PMHD_FI int64_t snx(int d) const { return nx + 2 * (d == 0); }
  PMHD_FI int64_t snz(int d) const { return nz + 2 * (d == 2); }
  PMHD_FI int64_t nslope(int d) const { return sny(d) * snx(d) * snz(d); }
  // index box (n0, n1, n2) = (y, x, z) extents of one launch kind along d
  PMHD_FI void extent(int kind, int d, int64_t n[3]) const {
    if (kind == EXT_FACE) {
      n[0] = fny(d); n[1] = fnx(d); n[2] = fnz(d);
    } else if (kind == EXT_SLOPE) {
      n[0] = sny(d); n[1] = snx(d); n[2] = snz(d);
    } else if (kind == EXT_EDGE) {
      n[0] = eny(); n[1] = enx(); n[2] = enz();
    } else if (kind == EXT_LINE || kind == EXT_ENDS) {
      n[0] = fny(d); n[1] = fnx(d); n[2] = fnz(d);
      n[d == 0 ? 1 : (d == 1 ? 0 : 2)] = kind == EXT_LINE ? 1 : 2;
    } else if (kind == EXT_SEG) {
      n[0] = fny(d); n[1] = fnx(d); n[2] = fnz(d);
      const int64_t s = seg[d] > 0 ? seg[d] : 1;
      n[d == 0 ? 1 : (d == 1 ? 0 : 2)] = (n_along(d) + s) / s;   // runs over n + 1 faces
    } else {
      n[0] = ny; n[1] = nx; n[2] = nz;
    }
  }
};

struct I3 {
  int64_t j, i, k;
};

// Floor call sites, in the order of the reference's applyFloors calls: the
// predictor's left and right states along x, y, z (sites 0-5;
// Source/EvolveFunctions_MHD.py:278-280, 513-519, 911-922), then the 2D post-CTU
// left and right states along x and y (sites 6-9; :562-576).  Word 2*site counts
// floored densities, 2*site+1 floored pressures (predictor) or energies (post-CTU).
enum { FLOOR_SITES = 10, FLOOR_CUR = 0, FLOOR_TOT = 2 * FLOOR_SITES, FLOOR_MAX = 4 * FLOOR_SITES,
       FLOOR_WORDS = 64 };

PMHD_FI void count_add(unsigned long long* p, unsigned long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicAdd(p, v);
#else
  __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
#endif
}

// the end of a cycle's counting (run by one thread before the update of the cycle,
// after every launch that floors): totals and per-cycle maxima, then a clear
PMHD_FI void fold_floor_counts(unsigned long long* f) {
  for (int k = 0; k < 2 * FLOOR_SITES; ++k) {
    const unsigned long long v = f[FLOOR_CUR + k];
    f[FLOOR_TOT + k] += v;
    if (v > f[FLOOR_MAX + k]) f[FLOOR_MAX + k] = v;
    f[FLOOR_CUR + k] = 0ull;
  }
}

// apply_floors to a 7-variable state, counting the floored values at `site`
template <class T>
PMHD_FI void floors_counted(const Ctx<T>& c, int site, T v[7], T min_last) {
  const bool lo0 = v[0] < c.min_dens;
  const bool lo1 = v[6] < min_last;
  apply_floors(v, c.min_dens, min_last);
  if (c.fcnt && (lo0 || lo1)) {
    if (lo0) count_add(c.fcnt + FLOOR_CUR + 2 * site, 1ull);
    if (lo1) count_add(c.fcnt + FLOOR_CUR + 2 * site + 1, 1ull);
  }
}
PMHD_FI int64_t co(const I3& p, int d) { return d == 0 ? p.i : (d == 1 ? p.j : p.k); }
PMHD_FI I3 with(I3 p, int d, int64_t v) {
  if (d == 0) p.i = v; else if (d == 1) p.j = v; else p.k = v;
  return p;
}
PMHD_FI I3 unflat(int64_t idx, int64_t ny, int64_t nx, int64_t nz) {
  (void)ny;
  I3 p;
  p.k = idx % nz;
  int64_t t = idx / nz;
  p.i = t % nx;
  p.j = t / nx;
  return p;
}

// ghost maps along one axis of n cells
// reconstruction (2 ghosts): callers pass -2n <= e < 3n, so two conditional wraps
// give e mod n without a division
PMHD_FI int64_t map2(int64_t e, int64_t n, int bc) {
  if (bc == OUTFLOW) return e < 0 ? 0 : (e >= n ? n - 1 : e);
  if (e < 0) e += n;
  if (e < 0) e += n;
  if (e >= n) e -= n;
  if (e >= n) e -= n;
  return e;
}
PMHD_FI int64_t map1(int64_t e, int64_t n, int bc) {  // one ghost each side
  if (e < 0) return bc == OUTFLOW ? 0 : n - 1;
  if (e >= n) return bc == OUTFLOW ? n - 1 : 0;
  return e;
}
PMHD_FI int64_t edge_lo(int64_t e, int64_t n, int bc) { return e == 0 ? (bc == PERIODIC ? n - 1 : 0) : e - 1; }
PMHD_FI int64_t edge_hi(int64_t e, int64_t n, int bc) { return e == n ? (bc == PERIODIC ? 0 : n - 1) : e; }

// rotated slot r <- 8-variable index, direction d (0 x, 1 y, 2 z)
PMHD_FI int rot8(int d, int r) {
  switch (r) {
    case 0: return 0;
    case 1: return 1 + d;
    case 2: return 1 + (d + 1) % 3;
    case 3: return 1 + (d + 2) % 3;
    case 4: return 4 + (d + 1) % 3;
    case 5: return 4 + (d + 2) % 3;
    default: return 7;
  }
}
// rotated slot r -> directional 7-vector slot
PMHD_FI int rot_to_dir(int d, int r) {
  switch (r) {
    case 1: return 1 + d;
    case 2: return 1 + (d + 1) % 3;
    case 3: return 1 + (d + 2) % 3;
    default: return r;
  }
}

// Element i of the array that starts at base (one variable plane of a state or work
// array).  On the GPU the byte offset is formed in 32 bits -- every plane is below
// 4 GiB (checked when a grid is created) -- so a load or store takes the plane's
// uniform base address and a 32-bit offset, with no 64-bit address arithmetic per
// element.  Accessor types (the adjoint's dual arrays) index directly.
template <class T>
PMHD_FI T ld(const T* base, int64_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + uint32_t(uint64_t(i) * sizeof(T)));
#else
  return base[i];
#endif
}
template <class T>
PMHD_FI T ld(T* base, int64_t i) {
  return ld(static_cast<const T*>(base), i);
}
template <class P, class = typename std::enable_if<std::is_class<P>::value>::type>
PMHD_FI auto ld(const P& base, int64_t i) -> decltype(base[i]) {
  return base[i];
}
template <class T, class V>
PMHD_FI void st(T* base, int64_t i, const V& v) {
#if defined(__HIP_DEVICE_COMPILE__)
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + uint32_t(uint64_t(i) * sizeof(T))) = T(v);
#else
  base[i] = T(v);
#endif
}
template <class P, class V, class = typename std::enable_if<std::is_class<P>::value>::type>
PMHD_FI void st(const P& base, int64_t i, const V& v) {
  base[i] = v;
}

// A kernel's result: in a unit that defines PMHD_NT_STORES (built for grids whose
// launch outputs are far larger than the caches) a non-temporal store, whose value
// streams past the caches and leaves them to the neighbour re-reads of the launch's
// inputs; else an ordinary store (the next launch may find the value cached).
template <class C, class T, class V>
PMHD_FI void put(const C& c, T* base, int64_t i, const V& v) {
  (void)c;
#if defined(__HIP_DEVICE_COMPILE__) && defined(PMHD_NT_STORES)
  __builtin_nontemporal_store(T(v), reinterpret_cast<T*>(reinterpret_cast<char*>(base) +
                                                        uint32_t(uint64_t(i) * sizeof(T))));
#else
  st(base, i, v);
#endif
}
template <class C, class P, class V, class = typename std::enable_if<std::is_class<P>::value>::type>
PMHD_FI void put(const C&, const P& base, int64_t i, const V& v) {
  base[i] = v;
}

// An input only this element reads (no neighbour re-reads it): in the
// non-temporal unit it streams past the caches like the results
template <class T>
PMHD_FI T ld_once(const T* base, int64_t i) {
#if defined(__HIP_DEVICE_COMPILE__) && defined(PMHD_NT_STORES)
  return __builtin_nontemporal_load(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                                               uint32_t(uint64_t(i) * sizeof(T))));
#else
  return ld(base, i);
#endif
}
template <class T>
PMHD_FI T ld_once(T* base, int64_t i) {
  return ld_once(static_cast<const T*>(base), i);
}
template <class P, class = typename std::enable_if<std::is_class<P>::value>::type>
PMHD_FI auto ld_once(const P& base, int64_t i) -> decltype(base[i]) {
  return base[i];
}

template <class P, class T>
PMHD_FI void load_rot8(const P& A, int64_t N, int64_t c, int d, T w[7]) {
#pragma unroll
  for (int r = 0; r < 7; ++r) w[r] = ld(A + rot8(d, r) * N, c);
}
template <class T>
PMHD_FI void to_rot(const T dir[7], int d, T rot[7]) {
#pragma unroll
  for (int r = 0; r < 7; ++r) rot[r] = dir[rot_to_dir(d, r)];
}
template <class T>
PMHD_FI void from_rot(const T rot[7], int d, T dir[7]) {
#pragma unroll
  for (int r = 0; r < 7; ++r) dir[rot_to_dir(d, r)] = rot[r];
}
template <class P, class T>
PMHD_FI void load7(const P& A, int64_t NF, int64_t f, T v[7]) {
#pragma unroll
  for (int r = 0; r < 7; ++r) v[r] = ld(A + r * NF, f);
}
template <class P, class T>
PMHD_FI void load7_once(const P& A, int64_t NF, int64_t f, T v[7]) {
#pragma unroll
  for (int r = 0; r < 7; ++r) v[r] = ld_once(A + r * NF, f);
}
template <class C, class P, class T>
PMHD_FI void store7(const C& c, const P& A, int64_t NF, int64_t f, const T v[7]) {
#pragma unroll
  for (int r = 0; r < 7; ++r) put(c, A + r * NF, f, v[r]);
}

// ===========================================================================
// CFL speeds of one cell
// ===========================================================================
// of the state (U, fb): the current one, or the one an update has just written
template <class T, class PU, class PF>
PMHD_FI void k_cfl(const Ctx<T>& c, const PU& U, const PF* fb, I3 p, T out[3]) {
  const int64_t N = c.ncell();
  const int64_t idx = c.cid(p.j, p.i, p.k);
  if (c.physics == 1) {
    T h[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) h[v] = ld(U + v * N, idx);
    hydro::cell_speeds(h, c.gamma, out);
    return;
  }
  T u[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) u[v] = ld(U + v * N, idx);
  T bf[3] = {T(0), T(0), T(0)};
  if (c.ndim >= 2) {
    bf[0] = T(ld(fb[0], c.fid(0, p.j, p.i, p.k)));
    bf[1] = T(ld(fb[1], c.fid(1, p.j, p.i, p.k)));
  }
  if (c.ndim == 3) bf[2] = T(ld(fb[2], c.fid(2, p.j, p.i, p.k)));
  cell_speeds(u, bf, c.ndim, c.gamma, out);
}

// ===========================================================================
// reconstruction: characteristic-limited slopes on the cells -1..n along d
// ===========================================================================
template <class T>
PMHD_FI void k_slopes(const Ctx<T>& c, int d, int bc, I3 q) {
  const int64_t N = c.ncell();
  const int64_t idx = (q.j * c.snx(d) + q.i) * c.snz(d) + q.k;
  const int64_t n = c.n_along(d);
  const int64_t e = co(q, d) - 1;
  I3 pc = with(q, d, map2(e, n, bc));
  I3 pl = with(q, d, map2(e - 1, n, bc));
  I3 pr = with(q, d, map2(e + 1, n, bc));
  T wl[7], wc[7], wr[7];
  load_rot8(c.W, N, c.cid(pl.j, pl.i, pl.k), d, wl);
  load_rot8(c.W, N, c.cid(pc.j, pc.i, pc.k), d, wc);
  load_rot8(c.W, N, c.cid(pr.j, pr.i, pr.k), d, wr);
  T bn = ld(c.W + (4 + d) * N, c.cid(pc.j, pc.i, pc.k));
  PrimEig<T> E;
  eigen_prim(wc, bn, c.gamma, E);
  T ps[7];
  limited_slope(E, wl, wc, wr, ps);
  store7(c, c.S, c.nslope(d), idx, ps);
}

// traced edge states of cell p (coordinate e along d, 0 <= e < n) in directional
// order: dla is the right state of face e, dra the left state of face e+1.  The
// slopes of cells e-1, e, e+1 come from the slope array (PPM); with from_s false
// (PLM) the cell's own limited slope is formed here from the same eigen-system as
// the tracing, which is what the slope launch computes for that cell.
// the edge states of cell p from its own and its neighbours' primitive vectors
// (rotated) and its normal field bn
template <class T>
PMHD_FI void cell_edge_core(const Ctx<T>& c, int d, I3 p, bool from_s, const T wl[7], const T wc[7], const T wr[7],
                            T bn, T dla[7], T dra[7]) {
  const int64_t e = co(p, d);
  PrimEig<T> E;
  eigen_prim(wc, bn, c.gamma, E);
  T la[7], ra[7];
  if (from_s) {
    const int64_t NS = c.nslope(d);
    auto sidx = [&](int64_t ee) {
      I3 s = with(p, d, ee + 1);
      return (s.j * c.snx(d) + s.i) * c.snz(d) + s.k;
    };
    T sl[7], sc[7], sr[7];
    load7(c.S, NS, sidx(e), sc);
    if (c.order == 2) {
      load7(c.S, NS, sidx(e - 1), sl);
      load7(c.S, NS, sidx(e + 1), sr);
    } else {
      // PLM traces with the cell's own slope only (its neighbours' slopes are not
      // read, so the adjoint does not differentiate with respect to them)
#pragma unroll
      for (int v = 0; v < 7; ++v) sl[v] = sr[v] = sc[v];
    }
    trace_cell(E, wl, wc, wr, sl, sc, sr, c.dtd[d], c.order, la, ra);
  } else {
    T sc[7];
    limited_slope(E, wl, wc, wr, sc);
    trace_cell(E, wl, wc, wr, sc, sc, sc, c.dtd[d], c.order, la, ra);
  }
  from_rot(la, d, dla);
  from_rot(ra, d, dra);
}

template <class T>
PMHD_FI void cell_edge_states(const Ctx<T>& c, int d, int bc, I3 p, bool from_s, T dla[7], T dra[7]) {
  const int64_t N = c.ncell();
  const int64_t idx = c.cid(p.j, p.i, p.k);
  const int64_t n = c.n_along(d);
  const int64_t e = co(p, d);
  I3 pl = with(p, d, map2(e - 1, n, bc));
  I3 pr = with(p, d, map2(e + 1, n, bc));
  T wl[7], wc[7], wr[7];
  load_rot8(c.W, N, c.cid(pl.j, pl.i, pl.k), d, wl);
  load_rot8(c.W, N, idx, d, wc);
  load_rot8(c.W, N, c.cid(pr.j, pr.i, pr.k), d, wr);
  T bn = ld(c.W + (4 + d) * N, idx);
  cell_edge_core(c, d, p, from_s, wl, wc, wr, bn, dla, dra);
}

// edges and characteristic tracing of each cell; writes the two faces it owns
// (from_s false: PLM with the cell's own limited slope formed here)
template <class T>
PMHD_FI void k_edges(const Ctx<T>& c, int d, int bc, I3 p, bool from_s = true) {
  const int64_t n = c.n_along(d);
  const int64_t e = co(p, d);
  T dla[7], dra[7];
  cell_edge_states(c, d, bc, p, from_s, dla, dra);
  const int64_t NF = c.nface(d);
  I3 f0 = with(p, d, e);
  I3 f1 = with(p, d, e + 1);
  store7(c, c.WR, NF, c.fid(d, f0.j, f0.i, f0.k), dla);
  store7(c, c.WL, NF, c.fid(d, f1.j, f1.i, f1.k), dra);
  if (e == 0) {
    if (bc == OUTFLOW) {
      store7(c, c.WL, NF, c.fid(d, f0.j, f0.i, f0.k), dla);
    } else {
      I3 fn = with(p, d, n);
      store7(c, c.WR, NF, c.fid(d, fn.j, fn.i, fn.k), dla);
    }
  }
  if (e == n - 1) {
    if (bc == OUTFLOW) {
      store7(c, c.WR, NF, c.fid(d, f1.j, f1.i, f1.k), dra);
    } else {
      I3 fz = with(p, d, 0);
      store7(c, c.WL, NF, c.fid(d, fz.j, fz.i, fz.k), dra);
    }
  }
}

// first-order interface states (reconstructOrder 0)
template <class T>
PMHD_FI void k_copy_states(const Ctx<T>& c, int d, int bc, I3 f) {
  const int64_t N = c.ncell();
  const int64_t idx = c.fid(d, f.j, f.i, f.k);
  const int64_t n = c.n_along(d);
  const int64_t e = co(f, d);
  I3 pl = with(f, d, map1(e - 1, n, bc));
  I3 pr = with(f, d, map1(e, n, bc));
  T wl[7], wr[7], dl[7], dr[7];
  load_rot8(c.W, N, c.cid(pl.j, pl.i, pl.k), d, wl);
  load_rot8(c.W, N, c.cid(pr.j, pr.i, pr.k), d, wr);
  from_rot(wl, d, dl);
  from_rot(wr, d, dr);
  const int64_t NF = c.nface(d);
  store7(c, c.WL, NF, idx, dl);
  store7(c, c.WR, NF, idx, dr);
}

// ===========================================================================
// predictor: floors, MHD source terms on the reconstructed transverse B,
// prim -> cons, Roe flux
// ===========================================================================
template <class T, class P>
PMHD_FI T face_val(const Ctx<T>& c, const P& fb, int a, I3 p, int64_t fa) {
  // value of face array of axis a at face coordinate fa along a, other coords of p
  I3 q = with(p, a, fa);
  return T(ld(fb, c.fid(a, q.j, q.i, q.k)));
}

// source term of the 3D reconstruction on the (ghost-mapped) cell e along d
template <class T>
PMHD_FI void recon_b_term_3d(const Ctx<T>& c, int d, int bc, I3 p, int64_t e, T out[2]) {
  const int64_t N = c.ncell();
  const int64_t n = c.n_along(d);
  const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
  I3 pm = with(p, d, map1(e, n, bc));
  // normal derivative with ghost faces
  int64_t flo, fhi;
  if (e < 0) flo = bc == OUTFLOW ? 0 : n - 1; else flo = e;
  if (e + 1 > n) fhi = bc == OUTFLOW ? n : 1; else fhi = e + 1;
  T dbn = (face_val(c, c.fb[d], d, p, fhi) - face_val(c, c.fb[d], d, p, flo)) * c.invd[d];
  T db1 = (face_val(c, c.fb[t1], t1, pm, co(pm, t1) + 1) - face_val(c, c.fb[t1], t1, pm, co(pm, t1))) * c.invd[t1];
  T db2 = (face_val(c, c.fb[t2], t2, pm, co(pm, t2) + 1) - face_val(c, c.fb[t2], t2, pm, co(pm, t2))) * c.invd[t2];
  T term1 = minmod_recon(dbn, db1, d == 1);
  T term2 = minmod_recon(dbn, db2, false);
  int64_t cm = c.cid(pm.j, pm.i, pm.k);
  T rho = ld(c.W, cm);
  T v1 = ld(c.U + (1 + t1) * N, cm) / rho;
  T v2 = ld(c.U + (1 + t2) * N, cm) / rho;
  out[0] = c.hdt * (v1 * term1);
  out[1] = c.hdt * (v2 * term2);
}

// The reconstruction source term of cell m (ghost-mapped) on the line of faces
// through f along d, added to the edge state of that cell at the face between m
// and its neighbour: in 3D the two transverse-B terms of recon_b_term_3d; in 2D
// the B_y (x faces) or B_x (y faces) term v_t dB_n/dn of the cell
// (getByReconSourceTerms_2D / getBxReconSourceTerms_2D), from the normal field on
// the cell's two faces.  The same value serves as the left term of face m+1 and
// the right term of face m, so a sweep forms it once per cell.
template <class T>
PMHD_FI void cell_recon_src(const Ctx<T>& c, int d, I3 f, int64_t m, T out[2]) {
  if (c.ndim == 3) {
    recon_b_term_3d(c, d, d == 2 ? c.bc_zrec : c.bc[d], f, m, out);
  } else if (c.ndim == 2) {
    const int64_t N = c.ncell();
    const int64_t n = c.n_along(d);
    const int bc = c.bc[d];
    I3 pm = with(f, d, map1(m, n, bc));
    const int vslot = d == 0 ? 2 : 1;  // x faces use v_y, y faces use v_x
    const T v = ld(c.W + vslot * N, c.cid(pm.j, pm.i, pm.k));
    const int64_t flo = m < 0 ? (bc == OUTFLOW ? 0 : n - 1) : m;
    const int64_t fhi = m + 1 > n ? (bc == OUTFLOW ? n : 1) : m + 1;
    out[0] = c.hdt * v * ((face_val(c, c.fb[d], d, f, fhi) - face_val(c, c.fb[d], d, f, flo)) * c.invd[d]);
    out[1] = T(0);
  } else {
    out[0] = out[1] = T(0);
  }
}

// the predictor's conserved states and flux of a face of direction d from its
// prepared primitive states (floors and source terms applied) and normal field
// (the context supplies gamma and the entropy fix; its number type may differ)
template <class T, class U>
PMHD_FI void predict_states(const Ctx<U>& c, int ke, const T wl[7], const T wr[7], T bn, T ul[7], T ur[7]) {
  prim_to_cons(wl, bn, T(c.gamma), ke, ul);
  prim_to_cons(wr, bn, T(c.gamma), ke, ur);
}
template <int S, class T, class U>
PMHD_FI void predict_roe(const Ctx<U>& c, int d, const T wl[7], const T wr[7], const T ul[7], const T ur[7], T bn,
                         T df[7]) {
  T rwl[7], rwr[7], rul[7], rur[7], rf[7];
  to_rot(wl, d, rwl);
  to_rot(wr, d, rwr);
  to_rot(ul, d, rul);
  to_rot(ur, d, rur);
  roe_flux<S>(rwl, rwr, rul, rur, bn, T(c.gamma), c.efix != 0, T(c.eps), rf);
  from_rot(rf, d, df);
}
template <int S, class T, class U>
PMHD_FI void predict_flux(const Ctx<U>& c, int d, int ke, const T wl[7], const T wr[7], T bn, T ul[7], T ur[7],
                          T df[7]) {
  predict_states(c, ke, wl, wr, bn, ul, ur);
  predict_roe<S>(c, d, wl, wr, ul, ur, bn, df);
}

// predictor of face f from its reconstructed states wl, wr (directional order;
// modified in place) and the source terms of its left and right cells
// (cell_recon_src of cells e-1 and e)
template <int S, class T>
PMHD_FI void predict_face_src(const Ctx<T>& c, int d, I3 f, T wl[7], T wr[7], const T cl[2], const T cr[2]) {
  const int64_t NF = c.nface(d);
  const int64_t idx = c.fid(d, f.j, f.i, f.k);
  floors_counted(c, 2 * d, wl, c.min_pres);
  floors_counted(c, 2 * d + 1, wr, c.min_pres);
  T bn;
  int ke = 0;
  if (c.ndim == 3) {
    wl[4] = wl[4] + cl[0];
    wl[5] = wl[5] + cl[1];
    wr[4] = wr[4] + cr[0];
    wr[5] = wr[5] + cr[1];
    bn = T(ld(c.fb[d], idx));
    ke = d;
  } else if (c.ndim == 2) {
    const int slot = d == 0 ? 4 : 5;
    wl[slot] = wl[slot] + cl[0];
    wr[slot] = wr[slot] + cr[0];
    bn = T(ld(c.fb[d], idx));
  } else {
    bn = c.bx1d;
  }
  // the states are stored before the flux is formed (they need not stay live
  // through the Roe solver)
  T ul[7], ur[7], df[7];
  predict_states(c, ke, wl, wr, bn, ul, ur);
  store7(c, c.UL[d], NF, idx, ul);
  store7(c, c.UR[d], NF, idx, ur);
  predict_roe<S>(c, d, wl, wr, ul, ur, bn, df);
  store7(c, c.F[d], NF, idx, df);
}

template <int S, class T>
PMHD_FI void predict_face(const Ctx<T>& c, int d, I3 f, T wl[7], T wr[7]) {
  const int64_t e = co(f, d);
  T cl[2], cr[2];
  cell_recon_src(c, d, f, e - 1, cl);
  cell_recon_src(c, d, f, e, cr);
  predict_face_src<S>(c, d, f, wl, wr, cl, cr);
}

template <int S, class T>
PMHD_FI void k_predict(const Ctx<T>& c, int d, I3 f) {
  const int64_t NF = c.nface(d);
  const int64_t idx = c.fid(d, f.j, f.i, f.k);
  T wl[7], wr[7];
  load7(c.WL, NF, idx, wl);
  load7(c.WR, NF, idx, wr);
  predict_face<S>(c, d, f, wl, wr);
}

// ===========================================================================
// fused predictor sweep: reconstruction, tracing and the predictor of a run of
// seg[d] consecutive faces along d.  q's coordinate along d is the run index;
// the run walks its faces in order and keeps the left state of the next face
// (the right edge state of the cell before it) in registers, so each face costs
// one cell reconstruction, plus one for the cell before the run.  The face
// states are those the edge launch would store: face 0 takes the left edge of
// cell 0 on both sides with outflow and the right edge of cell n-1 on its left
// when periodic; face n takes the right edge of cell n-1 on both sides with
// outflow and the left edge of cell 0 on its right when periodic.  order 2
// reads the slopes of the slope launch; order 1 forms them here.
// ===========================================================================
// the faces [f0, f1) of the line through q along d.  The run walks the cells
// (logical indices, ghost-mapped) that those faces touch and keeps a window of
// primitive vectors in registers: each cell's own and right neighbour's vectors
// come from the previous step, and the vector two cells ahead is loaded one step
// early, so its latency overlaps the current cell's arithmetic (the loads of the
// window are issued before the step's stores, which the compiler could not move
// them across).
template <int S, class T>
PMHD_FI void k_sweep_run(const Ctx<T>& c, int d, int bc, I3 q, int64_t f0, int64_t f1) {
  const int64_t n = c.n_along(d);
  const int64_t N = c.ncell();
  const bool from_s = c.order == 2;
  const bool out = bc == OUTFLOW;
  // rotated primitive vector (slots 0-6) and normal field (slot 7) of logical cell e
  auto load_cell = [&](int64_t e, T w[8]) {
    const I3 pe = with(q, d, map2(e, n, bc));
    const int64_t id = c.cid(pe.j, pe.i, pe.k);
    load_rot8(c.W, N, id, d, w);
    w[7] = ld(c.W + (4 + d) * N, id);
  };
  // cells e0..e1: the left cell of face f0 (cell 0 itself for an outflow face 0)
  // through the right cell of face f1-1 (cell n-1 for an outflow face n)
  const int64_t e0 = (f0 == 0 && out) ? 0 : f0 - 1;
  const bool last_out = f1 - 1 == n && out;
  const int64_t e1 = last_out ? n - 1 : f1 - 1;
  T wm[8], w0[8], wp[8];
  load_cell(e0 - 1, wm);
  load_cell(e0, w0);
  load_cell(e0 + 1, wp);
  T prev[7], src_prev[2] = {T(0), T(0)};
  if (f0 == 0 && out) cell_recon_src(c, d, with(q, d, 0), -1, src_prev);
  for (int64_t e = e0; e <= e1; ++e) {
    T wn[8];
    if (e < e1) load_cell(e + 2, wn);
    T la[7], ra[7];
    cell_edge_core(c, d, with(q, d, map2(e, n, bc)), from_s, wm, w0, wp, w0[7], la, ra);
    T src[2];
    cell_recon_src(c, d, with(q, d, e < 0 ? 0 : e), e, src);
    if (e >= f0) {
      T wl[7];
#pragma unroll
      for (int v = 0; v < 7; ++v) wl[v] = (e == 0 && out) ? la[v] : prev[v];
      T wr[7];
#pragma unroll
      for (int v = 0; v < 7; ++v) wr[v] = la[v];
      predict_face_src<S>(c, d, with(q, d, e), wl, wr, src_prev, src);
    }
#pragma unroll
    for (int v = 0; v < 7; ++v) prev[v] = ra[v];
    src_prev[0] = src[0];
    src_prev[1] = src[1];
    if (e < e1) {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        wm[v] = w0[v];
        w0[v] = wp[v];
        wp[v] = wn[v];
      }
    }
  }
  if (last_out) {
    // face n with outflow: the right edge state of cell n-1 on both sides
    T wl[7], wr[7], src[2];
#pragma unroll
    for (int v = 0; v < 7; ++v) wl[v] = wr[v] = prev[v];
    cell_recon_src(c, d, with(q, d, n), n, src);
    predict_face_src<S>(c, d, with(q, d, n), wl, wr, src_prev, src);
  }
}

// Lines along the contiguous axis d (z in 3D, x in 2D): a line is numbered by its
// other two coordinates, slowest first.
template <class T>
PMHD_FI int64_t line_count(const Ctx<T>& c, int d) {
  return d == 2 ? c.ny * c.nx : (d == 0 ? c.ny * c.nz : c.nx * c.nz);
}
template <class T>
PMHD_FI int64_t line_of(const Ctx<T>& c, int d, I3 q) {
  return d == 2 ? q.j * c.nx + q.i : (d == 0 ? q.j * c.nz + q.k : q.i * c.nz + q.k);
}
// keep the left edge state of the first cell and the right edge state of the last
// cell of the line through q along d for the end faces (k_sweep_end)
template <class T>
PMHD_FI void store_line_end(const Ctx<T>& c, int d, I3 q, int side, const T s[7]) {
  const int64_t L = line_count(c, d);
  const int64_t l = line_of(c, d, q);
#pragma unroll
  for (int v = 0; v < 7; ++v) put(c, c.zend + (7 * side + v) * L, l, s[v]);
}
template <class T>
PMHD_FI void k_line_ends(const Ctx<T>& c, int d, int bc, I3 q) {
  const bool from_s = c.order == 2;
  T la[7], ra[7];
  cell_edge_states(c, d, bc, with(q, d, 0), from_s, la, ra);
  store_line_end(c, d, q, 0, la);
  cell_edge_states(c, d, bc, with(q, d, c.n_along(d) - 1), from_s, la, ra);
  store_line_end(c, d, q, 1, ra);
}
// end face 0 (co(q, d) == 0) or n (co(q, d) == 1) of the line through q along d
// from the stored edge states: face 0 takes the left edge of cell 0 on both sides
// (outflow) or the right edge of cell n-1 and the left edge of cell 0 (periodic);
// face n takes the right edge of cell n-1 on both sides (outflow) or the right edge
// of cell n-1 and the left edge of cell 0 (periodic)
template <int S, class T>
PMHD_FI void k_sweep_end(const Ctx<T>& c, int d, int bc, I3 q) {
  const int64_t L = line_count(c, d);
  const int64_t l = line_of(c, d, q);
  T l0[7], rn[7], wl[7], wr[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    l0[v] = ld(c.zend + v * L, l);
    rn[v] = ld(c.zend + (7 + v) * L, l);
  }
  const bool face0 = co(q, d) == 0;
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    wl[v] = face0 ? (bc == OUTFLOW ? l0[v] : rn[v]) : rn[v];
    wr[v] = face0 ? l0[v] : (bc == OUTFLOW ? rn[v] : l0[v]);
  }
  predict_face<S>(c, d, with(q, d, face0 ? 0 : c.n_along(d)), wl, wr);
}

template <int S, class T>
PMHD_FI void k_sweep(const Ctx<T>& c, int d, int bc, I3 q) {
  const int64_t n = c.n_along(d);
  const int64_t len = c.seg[d] > 0 ? c.seg[d] : 1;
  const int64_t f0 = co(q, d) * len;
  k_sweep_run<S>(c, d, bc, q, f0, f0 + len < n + 1 ? f0 + len : n + 1);
}

// ===========================================================================
// cell-centred electric field
// ===========================================================================
// cell E of a cell from its density, momenta, velocities and cell B (3D: (B x M)/rho;
// 2D: v_y B_x - v_x B_y)
template <class T>
PMHD_FI void store_cell_e(const Ctx<T>& c, int64_t idx, T rho, T mx, T my, T mz, T vx, T vy, T bx, T by, T bz) {
  const int64_t N = c.ncell();
  if (c.ndim == 3) {
    put(c, c.Ec, idx, (by * mz - bz * my) / rho);
    put(c, c.Ec + N, idx, (bz * mx - bx * mz) / rho);
    put(c, c.Ec + 2 * N, idx, (bx * my - by * mx) / rho);
  } else {
    put(c, c.Ec, idx, vy * bx - vx * by);
  }
}

template <class T>
PMHD_FI void k_cell_e(const Ctx<T>& c, I3 p) {
  const int64_t N = c.ncell();
  const int64_t idx = c.cid(p.j, p.i, p.k);
  if (c.ndim == 3) {
    T rho = ld(c.U, idx), mx = ld(c.U + N, idx), my = ld(c.U + 2 * N, idx), mz = ld(c.U + 3 * N, idx);
    T bx = ld(c.W + 4 * N, idx), by = ld(c.W + 5 * N, idx), bz = ld(c.W + 6 * N, idx);
    store_cell_e<T>(c, idx, rho, mx, my, mz, T(0), T(0), bx, by, bz);
  } else {
    store_cell_e<T>(c, idx, T(0), T(0), T(0), T(0), ld(c.W + N, idx), ld(c.W + 2 * N, idx), ld(c.W + 4 * N, idx),
                 ld(c.W + 5 * N, idx), T(0));
  }
}

// half-step cell electric field from the predictor fluxes and half-step faces
template <class T>
PMHD_FI void k_half_e(const Ctx<T>& c, I3 p) {
  const int64_t N = c.ncell();
  const int64_t idx = c.cid(p.j, p.i, p.k);
  const int64_t NX = c.nface(0), NY = c.nface(1);
  int64_t fx0 = c.fid(0, p.j, p.i, p.k), fx1 = c.fid(0, p.j, p.i + 1, p.k);
  int64_t fy0 = c.fid(1, p.j, p.i, p.k), fy1 = c.fid(1, p.j + 1, p.i, p.k);
  if (c.ndim == 3) {
    const int64_t NZ = c.nface(2);
    int64_t fz0 = c.fid(2, p.j, p.i, p.k), fz1 = c.fid(2, p.j, p.i, p.k + 1);
    T h[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      h[v] = ld_once(c.U + v * N, idx) - c.hdtd[0] * (ld(c.F[0] + v * NX, fx1) - ld(c.F[0] + v * NX, fx0))
             - c.hdtd[1] * (ld(c.F[1] + v * NY, fy1) - ld(c.F[1] + v * NY, fy0))
             - c.hdtd[2] * (ld(c.F[2] + v * NZ, fz1) - ld(c.F[2] + v * NZ, fz0));
    }
    T bx = T(0.5) * (T(ld(c.nfb[0], fx0)) + T(ld(c.nfb[0], fx1)));
    T by = T(0.5) * (T(ld(c.nfb[1], fy0)) + T(ld(c.nfb[1], fy1)));
    T bz = T(0.5) * (T(ld(c.nfb[2], fz0)) + T(ld(c.nfb[2], fz1)));
    put(c, c.Ec, idx, (by * h[3] - bz * h[2]) / h[0]);
    put(c, c.Ec + N, idx, (bz * h[1] - bx * h[3]) / h[0]);
    put(c, c.Ec + 2 * N, idx, (bx * h[2] - by * h[1]) / h[0]);
  } else {
    T h[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      T a = ld(c.U + v * N, idx) - c.hdtd[0] * (ld(c.F[0] + v * NX, fx1) - ld(c.F[0] + v * NX, fx0));
      h[v] = a - c.hdtd[1] * (ld(c.F[1] + v * NY, fy1) - ld(c.F[1] + v * NY, fy0));
    }
    T bx = (T(ld(c.nfb[0], fx0)) + T(ld(c.nfb[0], fx1))) / T(2);
    T by = (T(ld(c.nfb[1], fy0)) + T(ld(c.nfb[1], fy1))) / T(2);
    T vx = h[1] / h[0];
    T vy = h[2] / h[0];
    put(c, c.Ec, idx, vy * bx - vx * by);
  }
}

// ===========================================================================
// corner EMFs (Source/CornerEMF_HelperModule.py:440-952)
// ===========================================================================
template <class T>
PMHD_FI T fluxv(const Ctx<T>& c, const Ptr<T>* Fs, int d, int v, int64_t j, int64_t i, int64_t k) {
  return ld(Fs[d] + v * c.nface(d), c.fid(d, j, i, k));
}

// Fs: the flux set (predictor c.F or corrector c.Fc) the EMFs are formed from
template <class T>
PMHD_FI void k_emf(const Ctx<T>& c, const Ptr<T>* Fs, I3 p) {
  const int64_t N = c.ncell();
  const int64_t idx = c.eid(p.j, p.i, p.k);
  const int bx = c.bc[0], by = c.bc[1], bz = c.bc[2];
  const int64_t xl = edge_lo(p.i, c.nx, bx), xr = edge_hi(p.i, c.nx, bx);
  const int64_t yl = edge_lo(p.j, c.ny, by), yr = edge_hi(p.j, c.ny, by);
  if (c.ndim == 2) {
    T t0 = fluxv(c, Fs, 1, 0, p.j, xr, 0), t5 = fluxv(c, Fs, 1, 5, p.j, xr, 0);
    T lt0 = fluxv(c, Fs, 1, 0, p.j, xl, 0), lt5 = fluxv(c, Fs, 1, 5, p.j, xl, 0);
    T lf0 = fluxv(c, Fs, 0, 0, yr, p.i, 0), lf4 = fluxv(c, Fs, 0, 4, yr, p.i, 0);
    T ul0 = fluxv(c, Fs, 0, 0, yl, p.i, 0), ul4 = fluxv(c, Fs, 0, 4, yl, p.i, 0);
    T cE = ld(c.Ec, c.cid(yr, xr, 0)), lE = ld(c.Ec, c.cid(yr, xl, 0));
    T uE = ld(c.Ec, c.cid(yl, xr, 0)), ulE = ld(c.Ec, c.cid(yl, xl, 0));
    T x1 = upwind(lt0, lt5 - ul4 - ulE, lt5 - lf4 - lE, lt5 - T(0.5) * (lf4 + ul4 + lE + ulE));
    T x2 = upwind(t0, t5 - ul4 - uE, t5 - lf4 - cE, t5 - T(0.5) * (lf4 + ul4 + cE + uE));
    T y1 = upwind(ul0, -ul4 + lt5 - ulE, -ul4 + t5 - uE, -ul4 + T(0.5) * (lt5 + t5 - ulE - uE));
    T y2 = upwind(lf0, -lf4 + lt5 - lE, -lf4 + t5 - cE, -lf4 + T(0.5) * (lt5 + t5 - lE - cE));
    put(c, c.E[2], idx, T(0.25) * (x1 + x2 + y1 + y2));
    return;
  }
  const int64_t zl = edge_lo(p.k, c.nz, bz), zr = edge_hi(p.k, c.nz, bz);
  const auto Exc = c.Ec;
  const auto Eyc = c.Ec + N;
  const auto Ezc = c.Ec + 2 * N;
  {  // E_x (getXCornerEMF_3D): faces y (G) and z (H)
    T b0 = fluxv(c, Fs, 2, 0, yr, xr, p.k), b5 = fluxv(c, Fs, 2, 5, yr, xr, p.k);
    T ub0 = fluxv(c, Fs, 2, 0, yl, xr, p.k), ub5 = fluxv(c, Fs, 2, 5, yl, xr, p.k);
    T t0 = fluxv(c, Fs, 1, 0, p.j, xr, zr), t4 = fluxv(c, Fs, 1, 4, p.j, xr, zr);
    T bt0 = fluxv(c, Fs, 1, 0, p.j, xr, zl), bt4 = fluxv(c, Fs, 1, 4, p.j, xr, zl);
    T cE = ld(Exc, c.cid(yr, xr, zr)), bE = ld(Exc, c.cid(yr, xr, zl));
    T uE = ld(Exc, c.cid(yl, xr, zr)), ubE = ld(Exc, c.cid(yl, xr, zl));
    T z1 = upwind(bt0, ub5 - ubE, b5 - bE, T(0.5) * (ub5 + b5 - ubE - bE));
    T z2 = upwind(t0, ub5 - uE, b5 - cE, T(0.5) * (ub5 + b5 - uE - cE));
    T y1 = upwind(ub0, -bt4 - ubE, -t4 - uE, T(0.5) * (-bt4 - t4 - ubE - uE));
    T y2 = upwind(b0, -bt4 - bE, -t4 - cE, T(0.5) * (-bt4 - t4 - bE - cE));
    put(c, c.E[0], idx, T(0.25) * (b5 + ub5 - t4 - bt4 + y1 + y2 + z1 + z2));
  }
  {  // E_y (getYCornerEMF_3D): faces x (F) and z (H)
    T b0 = fluxv(c, Fs, 2, 0, yr, xr, p.k), b4 = fluxv(c, Fs, 2, 4, yr, xr, p.k);
    T lb0 = fluxv(c, Fs, 2, 0, yr, xl, p.k), lb4 = fluxv(c, Fs, 2, 4, yr, xl, p.k);
    T lf0 = fluxv(c, Fs, 0, 0, yr, p.i, zr), lf5 = fluxv(c, Fs, 0, 5, yr, p.i, zr);
    T blf0 = fluxv(c, Fs, 0, 0, yr, p.i, zl), blf5 = fluxv(c, Fs, 0, 5, yr, p.i, zl);
    T cE = ld(Eyc, c.cid(yr, xr, zr)), bE = ld(Eyc, c.cid(yr, xr, zl));
    T lE = ld(Eyc, c.cid(yr, xl, zr)), blE = ld(Eyc, c.cid(yr, xl, zl));
    T z1 = upwind(blf0, -lb4 - blE, -b4 - bE, T(0.5) * (-lb4 - b4 - blE - bE));
    T z2 = upwind(lf0, -lb4 - lE, -b4 - cE, T(0.5) * (-lb4 - b4 - lE - cE));
    T x1 = upwind(lb0, blf5 - blE, lf5 - lE, T(0.5) * (blf5 + lf5 - blE - lE));
    T x2 = upwind(b0, blf5 - bE, lf5 - cE, T(0.5) * (blf5 + lf5 - bE - cE));
    put(c, c.E[1], idx, T(0.25) * (lf5 + blf5 - b4 - lb4 + x1 + x2 + z1 + z2));
  }
  {  // E_z (getZCornerEMF_3D): faces x (F) and y (G)
    T t0 = fluxv(c, Fs, 1, 0, p.j, xr, zr), t5 = fluxv(c, Fs, 1, 5, p.j, xr, zr);
    T lt0 = fluxv(c, Fs, 1, 0, p.j, xl, zr), lt5 = fluxv(c, Fs, 1, 5, p.j, xl, zr);
    T lf0 = fluxv(c, Fs, 0, 0, yr, p.i, zr), lf4 = fluxv(c, Fs, 0, 4, yr, p.i, zr);
    T ul0 = fluxv(c, Fs, 0, 0, yl, p.i, zr), ul4 = fluxv(c, Fs, 0, 4, yl, p.i, zr);
    T cE = ld(Ezc, c.cid(yr, xr, zr)), lE = ld(Ezc, c.cid(yr, xl, zr));
    T uE = ld(Ezc, c.cid(yl, xr, zr)), ulE = ld(Ezc, c.cid(yl, xl, zr));
    T x1 = upwind(lt0, -ul4 - ulE, -lf4 - lE, T(-0.5) * (lf4 + ul4 + lE + ulE));
    T x2 = upwind(t0, -ul4 - uE, -lf4 - cE, T(-0.5) * (lf4 + ul4 + cE + uE));
    T y1 = upwind(ul0, lt5 - ulE, t5 - uE, T(0.5) * (lt5 + t5 - ulE - uE));
    T y2 = upwind(lf0, lt5 - lE, t5 - cE, T(0.5) * (lt5 + t5 - lE - cE));
    put(c, c.E[2], idx, T(0.25) * (lt5 + t5 - ul4 - lf4 + x1 + x2 + y1 + y2));
  }
}

// ===========================================================================
// CTU transverse corrections of the interface states
// ===========================================================================
template <class T>
PMHD_FI void ctu_source_3d(const Ctx<T>& c, int d, int bc, I3 f, int64_t e, T src[7]) {
  const int64_t N = c.ncell();
  const int64_t n = c.n_along(d);
  const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
  I3 pm = with(f, d, map1(e, n, bc));
  int64_t flo, fhi;
  if (e < 0) flo = bc == OUTFLOW ? 0 : n - 1; else flo = e;
  // the reference's z-outflow forward ghost face is face nz-1 (Source/CTU_HelperModule.py:897-898)
  const bool zq = d == 2 && !(c.fixq & PMHD_FIX_OUTFLOW_GHOSTS);
  if (e + 1 > n) fhi = bc == OUTFLOW ? (zq ? n - 1 : n) : 1; else fhi = e + 1;
  T dd = face_val(c, c.fb[d], d, f, fhi) - face_val(c, c.fb[d], d, f, flo);
  T d1 = face_val(c, c.fb[t1], t1, pm, co(pm, t1) + 1) - face_val(c, c.fb[t1], t1, pm, co(pm, t1));
  T d2 = face_val(c, c.fb[t2], t2, pm, co(pm, t2) + 1) - face_val(c, c.fb[t2], t2, pm, co(pm, t2));
  T dbn, db1, db2;
  if (d == 0) {
    dbn = dd * c.invd[d]; db1 = d1 * c.invd[t1]; db2 = d2 * c.invd[t2];
  } else {
    dbn = dd / c.width[d]; db1 = d1 / c.width[t1]; db2 = d2 / c.width[t2];
  }
  T mm1 = minmod_ctu(db1, dbn);
  T mm2 = minmod_ctu(db2, dbn);
  int64_t cm = c.cid(pm.j, pm.i, pm.k);
  T cb[3] = {ld(c.W + 4 * N, cm), ld(c.W + 5 * N, cm), ld(c.W + 6 * N, cm)};
  T cv[3] = {ld(c.W + 1 * N, cm), ld(c.W + 2 * N, cm), ld(c.W + 3 * N, cm)};
  T v1 = cv[t1], v2 = cv[t2], b1 = cb[t1], b2 = cb[t2];
  src[0] = T(0);
  src[1] = c.hdt * cb[0] * dbn;
  src[2] = c.hdt * cb[1] * dbn;
  src[3] = c.hdt * cb[2] * dbn;
  src[4] = c.hdt * v1 * (-mm2);
  src[5] = c.hdt * v2 * (-mm1);
  src[6] = c.hdt * (b1 * v1 * (-mm2) + b2 * v2 * (-mm1));
}

// one transverse direction tc of the 3D CTU update of a d-face state sitting in
// cell column cell_e (already ghost-mapped) -- flux difference plus EMF term
template <class T>
PMHD_FI void ctu_apply_3d(const Ctx<T>& c, int d, int tc, I3 pcell, int64_t hi_override,
                          I3 fedge, int64_t layer, T u[7]) {
  const int t1 = (d + 1) % 3;
  const int64_t NF = c.nface(tc);
  I3 lo = pcell;
  I3 hi = with(pcell, tc, co(pcell, tc) + 1);
  if (hi_override >= 0) hi = with(hi, 1, hi_override);
  int64_t flo = c.fid(tc, lo.j, lo.i, lo.k);
  int64_t fhi = c.fid(tc, hi.j, hi.i, hi.k);
  T h = c.hdtd[tc];
  const int vs[5] = {0, 1, 2, 3, 6};
  // x faces, y fluxes: the reference's 80-bit factor (FP64 instantiation only)
  const bool ext80 = is_f64<T>() && d == 0 && tc == 1 && !(c.fixq & PMHD_FIX_DTDY_80BIT);
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    int v = vs[q];
    T x = ld(c.F[tc] + v * NF, fhi) - ld(c.F[tc] + v * NF, flo);
    if (ext80)
      u[v] = T(ld_fms(double(u[v]), c.ldh[0], c.ldh[1], double(x), false));
    else
      u[v] = u[v] - h * x;
  }
  // EMF stencil E(+t1, +t2) around the face in the layer of this state
  const int t2 = (d + 2) % 3;
  I3 b = with(fedge, d, layer);
  I3 b10 = with(b, t1, co(b, t1) + 1);
  I3 b01 = with(b, t2, co(b, t2) + 1);
  I3 b11 = with(b10, t2, co(b, t2) + 1);
  const auto E = c.E[d];
  T e00 = ld(E, c.eid(b.j, b.i, b.k)), e10 = ld(E, c.eid(b10.j, b10.i, b10.k));
  T e01 = ld(E, c.eid(b01.j, b01.i, b01.k)), e11 = ld(E, c.eid(b11.j, b11.i, b11.k));
  T q = c.emfq[tc];
  if (tc == t1) {
    T y = e10 - e00 + e11 - e01;
    if (ext80)
      u[5] = T(ld_fms(double(u[5]), 0.5 * c.ldh[0], 0.5 * c.ldh[1], double(y), true));
    else
      u[5] = u[5] + q * y;
  } else {
    u[4] = u[4] - q * (e01 - e00 + e11 - e10);
  }
}

template <class T>
PMHD_FI void ctu_states(const Ctx<T>& c, int d, I3 f, T ul[7], T ur[7]) {
  const int64_t N = c.ncell();
  const int64_t NF = c.nface(d);
  const int64_t idx = c.fid(d, f.j, f.i, f.k);
  const int64_t n = c.n_along(d);
  const int bc = c.bc[d];
  const int64_t e = co(f, d);
  load7_once(c.UL[d], NF, idx, ul);
  load7_once(c.UR[d], NF, idx, ur);
  I3 pl = with(f, d, map1(e - 1, n, bc));
  I3 pr = with(f, d, map1(e, n, bc));
  if (c.ndim == 3) {
    const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
    const int first = d == 1 ? t2 : t1;
    const int second = d == 1 ? t1 : t2;
    int64_t layer_l = e == 0 ? (bc == PERIODIC ? n - 1 : 0) : e - 1;
    int64_t layer_r = e;
    // y-outflow: the bottom ghost of the high x-face flux sits in row nx-1
    // (Source/CTU_HelperModule.py:724); where that row does not exist (nx > ny, the
    // reference raises IndexError) the last row ny-1 is used
    int64_t hov = (d == 1 && bc == OUTFLOW && e == n && !(c.fixq & PMHD_FIX_OUTFLOW_GHOSTS))
                      ? (c.nx < c.ny ? c.nx - 1 : c.ny - 1) : -1;
    for (int s = 0; s < 2; ++s) {
      int tc = s == 0 ? first : second;
      ctu_apply_3d(c, d, tc, pl, -1, f, layer_l, ul);
      ctu_apply_3d(c, d, tc, pr, tc == 0 ? hov : -1, f, layer_r, ur);
    }
    T sl[7], sr[7];
    ctu_source_3d(c, d, bc, f, e - 1, sl);
    ctu_source_3d(c, d, bc, f, e, sr);
#pragma unroll
    for (int v = 1; v < 7; ++v) {
      ul[v] = ul[v] + sl[v];
      ur[v] = ur[v] + sr[v];
    }
  } else {
    // 2D (getXTransverseUpdates_2D / getYTransverseUpdates_2D)
    const int tc = 1 - d;
    const int64_t NT = c.nface(tc);
    const T h = c.hdtd[tc];
    int64_t cl = map1(e - 1, n, bc), cr = map1(e, n, bc);
    int64_t el = e - 1, er = e;  // ext cells for the source term
    auto src2 = [&](int64_t ee, T s[7]) {
      int64_t cm;
      int64_t flo, fhi;
      if (d == 0 || (c.fixq & PMHD_FIX_OUTFLOW_GHOSTS)) {
        cm = map1(ee, n, bc);
        flo = ee < 0 ? (bc == OUTFLOW ? 0 : n - 1) : ee;
        fhi = ee + 1 > n ? (bc == OUTFLOW ? n : 1) : ee + 1;
      } else if (bc == OUTFLOW) {
        // the reference's top ghost is cell row 1 (row 0 when ny == 1, where it raises)
        cm = ee < 0 ? (n > 1 ? 1 : 0) : (ee >= n ? n - 1 : ee);
        flo = ee < 0 ? 1 : ee;
        fhi = ee + 1 > n ? n - 1 : ee + 1;
      } else {
        cm = map1(ee, n, bc);
        flo = ee < 0 ? n - 1 : ee;
        fhi = ee + 1 > n ? 1 : ee + 1;
      }
      I3 pc = with(f, d, cm);
      int64_t ci = c.cid(pc.j, pc.i, pc.k);
      T db = face_val(c, c.fb[d], d, f, fhi) - face_val(c, c.fb[d], d, f, flo);
      T cbx = ld(c.W + 4 * N, ci), cby = ld(c.W + 5 * N, ci), cbz = ld(c.W + 6 * N, ci), cvz = ld(c.W + 3 * N, ci);
      T inv = c.invd[d];
      s[0] = T(0);
      s[1] = cbx * db * inv;
      s[2] = cby * db * inv;
      s[3] = cbz * db * inv;
      s[6] = (cbz * cvz) * db * inv;
      if (d == 0) { s[4] = T(0); s[5] = cvz * db * inv; }
      else { s[4] = cvz * db * inv; s[5] = T(0); }
    };
    T sl[7], sr[7];
    src2(el, sl);
    src2(er, sr);
    for (int side = 0; side < 2; ++side) {
      I3 pc = with(f, d, side == 0 ? cl : cr);
      I3 hi = with(pc, tc, co(pc, tc) + 1);
      int64_t flo = c.fid(tc, pc.j, pc.i, pc.k), fhi = c.fid(tc, hi.j, hi.i, hi.k);
      T g[7];
#pragma unroll
      for (int v = 0; v < 7; ++v) g[v] = ld(c.F[tc] + v * NT, fhi) - ld(c.F[tc] + v * NT, flo);
      // slot remap: the transverse flux's in-plane B component moves to the
      // out-of-plane B_z slot of this direction, its own B slot is zeroed
      T dr[7] = {g[0], g[1], g[2], g[3], T(0), T(0), g[6]};
      if (d == 0) dr[5] = g[4]; else dr[4] = g[5];
      T* u = side == 0 ? ul : ur;
      const T* s = side == 0 ? sl : sr;
#pragma unroll
      for (int v = 0; v < 7; ++v) u[v] = u[v] + ((-h) * dr[v] + c.hdt * s[v]);
      floors_counted(c, 6 + 2 * d + side, u, c.min_energy);
    }
  }
}

// ===========================================================================
// constrained transport of the face field (Source/EvolveFunctions_MHD.py:1003-1028,
// :1082-1104 for 3D; :579-586, :643-649 for 2D)
// ===========================================================================
// fin/fout: face arrays (high parts); lin/lout: their low parts (split faces only)
template <class T, class PI, class PO>
PMHD_FI void k_ct(const Ctx<T>& c, int d, const PI& fin, const PI& lin, const PO& fout, const PO& lout,
                  const T h[3], I3 f) {
  // F: double for split faces (the curl differences of the edge EMFs and the
  // update then round at double precision), else the cell type
  using F = typename std::conditional<split_faces<T>(), double, T>::type;
  const int64_t idx = c.fid(d, f.j, f.i, f.k);
  F old;
  if constexpr (split_faces<T>()) old = F(ld(fin, idx)) + F(ld(lin, idx));
  else old = ld(fin, idx);
  F res;
  if (c.ndim == 2) {
    const auto E = c.E[2];
    if (d == 0) {
      res = old - F(h[1]) * (F(ld(E, c.eid(f.j + 1, f.i, 0))) - F(ld(E, c.eid(f.j, f.i, 0))));
    } else {
      res = old + F(h[0]) * (F(ld(E, c.eid(f.j, f.i + 1, 0))) - F(ld(E, c.eid(f.j, f.i, 0))));
    }
  } else {
    const auto Ex = c.E[0];
    const auto Ey = c.E[1];
    const auto Ez = c.E[2];
    F a, b;
    if (d == 0) {
      a = F(h[2]) * (F(ld(Ey, c.eid(f.j, f.i, f.k + 1))) - F(ld(Ey, c.eid(f.j, f.i, f.k))));
      b = F(h[1]) * (F(ld(Ez, c.eid(f.j + 1, f.i, f.k))) - F(ld(Ez, c.eid(f.j, f.i, f.k))));
    } else if (d == 1) {
      a = F(h[0]) * (F(ld(Ez, c.eid(f.j, f.i + 1, f.k))) - F(ld(Ez, c.eid(f.j, f.i, f.k))));
      b = F(h[2]) * (F(ld(Ex, c.eid(f.j, f.i, f.k + 1))) - F(ld(Ex, c.eid(f.j, f.i, f.k))));
    } else {
      a = F(h[1]) * (F(ld(Ex, c.eid(f.j + 1, f.i, f.k))) - F(ld(Ex, c.eid(f.j, f.i, f.k))));
      b = F(h[0]) * (F(ld(Ey, c.eid(f.j, f.i + 1, f.k))) - F(ld(Ey, c.eid(f.j, f.i, f.k))));
    }
    res = old + (a - b);
  }
  // (an ordinary store: the next launch reads the faces at once)
  if constexpr (split_faces<T>()) {
    const T hi = T(res);
    st(fout, idx, hi);
    st(lout, idx, T(res - F(hi)));
  } else {
    (void)lin;
    (void)lout;
    st(fout, idx, res);
  }
}

// ===========================================================================
// corrector: cons -> prim with the half-step normal field, Roe flux
// ===========================================================================
// CTU-corrected states -> primitive with the half-step normal field -> Roe flux
// the corrector flux of face f along d from its CTU-corrected states and half-step
// normal field (directional order)
template <int S, class T, class U>
PMHD_FI void correct_flux(const Ctx<U>& c, int d, const T ul[7], const T ur[7], T bn, T df[7]) {
  T wl[7], wr[7];
  const int ke = c.ndim == 3 ? d : 0;
  const T g = T(c.gamma), eps = T(c.eps);
  cons_to_prim(ul, bn, g, ke, wl);
  cons_to_prim(ur, bn, g, ke, wr);
  T rwl[7], rwr[7], rul[7], rur[7], rf[7];
  to_rot(wl, d, rwl);
  to_rot(wr, d, rwr);
  to_rot(ul, d, rul);
  to_rot(ur, d, rur);
  roe_flux<S>(rwl, rwr, rul, rur, bn, g, c.efix != 0, eps, rf);
  from_rot(rf, d, df);
}

template <int S, class T>
PMHD_FI void k_correct(const Ctx<T>& c, int d, I3 f) {
  const int64_t NF = c.nface(d);
  const int64_t idx = c.fid(d, f.j, f.i, f.k);
  T ul[7], ur[7], df[7];
  ctu_states(c, d, f, ul, ur);
  T bn = T(ld(c.nfb[d], idx));
  correct_flux<S>(c, d, ul, ur, bn, df);
  store7(c, c.Fc[d], NF, idx, df);
}

// ===========================================================================
// conservative update of the cells and the new primitive state
// ===========================================================================
template <class T>
PMHD_FI void k_update(const Ctx<T>& c, I3 p) {
  const int64_t N = c.ncell();
  const int64_t idx = c.cid(p.j, p.i, p.k);
  // X-directional conserved vector (rho, Mx, My, Mz, By, Bz, E)
  const int dv[7] = {0, 1, 2, 3, 5, 6, 7};
  T u[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) u[v] = ld_once(c.U + dv[v] * N, idx);
  const int64_t NX = c.nface(0);
  int64_t fx0 = c.fid(0, p.j, p.i, p.k), fx1 = c.fid(0, p.j, p.i + 1, p.k);
  T bx, w[7];
  if (c.ndim == 1) {
#pragma unroll
    for (int v = 0; v < 7; ++v) u[v] = u[v] - c.dtd[0] * (ld(c.F[0] + v * NX, fx1) - ld(c.F[0] + v * NX, fx0));
    bx = c.bx1d;
    cons_to_prim(u, bx, c.gamma, 0, w);
  } else if (c.ndim == 2) {
    const int64_t NY = c.nface(1);
    int64_t fy0 = c.fid(1, p.j, p.i, p.k), fy1 = c.fid(1, p.j + 1, p.i, p.k);
    T nbz = u[5] - c.dtd[0] * (ld(c.Fc[0] + 5 * NX, fx1) - ld(c.Fc[0] + 5 * NX, fx0))
            - c.dtd[1] * (ld(c.Fc[1] + 4 * NY, fy1) - ld(c.Fc[1] + 4 * NY, fy0));
#pragma unroll
    for (int v = 0; v < 7; ++v) {
      u[v] = u[v] - c.dtd[0] * (ld(c.Fc[0] + v * NX, fx1) - ld(c.Fc[0] + v * NX, fx0));
      u[v] = u[v] - c.dtd[1] * (ld(c.Fc[1] + v * NY, fy1) - ld(c.Fc[1] + v * NY, fy0));
    }
    bx = (T(ld(c.fb2[0], fx0)) + T(ld(c.fb2[0], fx1))) / T(2);
    T by = (T(ld(c.fb2[1], fy0)) + T(ld(c.fb2[1], fy1))) / T(2);
    u[5] = nbz;
    u[4] = by;
    cons_to_prim(u, bx, c.gamma, 0, w);
  } else {
    const int64_t NY = c.nface(1), NZ = c.nface(2);
    int64_t fy0 = c.fid(1, p.j, p.i, p.k), fy1 = c.fid(1, p.j + 1, p.i, p.k);
    int64_t fz0 = c.fid(2, p.j, p.i, p.k), fz1 = c.fid(2, p.j, p.i, p.k + 1);
#pragma unroll
    for (int v = 0; v < 7; ++v) {
      u[v] = u[v] - c.dtd[0] * (ld(c.Fc[0] + v * NX, fx1) - ld(c.Fc[0] + v * NX, fx0));
      u[v] = u[v] - c.dtd[1] * (ld(c.Fc[1] + v * NY, fy1) - ld(c.Fc[1] + v * NY, fy0));
      u[v] = u[v] - c.dtd[2] * (ld(c.Fc[2] + v * NZ, fz1) - ld(c.Fc[2] + v * NZ, fz0));
    }
    bx = T(0.5) * (T(ld(c.fb2[0], fx0)) + T(ld(c.fb2[0], fx1)));
    T by = T(0.5) * (T(ld(c.fb2[1], fy0)) + T(ld(c.fb2[1], fy1)));
    T bz = T(0.5) * (T(ld(c.fb2[2], fz0)) + T(ld(c.fb2[2], fz1)));
    u[4] = by;
    u[5] = bz;
    cons_to_prim(u, bx, c.gamma, 0, w);
  }
  put(c, c.U2, idx, u[0]);
  put(c, c.U2 + N, idx, u[1]);
  put(c, c.U2 + 2 * N, idx, u[2]);
  put(c, c.U2 + 3 * N, idx, u[3]);
  put(c, c.U2 + 4 * N, idx, bx);
  put(c, c.U2 + 5 * N, idx, u[4]);
  put(c, c.U2 + 6 * N, idx, u[5]);
  put(c, c.U2 + 7 * N, idx, u[6]);
  put(c, c.W2, idx, w[0]);
  put(c, c.W2 + N, idx, w[1]);
  put(c, c.W2 + 2 * N, idx, w[2]);
  put(c, c.W2 + 3 * N, idx, w[3]);
  put(c, c.W2 + 4 * N, idx, bx);
  put(c, c.W2 + 5 * N, idx, w[4]);
  put(c, c.W2 + 6 * N, idx, w[5]);
  put(c, c.W2 + 7 * N, idx, w[6]);
  // the next cycle's cell E from the stored values (what CellE would read back)
  if (c.ec_next && c.ndim >= 2) store_cell_e<T>(c, idx, u[0], u[1], u[2], u[3], w[1], w[2], bx, w[4], w[5]);
}

// ===========================================================================
// the Riemann flux of one face from float64 arrays of directional 7-vectors
// (planes of n faces; riemannSolver{X,Y,Z}_mhd_roe's layout,
// Source/RiemannSolvers_MHD_Roe_HelperModule.py:74, 595, 1093), computed in T
// ===========================================================================
template <class T>
PMHD_FI void riemann_face(int d, int solver, int64_t n, int64_t i, const double* wl, const double* wr,
                          const double* ul, const double* ur, const double* bn, T g, bool efix, T eps,
                          double* F) {
  T a[7], b[7], c[7], e[7], ra[7], rb[7], rc[7], re[7], rf[7], df[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    a[v] = T(wl[v * n + i]);
    b[v] = T(wr[v * n + i]);
    c[v] = T(ul[v * n + i]);
    e[v] = T(ur[v * n + i]);
  }
  to_rot(a, d, ra);
  to_rot(b, d, rb);
  to_rot(c, d, rc);
  to_rot(e, d, re);
  roe_flux_solver(ra, rb, rc, re, T(bn[i]), g, efix, eps, rf, solver);
  from_rot(rf, d, df);
#pragma unroll
  for (int v = 0; v < 7; ++v) F[v * n + i] = double(df[v]);
}

// ===========================================================================
// diagnostics of one cell, in double: the 8 conserved values (summed over the
// grid), |div B| of the face field times the smallest width, |B| of the faces
// (both maximised) and the number of non-finite conserved values (summed)
// ===========================================================================
template <class T>
PMHD_FI void k_diag(const Ctx<T>& c, int64_t idx, double sums[9], double maxes[2]) {
  const int64_t N = c.ncell();
  I3 p = unflat(idx, c.ny, c.nx, c.nz);
  double bad = 0.0;
  const int nv = c.physics == 1 ? 5 : 8;
  for (int v = 0; v < 8; ++v) {
    double u = v < nv ? double(c.U[v * N + idx]) : 0.0;
    sums[v] = u;
    if (!(u == u) || u - u != 0.0) bad += 1.0;
  }
  sums[8] = bad;
  double div = 0.0, bmax = 0.0, wmin = c.wd[0];
  if (c.ndim >= 2 && c.physics == 0) {
    auto face = [&](int a, int64_t i) {
      double v = double(ld(c.fb[a], i));
      if constexpr (split_faces<T>()) v += double(ld(c.fblo[a], i));
      return v;
    };
    double fx0 = face(0, c.fid(0, p.j, p.i, p.k)), fx1 = face(0, c.fid(0, p.j, p.i + 1, p.k));
    double fy0 = face(1, c.fid(1, p.j, p.i, p.k)), fy1 = face(1, c.fid(1, p.j + 1, p.i, p.k));
    div = (fx1 - fx0) / c.wd[0] + (fy1 - fy0) / c.wd[1];
    bmax = npmax(npmax(fabs(fx0), fabs(fx1)), npmax(fabs(fy0), fabs(fy1)));
    wmin = npmin(wmin, c.wd[1]);
    if (c.ndim == 3) {
      double fz0 = face(2, c.fid(2, p.j, p.i, p.k)), fz1 = face(2, c.fid(2, p.j, p.i, p.k + 1));
      div += (fz1 - fz0) / c.wd[2];
      bmax = npmax(bmax, npmax(fabs(fz0), fabs(fz1)));
      wmin = npmin(wmin, c.wd[2]);
    }
  }
  maxes[0] = fabs(div) * wmin;
  maxes[1] = bmax;
}

PMHD_NS_END
