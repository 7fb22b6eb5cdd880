// One CTU+CT cycle as a sequence of named phases.  A phase is a functor type
// with a launch id, the kind of element it runs over (cells, faces, slope cells or
// edge points) and a direction; an executor X runs it over every element:
//   X::run<Ph>(c)              -- Ph()(c, I3{j, i, k}) over the phase's index box
//   X::run_max<Ph>(c)          -- the same for a phase that also yields three speeds
//                                 per element, maximised (kept for the next dt)
//   X::cfl(c, out)             -- out[d] = max over cells of the CFL speed along d
//   X::reduce(id, n, f, sums, maxes)
// The single-grid executors take c as the grid; the batch executors run the same
// phase over every box of a batch, each with its own context.  The direction is a
// template parameter so the rotations and index arithmetic fold at compile time.
// (Source/EvolveFunctions_MHD.py:252-317 for 1D, :464-692 for 2D, :840-1152 for 3D)
#pragma once
#include <type_traits>
#include "pmhd_kernels.h"

PMHD_NS_BEGIN

// ---------------------------------------------------------------------------
// phases
// ---------------------------------------------------------------------------
#define PMHD_PHASE_TRAITS(ID, KIND, DIR) \
  static constexpr int id = (ID);        \
  static constexpr int kind = (KIND);    \
  static constexpr int dir = (DIR);
// phases that form Riemann fluxes carry their solver (RIEMANN_*) as a parameter

template <int D>
struct Slopes {
  PMHD_PHASE_TRAITS(K_SLOPES + D, EXT_SLOPE, D)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { k_slopes(c, D, c.bc[D], p); }
};
template <int D>
struct Edges {
  PMHD_PHASE_TRAITS(K_EDGES + D, EXT_CELL, D)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { k_edges(c, D, c.bc[D], p); }
};
// PLM edge states with each cell's own limited slope, formed from the same
// eigen-system as the tracing (the value the slope launch would store): no slope
// launch and one eigen-system per cell
template <int D>
struct EdgesOwn {
  PMHD_PHASE_TRAITS(K_EDGES + D, EXT_CELL, D)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { k_edges(c, D, c.bc[D], p, false); }
};
template <int D>
struct CopyStates {
  PMHD_PHASE_TRAITS(K_COPY + D, EXT_FACE, D)
  // the reference's 3D first-order z states use the x boundary condition
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const {
    k_copy_states(c, D, D == 2 ? c.bc_zrec : c.bc[D], p);
  }
};
template <int D, int S = RIEMANN_ROE>
struct Predict {
  PMHD_PHASE_TRAITS(K_PREDICT + D, EXT_FACE, D)
  static constexpr int solver = S;
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { k_predict<S>(c, D, p); }
};
template <int D, int S = RIEMANN_ROE>
struct Sweep {  // fused slopes (PLM) + edges + predictor over runs of faces
  PMHD_PHASE_TRAITS(K_SWEEP + D, EXT_SEG, D)
  static constexpr int solver = S;
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 q) const { k_sweep<S>(c, D, c.bc[D], q); }
};
// the fused sweep over whole lines along the contiguous axis D (z in 3D, x in 2D),
// faces 1..n-1, keeping the edge states of both end cells (one workgroup per line
// on the GPU; here the reference order of one thread per line)
template <int D, int S = RIEMANN_ROE>
struct SweepLine {
  static_assert(D == 0 || D == 2, "line sweeps run along the contiguous axis");
  PMHD_PHASE_TRAITS(D == 2 ? K_SWEEP_LINE_Z : K_SWEEP_LINE_X, EXT_LINE, D)
  static constexpr int solver = S;
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 q) const {
    k_sweep_run<S>(c, D, c.bc[D], q, 1, c.n_along(D));
    k_line_ends(c, D, c.bc[D], q);
  }
};
// the two end faces (0 and n) of every line along D from the kept end states
template <int D, int S = RIEMANN_ROE>
struct SweepEnds {
  PMHD_PHASE_TRAITS(D == 2 ? K_SWEEP_ENDS_Z : K_SWEEP_ENDS_X, EXT_ENDS, D)
  static constexpr int solver = S;
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 q) const { k_sweep_end<S>(c, D, c.bc[D], q); }
};
struct CellE {
  PMHD_PHASE_TRAITS(K_CELL_E, EXT_CELL, 0)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { k_cell_e(c, p); }
};
struct EmfPredictor {
  PMHD_PHASE_TRAITS(K_EMF, EXT_EDGE, 0)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { k_emf(c, c.F, p); }
};
struct EmfCorrector {
  PMHD_PHASE_TRAITS(K_EMF2, EXT_EDGE, 0)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { k_emf(c, c.Fc, p); }
};
struct CtHalf {  // half-step face field
  PMHD_PHASE_TRAITS(K_CT_HALF, EXT_EDGE, 0)
  // one launch over the edge points for all directions: element (j, i, k) updates
  // the faces (j, i, k) of every direction that has one there
  template <int D, class T>
  PMHD_FI static void face(const Ctx<T>& c, I3 p) {
    if (D < c.ndim && p.j < c.fny(D) && p.i < c.fnx(D) && p.k < c.fnz(D))
      k_ct(c, D, c.fb[D], c.fblo[D], c.nfb[D], c.nfblo[D], c.hdtd, p);
  }
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const {
    // (compile-time directions: the context's arrays are never indexed at run time)
    face<0>(c, p);
    face<1>(c, p);
    face<2>(c, p);
  }
};
struct HalfE {
  PMHD_PHASE_TRAITS(K_HALF_E, EXT_CELL, 0)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { k_half_e(c, p); }
};
template <int D, int S = RIEMANN_ROE>
struct Correct {  // CTU transverse corrections + corrector flux
  PMHD_PHASE_TRAITS(K_CORRECT + D, EXT_FACE, D)
  static constexpr int solver = S;
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { k_correct<S>(c, D, p); }
};
struct CtFull {  // full-step face field (one launch for all directions, as CtHalf)
  PMHD_PHASE_TRAITS(K_CT_FULL, EXT_EDGE, 0)
  template <int D, class T>
  PMHD_FI static void face(const Ctx<T>& c, I3 p) {
    if (D < c.ndim && p.j < c.fny(D) && p.i < c.fnx(D) && p.k < c.fnz(D))
      k_ct(c, D, c.fb[D], c.fblo[D], c.fb2[D], c.fb2lo[D], c.dtd, p);
  }
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const {
    face<0>(c, p);
    face<1>(c, p);
    face<2>(c, p);
  }
};
template <class T>
PMHD_FI void speeds_out(const T t[3], double o[3]) {
  o[0] = double(t[0]);
  o[1] = double(t[1]);
  o[2] = double(t[2]);
}
// conservative update, then the CFL speeds of the new state for the next
// cycle's timestep (maximised by X::run_max)
struct Update {
  PMHD_PHASE_TRAITS(K_UPDATE, EXT_CELL, 0)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p, double o[3]) const {
    k_update(c, p);
    if constexpr (std::is_floating_point<T>::value) {
      T t[3];
      k_cfl(c, c.U2, c.fb2, p, t);
      speeds_out(t, o);
    } else {
      // the adjoint's dual evaluation: the speeds only pick the next dt, which the
      // adjoint keeps fixed
      o[0] = o[1] = o[2] = 0.0;
    }
  }
};
struct CflSpeeds {  // per-cell speeds of the current state, maximised by X::cfl
  PMHD_PHASE_TRAITS(K_CFL, EXT_CELL, 0)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p, double o[3]) const {
    T t[3];
    k_cfl(c, c.U, c.fb, p, t);
    speeds_out(t, o);
  }
};

// FP64 diagnostics: sums[0..7] conserved totals, sums[8] non-finite count,
// maxes[0] max |div B| * min width, maxes[1] max |B| on faces
template <class T, class X>
void diagnostics(X& ex, const Ctx<T>& c, double sums[9], double maxes[2]) {
  ex.reduce(K_DIAG, c.ncell(), [=] PMHD_HD(int64_t i, double* s, double* m) { k_diag(c, i, s, m); },
            sums, maxes);
}

// interface states and predictor fluxes of direction D: with c.seg[D] > 0 the
// reconstruction and the predictor are one launch (PPM keeps its slope launch)
template <int D, int S, class T, class X>
void predictor_sweep(X& ex, const Ctx<T>& c) {
  // line launches exist along the contiguous axis only (z in 3D, x in 2D)
  constexpr bool can_line = D == 2 || D == 0;
  const bool line = can_line && c.seg[D] < 0 && (D == 2 ? c.ndim == 3 : c.ndim == 2);
  if (c.order > 0 && (c.seg[D] > 0 || line)) {
    if (c.order == 2) ex.template run<Slopes<D>>(c);
    if constexpr (can_line) {
      if (line) {
        ex.template run<SweepLine<D, S>>(c);
        // the end faces overlap the other directions' sweeps (joined before the EMFs)
        ex.template run_side<SweepEnds<D, S>>(c);
        return;
      }
    }
    ex.template run<Sweep<D, S>>(c);
    return;
  }
  // PLM in FP32: one launch per cell for slope and edges; FP64 keeps the two (its
  // combined body spills: 1.9 -> 2.0 ms per direction at 256^3)
  if (c.order == 1 && sizeof(T) == sizeof(float)) {
    ex.template run<EdgesOwn<D>>(c);
  } else if (c.order > 0) {
    ex.template run<Slopes<D>>(c);
    ex.template run<Edges<D>>(c);
  } else {
    ex.template run<CopyStates<D>>(c);
  }
  ex.template run<Predict<D, S>>(c);
}

// c supplies the launch-invariant decisions (ndim, order, solver), which a batch shares
template <int S, class T, class X>
void cycle_solver(X& ex, const Ctx<T>& c) {
  const int nd = c.ndim;
  // the contiguous axis first (z in 3D): its line sweep's end-face launch then
  // overlaps the other sweeps
  if (nd == 3) predictor_sweep<2, S>(ex, c);
  predictor_sweep<0, S>(ex, c);
  if (nd >= 2) predictor_sweep<1, S>(ex, c);
  ex.join_side();
  if (nd == 1) {
    ex.template run_max<Update>(c);
    return;
  }
  if (!c.ec_valid) ex.template run<CellE>(c);   // else the last update stored it
  ex.template run<EmfPredictor>(c);   // corner EMFs of the predictor fluxes
  ex.template run<CtHalf>(c);         // half-step face field
  ex.template run<HalfE>(c);
  ex.template run<Correct<0, S>>(c);  // CTU transverse corrections + corrector fluxes
  ex.template run<Correct<1, S>>(c);
  if (nd == 3) ex.template run<Correct<2, S>>(c);
  ex.template run<EmfCorrector>(c);   // corner EMFs of the corrector fluxes
  ex.template run<CtFull>(c);         // full-step face field
  ex.template run_max<Update>(c);
}

// one cycle with the grid's Riemann solver (a unit that defines PMHD_ROE_ONLY, the
// adjoint's, instantiates the Roe cycle only)
template <class T, class X>
void cycle(X& ex, const Ctx<T>& c) {
#ifndef PMHD_ROE_ONLY
  if (c.riemann == RIEMANN_HLLE) {
    cycle_solver<RIEMANN_HLLE>(ex, c);
    return;
  }
  if (c.riemann == RIEMANN_ROE_HLLE) {
    cycle_solver<RIEMANN_ROE_HLLE>(ex, c);
    return;
  }
#endif
  cycle_solver<RIEMANN_ROE>(ex, c);
}

// ---------------------------------------------------------------------------
// timestep (Source/TimestepCalculations_HelperModule.py:222-225, 285-289, 354-361)
// and the per-cycle scalars of a context
// ---------------------------------------------------------------------------
PMHD_FI double dt_from_maxima(int ndim, double cfl, const double m[3], const double w[3]) {
  if (ndim == 1) return cfl / (m[0] / w[0]);
  if (ndim == 2) return cfl / npmax(m[0] / w[0], m[1] / w[1]);
  return cfl / npmax(m[0] / w[0], npmax(m[1] / w[1], m[2] / w[2]));
}

namespace hydro {
// the minimum of the per-direction dt (Source/TimestepCalculations_HelperModule.py:35-170)
PMHD_FI double dt_from_maxima(int ndim, double cfl, const double m[3], const double w[3]) {
  const double dx = cfl / (m[0] / w[0]);
  if (ndim == 1) return dx;
  const double dy = cfl / (m[1] / w[1]);
  if (ndim == 2) return npmin(dx, dy);
  const double dz = cfl / (m[2] / w[2]);
  return npmin(dx, npmin(dy, dz));
}
}  // namespace hydro

template <class T>
PMHD_FI double dt_of(const Ctx<T>& c, double cfl, const double m[3]) {
  return c.physics == 1 ? hydro::dt_from_maxima(c.ndim, cfl, m, c.wd) : dt_from_maxima(c.ndim, cfl, m, c.wd);
}

// the reference's end-of-interval clipping (Source/EvolveFunctions_MHD.py:259-260,
// 472-473 for 1D/2D MHD; :851-852 for 3D MHD; Source/EvolveFunctions_Hydro.py:268-269,
// 398-399, 616-617 for hydro)
PMHD_FI double clip_dt(int ndim, int physics, double t, double t_end, double dt) {
  if (ndim == 3 && physics == 0) {
    if (t_end - t < dt) dt = t_end - t;
  } else {
    if (t + dt > t_end) dt = t_end - t;
  }
  return dt;
}

// 0.5*dt/dy as the reference's 80-bit value (Source/CTU_HelperModule.py:455),
// from a double-double quotient rounded to a 64-bit significand; equal to the
// host's long double result unless dt/dy lies within 2^-106 (relative) of a
// 64-bit rounding midpoint
PMHD_FI void half_ratio_80(double dt, double dy, double out[2]) {
  const double qh = dt / dy;
  const double r = fma(-qh, dy, dt);
  const double ql = r / dy;
  const double lo64 = round64_lo(qh, ql);
  double s, e;
  two_sum(qh, lo64, s, e);
  out[0] = 0.5 * s;
  out[1] = 0.5 * e;
}

template <class T>
PMHD_FI void set_step_scalars(Ctx<T>& c, double dt) {
  c.dt = T(dt);
  c.hdt = T(0.5 * dt);
  for (int d = 0; d < 3; ++d) {
    const double dtd = dt / c.wd[d];
    const double h = 0.5 * dtd;
    c.dtd[d] = T(dtd);
    c.hdtd[d] = T(h);
    c.emfq[d] = T(h * 0.5);
  }
}

// the double-buffered state of a context after a completed cycle
template <class T>
PMHD_FI void swap_buffers(Ctx<T>& c) {
  T* u = c.U2;
  T* w = c.W2;
  c.U2 = const_cast<T*>(c.U);
  c.W2 = const_cast<T*>(c.W);
  c.U = u;
  c.W = w;
  for (int d = 0; d < 3; ++d) {
    T* f = c.fb2[d];
    c.fb2[d] = const_cast<T*>(c.fb[d]);
    c.fb[d] = f;
    T* l = c.fb2lo[d];
    c.fb2lo[d] = const_cast<T*>(c.fblo[d]);
    c.fblo[d] = l;
  }
}

// the element of thread `tid` in tile vb of a batch launch: box b owns tiles
// [prefix[b], prefix[b+1]); a tile is bk (along z) x bi (along x) elements of one
// y row, tiles ordered z-fastest, then x, then y
struct TileRef {
  int box, j, i, k;
  bool in;
};

template <class T>
PMHD_FI TileRef locate_tile(const Ctx<T>* ctxs, const int* prefix, int nbox, int vb, int kind, int dir, int bk,
                            int bi, int tid) {
  int lo = 0, hi = nbox;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= vb) lo = mid;
    else hi = mid;
  }
  TileRef r;
  r.box = lo;
  int64_t n[3];
  ctxs[lo].extent(kind, dir, n);
  const int gk = (int(n[2]) + bk - 1) / bk;
  const int gi = (int(n[1]) + bi - 1) / bi;
  const int local = vb - prefix[lo];
  const int tk = local % gk;
  const int rest = local / gk;
  const int ti = rest % gi;
  r.j = rest / gi;
  r.k = tk * bk + tid % bk;
  r.i = ti * bi + tid / bk;
  r.in = r.k < n[2] && r.i < n[1] && r.j < n[0];
  return r;
}

// clock of one batch box; the second-last state is the other buffer once done > 0
struct BoxClock {
  double t, t2, dt;
  int64_t cyc;        // cycle number
  int64_t done;       // cycles completed since the box's state was set
  int64_t nonfinite;  // cycles whose CFL timestep was not finite (before clipping)
};

// start of a cycle of one batch box, from its CFL maxima m: decide whether the box
// still runs (the reference's loop condition), clip dt to the interval end and form
// the cycle's scalars.  commit = 0 only forms the scalars (used by tile tuning).
template <class T>
PMHD_FI void box_begin(Ctx<T>& c, BoxClock& k, const double m[3], double cfl, double t_end, int64_t max_cycles,
                       int commit) {
  if (commit && !(k.t < t_end && k.cyc < max_cycles)) {
    c.active = 0;
    return;
  }
  double dt = dt_of(c, cfl, m);
  if (commit) {
    // a non-finite state gives a NaN timestep; the reference still runs that cycle
    // (its clock becomes NaN and its loop ends)
    if (!(dt - dt == 0.0)) k.nonfinite += 1;
    dt = clip_dt(c.ndim, c.physics, k.t, t_end, dt);
    k.t2 = k.t;
    k.dt = dt;
  }
  c.active = 1;
  set_step_scalars(c, dt);
  if (sizeof(T) == sizeof(double) && c.ndim == 3 && c.physics == 0) half_ratio_80(dt, c.wd[1], c.ldh);
}

template <class T>
PMHD_FI void box_end(Ctx<T>& c, BoxClock& k) {
  if (!c.active) return;
  swap_buffers(c);
  k.t += k.dt;
  k.cyc += 1;
  k.done += 1;
}

PMHD_NS_END
