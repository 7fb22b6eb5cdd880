// Adjoint of a run of fixed-dt steps (no counterpart in the reference), with
// checkpoint-and-recompute.
//
// forward(state0, dt, nsteps): runs the steps, keeping the state at every
//   checkpoint_every-th step.
// backward(cotangent of the final state): for each checkpoint segment, last
//   first, runs the segment's steps forward from its checkpoint, recording each
//   step's intermediates into value buffers of its own (or, beyond a memory budget,
//   replays the states and recomputes each step's intermediates when its turn
//   comes); then for each step of the segment, last first, runs the adjoint of
//   every phase in reverse order.  A phase's adjoint evaluates the phase's own kernel body on dual
//   numbers (pmhd_dual.h): per output element, a recording pass (plain values)
//   finds the input slots (the HIP backend keeps these lists after a launch's first
//   evaluation), one seeded pass per slot forms that slot's share of the
//   vector-Jacobian product, and the element's output cotangents are then cleared
//   (the values they belong to are dead before the phase).  The corrector faces are
//   differentiated in two stages (adjoint_correct): through the Riemann flux per
//   intermediate value, then through the hand-written reverse of the CTU correction;
//   the update cells through a hand-written reverse (adjoint_update).  The shares
//   reach the cotangent arrays through the backend's accumulation (a fixed-order
//   sum per target on the GPU, pmhd_hip_adjdet.h, or FP64 atomics).  Returns the
//   cotangent of the initial state (prim, cons, faces).
//
// The loss's dependence on dt is not differentiated: dt is a fixed input.
#pragma once
#include <cstdlib>

#include "pmhd_api.h"
#include "pmhd_dual.h"

struct pmhd_adjoint {
  virtual ~pmhd_adjoint() {}
  std::string err;
  virtual int init(const pmhd_grid& g, const pmhd_params& p, int checkpoint_every) = 0;
  // dev: the arrays are device float64 arrays (else host float64 arrays)
  virtual int forward(const double* prim, const double* cons, const double* const fin[3], double dt,
                      int64_t nsteps, double* prim_out, double* cons_out, double* const fout[3], bool dev) = 0;
  virtual int backward(const double* gprim, const double* gcons, const double* const gf[3], double* gprim0,
                       double* gcons0, double* const gf0[3], bool dev) = 0;
  virtual int set_stream(void* stream) = 0;
  virtual int set_deterministic(int on) = 0;
  virtual int set_profiling(int on) = 0;
  virtual int get_profile(int maxn, char* names, double* ms, int64_t* launches, int64_t* elems, int32_t* n) = 0;
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

// constructors of the adjoint handles (defined in the adjoint's translation unit)
pmhd_adjoint* pmhd_new_adjoint_f64();
pmhd_adjoint* pmhd_new_adjoint_f32();

PMHD_NS_BEGIN

// the element function of a phase (the update also yields CFL speeds, unused here)
template <class Ph, class T>
PMHD_FI void call_phase(const Ctx<T>& c, I3 p) {
  if constexpr (Ph::id == K_UPDATE) {
    double o[3];
    Ph()(c, p, o);
  } else {
    Ph()(c, p);
  }
}



// the context of a phase on number type D: values from c, cotangents from a, each
// array wrapped by mk(values, cotangents, id) with a fixed id per array member
// (0 .. kAdIds-1)
template <class D, class V, class Mk>
PMHD_FI void make_ctx_of(const Ctx<V>& c, const Ctx<double>& a, Mk mk, Ctx<D>& d) {
  d.ndim = c.ndim;
  d.active = c.active;
  d.physics = c.physics;
  d.ny = c.ny;
  d.nx = c.nx;
  d.nz = c.nz;
  d.bc_zrec = c.bc_zrec;
  d.order = c.order;
  d.efix = c.efix;
  d.ec_next = c.ec_next;
  d.ec_valid = c.ec_valid;
  d.riemann = c.riemann;
  d.fixq = c.fixq;
  d.gamma = D(double(c.gamma));
  d.eps = D(double(c.eps));
  d.min_dens = D(double(c.min_dens));
  d.min_pres = D(double(c.min_pres));
  d.min_energy = D(double(c.min_energy));
  d.dt = D(double(c.dt));
  d.hdt = D(double(c.hdt));
  d.bx1d = D(double(c.bx1d));
  for (int q = 0; q < 3; ++q) {
    d.bc[q] = c.bc[q];
    d.seg[q] = c.seg[q];
    d.dtd[q] = D(double(c.dtd[q]));
    d.hdtd[q] = D(double(c.hdtd[q]));
    d.invd[q] = D(double(c.invd[q]));
    d.width[q] = D(double(c.width[q]));
    d.wd[q] = c.wd[q];
    d.emfq[q] = D(double(c.emfq[q]));
    d.fb[q] = mk(c.fb[q], const_cast<double*>(a.fb[q]), 8 * q);
    d.fb2[q] = mk(c.fb2[q], a.fb2[q], 8 * q + 1);
    d.nfb[q] = mk(c.nfb[q], a.nfb[q], 8 * q + 2);
    d.UL[q] = mk(c.UL[q], a.UL[q], 8 * q + 3);
    d.UR[q] = mk(c.UR[q], a.UR[q], 8 * q + 4);
    d.F[q] = mk(c.F[q], a.F[q], 8 * q + 5);
    d.Fc[q] = mk(c.Fc[q], a.Fc[q], 8 * q + 6);
    d.E[q] = mk(c.E[q], a.E[q], 8 * q + 7);
  }
  d.ldh[0] = c.ldh[0];
  d.ldh[1] = c.ldh[1];
  d.U = mk(c.U, const_cast<double*>(a.U), 24);
  d.W = mk(c.W, const_cast<double*>(a.W), 25);
  d.U2 = mk(c.U2, a.U2, 26);
  d.W2 = mk(c.W2, a.W2, 27);
  d.WL = mk(c.WL, a.WL, 28);
  d.WR = mk(c.WR, a.WR, 29);
  d.S = mk(c.S, a.S, 30);
  d.Ec = mk(c.Ec, a.Ec, 31);
  static_assert(kAdIds == 32, "array ids 0..31");
  d.fcnt = nullptr;   // the dual evaluations floor without counting
}

// ---------------------------------------------------------------------------
// Reverse of the CTU correction of a face (ctu_states, pmhd_kernels.h): given the
// cotangents gu of the corrected left and right states, add each input's share to
// its cotangent array.  The corrected states are linear in the interface states,
// the transverse flux differences and the EMFs, and bilinear in the source terms'
// fields; the floors (2D) and minmod selections (3D) take the branches the forward
// values take.  The derivatives are those of the dual-number evaluation (the FP64
// step's 80-bit factor enters as its double value).
// ---------------------------------------------------------------------------
template <class V, class Add>
PMHD_FI void ctu_apply_vjp(const Ctx<V>& c, const Ctx<double>& a, int d, int tc, I3 pcell, int64_t hi_override,
                           I3 fedge, int64_t layer, const double gu[7], Add put) {
  const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
  const int64_t NF = c.nface(tc);
  const I3 lo = pcell;
  I3 hi = with(pcell, tc, co(pcell, tc) + 1);
  if (hi_override >= 0) hi = with(hi, 1, hi_override);
  const int64_t flo = c.fid(tc, lo.j, lo.i, lo.k), fhi = c.fid(tc, hi.j, hi.i, hi.k);
  const double h = double(c.hdtd[tc]);
  const int vs[5] = {0, 1, 2, 3, 6};
  for (int q = 0; q < 5; ++q) {
    const int v = vs[q];
    put(a.F[tc], v * NF + fhi, -h * gu[v]);
    put(a.F[tc], v * NF + flo, h * gu[v]);
  }
  const I3 b = with(fedge, d, layer);
  const I3 b10 = with(b, t1, co(b, t1) + 1);
  const I3 b01 = with(b, t2, co(b, t2) + 1);
  const I3 b11 = with(b10, t2, co(b, t2) + 1);
  const double q = double(c.emfq[tc]);
  const double* E = a.E[d];
  if (tc == t1) {
    const double gy = q * gu[5];   // u5 += q (e10 - e00 + e11 - e01)
    put(E, c.eid(b10.j, b10.i, b10.k), gy);
    put(E, c.eid(b.j, b.i, b.k), -gy);
    put(E, c.eid(b11.j, b11.i, b11.k), gy);
    put(E, c.eid(b01.j, b01.i, b01.k), -gy);
  } else {
    const double gz = -q * gu[4];  // u4 -= q (e01 - e00 + e11 - e10)
    put(E, c.eid(b01.j, b01.i, b01.k), gz);
    put(E, c.eid(b.j, b.i, b.k), -gz);
    put(E, c.eid(b11.j, b11.i, b11.k), gz);
    put(E, c.eid(b10.j, b10.i, b10.k), -gz);
  }
}

// minmod_ctu(t, n)'s choice: 1 the first argument, -1 minus the second, 0 zero
template <class V>
PMHD_FI int minmod_ctu_pick(V t, V n) {
  if (n > V(0) && t < V(0)) return (t > -n) ? 1 : -1;
  if (n < V(0) && t > V(0)) return (t < -n) ? 1 : -1;
  return 0;
}

template <class V, class Add>
PMHD_FI void ctu_source_vjp(const Ctx<V>& c, const Ctx<double>& a, int d, int bc, I3 f, int64_t e,
                            const double gu[7], Add put) {
  const int64_t N = c.ncell();
  const int64_t n = c.n_along(d);
  const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
  const I3 pm = with(f, d, map1(e, n, bc));
  int64_t flo, fhi;
  if (e < 0) flo = bc == OUTFLOW ? 0 : n - 1; else flo = e;
  const bool zq = d == 2 && !(c.fixq & PMHD_FIX_OUTFLOW_GHOSTS);
  if (e + 1 > n) fhi = bc == OUTFLOW ? (zq ? n - 1 : n) : 1; else fhi = e + 1;
  const I3 fh = with(f, d, fhi), fl = with(f, d, flo);
  const I3 p1h = with(pm, t1, co(pm, t1) + 1), p2h = with(pm, t2, co(pm, t2) + 1);
  // the forward values, in the forward's arithmetic (for the minmod choices)
  const V dd = face_val(c, c.fb[d], d, fh, fhi) - face_val(c, c.fb[d], d, fl, flo);
  const V d1 = face_val(c, c.fb[t1], t1, p1h, co(p1h, t1)) - face_val(c, c.fb[t1], t1, pm, co(pm, t1));
  const V d2 = face_val(c, c.fb[t2], t2, p2h, co(p2h, t2)) - face_val(c, c.fb[t2], t2, pm, co(pm, t2));
  V dbn, db1, db2;
  double sn, s1, s2;   // d(db*)/d(d*)
  if (d == 0) {
    dbn = dd * c.invd[d]; db1 = d1 * c.invd[t1]; db2 = d2 * c.invd[t2];
    sn = double(c.invd[d]); s1 = double(c.invd[t1]); s2 = double(c.invd[t2]);
  } else {
    dbn = dd / c.width[d]; db1 = d1 / c.width[t1]; db2 = d2 / c.width[t2];
    sn = 1.0 / double(c.width[d]); s1 = 1.0 / double(c.width[t1]); s2 = 1.0 / double(c.width[t2]);
  }
  const int pick1 = minmod_ctu_pick(db1, dbn), pick2 = minmod_ctu_pick(db2, dbn);
  const double mm1 = pick1 == 1 ? double(db1) : (pick1 == -1 ? -double(dbn) : 0.0);
  const double mm2 = pick2 == 1 ? double(db2) : (pick2 == -1 ? -double(dbn) : 0.0);
  const int64_t cm = c.cid(pm.j, pm.i, pm.k);
  double cb[3], cv[3];
  for (int k = 0; k < 3; ++k) {
    cb[k] = double(ld(c.W + (4 + k) * N, cm));
    cv[k] = double(ld(c.W + (1 + k) * N, cm));
  }
  const double v1 = cv[t1], v2 = cv[t2], b1 = cb[t1], b2 = cb[t2];
  const double hdt = double(c.hdt);
  // src1..3 = hdt cb dbn, src4 = -hdt v1 mm2, src5 = -hdt v2 mm1,
  // src6 = hdt (-b1 v1 mm2 - b2 v2 mm1); each is added to u
  double gcb[3] = {0, 0, 0}, gcv[3] = {0, 0, 0};
  double gdbn = 0.0, gmm1 = 0.0, gmm2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    gcb[k] += hdt * double(dbn) * gu[1 + k];
    gdbn += hdt * cb[k] * gu[1 + k];
  }
  gcv[t1] += -hdt * mm2 * gu[4];
  gmm2 += -hdt * v1 * gu[4];
  gcv[t2] += -hdt * mm1 * gu[5];
  gmm1 += -hdt * v2 * gu[5];
  gcb[t1] += -hdt * v1 * mm2 * gu[6];
  gcv[t1] += -hdt * b1 * mm2 * gu[6];
  gmm2 += -hdt * b1 * v1 * gu[6];
  gcb[t2] += -hdt * v2 * mm1 * gu[6];
  gcv[t2] += -hdt * b2 * mm1 * gu[6];
  gmm1 += -hdt * b2 * v2 * gu[6];
  double gdb1 = 0.0, gdb2 = 0.0;
  if (pick1 == 1) gdb1 += gmm1; else if (pick1 == -1) gdbn -= gmm1;
  if (pick2 == 1) gdb2 += gmm2; else if (pick2 == -1) gdbn -= gmm2;
  const double gdd = gdbn * sn, gd1 = gdb1 * s1, gd2 = gdb2 * s2;
  put(a.fb[d], c.fid(d, fh.j, fh.i, fh.k), gdd);
  put(a.fb[d], c.fid(d, fl.j, fl.i, fl.k), -gdd);
  put(a.fb[t1], c.fid(t1, p1h.j, p1h.i, p1h.k), gd1);
  put(a.fb[t1], c.fid(t1, pm.j, pm.i, pm.k), -gd1);
  put(a.fb[t2], c.fid(t2, p2h.j, p2h.i, p2h.k), gd2);
  put(a.fb[t2], c.fid(t2, pm.j, pm.i, pm.k), -gd2);
  for (int k = 0; k < 3; ++k) {
    put(a.W, (4 + k) * N + cm, gcb[k]);
    put(a.W, (1 + k) * N + cm, gcv[k]);
  }
}

// g[0..6], g[7..13]: cotangents of the corrected left and right states of face f
template <class V, class Add>
PMHD_FI void ctu_vjp(const Ctx<V>& c, const Ctx<double>& a, int d, I3 f, const double g[14], Add add) {
  auto put = [&](const double* arr, int64_t i, double v) { add(const_cast<double*>(arr) + i, v); };
  const int64_t N = c.ncell();
  const int64_t NF = c.nface(d);
  const int64_t idx = c.fid(d, f.j, f.i, f.k);
  const int64_t n = c.n_along(d);
  const int bc = c.bc[d];
  const int64_t e = co(f, d);
  if (c.ndim == 3) {
    for (int v = 0; v < 7; ++v) {
      put(a.UL[d], v * NF + idx, g[v]);
      put(a.UR[d], v * NF + idx, g[7 + v]);
    }
    const I3 pl = with(f, d, map1(e - 1, n, bc));
    const I3 pr = with(f, d, map1(e, n, bc));
    const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
    const int first = d == 1 ? t2 : t1;
    const int second = d == 1 ? t1 : t2;
    const int64_t layer_l = e == 0 ? (bc == PERIODIC ? n - 1 : 0) : e - 1;
    const int64_t layer_r = e;
    const int64_t hov = (d == 1 && bc == OUTFLOW && e == n && !(c.fixq & PMHD_FIX_OUTFLOW_GHOSTS))
                            ? (c.nx < c.ny ? c.nx - 1 : c.ny - 1) : -1;
    for (int s = 0; s < 2; ++s) {
      const int tc = s == 0 ? first : second;
      ctu_apply_vjp(c, a, d, tc, pl, -1, f, layer_l, g, put);
      ctu_apply_vjp(c, a, d, tc, pr, tc == 0 ? hov : -1, f, layer_r, g + 7, put);
    }
    ctu_source_vjp(c, a, d, bc, f, e - 1, g, put);
    ctu_source_vjp(c, a, d, bc, f, e, g + 7, put);
    return;
  }
  // 2D (getXTransverseUpdates_2D / getYTransverseUpdates_2D)
  const int tc = 1 - d;
  const int64_t NT = c.nface(tc);
  const V h = c.hdtd[tc];
  const int64_t cl = map1(e - 1, n, bc), cr = map1(e, n, bc);
  for (int side = 0; side < 2; ++side) {
    const int64_t ee = side == 0 ? e - 1 : e;
    // the source term's cell and faces (as ctu_states)
    int64_t cm, flo, fhi;
    if (d == 0 || (c.fixq & PMHD_FIX_OUTFLOW_GHOSTS)) {
      cm = map1(ee, n, bc);
      flo = ee < 0 ? (bc == OUTFLOW ? 0 : n - 1) : ee;
      fhi = ee + 1 > n ? (bc == OUTFLOW ? n : 1) : ee + 1;
    } else if (bc == OUTFLOW) {
      cm = ee < 0 ? (n > 1 ? 1 : 0) : (ee >= n ? n - 1 : ee);
      flo = ee < 0 ? 1 : ee;
      fhi = ee + 1 > n ? n - 1 : ee + 1;
    } else {
      cm = map1(ee, n, bc);
      flo = ee < 0 ? n - 1 : ee;
      fhi = ee + 1 > n ? 1 : ee + 1;
    }
    const I3 ps = with(f, d, cm);
    const int64_t ci = c.cid(ps.j, ps.i, ps.k);
    const V fvh = face_val(c, c.fb[d], d, f, fhi), fvl = face_val(c, c.fb[d], d, f, flo);
    const V db = fvh - fvl;
    const V cbx = ld(c.W + 4 * N, ci), cby = ld(c.W + 5 * N, ci), cbz = ld(c.W + 6 * N, ci);
    const V cvz = ld(c.W + 3 * N, ci);
    const V inv = c.invd[d];
    V s[7];
    s[0] = V(0);
    s[1] = cbx * db * inv;
    s[2] = cby * db * inv;
    s[3] = cbz * db * inv;
    s[6] = (cbz * cvz) * db * inv;
    if (d == 0) { s[4] = V(0); s[5] = cvz * db * inv; }
    else { s[4] = cvz * db * inv; s[5] = V(0); }
    // the transverse flux difference of the side's cell, its slots remapped
    const I3 pc = with(f, d, side == 0 ? cl : cr);
    const I3 hi = with(pc, tc, co(pc, tc) + 1);
    const int64_t tlo = c.fid(tc, pc.j, pc.i, pc.k), thi = c.fid(tc, hi.j, hi.i, hi.k);
    V gflux[7];
    for (int v = 0; v < 7; ++v) gflux[v] = ld(c.F[tc] + v * NT, thi) - ld(c.F[tc] + v * NT, tlo);
    V dr[7] = {gflux[0], gflux[1], gflux[2], gflux[3], V(0), V(0), gflux[6]};
    if (d == 0) dr[5] = gflux[4]; else dr[4] = gflux[5];
    // the corrected state before its floors (which floors act)
    const auto& U = side == 0 ? c.UL[d] : c.UR[d];
    V u0 = ld(U, idx), u6 = ld(U + 6 * NF, idx);
    u0 = u0 + ((-h) * dr[0] + c.hdt * s[0]);
    u6 = u6 + ((-h) * dr[6] + c.hdt * s[6]);
    double gp[7];
    for (int v = 0; v < 7; ++v) gp[v] = g[7 * side + v];
    if (u0 < c.min_dens) gp[0] = 0.0;
    if (u6 < c.min_energy) gp[6] = 0.0;
    const double* gU = side == 0 ? a.UL[d] : a.UR[d];
    for (int v = 0; v < 7; ++v) put(gU, v * NF + idx, gp[v]);
    // u += -h dr
    const double hd = double(h);
    const int fvar[7] = {0, 1, 2, 3, d == 0 ? -1 : 5, d == 0 ? 4 : -1, 6};
    for (int v = 0; v < 7; ++v) {
      if (fvar[v] < 0) continue;
      put(a.F[tc], fvar[v] * NT + thi, -hd * gp[v]);
      put(a.F[tc], fvar[v] * NT + tlo, hd * gp[v]);
    }
    // u += hdt s
    const double hdt = double(c.hdt), id = double(inv), dbd = double(db);
    const double gs1 = hdt * gp[1], gs2 = hdt * gp[2], gs3 = hdt * gp[3], gs6 = hdt * gp[6];
    const double gsv = hdt * (d == 0 ? gp[5] : gp[4]);
    const double bx = double(cbx), by = double(cby), bz = double(cbz), vz = double(cvz);
    const double gbx = dbd * id * gs1, gby = dbd * id * gs2;
    const double gbz = dbd * id * gs3 + vz * dbd * id * gs6;
    const double gvz = bz * dbd * id * gs6 + dbd * id * gsv;
    const double gdb = (bx * gs1 + by * gs2 + bz * gs3 + bz * vz * gs6 + vz * gsv) * id;
    put(a.W, 4 * N + ci, gbx);
    put(a.W, 5 * N + ci, gby);
    put(a.W, 6 * N + ci, gbz);
    put(a.W, 3 * N + ci, gvz);
    put(a.fb[d], c.fid(d, with(f, d, fhi).j, with(f, d, fhi).i, with(f, d, fhi).k), gdb);
    put(a.fb[d], c.fid(d, with(f, d, flo).j, with(f, d, flo).i, with(f, d, flo).k), -gdb);
  }
}

template <class Ph>
struct is_correct : std::false_type {};
template <int D, int S>
struct is_correct<Correct<D, S>> : std::true_type {};

// Adjoint of one corrector face in two stages.  The face's flux depends on its
// input slots only through the CTU-corrected states u_L, u_R and the half-step
// normal field b_n (15 values x).  First one forward pass per x_q through the flux
// (cons->prim and the Roe solver) gives g_q = sum_v cotangent(F_v) dF_v/dx_q; then
// the hand-written reverse of the CTU correction (ctu_vjp) distributes g to the
// inputs.  15 passes through the Roe solver and one reverse sweep, instead of one
// pass through the whole body per input slot (about 55).
// One face may be split over `lanes` threads (lanes of one wave): lane L forms g_q
// for q = L, L + lanes, ..., share(g) sums the lanes' g (zero where a lane formed
// none, so the sums are exact), and lane 0 runs the reverse of the CTU correction.
struct NoShare {
  PMHD_FI void operator()(double*) const {}
};
template <class Ph, class V, class Add, class Share = NoShare>
PMHD_FI bool adjoint_correct(const Ctx<V>& c, const Ctx<double>& a, I3 f, Add add, int lane = 0, int lanes = 1,
                             Share share = Share()) {
  constexpr int Dir = Ph::dir;
  using D1 = ADual<V>;
  AdHot hot{-1, -1, 0.0};
  Ctx<D1> d;
  make_ctx_of<D1>(c, a, [&](const V* val, double* adj, int id) { return AdPtr<V>{val, adj, &hot, 0, id}; }, d);
  const int64_t NF = c.nface(Dir);
  const int64_t idx = c.fid(Dir, f.j, f.i, f.k);
  // the corrected states and normal field (the dual evaluation with no seed)
  D1 x[15];
  ctu_states(d, Dir, f, x, x + 7);
  x[14] = D1(ld(d.nfb[Dir], idx));
  double gF[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) gF[v] = a.Fc[Dir][v * NF + idx];
  // stage 1: the flux's sensitivities to the 15 intermediate values
  double g[15];
#pragma unroll
  for (int q = 0; q < 15; ++q) g[q] = 0.0;
  for (int q = lane; q < 15; q += lanes) {
    D1 y[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) y[k] = D1(x[k].v, k == q ? V(1) : V(0), 0);
    D1 F[7];
    correct_flux<Ph::solver>(d, Dir, y, y + 7, y[14], F);
    double acc = 0.0;
#pragma unroll
    for (int v = 0; v < 7; ++v) acc += gF[v] * double(F[v].t);
    g[q] = acc;
  }
  share(g);
  if (lane != 0) return true;
  // stage 2: the reverse of the CTU correction, and the normal field's share
  Ctx<V> cv = c;
  cv.fcnt = nullptr;
  ctu_vjp(cv, a, Dir, f, g, add);
  add(a.nfb[Dir] + idx, g[14]);
#pragma unroll
  for (int v = 0; v < 7; ++v) a.Fc[Dir][v * NF + idx] = 0.0;
  return true;
}

// the one-launch PLM slope and edge body: its seeded passes need every VGPR a wave
// has, and the two-lane form spills, so it keeps one thread per cell
template <class Ph>
struct is_edges_own : std::false_type {};
template <int D>
struct is_edges_own<EdgesOwn<D>> : std::true_type {};

// ---------------------------------------------------------------------------
// Predictor faces in two stages (as the corrector): the face's outputs (U_L, U_R
// and the flux) depend on its input slots only through the prepared primitive
// states w_L, w_R (floors and reconstruction source terms applied) and the normal
// field (15 values).  15 dual passes through prim->cons and the Roe flux give the
// outputs' sensitivities g; the hand-written reverse of the preparation
// (predict_prep_vjp) distributes them to the inputs.
// ---------------------------------------------------------------------------
template <class Ph>
struct is_predict : std::false_type {};
template <int D, int S>
struct is_predict<Predict<D, S>> : std::true_type {};

// minmod_recon(dn, dt, zero_fix)'s choice: 1 dn, -1 minus dt, 0 zero
template <class V>
PMHD_FI int minmod_recon_pick(V dn, V dt_, bool zero_fix) {
  bool took_t;
  V t;
  if (dn >= V(0)) {
    took_t = !(dn < -dt_);
    t = took_t ? -dt_ : dn;
  } else {
    took_t = !(dn > -dt_);
    t = took_t ? -dt_ : dn;
  }
  if (zero_fix && took_t && dt_ == V(0)) return 0;
  if (dn >= V(0) && t < V(0)) return 0;
  if (!(dn >= V(0)) && t > V(0)) return 0;
  return took_t ? -1 : 1;
}

// reverse of the source term of the reconstruction on cell m of the line of faces
// through f along d (cell_recon_src): go[0], go[1] the cotangents of its terms
template <class V, class Put>
PMHD_FI void recon_src_vjp(const Ctx<V>& c, const Ctx<double>& a, int d, I3 f, int64_t m, const double go[2],
                           Put put) {
  const int64_t N = c.ncell();
  const double hdt = double(c.hdt);
  if (c.ndim == 2) {
    const int64_t n = c.n_along(d);
    const int bc = c.bc[d];
    const I3 pm = with(f, d, map1(m, n, bc));
    const int vslot = d == 0 ? 2 : 1;
    const int64_t cm = c.cid(pm.j, pm.i, pm.k);
    const V v = ld(c.W + vslot * N, cm);
    const int64_t flo = m < 0 ? (bc == OUTFLOW ? 0 : n - 1) : m;
    const int64_t fhi = m + 1 > n ? (bc == OUTFLOW ? n : 1) : m + 1;
    const double dd = double((face_val(c, c.fb[d], d, f, fhi) - face_val(c, c.fb[d], d, f, flo)) * c.invd[d]);
    // out0 = hdt v ((fb[fhi] - fb[flo]) invd)
    put(a.W, vslot * N + cm, hdt * dd * go[0]);
    const double gd = hdt * double(v) * double(c.invd[d]) * go[0];
    const I3 qh = with(f, d, fhi), ql = with(f, d, flo);
    put(a.fb[d], c.fid(d, qh.j, qh.i, qh.k), gd);
    put(a.fb[d], c.fid(d, ql.j, ql.i, ql.k), -gd);
    return;
  }
  // 3D (recon_b_term_3d)
  const int bc = d == 2 ? c.bc_zrec : c.bc[d];
  const int64_t n = c.n_along(d);
  const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
  const I3 pm = with(f, d, map1(m, n, bc));
  const int64_t flo = m < 0 ? (bc == OUTFLOW ? 0 : n - 1) : m;
  const int64_t fhi = m + 1 > n ? (bc == OUTFLOW ? n : 1) : m + 1;
  const V dbn = (face_val(c, c.fb[d], d, f, fhi) - face_val(c, c.fb[d], d, f, flo)) * c.invd[d];
  const V db1 = (face_val(c, c.fb[t1], t1, pm, co(pm, t1) + 1) - face_val(c, c.fb[t1], t1, pm, co(pm, t1))) *
                c.invd[t1];
  const V db2 = (face_val(c, c.fb[t2], t2, pm, co(pm, t2) + 1) - face_val(c, c.fb[t2], t2, pm, co(pm, t2))) *
                c.invd[t2];
  const int p1 = minmod_recon_pick(dbn, db1, d == 1), p2 = minmod_recon_pick(dbn, db2, false);
  const double term1 = p1 == 1 ? double(dbn) : (p1 == -1 ? -double(db1) : 0.0);
  const double term2 = p2 == 1 ? double(dbn) : (p2 == -1 ? -double(db2) : 0.0);
  const int64_t cm = c.cid(pm.j, pm.i, pm.k);
  const V rho = ld(c.W, cm);
  const V v1 = ld(c.U + (1 + t1) * N, cm) / rho;
  const V v2 = ld(c.U + (1 + t2) * N, cm) / rho;
  const double r = double(rho), w1 = double(v1), w2 = double(v2);
  // out0 = hdt (v1 term1), out1 = hdt (v2 term2), v_k = M_k / rho
  const double gv1 = hdt * term1 * go[0], gv2 = hdt * term2 * go[1];
  const double gt1 = hdt * w1 * go[0], gt2 = hdt * w2 * go[1];
  put(a.U, (1 + t1) * N + cm, gv1 / r);
  put(a.U, (1 + t2) * N + cm, gv2 / r);
  put(a.W, cm, -(gv1 * w1 + gv2 * w2) / r);
  double gdbn = 0.0, gdb1 = 0.0, gdb2 = 0.0;
  if (p1 == 1) gdbn += gt1; else if (p1 == -1) gdb1 -= gt1;
  if (p2 == 1) gdbn += gt2; else if (p2 == -1) gdb2 -= gt2;
  const double gn = gdbn * double(c.invd[d]), g1 = gdb1 * double(c.invd[t1]), g2 = gdb2 * double(c.invd[t2]);
  const I3 qh = with(f, d, fhi), ql = with(f, d, flo);
  put(a.fb[d], c.fid(d, qh.j, qh.i, qh.k), gn);
  put(a.fb[d], c.fid(d, ql.j, ql.i, ql.k), -gn);
  const I3 p1h = with(pm, t1, co(pm, t1) + 1), p2h = with(pm, t2, co(pm, t2) + 1);
  put(a.fb[t1], c.fid(t1, p1h.j, p1h.i, p1h.k), g1);
  put(a.fb[t1], c.fid(t1, pm.j, pm.i, pm.k), -g1);
  put(a.fb[t2], c.fid(t2, p2h.j, p2h.i, p2h.k), g2);
  put(a.fb[t2], c.fid(t2, pm.j, pm.i, pm.k), -g2);
}

// (MHD in 2D and 3D: the adjoint takes no 1D MHD run)
template <class Ph, class V, class Add, class Share = NoShare>
PMHD_FI bool adjoint_predict(const Ctx<V>& c, const Ctx<double>& a, I3 f, Add add, int lane = 0, int lanes = 1,
                             Share share = Share()) {
  constexpr int Dir = Ph::dir;
  using D1 = ADual<V>;
  auto put = [&](const double* arr, int64_t i, double v) { add(const_cast<double*>(arr) + i, v); };
  const int64_t NF = c.nface(Dir);
  const int64_t idx = c.fid(Dir, f.j, f.i, f.k);
  const int64_t e = co(f, Dir);
  // the prepared states (predict_face_src's values)
  V wl[7], wr[7];
  load7(c.WL, NF, idx, wl);
  load7(c.WR, NF, idx, wr);
  const bool fl0 = wl[0] < c.min_dens, fl6 = wl[6] < c.min_pres;
  const bool fr0 = wr[0] < c.min_dens, fr6 = wr[6] < c.min_pres;
  apply_floors(wl, c.min_dens, c.min_pres);
  apply_floors(wr, c.min_dens, c.min_pres);
  Ctx<V> cv = c;
  cv.fcnt = nullptr;
  V cl[2], cr[2];
  cell_recon_src(cv, Dir, f, e - 1, cl);
  cell_recon_src(cv, Dir, f, e, cr);
  int ke = 0;
  if (c.ndim == 3) {
    wl[4] = wl[4] + cl[0];
    wl[5] = wl[5] + cl[1];
    wr[4] = wr[4] + cr[0];
    wr[5] = wr[5] + cr[1];
    ke = Dir;
  } else {
    const int slot = Dir == 0 ? 4 : 5;
    wl[slot] = wl[slot] + cl[0];
    wr[slot] = wr[slot] + cr[0];
  }
  const V bn = V(ld(c.fb[Dir], idx));
  // cotangents of the outputs U_L, U_R, F
  double gUL[7], gUR[7], gF[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    gUL[v] = a.UL[Dir][v * NF + idx];
    gUR[v] = a.UR[Dir][v * NF + idx];
    gF[v] = a.F[Dir][v * NF + idx];
  }
  AdHot hot{-1, -1, 0.0};
  Ctx<D1> d;
  make_ctx_of<D1>(c, a, [&](const V* val, double* adj, int id) { return AdPtr<V>{val, adj, &hot, 0, id}; }, d);
  // stage 1: the outputs' sensitivities to the 15 prepared values
  double g[15];
#pragma unroll
  for (int q = 0; q < 15; ++q) g[q] = 0.0;
  for (int q = lane; q < 15; q += lanes) {
    D1 y[15];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      y[k] = D1(wl[k], k == q ? V(1) : V(0), 0);
      y[7 + k] = D1(wr[k], 7 + k == q ? V(1) : V(0), 0);
    }
    y[14] = D1(bn, q == 14 ? V(1) : V(0), 0);
    D1 ul[7], ur[7], F[7];
    predict_flux<Ph::solver>(d, Dir, ke, y, y + 7, y[14], ul, ur, F);
    double acc = 0.0;
#pragma unroll
    for (int v = 0; v < 7; ++v) acc += gUL[v] * double(ul[v].t) + gUR[v] * double(ur[v].t) + gF[v] * double(F[v].t);
    g[q] = acc;
  }
  share(g);
  if (lane != 0) return true;
  // stage 2: the reverse of the preparation
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    put(a.WL, v * NF + idx, ((v == 0 && fl0) || (v == 6 && fl6)) ? 0.0 : g[v]);
    put(a.WR, v * NF + idx, ((v == 0 && fr0) || (v == 6 && fr6)) ? 0.0 : g[7 + v]);
  }
  put(a.fb[Dir], idx, g[14]);
  double gl[2], gr[2];
  if (c.ndim == 3) {
    gl[0] = g[4];
    gl[1] = g[5];
    gr[0] = g[11];
    gr[1] = g[12];
  } else {
    const int slot = Dir == 0 ? 4 : 5;
    gl[0] = g[slot];
    gr[0] = g[7 + slot];
    gl[1] = gr[1] = 0.0;
  }
  recon_src_vjp(cv, a, Dir, f, e - 1, gl, put);
  recon_src_vjp(cv, a, Dir, f, e, gr, put);
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    a.UL[Dir][v * NF + idx] = 0.0;
    a.UR[Dir][v * NF + idx] = 0.0;
    a.F[Dir][v * NF + idx] = 0.0;
  }
  return true;
}

// the two-stage element adjoints (corrector and predictor faces), for a kernel that
// splits their first stage over lanes
template <class Ph, class V, class Add, class Share>
PMHD_FI bool adjoint_staged(const Ctx<V>& c, const Ctx<double>& a, I3 f, Add add, int lane, int lanes, Share share) {
  if constexpr (is_predict<Ph>::value) {
    return adjoint_predict<Ph>(c, a, f, add, lane, lanes, share);
  } else {
    return adjoint_correct<Ph>(c, a, f, add, lane, lanes, share);
  }
}

template <class Ph>
struct is_update : std::false_type {};
template <>
struct is_update<Update> : std::true_type {};

// Adjoint of one cell of the MHD update (k_update, 2D and 3D; the adjoint's steps
// store no cell E in it): the new conserved state is linear in the old one, the
// corrector fluxes and the face field, and the new primitive state is cons->prim
// of it.  The reverse is written out: the cotangents of U2 and W2 go back through
// cons->prim to the new state, then through the flux differences and face averages.
template <class V, class Add>
PMHD_FI bool adjoint_update(const Ctx<V>& c, const Ctx<double>& a, I3 p, Add add) {
  auto put = [&](const double* arr, int64_t i, double v) { add(const_cast<double*>(arr) + i, v); };
  const int64_t N = c.ncell();
  const int64_t idx = c.cid(p.j, p.i, p.k);
  const bool d3 = c.ndim == 3;
  const int dv[7] = {0, 1, 2, 3, 5, 6, 7};
  const int64_t NX = c.nface(0), NY = c.nface(1), NZ = d3 ? c.nface(2) : 0;
  const int64_t fx0 = c.fid(0, p.j, p.i, p.k), fx1 = c.fid(0, p.j, p.i + 1, p.k);
  const int64_t fy0 = c.fid(1, p.j, p.i, p.k), fy1 = c.fid(1, p.j + 1, p.i, p.k);
  const int64_t fz0 = d3 ? c.fid(2, p.j, p.i, p.k) : 0, fz1 = d3 ? c.fid(2, p.j, p.i, p.k + 1) : 0;
  // the forward values (k_update's expressions)
  V u[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) u[v] = ld(c.U + dv[v] * N, idx);
  V nbz = V(0);
  if (!d3)
    nbz = u[5] - c.dtd[0] * (ld(c.Fc[0] + 5 * NX, fx1) - ld(c.Fc[0] + 5 * NX, fx0))
          - c.dtd[1] * (ld(c.Fc[1] + 4 * NY, fy1) - ld(c.Fc[1] + 4 * NY, fy0));
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    u[v] = u[v] - c.dtd[0] * (ld(c.Fc[0] + v * NX, fx1) - ld(c.Fc[0] + v * NX, fx0));
    u[v] = u[v] - c.dtd[1] * (ld(c.Fc[1] + v * NY, fy1) - ld(c.Fc[1] + v * NY, fy0));
    if (d3) u[v] = u[v] - c.dtd[2] * (ld(c.Fc[2] + v * NZ, fz1) - ld(c.Fc[2] + v * NZ, fz0));
  }
  V bx;
  if (!d3) {
    bx = (V(ld(c.fb2[0], fx0)) + V(ld(c.fb2[0], fx1))) / V(2);
    u[4] = (V(ld(c.fb2[1], fy0)) + V(ld(c.fb2[1], fy1))) / V(2);
    u[5] = nbz;
  } else {
    bx = V(0.5) * (V(ld(c.fb2[0], fx0)) + V(ld(c.fb2[0], fx1)));
    u[4] = V(0.5) * (V(ld(c.fb2[1], fy0)) + V(ld(c.fb2[1], fy1)));
    u[5] = V(0.5) * (V(ld(c.fb2[2], fz0)) + V(ld(c.fb2[2], fz1)));
  }
  const V invv = V(1) / u[0];
  const double inv = double(invv);
  const double ke = double(ke_sum(u, 0));
  const double gm1 = double(c.gamma - V(1));
  double x[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) x[v] = double(u[v]);
  const double b = double(bx);
  // cotangents of the outputs U2 = (u0..u3, bx, u4, u5, u6), W2 = (w0..w3, bx, w4, w5, w6)
  double gU[8], gW[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    gU[q] = a.U2[q * N + idx];
    gW[q] = a.W2[q * N + idx];
  }
  double gu[7] = {gU[0] + gW[0], gU[1], gU[2], gU[3], gU[5] + gW[5], gU[6] + gW[6], gU[7]};
  double gb = gU[4] + gW[4];
  double ginv = 0.0;
#pragma unroll
  for (int k = 1; k <= 3; ++k) {
    gu[k] += gW[k] * inv;
    ginv += gW[k] * x[k];
  }
  // w6 = (gamma - 1) (u6 - ke inv / 2 - (bx^2 + u4^2 + u5^2) / 2), ke = u1^2 + u2^2 + u3^2
  const double gp = gm1 * gW[7];
  gu[6] += gp;
  const double gke = -0.5 * inv * gp;
  ginv += -0.5 * ke * gp;
  gb += -b * gp;
  gu[4] += -x[4] * gp;
  gu[5] += -x[5] * gp;
#pragma unroll
  for (int k = 1; k <= 3; ++k) gu[k] += 2.0 * x[k] * gke;
  gu[0] += -inv * inv * ginv;
  // back through the update
  const double hx = double(c.dtd[0]), hy = double(c.dtd[1]), hz = d3 ? double(c.dtd[2]) : 0.0;
  const int lin[5] = {0, 1, 2, 3, 6};
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const int v = lin[q];
    put(a.U, dv[v] * N + idx, gu[v]);
    put(a.Fc[0], v * NX + fx1, -hx * gu[v]);
    put(a.Fc[0], v * NX + fx0, hx * gu[v]);
    put(a.Fc[1], v * NY + fy1, -hy * gu[v]);
    put(a.Fc[1], v * NY + fy0, hy * gu[v]);
    if (d3) {
      put(a.Fc[2], v * NZ + fz1, -hz * gu[v]);
      put(a.Fc[2], v * NZ + fz0, hz * gu[v]);
    }
  }
  put(a.fb2[0], fx0, 0.5 * gb);
  put(a.fb2[0], fx1, 0.5 * gb);
  put(a.fb2[1], fy0, 0.5 * gu[4]);
  put(a.fb2[1], fy1, 0.5 * gu[4]);
  if (d3) {
    put(a.fb2[2], fz0, 0.5 * gu[5]);
    put(a.fb2[2], fz1, 0.5 * gu[5]);
  } else {
    // u5 = the reference's new B_z: old B_z minus the x flux of its slot 5 and the y
    // flux of its slot 4
    put(a.U, 6 * N + idx, gu[5]);
    put(a.Fc[0], 5 * NX + fx1, -hx * gu[5]);
    put(a.Fc[0], 5 * NX + fx0, hx * gu[5]);
    put(a.Fc[1], 4 * NY + fy1, -hy * gu[5]);
    put(a.Fc[1], 4 * NY + fy0, hy * gu[5]);
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    a.U2[q * N + idx] = 0.0;
    a.W2[q * N + idx] = 0.0;
  }
  return true;
}

// the slot lists of one element kept from an earlier recording pass of the same
// launch (rows of E elements: inputs in rows 0 .. nin-1, outputs from row rin), so
// later evaluations skip the recording pass; write: record and keep them
struct SlotView {
  uint64_t* slots;
  int32_t* nin;
  int32_t* nout;
  int64_t E, e;
  int rin;
  bool write;
};
// the cotangent array of array id (the ids make_ctx_of gives)
PMHD_FI double* adj_array(const Ctx<double>& a, int id) {
  if (id < 24) {
    const int q = id >> 3;
    switch (id & 7) {
      case 0: return const_cast<double*>(a.fb[q]);
      case 1: return a.fb2[q];
      case 2: return a.nfb[q];
      case 3: return a.UL[q];
      case 4: return a.UR[q];
      case 5: return a.F[q];
      case 6: return a.Fc[q];
      default: return a.E[q];
    }
  }
  switch (id) {
    case 24: return const_cast<double*>(a.U);
    case 25: return const_cast<double*>(a.W);
    case 26: return a.U2;
    case 27: return a.W2;
    case 28: return a.WL;
    case 29: return a.WR;
    case 30: return a.S;
    default: return a.Ec;
  }
}
PMHD_FI uint64_t slot_code(const AdSlot& s) { return (uint64_t(s.id) << 56) | uint64_t(s.i); }
PMHD_FI AdSlot slot_of(uint64_t v) { return AdSlot{int(v >> 56), int64_t(v & ((uint64_t(1) << 56) - 1))}; }

// the adjoint of one output element through its whole body: a recording pass (or
// the lists an earlier one kept), then one seeded pass per input slot
template <class Ph, class V, class Add>
PMHD_FI bool adjoint_generic(const Ctx<V>& c, const Ctx<double>& a, I3 p, AdState& st, Add add,
                             const SlotView* sv) {
  const bool kept = sv && !sv->write;
  if (!kept) {
    // recording pass on plain values
    Ctx<ARec<V>> r;
    make_ctx_of<ARec<V>>(c, a, [&](const V* val, double*, int id) { return RecPtr<V>{val, &st, 0, id}; }, r);
    st.nin = 0;
    st.nout = 0;
    st.overflow = 0;
    call_phase<Ph>(r, p);
    if (st.overflow) return false;
    ad_prune(st);
    if (sv) {
      for (int k = 0; k < st.nin; ++k) sv->slots[k * sv->E + sv->e] = slot_code(st.in[k]);
      for (int k = 0; k < st.nout; ++k) sv->slots[(sv->rin + k) * sv->E + sv->e] = slot_code(st.out[k]);
      sv->nin[sv->e] = st.nin;
      sv->nout[sv->e] = st.nout;
    }
  }
  const int nin = kept ? sv->nin[sv->e] : st.nin;
  const int nout = kept ? sv->nout[sv->e] : st.nout;
  AdHot hot{-1, -1, 0.0};
  Ctx<ADual<V>> d;
  make_ctx_of<ADual<V>>(c, a, [&](const V* val, double* adj, int id) { return AdPtr<V>{val, adj, &hot, 0, id}; }, d);
  for (int k = 0; k < nin; ++k) {
    const AdSlot s = kept ? slot_of(sv->slots[k * sv->E + sv->e]) : st.in[k];
    hot.seed_id = s.id;
    hot.seed_i = s.i;
    hot.acc = 0.0;
    call_phase<Ph>(d, p);
    add(adj_array(a, s.id) + s.i, hot.acc);
  }
  for (int k = 0; k < nout; ++k) {
    const AdSlot s = kept ? slot_of(sv->slots[(sv->rin + k) * sv->E + sv->e]) : st.out[k];
    adj_array(a, s.id)[s.i] = 0.0;
  }
  return true;
}

// the generic element adjoint from kept slot lists, split over `lanes` threads: the
// thread of lane L seeds the element's input slots L, L + lanes, ... and writes the
// share of slot k to vals[k * E + e].  The threads of one element must be lanes of
// one wave: after the passes of all of them, lane 0 clears the cotangents of the
// element's outputs.
template <class Ph, class V>
PMHD_FI void adjoint_generic_lanes(const Ctx<V>& c, const Ctx<double>& a, I3 p, const SlotView& sv, int lane,
                                   int lanes, double* vals) {
  const int nin = sv.nin[sv.e];
  AdHot hot{-1, -1, 0.0};
  Ctx<ADual<V>> d;
  make_ctx_of<ADual<V>>(c, a, [&](const V* val, double* adj, int id) { return AdPtr<V>{val, adj, &hot, 0, id}; }, d);
  for (int k = lane; k < nin; k += lanes) {
    const AdSlot s = slot_of(sv.slots[k * sv.E + sv.e]);
    hot.seed_id = s.id;
    hot.seed_i = s.i;
    hot.acc = 0.0;
    call_phase<Ph>(d, p);
    vals[k * sv.E + sv.e] = hot.acc;
  }
  if (lane == 0) {
    const int nout = sv.nout[sv.e];
    for (int k = 0; k < nout; ++k) {
      const AdSlot s = slot_of(sv.slots[(sv.rin + k) * sv.E + sv.e]);
      adj_array(a, s.id)[s.i] = 0.0;
    }
  }
}

// whether the element adjoint of Ph records slot lists (the generic path)
// (PMHD_ADJ_DUAL_CORRECTOR: every phase through the per-slot dual passes, the
// reference the hand-written reverses are tested against)
template <class Ph>
constexpr bool uses_slots() {
#ifdef PMHD_ADJ_DUAL_CORRECTOR
  return true;
#else
  return !is_correct<Ph>::value && !is_update<Ph>::value && !is_predict<Ph>::value;
#endif
}

// at most this many contributions per element (input slots of the generic path;
// the corrector's reverse adds at most 95)
constexpr int kAdjMaxContrib = 160;

// adjoint of one output element of phase Ph; returns false when the element has
// more input or output slots than the state holds.  add(p, v) receives each of the
// element's contributions to the cotangent arrays, zeros included, in an order and
// to targets that depend only on the element's position and the configuration (the
// deterministic accumulation relies on this; atomic sinks skip the zeros)
template <class Ph, class V, class Add>
PMHD_FI bool adjoint_element(const Ctx<V>& c, const Ctx<double>& a, I3 p, AdState& st, Add add,
                             const SlotView* sv = nullptr) {
  if constexpr (!uses_slots<Ph>() && is_correct<Ph>::value) {
    (void)st;
    return adjoint_correct<Ph>(c, a, p, add);
  } else if constexpr (!uses_slots<Ph>() && is_predict<Ph>::value) {
    (void)st;
    return adjoint_predict<Ph>(c, a, p, add);
  } else if constexpr (!uses_slots<Ph>() && is_update<Ph>::value) {
    (void)st;
    return adjoint_update(c, a, p, add);
  } else {
    return adjoint_generic<Ph>(c, a, p, st, add, sv);
  }
}

// ---------------------------------------------------------------------------
// buffers of one recomputed step: each phase output has its own array, so
// every phase's inputs are still intact when its adjoint runs
// ---------------------------------------------------------------------------
template <class T>
struct StepBufs {
  T* WL[3] = {};
  T* WR[3] = {};
  T* S[3] = {};
  T* UL[3] = {};
  T* UR[3] = {};
  T* F[3] = {};
  T* Fc[3] = {};
  T* nfb[3] = {};
  T* Ec_p = nullptr;   // cell E of the predictor (CellE)
  T* Ec_h = nullptr;   // cell E of the half step (HalfE)
  T* E_p[3] = {};      // corner EMFs of the predictor fluxes
  T* E_c[3] = {};      // corner EMFs of the corrector fluxes
  T* U2 = nullptr;
  T* W2 = nullptr;
  T* fb2[3] = {};
};

template <class T, class U, class A>
void alloc_step(const Ctx<T>& c, StepBufs<U>& s, A&& alloc) {
  const int nd = c.ndim;
  const int64_t N = c.ncell();
  const int64_t cv = cell_vars(c.physics), fv = face_vars(c.physics);
  const bool mhd = c.physics == 0;
  for (int d = 0; d < 3; ++d) {
    const bool used = d < nd;
    s.WL[d] = alloc(used ? fv * c.nface(d) : 1);
    s.WR[d] = alloc(used ? fv * c.nface(d) : 1);
    s.S[d] = alloc(used && c.order > 0 ? fv * c.nslope(d) : 1);
    s.UL[d] = alloc(used ? fv * c.nface(d) : 1);
    s.UR[d] = alloc(used ? fv * c.nface(d) : 1);
    s.F[d] = alloc(used ? fv * c.nface(d) : 1);
    s.Fc[d] = alloc(used && nd >= 2 ? fv * c.nface(d) : 1);
    s.nfb[d] = alloc(mhd && used && nd >= 2 ? c.nface(d) : 1);
    s.E_p[d] = alloc(mhd && (nd == 3 || (nd == 2 && d == 2)) ? c.nedge() : 1);
    s.E_c[d] = alloc(mhd && (nd == 3 || (nd == 2 && d == 2)) ? c.nedge() : 1);
    s.fb2[d] = alloc(mhd && used && nd >= 2 ? c.nface(d) : 1);
  }
  s.Ec_p = alloc(mhd ? (nd == 3 ? 3 * N : N) : 1);
  s.Ec_h = alloc(mhd ? (nd == 3 ? 3 * N : N) : 1);
  s.U2 = alloc(cv * N);
  s.W2 = alloc(cv * N);
}

// the arrays a launch of id sees (values or cotangents of the same layout)
template <class T>
void bind_step(Ctx<T>& c, const StepBufs<T>& s, int id) {
  int kind, dir;
  phase_shape(id, kind, dir);
  for (int d = 0; d < 3; ++d) {
    c.UL[d] = s.UL[d];
    c.UR[d] = s.UR[d];
    c.F[d] = s.F[d];
    c.Fc[d] = s.Fc[d];
    c.nfb[d] = s.nfb[d];
    c.fb2[d] = s.fb2[d];
  }
  c.U2 = s.U2;
  c.W2 = s.W2;
  c.WL = s.WL[dir];
  c.WR = s.WR[dir];
  c.S = s.S[dir];
  const bool half = id == K_HALF_E || id == K_EMF2 || (id >= K_CT_FULL && id < K_CT_FULL + 3) || id == K_UPDATE;
  c.Ec = half ? s.Ec_h : s.Ec_p;
  for (int d = 0; d < 3; ++d) c.E[d] = half ? s.E_c[d] : s.E_p[d];
}

// runs a cycle forward while writing down the adjoint launches it implies
template <class T, class B>
struct TapeExec {
  B* be = nullptr;
  const StepBufs<T>* vals = nullptr;
  const StepBufs<double>* adjs = nullptr;
  Ctx<double> abase;                       // state cotangent pointers of this step
  std::vector<std::function<void()>> tape;

  template <class Ph>
  void record(const Ctx<T>& c, Ctx<T>& cv, Ctx<double>& av) {
    cv = c;
    bind_step(cv, *vals, Ph::id);
    av = abase;
    bind_step(av, *adjs, Ph::id);
  }
  bool fused = false;   // a fused launch was met (it has no per-element adjoint)
  template <class Ph>
  void run(const Ctx<T>& c) {
    if constexpr (Ph::kind == EXT_SEG || Ph::kind == EXT_LINE || Ph::kind == EXT_ENDS) {
      // the adjoint's contexts keep seg = 0: separate slope, edge and predictor
      // launches, whose elements are small enough to differentiate one by one
      (void)c;
      fused = true;
    } else {
      Ctx<T> cv;
      Ctx<double> av;
      record<Ph>(c, cv, av);
      be->exec().template run<Ph>(cv);
      B* b = be;
      tape.push_back([b, cv, av] { b->template adjoint_phase<Ph>(cv, av); });
    }
  }
  template <class Ph>
  void run_side(const Ctx<T>& c) {
    run<Ph>(c);
  }
  void join_side() {}
  template <class Ph>
  void run_max(const Ctx<T>& c) {
    Ctx<T> cv;
    Ctx<double> av;
    record<Ph>(c, cv, av);
    be->exec().template run_max<Ph>(cv);
    B* b = be;
    tape.push_back([b, cv, av] { b->template adjoint_phase<Ph>(cv, av); });
  }
};

template <class T, class B>
struct AdjointHandle : public pmhd_adjoint {
  // the adjoint keeps whole faces in the cell type (its unit does not define PMHD_FACES_SPLIT)
  static_assert(!split_faces<T>(), "adjoint contexts store whole faces");
  B be;
  pmhd_grid g;
  pmhd_params p;
  Ctx<T> c;                 // forward context (single-grid layout)
  StateBufs<T> sb;
  int every = 1;
  std::vector<void*> owned;
  // checkpoints and the states of the segment being reversed: (U, W, fb) each
  struct Snap {
    T* U;
    T* W;
    T* fb[3];
  };
  std::vector<Snap> ckpt;
  std::vector<Snap> seg;
  StepBufs<T> vals;
  // value buffers of the further steps of a segment, when they fit the budget below
  // (then one forward pass over a segment records every step's tape)
  std::vector<StepBufs<T>> vals_more;
  bool vals_more_failed = false;
  // the budget (bytes; environment PMHD_ADJ_KEEP_BYTES at creation overrides it)
  int64_t keep_bytes = int64_t(4) << 30;
  StepBufs<double> adjs;
  double* adj_state[2][5] = {};   // cotangents of (U, W, fb0, fb1, fb2), two sets
  double dt = 0.0;
  int64_t nsteps = 0;
  bool have_forward = false;

  ~AdjointHandle() override {
    for (void* ptr : owned) be.release(ptr);
  }
  void* dalloc_bytes(size_t bytes) {
    void* ptr = be.alloc(bytes > 0 ? bytes : 8);
    if (!ptr) throw std::bad_alloc();
    owned.push_back(ptr);
    return ptr;
  }
  T* dalloc(int64_t n) { return static_cast<T*>(dalloc_bytes(size_t(n > 0 ? n : 1) * sizeof(T))); }
  double* dalloc_d(int64_t n) {
    double* ptr = static_cast<double*>(dalloc_bytes(size_t(n > 0 ? n : 1) * sizeof(double)));
    be.zero(ptr, size_t(n > 0 ? n : 1) * sizeof(double));
    return ptr;
  }
  int64_t nv() const { return cell_vars(p.physics); }
  int64_t nface(int d) const { return p.physics == 0 && d < g.ndim && g.ndim >= 2 ? c.nface(d) : 0; }

  Snap new_snap() {
    Snap s;
    s.U = dalloc(nv() * c.ncell());
    s.W = dalloc(nv() * c.ncell());
    for (int d = 0; d < 3; ++d) s.fb[d] = dalloc(nface(d) > 0 ? nface(d) : 1);
    return s;
  }
  void copy_snap(T* U, T* W, T* const fb[3], const T* U0, const T* W0, const T* const fb0[3]) {
    be.d2d(U, U0, sizeof(T) * size_t(nv() * c.ncell()));
    be.d2d(W, W0, sizeof(T) * size_t(nv() * c.ncell()));
    for (int d = 0; d < 3; ++d)
      if (nface(d) > 0) be.d2d(fb[d], fb0[d], sizeof(T) * size_t(nface(d)));
  }

  int init(const pmhd_grid& g_, const pmhd_params& p_, int every_) override {
    g = g_;
    p = p_;
    every = every_ > 0 ? every_ : 1;
    if (const char* kb = std::getenv("PMHD_ADJ_KEEP_BYTES")) keep_bytes = std::atoll(kb);
    init_ctx(c, g, p);
    try {
      alloc_grid(c, sb, p.reconstruct_order, [&](int64_t n) { return dalloc(n); });
      alloc_step(c, vals, [&](int64_t n) { return dalloc(n); });
      // every cotangent array in one zeroed block: a contribution's target is then
      // an index into it (the deterministic accumulation sorts by that index)
      int64_t total = 0;
      double* arena = nullptr;
      auto sub = [&](int64_t n) {
        double* ptr = arena ? arena + total : nullptr;
        total += (std::max<int64_t>(n, 1) + 63) / 64 * 64;
        return ptr;
      };
      auto carve = [&] {
        alloc_step(c, adjs, sub);
        for (int s = 0; s < 2; ++s) {
          adj_state[s][0] = sub(nv() * c.ncell());
          adj_state[s][1] = sub(nv() * c.ncell());
          for (int d = 0; d < 3; ++d) adj_state[s][2 + d] = sub(nface(d) > 0 ? nface(d) : 1);
        }
      };
      carve();
      const int64_t len = total;
      arena = dalloc_d(len);
      total = 0;
      carve();
      be.set_arena(arena, len);
      for (int k = 0; k < every; ++k) seg.push_back(new_snap());
    } catch (const std::bad_alloc&) {
      return fail(PMHD_ENOMEM, "device allocation failed");
    }
    bind_bufs(c, sb, 0);
    be.zero(c.fcnt, FLOOR_WORDS * sizeof(unsigned long long));
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  void set_dt_scalars() {
    set_step_scalars(c, dt);
    const long double hl = 0.5L * (static_cast<long double>(dt) / static_cast<long double>(g.dy));
    c.ldh[0] = static_cast<double>(hl);
    c.ldh[1] = static_cast<double>(hl - static_cast<long double>(c.ldh[0]));
  }

  // one forward step of the single-grid layout; the state ends in buffer set cur
  int cur = 0;
  void step_forward() {
    bind_bufs(c, sb, cur);
    run_cycle(be.exec(), c);
    cur = 1 - cur;
    bind_bufs(c, sb, cur);
  }

  int forward(const double* prim, const double* cons, const double* const fin[3], double dt_, int64_t n,
              double* prim_out, double* cons_out, double* const fout[3], bool dev) override {
    if (!(dt_ > 0) || n < 1) return fail(PMHD_EINVAL, "dt must be positive and nsteps at least 1");
    dt = dt_;
    nsteps = n;
    cur = 0;
    bind_bufs(c, sb, 0);
    int rc = dev ? load_state_dev(be, c, g, prim, cons, fin, sb.U[0], sb.W[0], sb.FB[0], sb.FBLO[0], err)
                 : load_state(be, c, g, prim, cons, fin, sb.U[0], sb.W[0], sb.FB[0], sb.FBLO[0], err);
    if (rc) return rc;
    set_dt_scalars();
    const int64_t nck = (n + every - 1) / every;
    try {
      while (int64_t(ckpt.size()) < nck) ckpt.push_back(new_snap());
    } catch (const std::bad_alloc&) {
      return fail(PMHD_ENOMEM, "device allocation failed (checkpoints)");
    }
    for (int64_t s = 0; s < n; ++s) {
      if (s % every == 0) {
        const Snap& k = ckpt[size_t(s / every)];
        copy_snap(k.U, k.W, k.fb, sb.U[cur], sb.W[cur], sb.FB[cur]);
      }
      step_forward();
    }
    double* fo[3] = {fout[0], fout[1], fout[2]};
    if (dev) store_state_dev(be, c, g, sb.U[cur], sb.W[cur], sb.FB[cur], sb.FBLO[cur], prim_out, cons_out, fo);
    else store_state(be, c, g, sb.U[cur], sb.W[cur], sb.FB[cur], sb.FBLO[cur], prim_out, cons_out, fo);
    be.sync();
    have_forward = true;
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  void upload_adj(double* dst, const double* src, int64_t n, bool dev) {
    if (src && dev) be.d2d(dst, src, sizeof(double) * size_t(n));
    else if (src) be.h2d(dst, src, sizeof(double) * size_t(n));
    else be.zero(dst, sizeof(double) * size_t(n));
  }

  int set_deterministic(int on) override {
    be.set_deterministic(on != 0);
    return PMHD_OK;
  }

  int set_stream(void* stream) override {
    be.set_stream(stream);
    const char* e = be.error();
    return e ? fail(PMHD_EHIP, e) : PMHD_OK;
  }

  // the cotangent set a step's inputs use: the last step (nsteps - 1) takes set 1,
  // and the sets alternate from there
  int in_parity(int64_t s) const { return 1 - int((nsteps - 1 - s) % 2); }

  // runs step cs forward into the value buffers vb while recording its adjoint tape
  // (the tape's cotangent pointers: the state set `in`)
  void record_step(TapeExec<T, B>& tx, const Ctx<T>& cs, const StepBufs<T>& vb, int in) {
    tx.be = &be;
    tx.vals = &vb;
    tx.adjs = &adjs;
    std::memset(&tx.abase, 0, sizeof(tx.abase));
    tx.abase.U = adj_state[in][0];
    tx.abase.W = adj_state[in][1];
    for (int d = 0; d < 3; ++d) tx.abase.fb[d] = adj_state[in][2 + d];
    run_cycle(tx, cs);
  }

  // value buffers of the i-th step of a segment
  const StepBufs<T>& step_vals(int64_t i) const { return i == 0 ? vals : vals_more[size_t(i - 1)]; }

  // whether every step of a segment of len steps can have its own value buffers
  // (within keep_bytes; allocated on first use)
  bool keep_values(int64_t len) {
    if (len <= 1) return true;
    if (vals_more_failed) return false;
    if (int64_t(vals_more.size()) >= len - 1) return true;
    int64_t words = 0;
    StepBufs<T> probe;
    alloc_step(c, probe, [&](int64_t n) {
      words += n > 0 ? n : 1;
      return static_cast<T*>(nullptr);
    });
    if (words * int64_t(sizeof(T)) * len > keep_bytes) {
      vals_more_failed = true;
      return false;
    }
    try {
      while (int64_t(vals_more.size()) < len - 1) {
        StepBufs<T> b;
        alloc_step(c, b, [&](int64_t n) { return dalloc(n); });
        vals_more.push_back(b);
      }
    } catch (const std::bad_alloc&) {
      vals_more_failed = true;
      return false;
    }
    return true;
  }

  int backward(const double* gprim, const double* gcons, const double* const gf[3], double* gprim0,
               double* gcons0, double* const gf0[3], bool dev) override {
    if (!have_forward) return fail(PMHD_EINVAL, "backward needs a preceding forward");
    const int64_t N = c.ncell();
    int out = 0;   // adj_state[out]: cotangent of the current step's output state
    upload_adj(adj_state[out][0], gcons, nv() * N, dev);
    upload_adj(adj_state[out][1], gprim, nv() * N, dev);
    for (int d = 0; d < 3; ++d)
      if (nface(d) > 0) upload_adj(adj_state[out][2 + d], gf[d], nface(d), dev);
    const int64_t nck = (nsteps + every - 1) / every;
    for (int64_t sgi = nck - 1; sgi >= 0; --sgi) {
      const int64_t s0 = sgi * every;
      const int64_t s1 = std::min<int64_t>(s0 + every, nsteps);
      const Snap& k = ckpt[size_t(sgi)];
      const bool keep = keep_values(s1 - s0);
      std::vector<TapeExec<T, B>> txs(size_t(keep ? s1 - s0 : 0));
      if (keep) {
        // one forward pass over the segment records every step's tape, each step with
        // its own value buffers; step s reads the outputs of step s - 1 where they are
        const Snap& q = seg[0];
        copy_snap(q.U, q.W, q.fb, k.U, k.W, k.fb);
        for (int64_t s = s0; s < s1; ++s) {
          const int64_t i = s - s0;
          Ctx<T> cs = c;
          if (i == 0) {
            cs.U = q.U;
            cs.W = q.W;
            for (int d = 0; d < 3; ++d) cs.fb[d] = q.fb[d];
          } else {
            const StepBufs<T>& pv = step_vals(i - 1);
            cs.U = pv.U2;
            cs.W = pv.W2;
            for (int d = 0; d < 3; ++d) cs.fb[d] = pv.fb2[d];
          }
          record_step(txs[size_t(i)], cs, step_vals(i), in_parity(s));
          if (txs[size_t(i)].fused) return fail(PMHD_EINVAL, "the adjoint needs separate predictor launches (seg = 0)");
        }
      } else {
        // the states of the segment, from its checkpoint
        cur = 0;
        copy_snap(sb.U[0], sb.W[0], sb.FB[0], k.U, k.W, k.fb);
        for (int64_t s = s0; s < s1; ++s) {
          const Snap& q = seg[size_t(s - s0)];
          copy_snap(q.U, q.W, q.fb, sb.U[cur], sb.W[cur], sb.FB[cur]);
          if (s + 1 < s1) step_forward();
        }
      }
      for (int64_t s = s1 - 1; s >= s0; --s) {
        const int in = 1 - out;
        for (int a = 0; a < 5; ++a) {
          const int64_t len = a < 2 ? nv() * N : nface(a - 2);
          if (len > 0) be.zero(adj_state[in][a], sizeof(double) * size_t(len));
        }
        // the output state's cotangents live where this step writes U2, W2, fb2
        be.d2d(adjs.U2, adj_state[out][0], sizeof(double) * size_t(nv() * N));
        be.d2d(adjs.W2, adj_state[out][1], sizeof(double) * size_t(nv() * N));
        for (int d = 0; d < 3; ++d)
          if (nface(d) > 0) be.d2d(adjs.fb2[d], adj_state[out][2 + d], sizeof(double) * size_t(nface(d)));
        TapeExec<T, B> one;
        if (!keep) {
          // the step's context: inputs from the stored state
          const Snap& q = seg[size_t(s - s0)];
          Ctx<T> cs = c;
          cs.U = q.U;
          cs.W = q.W;
          for (int d = 0; d < 3; ++d) cs.fb[d] = q.fb[d];
          record_step(one, cs, vals, in);
          if (one.fused) return fail(PMHD_EINVAL, "the adjoint needs separate predictor launches (seg = 0)");
        }
        TapeExec<T, B>& tx = keep ? txs[size_t(s - s0)] : one;
        for (size_t t = tx.tape.size(); t-- > 0;) {
          be.adjoint_select(in, t);   // the launch's accumulation plan (its targets depend on in)
          tx.tape[t]();
        }
        out = in;
      }
      if (be.adjoint_overflow())
        return fail(PMHD_EINVAL, "adjoint element with more contributions than supported, or whose "
                                 "contributions changed between steps");
      const char* ae = be.error();
      if (ae) return fail(PMHD_EHIP, ae);
    }
    be.adjoint_end();
    // cotangent of the initial state
    std::vector<double> tmp;
    auto down = [&](double* dst, const double* src, int64_t n) {
      if (dst && dev) be.d2d(dst, src, sizeof(double) * size_t(n));
      else if (dst) be.d2h(dst, src, sizeof(double) * size_t(n));
    };
    down(gcons0, adj_state[out][0], nv() * N);
    down(gprim0, adj_state[out][1], nv() * N);
    for (int d = 0; d < 3; ++d)
      if (nface(d) > 0) down(gf0[d], adj_state[out][2 + d], nface(d));
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
};

inline int create_adjoint(const pmhd_grid* g, const pmhd_params* p, int every, pmhd_adjoint** out,
                          std::string& err) {
  if (!g || !p || !out) {
    err = "null argument";
    return PMHD_EINVAL;
  }
  *out = nullptr;
  int rc = check_grid(g, err);
  if (rc) return rc;
  // the phase adjoints run 64-thread blocks
  rc = check_launch_extents(g, err, 64);
  if (rc) return rc;
  rc = check_params(p, err);
  if (rc) return rc;
  if (p->riemann != 0) {
    err = "the adjoint differentiates the Roe step only (riemann 0)";
    return PMHD_EINVAL;
  }
  if (g->ndim == 1 && p->physics == 0) {
    // the 1D MHD step holds B_x as a scalar of the context (the reference's uniform
    // Bx, Source/EvolveFunctions_MHD.py:209-212), which the dual evaluation cannot
    // differentiate: its cotangent would be silently zero
    err = "the adjoint supports 2D and 3D MHD and 1D-3D hydrodynamics, not 1D MHD";
    return PMHD_EINVAL;
  }
  if (every < 1) {
    err = "checkpoint_every must be at least 1";
    return PMHD_EINVAL;
  }
  const pmhd_grid gg = normal_grid(*g);
  pmhd_adjoint* h = p->precision == 64 ? pmhd_new_adjoint_f64() : pmhd_new_adjoint_f32();
  if (!h) { err = "out of host memory"; return PMHD_ENOMEM; }
  rc = h->init(gg, *p, every);
  if (rc != PMHD_OK) {
    err = h->err;
    delete h;
    return rc;
  }
  *out = h;
  return PMHD_OK;
}

PMHD_NS_END

#define PMHD_DEFINE_ADJOINT_FACTORIES(BACKEND)                                                      \
  pmhd_adjoint* pmhd_new_adjoint_f64() { return new (std::nothrow) pmhd::AdjointHandle<double, BACKEND>(); } \
  pmhd_adjoint* pmhd_new_adjoint_f32() { return new (std::nothrow) pmhd::AdjointHandle<float, BACKEND>(); }

#define PMHD_DEFINE_ADJOINT_ABI()                                                                   \
  static std::string g_pmhd_adj_error;                                                              \
  extern "C" int pmhd_adjoint_create(const pmhd_grid* g, const pmhd_params* p,                      \
                                     int32_t checkpoint_every, pmhd_adjoint** out) {                \
    try {                                                                                           \
      return pmhd::create_adjoint(g, p, checkpoint_every, out, g_pmhd_adj_error);                   \
    } catch (...) {                                                                                 \
      g_pmhd_adj_error = "unexpected exception in pmhd_adjoint_create";                             \
      return PMHD_EINVAL;                                                                           \
    }                                                                                               \
  }                                                                                                 \
  extern "C" int pmhd_adjoint_destroy(pmhd_adjoint* a) {                                            \
    delete a;                                                                                       \
    return PMHD_OK;                                                                                 \
  }                                                                                                 \
  extern "C" int pmhd_adjoint_forward(pmhd_adjoint* a, const double* prim, const double* cons,      \
                                      const double* fbx, const double* fby, const double* fbz,      \
                                      double dt, int64_t nsteps, double* prim_out,                  \
                                      double* cons_out, double* fbx_out, double* fby_out,           \
                                      double* fbz_out) {                                            \
    if (!a) return PMHD_EINVAL;                                                                     \
    const double* fin[3] = {fbx, fby, fbz};                                                         \
    double* fo[3] = {fbx_out, fby_out, fbz_out};                                                    \
    try { return a->forward(prim, cons, fin, dt, nsteps, prim_out, cons_out, fo, false); }          \
    catch (...) { return a->fail(PMHD_EINVAL, "exception in pmhd_adjoint_forward"); }               \
  }                                                                                                 \
  extern "C" int pmhd_adjoint_forward_device(pmhd_adjoint* a, const double* prim,                   \
                                             const double* cons, const double* fbx,                 \
                                             const double* fby, const double* fbz, double dt,       \
                                             int64_t nsteps, double* prim_out, double* cons_out,    \
                                             double* fbx_out, double* fby_out, double* fbz_out) {   \
    if (!a) return PMHD_EINVAL;                                                                     \
    const double* fin[3] = {fbx, fby, fbz};                                                         \
    double* fo[3] = {fbx_out, fby_out, fbz_out};                                                    \
    try { return a->forward(prim, cons, fin, dt, nsteps, prim_out, cons_out, fo, true); }           \
    catch (...) { return a->fail(PMHD_EINVAL, "exception in pmhd_adjoint_forward_device"); }        \
  }                                                                                                 \
  extern "C" int pmhd_adjoint_backward(pmhd_adjoint* a, const double* gprim, const double* gcons,   \
                                       const double* gfbx, const double* gfby, const double* gfbz,  \
                                       double* gprim0, double* gcons0, double* gfbx0,               \
                                       double* gfby0, double* gfbz0) {                              \
    if (!a) return PMHD_EINVAL;                                                                     \
    const double* gf[3] = {gfbx, gfby, gfbz};                                                       \
    double* g0[3] = {gfbx0, gfby0, gfbz0};                                                          \
    try { return a->backward(gprim, gcons, gf, gprim0, gcons0, g0, false); }                        \
    catch (...) { return a->fail(PMHD_EINVAL, "exception in pmhd_adjoint_backward"); }              \
  }                                                                                                 \
  extern "C" int pmhd_adjoint_backward_device(pmhd_adjoint* a, const double* gprim,                 \
                                              const double* gcons, const double* gfbx,              \
                                              const double* gfby, const double* gfbz,               \
                                              double* gprim0, double* gcons0, double* gfbx0,        \
                                              double* gfby0, double* gfbz0) {                       \
    if (!a) return PMHD_EINVAL;                                                                     \
    const double* gf[3] = {gfbx, gfby, gfbz};                                                       \
    double* g0[3] = {gfbx0, gfby0, gfbz0};                                                          \
    try { return a->backward(gprim, gcons, gf, gprim0, gcons0, g0, true); }                         \
    catch (...) { return a->fail(PMHD_EINVAL, "exception in pmhd_adjoint_backward_device"); }       \
  }                                                                                                 \
  extern "C" int pmhd_adjoint_set_stream(pmhd_adjoint* a, void* stream) {                           \
    if (!a) return PMHD_EINVAL;                                                                     \
    return a->set_stream(stream);                                                                   \
  }                                                                                                 \
  extern "C" int pmhd_adjoint_set_deterministic(pmhd_adjoint* a, int32_t enable) {                  \
    if (!a) return PMHD_EINVAL;                                                                     \
    return a->set_deterministic(enable);                                                            \
  }                                                                                                 \
  extern "C" int pmhd_adjoint_set_profiling(pmhd_adjoint* a, int32_t enable) {                      \
    if (!a) return PMHD_EINVAL;                                                                     \
    return a->set_profiling(enable);                                                                \
  }                                                                                                 \
  extern "C" int pmhd_adjoint_get_profile(pmhd_adjoint* a, int32_t max_entries, char* names,        \
                                          double* total_ms, int64_t* launches, int64_t* elements,   \
                                          int32_t* n_entries) {                                     \
    if (!a || !names || !total_ms || !launches || !elements || !n_entries) return PMHD_EINVAL;      \
    return a->get_profile(max_entries, names, total_ms, launches, elements, n_entries);             \
  }                                                                                                 \
  extern "C" const char* pmhd_adjoint_last_error(const pmhd_adjoint* a) {                           \
    return a ? a->err.c_str() : g_pmhd_adj_error.c_str();                                           \
  }
