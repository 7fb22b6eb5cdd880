// Hydrodynamic CTU step (Source/EvolveFunctions_Hydro.py:232-806): the kernel
// bodies and phases of one cycle, on the same contexts and executors as the MHD
// step.  State: five variables stored unrotated, primitive (rho, v_x, v_y, v_z, p)
// and conserved (rho, M_x, M_y, M_z, E); direction D has its normal velocity in
// slot 1 + D.  Interface states and fluxes are 5-vectors in that storage order.
//
// The reference builds every hydro eigen-system densely in storage order and sums
// every projection over all five components, zeros included
// (Source/Reconstruction_PPM_Hydro_HelperModule.py:178-210,
// Source/RiemannSolvers_Hydro_Roe_HelperModule.py:282-308); these bodies do the same,
// so the FP64 instantiation reproduces it bit for bit, signed zeros included.
#pragma once
#include "pmhd_step.h"

PMHD_NS_BEGIN
namespace hydro {

constexpr int NV = 5;

// kinetic-energy order of the _x/_y/_z conversions (Source/PrimCons_HelperModule.py:56-190)
PMHD_FI int ke_a(int d) { return 1 + d; }
PMHD_FI int ke_b(int d) { return 1 + (d + 1) % 3; }
PMHD_FI int ke_c(int d) { return 1 + (d + 2) % 3; }

template <class T>
PMHD_FI void prim_to_cons(const T w[NV], T gamma, int d, T u[NV]) {
  const T a = w[ke_a(d)], b = w[ke_b(d)], c = w[ke_c(d)];
  u[0] = w[0];
  u[1] = w[0] * w[1];
  u[2] = w[0] * w[2];
  u[3] = w[0] * w[3];
  u[4] = w[4] / (gamma - T(1)) + T(0.5) * w[0] * (a * a + b * b + c * c);
}

template <class T>
PMHD_FI void cons_to_prim(const T u[NV], T gamma, int d, T w[NV]) {
  const T a = u[ke_a(d)], b = u[ke_b(d)], c = u[ke_c(d)];
  const T inv = T(1) / u[0];
  w[0] = u[0];
  w[1] = u[1] * inv;
  w[2] = u[2] * inv;
  w[3] = u[3] * inv;
  const T e = u[4] - T(0.5) * (a * a + b * b + c * c) * inv;
  w[4] = e * (gamma - T(1));
}

template <class T>
PMHD_FI void floors(T v[NV], T min_dens, T min_pres) {
  v[0] = v[0] < min_dens ? min_dens : v[0];
  v[4] = v[4] < min_pres ? min_pres : v[4];
}

// dense primitive eigen-system of direction d
// (Source/ReconstructionEigenmatrices_Hydro_HelperModule.py:24-280)
template <class T>
struct Eig {
  T lam[NV];
  T L[NV][NV];
  T R[NV][NV];
};

template <class T>
PMHD_FI void eigen(const T w[NV], T gamma, int d, Eig<T>& E) {
  const int n = 1 + d;
  const T c2 = gamma * w[4] / w[0];
  const T c = sqrt_(c2);
#pragma unroll
  for (int a = 0; a < NV; ++a)
#pragma unroll
    for (int b = 0; b < NV; ++b) {
      E.L[a][b] = T(0);
      E.R[a][b] = T(0);
    }
  E.lam[0] = w[n] - c;
  E.lam[1] = w[n];
  E.lam[2] = w[n];
  E.lam[3] = w[n];
  E.lam[4] = w[n] + c;
  E.R[0][0] = T(1);
  E.R[n][0] = -c / w[0];
  E.R[4][0] = c2;
  E.R[0][n] = T(1);
  E.R[0][4] = T(1);
  E.R[n][4] = -E.R[n][0];
  E.R[4][4] = c2;
  E.L[0][n] = T(-0.5) * w[0] / c;
  E.L[0][4] = T(0.5) / c2;
  E.L[n][0] = T(1);
  E.L[n][4] = T(-1) / c2;
  E.L[4][n] = -E.L[0][n];
  E.L[4][4] = E.L[0][4];
#pragma unroll
  for (int k = 1; k <= 3; ++k)
    if (k != n) {
      E.R[k][k] = T(1);
      E.L[k][k] = T(1);
    }
}

template <class T>
PMHD_FI void project(const T L[NV][NV], const T v[NV], T out[NV]) {
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < NV; ++j) acc = acc + L[i][j] * v[j];
    out[i] = acc;
  }
}

// characteristic-limited slope of one cell (Source/Reconstruction_PPM_Hydro_HelperModule.py:166-214)
template <class T>
PMHD_FI void limited_slope(const Eig<T>& E, const T wl[NV], const T wc[NV], const T wr[NV], T ps[NV]) {
  T dc[NV], dl[NV], dr[NV], wg[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    dc[v] = wr[v] - wl[v];
    dl[v] = wc[v] - wl[v];
    dr[v] = wr[v] - wc[v];
    wg[v] = (dl[v] * dr[v] > T(0)) ? (T(2) * dl[v] * dr[v]) / (dl[v] + dr[v]) : T(0);
  }
  T cc[NV], lc[NV], rc[NV], gc[NV];
  project(E.L, dc, cc);
  project(E.L, dl, lc);
  project(E.L, dr, rc);
  project(E.L, wg, gc);
  T s[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const T lim1 = npmin(npabs(lc[v]), npabs(rc[v]));
    const T lim2 = npmin(T(0.5) * npabs(cc[v]), npabs(gc[v]));
    s[v] = (lc[v] * rc[v] > T(0)) ? npsign(cc[v]) * npmin(T(2) * lim1, lim2) : T(0);
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < NV; ++j) acc = acc + s[j] * E.R[i][j];
    ps[i] = acc;
  }
}

// edge values and characteristic tracing of one cell, order 2 (PPM) or 1 (PLM)
// (Source/Reconstruction_PPM_Hydro_HelperModule.py:215-355)
template <class T>
PMHD_FI void trace_cell(const Eig<T>& E, const T wl[NV], const T wc[NV], const T wr[NV], const T sl[NV],
                        const T sc[NV], const T sr[NV], T dtdx, int order, T la[NV], T ra[NV]) {
  T c1[NV], c2[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    T l, r;
    if (order == 2) {
      l = T(0.5) * (wc[v] + wl[v]) - (sc[v] - sl[v]) / T(6);
      r = T(0.5) * (wr[v] + wc[v]) - (sr[v] - sc[v]) / T(6);
      const T dprod = (r - wc[v]) * (wc[v] - l);
      const T idiff = r - l;
      const T adiff = T(6) * (wc[v] - T(0.5) * (l + r));
      if (dprod <= T(0)) {
        l = wc[v];
        r = wc[v];
      } else if (dprod > T(0)) {
        if ((idiff * adiff) > idiff * idiff) {
          l = (T(6) * wc[v] - T(4) * r) / T(2);
        } else if ((idiff * adiff) < (T(-1) * (idiff * idiff))) {
          r = (T(6) * wc[v] - T(4) * l) / T(2);
        }
      }
    } else {
      l = wc[v] - T(0.5) * sc[v];
      r = wc[v] + T(0.5) * sc[v];
    }
    l = npmax(npmin(wc[v], wl[v]), l);
    l = npmin(npmax(wc[v], wl[v]), l);
    r = npmax(npmin(wc[v], wr[v]), r);
    r = npmin(npmax(wc[v], wr[v]), r);
    la[v] = l;
    ra[v] = r;
    c1[v] = r - l;
    c2[v] = order == 2 ? T(6) * (wc[v] - T(0.5) * (l + r)) : T(0);
  }
  const T c3 = T(0.5) * npmax(E.lam[4], T(0)) * dtdx;
  const T c4 = T(-0.5) * npmin(E.lam[0], T(0)) * dtdx;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    la[v] = la[v] + c4 * (c1[v] + (T(1) - T(kFourThirds) * c4) * c2[v]);
    ra[v] = ra[v] - c3 * (c1[v] - (T(1) - T(kFourThirds) * c3) * c2[v]);
  }
  // every wave contributes (times zero when it cannot reach the face), as the
  // reference's masked sums do
  const T hdt = T(0.5) * dtdx;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const T ind = E.lam[i] <= T(0) ? T(1) : T(0);
    const T v1 = hdt * E.lam[0];
    const T v2 = hdt * E.lam[i];
    const T v3 = v1 - v2;
    const T v4 = T(kFourThirds) * (v1 * v1 - v2 * v2);
    T fac = T(0);
#pragma unroll
    for (int j = 0; j < NV; ++j) fac = fac + ind * E.L[i][j] * (v3 * (c1[j] + c2[j]) + v4 * c2[j]);
#pragma unroll
    for (int j = 0; j < NV; ++j) la[j] = la[j] + fac * E.R[j][i];
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const T ind = E.lam[i] >= T(0) ? T(1) : T(0);
    const T v1 = hdt * E.lam[4];
    const T v2 = hdt * E.lam[i];
    const T v3 = v1 - v2;
    const T v4 = T(kFourThirds) * (v1 * v1 - v2 * v2);
    T fac = T(0);
#pragma unroll
    for (int j = 0; j < NV; ++j) fac = fac + ind * E.L[i][j] * (v3 * (c1[j] - c2[j]) + v4 * c2[j]);
#pragma unroll
    for (int j = 0; j < NV; ++j) ra[j] = ra[j] + fac * E.R[j][i];
  }
}

// Roe flux of direction d from the two primitive states
// (Source/RiemannSolvers_Hydro_Roe_HelperModule.py:68-328 x, :342-596 y, :614-869 z).
// Waves 1..3 are the entropy wave (slot 1 + d) and the shear waves of the
// transverse velocities, in the reference's slot order of each direction.
template <class T>
PMHD_FI void roe_flux(const T wl[NV], const T wr[NV], T g, bool efix, T eps, int d, T F[NV]) {
  const int n = 1 + d;
  T ul[NV], ur[NV];
  prim_to_cons(wl, g, d, ul);
  prim_to_cons(wr, g, d, ur);
  const T sl = sqrt_(wl[0]);
  const T sr = sqrt_(wr[0]);
  const T inv = T(1) / (sl + sr);
  T v[4];
  v[0] = T(0);
#pragma unroll
  for (int k = 1; k <= 3; ++k) v[k] = (sl * wl[k] + sr * wr[k]) * inv;
  const T va = v[ke_a(d)], vb = v[ke_b(d)], vc = v[ke_c(d)];
  const T vsq = va * va + vb * vb + vc * vc;
  const T H = ((ul[4] + wl[4]) / sl + (ur[4] + wr[4]) / sr) * inv;
  const T g1 = g - T(1);
  const T cs2 = g1 * npmax(H - T(0.5) * vsq, T(kSmall));
  const T cs = sqrt_(cs2);
  const T vn = v[n];
  T lam[NV] = {vn - cs, vn, vn, vn, vn + cs};
  if (efix) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lam[k] = entropy_fixed(lam[k], eps);
  }
  T R[NV][NV], L[NV][NV];
#pragma unroll
  for (int a = 0; a < NV; ++a)
#pragma unroll
    for (int b = 0; b < NV; ++b) {
      R[a][b] = T(0);
      L[a][b] = T(0);
    }
  const T nf = T(0.5) / cs2;
  const T gn = g1 / cs2;
  R[0][0] = T(1);
  R[0][4] = T(1);
#pragma unroll
  for (int k = 1; k <= 3; ++k) {
    R[k][0] = k == n ? vn - cs : v[k];
    R[k][4] = k == n ? vn + cs : v[k];
  }
  R[4][0] = H - vn * cs;
  R[4][4] = H + vn * cs;
  L[0][0] = nf * (T(0.5) * g1 * vsq + vn * cs);
  L[4][0] = nf * (T(0.5) * g1 * vsq - vn * cs);
#pragma unroll
  for (int k = 1; k <= 3; ++k) {
    if (k == n) {
      L[0][k] = -nf * (g1 * vn + cs);
      L[4][k] = -nf * (g1 * vn - cs);
    } else {
      L[0][k] = -nf * g1 * v[k];
      L[4][k] = L[0][k];
    }
  }
  L[0][4] = nf * g1;
  L[4][4] = L[0][4];
  // middle slots: x (v_y, v_z, entropy), y (entropy, v_z, v_x), z (v_y, entropy, v_x)
  const int mid[3][3] = {{2, 3, 0}, {0, 3, 1}, {2, 0, 1}};
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int slot = s + 1;
    const int kind = mid[d][s];
    if (kind == 0) {
      R[0][slot] = T(1);
      R[1][slot] = v[1];
      R[2][slot] = v[2];
      R[3][slot] = v[3];
      R[4][slot] = T(0.5) * vsq;
      L[slot][0] = T(1) - nf * g1 * vsq;
      L[slot][1] = gn * v[1];
      L[slot][2] = gn * v[2];
      L[slot][3] = gn * v[3];
      L[slot][4] = -gn;
    } else {
      R[kind][slot] = T(1);
      R[4][slot] = v[kind];
      L[slot][0] = -v[kind];
      L[slot][kind] = T(1);
    }
  }
  T FL[NV], FR[NV];
  FL[0] = ul[n];
  FR[0] = ur[n];
#pragma unroll
  for (int k = 1; k <= 3; ++k) {
    FL[k] = k == n ? ul[n] * wl[k] + wl[4] : ul[n] * wl[k];
    FR[k] = k == n ? ur[n] * wr[k] + wr[4] : ur[n] * wr[k];
  }
  FL[4] = (ul[4] + wl[4]) * wl[n];
  FR[4] = (ur[4] + wr[4]) * wr[n];
  if (lam[4] <= T(0)) {
#pragma unroll
    for (int k = 0; k < NV; ++k) F[k] = FR[k];
    return;
  }
  if (lam[0] >= T(0)) {
#pragma unroll
    for (int k = 0; k < NV; ++k) F[k] = FL[k];
    return;
  }
  T du[NV], proj[NV], amp[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) du[k] = ur[k] - ul[k];
  project(L, du, proj);
#pragma unroll
  for (int k = 0; k < NV; ++k) amp[k] = T(0.5) * npabs(lam[k]) * proj[k];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    T f = T(0.5) * (FL[i] + FR[i]);
#pragma unroll
    for (int j = 0; j < NV; ++j) f = f - amp[j] * R[i][j];
    F[i] = f;
  }
}

template <class P, class T>
PMHD_FI void load5(const P& A, int64_t N, int64_t i, T v[NV]) {
#pragma unroll
  for (int r = 0; r < NV; ++r) v[r] = A[r * N + i];
}
template <class P, class T>
PMHD_FI void store5(const P& A, int64_t N, int64_t i, const T v[NV]) {
#pragma unroll
  for (int r = 0; r < NV; ++r) A[r * N + i] = v[r];
}

// ---------------------------------------------------------------------------
// kernel bodies
// ---------------------------------------------------------------------------
template <class T>
PMHD_FI void k_slopes(const Ctx<T>& c, int d, int bc, I3 q) {
  const int64_t N = c.ncell();
  const int64_t idx = (q.j * c.snx(d) + q.i) * c.snz(d) + q.k;
  const int64_t n = c.n_along(d);
  const int64_t e = co(q, d) - 1;
  const I3 pc = with(q, d, map2(e, n, bc));
  const I3 pl = with(q, d, map2(e - 1, n, bc));
  const I3 pr = with(q, d, map2(e + 1, n, bc));
  T wl[NV], wc[NV], wr[NV];
  load5(c.W, N, c.cid(pl.j, pl.i, pl.k), wl);
  load5(c.W, N, c.cid(pc.j, pc.i, pc.k), wc);
  load5(c.W, N, c.cid(pr.j, pr.i, pr.k), wr);
  Eig<T> E;
  eigen(wc, c.gamma, d, E);
  T ps[NV];
  limited_slope(E, wl, wc, wr, ps);
  store5(c.S, c.nslope(d), idx, ps);
}

template <class T>
PMHD_FI void k_edges(const Ctx<T>& c, int d, int bc, I3 p) {
  const int64_t N = c.ncell();
  const int64_t idx = c.cid(p.j, p.i, p.k);
  const int64_t n = c.n_along(d);
  const int64_t e = co(p, d);
  const I3 pl = with(p, d, map2(e - 1, n, bc));
  const I3 pr = with(p, d, map2(e + 1, n, bc));
  T wl[NV], wc[NV], wr[NV];
  load5(c.W, N, c.cid(pl.j, pl.i, pl.k), wl);
  load5(c.W, N, idx, wc);
  load5(c.W, N, c.cid(pr.j, pr.i, pr.k), wr);
  Eig<T> E;
  eigen(wc, c.gamma, d, E);
  const int64_t NS = c.nslope(d);
  auto sidx = [&](int64_t ee) {
    const I3 s = with(p, d, ee + 1);
    return (s.j * c.snx(d) + s.i) * c.snz(d) + s.k;
  };
  T sl[NV], sc[NV], sr[NV];
  load5(c.S, NS, sidx(e - 1), sl);
  load5(c.S, NS, sidx(e), sc);
  load5(c.S, NS, sidx(e + 1), sr);
  T la[NV], ra[NV];
  trace_cell(E, wl, wc, wr, sl, sc, sr, c.dtd[d], c.order, la, ra);
  const int64_t NF = c.nface(d);
  const I3 f0 = with(p, d, e);
  const I3 f1 = with(p, d, e + 1);
  store5(c.WR, NF, c.fid(d, f0.j, f0.i, f0.k), la);
  store5(c.WL, NF, c.fid(d, f1.j, f1.i, f1.k), ra);
  // edge faces: outflow keeps the edge cell's own traced values
  // (Source/Reconstruction_PPM_Hydro_HelperModule.py:317-330); periodic wraps
  if (e == 0) {
    if (bc == OUTFLOW) {
      store5(c.WL, NF, c.fid(d, f0.j, f0.i, f0.k), la);
    } else {
      const I3 fn = with(p, d, n);
      store5(c.WR, NF, c.fid(d, fn.j, fn.i, fn.k), la);
    }
  }
  if (e == n - 1) {
    if (bc == OUTFLOW) {
      store5(c.WR, NF, c.fid(d, f1.j, f1.i, f1.k), ra);
    } else {
      const I3 fz = with(p, d, 0);
      store5(c.WL, NF, c.fid(d, fz.j, fz.i, fz.k), ra);
    }
  }
}

template <class T>
PMHD_FI void k_copy_states(const Ctx<T>& c, int d, int bc, I3 f) {
  const int64_t N = c.ncell();
  const int64_t idx = c.fid(d, f.j, f.i, f.k);
  const int64_t n = c.n_along(d);
  const int64_t e = co(f, d);
  const I3 pl = with(f, d, map1(e - 1, n, bc));
  const I3 pr = with(f, d, map1(e, n, bc));
  T wl[NV], wr[NV];
  load5(c.W, N, c.cid(pl.j, pl.i, pl.k), wl);
  load5(c.W, N, c.cid(pr.j, pr.i, pr.k), wr);
  const int64_t NF = c.nface(d);
  store5(c.WL, NF, idx, wl);
  store5(c.WR, NF, idx, wr);
}

// floors of a predictor state, counting the floored values at `site` (the call
// sites of Source/EvolveFunctions_Hydro.py:286-288, 437-443, 675-685)
template <class T>
PMHD_FI void floors_counted5(const Ctx<T>& c, int site, T v[NV]) {
  const bool lo0 = v[0] < c.min_dens;
  const bool lo1 = v[4] < c.min_pres;
  floors(v, c.min_dens, c.min_pres);
  if (c.fcnt && (lo0 || lo1)) {
    if (lo0) count_add(c.fcnt + FLOOR_CUR + 2 * site, 1ull);
    if (lo1) count_add(c.fcnt + FLOOR_CUR + 2 * site + 1, 1ull);
  }
}

// floors, conserved interface states (kept for the transverse update), Roe flux
template <class T>
PMHD_FI void k_predict(const Ctx<T>& c, int d, I3 f) {
  const int64_t NF = c.nface(d);
  const int64_t idx = c.fid(d, f.j, f.i, f.k);
  T wl[NV], wr[NV];
  load5(c.WL, NF, idx, wl);
  load5(c.WR, NF, idx, wr);
  floors_counted5(c, 2 * d, wl);
  floors_counted5(c, 2 * d + 1, wr);
  if (c.ndim >= 2) {
    T ul[NV], ur[NV];
    prim_to_cons(wl, c.gamma, d, ul);
    prim_to_cons(wr, c.gamma, d, ur);
    store5(c.UL[d], NF, idx, ul);
    store5(c.UR[d], NF, idx, ur);
  }
  T F[NV];
  roe_flux(wl, wr, c.gamma, c.efix != 0, c.eps, d, F);
  store5(c.F[d], NF, idx, F);
}

// CTU: half-step transverse flux differences of the cells on each side, in the
// order x, y, z (Source/EvolveFunctions_Hydro.py:464-483, :724-766), then the
// corrector flux
template <class T>
PMHD_FI void k_correct(const Ctx<T>& c, int d, I3 f) {
  const int64_t NF = c.nface(d);
  const int64_t idx = c.fid(d, f.j, f.i, f.k);
  const int64_t n = c.n_along(d);
  const int bc = c.bc[d];
  const int64_t e = co(f, d);
  T ul[NV], ur[NV];
  load5(c.UL[d], NF, idx, ul);
  load5(c.UR[d], NF, idx, ur);
  const I3 pl = with(f, d, map1(e - 1, n, bc));
  const I3 pr = with(f, d, map1(e, n, bc));
  for (int t = 0; t < c.ndim; ++t) {
    if (t == d) continue;
    const int64_t NT = c.nface(t);
    const T h = c.hdtd[t];
    const I3 lh = with(pl, t, co(pl, t) + 1);
    const I3 rh = with(pr, t, co(pr, t) + 1);
    const int64_t l0 = c.fid(t, pl.j, pl.i, pl.k), l1 = c.fid(t, lh.j, lh.i, lh.k);
    const int64_t r0 = c.fid(t, pr.j, pr.i, pr.k), r1 = c.fid(t, rh.j, rh.i, rh.k);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      ul[v] = ul[v] - h * (c.F[t][v * NT + l1] - c.F[t][v * NT + l0]);
      ur[v] = ur[v] - h * (c.F[t][v * NT + r1] - c.F[t][v * NT + r0]);
    }
  }
  T wl[NV], wr[NV], F[NV];
  cons_to_prim(ul, c.gamma, d, wl);
  cons_to_prim(ur, c.gamma, d, wr);
  roe_flux(wl, wr, c.gamma, c.efix != 0, c.eps, d, F);
  store5(c.Fc[d], NF, idx, F);
}

template <class T>
PMHD_FI void k_update(const Ctx<T>& c, I3 p) {
  const int64_t N = c.ncell();
  const int64_t idx = c.cid(p.j, p.i, p.k);
  T u[NV];
  load5(c.U, N, idx, u);
  for (int d = 0; d < c.ndim; ++d) {
    const auto Fd = c.ndim >= 2 ? c.Fc[d] : c.F[d];
    const int64_t NF = c.nface(d);
    const I3 hi = with(p, d, co(p, d) + 1);
    const int64_t f0 = c.fid(d, p.j, p.i, p.k), f1 = c.fid(d, hi.j, hi.i, hi.k);
#pragma unroll
    for (int v = 0; v < NV; ++v) u[v] = u[v] - c.dtd[d] * (Fd[v * NF + f1] - Fd[v * NF + f0]);
  }
  T w[NV];
  cons_to_prim(u, c.gamma, 0, w);
  store5(c.U2, N, idx, u);
  store5(c.W2, N, idx, w);
}

// ---------------------------------------------------------------------------
// phases (launch ids shared with the MHD step; a handle runs one physics)
// ---------------------------------------------------------------------------
template <int D>
struct Slopes {
  PMHD_PHASE_TRAITS(K_SLOPES + D, EXT_SLOPE, D)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { hydro::k_slopes(c, D, c.bc[D], p); }
};
template <int D>
struct Edges {
  PMHD_PHASE_TRAITS(K_EDGES + D, EXT_CELL, D)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { hydro::k_edges(c, D, c.bc[D], p); }
};
template <int D>
struct CopyStates {
  PMHD_PHASE_TRAITS(K_COPY + D, EXT_FACE, D)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { hydro::k_copy_states(c, D, c.bc[D], p); }
};
template <int D>
struct Predict {
  PMHD_PHASE_TRAITS(K_PREDICT + D, EXT_FACE, D)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { hydro::k_predict(c, D, p); }
};
template <int D>
struct Correct {
  PMHD_PHASE_TRAITS(K_CORRECT + D, EXT_FACE, D)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p) const { hydro::k_correct(c, D, p); }
};
struct Update {  // update, then the CFL speeds of the new state
  PMHD_PHASE_TRAITS(K_UPDATE, EXT_CELL, 0)
  template <class T>
  PMHD_FI void operator()(const Ctx<T>& c, I3 p, double o[3]) const {
    hydro::k_update(c, p);
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

template <int D, class T, class X>
void predictor_sweep(X& ex, const Ctx<T>& c) {
  if (c.order > 0) {
    ex.template run<Slopes<D>>(c);
    ex.template run<Edges<D>>(c);
  } else {
    ex.template run<CopyStates<D>>(c);
  }
  ex.template run<Predict<D>>(c);
}

// one hydro cycle (Source/EvolveFunctions_Hydro.py:260-304 1D, :390-515 2D, :608-805 3D)
template <class T, class X>
void cycle(X& ex, const Ctx<T>& c) {
  const int nd = c.ndim;
  hydro::predictor_sweep<0>(ex, c);
  if (nd >= 2) hydro::predictor_sweep<1>(ex, c);
  if (nd == 3) hydro::predictor_sweep<2>(ex, c);
  if (nd >= 2) {
    ex.template run<Correct<0>>(c);
    ex.template run<Correct<1>>(c);
    if (nd == 3) ex.template run<Correct<2>>(c);
  }
  ex.template run_max<Update>(c);
}

}  // namespace hydro

// the cycle of a context's physics
template <class T, class X>
void run_cycle(X& ex, const Ctx<T>& c) {
  if (c.physics == 1) hydro::cycle(ex, c);
  else cycle(ex, c);
}

PMHD_NS_END
