// One CTU+CT cycle as a sequence of launches over an executor X that provides
//   X::run3(id, n0, n1, n2, f)    -- f(I3{j, i, k}) over a (n0, n1, n2) index box
//   X::max3(id, n, f, out)        -- out[d] = max over i in [0, n) of f(i)[d]
//   X::reduce(id, n, f, sums, maxes)
// The direction of each sweep is a template parameter so the rotations and index
// arithmetic of the kernel bodies fold at compile time.
// (Source/EvolveFunctions_MHD.py:252-317 for 1D, :464-692 for 2D, :840-1152 for 3D)
#pragma once
#include "pmhd_kernels.h"

namespace pmhd {

template <class T, class X>
void cfl_maxima(X& ex, const Ctx<T>& c, double out[3]) {
  ex.max3(K_CFL, c.ncell(), [=] PMHD_HD(int64_t i, double o[3]) {
    T t[3];
    k_cfl(c, i, t);
    o[0] = double(t[0]);
    o[1] = double(t[1]);
    o[2] = double(t[2]);
  }, out);
}

// FP64 diagnostics: sums[0..7] conserved totals, sums[8] non-finite count,
// maxes[0] max |div B| * min width, maxes[1] max |B| on faces
template <class T, class X>
void diagnostics(X& ex, const Ctx<T>& c, double sums[9], double maxes[2]) {
  ex.reduce(K_DIAG, c.ncell(), [=] PMHD_HD(int64_t i, double* s, double* m) { k_diag(c, i, s, m); },
            sums, maxes);
}

// interface states and predictor fluxes of direction D
template <int D, class T, class X>
void predictor_sweep(X& ex, const Ctx<T>& c) {
  const int bc = c.bc[D];
  if (c.order > 0) {
    ex.run3(K_SLOPES + D, c.sny(D), c.snx(D), c.snz(D), [=] PMHD_HD(I3 p) { k_slopes(c, D, bc, p); });
    ex.run3(K_EDGES + D, c.ny, c.nx, c.nz, [=] PMHD_HD(I3 p) { k_edges(c, D, bc, p); });
  } else {
    // the reference's 3D first-order z states use the x boundary condition
    const int bcq = D == 2 ? c.bc_zrec : bc;
    ex.run3(K_COPY + D, c.fny(D), c.fnx(D), c.fnz(D), [=] PMHD_HD(I3 p) { k_copy_states(c, D, bcq, p); });
  }
  ex.run3(K_PREDICT + D, c.fny(D), c.fnx(D), c.fnz(D), [=] PMHD_HD(I3 p) { k_predict(c, D, p); });
}

template <int D, class T, class X>
void face_pass(X& ex, const Ctx<T>& c, int which) {
  const int64_t n0 = c.fny(D), n1 = c.fnx(D), n2 = c.fnz(D);
  if (which == 1) {
    ex.run3(K_CT_HALF + D, n0, n1, n2, [=] PMHD_HD(I3 p) {
      T h[3] = {c.hdtd[0], c.hdtd[1], c.hdtd[2]};
      k_ct(c, D, c.fb[D], c.nfb[D], h, p);
    });
  } else if (which == 2) {
    ex.run3(K_CORRECT + D, n0, n1, n2, [=] PMHD_HD(I3 p) { k_correct(c, D, p); });
  } else {
    ex.run3(K_CT_FULL + D, n0, n1, n2, [=] PMHD_HD(I3 p) {
      T h[3] = {c.dtd[0], c.dtd[1], c.dtd[2]};
      k_ct(c, D, c.fb[D], c.fb2[D], h, p);
    });
  }
}

template <class T, class X>
void face_passes(X& ex, const Ctx<T>& c, int which) {
  face_pass<0>(ex, c, which);
  face_pass<1>(ex, c, which);
  if (c.ndim == 3) face_pass<2>(ex, c, which);
}

template <class T, class X>
void cycle(X& ex, const Ctx<T>& c) {
  const int nd = c.ndim;
  predictor_sweep<0>(ex, c);
  if (nd >= 2) predictor_sweep<1>(ex, c);
  if (nd == 3) predictor_sweep<2>(ex, c);
  if (nd == 1) {
    ex.run3(K_UPDATE, c.ny, c.nx, c.nz, [=] PMHD_HD(I3 p) { k_update(c, p); });
    return;
  }
  ex.run3(K_CELL_E, c.ny, c.nx, c.nz, [=] PMHD_HD(I3 p) { k_cell_e(c, p); });
  Ctx<T> cp = c;           // corner EMFs of the predictor fluxes
  for (int d = 0; d < 3; ++d) cp.Fe[d] = c.F[d];
  ex.run3(K_EMF, c.eny(), c.enx(), c.enz(), [=] PMHD_HD(I3 p) { k_emf(cp, p); });
  face_passes(ex, c, 1);   // half-step face field
  ex.run3(K_HALF_E, c.ny, c.nx, c.nz, [=] PMHD_HD(I3 p) { k_half_e(c, p); });
  face_passes(ex, c, 2);   // CTU transverse corrections + corrector fluxes
  Ctx<T> cc = c;           // corner EMFs of the corrector fluxes
  for (int d = 0; d < 3; ++d) cc.Fe[d] = c.Fc[d];
  ex.run3(K_EMF2, c.eny(), c.enx(), c.enz(), [=] PMHD_HD(I3 p) { k_emf(cc, p); });
  face_passes(ex, c, 3);   // full-step face field
  ex.run3(K_UPDATE, c.ny, c.nx, c.nz, [=] PMHD_HD(I3 p) { k_update(c, p); });
}

}  // namespace pmhd
