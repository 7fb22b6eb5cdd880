// One CTU+CT cycle as a sequence of launches over an executor X that provides
//   X::run(id, n, f)        -- f(i) for i in [0, n); id names the launch
//   X::max3(id, n, f, out)  -- out[d] = max over i of f(i)[d]
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

template <class T, class X>
void cycle(X& ex, const Ctx<T>& c) {
  const int nd = c.ndim;
  for (int d = 0; d < nd; ++d) {
    const int bc = c.bc[d];
    if (c.order > 0) {
      ex.run(K_SLOPES + d, c.nslope(d), [=] PMHD_HD(int64_t i) { k_slopes(c, d, bc, i); });
      ex.run(K_EDGES + d, c.ncell(), [=] PMHD_HD(int64_t i) { k_edges(c, d, bc, i); });
    } else {
      // the reference's 3D first-order z states use the x boundary condition
      const int bcq = d == 2 ? c.bc_zrec : bc;
      ex.run(K_COPY + d, c.nface(d), [=] PMHD_HD(int64_t i) { k_copy_states(c, d, bcq, i); });
    }
    ex.run(K_PREDICT + d, c.nface(d), [=] PMHD_HD(int64_t i) { k_predict(c, d, i); });
  }
  if (nd == 1) {
    ex.run(K_UPDATE, c.ncell(), [=] PMHD_HD(int64_t i) { k_update(c, i); });
    return;
  }
  ex.run(K_CELL_E, c.ncell(), [=] PMHD_HD(int64_t i) { k_cell_e(c, i); });
  ex.run(K_EMF, c.nedge(), [=] PMHD_HD(int64_t i) { k_emf(c, i); });
  for (int d = 0; d < nd; ++d) ex.run(K_CTU + d, c.nface(d), [=] PMHD_HD(int64_t i) { k_ctu(c, d, i); });
  for (int d = 0; d < nd; ++d) {
    ex.run(K_CT_HALF + d, c.nface(d), [=] PMHD_HD(int64_t i) {
      T h[3] = {c.hdtd[0], c.hdtd[1], c.hdtd[2]};
      k_ct(c, d, c.fb[d], c.nfb[d], h, i);
    });
  }
  ex.run(K_HALF_E, c.ncell(), [=] PMHD_HD(int64_t i) { k_half_e(c, i); });
  for (int d = 0; d < nd; ++d) ex.run(K_CORRECT + d, c.nface(d), [=] PMHD_HD(int64_t i) { k_correct(c, d, i); });
  ex.run(K_EMF, c.nedge(), [=] PMHD_HD(int64_t i) { k_emf(c, i); });
  for (int d = 0; d < nd; ++d) {
    ex.run(K_CT_FULL + d, c.nface(d), [=] PMHD_HD(int64_t i) {
      T h[3] = {c.dtd[0], c.dtd[1], c.dtd[2]};
      k_ct(c, d, c.fb[d], c.fb2[d], h, i);
    });
  }
  ex.run(K_UPDATE, c.ncell(), [=] PMHD_HD(int64_t i) { k_update(c, i); });
}

}  // namespace pmhd
