"""The adjoint's hand-written reverses against the per-slot dual-number evaluation
of the same elements (CPU, host builds).

The host library differentiates each corrector face as 15 dual passes through
the flux followed by the hand-written reverse of the transverse correction
(ctu_vjp in pythonmhd_amd/csrc/pmhd_adjoint.h), each predictor face as 15 dual
passes through prim->cons and the flux followed by the hand-written reverse of its
preparation (floors and reconstruction source terms, adjoint_predict), each update
cell through its hand-written reverse (adjoint_update), and forms the eigen-system
of passes that seed none of its inputs on plain values (pmhd_dual.h).
libpmhd_host_dualc.so is the same library built with every phase through one full
dual pass per input slot (no hand-written derivative or shortcut anywhere).

The two give the same gradient of a whole checkpointed run to FP64 rounding, on 2D
and 3D grids, periodic, outflow and mixed boundaries, PLM and first order, with and
without the reference's boundary quirks, and with the density, pressure and energy
floors acting on part of the grid (the reverses then take the floors' branches).
In FP32 the dual path carries its derivatives in FP32 and the reverses in FP64, so
they agree to FP32 rounding of the derivatives.
"""
import os

import numpy as np
import pytest

import oracle.mhd as M
from tests.parity_suite import GAMMA, _generic_state, _kinetic_cotangent, widths_of

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, 'hostbuild', 'libpmhd_host.so')
DUAL_LIB = os.path.join(HERE, 'hostbuild', 'libpmhd_host_dualc.so')


@pytest.fixture(scope='module')
def libs():
    if not (os.path.exists(HOST_LIB) and os.path.exists(DUAL_LIB)):
        from pythonmhd_amd import build
        build.build_host()
    from pythonmhd_amd._lib import load
    return load(HOST_LIB), load(DUAL_LIB)


# (ndim, boundaries, order, fix_quirks, precision, min_dens, min_energy, min_pres)
CASES = [
    (2, (1, 1), 2, 0, 64, 1e-12, 1e-12, 1e-12),
    (2, (0, 0), 1, 0, 64, 0.95, 1e-12, 1e-12),
    (2, (0, 0), 2, 7, 64, 1e-12, 1.6, 1e-12),
    (2, (1, 0), 2, 0, 32, 1e-12, 1e-12, 1e-12),
    (2, (1, 1), 1, 0, 64, 1e-12, 1e-12, 1.0),
    (3, (1, 1, 1), 2, 0, 64, 1e-12, 1e-12, 1e-12),
    (3, (0, 0, 0), 2, 0, 64, 0.97, 1.7, 1e-12),
    (3, (1, 0, 1), 1, 7, 64, 1e-12, 1e-12, 1e-12),
    (3, (0, 1, 0), 2, 7, 32, 1e-12, 1e-12, 1e-12),
    (3, (1, 1, 1), 1, 0, 64, 1e-12, 1e-12, 1.0),
]


@pytest.mark.parametrize('nd,bcs,order,fq,prec,md,me,mp', CASES)
def test_corrector_reverse_matches_dual_passes(libs, nd, bcs, order, fq, prec, md, me, mp):
    from pythonmhd_amd.adjoint import AdjointRun
    shape = (10, 12) if nd == 2 else (5, 6, 4)
    rng = np.random.default_rng(11 + nd + order + fq)
    prim, faces = _generic_state(rng, shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    gp0 = rng.standard_normal(prim.shape)
    gf0 = [rng.standard_normal(f.shape) for f in faces]
    grads = []
    for lib in libs:
        run = AdjointRun(shape, widths_of(shape), bcs, GAMMA, order=order, precision=prec, checkpoint_every=2,
                         lib=lib, fix_quirks=fq, min_dens=md, min_energy=me, min_pres=mp)
        _, co, _ = run.forward(prim, cons, faces, 0.004, 3)
        gp, gc, gf = run.backward(gp0, _kinetic_cotangent(co), gf0)
        run.close()
        grads.append([gp, gc] + list(gf))
    tol = 1e-13 if prec == 64 else 3e-8
    for a, b in zip(*grads):
        scale = float(np.max(np.abs(b)))
        assert scale > 0
        assert np.max(np.abs(a - b)) <= tol * scale, (float(np.max(np.abs(a - b))), scale)
