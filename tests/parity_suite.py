"""Parity suite of the step library against the golden vectors written from the
reference and against the FP64 oracle.  FP64 results are compared bit for bit (the
kernels keep the reference's operation order, including its 80-bit x-direction
CTU factor); FP32 results against the FP64 oracle within stated tolerances.

The test modules that import this suite provide the ``lib`` fixture: the HIP
library on the GPU (test_gpu_parity.py) or the host build of the same kernel
bodies (test_host_parity.py)."""
import numpy as np
import pytest

from oracle import mhd as M
from tests.parity_cases import (GOLD_2D, GOLD_3D, GAMMA, div_b, golden, oracle_run, run_solver,
                                smooth_state, widths_of)

def Solver():
    from pythonmhd_amd._lib import Solver as S
    return S


def same(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape
    assert np.array_equal(a, b), float(np.max(np.abs(a - b)))


@pytest.mark.parametrize('order', [0, 2])
def test_golden_1d_briowu(lib, order):
    e = golden()
    pin = e['bw_o%d_in' % order]
    print('case: 1D Brio-Wu order', order, flush=True)
    r = run_solver(Solver(), lib, (64,), (1 / 64,), (0,), pin, M.prim_to_cons_all(pin, 2.0), (),
                   0.1, 25, gamma=2.0, cfl=0.8, order=order)
    for i, n in ((0, 'prim'), (1, 'cons'), (3, 't'), (4, 'cyc'), (5, 'prim2'), (6, 'cons2'), (8, 't2')):
        same(r[i], e['bw_o%d_%s' % (order, n)])


@pytest.mark.parametrize('tag,shape,bc,order,ncyc', GOLD_2D)
def test_golden_2d(lib, tag, shape, bc, order, ncyc):
    e = golden()
    ny, nx = shape
    print('case: 2D', tag, flush=True)
    r = run_solver(Solver(), lib, shape, (1 / nx, 1 / ny), (bc, bc), e[tag + '_in_prim'], e[tag + '_in_cons'],
                   (e[tag + '_in_fbx'], e[tag + '_in_fby']), 10.0, ncyc, cfl=0.8, order=order)
    same(r[0], e[tag + '_prim'])
    same(r[1], e[tag + '_cons'])
    same(r[2][0], e[tag + '_fbx'])
    same(r[2][1], e[tag + '_fby'])
    same(r[3], e[tag + '_t'])
    same(r[5], e[tag + '_prim2'])
    same(r[7][1], e[tag + '_fby2'])


@pytest.mark.parametrize('tag,shape,bc,order,ncyc', GOLD_3D)
def test_golden_3d(lib, tag, shape, bc, order, ncyc):
    e = golden()
    print('case: 3D', tag, flush=True)
    r = run_solver(Solver(), lib, shape, (1 / shape[1], 1 / shape[0], 1 / shape[2]), (bc, bc, bc),
                   e[tag + '_in_prim'], e[tag + '_in_cons'],
                   (e[tag + '_in_fbx'], e[tag + '_in_fby'], e[tag + '_in_fbz']), 10.0, ncyc,
                   cfl=0.4, order=order, entropy_fix=True)
    same(r[0], e[tag + '_prim'])
    same(r[1], e[tag + '_cons'])
    for i, n in enumerate(('fbx', 'fby', 'fbz')):
        same(r[2][i], e[tag + '_' + n])
    same(r[3], e[tag + '_t'])
    same(r[6], e[tag + '_cons2'])


@pytest.mark.parametrize('nd,order', [(1, 1), (2, 1), (3, 1), (2, 2), (3, 2)])
def test_oracle_fp64_larger(lib, nd, order):
    """Larger periodic grids (PLM and PPM) against the oracle, bit for bit."""
    rng = np.random.default_rng(100 + 10 * nd + order)
    shape = {1: (96,), 2: (40, 48), 3: (20, 24, 28)}[nd]
    prim, faces = smooth_state(rng, shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    widths = widths_of(shape)
    cfl = 0.4 if nd == 3 else 0.8
    P = M.Params(GAMMA, cfl, order=order, entropy_fix=(nd == 3))
    ncyc = 4
    print('case: fp64 vs oracle nd=%d order=%d shape=%s' % (nd, order, shape), flush=True)
    ref = oracle_run(shape, widths, (1,) * nd, prim, cons, faces, 10.0, ncyc, P)
    got = run_solver(Solver(), lib, shape, widths, (1,) * nd, prim, cons, faces, 10.0, ncyc, cfl=cfl,
                     order=order, entropy_fix=(nd == 3))
    same(got[0], ref[0])
    same(got[1], ref[1])
    for a, b in zip(got[2], ref[2]):
        same(a, b)


@pytest.mark.parametrize('nd', [2, 3])
def test_fp32_against_fp64_oracle(lib, nd):
    """FP32 bulk with FP64 islands: density L1 error against the FP64 oracle and
    the dimensionless div B of the faces.  Tolerances: L1(rho) < 1e-5 (relative
    to mean rho), div B < 1e-6."""
    rng = np.random.default_rng(7 + nd)
    shape = {2: (48, 48), 3: (24, 24, 24)}[nd]
    prim, faces = smooth_state(rng, shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    widths = widths_of(shape)
    cfl = 0.4 if nd == 3 else 0.8
    P = M.Params(GAMMA, cfl, order=1, entropy_fix=False)
    ncyc = 10
    print('case: fp32 nd=%d shape=%s' % (nd, shape), flush=True)
    ref = oracle_run(shape, widths, (1,) * nd, prim, cons, faces, 10.0, ncyc, P)
    got = run_solver(Solver(), lib, shape, widths, (1,) * nd, prim, cons, faces, 10.0, ncyc, cfl=cfl,
                     order=1, precision=32)
    l1 = float(np.mean(np.abs(got[1][0] - ref[1][0])) / np.mean(ref[1][0]))
    print('fp32 L1(rho) = %.3e, divB = %.3e' % (l1, div_b(got[2], widths)))
    assert l1 < 1e-5
    assert div_b(got[2], widths) < 1e-6


def test_conservation_and_divb_fp64(lib):
    """Periodic 3D PPM: total mass, momentum and energy conserved to roundoff; the
    face field stays divergence-free to roundoff."""
    rng = np.random.default_rng(3)
    shape = (16, 16, 16)
    prim, faces = smooth_state(rng, shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    widths = (1 / 16, 1 / 16, 1 / 16)
    d0 = div_b(faces, widths)
    got = run_solver(Solver(), lib, shape, widths, (1, 1, 1), prim, cons, faces, 10.0, 6, cfl=0.4, order=2,
                     entropy_fix=True)
    for v in (0, 1, 2, 3, 7):
        tot0 = np.sum(cons[v])
        tot1 = np.sum(got[1][v])
        assert abs(tot1 - tot0) <= 1e-12 * max(1.0, np.sum(np.abs(cons[v]))), (v, tot0, tot1)
    assert div_b(got[2], widths) < 1e-13 + d0


@pytest.mark.parametrize('precision', [64, 32])
def test_diagnostics(lib, precision):
    """pmhd_diagnostics: FP64 conserved totals and the dimensionless max |div B|
    against NumPy on the downloaded state."""
    from pythonmhd_amd._lib import Solver as S
    rng = np.random.default_rng(11)
    shape = (12, 14, 10)
    prim, faces = smooth_state(rng, shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    w = widths_of(shape)
    s = S(shape, w, (1, 1, 1), GAMMA, 0.4, order=1, precision=precision, lib=lib)
    s.set_state(prim, cons, faces)
    s.advance(0.0, 10.0, 0, 3)
    d = s.diagnostics()
    p, c, f = s.get_state()
    s.close()
    tot = np.array([np.sum(c[v]) for v in range(8)])
    assert np.allclose(d['totals'], tot, rtol=1e-13, atol=1e-12)
    assert d['nonfinite'] == 0
    assert abs(d['max_divb'] - div_b(f, w)) <= 1e-12 * max(1.0, div_b(f, w))
    assert d['max_b'] == max(np.max(np.abs(a)) for a in f)
