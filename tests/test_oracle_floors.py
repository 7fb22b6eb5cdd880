"""The oracle against the reference's floor-firing golden trajectories (CPU).

tests/golden/floors.npz holds evolve trajectories of the reference in which the
density, pressure and energy floors bite (Source/PrimCons_HelperModule.py:499-554):
the 1D and 3D predictor floors, the 2D predictor floors and the 2D post-CTU density
and energy floors (Source/EvolveFunctions_MHD.py:278-280, 513-519, 562-576,
911-922), and the hydro predictor floors (Source/EvolveFunctions_Hydro.py:286-288,
437-443, 675-685).  The fixture also counts the values each applyFloors call
floored; a mistake such as the pressure floor in place of the energy floor after
the 2D CTU update, ``<=`` for ``<``, or a floor in the 3D corrector changes the
trajectories, which are compared bit for bit.
"""
import os

import numpy as np
import pytest

from oracle import hydro as H
from oracle import mhd as M

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'floors.npz')

# (tag, ndim, physics, gamma, cfl, order, entropy_fix, bcs, t_end, max_cycles)
MHD_CASES = [
    ('m1o0', 1, 2.0, 0.8, 0, False, (0,), 0.1, 12),
    ('m1o2', 1, 2.0, 0.8, 2, False, (0,), 0.1, 12),
    ('m1p', 1, 5.0 / 3.0, 0.8, 2, True, (1,), 10.0, 8),
    ('m2o2', 2, 5.0 / 3.0, 0.8, 2, False, (1, 1), 10.0, 4),
    ('m2o0', 2, 5.0 / 3.0, 0.8, 0, False, (0, 0), 10.0, 4),
    ('m3o2', 3, 5.0 / 3.0, 0.4, 2, True, (1, 1, 1), 10.0, 2),
    ('m3o0', 3, 5.0 / 3.0, 0.4, 0, True, (1, 1, 1), 10.0, 2),
    ('m3o1', 3, 5.0 / 3.0, 0.4, 0, True, (1, 1, 1), 10.0, 1),
]
HYDRO_CASES = [
    ('h1o0', 1.4, 0.8, 0, False, (0,), 0.2, 20),
    ('h1o2', 1.4, 0.8, 2, True, (0,), 0.2, 20),
    ('h2o2', 5.0 / 3.0, 0.8, 2, False, (1, 1), 10.0, 5),
    ('h3o2', 5.0 / 3.0, 0.4, 2, True, (1, 1, 1), 10.0, 2),
]


def golden():
    return np.load(GOLD)


def widths_of(shape):
    if len(shape) == 1:
        return (1.0 / shape[0],)
    if len(shape) == 2:
        return (1.0 / shape[1], 1.0 / shape[0])
    return (1.0 / shape[1], 1.0 / shape[0], 1.0 / shape[2])


def mhd_inputs(d, tag, nd):
    """prim, cons, faces of an MHD case (1D: cons from the reference's primToCons)."""
    if nd == 1:
        prim = d[tag + '_in']
        return prim, None, ()
    prim, cons = d[tag + '_in_prim'], d[tag + '_in_cons']
    names = ('fbx', 'fby', 'fbz')[:nd]
    return prim, cons, tuple(d[tag + '_in_' + n] for n in names)


def mhd_expected(d, tag, nd):
    """The golden evolve tuple in the oracle's / solver's order:
    (prim, cons, faces, t, cyc, prim2, cons2, faces2, t2)."""
    g = lambda n: d[tag + '_' + n]  # noqa: E731
    fn = ('fbx', 'fby', 'fbz')[:nd] if nd >= 2 else ()
    return (g('prim'), g('cons'), [g(n) for n in fn], g('t'), g('cyc'), g('prim2'), g('cons2'),
            [g(n + '2') for n in fn], g('t2'))


def oracle_mhd(prim, cons, faces, nd, gamma, cfl, order, efix, bcs, t_end, ncyc, floors):
    P = M.Params(gamma, cfl, floors[0], floors[1], floors[2], order=order, entropy_fix=efix)
    shape = prim.shape[1:]
    w = widths_of(shape)
    if nd == 1:
        r = M.evolve_1d(prim, cons, w[0], bcs[0], 0.0, t_end, 0, ncyc, P)
        return r[0], r[1], [], r[2], r[3], r[4], r[5], [], r[6]
    if nd == 2:
        r = M.evolve_2d(prim, cons, faces[0], faces[1], w[0], w[1], bcs[0], bcs[1], 0.0, t_end, 0, ncyc, P)
        return r[0], r[1], [r[2], r[3]], r[4], r[5], r[6], r[7], [r[8], r[9]], r[10]
    r = M.evolve_3d(prim, cons, *faces, w[0], w[1], w[2], *bcs, 0.0, t_end, 0, ncyc, P)
    return r[0], r[1], [r[2], r[3], r[4]], r[5], r[6], r[7], r[8], [r[9], r[10], r[11]], r[12]


def _same(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape
    assert np.array_equal(a, b), float(np.max(np.abs(a - b)))


def _cons1d(prim, gamma):
    return M.prim_to_cons_all(prim, gamma)


@pytest.mark.parametrize('tag,nd,gamma,cfl,order,efix,bcs,t_end,ncyc', MHD_CASES)
def test_oracle_mhd_floors(tag, nd, gamma, cfl, order, efix, bcs, t_end, ncyc):
    d = golden()
    prim, cons, faces = mhd_inputs(d, tag, nd)
    if cons is None:
        cons = _cons1d(prim, gamma)
    floors = tuple(float(x) for x in d[tag + '_floors'])
    got = oracle_mhd(prim, cons, faces, nd, gamma, cfl, order, efix, bcs, t_end, ncyc, floors)
    want = mhd_expected(d, tag, nd)
    for i in (0, 1, 3, 4, 5, 6, 8):
        _same(got[i], want[i])
    for a, b in zip(got[2], want[2]):
        _same(a, b)
    for a, b in zip(got[7], want[7]):
        _same(a, b)


@pytest.mark.parametrize('tag,gamma,cfl,order,efix,bcs,t_end,ncyc', HYDRO_CASES)
def test_oracle_hydro_floors(tag, gamma, cfl, order, efix, bcs, t_end, ncyc):
    d = golden()
    prim = d[tag + '_in']
    floors = tuple(float(x) for x in d[tag + '_floors'])
    P = H.Params(gamma, cfl, floors[0], floors[1], floors[2], order=order, entropy_fix=efix)
    r = H.evolve(prim.copy(), H.prim_to_cons(prim, gamma, 0), widths_of(prim.shape[1:]), bcs, 0.0, t_end, 0,
                 ncyc, P)
    for a, k in zip(r, ('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')):
        _same(a, d[tag + '_' + k])


def test_fixture_floors_fire_everywhere_they_can():
    """Every floor call site of the reference fires in some trajectory: the
    predictor density and pressure floors in 1D, 2D and 3D, the 2D post-CTU density
    and energy floors; 3D calls never floor more than two values of one variable."""
    d = golden()
    c1 = d['m1o2_counts'].sum(axis=0)
    assert c1[0].all() or c1[1].all()
    c2 = sum(d[t + '_counts'].sum(axis=0) for t in ('m2o2', 'm2o0'))
    assert (c2[0:4] > 0).any(axis=0).all()          # predictor: density and pressure
    assert (c2[6:10, 1] > 0).all()                    # post-CTU energy floor at all four sites
    assert (c2[6:10, 0] > 0).any()                    # post-CTU density floor
    c3 = sum(d[t + '_counts'].sum(axis=0) for t in ('m3o2', 'm3o0', 'm3o1'))
    assert (c3[0:6] > 0).all()
    assert max(d[t + '_counts'].max() for t in ('m3o2', 'm3o0', 'm3o1', 'h3o2')) <= 2


def test_floor_semantics_of_the_oracle():
    """apply_floors: strict comparison (a value equal to the floor stays), density on
    the first and pressure or energy on the last variable, NaN passes through."""
    v = np.array([[0.1, 0.2, 0.3, np.nan], [1, 2, 3, 4], [0.5, 0.05, 0.2, 0.1]], dtype=np.float64)
    out = M.apply_floors(v, True, 0.2, 0.1, 7.0)
    assert np.array_equal(out[0][:3], [0.2, 0.2, 0.3]) and np.isnan(out[0][3])
    assert np.array_equal(out[1], v[1])
    assert np.array_equal(out[2], [0.5, 0.1, 0.2, 0.1])
    out = M.apply_floors(v, False, 0.2, 0.1, 0.3)
    assert np.array_equal(out[2], [0.5, 0.3, 0.3, 0.3])
