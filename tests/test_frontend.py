"""The drop-in evolve functions (reference signatures and return tuples) driven
through the host build of the kernels, against the golden trajectories."""
import os

import numpy as np
import pytest

from oracle import mhd as M
from tests.parity_cases import golden

HOST_LIB = os.path.join(os.path.dirname(__file__), 'hostbuild', 'libpmhd_host.so')


class Grid:
    pass


@pytest.fixture
def evolve(monkeypatch):
    if not os.path.exists(HOST_LIB):
        from pythonmhd_amd import build
        build.build_host()
    from pythonmhd_amd import _lib, evolve as ev
    host = _lib.load(HOST_LIB)

    class HostSolver(_lib.Solver):
        def __init__(self, *a, **k):
            k['lib'] = host
            super().__init__(*a, **k)
    monkeypatch.setattr(ev, 'Solver', HostSolver)
    ev.release_device_state()
    yield ev
    ev.release_device_state()


def test_evolve_1d_tuple(evolve):
    e = golden()
    g = Grid()
    g.primVars = e['bw_o2_in']
    g.consVars = M.prim_to_cons_all(g.primVars, 2.0)
    g.dx, g.BcX = 1 / 64, 0
    r = evolve.evolve_1D_mhd(g, 0.0, 0.1, 0, 25, 2.0, 0.8, 1e-12, 1e-12, 1e-12, 2, False, 0.5)
    assert len(r) == 7
    for i, n in enumerate(('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')):
        assert np.array_equal(np.asarray(r[i]), e['bw_o2_' + n]), n


def test_evolve_2d_tuple(evolve):
    e = golden()
    tag = 'o2o2'
    g = Grid()
    g.primVars, g.consVars = e[tag + '_in_prim'], e[tag + '_in_cons']
    g.faceBx, g.faceBy = e[tag + '_in_fbx'], e[tag + '_in_fby']
    g.dx, g.dy, g.BcX, g.BcY = 1 / 14, 1 / 12, 0, 0
    r = evolve.evolve_2D_mhd(g, 0.0, 10.0, 0, 5, 5 / 3, 0.8, 1e-12, 1e-12, 1e-12, 2, False, 0.5)
    names = ('prim', 'cons', 'fbx', 'fby', 't', 'cyc', 'prim2', 'cons2', 'fbx2', 'fby2', 't2')
    assert len(r) == len(names)
    for i, n in enumerate(names):
        assert np.array_equal(np.asarray(r[i]), e[tag + '_' + n]), n


def test_evolve_3d_tuple(evolve):
    e = golden()
    tag = 'n3o2'
    g = Grid()
    g.primVars, g.consVars = e[tag + '_in_prim'], e[tag + '_in_cons']
    g.faceBx, g.faceBy, g.faceBz = e[tag + '_in_fbx'], e[tag + '_in_fby'], e[tag + '_in_fbz']
    g.dx, g.dy, g.dz = 1 / 8, 1 / 6, 1 / 10
    g.BcX = g.BcY = g.BcZ = 1
    r = evolve.evolve_3D_mhd(g, 0.0, 10.0, 0, 2, 5 / 3, 0.4, 1e-12, 1e-12, 1e-12, 2, True, 0.5)
    names = ('prim', 'cons', 'fbx', 'fby', 'fbz', 't', 'cyc', 'prim2', 'cons2', 'fbx2', 'fby2', 'fbz2', 't2')
    assert len(r) == len(names)
    for i, n in enumerate(names):
        assert np.array_equal(np.asarray(r[i]), e[tag + '_' + n]), n


def test_zero_cycles_returns_input_as_second_last(evolve):
    e = golden()
    g = Grid()
    g.primVars = e['bw_o0_in']
    g.consVars = M.prim_to_cons_all(g.primVars, 2.0)
    g.dx, g.BcX = 1 / 64, 0
    r = evolve.evolve_1D_mhd(g, 0.5, 0.5, 3, 10, 2.0, 0.8, 1e-12, 1e-12, 1e-12, 0)
    assert r[2] == 0.5 and r[3] == 3 and r[6] == 0.5
    assert np.array_equal(r[0], g.primVars) and np.array_equal(r[4], g.primVars)


def test_assertions_match_reference(evolve):
    g = Grid()
    g.primVars = np.ones((8, 16))
    g.consVars = np.ones((8, 16))
    g.dx, g.BcX = 0.1, 2
    with pytest.raises(AssertionError):
        evolve.evolve_1D_mhd(g, 0.0, 1.0, 0, 10, 1.4, 0.8, 1e-12, 1e-12, 1e-12)
    g.BcX = 0
    with pytest.raises(AssertionError):
        evolve.evolve_1D_mhd(g, 1.0, 0.0, 0, 10, 1.4, 0.8, 1e-12, 1e-12, 1e-12)


def test_raise_like_reference(evolve):
    """raiseLikeReference: the reference's exceptions on the inputs where it raises
    (3D floors with three floored values in one call; 3D first-order outflow y; 3D
    hydro outflow x; 1D hydro periodic PPM); by default the step runs on."""
    from tests.test_oracle_floors import golden as floor_golden
    d = floor_golden()
    g = Grid()
    g.primVars, g.consVars = d['m3raise_in_prim'], d['m3raise_in_cons']
    g.faceBx, g.faceBy, g.faceBz = d['m3raise_in_fbx'], d['m3raise_in_fby'], d['m3raise_in_fbz']
    g.dx = g.dy = g.dz = 1 / 8
    g.BcX = g.BcY = g.BcZ = 1
    fl = [float(x) for x in d['m3raise_floors']]
    r = evolve.evolve_3D_mhd(g, 0.0, 10.0, 0, 1, 5 / 3, 0.4, *fl, 0, True, 0.5)
    assert r[6] == 1
    with pytest.raises(IndexError):
        evolve.evolve_3D_mhd(g, 0.0, 10.0, 0, 1, 5 / 3, 0.4, *fl, 0, True, 0.5, raiseLikeReference=True)
    g.BcY = 0
    with pytest.raises(ValueError):
        evolve.evolve_3D_mhd(g, 0.0, 10.0, 0, 1, 5 / 3, 0.4, 1e-12, 1e-12, 1e-12, 0, True, 0.5,
                             raiseLikeReference=True)
    h = Grid()
    h.primVars = np.ones((5, 4, 5, 6))
    h.consVars = np.ones((5, 4, 5, 6))
    h.dx = h.dy = h.dz = 0.2
    h.BcX, h.BcY, h.BcZ = 0, 1, 1
    h.numYCells, h.numXCells, h.numZCells = 4, 5, 6
    with pytest.raises(ValueError):
        evolve.evolve_3D_hydro(h, 0.0, 1.0, 0, 1, 1.4, 0.4, 1e-12, 1e-12, 1e-12, 2, raiseLikeReference=True)
    h1 = Grid()
    h1.primVars = np.ones((5, 16))
    h1.consVars = np.ones((5, 16))
    h1.dx, h1.BcX = 1 / 16, 1
    with pytest.raises(ValueError):
        evolve.evolve_1D_hydro(h1, 0.0, 1.0, 0, 1, 1.4, 0.8, 1e-12, 1e-12, 1e-12, 2, raiseLikeReference=True)


def test_hlle_keyword_and_simulation_key(evolve, tmp_path):
    """riemannSolver=HLLE through the evolve function and the simPar key: the
    oracle's HLLE trajectory bit for bit."""
    from pythonmhd_amd import constants as C
    from pythonmhd_amd.simulation import Simulation
    e = golden()
    tag = 'o2o2'
    g = Grid()
    g.primVars, g.consVars = e[tag + '_in_prim'], e[tag + '_in_cons']
    g.faceBx, g.faceBy = e[tag + '_in_fbx'], e[tag + '_in_fby']
    g.dx, g.dy, g.BcX, g.BcY = 1 / 14, 1 / 12, 0, 0
    r = evolve.evolve_2D_mhd(g, 0.0, 10.0, 0, 3, 5 / 3, 0.8, 1e-12, 1e-12, 1e-12, 2, False, 0.5,
                             riemannSolver=C.HLLE)
    P = M.Params(5 / 3, 0.8, order=2, riemann=1)
    o = M.evolve_2d(g.primVars, g.consVars, g.faceBx, g.faceBy, 1 / 14, 1 / 12, 0, 0, 0.0, 10.0, 0, 3, P)
    assert np.array_equal(r[1], o[1]) and np.array_equal(r[2], o[2])
    gridPar = {C.NUM_X_CELLS: 14, C.MIN_X: 0.0, C.MAX_X: 1.0, C.NUM_Y_CELLS: 12, C.MIN_Y: 0.0, C.MAX_Y: 1.0}
    simPar = {C.IS_MHD: True, C.GAMMA: 5 / 3, C.TLIM: 0.01, C.CFL: 0.8, C.RECONSTRUCT_ORDER: 2,
              C.RIEMANN_SOLVER: C.ROE_HLLE_FALLBACK}
    grid = Simulation(simPar, gridPar, g.primVars, {}, g.consVars, g.faceBx, g.faceBy).run()
    assert np.isfinite(grid.consVars).all()
