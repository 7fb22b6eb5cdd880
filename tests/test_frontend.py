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
