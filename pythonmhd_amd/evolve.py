"""Drop-in evolve functions with PythonMHD's signatures and return tuples
(Source/EvolveFunctions_MHD.py:195, :393, :750; Source/EvolveFunctions_Hydro.py:232,
:350, :559), running the step on the MI355X through libpmhd (include/pmhd.h).

Each function reads the duck-typed grid attributes the reference reads
(primVars, consVars, faceBx/faceBy/faceBz, dx/dy/dz, BcX/BcY/BcZ), advances from
tStart until tEnd or maxCycles with the reference's clipping of the last dt, and
returns the final state, the time and cycle reached, and the second-last state.

Extensions (keywords the reference does not have):
- ``reconstructOrder=1`` selects PLM (absent from the reference);
- ``precision=32`` runs the FP32-bulk step;
- ``printCycles=True`` prints the reference's per-cycle line "Cycle = .. Time = ..
  Dt = .." (Source/EvolveFunctions_MHD.py:317, 692, 1152); the loop then runs cycle
  by cycle from the host, with the same results;
- ``riemannSolver``: C.ROE (the reference's), C.HLLE, or C.ROE_HLLE_FALLBACK (Roe,
  HLLE at faces whose Roe intermediate states are unphysical); MHD only;
- ``fixQuirks``: bits of the reference behaviours to correct instead of reproduce
  (FIX_Z_BC, FIX_OUTFLOW_GHOSTS, FIX_DTDY_80BIT, as in include/pmhd.h);
- ``raiseLikeReference=True`` raises where the reference raises instead of running
  on: IndexError when a 3D floor call floors three or more values of one variable
  (Source/PrimCons_HelperModule.py:539-543), ValueError for 3D MHD first-order
  states with outflow y (Source/EvolveFunctions_MHD.py:885), 3D hydro with outflow
  x (Source/EvolveFunctions_Hydro.py:714-716) and 1D hydro periodic PPM
  (Source/Reconstruction_PPM_Hydro_HelperModule.py:355).  By default the step runs
  on (the floors as in 1D/2D, the edge copies for the ghost layers).
"""
import numpy as np

from . import constants as C
from ._lib import HYDRO, MHD, Solver

_solver_cache = {}

# values of the riemannSolver keyword / simPar key -> PMHD_RIEMANN_*
RIEMANN_SOLVERS = (C.ROE, C.HLLE, C.ROE_HLLE_FALLBACK)
_RIEMANN_CODE = {C.ROE: 0, C.HLLE: 1, C.ROE_HLLE_FALLBACK: 2}
FIX_Z_BC, FIX_OUTFLOW_GHOSTS, FIX_DTDY_80BIT = 1, 2, 4


def _solver(key, make):
    s = _solver_cache.get(key)
    if s is None:
        _solver_cache.clear()   # one resident grid at a time
        s = make()
        _solver_cache[key] = s
    return s


def last_stats():
    """Floor counters and health of the last evolve call (``Solver.stats``), or
    None before the first call.  The reference floors silently; these counts say
    how often it would have.  Keys: floored / floored_max (10 call sites x density,
    pressure-or-energy), floored_total, cycles, nonfinite."""
    for s in _solver_cache.values():
        return s.stats()
    return None


def release_device_state():
    """Free the device buffers of the cached solver."""
    for s in _solver_cache.values():
        s.close()
    _solver_cache.clear()


def _check_common(tStart, tEnd, startCycle):
    assert tStart <= tEnd
    assert startCycle >= 0


def _loop_printing(s, ndim, physics, tStart, tEnd, startCycle, maxCycles):
    """The evolve loop driven from the host, printing each cycle as the reference
    does; the same sequence of timesteps and clipping as pmhd_advance."""
    t, cyc, t2 = tStart, startCycle, tStart
    while t < tEnd and cyc < maxCycles:
        t2 = t
        dt = s.compute_dt()
        if ndim == 3 and physics == MHD:
            if tEnd - t < dt:
                dt = tEnd - t
        elif t + dt > tEnd:
            dt = tEnd - t
        s.step(dt)
        t += dt
        cyc += 1
        print("Cycle = " + str(cyc) + " Time = " + str(t) + " Dt = " + str(dt))
    return t, cyc, t2


def _reference_raises(shape, bcs, physics, reconstructOrder):
    """The exception the reference raises in its first cycle on this input (or None)."""
    nd = len(shape)
    if physics == MHD and nd == 3 and reconstructOrder == 0 and bcs[1] == C.OUTFLOW:
        return ValueError('the reference cannot append its first-order outflow y ghost states in 3D '
                          '(Source/EvolveFunctions_MHD.py:885)')
    if physics == HYDRO and nd == 3 and bcs[0] == C.OUTFLOW:
        return ValueError('the reference cannot reshape its outflow x transverse fluxes in 3D hydro '
                          '(Source/EvolveFunctions_Hydro.py:714-716)')
    if physics == HYDRO and nd == 1 and bcs[0] == C.PERIODIC and reconstructOrder == 2:
        return ValueError('the reference cannot append its periodic 1D hydro PPM states '
                          '(Source/Reconstruction_PPM_Hydro_HelperModule.py:355)')
    return None


def _run(shape, widths, bcs, prim, cons, faces, tStart, tEnd, startCycle, maxCycles, gamma, cfl,
         minDens, minPres, minEnergy, reconstructOrder, entropyFix, epsilon, precision, physics=MHD,
         printCycles=False, riemannSolver=C.ROE, fixQuirks=0, raiseLikeReference=False):
    if riemannSolver not in RIEMANN_SOLVERS:
        raise RuntimeError('riemannSolver must be one of %s' % (RIEMANN_SOLVERS,))
    runs = tStart < tEnd and startCycle < maxCycles
    if raiseLikeReference and runs:
        e = _reference_raises(shape, bcs, physics, reconstructOrder)
        if e is not None:
            raise e
    riemann = _RIEMANN_CODE[riemannSolver]
    key = (shape, tuple(widths), tuple(bcs), float(gamma), float(cfl), float(minDens), float(minPres),
           float(minEnergy), int(reconstructOrder), bool(entropyFix), float(epsilon), int(precision),
           int(physics), riemann, int(fixQuirks))
    s = _solver(key, lambda: Solver(shape, widths, bcs, gamma, cfl, minDens, minPres, minEnergy,
                                    reconstructOrder, entropyFix, epsilon, precision, physics=physics,
                                    riemann=riemann, fix_quirks=int(fixQuirks)))
    s.set_state(prim, cons, faces)
    # a state that turns non-finite ends the loop with a NaN clock, as the
    # reference's does; the results are returned all the same
    if printCycles:
        t, cyc, t2 = _loop_printing(s, len(shape), physics, tStart, tEnd, startCycle, maxCycles)
    else:
        t, cyc, t2 = s.advance(tStart, tEnd, startCycle, maxCycles, allow_nonfinite=True)
    if raiseLikeReference and len(shape) == 3:
        st = s.stats()
        if st['floored_max'][:6].max() >= 3:
            raise IndexError('index 12 is out of bounds for axis 0 (the reference\'s 3D applyFloors with three or '
                             'more floored values in one call, Source/PrimCons_HelperModule.py:539-543)')
    p, c, f = s.get_state(0)
    p2, c2, f2 = s.get_state(1)
    return p, c, f, t, cyc, p2, c2, f2, t2


def _opts(precision, printCycles, riemannSolver, fixQuirks, raiseLikeReference):
    return dict(printCycles=printCycles, riemannSolver=riemannSolver, fixQuirks=fixQuirks,
                raiseLikeReference=raiseLikeReference)


def evolve_1D_mhd(simulationGrid, tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens, minPres,
                  minEnergy, reconstructOrder=0, entropyFix=False, epsilon=0.5,
                  precision=C.DEFAULT_PRECISION, printCycles=False, riemannSolver=C.ROE, fixQuirks=0,
                  raiseLikeReference=False):
    prim = np.asarray(simulationGrid.primVars, dtype=np.float64)
    assert prim.ndim == 2 and prim.shape[0] == 8 and prim.shape[1] > 0
    cons = np.asarray(simulationGrid.consVars, dtype=np.float64)
    assert prim.shape == cons.shape
    dx = simulationGrid.dx
    assert dx > 0
    bcx = simulationGrid.BcX
    assert bcx == C.OUTFLOW or bcx == C.PERIODIC
    _check_common(tStart, tEnd, startCycle)
    p, c, _, t, cyc, p2, c2, _, t2 = _run(
        (prim.shape[1],), (dx,), (bcx,), prim, cons, (), tStart, tEnd, startCycle, maxCycles, gamma, cfl,
        minDens, minPres, minEnergy, reconstructOrder, entropyFix, epsilon, precision,
        **_opts(precision, printCycles, riemannSolver, fixQuirks, raiseLikeReference))
    return p, c, t, cyc, p2, c2, t2


def evolve_2D_mhd(simulationGrid, tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens, minPres,
                  minEnergy, reconstructOrder=0, entropyFix=False, epsilon=0.5,
                  precision=C.DEFAULT_PRECISION, printCycles=False, riemannSolver=C.ROE, fixQuirks=0,
                  raiseLikeReference=False):
    prim = np.asarray(simulationGrid.primVars, dtype=np.float64)
    bcx, bcy = simulationGrid.BcX, simulationGrid.BcY
    assert bcx == C.OUTFLOW or bcx == C.PERIODIC
    assert bcy == C.OUTFLOW or bcy == C.PERIODIC
    assert prim.ndim == 3 and prim.shape[0] == 8 and prim.shape[1] > 0
    cons = np.asarray(simulationGrid.consVars, dtype=np.float64)
    assert prim.shape == cons.shape
    dx, dy = simulationGrid.dx, simulationGrid.dy
    assert dx > 0
    assert dy > 0
    _check_common(tStart, tEnd, startCycle)
    faces = (simulationGrid.faceBx, simulationGrid.faceBy)
    p, c, f, t, cyc, p2, c2, f2, t2 = _run(
        prim.shape[1:], (dx, dy), (bcx, bcy), prim, cons, faces, tStart, tEnd, startCycle, maxCycles,
        gamma, cfl, minDens, minPres, minEnergy, reconstructOrder, entropyFix, epsilon, precision,
        **_opts(precision, printCycles, riemannSolver, fixQuirks, raiseLikeReference))
    return p, c, f[0], f[1], t, cyc, p2, c2, f2[0], f2[1], t2


def evolve_3D_mhd(simulationGrid, tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens, minPres,
                  minEnergy, reconstructOrder=0, entropyFix=True, epsilon=0.5,
                  precision=C.DEFAULT_PRECISION, printCycles=False, riemannSolver=C.ROE, fixQuirks=0,
                  raiseLikeReference=False):
    prim = np.asarray(simulationGrid.primVars, dtype=np.float64)
    bcx, bcy, bcz = simulationGrid.BcX, simulationGrid.BcY, simulationGrid.BcZ
    for b in (bcx, bcy, bcz):
        assert b == C.OUTFLOW or b == C.PERIODIC
    assert prim.ndim == 4 and prim.shape[0] == 8 and prim.shape[1] > 0
    cons = np.asarray(simulationGrid.consVars, dtype=np.float64)
    assert prim.shape == cons.shape
    dx, dy, dz = simulationGrid.dx, simulationGrid.dy, simulationGrid.dz
    assert dx > 0
    assert dy > 0
    assert dz > 0
    _check_common(tStart, tEnd, startCycle)
    faces = (simulationGrid.faceBx, simulationGrid.faceBy, simulationGrid.faceBz)
    p, c, f, t, cyc, p2, c2, f2, t2 = _run(
        prim.shape[1:], (dx, dy, dz), (bcx, bcy, bcz), prim, cons, faces, tStart, tEnd, startCycle,
        maxCycles, gamma, cfl, minDens, minPres, minEnergy, reconstructOrder, entropyFix, epsilon,
        precision, **_opts(precision, printCycles, riemannSolver, fixQuirks, raiseLikeReference))
    return p, c, f[0], f[1], f[2], t, cyc, p2, c2, f2[0], f2[1], f2[2], t2


# ---------------------------------------------------------------------------
# hydrodynamics: five variables (rho, v, p) / (rho, M, E), no face field; the
# same tuple in 1D, 2D and 3D (Source/EvolveFunctions_Hydro.py:305, 516, 806)
# ---------------------------------------------------------------------------
def _hydro(simulationGrid, nd, tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens, minPres,
           minEnergy, reconstructOrder, entropyFix, epsilon, precision, printCycles=False, riemannSolver=C.ROE,
           fixQuirks=0, raiseLikeReference=False):
    prim = np.asarray(simulationGrid.primVars, dtype=np.float64)
    assert prim.ndim == nd + 1 and prim.shape[0] == 5
    cons = np.asarray(simulationGrid.consVars, dtype=np.float64)
    assert prim.shape == cons.shape
    names = ('dx', 'dy', 'dz')[:nd]
    bcn = ('BcX', 'BcY', 'BcZ')[:nd]
    if nd >= 2:
        # the reference reads the cell counts from the grid (EvolveFunctions_Hydro.py:363-368)
        assert simulationGrid.numXCells > 1
        assert simulationGrid.numYCells > 1
        if nd == 3:
            assert simulationGrid.numZCells > 1
    widths = tuple(getattr(simulationGrid, n) for n in names)
    for w in widths:
        assert w > 0
    bcs = tuple(getattr(simulationGrid, n) for n in bcn)
    for b in bcs:
        assert b == C.OUTFLOW or b == C.PERIODIC
    _check_common(tStart, tEnd, startCycle)
    assert maxCycles > 0
    shape = prim.shape[1:]
    p, c, _, t, cyc, p2, c2, _, t2 = _run(
        shape, widths, bcs, prim, cons, (), tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens,
        minPres, minEnergy, reconstructOrder, entropyFix, epsilon, precision, physics=HYDRO,
        **_opts(precision, printCycles, riemannSolver, fixQuirks, raiseLikeReference))
    return p, c, t, cyc, p2, c2, t2


def evolve_1D_hydro(simulationGrid, tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens, minPres,
                    minEnergy, reconstructOrder=0, entropyFix=False, epsilon=0.5,
                    precision=C.DEFAULT_PRECISION, printCycles=False, riemannSolver=C.ROE, fixQuirks=0,
                  raiseLikeReference=False):
    return _hydro(simulationGrid, 1, tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens, minPres,
                  minEnergy, reconstructOrder, entropyFix, epsilon, precision, printCycles, riemannSolver,
                  fixQuirks, raiseLikeReference)


def evolve_2D_hydro(simulationGrid, tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens, minPres,
                    minEnergy, reconstructOrder=0, entropyFix=False, epsilon=0.5,
                    precision=C.DEFAULT_PRECISION, printCycles=False, riemannSolver=C.ROE, fixQuirks=0,
                  raiseLikeReference=False):
    return _hydro(simulationGrid, 2, tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens, minPres,
                  minEnergy, reconstructOrder, entropyFix, epsilon, precision, printCycles, riemannSolver,
                  fixQuirks, raiseLikeReference)


def evolve_3D_hydro(simulationGrid, tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens, minPres,
                    minEnergy, reconstructOrder=0, entropyFix=False, epsilon=0.5,
                    precision=C.DEFAULT_PRECISION, printCycles=False, riemannSolver=C.ROE, fixQuirks=0,
                  raiseLikeReference=False):
    return _hydro(simulationGrid, 3, tStart, tEnd, startCycle, maxCycles, gamma, cfl, minDens, minPres,
                  minEnergy, reconstructOrder, entropyFix, epsilon, precision, printCycles, riemannSolver,
                  fixQuirks, raiseLikeReference)
