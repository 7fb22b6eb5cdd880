"""The simulation driver (Source/Simulation.py:47-434, Source/Simulation_HelperModule.py:28-480)
over the MI355X evolve functions.

``Simulation(simPar, gridPar, primVars, visPar={}, consVars=..., faceBx=..., faceBy=...,
faceBz=...).run()`` takes an init file's dictionaries unchanged, checks them with
the reference's rules, builds the SimulationGrid and runs the reference's loop:
the evolve function of the grid's dimension and physics advances to the next stop
(the next data-save or plot time, or tLim), the snapshot is written at the stop,
and unless the run is over the grid rolls back to the state before the last
(shortened) cycle, so the next leg repeats that cycle with an unclipped timestep;
the cycle count is not rolled back.  Returns the grid.

Differences from the reference, all stated where they apply:
- reconstructOrder also accepts 1 (PLM), and simPar may hold ``precision`` (64 or
  32) and ``riemannSolver`` (C.ROE; the reference names the key but defines no
  value, so any value raises there);
- plotData keeps the plot stops (they shape the trajectory through the rollback)
  but draws nothing; .vtk output is refused (the reference needs pyevtk);
- ``saveCheckpoints`` (new) writes a restart point at every data save, and
  ``Simulation.resume(checkpoint, ...)`` continues from it bit for bit;
- ``printCycles`` (new, default False) prints the reference's per-cycle line.
"""
import os
import time as timeLib

import numpy as np

from . import constants as C
from . import evolve as E
from .filesaver import FileSaver, load_checkpoint, time_precision
from .simulation_grid import SimulationGrid

SAVE_CHECKPOINTS = "saveCheckpoints"


def _is_num(x):
    return type(x) == int or type(x) == float  # noqa: E721 -- the reference's test (bool is not numeric)


def validateSimPar(simPar, primVars):
    """Source/Simulation_HelperModule.py:28-480: types and ranges of the simulation
    parameters; raises RuntimeError on the first violation."""
    k = simPar.keys()
    if C.IS_MHD in k and type(simPar[C.IS_MHD]) != bool:  # noqa: E721
        raise RuntimeError("\nYou passed an invalid value for isMHD; it must be True or False.")
    if C.TLIM not in k:
        raise RuntimeError("You did not specify a time limit (tLim) in your simulation parameters.")
    if not _is_num(simPar[C.TLIM]):
        raise RuntimeError("\nYou did not provide a valid data type for tLim; it must be numeric and greater than 0.")
    if simPar[C.TLIM] <= 0.0:
        raise RuntimeError("\nThe time limit tLim must be greater than 0.")
    if C.MAX_CYCLES in k:
        if type(simPar[C.MAX_CYCLES]) != int:  # noqa: E721
            raise RuntimeError("\nYou passed an invalid data type for maxCycles; it must be an integer.")
        if simPar[C.MAX_CYCLES] < 1:
            raise RuntimeError("\nmaxCycles must be at least 1.")
    if not isinstance(primVars, np.ndarray):
        raise RuntimeError("\nThe primitive variables must be a NumPy array.")
    nd = len(np.squeeze(primVars).shape) - 1
    if nd not in (1, 2, 3):
        raise RuntimeError("\nYou have an invalid number of spatial dimensions in your primitive variables matrix.")
    if C.CFL in k:
        cfl = simPar[C.CFL]
        if type(cfl) != float:  # noqa: E721
            raise RuntimeError("\nYou passed an invalid data type for cfl; it must be a float.")
        if cfl <= 0:
            raise RuntimeError("\ncfl must be greater than 0.")
        if nd == 3 and cfl >= 0.5:
            raise RuntimeError("\ncfl must be below 0.5 in 3D.")
        if nd != 3 and cfl >= 1.0:
            raise RuntimeError("\ncfl must be below 1.0 in 1D and 2D.")
    if C.MIN_DENSITY in k:
        if type(simPar[C.MIN_DENSITY]) != float:  # noqa: E721
            raise RuntimeError("You passed an invalid data type for minDens; it must be a float.")
        if simPar[C.MIN_DENSITY] <= 0:
            raise RuntimeError("minDens must be greater than 0.")
        if simPar[C.MIN_DENSITY] != C.DEFAULT_MIN_DENSITY:
            print("Runtime Warning: minDens (%s) differs from the default minimum density (%s)."
                  % (simPar[C.MIN_DENSITY], C.DEFAULT_MIN_DENSITY))
    if C.RECONSTRUCT_ORDER in k:
        order = simPar[C.RECONSTRUCT_ORDER]
        if type(order) != int:  # noqa: E721
            raise RuntimeError("\nYou passed an invalid data type for reconstructOrder; it must be an integer.")
        # the reference accepts 0 and 2; 1 (PLM) is this build's addition
        if order not in (C.NO_SPATIAL_RECONSTRUCTION, C.PLM_SPATIAL_RECONSTRUCTION, C.PPM_SPATIAL_RECONSTRUCTION):
            raise RuntimeError("\nreconstructOrder must be 0 (none), 1 (PLM) or 2 (PPM).")
    if C.RIEMANN_SOLVER in k and simPar[C.RIEMANN_SOLVER] not in E.RIEMANN_SOLVERS:
        raise RuntimeError("You passed an invalid value for riemannSolver; valid values: %s" % (E.RIEMANN_SOLVERS,))
    if C.GAMMA in k:
        if type(simPar[C.GAMMA]) != float:  # noqa: E721
            raise RuntimeError("\nYou passed an invalid data type for gamma; it must be a float.")
        if simPar[C.GAMMA] <= 1.0:
            raise RuntimeError("\ngamma must be greater than 1.")
    if C.ENTROPY_FIX in k:
        if type(simPar[C.ENTROPY_FIX]) != bool:  # noqa: E721
            raise RuntimeError("\nYou passed an invalid value for entropyFix; it must be True or False.")
        if simPar[C.ENTROPY_FIX] and C.EPSILON in k:
            if not _is_num(simPar[C.EPSILON]):
                raise RuntimeError("\nYou passed an invalid data type for epsilon; it must be numeric.")
            if simPar[C.EPSILON] <= 0.0:
                raise RuntimeError("\nepsilon must be greater than 0.")
    if C.RIEMANN_SOLVER in k and simPar[C.RIEMANN_SOLVER] != C.ROE and not simPar.get(C.IS_MHD, C.DEFAULT_IS_MHD):
        raise RuntimeError("The hydrodynamic step has the Roe riemannSolver only.")
    if C.FIX_QUIRKS in k and (type(simPar[C.FIX_QUIRKS]) != int or simPar[C.FIX_QUIRKS] & ~7):  # noqa: E721
        raise RuntimeError("fixQuirks must be an integer of FIX_Z_BC | FIX_OUTFLOW_GHOSTS | FIX_DTDY_80BIT bits.")
    if C.PRECISION in k and simPar[C.PRECISION] not in (32, 64):
        raise RuntimeError("\nprecision must be 64 (FP64 throughout) or 32 (FP32 bulk with FP64 islands).")
    if C.SAVE_DATA in k:
        if type(simPar[C.SAVE_DATA]) != bool:  # noqa: E721
            raise RuntimeError("\nYou passed an invalid value for saveData; it must be True or False.")
        if simPar[C.SAVE_DATA]:
            plotting = C.PLOT_DATA in k and type(simPar[C.PLOT_DATA]) == bool and simPar[C.PLOT_DATA]  # noqa: E721
            if not plotting and C.SAVE_DATA_DT not in k:
                raise RuntimeError("\nYou did not specify how often to save data (saveDt).")
            if C.SAVE_DATA_DT in k:
                if not _is_num(simPar[C.SAVE_DATA_DT]):
                    raise RuntimeError("\nYou passed an invalid data type for saveDt; it must be numeric.")
                if simPar[C.SAVE_DATA_DT] <= 0:
                    raise RuntimeError("\nsaveDt must be greater than 0.")
                if C.SAVE_DATA_FORMAT in k:
                    if not isinstance(simPar[C.SAVE_DATA_FORMAT], str):
                        raise RuntimeError("\nThe data format must be one of %s." % (C.VALID_DATA_FILE_FORMATS,))
                    if simPar[C.SAVE_DATA_FORMAT] not in C.VALID_DATA_FILE_FORMATS:
                        raise RuntimeError("\nUnsupported data format; valid: %s." % (C.VALID_DATA_FILE_FORMATS,))
    if C.PLOT_DATA in k:
        if type(simPar[C.PLOT_DATA]) != bool:  # noqa: E721
            raise RuntimeError("\nYou passed an invalid data type for plotData; it must be True or False.")
        if simPar[C.PLOT_DATA]:
            if C.PLOT_DATA_DT not in k:
                raise RuntimeError("\nWith plotData True you must provide plotDt.")
            if not _is_num(simPar[C.PLOT_DATA_DT]):
                raise RuntimeError("\nYou passed an invalid data type for plotDt; it must be numeric.")
            if simPar[C.PLOT_DATA_DT] <= 0:
                raise RuntimeError("\nplotDt must be greater than 0.")
            breaks = int(simPar[C.TLIM] / simPar[C.PLOT_DATA_DT]) + 1
            if breaks > C.MAX_PLOT_BREAKS:
                raise RuntimeError("\nYour plotDt gives %d plot updates, more than %d." % (breaks, C.MAX_PLOT_BREAKS))
            if C.MATPLOTLIB_BACKEND in k:
                if not isinstance(simPar[C.MATPLOTLIB_BACKEND], str) or \
                        simPar[C.MATPLOTLIB_BACKEND] not in C.MATPLOTLIB_INTERACTIVE_BACKENDS:
                    raise RuntimeError("\nmatplotlibBackend must be one of %s." % (C.MATPLOTLIB_INTERACTIVE_BACKENDS,))
            if C.SAVE_FIGS in k:
                if type(simPar[C.SAVE_FIGS]) != bool:  # noqa: E721
                    raise RuntimeError("\nYou passed an invalid value for saveFigs; it must be True or False.")
                if simPar[C.SAVE_FIGS]:
                    if C.MAKE_MOVIE in k and type(simPar[C.MAKE_MOVIE]) != bool:  # noqa: E721
                        raise RuntimeError("\nYou passed an invalid value for makeMovie; it must be True or False.")
                    if C.SAVE_FIGS_FORMAT in k:
                        if not isinstance(simPar[C.SAVE_FIGS_FORMAT], str):
                            raise RuntimeError("\nfigFormat must be a string.")
                        if simPar.get(C.MAKE_MOVIE) and simPar[C.SAVE_FIGS_FORMAT] != C.IMAGE_FORMAT_PNG:
                            raise RuntimeError("\nMovies are made from png figures only.")
                    if C.SAVE_FIGS_DPI in k:
                        dpi = simPar[C.SAVE_FIGS_DPI]
                        if type(dpi) != int:  # noqa: E721
                            raise RuntimeError("\nsaveFigsDPI must be an integer.")
                        if dpi <= 0 or dpi < C.MIN_RECOMMENDED_SAVE_FIGS_DPI or dpi > C.MAX_RECOMMENDED_SAVE_FIGS_DPI:
                            raise RuntimeError("\nsaveFigsDPI must lie in [%d, %d]."
                                               % (C.MIN_RECOMMENDED_SAVE_FIGS_DPI, C.MAX_RECOMMENDED_SAVE_FIGS_DPI))
    saving = (C.SAVE_FIGS in k) or (C.SAVE_DATA in k)
    if C.SAVE_FIGS in k:
        saving = simPar[C.SAVE_FIGS]
    if not saving and C.SAVE_DATA in k:
        saving = simPar[C.SAVE_DATA]
    if saving:
        if C.OUTPUT_FOLDER_PATH in k:
            if not isinstance(simPar[C.OUTPUT_FOLDER_PATH], str):
                raise RuntimeError("\noutputFolderPath must be a string.")
            if not os.path.exists(simPar[C.OUTPUT_FOLDER_PATH]):
                raise RuntimeError("\nThe outputFolderPath %s does not exist." % simPar[C.OUTPUT_FOLDER_PATH])
        if C.OUTPUT_FOLDER_NAME in k:
            name = simPar[C.OUTPUT_FOLDER_NAME]
            if not isinstance(name, str) or len(name.strip()) == 0 or len(name) > C.MAX_OUTPUT_FOLDER_NAME_CHARACTERS:
                raise RuntimeError("\noutputFolderName must be a non-empty string of at most %d characters."
                                   % C.MAX_OUTPUT_FOLDER_NAME_CHARACTERS)
        if C.OUTPUT_FILE_NAME in k:
            name = simPar[C.OUTPUT_FILE_NAME]
            if not isinstance(name, str) or len(name) > C.MAX_OUTPUT_FILE_NAME_CHARACTERS:
                raise RuntimeError("\noutputFileName must be a string of at most %d characters."
                                   % C.MAX_OUTPUT_FILE_NAME_CHARACTERS)


class Simulation:
    """Source/Simulation.py:47-434 on the MI355X evolve functions."""

    def __init__(self, simPar, gridPar, primVars, visPar={}, consVars=np.zeros(shape=(1)),  # noqa: B006
                 faceBx=np.zeros(shape=(1)), faceBy=np.zeros(shape=(1)), faceBz=np.zeros(shape=(1))):
        print("PythonMHD on the MI355X\n")
        print("Validating Simulation Parameters\n")
        validateSimPar(simPar, primVars)
        g = simPar.get
        self.isMHD = g(C.IS_MHD, C.DEFAULT_IS_MHD)
        self.tLim = simPar[C.TLIM]
        self.maxCycles = g(C.MAX_CYCLES, C.DEFAULT_MAX_CYCLES)
        self.reconstructOrder = g(C.RECONSTRUCT_ORDER, C.NO_SPATIAL_RECONSTRUCTION)
        # (the reference reads gamma only when given and then fails without it)
        self.gamma = g(C.GAMMA, C.DEFAULT_GAMMA)
        self.entropyFix = g(C.ENTROPY_FIX, False)
        self.epsilon = g(C.EPSILON, C.DEFAULT_EPSILON)
        self.minDens = g(C.MIN_DENSITY, C.DEFAULT_MIN_DENSITY)
        self.minPres = g(C.MIN_PRESSURE, C.DEFAULT_MIN_PRESSURE)
        self.minEnergy = g(C.MIN_ENERGY, C.DEFAULT_MIN_ENERGY)
        self.precision = g(C.PRECISION, C.DEFAULT_PRECISION)
        self.riemannSolver = g(C.RIEMANN_SOLVER, C.ROE)
        self.printCycles = bool(g(C.PRINT_CYCLES, False))
        self.fixQuirks = g(C.FIX_QUIRKS, 0)
        self.raiseLikeReference = bool(g(C.RAISE_LIKE_REFERENCE, False))
        print("Building Simulation Grid\n")
        self.grid = SimulationGrid(primVars, gridPar, self.isMHD, self.gamma, self.minDens, self.minPres,
                                   self.minEnergy, consVars, faceBx, faceBy, faceBz)
        self.cfl = {1: C.DEFAULT_CFL_1D, 2: C.DEFAULT_CFL_2D, 3: C.DEFAULT_CFL_3D}[self.grid.nDim]
        self.cfl = g(C.CFL, self.cfl)
        self.plotData = g(C.PLOT_DATA, C.DEFAULT_PLOT_DATA)
        self.saveFigs = bool(self.plotData and g(C.SAVE_FIGS, C.DEFAULT_SAVE_FIGS))
        if self.plotData:
            self.plotDt = simPar[C.PLOT_DATA_DT]
            self.plotTimePrec = time_precision(self.plotDt, self.tLim)
            print("Note: plotData is set; this package keeps the plot times as stops of the run but does not draw.")
        self.saveData = g(C.SAVE_DATA, False)
        self.saveCheckpoints = bool(g(SAVE_CHECKPOINTS, False))
        self.FileSaver = None
        if self.saveData:
            self.saveDt = simPar[C.SAVE_DATA_DT] if C.SAVE_DATA_DT in simPar else self.plotDt
            self.saveDataFormat = g(C.SAVE_DATA_FORMAT, C.DEFAULT_DATA_FORMAT)
            if self.saveDataFormat != C.DATA_FORMAT_MAT:
                raise RuntimeError("Only '%s' data files are written by this package" % C.DATA_FORMAT_MAT)
            self.FileSaver = FileSaver(False, True, self.tLim, -1.0, self.saveDt, g(C.OUTPUT_FOLDER_PATH, ""),
                                       g(C.OUTPUT_FOLDER_NAME, ""), g(C.OUTPUT_FILE_NAME, ""))
        self.savedFiles = []
        self.startTime = 0.0
        self.startCycle = 0

    @classmethod
    def resume(cls, checkpoint, simPar, gridPar, visPar={}):  # noqa: B006
        """Continue a run from a checkpoint written with saveCheckpoints: the state,
        face field, time and cycle count the interrupted run continued from, so the
        resumed run gives the same results bit for bit."""
        ck = load_checkpoint(checkpoint) if isinstance(checkpoint, str) else checkpoint
        faces = list(ck['faces']) + [np.zeros(shape=(1))] * (3 - len(ck['faces']))
        sim = cls(simPar, gridPar, ck['primVars'], visPar, ck['consVars'], *faces)
        sim.startTime = ck['simTime']
        sim.startCycle = ck['numCycles']
        sim.resumeStops = (ck['lastPlotTime'], ck['lastSaveTime'])
        return sim

    def run(self):
        return self.__run_sim()

    def _evolve(self):
        nd, mhd = self.grid.nDim, self.isMHD
        if not mhd:
            return {1: E.evolve_1D_hydro, 2: E.evolve_2D_hydro, 3: E.evolve_3D_hydro}[nd]
        return {1: E.evolve_1D_mhd, 2: E.evolve_2D_mhd, 3: E.evolve_3D_mhd}[nd]

    def __run_sim(self):
        print("Starting simulation\n")
        t_start = timeLib.time()
        simTime = self.startTime
        numCycles = self.startCycle
        evolve = self._evolve()
        extra = dict(precision=self.precision, printCycles=self.printCycles, riemannSolver=self.riemannSolver,
                     fixQuirks=self.fixQuirks, raiseLikeReference=self.raiseLikeReference)
        # the stops: the next plot time and the next save time (Simulation.py:330-346)
        resumed = hasattr(self, 'resumeStops')
        lastPlotTime, lastSaveTime = self.resumeStops if resumed else (0.0, 0.0)
        if self.saveData and not resumed:
            self.savedFiles.append(self.FileSaver.saveDataFiles(self.grid, simTime, self.saveDataFormat))
        nextPlotTime = nextSaveTime = self.tLim
        if self.plotData:
            nextPlotTime = np.minimum(nextPlotTime, lastPlotTime + self.plotDt) if not resumed else \
                np.round(np.minimum(nextPlotTime, lastPlotTime + self.plotDt), self.plotTimePrec)
        if self.saveData:
            nextSaveTime = np.minimum(nextSaveTime, lastSaveTime + self.saveDt) if not resumed else \
                np.round(np.minimum(nextSaveTime, lastSaveTime + self.saveDt), self.FileSaver.dataTimePrec)
        nextStopTime = np.minimum(nextPlotTime, nextSaveTime)
        new = None
        g = self.grid
        while simTime < self.tLim and numCycles < self.maxCycles:
            r = evolve(g, simTime, nextStopTime, numCycles, self.maxCycles, self.gamma, self.cfl, self.minDens,
                       self.minPres, self.minEnergy, self.reconstructOrder, self.entropyFix, self.epsilon, **extra)
            nf = len(g.faces())
            # (prim, cons, faces..., t, cycle, prim2, cons2, faces2..., t2)
            new = (r[0], r[1], tuple(r[2:2 + nf]))
            newTime, newCycle = r[2 + nf], r[3 + nf]
            second = (r[4 + nf], r[5 + nf], tuple(r[6 + nf:6 + 2 * nf]))
            secondTime = r[6 + 2 * nf]
            self._set_state(new)
            simTime = newTime
            numCycles = newCycle
            saved = self.saveData and nextStopTime == nextSaveTime
            if saved:
                lastSaveTime = simTime
                self.savedFiles.append(self.FileSaver.saveDataFiles(g, simTime, self.saveDataFormat))
            if self.plotData and nextStopTime == nextPlotTime:
                lastPlotTime = simTime
            nextPlotTime = nextSaveTime = self.tLim
            if self.plotData:
                nextPlotTime = np.round(np.minimum(nextPlotTime, lastPlotTime + self.plotDt), self.plotTimePrec)
            if self.saveData:
                nextSaveTime = np.round(np.minimum(nextSaveTime, lastSaveTime + self.saveDt),
                                        self.FileSaver.dataTimePrec)
            nextStopTime = np.minimum(nextPlotTime, nextSaveTime)
            if simTime < self.tLim and numCycles < self.maxCycles:
                # the rollback: repeat the shortened last cycle in the next leg
                self._set_state(second)
                if saved and self.saveCheckpoints:
                    self.savedFiles.append(self.FileSaver.saveCheckpoint(
                        second, secondTime, numCycles, simTime, lastPlotTime=lastPlotTime, lastSaveTime=lastSaveTime))
                simTime = secondTime
        if new is not None:
            self._set_state(new)
        self.simTime, self.numCycles = simTime, numCycles
        print("Simulation took: ", timeLib.time() - t_start, " seconds")
        print("SIMULATION PROCESSING COMPLETE. SAFE TO END PROGRAM.")
        return self.grid

    def _set_state(self, state):
        prim, cons, faces = state
        self.grid.primVars = prim
        self.grid.consVars = cons
        for name, f in zip(('faceBx', 'faceBy', 'faceBz'), faces):
            setattr(self.grid, name, f)
