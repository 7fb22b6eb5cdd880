"""On-disk output of a simulation: the reference's .mat snapshots
(Source/FileSaver.py:120-228) and, new in this build, checkpoints that carry the
face-centred field, time and cycle count so that a run restarts bit for bit.

Snapshot files: ``<fullPath>/Data/<outputFileName>_Output_Time_<t>.mat`` with the
keys Density, Vx, Vy, Vz, [Bx, By, Bz,] Pressure, Energy (cell-centred arrays of
the reference layout), the time printed with the reference's number of decimals.
Checkpoints: ``<fullPath>/Data/<outputFileName>_Checkpoint_Time_<t>.npz`` (NumPy,
no pickle): primVars, consVars, faceBx/By/Bz, simTime, numCycles, and the output
time it was written at.
"""
import os
from datetime import datetime

import numpy as np

from . import constants as C


def time_precision(step, t_lim, limit=C.MAX_FILE_NAME_TIME_PREC):
    """Decimals that show multiples of ``step`` and ``t_lim`` exactly (at least one,
    at most ``limit``): Source/FileSaver.py:94-118, Source/SimVisualizer.py:58-67."""
    def prec(x):
        p = 1
        while abs(10.0 ** p * x - int(10.0 ** p * x)) > C.TIME_PRECISION_SMALL_VAL and p < limit:
            p += 1
        return p
    return max(prec(step), prec(t_lim))


class FileSaver:
    """Output folder and snapshot writer (Source/FileSaver.py:43-228).  Unlike the
    reference it does not change the working directory (Source/FileSaver.py:79);
    the default output folder is ./Outputs under the working directory."""

    def __init__(self, saveFigs, saveData, tLim, saveFigDt=-1, saveDataDt=-1, outputFolderPathInput="",
                 outputFolderNameInput="", outputFileNameInput=""):
        if not (saveFigs or saveData):
            raise AssertionError('FileSaver created although the simulation saves nothing')
        if outputFolderPathInput != "":
            if not os.path.exists(outputFolderPathInput):
                raise AssertionError('The outputFolderPath provided in simPar, %s, does not exist'
                                     % outputFolderPathInput)
            self.outputFolderPath = outputFolderPathInput
        else:
            self.outputFolderPath = os.path.join(os.getcwd(), 'Outputs')
            os.makedirs(self.outputFolderPath, exist_ok=True)
        if outputFolderNameInput != "":
            self.outputFolderName = outputFolderNameInput
        else:
            now = datetime.now()
            self.outputFolderName = ("PythonMHD_Simulation_Date_" + now.strftime("month%m_day%d_year%Y_") + "Time_"
                                     + now.strftime("%Hoclock%Mmin") + "_")
        self.saveFigs = saveFigs
        self.saveData = saveData
        # a folder that exists already gets the first free suffix _1, _2, ...
        name = self.outputFolderName
        k = 0
        while os.path.exists(os.path.join(self.outputFolderPath, name)):
            k += 1
            name = self.outputFolderName + "_" + str(k)
        self.outputFolderName = name
        self.fullPath = os.path.join(self.outputFolderPath, name)
        os.mkdir(self.fullPath)
        if saveFigs:
            os.mkdir(os.path.join(self.fullPath, 'Figures'))
        if saveData:
            os.mkdir(os.path.join(self.fullPath, 'Data'))
        self.outputFileName = outputFileNameInput if outputFileNameInput != "" else self.outputFolderName
        self.saveFigDt = saveFigDt
        self.saveDataDt = saveDataDt
        self.figTimePrec = time_precision(saveFigDt, tLim) if saveFigs else -1
        self.dataTimePrec = time_precision(saveDataDt, tLim) if saveData else -1

    def _path(self, kind, simTime, ext):
        return os.path.join(self.fullPath, 'Data', self.outputFileName + ("_%s_Time_{:.%df}" % (kind, self.dataTimePrec))
                            .format(simTime) + ext)

    def saveDataFiles(self, simGrid, simTime, dataFormat):
        """One snapshot of the grid's cell state (Source/FileSaver.py:191-228)."""
        print("Saving Data")
        if dataFormat != C.DATA_FORMAT_MAT:
            raise RuntimeError("Only '%s' data files are written by this package; the reference's '%s' output needs "
                               "the pyevtk package" % (C.DATA_FORMAT_MAT, dataFormat))
        import scipy.io
        w, u = simGrid.primVars, simGrid.consVars
        if simGrid.isMHD:
            data = {C.DENSITY: w[0], C.VX: w[1], C.VY: w[2], C.VZ: w[3], C.BX: w[4], C.BY: w[5], C.BZ: w[6],
                    C.PRESSURE: w[7], C.ENERGY: u[7]}
        else:
            data = {C.DENSITY: w[0], C.VX: w[1], C.VY: w[2], C.VZ: w[3], C.PRESSURE: w[4], C.ENERGY: u[4]}
        path = self._path('Output', simTime, '.mat')
        scipy.io.savemat(path, data)
        return path

    def saveCheckpoint(self, state, simTime, numCycles, outputTime, lastPlotTime=0.0, lastSaveTime=0.0):
        """A restart point: ``state`` = (primVars, consVars, faces) the run continues
        from at time simTime after numCycles cycles, written at outputTime, with the
        last plot and save times that set the run's next stops."""
        prim, cons, faces = state
        arrays = {'primVars': prim, 'consVars': cons, 'simTime': np.float64(simTime),
                  'numCycles': np.int64(numCycles), 'outputTime': np.float64(outputTime),
                  'lastPlotTime': np.float64(lastPlotTime), 'lastSaveTime': np.float64(lastSaveTime)}
        for name, f in zip(('faceBx', 'faceBy', 'faceBz'), faces):
            arrays[name] = f
        path = self._path('Checkpoint', outputTime, '.npz')
        np.savez(path, **arrays)
        return path


def load_checkpoint(path):
    """A checkpoint written by FileSaver.saveCheckpoint: dict with primVars,
    consVars, faces (tuple), simTime, numCycles, outputTime, lastPlotTime,
    lastSaveTime."""
    with np.load(path, allow_pickle=False) as z:
        faces = tuple(np.array(z[n]) for n in ('faceBx', 'faceBy', 'faceBz') if n in z.files)
        return {'primVars': np.array(z['primVars']), 'consVars': np.array(z['consVars']), 'faces': faces,
                'simTime': float(z['simTime']), 'numCycles': int(z['numCycles']),
                'outputTime': float(z['outputTime']), 'lastPlotTime': float(z['lastPlotTime']),
                'lastSaveTime': float(z['lastSaveTime'])}
