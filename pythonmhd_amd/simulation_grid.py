"""The grid of a simulation (Source/SimulationGrid.py:47-169) and its checks
(Source/SimulationGrid_HelperModule.py:25-407).

``SimulationGrid`` holds what the evolve functions read -- primVars, consVars,
faceBx/By/Bz, dx/dy/dz, BcX/BcY/BcZ, numX/Y/ZCells -- plus the coordinates,
built the reference's way: the primitive matrix is squeezed and checked against
gridPar, a hydro matrix with fewer variables is padded to five, densities,
pressures and energies below the floors are rejected, consVars and the face
fields are formed from primVars when not given (prim -> cons with the reference's
sum order, faces as means of the neighbouring cells).  The checks raise the
reference's RuntimeErrors on the same conditions (messages are this package's).
"""
import numpy as np

from . import constants as C
from . import ct
from .initialization import createCoordinates1D_Cartesian, createCoordinates2D_Cartesian, \
    createCoordinates3D_Cartesian
from .primcons import primToCons_hydro, primToCons_mhd_allVars

_AXES = ((C.NUM_X_CELLS, C.MIN_X, C.MAX_X, 'x'), (C.NUM_Y_CELLS, C.MIN_Y, C.MAX_Y, 'y'),
         (C.NUM_Z_CELLS, C.MIN_Z, C.MAX_Z, 'z'))


def validateGridPar(primVars, gridPar):
    """Source/SimulationGrid_HelperModule.py:25-197: cell counts present, integral,
    at least 2 and equal to the matrix's extents; coordinate ranges numeric and
    increasing."""
    nd = len(np.squeeze(primVars).shape) - 1
    if nd < 1:
        raise RuntimeError("\nZero spatial dimensions in primitive variable matrix. Simulation must be 1D, 2D, or 3D.")
    if nd > 3:
        raise RuntimeError("\nToo many spatial dimensions in primitive variable matrix. "
                           "Simulation must be 1D, 2D, or 3D.")
    for key, _, _, _ in _AXES[:nd]:
        if key not in gridPar.keys():
            raise RuntimeError("\nYou did not specify " + key + " in your grid parameters dictionary.")
        if not isinstance(gridPar[key], int):
            raise RuntimeError("\nThe " + key + " value in your grid parameters dictionary is not an integer.")
    for key, _, _, axis in _AXES[:nd]:
        if gridPar[key] < 2:
            raise RuntimeError("\nThe " + key + " value in your grid parameters dictionary is too small.\nYou need "
                               "at least two cells in the %s-direction for a %dD simulation." % (axis, nd))
    # extents of the matrix: (v, nx) in 1D, (v, ny, nx[, nz]) otherwise
    ext = {'x': primVars.shape[1] if nd == 1 else primVars.shape[2]}
    if nd >= 2:
        ext['y'] = primVars.shape[1]
    if nd == 3:
        ext['z'] = primVars.shape[3]
    for key, _, _, axis in _AXES[:nd]:
        if ext[axis] > gridPar[key]:
            raise RuntimeError("\nPrimitive variables matrix has too many cells in the %s-direction.\nPlease adjust "
                               "the matrix or change the value of %s in your grid parameters dictionary." % (axis, key))
        if ext[axis] < gridPar[key]:
            raise RuntimeError("\nPrimitive variables matrix does not have enough cells in the %s-direction.\nPlease "
                               "adjust the matrix or change the value of %s in your grid parameters dictionary."
                               % (axis, key))
    for _, lo, hi, axis in _AXES[:nd]:
        for key in (lo, hi):
            if key not in gridPar.keys():
                raise RuntimeError("\nYou did not specify " + key + " in your grid parameters dictionary.")
            if type(gridPar[key]) != int and type(gridPar[key]) != float:  # noqa: E721 -- the reference's test
                raise RuntimeError("\nYou passed an invalid data type for " + key + " in your grid parameters "
                                   "dictionary.\nThe %s-coordinates must be numeric (an integer or a float)." % axis)
        if gridPar[lo] >= gridPar[hi]:
            raise RuntimeError("\nInvalid values for " + lo + " and " + hi + ".\nThe minimum %s-coordinate must be "
                               "smaller than the maximum %s-coordinate." % (axis, axis))


def formatPrimVars(primVars, isMHD):
    """Source/SimulationGrid_HelperModule.py:209-270: 1D hydro (rho, vx, p) and 1D/2D
    hydro (rho, vx, vy, p) padded to the five hydro variables; MHD needs exactly
    eight."""
    nv = primVars.shape[0]
    nd = len(primVars.shape) - 1
    if nv == 3 and not isMHD and nd == 1:
        out = np.zeros((5, primVars.shape[1]))
        out[0], out[1], out[4] = primVars[0], primVars[1], primVars[2]
        return out
    if nv == 4 and not isMHD and nd in (1, 2):
        # (the reference allocates (5, shape[1]): a 2D matrix would not fit, and it raises)
        out = np.zeros((5,) + primVars.shape[1:])
        out[0], out[1], out[2], out[4] = primVars[0], primVars[1], primVars[2], primVars[3]
        return out
    if nv == 5 and not isMHD:
        return np.array(primVars, dtype=np.float64, copy=True)
    if nv > 5 and not isMHD:
        raise RuntimeError("\nYou cannot have more than five variables (dens, vx, vy, vz, pres) for each cell in a "
                           "hydrodynamic simulation.\nTo include magnetic fields set " + C.IS_MHD + " = True in your "
                           "simulation parameters dictionary.")
    if not isMHD and ((nv < 3 and nd == 1) or (nv < 4 and nd == 2) or (nv < 5 and nd == 3)):
        raise RuntimeError("\nToo few variables for each cell in a %dD hydrodynamic simulation." % nd)
    if isMHD and nv > 8:
        raise RuntimeError("\nYou provided more than eight variables for the cells in your MHD simulation; an MHD "
                           "simulation needs exactly eight (dens, vx, vy, vz, Bx, By, Bz, pres).")
    if isMHD and nv < 8:
        raise RuntimeError("\nYou provided fewer than eight variables for the cells in your MHD simulation; an MHD "
                           "simulation needs exactly eight (dens, vx, vy, vz, Bx, By, Bz, pres).\nFor a "
                           "hydrodynamic simulation set " + C.IS_MHD + " = False.")
    return np.array(primVars, dtype=np.float64, copy=True)


def _where_text(idx, nd, limit):
    n = min(idx[0].shape[0], limit)
    names = ('x_index',) if nd == 1 else ('y_index', 'x_index', 'z_index')[:nd]
    return ', '.join('(' + ', '.join('%s = %d' % (names[a], idx[a][k]) for a in range(nd)) + ')' for k in range(n))


def validatePrimVars(primVars, gamma, minDens, minPres, minEnergy):
    """Source/SimulationGrid_HelperModule.py:281-407: density, pressure and the
    total energy formed from them must be non-negative and not below their floors
    (checked in that order; the first failure raises)."""
    nd = len(primVars.shape) - 1
    mhd = primVars.shape[0] == 8
    w = primVars
    energy = 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3])
    if mhd:
        energy = energy + w[7] / (gamma - 1.0)
        energy = energy + 0.5 * (w[4] * w[4] + w[5] * w[5] + w[6] * w[6])
    else:
        energy = energy + w[4] / (gamma - 1.0)
    checks = ((w[0], minDens, C.DENSITY, C.PRIMITIVE), (w[7] if mhd else w[4], minPres, C.PRESSURE, C.PRIMITIVE),
              (energy, minEnergy, C.ENERGY, C.CONSERVATIVE))
    for var, lo, name, kind in checks:
        if np.any(var < 0):
            raise RuntimeError("You have at least one negative %s value in your %s variables matrix.\n The negative "
                               "values are at: %s" % (name, kind, _where_text(np.where(var < 0), nd,
                                                                              C.MAX_ERROR_CELLS_TO_PRINT)))
        if np.any(var < lo):
            raise RuntimeError("\nYou have at least one %s value in your %s variables matrix that is less than \nthe "
                               "minimum %s value for the simulation (%s), at: %s"
                               % (name, kind, name, lo, _where_text(np.where(var < lo), nd,
                                                                    C.MAX_ERROR_CELLS_TO_PRINT)))


class SimulationGrid:
    """Source/SimulationGrid.py:47-169 (Cartesian grids)."""

    def __init__(self, primVars, gridPar, isMHD, gamma, minDens, minPres, minEnergy, consVars, faceBx, faceBy,
                 faceBz):
        primVars = np.squeeze(primVars)
        validateGridPar(primVars, gridPar)
        self.isMHD = isMHD
        primVars = formatPrimVars(primVars, isMHD)
        validatePrimVars(primVars, gamma, minDens, minPres, minEnergy)
        self.nDim = len(primVars.shape) - 1
        self.primVars = primVars
        if np.asarray(consVars).shape[0] > 1:
            self.consVars = np.asarray(consVars, dtype=np.float64)
        else:
            self.consVars = primToCons_mhd_allVars(primVars, gamma) if isMHD else primToCons_hydro(primVars, gamma)
        self.minX, self.maxX = gridPar[C.MIN_X], gridPar[C.MAX_X]
        self.numXCells = gridPar[C.NUM_X_CELLS]
        self.numYCells = 1
        self.numZCells = 1
        if self.nDim >= 2:
            self.minY, self.maxY = gridPar[C.MIN_Y], gridPar[C.MAX_Y]
            self.numYCells = gridPar[C.NUM_Y_CELLS]
        if self.nDim == 3:
            self.minZ, self.maxZ = gridPar[C.MIN_Z], gridPar[C.MAX_Z]
            self.numZCells = gridPar[C.NUM_Z_CELLS]
        self.dx = (self.maxX - self.minX) / self.numXCells
        if self.nDim >= 2:
            self.dy = (self.maxY - self.minY) / self.numYCells
        if self.nDim == 3:
            self.dz = (self.maxZ - self.minZ) / self.numZCells
        if self.nDim == 1:
            self.xCoords = createCoordinates1D_Cartesian(gridPar)
        elif self.nDim == 2:
            self.xCoords, self.yCoords = createCoordinates2D_Cartesian(gridPar)
        else:
            self.xCoords, self.yCoords, self.zCoords = createCoordinates3D_Cartesian(gridPar)
        # a boundary condition other than periodic is outflow (SimulationGrid.py:126-139)
        self.BcX = C.PERIODIC if gridPar.get(C.BC_X, C.DEFAULT_BC_X) == C.PERIODIC else C.OUTFLOW
        self.BcY = C.PERIODIC if gridPar.get(C.BC_Y, C.DEFAULT_BC_Y) == C.PERIODIC else C.OUTFLOW
        self.BcZ = C.PERIODIC if gridPar.get(C.BC_Z, C.DEFAULT_BC_Z) == C.PERIODIC else C.OUTFLOW
        if isMHD and self.nDim == 2:
            self.faceBx = self._given(faceBx) if self._has(faceBx) else ct.getFaceBx_2D(primVars[4], self.BcX)
            self.faceBy = self._given(faceBy) if self._has(faceBy) else ct.getFaceBy_2D(primVars[5], self.BcY)
        elif isMHD and self.nDim == 3:
            self.faceBx = self._given(faceBx) if self._has(faceBx) else ct.getFaceBx_3D(primVars[4], self.BcX)
            self.faceBy = self._given(faceBy) if self._has(faceBy) else ct.getFaceBy_3D(primVars[5], self.BcY)
            self.faceBz = self._given(faceBz) if self._has(faceBz) else ct.getFaceBz_3D(primVars[6], self.BcZ)

    @staticmethod
    def _has(f):
        # the reference's test for a face array passed by the caller (shape[0] > 1)
        return np.asarray(f).shape[0] > 1

    @staticmethod
    def _given(f):
        return np.asarray(f, dtype=np.float64)

    def faces(self):
        """The face arrays the grid has (empty for hydro and 1D)."""
        if not self.isMHD or self.nDim == 1:
            return ()
        return (self.faceBx, self.faceBy) if self.nDim == 2 else (self.faceBx, self.faceBy, self.faceBz)

    def widths(self):
        return (self.dx,) if self.nDim == 1 else ((self.dx, self.dy) if self.nDim == 2 else (self.dx, self.dy, self.dz))
