"""Initialisation helpers of init files (Source/Initialization_HelperModule.py):
cell-centre coordinates of a Cartesian grid, with the reference's checks of the
grid parameters, and (new) face fields from a vector potential whose discrete
divergence is checked.

Init files import this module in place of Source.Initialization_HelperModule.
"""
import numpy as np

from . import constants as C
from .ct import div_b


def _need_cells(gridPar, key, nd):
    if key not in gridPar.keys():
        raise RuntimeError("\nYou did not specify " + key + " in your grid parameters dictionary.")
    if not isinstance(gridPar[key], int):
        raise RuntimeError("\nThe " + key + " value in your grid parameters dictionary is not an integer.")
    if gridPar[key] < 1:
        raise RuntimeError("\nThe " + key + " value in your grid parameters dictionary is too small for a %dD "
                           "simulation." % nd)


def _need_range(gridPar, lo, hi, axis):
    for key, what in ((lo, 'minimum'), (hi, 'maximum')):
        if key not in gridPar.keys():
            raise RuntimeError("\nYou did not specify " + key + " in your grid parameters dictionary.")
        if type(gridPar[key]) != int and type(gridPar[key]) != float:  # noqa: E721 -- the reference's type test
            raise RuntimeError("\nYou passed an invalid data type for " + key + " in your grid parameters dictionary."
                               "\nThe %s %s-coordinate must be a numeric value (an integer or a float)." % (what, axis))
    if gridPar[lo] >= gridPar[hi]:
        raise RuntimeError("\nInvalid values for " + lo + " and " + hi + ".\nThe minimum %s-coordinate must be "
                           "smaller than the maximum." % axis)


def _centres(gridPar, n_key, lo_key, hi_key):
    """Cell centres min + dx/2, min + 3dx/2, ... as np.r_ steps them (start, stop
    max + dx/2, step dx), the reference's construction."""
    d = (gridPar[hi_key] - gridPar[lo_key]) / gridPar[n_key]
    return np.r_[gridPar[lo_key] + d / 2.0:gridPar[hi_key] + d / 2.0:d].reshape(-1)


def createCoordinates1D_Cartesian(gridPar):
    """Source/Initialization_HelperModule.py (1D): x centres, shape (nx,)."""
    _need_cells(gridPar, C.NUM_X_CELLS, 1)
    _need_range(gridPar, C.MIN_X, C.MAX_X, 'x')
    return _centres(gridPar, C.NUM_X_CELLS, C.MIN_X, C.MAX_X)


def createCoordinates2D_Cartesian(gridPar):
    """(xCoords, yCoords), each (ny, nx)."""
    for k in (C.NUM_X_CELLS, C.NUM_Y_CELLS):
        _need_cells(gridPar, k, 2)
    _need_range(gridPar, C.MIN_X, C.MAX_X, 'x')
    _need_range(gridPar, C.MIN_Y, C.MAX_Y, 'y')
    nx, ny = gridPar[C.NUM_X_CELLS], gridPar[C.NUM_Y_CELLS]
    x = _centres(gridPar, C.NUM_X_CELLS, C.MIN_X, C.MAX_X)
    y = _centres(gridPar, C.NUM_Y_CELLS, C.MIN_Y, C.MAX_Y)
    return np.repeat(x.reshape(1, -1), ny, axis=0), np.repeat(y.reshape(-1, 1), nx, axis=1)


def createCoordinates3D_Cartesian(gridPar):
    """(xCoords, yCoords, zCoords), each (ny, nx, nz)."""
    for k in (C.NUM_X_CELLS, C.NUM_Y_CELLS, C.NUM_Z_CELLS):
        _need_cells(gridPar, k, 3)
    _need_range(gridPar, C.MIN_X, C.MAX_X, 'x')
    _need_range(gridPar, C.MIN_Y, C.MAX_Y, 'y')
    _need_range(gridPar, C.MIN_Z, C.MAX_Z, 'z')
    nx, ny, nz = gridPar[C.NUM_X_CELLS], gridPar[C.NUM_Y_CELLS], gridPar[C.NUM_Z_CELLS]
    x = _centres(gridPar, C.NUM_X_CELLS, C.MIN_X, C.MAX_X).reshape(1, -1, 1)
    y = _centres(gridPar, C.NUM_Y_CELLS, C.MIN_Y, C.MAX_Y).reshape(-1, 1, 1)
    z = _centres(gridPar, C.NUM_Z_CELLS, C.MIN_Z, C.MAX_Z).reshape(1, 1, -1)
    shape = (ny, nx, nz)
    return np.broadcast_to(x, shape).copy(), np.broadcast_to(y, shape).copy(), np.broadcast_to(z, shape).copy()


# ---------------------------------------------------------------------------
# face fields from a vector potential (no counterpart in the reference, whose
# init files average cell B to faces): the curl of a potential sampled on the
# grid's corners (2D) or edges (3D) has zero discrete divergence, so a
# constrained-transport run starts divergence-free; the divergence is checked
# ---------------------------------------------------------------------------
def face_b_from_potential_2d(Az, dx, dy, check=1e-12):
    """Faces of B = curl(A_z e_z) from A_z at the cell corners, shape (ny+1, nx+1)
    (corner (j, i) at (x_i - dx/2, y_j - dy/2)): faceBx = dA/dy, faceBy = -dA/dx.
    Raises ValueError when the dimensionless divergence exceeds ``check``."""
    Az = np.asarray(Az, dtype=np.float64)
    fbx = (Az[1:, :] - Az[:-1, :]) / dy
    fby = -(Az[:, 1:] - Az[:, :-1]) / dx
    if check is not None:
        d = div_b((fbx, fby), (dx, dy))
        if d > check:
            raise ValueError('face field from the potential has div B %.3e > %.3e' % (d, check))
    return fbx, fby


def face_b_from_potential_3d(Ax, Ay, Az, dx, dy, dz, check=1e-12):
    """Faces of B = curl A from A on the cell edges: Ax (ny+1, nx, nz+1), Ay (ny, nx+1,
    nz+1), Az (ny+1, nx+1, nz), each component along its own axis through the cell
    corners.  Returns (faceBx, faceBy, faceBz); raises ValueError when the
    dimensionless divergence exceeds ``check``."""
    Ax, Ay, Az = (np.asarray(a, dtype=np.float64) for a in (Ax, Ay, Az))
    fbx = (Az[1:, :, :] - Az[:-1, :, :]) / dy - (Ay[:, :, 1:] - Ay[:, :, :-1]) / dz
    fby = (Ax[:, :, 1:] - Ax[:, :, :-1]) / dz - (Az[:, 1:, :] - Az[:, :-1, :]) / dx
    fbz = (Ay[:, 1:, :] - Ay[:, :-1, :]) / dx - (Ax[1:, :, :] - Ax[:-1, :, :]) / dy
    if check is not None:
        d = div_b((fbx, fby, fbz), (dx, dy, dz))
        if d > check:
            raise ValueError('face field from the potential has div B %.3e > %.3e' % (d, check))
    return fbx, fby, fbz
