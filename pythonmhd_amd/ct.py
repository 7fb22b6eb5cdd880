"""Cell <-> face magnetic field of the initialisation path (Source/CT_HelperModule.py:29-136).

A face value is the mean of the two cells it separates; the ghost cell beyond an
outflow boundary repeats the edge cell, beyond a periodic boundary it is the
opposite edge cell.  Face arrays: 2D faceBx (ny, nx+1), faceBy (ny+1, nx); 3D
faceBx (ny, nx+1, nz), faceBy (ny+1, nx, nz), faceBz (ny, nx, nz+1).  Host code
(NumPy), run once when a SimulationGrid is built; the same names as the
reference module so init files and scripts can call them.
"""
import numpy as np

from . import constants as C


def _face_average(cell, axis, bc):
    """Mean of the cells on both sides of every face along ``axis`` (n+1 faces)."""
    n = cell.shape[axis]
    lo_ghost = np.take(cell, [0 if bc == C.OUTFLOW else n - 1], axis=axis)
    hi_ghost = np.take(cell, [n - 1 if bc == C.OUTFLOW else 0], axis=axis)
    left = np.concatenate([lo_ghost, cell], axis=axis)
    right = np.concatenate([cell, hi_ghost], axis=axis)
    return (left + right) / 2.0


def getFaceBx_2D(centBx, BcX):
    """Source/CT_HelperModule.py:29-44: x faces of a (ny, nx) cell field."""
    return _face_average(np.asarray(centBx, dtype=np.float64), 1, BcX)


def getFaceBy_2D(centBy, BcY):
    """Source/CT_HelperModule.py:63-78: y faces of a (ny, nx) cell field."""
    return _face_average(np.asarray(centBy, dtype=np.float64), 0, BcY)


def getFaceBx_3D(centBx, BcX):
    """Source/CT_HelperModule.py:46-61: x faces of a (ny, nx, nz) cell field."""
    return _face_average(np.asarray(centBx, dtype=np.float64), 1, BcX)


def getFaceBy_3D(centBy, BcY):
    """Source/CT_HelperModule.py:80-95."""
    return _face_average(np.asarray(centBy, dtype=np.float64), 0, BcY)


def getFaceBz_3D(centBz, BcZ):
    """Source/CT_HelperModule.py:97-103."""
    return _face_average(np.asarray(centBz, dtype=np.float64), 2, BcZ)


def _cell_average(face, axis):
    n = face.shape[axis]
    return (np.take(face, range(0, n - 1), axis=axis) + np.take(face, range(1, n), axis=axis)) / 2.0


def getCentBx(faceBx):
    """Source/CT_HelperModule.py:110-115: cell B_x as the mean of the cell's x faces."""
    return _cell_average(np.asarray(faceBx, dtype=np.float64), 1)


def getCentBy(faceBy):
    """Source/CT_HelperModule.py:118-123."""
    return _cell_average(np.asarray(faceBy, dtype=np.float64), 0)


def getCentBz(faceBz):
    """Source/CT_HelperModule.py:126-131."""
    return _cell_average(np.asarray(faceBz, dtype=np.float64), 2)


def div_b(faces, widths):
    """max |div B| * min(width) / max |B| of a face field (2 or 3 face arrays in the
    reference layout; widths (dx, dy[, dz])): dimensionless, 0 for a field built from
    a vector potential up to rounding."""
    if len(faces) == 2:
        fbx, fby = faces
        d = (fbx[:, 1:] - fbx[:, :-1]) / widths[0] + (fby[1:, :] - fby[:-1, :]) / widths[1]
    else:
        fbx, fby, fbz = faces
        d = ((fbx[:, 1:, :] - fbx[:, :-1, :]) / widths[0] + (fby[1:, :, :] - fby[:-1, :, :]) / widths[1]
             + (fbz[:, :, 1:] - fbz[:, :, :-1]) / widths[2])
    bmax = max(float(np.max(np.abs(f))) for f in faces)
    return float(np.max(np.abs(d))) * min(widths[:len(faces)]) / bmax if bmax > 0 else 0.0
