"""Primitive -> conservative conversion of the initialisation path
(Source/PrimCons_HelperModule.py:23-37 hydro, :314-340 MHD) with the reference's
sum order, so a SimulationGrid built here holds the same consVars bits as the
reference's.  Host code (NumPy); the step itself converts on the device.
"""
import numpy as np


def primToCons_mhd_allVars(primVarsAll, gamma):
    """(rho, v, B, p) -> (rho, M, B, E) with E = p/(gamma-1) + rho v^2/2 + B^2/2,
    kinetic term summed x, y, z and added before the magnetic term."""
    w = np.asarray(primVarsAll, dtype=np.float64)
    if w.shape[0] != 8:
        raise AssertionError('primToCons_mhd_allVars needs the eight MHD variables')
    u = np.empty_like(w)
    u[0] = w[0]
    u[1:4] = w[0] * w[1:4]
    u[4:7] = w[4:7]
    e = w[7] / (gamma - 1.0) + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3])
    u[7] = e + 0.5 * (w[4] * w[4] + w[5] * w[5] + w[6] * w[6])
    return u


def primToCons_hydro(primVars, gamma):
    """(rho, v, p) -> (rho, M, E), E = p/(gamma-1) + rho v^2/2."""
    w = np.asarray(primVars, dtype=np.float64)
    u = np.empty_like(w)
    u[0] = w[0]
    u[1:4] = w[0] * w[1:4]
    u[4] = (w[4] / (gamma - 1.0)) + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3])
    return u
