"""The MHD Riemann solvers as NumPy-in, NumPy-out functions with the reference's
names and signatures (Source/RiemannSolvers_MHD_Roe_HelperModule.py:74, 595, 1093),
computed on the MI355X by the step's own flux code (pmhd_riemann_flux).

    riemannSolverX_mhd_roe(leftPrimVars, rightPrimVars, leftConsVars, rightConsVars,
                           faceBx, gamma, entropyFix, epsilon) -> intercell flux

Arrays are the reference's directional 7-vectors over any face shape; the result
has the same shape.  ``riemann_flux`` adds the precision (64: the reference's
arithmetic bit for bit; 32: the FP32 unit) and the solver (ROE, HLLE, ROE_HLLE).
"""
import ctypes

import numpy as np

from ._lib import ROE, PmhdError, _ptr, load


def riemann_flux(wl, wr, ul, ur, bn, gamma, entropy_fix, epsilon, direction, precision=64, solver=ROE, lib=None):
    lib = lib if lib is not None else load()
    wl, wr, ul, ur = (np.ascontiguousarray(a, dtype=np.float64) for a in (wl, wr, ul, ur))
    if wl.shape[0] != 7 or any(a.shape != wl.shape for a in (wr, ul, ur)):
        raise ValueError('states must be (7, ...) arrays of one shape')
    face_shape = wl.shape[1:]
    bn = np.ascontiguousarray(np.broadcast_to(np.asarray(bn, dtype=np.float64), face_shape))
    n = int(np.prod(face_shape))
    out = np.empty_like(wl)
    rc = lib.pmhd_riemann_flux(int(precision), int(direction), int(solver), n, _ptr(wl), _ptr(wr), _ptr(ul),
                               _ptr(ur), _ptr(bn), float(gamma), 1 if entropy_fix else 0, float(epsilon), _ptr(out))
    if rc != 0:
        raise PmhdError('pmhd_riemann_flux failed (%d): %s' % (rc, lib.pmhd_last_error(None).decode()), rc)
    return out


def riemannSolverX_mhd_roe(leftPrimVars, rightPrimVars, leftConsVars, rightConsVars, faceBx, gamma, entropyFix,
                           epsilon):
    """Source/RiemannSolvers_MHD_Roe_HelperModule.py:74."""
    return riemann_flux(leftPrimVars, rightPrimVars, leftConsVars, rightConsVars, faceBx, gamma, entropyFix,
                        epsilon, 0)


def riemannSolverY_mhd_roe(topPrimVars, bottomPrimVars, topConsVars, bottomConsVars, faceBy, gamma, entropyFix,
                           epsilon):
    """Source/RiemannSolvers_MHD_Roe_HelperModule.py:595."""
    return riemann_flux(topPrimVars, bottomPrimVars, topConsVars, bottomConsVars, faceBy, gamma, entropyFix,
                        epsilon, 1)


def riemannSolverZ_mhd_roe(backPrimVars, forwPrimVars, backConsVars, forwConsVars, faceBz, gamma, entropyFix,
                           epsilon):
    """Source/RiemannSolvers_MHD_Roe_HelperModule.py:1093."""
    return riemann_flux(backPrimVars, forwPrimVars, backConsVars, forwConsVars, faceBz, gamma, entropyFix,
                        epsilon, 2)
