"""FP64 NumPy restatement of PythonMHD's hydrodynamic CTU step (the hydro oracle).

TEST INFRASTRUCTURE ONLY -- used by tests/ and never by the product path.

State: five variables stored unrotated, primitive (rho, v_x, v_y, v_z, p) and
conserved (rho, M_x, M_y, M_z, E).  Sweep direction d (0 = x, 1 = y, 2 = z) has
its normal velocity in slot n = 1 + d.  Unlike the MHD oracle, which works in a
rotated frame, every eigen-system here is built in the storage order and every
projection sums over the storage index 0..4 *including* structurally zero
entries, exactly as the reference's loops do
(Source/Reconstruction_PPM_Hydro_HelperModule.py:178-210,
Source/RiemannSolvers_Hydro_Roe_HelperModule.py:282-308).

Parity of each part (what pins it):
  calc_dt          golden vectors of calcDt_{1,2,3}D_hydro; closed form on uniform states
  eigen            L.R = I and A.R = R.Lambda with the closed-form primitive Jacobian
  recon_states     order 2: golden vectors of ppmReconstruct{X,Y,Z}_hydro;
                   order 1 (PLM, absent from the reference): invariants only --
                   parity unpinned against the reference, which has none
  roe_flux         golden vectors of riemannSolver{X,Y,Z}_hydro_roe; consistency,
                   supersonic upwinding, the Roe property (exact for the hydro
                   linearisation)
  evolve_{1,2,3}d  golden trajectories of evolve_{1,2,3}D_hydro

Deviations (inputs where the reference raises instead of returning a result):
  * 1D periodic PPM: the reference's final ``np.append`` lacks a reshape
    (Source/Reconstruction_PPM_Hydro_HelperModule.py:355) and raises ValueError;
    here the periodic wrap of the other directions is used.
  * 3D with outflow x: the reference reshapes the z fluxes to a y-face shape
    (Source/EvolveFunctions_Hydro.py:717) and raises ValueError; here the edge
    column nx-1 is used, as for every other outflow ghost.
  * reconstructOrder 1 runs first order in the reference (only 2 selects PPM,
    Source/EvolveFunctions_Hydro.py:272, 408, 628); here it selects PLM.
"""
import numpy as np

from .mhd import FOUR_THIRDS, OUTFLOW, SMALL_VALUE, _recon_ghost_index, _take, apply_floors, copy_states

NVAR = 5
# kinetic-energy summation order of the x / y / z conversions
# (Source/PrimCons_HelperModule.py:56-190: v_n first, then the cyclic transverse pair)
KE_ORDER = ((1, 2, 3), (2, 3, 1), (3, 1, 2))
# Roe waves 1..3 of each direction: a velocity slot k (shear wave carrying v_k) or
# 'e' (entropy wave); Source/RiemannSolvers_Hydro_Roe_HelperModule.py:262-284 (x),
# :517-539 (y), :771-793 (z)
ROE_MIDDLE = ((2, 3, 'e'), ('e', 3, 1), (2, 'e', 1))


def prim_to_cons(w, gamma, d=0):
    """Source/PrimCons_HelperModule.py:23-110 (d selects the _x/_y/_z sum order)."""
    a, b, c = KE_ORDER[d]
    u = np.zeros_like(w)
    u[0] = w[0]
    u[1] = w[0] * w[1]
    u[2] = w[0] * w[2]
    u[3] = w[0] * w[3]
    u[4] = w[4] / (gamma - 1.0) + 0.5 * w[0] * (w[a] * w[a] + w[b] * w[b] + w[c] * w[c])
    return u


def cons_to_prim(u, gamma, d=0):
    """Source/PrimCons_HelperModule.py:112-190."""
    a, b, c = KE_ORDER[d]
    w = np.zeros_like(u)
    inv = 1.0 / u[0]
    w[0] = u[0]
    w[1] = u[1] * inv
    w[2] = u[2] * inv
    w[3] = u[3] * inv
    w[4] = u[4] - 0.5 * (u[a] * u[a] + u[b] * u[b] + u[c] * u[c]) * inv
    w[4] = w[4] * (gamma - 1.0)
    return w


# ---------------------------------------------------------------------------
# CFL timestep (Source/TimestepCalculations_HelperModule.py:24-170)
# ---------------------------------------------------------------------------
def max_speeds(cons, gamma, nd):
    inv = 1.0 / cons[0]
    v = (cons[1] * inv, cons[2] * inv, cons[3] * inv)
    vsq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
    pres = np.maximum((gamma - 1.0) * (cons[4] - 0.5 * cons[0] * vsq), SMALL_VALUE)
    cs = np.sqrt(gamma * pres * inv)
    return [np.max(np.abs(v[d]) + cs) for d in range(nd)]


def dt_from_speeds(speeds, cfl, widths):
    """Per-direction dt, combined with nested minima as the reference does."""
    dts = [cfl / (s / w) for s, w in zip(speeds, widths)]
    if len(dts) == 1:
        return dts[0]
    if len(dts) == 2:
        return np.minimum(dts[0], dts[1])
    return np.minimum(dts[0], np.minimum(dts[1], dts[2]))


def calc_dt(cons, gamma, cfl, widths):
    return dt_from_speeds(max_speeds(cons, gamma, len(widths)), cfl, widths)


# ---------------------------------------------------------------------------
# primitive eigen-system (Source/ReconstructionEigenmatrices_Hydro_HelperModule.py:24-280)
# wave 0: v_n - c; slot n: entropy; the other two middle slots: shear of v_k; 4: v_n + c
# ---------------------------------------------------------------------------
def eigen(w, gamma, d):
    n = 1 + d
    shp = w.shape[1:]
    c2 = gamma * w[4] / w[0]
    c = np.sqrt(c2)
    lam = np.empty((NVAR,) + shp)
    lam[0] = w[n] - c
    lam[1] = w[n]
    lam[2] = w[n]
    lam[3] = w[n]
    lam[4] = w[n] + c
    L = np.zeros((NVAR, NVAR) + shp)
    R = np.zeros((NVAR, NVAR) + shp)
    R[0, 0] = 1.0
    R[n, 0] = -c / w[0]
    R[4, 0] = c2
    R[0, n] = 1.0
    R[0, 4] = 1.0
    R[n, 4] = -R[n, 0]
    R[4, 4] = c2
    L[0, n] = -0.5 * w[0] / c
    L[0, 4] = 0.5 / c2
    L[n, 0] = 1.0
    L[n, 4] = -1.0 / c2
    L[4, n] = -L[0, n]
    L[4, 4] = L[0, 4]
    for k in (1, 2, 3):
        if k != n:
            R[k, k] = 1.0
            L[k, k] = 1.0
    return lam, L, R


def _project(L, v):
    """out[i] = sum over j = 0..4 of L[i, j] * v[j], accumulated from zero."""
    out = np.zeros_like(v)
    for i in range(NVAR):
        acc = np.zeros(v.shape[1:])
        for j in range(NVAR):
            acc = acc + L[i, j] * v[j]
        out[i] = acc
    return out


def char_limited_slopes(cent, left, right, L, R):
    """Characteristic-limited primitive slopes (Colella 1990; Stone et al. 2008 sec 4);
    Source/Reconstruction_PPM_Hydro_HelperModule.py:166-214."""
    dc = right - left
    dl = cent - left
    dr = right - cent
    with np.errstate(divide='ignore', invalid='ignore'):
        wgt = np.where(dl * dr > 0.0, (2.0 * dl * dr) / (dl + dr), 0.0)
    cc = _project(L, dc)
    lc = _project(L, dl)
    rc = _project(L, dr)
    wc = _project(L, wgt)
    lim1 = np.minimum(np.abs(lc), np.abs(rc))
    lim2 = np.minimum(0.5 * np.abs(cc), np.abs(wc))
    s = np.where(lc * rc > 0.0, np.sign(cc) * np.minimum(2.0 * lim1, lim2), 0.0)
    ps = np.zeros_like(s)
    for i in range(NVAR):
        acc = np.zeros(s.shape[1:])
        for j in range(NVAR):
            acc = acc + s[j] * R[i, j]
        ps[i] = acc
    return ps


def recon_states(w, gamma, dt, dn, bc, d, axis, order):
    """Interface primitive states (W_L, W_R) along ``axis`` for order 2 (PPM,
    Source/Reconstruction_PPM_Hydro_HelperModule.py:86-385) or 1 (PLM, new)."""
    ax = axis + 1
    n = w.shape[ax]
    e = _take(w, _recon_ghost_index(n, bc), ax)
    m = n + 4
    sl = lambda a, b: tuple([slice(None)] * ax + [slice(a, b)])  # noqa: E731
    cent, left, right = e[sl(1, m - 1)], e[sl(0, m - 2)], e[sl(2, m)]
    lam, L, R = eigen(cent, gamma, d)
    ps = char_limited_slopes(cent, left, right, L, R)
    cS, lS, rS = ps[sl(1, n + 1)], ps[sl(0, n)], ps[sl(2, n + 2)]
    cP, lP, rP = cent[sl(1, n + 1)], left[sl(1, n + 1)], right[sl(1, n + 1)]
    if order == 2:
        la = 0.5 * (cP + lP) - (cS - lS) / 6.0
        ra = 0.5 * (rP + cP) - (rS - cS) / 6.0
        dprod = (ra - cP) * (cP - la)
        idiff = ra - la
        adiff = 6.0 * (cP - 0.5 * (la + ra))
        neg = dprod <= 0.0
        la = np.where(neg, cP, la)
        ra = np.where(neg, cP, ra)
        cl = ~neg & ((idiff * adiff) > idiff * idiff)
        la = np.where(cl, (6.0 * cP - 4.0 * ra) / 2.0, la)
        cr = ~neg & ~cl & ((idiff * adiff) < (-1.0 * (idiff * idiff)))
        ra = np.where(cr, (6.0 * cP - 4.0 * la) / 2.0, ra)
    else:
        la = cP - 0.5 * cS
        ra = cP + 0.5 * cS
    la = np.minimum(np.maximum(cP, lP), np.maximum(np.minimum(cP, lP), la))
    ra = np.minimum(np.maximum(cP, rP), np.maximum(np.minimum(cP, rP), ra))
    inner = tuple([slice(None)] * axis + [slice(1, n + 1)])
    lam = lam[(slice(None),) + inner]
    L = L[(slice(None), slice(None)) + inner]
    R = R[(slice(None), slice(None)) + inner]
    dtdn = dt / dn
    c1 = ra - la
    c2 = 6.0 * (cP - 0.5 * (la + ra)) if order == 2 else np.zeros_like(c1)
    c3 = 0.5 * np.maximum(lam[4], 0.0) * dtdn
    c4 = -0.5 * np.minimum(lam[0], 0.0) * dtdn
    la = la + c4 * (c1 + (1.0 - FOUR_THIRDS * c4) * c2)
    ra = ra - c3 * (c1 - (1.0 - FOUR_THIRDS * c3) * c2)
    # remove the contributions of waves that cannot reach each face in dt/2
    for side, edge_wave, sgn in ((la, 0, 1.0), (ra, 4, -1.0)):
        for i in range(NVAR):
            reach = (lam[i] <= 0.0) if sgn > 0 else (lam[i] >= 0.0)
            v1 = 0.5 * dtdn * lam[edge_wave]
            v2 = 0.5 * dtdn * lam[i]
            v3 = v1 - v2
            v4 = FOUR_THIRDS * (v1 * v1 - v2 * v2)
            fac = np.zeros(v1.shape)
            for j in range(NVAR):
                cj = c1[j] + c2[j] if sgn > 0 else c1[j] - c2[j]
                fac = fac + reach * L[i, j] * (v3 * cj + v4 * c2[j])
            for j in range(NVAR):
                side[j] = side[j] + fac * R[j, i]
    first = _take(la, [0], ax) if bc == OUTFLOW else _take(ra, [n - 1], ax)
    last = _take(ra, [n - 1], ax) if bc == OUTFLOW else _take(la, [0], ax)
    return np.concatenate([first, ra], axis=ax), np.concatenate([la, last], axis=ax)


def interface_states(w, gamma, dt, dn, bc, d, axis, order):
    if order == 0:
        return copy_states(w, bc, axis)
    return recon_states(w, gamma, dt, dn, bc, d, axis, order)


# ---------------------------------------------------------------------------
# Roe flux (Source/RiemannSolvers_Hydro_Roe_HelperModule.py:68-869)
# ---------------------------------------------------------------------------
def _efix(lam, eps):
    return lam + np.sign(lam) * (np.abs(lam) < 2.0 * eps) * (eps + lam * lam / (4.0 * eps))


def roe_parts(WL, WR, gamma, entropy_fix, eps, d):
    """Eigenvalues, left/right Roe eigenvectors (storage order), cell fluxes and the
    conserved states of one sweep direction."""
    n = 1 + d
    UL = prim_to_cons(WL, gamma, d)
    UR = prim_to_cons(WR, gamma, d)
    sl = np.sqrt(WL[0])
    sr = np.sqrt(WR[0])
    inv = 1.0 / (sl + sr)
    v = [None] + [(sl * WL[k] + sr * WR[k]) * inv for k in (1, 2, 3)]
    a, b, c = KE_ORDER[d]
    vsq = v[a] * v[a] + v[b] * v[b] + v[c] * v[c]
    H = ((UL[4] + WL[4]) / sl + (UR[4] + WR[4]) / sr) * inv
    g1 = gamma - 1.0
    cs2 = g1 * np.maximum(H - 0.5 * vsq, SMALL_VALUE)
    cs = np.sqrt(cs2)
    vn = v[n]
    lam = [vn - cs, vn, vn, vn, vn + cs]
    if entropy_fix:
        lam = [_efix(x, eps) for x in lam]
    shp = WL.shape[1:]
    R = np.zeros((NVAR, NVAR) + shp)
    Lm = np.zeros((NVAR, NVAR) + shp)
    nf = 0.5 / cs2
    gn = g1 / cs2
    # acoustic waves 0 and 4
    for col, sg in ((0, -1.0), (4, 1.0)):
        R[0, col] = 1.0
        for k in (1, 2, 3):
            R[k, col] = (vn - cs if sg < 0 else vn + cs) if k == n else v[k]
        R[4, col] = H - vn * cs if sg < 0 else H + vn * cs
    Lm[0, 0] = nf * (0.5 * g1 * vsq + vn * cs)
    Lm[4, 0] = nf * (0.5 * g1 * vsq - vn * cs)
    for k in (1, 2, 3):
        if k == n:
            Lm[0, k] = -nf * (g1 * vn + cs)
            Lm[4, k] = -nf * (g1 * vn - cs)
        else:
            Lm[0, k] = -nf * g1 * v[k]
            Lm[4, k] = Lm[0, k]
    Lm[0, 4] = nf * g1
    Lm[4, 4] = Lm[0, 4]
    # entropy and shear waves in the reference's slot order of this direction
    for slot, kind in zip((1, 2, 3), ROE_MIDDLE[d]):
        if kind == 'e':
            R[0, slot] = 1.0
            R[1, slot] = v[1]
            R[2, slot] = v[2]
            R[3, slot] = v[3]
            R[4, slot] = 0.5 * vsq
            Lm[slot, 0] = 1.0 - nf * g1 * vsq
            Lm[slot, 1] = gn * v[1]
            Lm[slot, 2] = gn * v[2]
            Lm[slot, 3] = gn * v[3]
            Lm[slot, 4] = -gn
        else:
            R[kind, slot] = 1.0
            R[4, slot] = v[kind]
            Lm[slot, 0] = -v[kind]
            Lm[slot, kind] = 1.0

    def cell_flux(w, u):
        f = np.empty_like(w)
        f[0] = u[n]
        for k in (1, 2, 3):
            f[k] = u[n] * w[k] + w[4] if k == n else u[n] * w[k]
        f[4] = (u[4] + w[4]) * w[n]
        return f

    return lam, Lm, R, cell_flux(WL, UL), cell_flux(WR, UR), UL, UR


def roe_flux(WL, WR, gamma, entropy_fix, eps, d):
    lam, Lm, R, FL, FR, UL, UR = roe_parts(WL, WR, gamma, entropy_fix, eps, d)
    proj = _project(Lm, UR - UL)
    amps = [0.5 * np.abs(lam[k]) * proj[k] for k in range(NVAR)]
    F = 0.5 * (FL + FR)
    for i in range(NVAR):
        for j in range(NVAR):
            F[i] = F[i] - amps[j] * R[i, j]
    F = np.where(lam[0] >= 0.0, FL, F)
    F = np.where(lam[4] <= 0.0, FR, F)
    return F


# ---------------------------------------------------------------------------
# one cycle (Source/EvolveFunctions_Hydro.py:260-304 1D, :390-515 2D, :608-805 3D)
# ---------------------------------------------------------------------------
class Params:
    def __init__(self, gamma, cfl, min_dens=1e-12, min_pres=1e-12, min_energy=1e-12,
                 order=2, entropy_fix=False, eps=0.5):
        self.gamma = gamma
        self.cfl = cfl
        self.min_dens = min_dens
        self.min_pres = min_pres
        self.min_energy = min_energy
        self.order = order
        self.entropy_fix = entropy_fix
        self.eps = eps


def _ghosted(f, ax, bc, lo):
    """Append the flux slab ``f`` a ghost layer along array axis ``ax``: at the low
    end (lo=True) or the high end; outflow repeats the edge layer, periodic wraps."""
    n = f.shape[ax]
    if lo:
        g = _take(f, [0 if bc == OUTFLOW else n - 1], ax)
        return np.concatenate([g, f], axis=ax)
    g = _take(f, [n - 1 if bc == OUTFLOW else 0], ax)
    return np.concatenate([f, g], axis=ax)


def _diff(f, ax):
    n = f.shape[ax]
    hi = tuple([slice(None)] * ax + [slice(1, n)])
    lo = tuple([slice(None)] * ax + [slice(0, n - 1)])
    return f[hi] - f[lo]


def step(prim, cons, dt, widths, bcs, P):
    """One hydro CTU cycle on a 1D, 2D or 3D grid; returns (prim, cons)."""
    g = P.gamma
    nd = prim.ndim - 1
    # array axis (with the variable axis counted) of each direction
    axes = {1: (1,), 2: (2, 1), 3: (2, 1, 3)}[nd]
    fl = lambda a: apply_floors(a, True, P.min_dens, P.min_pres, P.min_energy)  # noqa: E731
    states, cstates, F = [], [], []
    for d in range(nd):
        wl, wr = interface_states(prim, g, dt, widths[d], bcs[d], d, axes[d] - 1, P.order)
        wl, wr = fl(wl), fl(wr)
        states.append((wl, wr))
        cstates.append((prim_to_cons(wl, g, d), prim_to_cons(wr, g, d)))
        F.append(roe_flux(wl, wr, g, P.entropy_fix, P.eps, d))
    if nd >= 2:
        half = [0.5 * (dt / widths[d]) for d in range(nd)]
        newF = []
        for d in range(nd):
            ul, ur = cstates[d]
            for t in range(nd):      # transverse directions in x, y, z order
                if t == d:
                    continue
                # move the transverse flux onto the faces of direction d: the low
                # state takes the cell below the face, the high state the cell above
                ftl = _ghosted(F[t], axes[d], bcs[d], True)
                ftr = _ghosted(F[t], axes[d], bcs[d], False)
                ul = ul - half[t] * _diff(ftl, axes[t])
                ur = ur - half[t] * _diff(ftr, axes[t])
            wl = cons_to_prim(ul, g, d)
            wr = cons_to_prim(ur, g, d)
            newF.append(roe_flux(wl, wr, g, P.entropy_fix, P.eps, d))
        F = newF
    cons = cons.copy()
    for d in range(nd):
        cons = cons - (dt / widths[d]) * _diff(F[d], axes[d])
    return cons_to_prim(cons, g, 0), cons


def evolve(prim, cons, widths, bcs, t0, t_end, cycle0, max_cycles, P):
    """evolve_{1,2,3}D_hydro: (prim, cons, t, cycle, prim2, cons2, t2)."""
    t, cyc = t0, cycle0
    sl = (prim.copy(), cons.copy(), t0)
    while t < t_end and cyc < max_cycles:
        sl = (prim.copy(), cons.copy(), t)
        dt = calc_dt(cons, P.gamma, P.cfl, widths)
        if t + dt > t_end:
            dt = t_end - t
        prim, cons = step(prim, cons, dt, widths, bcs, P)
        t += dt
        cyc += 1
    return prim, cons, t, cyc, sl[0], sl[1], sl[2]
