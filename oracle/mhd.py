"""FP64 NumPy restatement of PythonMHD's unsplit CTU+CT Godunov step (the oracle).

TEST INFRASTRUCTURE ONLY -- used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py; never by the product path.

Written from the mathematics and the observable behaviour of the reference; every
floating-point expression keeps the reference's evaluation order so that results
agree bit for bit (the one exception being nothing: the 80-bit ``longdouble`` term
of the x-direction CTU update is reproduced with ``np.longdouble`` as well).

The three sweep directions are handled by one rotated formulation: a directional
7-vector (rho, v_x, v_y, v_z, B_t1, B_t2, p) is read in the frame
(rho, v_n, v_t1, v_t2, B_t1, B_t2, p) with (n, t1, t2) cyclic in (x, y, z).  The
reference writes x, y and z out separately with the same operation order in that
frame (Source/Reconstruction_PPM_MHD_HelperModule.py:341-373 vs :668-700 vs
:990-1005), which is why one rotated routine reproduces all three.

Parity of each part (what pins it):
  calc_dt_*            golden vectors of calcDt_{1,2,3}D_mhd + closed form on uniform states
  eigen_prim           L.R = I and A.R = R.Lambda checks + golden PPM vectors
  ppm_states           golden vectors of ppmReconstruct{X,Y,Z}_mhd
  plm_states           NEW (the reference has no PLM): pinned by invariants only
                       (constant data exact, linear data, no new extrema) --
                       "parity unpinned" against the reference, which has none.
  roe_flux             Roe-property and consistency checks + golden vectors via the
                       full evolve step
  recon_b_corrections  golden vectors of getReconBCorrections{X,Y,Z}_3D and the 2D terms
  ctu_transverse_*     golden vectors of get{X,Y,Z}TransverseUpdates_{2D,3D}
  corner_emf_*         golden vectors of getCornerEMF_2D / get{X,Y,Z}CornerEMF_3D
  evolve_{1,2,3}d      golden trajectories of evolve_{1,2,3}D_mhd
"""
import numpy as np

SMALL_VALUE = 0.000000000001          # Source/PythonMHD_Constants.py:7
FOUR_THIRDS = 1.333333333333333       # Source/PythonMHD_Constants.py:8
OUTFLOW = 0
PERIODIC = 1

# directional 7-vector <- 8-variable state (Source/EvolveFunctions_MHD.py:775-807)
DIRVARS = {'x': (0, 1, 2, 3, 5, 6, 7), 'y': (0, 1, 2, 3, 6, 4, 7), 'z': (0, 1, 2, 3, 4, 5, 7)}
BNVAR = {'x': 4, 'y': 5, 'z': 6}
# rotated index r -> directional index (velocity slots permuted to (n, t1, t2))
ROT = {'x': (0, 1, 2, 3, 4, 5, 6), 'y': (0, 2, 3, 1, 4, 5, 6), 'z': (0, 3, 1, 2, 4, 5, 6)}
# summation order of every 7x7 projection in the rotated frame
PROJ_ORDER = (0, 1, 2, 3, 6, 4, 5)


def spatial_axis(d, ndim):
    """Array axis (without the leading variable axis) of direction d."""
    if ndim == 1:
        return 0
    if ndim == 2:
        return {'y': 0, 'x': 1}[d]
    return {'y': 0, 'x': 1, 'z': 2}[d]


def _take(a, idx, axis):
    return np.take(a, np.asarray(idx, dtype=np.int64), axis=axis)


def _rot(w, d):
    return np.stack([w[ROT[d][r]] for r in range(7)])


def _unrot(wr, d):
    out = np.empty_like(wr)
    for r in range(7):
        out[ROT[d][r]] = wr[r]
    return out


def _sign(x):
    return np.sign(x)


# ---------------------------------------------------------------------------
# primitive <-> conservative (Source/PrimCons_HelperModule.py:208-493)
# ---------------------------------------------------------------------------
_KE = {'x': (1, 2, 3), 'y': (2, 3, 1), 'z': (3, 1, 2), None: (1, 2, 3)}


def prim_to_cons(w, bn, gamma, ke=None):
    """Directional 7-vector prim -> cons; ``ke`` picks the kinetic-energy sum order
    (None: v_x, v_y, v_z as primToCons_mhd; 'y'/'z' as primToCons_mhd_y/_z)."""
    a, b, c = _KE[ke]
    u = np.zeros_like(w)
    u[0] = w[0]
    u[1] = w[0] * w[1]
    u[2] = w[0] * w[2]
    u[3] = w[0] * w[3]
    u[4] = w[4]
    u[5] = w[5]
    u[6] = w[6] / (gamma - 1.0) + 0.5 * w[0] * (w[a] * w[a] + w[b] * w[b] + w[c] * w[c])
    u[6] = u[6] + 0.5 * (bn * bn + w[4] * w[4] + w[5] * w[5])
    return u


def cons_to_prim(u, bn, gamma, ke=None):
    a, b, c = _KE[ke]
    w = np.zeros_like(u)
    inv = 1.0 / u[0]
    w[0] = u[0]
    w[1] = u[1] * inv
    w[2] = u[2] * inv
    w[3] = u[3] * inv
    w[4] = u[4]
    w[5] = u[5]
    w[6] = u[6] - 0.5 * (u[a] * u[a] + u[b] * u[b] + u[c] * u[c]) * inv
    w[6] = w[6] - 0.5 * (bn * bn + w[4] * w[4] + w[5] * w[5])
    w[6] = w[6] * (gamma - 1.0)
    return w


def prim_to_cons_all(w, gamma):
    u = np.zeros_like(w)
    u[0] = w[0]
    u[1] = w[0] * w[1]
    u[2] = w[0] * w[2]
    u[3] = w[0] * w[3]
    u[4] = w[4]
    u[5] = w[5]
    u[6] = w[6]
    u[7] = w[7] / (gamma - 1.0) + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3])
    u[7] = u[7] + 0.5 * (w[4] * w[4] + w[5] * w[5] + w[6] * w[6])
    return u


def cons_to_prim_all(u, gamma):
    w = np.zeros_like(u)
    inv = 1.0 / u[0]
    w[0] = u[0]
    w[1] = u[1] * inv
    w[2] = u[2] * inv
    w[3] = u[3] * inv
    w[4] = u[4]
    w[5] = u[5]
    w[6] = u[6]
    w[7] = u[7] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * inv
    w[7] = w[7] - 0.5 * (w[4] * w[4] + w[5] * w[5] + w[6] * w[6])
    w[7] = w[7] * (gamma - 1.0)
    return w


def apply_floors(v, is_prim, min_dens, min_pres, min_energy):
    """Source/PrimCons_HelperModule.py:499-554 (the unused message is dropped)."""
    v = v.copy()
    v[0] = np.where(v[0] < min_dens, min_dens, v[0])
    lo = min_pres if is_prim else min_energy
    v[-1] = np.where(v[-1] < lo, lo, v[-1])
    return v


# ---------------------------------------------------------------------------
# CFL timestep (Source/TimestepCalculations_HelperModule.py:183-361)
# ---------------------------------------------------------------------------
def _fast_speed_sq(bsq, inv, cs2, btsq):
    ssum = bsq * inv + cs2
    sdiff = bsq * inv - cs2
    return 0.5 * (ssum + np.sqrt(sdiff * sdiff + 4.0 * cs2 * btsq * inv))


def max_speeds(cons, faces, gamma):
    """Per-direction max(|v_d| + c_f,d) over all cells; faces = () for 1D,
    (fBx, fBy) for 2D, (fBx, fBy, fBz) for 3D."""
    g1 = gamma - 1.0
    inv = 1.0 / cons[0]
    vx = cons[1] * inv
    vy = cons[2] * inv
    vz = cons[3] * inv
    vsq = vx * vx + vy * vy + vz * vz
    nd = len(faces)
    if nd == 0:
        bx, by, bz = cons[4], cons[5], cons[6]
    elif nd == 2:
        fbx, fby = faces
        bx = cons[4] + np.abs(fbx[:, :-1] - cons[4])
        by = cons[5] + np.abs(fby[:-1, :] - cons[5])
        bz = cons[6]
    else:
        fbx, fby, fbz = faces
        bx = cons[4] + np.abs(fbx[:, :-1, :] - cons[4])
        by = cons[5] + np.abs(fby[:-1, :, :] - cons[5])
        bz = cons[6] + np.abs(fbz[:, :, :-1] - cons[6])
    bsq = bx * bx + by * by + bz * bz
    pres = np.maximum(g1 * (cons[7] - 0.5 * cons[0] * vsq - 0.5 * bsq), SMALL_VALUE)
    cs2 = gamma * pres * inv
    out = [np.max(np.abs(vx) + np.sqrt(_fast_speed_sq(bsq, inv, cs2, by * by + bz * bz)))]
    if nd >= 2:
        out.append(np.max(np.abs(vy) + np.sqrt(_fast_speed_sq(bsq, inv, cs2, bx * bx + bz * bz))))
    if nd == 3:
        out.append(np.max(np.abs(vz) + np.sqrt(_fast_speed_sq(bsq, inv, cs2, bx * bx + by * by))))
    return out


def dt_from_speeds(speeds, cfl, widths):
    """Combine per-direction maxima exactly as the reference does."""
    if len(speeds) == 1:
        return cfl / (speeds[0] / widths[0])
    if len(speeds) == 2:
        return cfl / np.maximum(speeds[0] / widths[0], speeds[1] / widths[1])
    return cfl / np.maximum(speeds[0] / widths[0],
                            np.maximum(speeds[1] / widths[1], speeds[2] / widths[2]))


def calc_dt(cons, faces, gamma, cfl, widths):
    return dt_from_speeds(max_speeds(cons, faces, gamma), cfl, widths)


# ---------------------------------------------------------------------------
# primitive eigen-system in the rotated frame
# (Source/ReconstructionEigenmatrices_MHD_HelperModule.py:27-207, rotated)
# ---------------------------------------------------------------------------
def eigen_prim(wr, bn, gamma):
    """wr: rotated primitive 7-vector; bn: cell-centred normal B.
    Returns (lam[7], L[7,7], R[7,7]) in rotated indexing."""
    rho, vn, bt1, bt2, p = wr[0], wr[1], wr[4], wr[5], wr[6]
    shp = rho.shape
    btsq = bt1 * bt1 + bt2 * bt2
    bt = np.sqrt(btsq)
    gp = gamma * p
    inv = 1.0 / rho
    sqd = np.sqrt(rho)
    rbtsq = btsq * inv
    cs2 = gp * inv
    cs = np.sqrt(cs2)
    ca2 = bn * bn * inv
    ca = np.sqrt(ca2)
    ssum = ca2 + rbtsq + cs2
    sdiff = ca2 + rbtsq - cs2
    ssqrt = np.sqrt(sdiff * sdiff + 4.0 * cs2 * rbtsq)
    cf2 = 0.5 * (ssum + ssqrt)
    cf = np.sqrt(cf2)
    csl2 = cs2 * ca2 / cf2
    csl = np.sqrt(csl2)
    with np.errstate(divide='ignore', invalid='ignore'):
        b1 = np.where(bt == 0.0, 1.0, np.where(bt > 0.0, bt1 / bt, 0.0))
        b2 = np.where(bt > 0.0, bt2 / bt, 0.0)
        fa1 = ssqrt == 0.0
        sa1 = np.logical_and(~fa1, (cs2 - csl2) <= 0.0)
        fa2 = np.logical_and(~(fa1 | sa1), (cf2 - cs2) <= 0.0)
        calc = ~(fa1 | sa1 | fa2)
        af = np.where(calc, np.sqrt((cs2 - csl2) / ssqrt), np.where(fa1 | fa2, 1.0, 0.0))
        als = np.where(calc, np.sqrt((cf2 - cs2) / ssqrt), np.where(sa1, 1.0, 0.0))
    sgn = _sign(bn)
    sgn = np.where(sgn == 0.0, 1.0, sgn)
    qf = cf * af * sgn
    qs = csl * als * sgn
    a_f = cs * af * sqd
    a_s = cs * als * sqd
    lam = np.stack([vn - cf, vn - ca, vn - csl, vn, vn + csl, vn + ca, vn + cf])
    R = np.zeros((7, 7) + shp)
    R[0, 0] = rho * af
    R[0, 2] = rho * als
    R[0, 3] = 1.0
    R[0, 4] = R[0, 2]
    R[0, 6] = R[0, 0]
    R[1, 0] = -cf * af
    R[1, 2] = -csl * als
    R[1, 4] = -R[1, 2]
    R[1, 6] = -R[1, 0]
    R[2, 0] = qs * b1
    R[2, 1] = -b2
    R[2, 2] = -qf * b1
    R[2, 4] = -R[2, 2]
    R[2, 5] = b2
    R[2, 6] = -R[2, 0]
    R[3, 0] = qs * b2
    R[3, 1] = b1
    R[3, 2] = -qf * b2
    R[3, 4] = -R[3, 2]
    R[3, 5] = -b1
    R[3, 6] = -R[3, 0]
    R[6, 0] = rho * cs2 * af
    R[6, 2] = rho * cs2 * als
    R[6, 4] = R[6, 2]
    R[6, 6] = R[6, 0]
    R[4, 0] = a_s * b1
    R[4, 1] = -b2 * sgn * sqd
    R[4, 2] = -a_f * b1
    R[4, 4] = R[4, 2]
    R[4, 5] = R[4, 1]
    R[4, 6] = R[4, 0]
    R[5, 0] = a_s * b2
    R[5, 1] = b1 * sgn * sqd
    R[5, 2] = -a_f * b2
    R[5, 4] = R[5, 2]
    R[5, 5] = R[5, 1]
    R[5, 6] = R[5, 0]
    nf = 0.5 / cs2
    qf = nf * qf
    qs = nf * qs
    afp = nf * a_f * inv
    asp = nf * a_s * inv
    L = np.zeros((7, 7) + shp)
    L[0, 1] = -nf * cf * af
    L[0, 2] = qs * b1
    L[0, 3] = qs * b2
    L[0, 6] = nf * af * inv
    L[0, 4] = asp * b1
    L[0, 5] = asp * b2
    L[1, 2] = -0.5 * b2
    L[1, 3] = 0.5 * b1
    L[1, 4] = -0.5 * b2 * sgn / sqd
    L[1, 5] = 0.5 * b1 * sgn / sqd
    L[2, 1] = -1.0 * nf * csl * als
    L[2, 2] = -1.0 * qf * b1
    L[2, 3] = -1.0 * qf * b2
    L[2, 6] = nf * als * inv
    L[2, 4] = -1.0 * afp * b1
    L[2, 5] = -1.0 * afp * b2
    L[3, 0] = 1.0
    L[3, 6] = -1.0 / cs2
    L[4, 1] = -L[2, 1]
    L[4, 2] = -L[2, 2]
    L[4, 3] = -L[2, 3]
    L[4, 6] = L[2, 6]
    L[4, 4] = L[2, 4]
    L[4, 5] = L[2, 5]
    L[5, 2] = -L[1, 2]
    L[5, 3] = -L[1, 3]
    L[5, 4] = L[1, 4]
    L[5, 5] = L[1, 5]
    L[6, 1] = -L[0, 1]
    L[6, 2] = -L[0, 2]
    L[6, 3] = -L[0, 3]
    L[6, 6] = L[0, 6]
    L[6, 4] = L[0, 4]
    L[6, 5] = L[0, 5]
    return lam, L, R


def _project(L, dv):
    out = np.zeros_like(dv)
    for i in range(7):
        acc = np.zeros(dv.shape[1:])
        for j in PROJ_ORDER:
            acc = acc + L[i, j] * dv[j]
        out[i] = acc
    return out


# ---------------------------------------------------------------------------
# reconstruction (PPM: Source/Reconstruction_PPM_MHD_HelperModule.py:189-551;
# PLM: new, the linear-profile member of the same characteristic-tracing family)
# ---------------------------------------------------------------------------
def _recon_ghost_index(n, bc):
    if bc == OUTFLOW:
        return [0, 0] + list(range(n)) + [n - 1, n - 1]
    return [n - 2, n - 1] + list(range(n)) + [0, 1]


def char_limited_slopes(cent, left, right, lam, L, R):
    """Characteristic-limited primitive slopes (Colella 1990 / Stone+ 2008 sec 4)."""
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
    for i in range(7):
        acc = np.zeros(s.shape[1:])
        for j in range(7):
            acc = acc + s[j] * R[i, j]
        ps[i] = acc
    return ps


def recon_states(w_dir, bn_cell, gamma, dt, dn, bc, d, axis, order):
    """Interface states (W_L, W_R) along ``axis`` for reconstruction order 1 (PLM)
    or 2 (PPM).  w_dir: directional 7-vector (7, ...)."""
    ax = axis + 1
    n = w_dir.shape[ax]
    gi = _recon_ghost_index(n, bc)
    e = _take(_rot(w_dir, d), gi, ax)
    eb = _take(bn_cell, gi, axis)
    m = n + 4
    sl = lambda a, b, k=ax: tuple([slice(None)] * k + [slice(a, b)])  # noqa: E731
    cent = e[sl(1, m - 1)]
    left = e[sl(0, m - 2)]
    right = e[sl(2, m)]
    bc_cent = eb[sl(1, m - 1, axis)]
    lam, L, R = eigen_prim(cent, bc_cent, gamma)
    ps = char_limited_slopes(cent, left, right, lam, L, R)
    cS = ps[sl(1, n + 1)]
    lS = ps[sl(0, n)]
    rS = ps[sl(2, n + 2)]
    cP = cent[sl(1, n + 1)]
    lP = left[sl(1, n + 1)]
    rP = right[sl(1, n + 1)]
    if order == 2:
        la = 0.5 * (cP + lP) - (cS - lS) / 6.0
        ra = 0.5 * (rP + cP) - (rS - cS) / 6.0
        dprod = (ra - cP) * (cP - la)
        idiff = ra - la
        adiff = 6.0 * (cP - 0.5 * (la + ra))
        neg = dprod <= 0.0
        pos = ~neg
        la = np.where(neg, cP, la)
        ra = np.where(neg, cP, ra)
        cl = pos & ((idiff * adiff) > idiff * idiff)
        la = np.where(cl, (6.0 * cP - 4.0 * ra) / 2.0, la)
        cr = pos & ~cl & ((idiff * adiff) < (-1.0 * (idiff * idiff)))
        ra = np.where(cr, (6.0 * cP - 4.0 * la) / 2.0, ra)
    else:
        la = cP - 0.5 * cS
        ra = cP + 0.5 * cS
    la = np.maximum(np.minimum(cP, lP), la)
    la = np.minimum(np.maximum(cP, lP), la)
    ra = np.maximum(np.minimum(cP, rP), ra)
    ra = np.minimum(np.maximum(cP, rP), ra)
    ix = (slice(None),) + tuple([slice(None)] * axis + [slice(1, n + 1)])
    lam = lam[ix]
    L = L[(slice(None),) + ix]
    R = R[(slice(None),) + ix]
    dtdx = dt / dn
    c1 = ra - la
    if order == 2:
        c2 = 6.0 * (cP - 0.5 * (la + ra))
    else:
        c2 = np.zeros_like(c1)
    c3 = 0.5 * np.maximum(lam[6], 0.0) * dtdx
    c4 = -0.5 * np.minimum(lam[0], 0.0) * dtdx
    la = la + c4 * (c1 + (1.0 - FOUR_THIRDS * c4) * c2)
    ra = ra - c3 * (c1 - (1.0 - FOUR_THIRDS * c3) * c2)
    for i in range(7):
        ind = (lam[i] <= 0.0).astype(np.float64)
        v1 = 0.5 * dtdx * lam[0]
        v2 = 0.5 * dtdx * lam[i]
        v3 = v1 - v2
        v4 = FOUR_THIRDS * (v1 * v1 - v2 * v2)
        fac = np.zeros(v1.shape)
        for j in PROJ_ORDER:
            fac = fac + ind * L[i, j] * (v3 * (c1[j] + c2[j]) + v4 * c2[j])
        for j in range(7):
            la[j] = la[j] + fac * R[j, i]
    for i in range(7):
        ind = (lam[i] >= 0.0).astype(np.float64)
        v1 = 0.5 * dtdx * lam[6]
        v2 = 0.5 * dtdx * lam[i]
        v3 = v1 - v2
        v4 = FOUR_THIRDS * (v1 * v1 - v2 * v2)
        fac = np.zeros(v1.shape)
        for j in PROJ_ORDER:
            fac = fac + ind * L[i, j] * (v3 * (c1[j] - c2[j]) + v4 * c2[j])
        for j in range(7):
            ra[j] = ra[j] + fac * R[j, i]
    first = _take(la, [0], ax) if bc == OUTFLOW else _take(ra, [n - 1], ax)
    last = _take(ra, [n - 1], ax) if bc == OUTFLOW else _take(la, [0], ax)
    wl = np.concatenate([first, ra], axis=ax)
    wr = np.concatenate([la, last], axis=ax)
    return _unrot(wl, d), _unrot(wr, d)


def copy_states(w_dir, bc, axis):
    """First-order interface states (reconstructOrder 0)."""
    ax = axis + 1
    n = w_dir.shape[ax]
    g_lo = 0 if bc == OUTFLOW else n - 1
    g_hi = n - 1 if bc == OUTFLOW else 0
    wl = np.concatenate([_take(w_dir, [g_lo], ax), w_dir], axis=ax)
    wr = np.concatenate([w_dir, _take(w_dir, [g_hi], ax)], axis=ax)
    return wl, wr


def interface_states(w_dir, bn_cell, gamma, dt, dn, bc, d, axis, order):
    if order == 0:
        return copy_states(w_dir, bc, axis)
    return recon_states(w_dir, bn_cell, gamma, dt, dn, bc, d, axis, order)


# ---------------------------------------------------------------------------
# Roe flux for ideal MHD in the rotated frame
# (Source/RiemannSolvers_MHD_Roe_HelperModule.py:74-576 and rotations)
# ---------------------------------------------------------------------------
def fast_speed(w, bn, gamma):
    """Fast magnetosonic speed of rotated primitive states (normal field bn)."""
    inv = 1.0 / w[0]
    btsq = w[4] * w[4] + w[5] * w[5]
    cs2 = gamma * w[6] * inv
    bsq = bn * bn + btsq
    ssum = bsq * inv + cs2
    sdiff = bsq * inv - cs2
    return np.sqrt(0.5 * (ssum + np.sqrt(sdiff * sdiff + 4.0 * cs2 * btsq * inv)))


def hlle_flux(wl, wr, ul, ur, FL, FR, bn, gamma, lam_lo, lam_hi):
    """HLLE flux in the rotated frame (Harten-Lax-van Leer with Einfeldt's bounds:
    the smaller of the left state's and the Roe average's left-going fast speeds,
    the larger of the right state's and the Roe average's right-going ones; the
    upwind flux when both bounds have one sign).  New in the build (the reference
    names HLLE as missing, Source/RiemannSolvers_MHD_Roe_HelperModule.py:24-30)."""
    sl = np.minimum(wl[1] - fast_speed(wl, bn, gamma), lam_lo)
    sr = np.maximum(wr[1] + fast_speed(wr, bn, gamma), lam_hi)
    with np.errstate(divide='ignore', invalid='ignore'):
        inv = 1.0 / (sr - sl)
        F = (sr * FL - sl * FR + sl * sr * (ur - ul)) * inv
    F = np.where(sr <= 0.0, FR, F)
    return np.where(sl >= 0.0, FL, F)


def _cons_pressure(u, bn, gamma):
    ke = 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0]
    return (gamma - 1.0) * (u[6] - ke - 0.5 * (bn * bn + u[4] * u[4] + u[5] * u[5]))


def roe_flux(WL, WR, UL, UR, bn, gamma, entropy_fix, eps, d, solver=0):
    """Directional prim/cons 7-vectors at faces -> directional flux 7-vector.
    solver 0: the reference's Roe flux; 1: HLLE; 2: Roe, with HLLE at the faces
    whose Roe intermediate states U_L + sum_{j<=k} (L_j . dU) R_j (k = 0..5) have a
    non-positive density or pressure."""
    lam, Lm, R, FL, FR, ul, ur, lam0 = roe_parts(WL, WR, UL, UR, bn, gamma, entropy_fix, eps, d, raw=True)
    if solver == 1:
        wl, wr = _rot(WL, d), _rot(WR, d)
        return _unrot(hlle_flux(wl, wr, ul, ur, FL, FR, bn, gamma, lam0[0], lam0[6]), d)
    dU = ur - ul
    proj = _project(Lm, dU)
    amps = [0.5 * np.abs(lam[k]) * proj[k] for k in range(7)]
    F = 0.5 * (FL + FR)
    for i in range(7):
        for k in range(7):
            F[i] = F[i] - amps[k] * R[i, k]
    if solver == 2:
        us = ul.copy()
        bad = np.zeros(ul.shape[1:], dtype=bool)
        for k in range(6):
            us = us + proj[k] * R[:, k]
            with np.errstate(divide='ignore', invalid='ignore'):
                bad = bad | ~(us[0] > 0.0) | ~(_cons_pressure(us, bn, gamma) > 0.0)
        wl, wr = _rot(WL, d), _rot(WR, d)
        F = np.where(bad, hlle_flux(wl, wr, ul, ur, FL, FR, bn, gamma, lam0[0], lam0[6]), F)
    up = lam[0] >= 0.0
    F = np.where(up, FL, F)
    dn = lam[6] <= 0.0
    F = np.where(dn, FR, F)
    return _unrot(F, d)


def roe_parts(WL, WR, UL, UR, bn, gamma, entropy_fix, eps, d, raw=False):
    """Roe eigen-system in the rotated frame: (lam[7], L[7,7], R[7,7], F_L, F_R, U_L, U_R)."""
    wl = _rot(WL, d)
    wr = _rot(WR, d)
    ul = _rot(UL, d)
    ur = _rot(UR, d)
    g = gamma
    rl, rr = wl[0], wr[0]
    sl = np.sqrt(rl)
    sr = np.sqrt(rr)
    isum = 1.0 / (sl + sr)
    rd = sl * sr
    vn = (sl * wl[1] + sr * wr[1]) * isum
    vt1 = (sl * wl[2] + sr * wr[2]) * isum
    vt2 = (sl * wl[3] + sr * wr[3]) * isum
    vsq = vn * vn + vt1 * vt1 + vt2 * vt2
    bt1 = (sr * wl[4] + sl * wr[4]) * isum
    bt2 = (sr * wl[5] + sl * wr[5]) * isum
    hbl = 0.5 * (bn * bn + wl[4] * wl[4] + wl[5] * wl[5])
    hbr = 0.5 * (bn * bn + wr[4] * wr[4] + wr[5] * wr[5])
    pl, pr = wl[6], wr[6]
    H = ((ul[6] + pl + hbl) / sl + (ur[6] + pr + hbr) / sr) * isum
    ird = 1.0 / rd
    d4 = wl[4] - wr[4]
    d5 = wl[5] - wr[5]
    ssd = sl + sr
    btc = 0.5 * (d4 * d4 + d5 * d5) / (ssd * ssd)
    rhoc = 0.5 * (rl + rr) / rd
    btsq = bt1 * bt1 + bt2 * bt2
    bt = np.sqrt(btsq)
    btsqs = ((g - 1.0) - (g - 2.0) * rhoc) * btsq
    bts = np.sqrt(btsqs)
    rbtsqs = btsqs * ird
    ca2 = bn * bn * ird
    ca = np.sqrt(ca2)
    hmb = H - (ca2 + btsq * ird)
    cs2 = np.maximum((g - 1.0) * (hmb - 0.5 * vsq) - (g - 2.0) * btc, SMALL_VALUE)
    cs = np.sqrt(cs2)
    ssum = ca2 + rbtsqs + cs2
    sdiff = ca2 + rbtsqs - cs2
    ssqrt = np.sqrt(sdiff * sdiff + 4.0 * cs2 * rbtsqs)
    cf2 = 0.5 * (ssum + ssqrt)
    cf = np.sqrt(cf2)
    csl2 = cs2 * ca2 / cf2
    csl = np.sqrt(csl2)
    lam = [vn - cf, vn - ca, vn - csl, vn, vn + csl, vn + ca, vn + cf]
    lam_raw = list(lam)
    if entropy_fix:
        lam = [l + np.sign(l) * (np.abs(l) < 2.0 * eps) * (eps + l * l / (4.0 * eps)) for l in lam]
    with np.errstate(divide='ignore', invalid='ignore'):
        fa1 = (cf2 - csl2) == 0.0
        sa1 = ~fa1 & ((cs2 - csl2) <= 0.0)
        fa2 = ~(fa1 | sa1) & ((cf2 - cs2) <= 0.0)
        calc = ~(fa1 | sa1 | fa2)
        af = np.where(calc, np.sqrt((cs2 - csl2) / (cf2 - csl2)), np.where(fa1 | fa2, 1.0, 0.0))
        als = np.where(calc, np.sqrt((cf2 - cs2) / (cf2 - csl2)), np.where(sa1, 1.0, 0.0))
        sgn = np.sign(bn)
        sgn = np.where(sgn == 0, 1.0, sgn)
        srd = np.sqrt(rd)
        isrd = 1.0 / srd
        qf = cf * af * sgn
        qs = csl * als * sgn
        afp = cs * af * isrd
        asp = cs * als * isrd
        b1 = np.where(bt == 0.0, 1.0, np.where(bt > 0.0, bt1 / bt, 0.0))
        b2 = np.where(bt > 0.0, bt2 / bt, 0.0)
    b1s = b1 / np.sqrt((g - 1.0) - (g - 2.0) * rhoc)
    b2s = b2 / np.sqrt((g - 1.0) - (g - 2.0) * rhoc)
    sb = vt1 * b1s + vt2 * b2s
    bss = b1s * b1s + b2s * b2s
    afb = afp * bts * bss
    asb = asp * bts * bss
    shp = rl.shape
    R = np.zeros((7, 7) + shp)
    R[0, 0] = af
    R[0, 2] = als
    R[0, 3] = 1.0
    R[0, 4] = als
    R[0, 6] = af
    R[1, 0] = af * lam[0]
    R[1, 2] = als * lam[2]
    R[1, 3] = vn
    R[1, 4] = als * lam[4]
    R[1, 6] = af * lam[6]
    fav = af * vt1
    sav = als * vt1
    qsb = qs * b1s
    qfb = qf * b1s
    R[2, 0] = fav + qsb
    R[2, 1] = -b2
    R[2, 2] = sav - qfb
    R[2, 3] = vt1
    R[2, 4] = sav + qfb
    R[2, 5] = b2
    R[2, 6] = fav - qsb
    fav = af * vt2
    sav = als * vt2
    qsb = qs * b2s
    qfb = qf * b2s
    R[3, 0] = fav + qsb
    R[3, 1] = b1
    R[3, 2] = sav - qfb
    R[3, 3] = vt2
    R[3, 4] = sav + qfb
    R[3, 5] = -b1
    R[3, 6] = fav - qsb
    R[6, 0] = af * (hmb - vn * cf) + qs * sb + asb
    R[6, 1] = -(vt1 * b2 - vt2 * b1)
    R[6, 2] = als * (hmb - vn * csl) - qf * sb - afb
    R[6, 3] = 0.5 * vsq + (g - 2.0) * btc / (g - 1.0)
    R[6, 4] = als * (hmb + vn * csl) + qf * sb - afb
    R[6, 5] = -R[6, 1]
    R[6, 6] = af * (hmb + vn * cf) - qs * sb + asb
    R[4, 0] = asp * b1s
    R[4, 1] = -b2 * sgn * isrd
    R[4, 2] = -afp * b1s
    R[4, 4] = R[4, 2]
    R[4, 5] = R[4, 1]
    R[4, 6] = R[4, 0]
    R[5, 0] = asp * b2s
    R[5, 1] = b1 * sgn * isrd
    R[5, 2] = -afp * b2s
    R[5, 4] = R[5, 2]
    R[5, 5] = R[5, 1]
    R[5, 6] = R[5, 0]
    nf = 0.5 / cs2
    cff = nf * af * cf
    css = nf * als * csl
    qf = qf * nf
    qs = qs * nf
    a_f = nf * afp * rd
    a_s = nf * asp * rd
    afb = nf * afp * bts
    asb = nf * asp * bts
    nf = nf * (g - 1.0)
    af = af * nf
    als = als * nf
    b1c = b1s / bss
    b2c = b2s / bss
    sbs = vt1 * b1c + vt2 * b2c
    nf = nf * 2.0
    Lm = np.zeros((7, 7) + shp)
    Lm[0, 0] = af * (vsq - hmb) + cff * (cf + vn) - qs * sbs - asb
    Lm[0, 1] = -1.0 * af * vn - cff
    Lm[0, 2] = -1.0 * af * vt1 + qs * b1c
    Lm[0, 3] = -1.0 * af * vt2 + qs * b2c
    Lm[0, 6] = af
    Lm[0, 4] = a_s * b1c - af * bt1
    Lm[0, 5] = a_s * b2c - af * bt2
    Lm[1, 0] = 0.5 * (vt1 * b2 - vt2 * b1)
    Lm[1, 2] = -0.5 * b2
    Lm[1, 3] = 0.5 * b1
    Lm[1, 4] = -0.5 * srd * b2 * sgn
    Lm[1, 5] = 0.5 * srd * b1 * sgn
    Lm[2, 0] = als * (vsq - hmb) + css * (csl + vn) + qf * sbs + afb
    Lm[2, 1] = -als * vn - css
    Lm[2, 2] = -als * vt1 - qf * b1c
    Lm[2, 3] = -als * vt2 - qf * b2c
    Lm[2, 6] = als
    Lm[2, 4] = -a_f * b1c - als * bt1
    Lm[2, 5] = -a_f * b2c - als * bt2
    Lm[3, 0] = 1.0 - nf * (0.5 * vsq - (g - 2.0) * btc / (g - 1.0))
    Lm[3, 1] = nf * vn
    Lm[3, 2] = nf * vt1
    Lm[3, 3] = nf * vt2
    Lm[3, 6] = -1.0 * nf
    Lm[3, 4] = nf * bt1
    Lm[3, 5] = nf * bt2
    Lm[4, 0] = als * (vsq - hmb) + css * (csl - vn) - qf * sbs + afb
    Lm[4, 1] = -als * vn + css
    Lm[4, 2] = -als * vt1 + qf * b1c
    Lm[4, 3] = -als * vt2 + qf * b2c
    Lm[4, 6] = als
    Lm[4, 4] = Lm[2, 4]
    Lm[4, 5] = Lm[2, 5]
    Lm[5, 0] = -Lm[1, 0]
    Lm[5, 2] = -Lm[1, 2]
    Lm[5, 3] = -Lm[1, 3]
    Lm[5, 4] = Lm[1, 4]
    Lm[5, 5] = Lm[1, 5]
    Lm[6, 0] = af * (vsq - hmb) + cff * (cf - vn) + qs * sbs - asb
    Lm[6, 1] = -af * vn + cff
    Lm[6, 2] = -af * vt1 - qs * b1c
    Lm[6, 3] = -af * vt2 - qs * b2c
    Lm[6, 6] = af
    Lm[6, 4] = Lm[0, 4]
    Lm[6, 5] = Lm[0, 5]
    FL = _cell_flux(wl, ul, bn, hbl)
    FR = _cell_flux(wr, ur, bn, hbr)
    if raw:
        return lam, Lm, R, FL, FR, ul, ur, lam_raw
    return lam, Lm, R, FL, FR, ul, ur


def _cell_flux(w, u, bn, hb):
    f = np.zeros_like(w)
    vn, vt1, vt2, b1, b2, p = w[1], w[2], w[3], w[4], w[5], w[6]
    mn = u[1]
    f[0] = mn
    f[1] = mn * vn
    f[2] = mn * vt1
    f[3] = mn * vt2
    f[1] = f[1] + p
    f[6] = (u[6] + p) * vn
    f[1] = f[1] - 0.5 * (bn * bn - b1 * b1 - b2 * b2)
    f[2] = f[2] - bn * b1
    f[3] = f[3] - bn * b2
    f[6] = f[6] + (hb * vn - bn * (bn * vn + b1 * vt1 + b2 * vt2))
    f[4] = b1 * vn - bn * vt1
    f[5] = b2 * vn - bn * vt2
    return f


# ---------------------------------------------------------------------------
# ghost-index helpers along one axis
# ---------------------------------------------------------------------------
def _cells_ext(n, bc):
    """Cells -1..n mapped onto 0..n-1 (one ghost each side)."""
    if bc == OUTFLOW:
        return [0] + list(range(n)) + [n - 1]
    return [n - 1] + list(range(n)) + [0]


def _faces_ext(n, bc):
    """Faces -1..n+1 of an (n+1)-face array (one ghost face each side)."""
    if bc == OUTFLOW:
        return [0] + list(range(n + 1)) + [n]
    return [n - 1] + list(range(n + 1)) + [1]


# ---------------------------------------------------------------------------
# MHD source terms on reconstructed transverse B
# (Source/Reconstruction_PPM_MHD_HelperModule.py:1193-1239, :1351-1825)
# ---------------------------------------------------------------------------
def recon_b_terms_2d(v_t, face_bn, dt, dn, bc, axis):
    """2D: returns (C0, C1) to add to (W_L, W_R) B_t (axis 1 = x, axis 0 = y)."""
    inv = 1.0 / dn
    hdt = 0.5 * dt
    n = v_t.shape[axis]
    ce = _cells_ext(n, bc)
    fe = _faces_ext(n, bc)
    lv = _take(v_t, ce[0:n + 1], axis)
    rv = _take(v_t, ce[1:n + 2], axis)
    lf = _take(face_bn, fe[0:n + 1], axis)
    rf = _take(face_bn, fe[2:n + 3], axis)
    c0 = hdt * lv * ((face_bn - lf) * inv)
    c1 = hdt * rv * ((rf - face_bn) * inv)
    return c0, c1


def _minmod_recon(dn_, dt_, zero_fix=False):
    pos = dn_ >= 0.0
    t = np.where(pos, np.where(dn_ < -dt_, dn_, -dt_), np.where(dn_ > -dt_, dn_, -dt_))
    if zero_fix:
        took_t = np.where(pos, ~(dn_ < -dt_), ~(dn_ > -dt_))
        t = np.where(took_t & (dt_ == 0.0), 0.0, t)
    t = np.where(pos & (t < 0.0), 0.0, t)
    t = np.where(~pos & (t > 0.0), 0.0, t)
    return t


def recon_b_terms_3d(d, cons_x, dens, fbx, fby, fbz, dt, dx, dy, dz, bc):
    """Returns C[2,2,...face shape]: C[0] for the left/top/back state, C[1] for the
    right/bottom/forward state; [.,0] adds to B_t1, [.,1] to B_t2."""
    inv = {'x': 1.0 / dx, 'y': 1.0 / dy, 'z': 1.0 / dz}
    hdt = 0.5 * dt
    n_axis = {'y': 0, 'x': 1, 'z': 2}
    ax = n_axis[d]
    t1, t2 = {'x': ('y', 'z'), 'y': ('z', 'x'), 'z': ('x', 'y')}[d]
    mom = {'x': cons_x[1], 'y': cons_x[2], 'z': cons_x[3]}
    faces = {'x': fbx, 'y': fby, 'z': fbz}
    n = dens.shape[ax]
    ce = _cells_ext(n, bc)
    fe = _faces_ext(n, bc)

    def dB(c):
        f = faces[c]
        if c == d:
            ef = _take(f, fe, ax)
        else:
            ef = _take(f, ce, ax)
        a = n_axis[c]
        m = ef.shape[a]
        hi = tuple([slice(None)] * a + [slice(1, m)])
        lo = tuple([slice(None)] * a + [slice(0, m - 1)])
        return (ef[hi] - ef[lo]) * inv[c]

    dbn = dB(d)
    dbt1 = dB(t1)
    dbt2 = dB(t2)
    term1 = _minmod_recon(dbn, dbt1, zero_fix=(d == 'y'))
    term2 = _minmod_recon(dbn, dbt2)
    v1 = _take(mom[t1] / dens, ce, ax)
    v2 = _take(mom[t2] / dens, ce, ax)
    s1 = v1 * term1
    s2 = v2 * term2
    lo = tuple([slice(None)] * ax + [slice(0, n + 1)])
    hi = tuple([slice(None)] * ax + [slice(1, n + 2)])
    out = np.zeros((2, 2) + s1[lo].shape)
    out[0, 0] = hdt * s1[lo]
    out[0, 1] = hdt * s2[lo]
    out[1, 0] = hdt * s1[hi]
    out[1, 1] = hdt * s2[hi]
    return out


# ---------------------------------------------------------------------------
# corner EMFs (Source/CornerEMF_HelperModule.py:440-952)
# ---------------------------------------------------------------------------
def _edge_idx(n, bc):
    """Edge-point index 0..n onto cells: (lower neighbour, upper neighbour)."""
    lo = [(n - 1 if bc == PERIODIC else 0)] + list(range(n))
    hi = list(range(n)) + [(0 if bc == PERIODIC else n - 1)]
    return lo, hi


def _upwind(m, pos_val, neg_val, zero_val):
    out = np.zeros(m.shape)
    out = np.where(m > 0.0, pos_val, out)
    out = np.where(m < 0.0, neg_val, out)
    out = np.where(m == 0.0, zero_val, out)
    return out


def corner_emf_2d(F, G, Ec, bcx, bcy):
    """F: x-flux (7, ny, nx+1); G: y-flux (7, ny+1, nx); Ec: cell E_z (ny, nx)."""
    ny, nx = Ec.shape
    xl, xr = _edge_idx(nx, bcx)
    yl, yr = _edge_idx(ny, bcy)
    t = _take(G, xr, 2)            # G(j, i)
    lt = _take(G, xl, 2)           # G(j, i-1)
    lf = _take(F, yr, 1)           # F(j, i)
    ulf = _take(F, yl, 1)          # F(j-1, i)
    E = lambda yi, xi: _take(_take(Ec, yi, 0), xi, 1)  # noqa: E731
    cE = E(yr, xr)
    lE = E(yr, xl)
    uE = E(yl, xr)
    ulE = E(yl, xl)
    x1 = _upwind(lt[0], lt[5] - ulf[4] - ulE, lt[5] - lf[4] - lE,
                 lt[5] - 0.5 * (lf[4] + ulf[4] + lE + ulE))
    x2 = _upwind(t[0], t[5] - ulf[4] - uE, t[5] - lf[4] - cE,
                 t[5] - 0.5 * (lf[4] + ulf[4] + cE + uE))
    y1 = _upwind(ulf[0], -ulf[4] + lt[5] - ulE, -ulf[4] + t[5] - uE,
                 -ulf[4] + 0.5 * (lt[5] + t[5] - ulE - uE))
    y2 = _upwind(lf[0], -lf[4] + lt[5] - lE, -lf[4] + t[5] - cE,
                 -lf[4] + 0.5 * (lt[5] + t[5] - lE - cE))
    return 0.25 * (x1 + x2 + y1 + y2)


def _E3(Ec, yi, xi, zi):
    return _take(_take(_take(Ec, yi, 0), xi, 1), zi, 2)


def corner_emf_3d_z(F, G, Ez, bcx, bcy, bcz):
    ny, nx, nz = Ez.shape
    xl, xr = _edge_idx(nx, bcx)
    yl, yr = _edge_idx(ny, bcy)
    _, zr = _edge_idx(nz, bcz)
    t = _take(_take(G, xr, 2), zr, 3)
    lt = _take(_take(G, xl, 2), zr, 3)
    lf = _take(_take(F, yr, 1), zr, 3)
    ulf = _take(_take(F, yl, 1), zr, 3)
    cE = _E3(Ez, yr, xr, zr)
    lE = _E3(Ez, yr, xl, zr)
    uE = _E3(Ez, yl, xr, zr)
    ulE = _E3(Ez, yl, xl, zr)
    x1 = _upwind(lt[0], -ulf[4] - ulE, -lf[4] - lE, -0.5 * (lf[4] + ulf[4] + lE + ulE))
    x2 = _upwind(t[0], -ulf[4] - uE, -lf[4] - cE, -0.5 * (lf[4] + ulf[4] + cE + uE))
    y1 = _upwind(ulf[0], lt[5] - ulE, t[5] - uE, 0.5 * (lt[5] + t[5] - ulE - uE))
    y2 = _upwind(lf[0], lt[5] - lE, t[5] - cE, 0.5 * (lt[5] + t[5] - lE - cE))
    return 0.25 * (lt[5] + t[5] - ulf[4] - lf[4] + x1 + x2 + y1 + y2)


def corner_emf_3d_x(G, H, Ex, bcx, bcy, bcz):
    ny, nx, nz = Ex.shape
    _, xr = _edge_idx(nx, bcx)
    yl, yr = _edge_idx(ny, bcy)
    zl, zr = _edge_idx(nz, bcz)
    b = _take(_take(H, xr, 2), yr, 1)       # H(j, i, k)
    ub = _take(_take(H, xr, 2), yl, 1)      # H(j-1, i, k)
    t = _take(_take(G, xr, 2), zr, 3)       # G(j, i, k)
    bt = _take(_take(G, xr, 2), zl, 3)      # G(j, i, k-1)
    cE = _E3(Ex, yr, xr, zr)
    bE = _E3(Ex, yr, xr, zl)
    uE = _E3(Ex, yl, xr, zr)
    ubE = _E3(Ex, yl, xr, zl)
    z1 = _upwind(bt[0], ub[5] - ubE, b[5] - bE, 0.5 * (ub[5] + b[5] - ubE - bE))
    z2 = _upwind(t[0], ub[5] - uE, b[5] - cE, 0.5 * (ub[5] + b[5] - uE - cE))
    y1 = _upwind(ub[0], -bt[4] - ubE, -t[4] - uE, 0.5 * (-bt[4] - t[4] - ubE - uE))
    y2 = _upwind(b[0], -bt[4] - bE, -t[4] - cE, 0.5 * (-bt[4] - t[4] - bE - cE))
    return 0.25 * (b[5] + ub[5] - t[4] - bt[4] + y1 + y2 + z1 + z2)


def corner_emf_3d_y(F, H, Ey, bcx, bcy, bcz):
    ny, nx, nz = Ey.shape
    xl, xr = _edge_idx(nx, bcx)
    _, yr = _edge_idx(ny, bcy)
    zl, zr = _edge_idx(nz, bcz)
    b = _take(_take(H, yr, 1), xr, 2)       # H(j, i, k)
    lb = _take(_take(H, yr, 1), xl, 2)      # H(j, i-1, k)
    lf = _take(_take(F, yr, 1), zr, 3)      # F(j, i, k)
    blf = _take(_take(F, yr, 1), zl, 3)     # F(j, i, k-1)
    cE = _E3(Ey, yr, xr, zr)
    bE = _E3(Ey, yr, xr, zl)
    lE = _E3(Ey, yr, xl, zr)
    blE = _E3(Ey, yr, xl, zl)
    z1 = _upwind(blf[0], -lb[4] - blE, -b[4] - bE, 0.5 * (-lb[4] - b[4] - blE - bE))
    z2 = _upwind(lf[0], -lb[4] - lE, -b[4] - cE, 0.5 * (-lb[4] - b[4] - lE - cE))
    x1 = _upwind(lb[0], blf[5] - blE, lf[5] - lE, 0.5 * (blf[5] + lf[5] - blE - lE))
    x2 = _upwind(b[0], blf[5] - bE, lf[5] - cE, 0.5 * (blf[5] + lf[5] - bE - cE))
    return 0.25 * (lf[5] + blf[5] - b[4] - lb[4] + x1 + x2 + z1 + z2)


# ---------------------------------------------------------------------------
# CTU transverse half-step corrections
# (Source/CTU_HelperModule.py:282-444 for 2D, :451-1079 for 3D)
# ---------------------------------------------------------------------------
def ctu_transverse_2d_x(prim_x, prim_y, G, fbx, dt, dx, dy, bcx):
    inv = 1.0 / dx
    hdy = 0.5 * (dt / dy)
    hdt = 0.5 * dt
    ny, nx = prim_x.shape[1:]
    ce = _cells_ext(nx, bcx)
    fe = _faces_ext(nx, bcx)
    cbx = _take(prim_y[5], ce, 1)
    cby = _take(prim_x[4], ce, 1)
    cbz = _take(prim_x[5], ce, 1)
    cvz = _take(prim_x[3], ce, 1)
    lb = _take(fbx, fe[0:nx + 2], 1)
    rb = _take(fbx, fe[1:nx + 3], 1)
    db = rb - lb
    src = np.zeros((7, ny, nx + 2))
    src[1] = cbx * db * inv
    src[2] = cby * db * inv
    src[3] = cbz * db * inv
    src[6] = cbz * cvz * db * inv
    src[5] = cvz * db * inv
    gt = _take(G[:, 0:ny, :], ce, 2)
    gb = _take(G[:, 1:ny + 1, :], ce, 2)
    gt = _remap_y_for_x(gt)
    gb = _remap_y_for_x(gb)
    out = []
    for lo, hi in ((0, nx + 1), (1, nx + 2)):
        s = (slice(None), slice(None), slice(lo, hi))
        out.append(-hdy * (gb[s] - gt[s]) + hdt * src[s])
    return out


def _remap_y_for_x(g):
    g = g.copy()
    g[5] = g[4]
    g[4] = 0.0
    return g


def _remap_x_for_y(f):
    f = f.copy()
    f[4] = f[5]
    f[5] = 0.0
    return f


def ctu_transverse_2d_y(prim_x, prim_y, F, fby, dt, dx, dy, bcy, fix=0):
    inv = 1.0 / dy
    hdx = 0.5 * (dt / dx)
    hdt = 0.5 * dt
    ny, nx = prim_x.shape[1:]
    if bcy == OUTFLOW and (fix & 2):
        # corrected: the x-direction convention (edge cell, edge faces)
        ce = [0] + list(range(ny)) + [ny - 1]
        tb = _take(fby, [0] + list(range(ny + 1)), 0)
        bb = _take(fby, list(range(ny + 1)) + [ny], 0)
        fe_t = [0] + list(range(ny))
        fe_b = list(range(ny)) + [ny - 1]
    elif bcy == OUTFLOW:
        # the reference's top ghost is cell row 1 (it raises IndexError when ny == 1;
        # here row 0, the only row, is used instead)
        ce = [min(1, ny - 1)] + list(range(ny)) + [ny - 1]
        tb = _take(fby, [1] + list(range(ny + 1)), 0)
        bb = _take(fby, list(range(ny + 1)) + [ny - 1], 0)
        fe_t = [0] + list(range(ny))
        fe_b = list(range(ny)) + [ny - 1]
    else:
        ce = [ny - 1] + list(range(ny)) + [0]
        tb = _take(fby, [ny - 1] + list(range(ny + 1)), 0)
        bb = _take(fby, list(range(ny + 1)) + [1], 0)
        fe_t = [ny - 1] + list(range(ny))
        fe_b = list(range(ny)) + [0]
    cbx = _take(prim_y[5], ce, 0)
    cby = _take(prim_x[4], ce, 0)
    cbz = _take(prim_x[5], ce, 0)
    cvz = _take(prim_x[3], ce, 0)
    db = bb - tb
    src = np.zeros((7, ny + 2, nx))
    src[1] = cbx * db * inv
    src[2] = cby * db * inv
    src[3] = cbz * db * inv
    src[6] = (cbz * cvz) * db * inv
    src[4] = cvz * db * inv
    fl = F[:, :, 0:nx]
    fr = F[:, :, 1:nx + 1]
    out = []
    for rows, lo, hi in ((fe_t, 0, ny + 1), (fe_b, 1, ny + 2)):
        l = _remap_x_for_y(_take(fl, rows, 1))
        r = _remap_x_for_y(_take(fr, rows, 1))
        s = (slice(None), slice(lo, hi), slice(None))
        out.append(-hdx * (r - l) + hdt * src[s])
    return out


def _minmod_ctu(t, nrm):
    """Minmod of (t, -n) as coded in the 3D CTU source terms."""
    out = np.zeros(t.shape)
    a = (nrm > 0.0) & (t < 0.0)
    out = np.where(a & (t > -nrm), t, out)
    out = np.where(a & (t <= -nrm), -nrm, out)
    b = (nrm < 0.0) & (t > 0.0)
    out = np.where(b & (t < -nrm), t, out)
    out = np.where(b & (t >= -nrm), -nrm, out)
    return out


def ctu_transverse_3d(d, UL, UR, prim_x, prim_y, flux, faces, emf, dt, dx, dy, dz, bc, fix=0):
    """3D CTU update of the d-interface conserved states.
    flux: dict of the predictor fluxes of the two transverse directions;
    faces: (fBx, fBy, fBz); emf: the corner EMF along d."""
    fbx, fby, fbz = faces
    ny, nx, nz = prim_x.shape[1:]
    nax = {'y': 0, 'x': 1, 'z': 2}
    width = {'x': dx, 'y': dy, 'z': dz}
    ax = nax[d]
    n = prim_x.shape[1 + ax]
    t1, t2 = {'x': ('y', 'z'), 'y': ('z', 'x'), 'z': ('x', 'y')}[d]
    cent = {'bx': prim_y[5], 'by': prim_x[4], 'bz': prim_x[5],
            'vx': prim_x[1], 'vy': prim_x[2], 'vz': prim_x[3]}
    ce = _cells_ext(n, bc)
    # transverse flux pairs (low face, high face) of each transverse direction, at
    # the cell columns left (i-1) and right (i) of each d-face
    rows_l = ce[0:n + 1]
    rows_r = ce[1:n + 2]
    if d == 'y' and bc == OUTFLOW and not (fix & 2):
        # reference quirk: row index nx-1 (Source/CTU_HelperModule.py:724); it raises
        # IndexError when nx > ny, where the last row ny-1 is used instead
        rows_r_fx = list(range(ny)) + [min(nx - 1, ny - 1)]
    else:
        rows_r_fx = rows_r

    def pair(c):
        f = flux[c]
        a = nax[c]
        m = f.shape[1 + a]
        lo = f[tuple([slice(None)] * (1 + a) + [slice(0, m - 1)])]
        hi = f[tuple([slice(None)] * (1 + a) + [slice(1, m)])]
        return lo, hi

    newL = UL.copy()
    newR = UR.copy()
    order = (t1, t2) if d in ('x', 'z') else (t2, t1)
    for c in order:
        lo, hi = pair(c)
        # the bottom ghost row of the high x-face flux in a y-outflow sweep
        rr = rows_r_fx if (d == 'y' and c == 'x') else rows_r
        loL = _take(lo, rows_l, 1 + ax).copy()
        hiL = _take(hi, rows_l, 1 + ax).copy()
        loR = _take(lo, rows_r, 1 + ax).copy()
        hiR = _take(hi, rr, 1 + ax).copy()
        for a_ in (loL, hiL, loR, hiR):
            a_[4] = 0.0
            a_[5] = 0.0
        ext80 = d == 'x' and c == 'y' and not (fix & 4)
        if ext80:
            h = 0.5 * (np.longdouble(dt) / np.longdouble(dy))
            newL = (newL - h * (hiL - loL)).astype(np.float64)
            newR = (newR - h * (hiR - loR)).astype(np.float64)
        else:
            h = 0.5 * (dt / width[c])
            newL = newL - h * (hiL - loL)
            newR = newR - h * (hiR - loR)
        # EMF terms: E(+c1, +c2) stencil on the edge array along d
        eL, eR = _emf_stencil(d, emf, n, bc)
        if c == t1:
            # component t2 gets + h/2 (E(+t1,0) - E(0,0) + E(+t1,+t2) - E(0,+t2))
            termL = eL['10'] - eL['00'] + eL['11'] - eL['01']
            termR = eR['10'] - eR['00'] + eR['11'] - eR['01']
            if ext80:
                newL[5] = (newL[5] + h * 0.5 * termL).astype(np.float64)
                newR[5] = (newR[5] + h * 0.5 * termR).astype(np.float64)
            else:
                newL[5] = newL[5] + h * 0.5 * termL
                newR[5] = newR[5] + h * 0.5 * termR
        else:
            termL = eL['01'] - eL['00'] + eL['11'] - eL['10']
            termR = eR['01'] - eR['00'] + eR['11'] - eR['10']
            newL[4] = newL[4] - h * 0.5 * termL
            newR[4] = newR[4] - h * 0.5 * termR
    # source terms
    dB = _ctu_db(d, faces, n, bc, width, nax, fix)
    bn_ = {'x': 'bx', 'y': 'by', 'z': 'bz'}
    mm_t1 = _minmod_ctu(dB[t1], dB[d])
    mm_t2 = _minmod_ctu(dB[t2], dB[d])
    ec = lambda k: _take(cent[k], ce, ax)  # noqa: E731
    hdt = 0.5 * dt
    src = np.zeros((7,) + dB[d].shape)
    src[1] = hdt * ec('bx') * dB[d]
    src[2] = hdt * ec('by') * dB[d]
    src[3] = hdt * ec('bz') * dB[d]
    v1 = ec('v' + t1)
    v2 = ec('v' + t2)
    b1 = ec(bn_[t1])
    b2 = ec(bn_[t2])
    src[4] = hdt * v1 * (-mm_t2)
    src[5] = hdt * v2 * (-mm_t1)
    src[6] = hdt * (b1 * v1 * (-mm_t2) + b2 * v2 * (-mm_t1))
    lo = tuple([slice(None)] * (1 + ax) + [slice(0, n + 1)])
    hi = tuple([slice(None)] * (1 + ax) + [slice(1, n + 2)])
    newL = newL + src[lo]
    newR = newR + src[hi]
    return newL, newR


def _emf_stencil(d, emf, n, bc):
    """Edge EMF values around each d-face for the left (cell i-1) and right (cell i)
    states.  Keys 'ab': a = +t1 offset, b = +t2 offset."""
    nax = {'y': 0, 'x': 1, 'z': 2}
    ax = nax[d]
    t1, t2 = {'x': ('y', 'z'), 'y': ('z', 'x'), 'z': ('x', 'y')}[d]
    a1, a2 = nax[t1], nax[t2]
    ghost = 0 if bc == OUTFLOW else n - 1
    left_idx = [ghost] + list(range(n))          # cell i-1 layer of the edge array
    right_idx = list(range(n + 1))               # layer i (n is the redundant layer)
    out = []
    for idx in (left_idx, right_idx):
        e = _take(emf, idx, ax)
        m1 = e.shape[a1]
        m2 = e.shape[a2]
        res = {}
        for o1 in (0, 1):
            for o2 in (0, 1):
                s = [slice(None)] * 3
                s[a1] = slice(o1, m1 - 1 + o1)
                s[a2] = slice(o2, m2 - 1 + o2)
                res['%d%d' % (o1, o2)] = e[tuple(s)]
        out.append(res)
    return out


def _ctu_db(d, faces, n, bc, width, nax, fix=0):
    """Cell-centred face-B derivatives on the cells -1..n along d (one ghost)."""
    fbx, fby, fbz = faces
    f = {'x': fbx, 'y': fby, 'z': fbz}
    ax = nax[d]
    ce = _cells_ext(n, bc)
    res = {}
    for c in ('x', 'y', 'z'):
        a = nax[c]
        if c == d:
            m = f[c].shape[a]           # n+1 faces
            if bc == OUTFLOW:
                lo_idx = [0] + list(range(n + 1))
                hi_idx = list(range(n + 1)) + [n if (d != 'z' or fix & 2) else n - 1]
            else:
                lo_idx = [n - 1] + list(range(n + 1))
                hi_idx = list(range(n + 1)) + [1]
            diff = _take(f[c], hi_idx, a) - _take(f[c], lo_idx, a)
        else:
            m = f[c].shape[a]
            lo = f[c][tuple([slice(None)] * a + [slice(0, m - 1)])]
            hi = f[c][tuple([slice(None)] * a + [slice(1, m)])]
            diff = _take(hi, ce, ax) - _take(lo, ce, ax)
        if d == 'x':
            res[c] = diff * (1.0 / width[c])
        else:
            res[c] = diff / width[c]
    return res


# ---------------------------------------------------------------------------
# one cycle of each evolve function
# ---------------------------------------------------------------------------
class Params:
    """Step parameters; riemann selects the flux (0 the reference's Roe, 1 HLLE,
    2 Roe with the HLLE fallback) and fix_quirks the reference behaviours to
    correct (bits 1 z BC, 2 outflow CTU ghosts, 4 80-bit factor); both new in the
    build, 0 is the reference."""

    def __init__(self, gamma, cfl, min_dens=1e-12, min_pres=1e-12, min_energy=1e-12,
                 order=2, entropy_fix=False, eps=0.5, riemann=0, fix_quirks=0):
        self.gamma = gamma
        self.cfl = cfl
        self.min_dens = min_dens
        self.min_pres = min_pres
        self.min_energy = min_energy
        self.order = order
        self.entropy_fix = entropy_fix
        self.eps = eps
        self.riemann = riemann
        self.fix_quirks = fix_quirks


def _dirvec(a8, d):
    return np.stack([a8[k] for k in DIRVARS[d]])


def step_1d(prim, cons, dt, dx, bcx, P):
    """One 1D Godunov cycle (Source/EvolveFunctions_MHD.py:252-317)."""
    g = P.gamma
    n = prim.shape[1]
    bx = prim[4][0] * np.ones(n)
    fbx = prim[4][0] * np.ones(n + 1)
    wx = _dirvec(prim, 'x')
    ux = _dirvec(cons, 'x')
    wl, wr = interface_states(wx, bx, g, dt, dx, bcx, 'x', 0, P.order)
    wl = apply_floors(wl, True, P.min_dens, P.min_pres, P.min_energy)
    wr = apply_floors(wr, True, P.min_dens, P.min_pres, P.min_energy)
    ul = prim_to_cons(wl, fbx, g)
    ur = prim_to_cons(wr, fbx, g)
    F = roe_flux(wl, wr, ul, ur, fbx, g, P.entropy_fix, P.eps, 'x', P.riemann)
    ux = ux - (dt / dx) * (F[:, 1:] - F[:, :-1])
    wx = cons_to_prim(ux, bx, g)
    pa = np.stack([wx[0], wx[1], wx[2], wx[3], bx, wx[4], wx[5], wx[6]])
    ca = np.stack([ux[0], ux[1], ux[2], ux[3], bx, ux[4], ux[5], ux[6]])
    return pa, ca


def step_2d(prim, cons, fbx, fby, dt, dx, dy, bcx, bcy, P):
    """One 2D CTU+CT cycle (Source/EvolveFunctions_MHD.py:464-692)."""
    g = P.gamma
    wx = _dirvec(prim, 'x')
    wy = _dirvec(prim, 'y')
    ux = _dirvec(cons, 'x')
    dtdx = dt / dx
    dtdy = dt / dy
    hdx = 0.5 * dtdx
    hdy = 0.5 * dtdy
    wl, wr = interface_states(wx, wy[5], g, dt, dx, bcx, 'x', 1, P.order)
    wt, wb = interface_states(wy, wx[4], g, dt, dy, bcy, 'y', 0, P.order)
    fl = lambda a, pr: apply_floors(a, pr, P.min_dens, P.min_pres, P.min_energy)  # noqa: E731
    wl, wr, wt, wb = fl(wl, True), fl(wr, True), fl(wt, True), fl(wb, True)
    c0, c1 = recon_b_terms_2d(wx[2], fbx, dt, dx, bcx, 1)
    wl[4] = wl[4] + c0
    wr[4] = wr[4] + c1
    ul = prim_to_cons(wl, fbx, g)
    ur = prim_to_cons(wr, fbx, g)
    c0, c1 = recon_b_terms_2d(wx[1], fby, dt, dy, bcy, 0)
    wt[5] = wt[5] + c0
    wb[5] = wb[5] + c1
    ut = prim_to_cons(wt, fby, g)
    ub = prim_to_cons(wb, fby, g)
    Ec = wx[2] * wy[5] - wx[1] * wx[4]
    F = roe_flux(wl, wr, ul, ur, fbx, g, P.entropy_fix, P.eps, 'x', P.riemann)
    G = roe_flux(wt, wb, ut, ub, fby, g, P.entropy_fix, P.eps, 'y', P.riemann)
    E = corner_emf_2d(F, G, Ec, bcx, bcy)
    hx = ctu_transverse_2d_x(wx, wy, G, fbx, dt, dx, dy, bcx)
    ul = fl(ul + hx[0], False)
    ur = fl(ur + hx[1], False)
    vy_ = ctu_transverse_2d_y(wx, wy, F, fby, dt, dx, dy, bcy, P.fix_quirks)
    ut = fl(ut + vy_[0], False)
    ub = fl(ub + vy_[1], False)
    nfbx = fbx - hdy * (E[1:, :] - E[:-1, :])
    nfby = fby + hdx * (E[:, 1:] - E[:, :-1])
    wl = cons_to_prim(ul, nfbx, g)
    wr = cons_to_prim(ur, nfbx, g)
    wt = cons_to_prim(ut, nfby, g)
    wb = cons_to_prim(ub, nfby, g)
    h0 = ux - hdx * (F[:, :, 1:] - F[:, :, :-1])
    h0 = h0 - hdy * (G[:, 1:, :] - G[:, :-1, :])
    ncbx = (nfbx[:, :-1] + nfbx[:, 1:]) / 2.0
    ncby = (nfby[:-1, :] + nfby[1:, :]) / 2.0
    nvx = h0[1] / h0[0]
    nvy = h0[2] / h0[0]
    nEc = nvy * ncbx - nvx * ncby
    F = roe_flux(wl, wr, ul, ur, nfbx, g, P.entropy_fix, P.eps, 'x', P.riemann)
    G = roe_flux(wt, wb, ut, ub, nfby, g, P.entropy_fix, P.eps, 'y', P.riemann)
    E = corner_emf_2d(F, G, nEc, bcx, bcy)
    fbx = fbx - (dt / dy) * (E[1:, :] - E[:-1, :])
    fby = fby + (dt / dx) * (E[:, 1:] - E[:, :-1])
    cbx = (fbx[:, :-1] + fbx[:, 1:]) / 2.0
    cby = (fby[:-1, :] + fby[1:, :]) / 2.0
    nbz = ux[5] - (dt / dx) * (F[5, :, 1:] - F[5, :, :-1]) - (dt / dy) * (G[4, 1:, :] - G[4, :-1, :])
    ux = ux - (dt / dx) * (F[:, :, 1:] - F[:, :, :-1])
    ux = ux - (dt / dy) * (G[:, 1:, :] - G[:, :-1, :])
    ux[5] = nbz
    ux[4] = cby
    wx = cons_to_prim(ux, cbx, g)
    pa = np.stack([wx[0], wx[1], wx[2], wx[3], cbx, wx[4], wx[5], wx[6]])
    ca = np.stack([ux[0], ux[1], ux[2], ux[3], cbx, ux[4], ux[5], ux[6]])
    return pa, ca, fbx, fby


def _ct_update_3d(fbx, fby, fbz, Ex, Ey, Ez, hx, hy, hz):
    nfbx = fbx + (hz * (Ey[:-1, :, 1:] - Ey[:-1, :, :-1]) - hy * (Ez[1:, :, :-1] - Ez[:-1, :, :-1]))
    nfby = fby + (hx * (Ez[:, 1:, :-1] - Ez[:, :-1, :-1]) - hz * (Ex[:, :-1, 1:] - Ex[:, :-1, :-1]))
    nfbz = fbz + (hy * (Ex[1:, :-1, :] - Ex[:-1, :-1, :]) - hx * (Ey[:-1, 1:, :] - Ey[:-1, :-1, :]))
    return nfbx, nfby, nfbz


def step_3d(prim, cons, fbx, fby, fbz, dt, dx, dy, dz, bcx, bcy, bcz, P):
    """One 3D CTU+CT cycle (Source/EvolveFunctions_MHD.py:840-1152)."""
    g = P.gamma
    wx = _dirvec(prim, 'x')
    wy = _dirvec(prim, 'y')
    wz = _dirvec(prim, 'z')
    ux = _dirvec(cons, 'x')
    dtdx, dtdy, dtdz = dt / dx, dt / dy, dt / dz
    hdx, hdy, hdz = 0.5 * dtdx, 0.5 * dtdy, 0.5 * dtdz
    if P.order == 0:
        wl, wr = copy_states(wx, bcx, 1)
        wt, wb = copy_states(wy, bcy, 0)
        wk, wf = copy_states(wz, bcz if P.fix_quirks & 1 else bcx, 2)   # the reference uses BcX
    else:
        wl, wr = recon_states(wx, wy[5], g, dt, dx, bcx, 'x', 1, P.order)
        wt, wb = recon_states(wy, wx[4], g, dt, dy, bcy, 'y', 0, P.order)
        wk, wf = recon_states(wz, wx[5], g, dt, dz, bcz, 'z', 2, P.order)
    fl = lambda a, pr: apply_floors(a, pr, P.min_dens, P.min_pres, P.min_energy)  # noqa: E731
    wl, wr, wt, wb, wk, wf = [fl(a, True) for a in (wl, wr, wt, wb, wk, wf)]
    C = recon_b_terms_3d('x', ux, wx[0], fbx, fby, fbz, dt, dx, dy, dz, bcx)
    wl[4] = wl[4] + C[0, 0]
    wl[5] = wl[5] + C[0, 1]
    wr[4] = wr[4] + C[1, 0]
    wr[5] = wr[5] + C[1, 1]
    ul = prim_to_cons(wl, fbx, g, 'x')
    ur = prim_to_cons(wr, fbx, g, 'x')
    C = recon_b_terms_3d('y', ux, wx[0], fbx, fby, fbz, dt, dx, dy, dz, bcy)
    wt[4] = wt[4] + C[0, 0]
    wt[5] = wt[5] + C[0, 1]
    wb[4] = wb[4] + C[1, 0]
    wb[5] = wb[5] + C[1, 1]
    ut = prim_to_cons(wt, fby, g, 'y')
    ub = prim_to_cons(wb, fby, g, 'y')
    C = recon_b_terms_3d('z', ux, wx[0], fbx, fby, fbz, dt, dx, dy, dz,
                         bcz if P.fix_quirks & 1 else bcx)   # the reference uses BcX
    wk[4] = wk[4] + C[0, 0]
    wk[5] = wk[5] + C[0, 1]
    wf[4] = wf[4] + C[1, 0]
    wf[5] = wf[5] + C[1, 1]
    uk = prim_to_cons(wk, fbz, g, 'z')
    uf = prim_to_cons(wf, fbz, g, 'z')
    Exc = (wx[4] * ux[3] - wx[5] * ux[2]) / ux[0]
    Eyc = (wx[5] * ux[1] - wy[5] * ux[3]) / ux[0]
    Ezc = (wy[5] * ux[2] - wx[4] * ux[1]) / ux[0]
    ef, eps = P.entropy_fix, P.eps
    F = roe_flux(wl, wr, ul, ur, fbx, g, ef, eps, 'x', P.riemann)
    G = roe_flux(wt, wb, ut, ub, fby, g, ef, eps, 'y', P.riemann)
    H = roe_flux(wk, wf, uk, uf, fbz, g, ef, eps, 'z', P.riemann)
    Ex = corner_emf_3d_x(G, H, Exc, bcx, bcy, bcz)
    Ey = corner_emf_3d_y(F, H, Eyc, bcx, bcy, bcz)
    Ez = corner_emf_3d_z(F, G, Ezc, bcx, bcy, bcz)
    faces = (fbx, fby, fbz)
    flux = {'x': F, 'y': G, 'z': H}
    fq = P.fix_quirks
    ul, ur = ctu_transverse_3d('x', ul, ur, wx, wy, flux, faces, Ex, dt, dx, dy, dz, bcx, fq)
    ut, ub = ctu_transverse_3d('y', ut, ub, wx, wy, flux, faces, Ey, dt, dx, dy, dz, bcy, fq)
    uk, uf = ctu_transverse_3d('z', uk, uf, wx, wy, flux, faces, Ez, dt, dx, dy, dz, bcz, fq)
    nfbx, nfby, nfbz = _ct_update_3d(fbx, fby, fbz, Ex, Ey, Ez, hdx, hdy, hdz)
    wl = cons_to_prim(ul, nfbx, g, 'x')
    wr = cons_to_prim(ur, nfbx, g, 'x')
    wt = cons_to_prim(ut, nfby, g, 'y')
    wb = cons_to_prim(ub, nfby, g, 'y')
    wk = cons_to_prim(uk, nfbz, g, 'z')
    wf = cons_to_prim(uf, nfbz, g, 'z')

    def half(k):
        return (ux[k] - hdx * (F[k, :, 1:, :] - F[k, :, :-1, :])
                - hdy * (G[k, 1:, :, :] - G[k, :-1, :, :])
                - hdz * (H[k, :, :, 1:] - H[k, :, :, :-1]))
    nd, nmx, nmy, nmz = half(0), half(1), half(2), half(3)
    nbx = 0.5 * (nfbx[:, :-1, :] + nfbx[:, 1:, :])
    nby = 0.5 * (nfby[:-1, :, :] + nfby[1:, :, :])
    nbz = 0.5 * (nfbz[:, :, :-1] + nfbz[:, :, 1:])
    nExc = (nby * nmz - nbz * nmy) / nd
    nEyc = (nbz * nmx - nbx * nmz) / nd
    nEzc = (nbx * nmy - nby * nmx) / nd
    F = roe_flux(wl, wr, ul, ur, nfbx, g, ef, eps, 'x', P.riemann)
    G = roe_flux(wt, wb, ut, ub, nfby, g, ef, eps, 'y', P.riemann)
    H = roe_flux(wk, wf, uk, uf, nfbz, g, ef, eps, 'z', P.riemann)
    Ex = corner_emf_3d_x(G, H, nExc, bcx, bcy, bcz)
    Ey = corner_emf_3d_y(F, H, nEyc, bcx, bcy, bcz)
    Ez = corner_emf_3d_z(F, G, nEzc, bcx, bcy, bcz)
    fbx, fby, fbz = _ct_update_3d(fbx, fby, fbz, Ex, Ey, Ez, dtdx, dtdy, dtdz)
    ux = ux - dtdx * (F[:, :, 1:, :] - F[:, :, :-1, :])
    ux = ux - dtdy * (G[:, 1:, :, :] - G[:, :-1, :, :])
    ux = ux - dtdz * (H[:, :, :, 1:] - H[:, :, :, :-1])
    cbx = 0.5 * (fbx[:, :-1, :] + fbx[:, 1:, :])
    cby = 0.5 * (fby[:-1, :, :] + fby[1:, :, :])
    cbz = 0.5 * (fbz[:, :, :-1] + fbz[:, :, 1:])
    ux[4] = cby
    ux[5] = cbz
    wx = cons_to_prim(ux, cbx, g)
    pa = np.stack([wx[0], wx[1], wx[2], wx[3], cbx, wx[4], wx[5], wx[6]])
    ca = np.stack([ux[0], ux[1], ux[2], ux[3], cbx, ux[4], ux[5], ux[6]])
    return pa, ca, fbx, fby, fbz


# ---------------------------------------------------------------------------
# evolve loops with the reference's time clipping and return tuples
# ---------------------------------------------------------------------------
def evolve_1d(prim, cons, dx, bcx, t0, t_end, cycle0, max_cycles, P):
    t, cyc = t0, cycle0
    sl = (prim.copy(), cons.copy(), t0)
    while t < t_end and cyc < max_cycles:
        sl = (prim.copy(), cons.copy(), t)
        dt = calc_dt(cons, (), P.gamma, P.cfl, (dx,))
        if t + dt > t_end:
            dt = t_end - t
        prim, cons = step_1d(prim, cons, dt, dx, bcx, P)
        t += dt
        cyc += 1
    return prim, cons, t, cyc, sl[0], sl[1], sl[2]


def evolve_2d(prim, cons, fbx, fby, dx, dy, bcx, bcy, t0, t_end, cycle0, max_cycles, P):
    t, cyc = t0, cycle0
    sl = (prim.copy(), cons.copy(), fbx.copy(), fby.copy(), t0)
    while t < t_end and cyc < max_cycles:
        sl = (prim.copy(), cons.copy(), fbx.copy(), fby.copy(), t)
        dt = calc_dt(cons, (fbx, fby), P.gamma, P.cfl, (dx, dy))
        if t + dt > t_end:
            dt = t_end - t
        prim, cons, fbx, fby = step_2d(prim, cons, fbx, fby, dt, dx, dy, bcx, bcy, P)
        t += dt
        cyc += 1
    return prim, cons, fbx, fby, t, cyc, sl[0], sl[1], sl[2], sl[3], sl[4]


def evolve_3d(prim, cons, fbx, fby, fbz, dx, dy, dz, bcx, bcy, bcz, t0, t_end, cycle0,
              max_cycles, P):
    t, cyc = t0, cycle0
    sl = (prim.copy(), cons.copy(), fbx.copy(), fby.copy(), fbz.copy(), t0)
    while t < t_end and cyc < max_cycles:
        sl = (prim.copy(), cons.copy(), fbx.copy(), fby.copy(), fbz.copy(), t)
        dt = calc_dt(cons, (fbx, fby, fbz), P.gamma, P.cfl, (dx, dy, dz))
        if t_end - t < dt:
            dt = t_end - t
        prim, cons, fbx, fby, fbz = step_3d(prim, cons, fbx, fby, fbz, dt, dx, dy, dz,
                                            bcx, bcy, bcz, P)
        t += dt
        cyc += 1
    return (prim, cons, fbx, fby, fbz, t, cyc) + sl
