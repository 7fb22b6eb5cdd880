"""Write the golden input/output vectors of the reference's hot-path functions.

Run once, in the container that holds the reference checkout:

    python tests/golden/make_golden.py /root/reference/PythonMHD

It imports PythonMHD's NumPy helper modules (never ``Source.Simulation``, whose
file-output imports are unavailable), calls them on small seeded inputs and saves
inputs and outputs as ``tests/golden/*.npz``.  Only these numeric fixtures are
committed; nothing of the reference travels with them.
"""
import contextlib
import io
import os
import sys

sys.dont_write_bytecode = True    # import the reference read-only: no __pycache__ in its tree

import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))


def _load(ref_root):
    sys.path.insert(0, ref_root)
    import Source.EvolveFunctions_MHD as evo
    import Source.Reconstruction_PPM_MHD_HelperModule as ppm
    import Source.RiemannSolvers_MHD_Roe_HelperModule as roe
    import Source.TimestepCalculations_HelperModule as dtl
    import Source.CTU_HelperModule as ctu
    import Source.CornerEMF_HelperModule as emf
    import Source.PrimCons_HelperModule as pc
    return evo, ppm, roe, dtl, ctu, emf, pc


class Grid:
    pass


def rand_prim(rng, shape, bzero_frac=0.15):
    """A random but physical 8-variable primitive state with some exact zeros in B."""
    p = np.empty((8,) + shape)
    p[0] = rng.uniform(0.3, 2.0, shape)
    p[1:4] = rng.uniform(-1.0, 1.0, (3,) + shape)
    p[4:7] = rng.uniform(-1.0, 1.0, (3,) + shape)
    for k in (4, 5, 6):
        m = rng.random(shape) < bzero_frac
        p[k][m] = 0.0
    p[7] = rng.uniform(0.2, 2.0, shape)
    return p


def smooth_prim(rng, shape, amp=0.2):
    """A smooth periodic state (a few random Fourier modes) so limiters take both branches."""
    grids = np.meshgrid(*[np.arange(n) / n for n in shape], indexing='ij')
    p = np.empty((8,) + shape)
    base = [1.0, 0.0, 0.0, 0.0, 0.4, 0.3, 0.2, 1.0]
    for v in range(8):
        f = np.full(shape, base[v])
        for _ in range(3):
            k = rng.integers(1, 3, len(shape))
            ph = rng.uniform(0, 2 * np.pi)
            f = f + amp * rng.uniform(-1, 1) * np.sin(
                2 * np.pi * sum(ki * g for ki, g in zip(k, grids)) + ph)
        p[v] = f
    p[0] = np.abs(p[0]) + 0.3
    p[7] = np.abs(p[7]) + 0.3
    return p


def faces_from_cells(b, axis, bc):
    """Face B from cell B by averaging neighbours (periodic wrap / edge copy)."""
    n = b.shape[axis]
    lo = np.take(b, [(n - 1 if bc == 1 else 0)] + list(range(n)), axis=axis)
    hi = np.take(b, list(range(n)) + [(0 if bc == 1 else n - 1)], axis=axis)
    return (lo + hi) / 2.0


def main(ref_root):
    evo, ppm, roe, dtl, ctu, emf, pc = _load(ref_root)
    rng = np.random.default_rng(20240101)
    quiet = contextlib.redirect_stdout(io.StringIO())

    # ---------------- calcDt ----------------
    g = 5.0 / 3.0
    d = {}
    for nd, shape in ((1, (33,)), (2, (9, 11)), (3, (5, 6, 7))):
        pr = rand_prim(rng, shape)
        cons = pc.primToCons_mhd_allVars(pr, g)
        d['cons%d' % nd] = cons
        if nd == 1:
            d['dt1'] = dtl.calcDt_1D_mhd(cons, g, 0.8, 0.03)
        elif nd == 2:
            fbx = rng.uniform(-1, 1, (shape[0], shape[1] + 1))
            fby = rng.uniform(-1, 1, (shape[0] + 1, shape[1]))
            d['fbx2'], d['fby2'] = fbx, fby
            d['dt2'] = dtl.calcDt_2D_mhd(cons, fbx, fby, g, 0.8, 0.05, 0.07)
        else:
            fbx = rng.uniform(-1, 1, (shape[0], shape[1] + 1, shape[2]))
            fby = rng.uniform(-1, 1, (shape[0] + 1, shape[1], shape[2]))
            fbz = rng.uniform(-1, 1, (shape[0], shape[1], shape[2] + 1))
            d['fbx3'], d['fby3'], d['fbz3'] = fbx, fby, fbz
            d['dt3'] = dtl.calcDt_3D_mhd(cons, fbx, fby, fbz, g, 0.4, 0.1, 0.11, 0.09)
    np.savez_compressed(os.path.join(OUT, 'calc_dt.npz'), **d)

    # ---------------- Roe fluxes ----------------
    d = {}
    nf = 1600
    for axis, fn, perm in (('x', roe.riemannSolverX_mhd_roe, (0, 1, 2, 3, 5, 6, 7)),
                           ('y', roe.riemannSolverY_mhd_roe, (0, 1, 2, 3, 6, 4, 7)),
                           ('z', roe.riemannSolverZ_mhd_roe, (0, 1, 2, 3, 4, 5, 7))):
        pl = rand_prim(rng, (nf,), 0.3)
        pr = rand_prim(rng, (nf,), 0.3)
        bn = rng.uniform(-1, 1, nf)
        bn[rng.random(nf) < 0.2] = 0.0
        # a block of identical states and a block of supersonic flows
        pr[:, :100] = pl[:, :100]
        pl[1, 100:200] = 4.0
        pr[1, 100:200] = 4.0
        pl[1, 200:300] = -4.0
        pr[1, 200:300] = -4.0
        pl[4:7, 300:350] = 0.0
        pr[4:7, 300:350] = 0.0
        bn[300:350] = 0.0
        wl = pl[list(perm)]
        wr = pr[list(perm)]
        bidx = {'x': 4, 'y': 5, 'z': 6}[axis]
        pl[bidx] = bn
        pr[bidx] = bn
        ul = pc.primToCons_mhd_allVars(pl, g)[list(perm)]
        ur = pc.primToCons_mhd_allVars(pr, g)[list(perm)]
        shp = (8, 10, 20)    # the y and z solvers take 2D/3D arrays only
        rs = lambda a: a.reshape(a.shape[:-1] + shp)  # noqa: E731
        for ef in (False, True):
            f = fn(rs(wl), rs(wr), rs(ul), rs(ur), bn.reshape(shp), g, ef, 0.5)
            d['%s_ef%d' % (axis, ef)] = f.reshape(7, nf)
        d[axis + '_wl'], d[axis + '_wr'], d[axis + '_ul'], d[axis + '_ur'], d[axis + '_bn'] = wl, wr, ul, ur, bn
    np.savez_compressed(os.path.join(OUT, 'roe.npz'), **d)

    # ---------------- PPM reconstruction ----------------
    d = {}
    shape3 = (6, 7, 8)
    for bc in (0, 1):
        pr = smooth_prim(rng, shape3) + 0.05 * rand_prim(rng, shape3, 0.0)
        d['p3_bc%d' % bc] = pr
        px = pr[[0, 1, 2, 3, 5, 6, 7]]
        py = pr[[0, 1, 2, 3, 6, 4, 7]]
        pz = pr[[0, 1, 2, 3, 4, 5, 7]]
        d['x3_bc%d' % bc] = np.stack(ppm.ppmReconstructX_mhd(px, pr[4], g, 0.01, 0.1, bc))
        d['y3_bc%d' % bc] = np.stack(ppm.ppmReconstructY_mhd(py, pr[5], g, 0.01, 0.12, bc))
        d['z3_bc%d' % bc] = np.stack(ppm.ppmReconstructZ_mhd(pz, pr[6], g, 0.01, 0.09, bc))
        pr2 = smooth_prim(rng, (9, 10)) + 0.05 * rand_prim(rng, (9, 10), 0.0)
        d['p2_bc%d' % bc] = pr2
        d['x2_bc%d' % bc] = np.stack(ppm.ppmReconstructX_mhd(pr2[[0, 1, 2, 3, 5, 6, 7]], pr2[4], g, 0.02, 0.1, bc))
        d['y2_bc%d' % bc] = np.stack(ppm.ppmReconstructY_mhd(pr2[[0, 1, 2, 3, 6, 4, 7]], pr2[5], g, 0.02, 0.1, bc))
        pr1 = smooth_prim(rng, (40,)) + 0.05 * rand_prim(rng, (40,), 0.0)
        d['p1_bc%d' % bc] = pr1
        d['x1_bc%d' % bc] = np.stack(ppm.ppmReconstructX_mhd(pr1[[0, 1, 2, 3, 5, 6, 7]], pr1[4], 2.0, 0.003, 0.025, bc))
    np.savez_compressed(os.path.join(OUT, 'ppm.npz'), **d)

    # ---------------- 3D helpers: recon B terms, corner EMF, CTU ----------------
    d = {}
    ny, nx, nz = 6, 6, 6
    for bc in (0, 1):
        pr = smooth_prim(rng, (ny, nx, nz))
        cons = pc.primToCons_mhd_allVars(pr, g)
        ux = cons[[0, 1, 2, 3, 5, 6, 7]]
        px = pr[[0, 1, 2, 3, 5, 6, 7]]
        py = pr[[0, 1, 2, 3, 6, 4, 7]]
        fbx = rng.uniform(-1, 1, (ny, nx + 1, nz))
        fby = rng.uniform(-1, 1, (ny + 1, nx, nz))
        fbz = rng.uniform(-1, 1, (ny, nx, nz + 1))
        if bc == 1:
            fbx[:, nx, :] = fbx[:, 0, :]
            fby[ny, :, :] = fby[0, :, :]
            fbz[:, :, nz] = fbz[:, :, 0]
        # exact zero differences so tie-breaks are exercised
        fbx[:, 2, :] = fbx[:, 1, :]
        fby[3, :, :] = fby[2, :, :]
        fbz[:, :, 4] = fbz[:, :, 3]
        dt, dx, dy, dz = 0.013, 0.1, 0.12, 0.09
        k = 'bc%d_' % bc
        d[k + 'prim'], d[k + 'cons'] = pr, cons
        d[k + 'fbx'], d[k + 'fby'], d[k + 'fbz'] = fbx, fby, fbz
        d[k + 'recx'] = ppm.getReconBCorrectionsX_3D(ux[2], ux[3], px[0], fbx, fby, fbz, dt, dx, dy, dz, bc)
        d[k + 'recy'] = ppm.getReconBCorrectionsY_3D(ux[1], ux[3], px[0], fbx, fby, fbz, dt, dx, dy, dz, bc)
        d[k + 'recz'] = ppm.getReconBCorrectionsZ_3D(ux[1], ux[2], px[0], fbx, fby, fbz, dt, dx, dy, dz, bc)
        F = rng.uniform(-1, 1, (7, ny, nx + 1, nz))
        G = rng.uniform(-1, 1, (7, ny + 1, nx, nz))
        H = rng.uniform(-1, 1, (7, ny, nx, nz + 1))
        for A in (F, G, H):
            A[0][rng.random(A[0].shape) < 0.2] = 0.0
        Ec = rng.uniform(-1, 1, (3, ny, nx, nz))
        d[k + 'F'], d[k + 'G'], d[k + 'H'], d[k + 'Ec'] = F, G, H, Ec
        Ex = emf.getXCornerEMF_3D(G, H, Ec[0], bc, bc, bc)
        Ey = emf.getYCornerEMF_3D(F, H, Ec[1], bc, bc, bc)
        Ez = emf.getZCornerEMF_3D(F, G, Ec[2], bc, bc, bc)
        d[k + 'Ex'], d[k + 'Ey'], d[k + 'Ez'] = Ex, Ey, Ez
        UL = [rng.uniform(-1, 1, s) for s in ((7, ny, nx + 1, nz), (7, ny + 1, nx, nz), (7, ny, nx, nz + 1))]
        UR = [rng.uniform(-1, 1, s) for s in ((7, ny, nx + 1, nz), (7, ny + 1, nx, nz), (7, ny, nx, nz + 1))]
        d[k + 'ULx'], d[k + 'URx'], d[k + 'ULy'], d[k + 'URy'], d[k + 'ULz'], d[k + 'URz'] = \
            UL[0], UR[0], UL[1], UR[1], UL[2], UR[2]
        d[k + 'ctux'] = np.stack(ctu.getXTransverseUpdates_3D(UL[0], UR[0], px, py, G, H, fbx, fby, fbz, Ex,
                                                              dt, dx, dy, dz, bc))
        d[k + 'ctuy'] = np.stack(ctu.getYTransverseUpdates_3D(UL[1], UR[1], px, py, F, H, fbx, fby, fbz, Ey,
                                                              dt, dx, dy, dz, bc))
        d[k + 'ctuz'] = np.stack(ctu.getZTransverseUpdates_3D(UL[2], UR[2], px, py, F, G, fbx, fby, fbz, Ez,
                                                              dt, dx, dy, dz, bc))
    np.savez_compressed(os.path.join(OUT, 'helpers3d.npz'), **d)

    # ---------------- 2D helpers ----------------
    d = {}
    ny, nx = 7, 9
    for bc in (0, 1):
        pr = smooth_prim(rng, (ny, nx))
        px = pr[[0, 1, 2, 3, 5, 6, 7]]
        py = pr[[0, 1, 2, 3, 6, 4, 7]]
        fbx = rng.uniform(-1, 1, (ny, nx + 1))
        fby = rng.uniform(-1, 1, (ny + 1, nx))
        F = rng.uniform(-1, 1, (7, ny, nx + 1))
        G = rng.uniform(-1, 1, (7, ny + 1, nx))
        F[0][rng.random(F[0].shape) < 0.2] = 0.0
        G[0][rng.random(G[0].shape) < 0.2] = 0.0
        Ec = rng.uniform(-1, 1, (ny, nx))
        dt, dx, dy = 0.013, 0.1, 0.12
        k = 'bc%d_' % bc
        d[k + 'prim'], d[k + 'fbx'], d[k + 'fby'], d[k + 'F'], d[k + 'G'], d[k + 'Ec'] = pr, fbx, fby, F, G, Ec
        d[k + 'srcby'] = ppm.getByReconSourceTerms_2D(px[2], fbx, dt, dx, bc)
        d[k + 'srcbx'] = ppm.getBxReconSourceTerms_2D(px[1], fby, dt, dy, bc)
        d[k + 'emf'] = emf.getCornerEMF_2D(F, G, Ec, bc, bc)
        d[k + 'ctux'] = ctu.getXTransverseUpdates_2D(px, py, G, fbx, dt, dx, dy, bc)
        d[k + 'ctuy'] = ctu.getYTransverseUpdates_2D(px, py, F, fby, dt, dx, dy, bc)
    np.savez_compressed(os.path.join(OUT, 'helpers2d.npz'), **d)

    # ---------------- full evolve trajectories ----------------
    d = {}
    # 1D Brio-Wu (64 cells), orders 0 and 2, outflow; and a periodic smooth case
    nx = 64
    x = (np.arange(nx) + 0.5) / nx
    pr = np.zeros((8, nx))
    pr[0] = np.where(x <= 0.5, 1.0, 0.125)
    pr[7] = np.where(x <= 0.5, 1.0, 0.1)
    pr[5] = np.where(x <= 0.5, 1.0, -1.0)
    pr[4] = 0.75
    for order in (0, 2):
        gr = Grid()
        gr.primVars, gr.consVars = pr.copy(), pc.primToCons_mhd_allVars(pr, 2.0)
        gr.dx, gr.BcX = 1.0 / nx, 0
        with quiet:
            r = evo.evolve_1D_mhd(gr, 0.0, 0.1, 0, 25, 2.0, 0.8, 1e-12, 1e-12, 1e-12, order, False, 0.5)
        d['bw_o%d_in' % order] = pr
        for i, name in enumerate(('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')):
            d['bw_o%d_%s' % (order, name)] = np.asarray(r[i])
    pr1 = smooth_prim(rng, (48,))
    pr1[4] = 0.6
    gr = Grid()
    gr.primVars, gr.consVars = pr1.copy(), pc.primToCons_mhd_allVars(pr1, g)
    gr.dx, gr.BcX = 1.0 / 48, 1
    with quiet:
        r = evo.evolve_1D_mhd(gr, 0.0, 10.0, 0, 15, g, 0.8, 1e-12, 1e-12, 1e-12, 2, True, 0.5)
    d['p1_in'] = pr1
    for i, name in enumerate(('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')):
        d['p1_%s' % name] = np.asarray(r[i])

    # 2D: periodic smooth 16x16 (orders 0 and 2), and an outflow 12x14 case
    for tag, (ny, nx), bc, order, ncyc in (('p2o2', (16, 16), 1, 2, 6), ('p2o0', (16, 16), 1, 0, 6),
                                            ('o2o2', (12, 14), 0, 2, 5)):
        pr = smooth_prim(rng, (ny, nx))
        fbx = faces_from_cells(pr[4], 1, bc)
        fby = faces_from_cells(pr[5], 0, bc)
        pr[4] = (fbx[:, :-1] + fbx[:, 1:]) / 2.0
        pr[5] = (fby[:-1, :] + fby[1:, :]) / 2.0
        cons = pc.primToCons_mhd_allVars(pr, g)
        gr = Grid()
        gr.primVars, gr.consVars, gr.faceBx, gr.faceBy = pr.copy(), cons.copy(), fbx.copy(), fby.copy()
        gr.dx, gr.dy, gr.BcX, gr.BcY = 1.0 / nx, 1.0 / ny, bc, bc
        with quiet:
            r = evo.evolve_2D_mhd(gr, 0.0, 10.0, 0, ncyc, g, 0.8, 1e-12, 1e-12, 1e-12, order, False, 0.5)
        d[tag + '_in_prim'], d[tag + '_in_cons'], d[tag + '_in_fbx'], d[tag + '_in_fby'] = pr, cons, fbx, fby
        names = ('prim', 'cons', 'fbx', 'fby', 't', 'cyc', 'prim2', 'cons2', 'fbx2', 'fby2', 't2')
        for i, name in enumerate(names):
            d[tag + '_' + name] = np.asarray(r[i])

    # 3D: periodic smooth 8^3 (orders 0 and 2, entropy fix on), outflow 6^3
    for tag, shape, bc, order, ncyc in (('p3o2', (8, 8, 8), 1, 2, 3), ('p3o0', (8, 8, 8), 1, 0, 3),
                                        ('n3o2', (6, 8, 10), 1, 2, 2), ('o3o2', (6, 6, 6), 0, 2, 2)):
        pr = smooth_prim(rng, shape)
        fbx = faces_from_cells(pr[4], 1, bc)
        fby = faces_from_cells(pr[5], 0, bc)
        fbz = faces_from_cells(pr[6], 2, bc)
        pr[4] = 0.5 * (fbx[:, :-1, :] + fbx[:, 1:, :])
        pr[5] = 0.5 * (fby[:-1, :, :] + fby[1:, :, :])
        pr[6] = 0.5 * (fbz[:, :, :-1] + fbz[:, :, 1:])
        cons = pc.primToCons_mhd_allVars(pr, g)
        gr = Grid()
        gr.primVars, gr.consVars = pr.copy(), cons.copy()
        gr.faceBx, gr.faceBy, gr.faceBz = fbx.copy(), fby.copy(), fbz.copy()
        gr.dx, gr.dy, gr.dz = 1.0 / shape[1], 1.0 / shape[0], 1.0 / shape[2]
        gr.BcX = gr.BcY = gr.BcZ = bc
        with quiet:
            r = evo.evolve_3D_mhd(gr, 0.0, 10.0, 0, ncyc, g, 0.4, 1e-12, 1e-12, 1e-12, order, True, 0.5)
        for nm, a in (('prim', pr), ('cons', cons), ('fbx', fbx), ('fby', fby), ('fbz', fbz)):
            d[tag + '_in_' + nm] = a
        names = ('prim', 'cons', 'fbx', 'fby', 'fbz', 't', 'cyc',
                 'prim2', 'cons2', 'fbx2', 'fby2', 'fbz2', 't2')
        for i, name in enumerate(names):
            d[tag + '_' + name] = np.asarray(r[i])
    np.savez_compressed(os.path.join(OUT, 'evolve.npz'), **d)
    print('golden vectors written to', OUT)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/PythonMHD')
