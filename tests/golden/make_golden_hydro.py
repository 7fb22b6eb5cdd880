"""Write the golden input/output vectors of the reference's hydrodynamic hot path.

Run once, in the container that holds the reference checkout:

    python tests/golden/make_golden_hydro.py /root/reference/PythonMHD

It imports PythonMHD's hydro helper modules and evolve functions, calls them on
small seeded inputs and saves inputs and outputs as ``tests/golden/hydro_*.npz``.
Only these numeric fixtures are committed; nothing of the reference travels.
Inputs on which the reference raises (1D periodic PPM, 3D outflow x) are not
generated.
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
    import Source.EvolveFunctions_Hydro as evo
    import Source.Reconstruction_PPM_Hydro_HelperModule as ppm
    import Source.RiemannSolvers_Hydro_Roe_HelperModule as roe
    import Source.TimestepCalculations_HelperModule as dtl
    import Source.PrimCons_HelperModule as pc
    return evo, ppm, roe, dtl, pc


class Grid:
    pass


def rand_prim(rng, shape):
    p = np.empty((5,) + shape)
    p[0] = rng.uniform(0.3, 2.0, shape)
    p[1:4] = rng.uniform(-1.0, 1.0, (3,) + shape)
    p[4] = rng.uniform(0.2, 2.0, shape)
    return p


def smooth_prim(rng, shape, amp=0.2):
    grids = np.meshgrid(*[np.arange(n) / n for n in shape], indexing='ij')
    p = np.empty((5,) + shape)
    base = [1.0, 0.0, 0.0, 0.0, 1.0]
    for v in range(5):
        f = np.full(shape, base[v])
        for _ in range(3):
            k = rng.integers(1, 3, len(shape))
            ph = rng.uniform(0, 2 * np.pi)
            f = f + amp * rng.uniform(-1, 1) * np.sin(2 * np.pi * sum(ki * g for ki, g in zip(k, grids)) + ph)
        p[v] = f
    p[0] = np.abs(p[0]) + 0.3
    p[4] = np.abs(p[4]) + 0.3
    return p


def main(ref_root):
    evo, ppm, roe, dtl, pc = _load(ref_root)
    rng = np.random.default_rng(20240202)
    quiet = contextlib.redirect_stdout(io.StringIO())
    g = 1.4

    # ---------------- calcDt ----------------
    d = {}
    for nd, shape in ((1, (33,)), (2, (9, 11)), (3, (5, 6, 7))):
        cons = pc.primToCons_hydro(rand_prim(rng, shape), g)
        d['cons%d' % nd] = cons
        if nd == 1:
            d['dt1'] = dtl.calcDt_1D_hydro(cons, g, 0.8, 0.03)
        elif nd == 2:
            d['dt2'] = dtl.calcDt_2D_hydro(cons, g, 0.8, 0.05, 0.07)
        else:
            d['dt3'] = dtl.calcDt_3D_hydro(cons, g, 0.4, 0.1, 0.11, 0.09)
    np.savez_compressed(os.path.join(OUT, 'hydro_calc_dt.npz'), **d)

    # ---------------- Roe fluxes ----------------
    d = {}
    shape = (4, 5, 40)
    wl = rand_prim(rng, shape)
    wr = rand_prim(rng, shape)
    flat_l = wl.reshape(5, -1)
    flat_r = wr.reshape(5, -1)
    m = flat_l.shape[1]
    flat_r[:, 0:40] = flat_l[:, 0:40]                 # equal states
    for k in (1, 2, 3):                                # supersonic both ways along each axis
        flat_l[k, 40 * k:40 * k + 20] = 4.0
        flat_r[k, 40 * k:40 * k + 20] = 4.5
        flat_l[k, 40 * k + 20:40 * k + 40] = -4.0
        flat_r[k, 40 * k + 20:40 * k + 40] = -4.5
    flat_l[1:4, 160:200] = 0.0                         # fluid at rest
    flat_r[1:4, 160:200] = 0.0
    assert m >= 200
    d['wl'], d['wr'] = wl, wr
    for axis, fn in (('x', roe.riemannSolverX_hydro_roe), ('y', roe.riemannSolverY_hydro_roe),
                     ('z', roe.riemannSolverZ_hydro_roe)):
        for efix in (False, True):
            d['f%s_%d' % (axis, efix)] = fn(wl.copy(), wr.copy(), g, efix, 0.5)
    np.savez_compressed(os.path.join(OUT, 'hydro_roe.npz'), **d)

    # ---------------- PPM ----------------
    d = {}
    w1 = smooth_prim(rng, (24,))
    d['w1'] = w1
    d['x1_0'] = np.stack(ppm.ppmReconstructX_hydro(w1.copy(), g, 0.004, 1.0 / 24, 0))
    w2 = smooth_prim(rng, (12, 14))
    d['w2'] = w2
    w3 = smooth_prim(rng, (6, 7, 8))
    d['w3'] = w3
    for bc in (0, 1):
        d['x2_%d' % bc] = np.stack(ppm.ppmReconstructX_hydro(w2.copy(), g, 0.004, 1.0 / 14, bc))
        d['y2_%d' % bc] = np.stack(ppm.ppmReconstructY_hydro(w2.copy(), g, 0.004, 1.0 / 12, bc))
        d['x3_%d' % bc] = np.stack(ppm.ppmReconstructX_hydro(w3.copy(), g, 0.004, 1.0 / 7, bc))
        d['y3_%d' % bc] = np.stack(ppm.ppmReconstructY_hydro(w3.copy(), g, 0.004, 1.0 / 6, bc))
        d['z3_%d' % bc] = np.stack(ppm.ppmReconstructZ_hydro(w3.copy(), g, 0.004, 1.0 / 8, bc))
    np.savez_compressed(os.path.join(OUT, 'hydro_ppm.npz'), **d)

    # ---------------- evolve trajectories ----------------
    d = {}
    names = ('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')
    # 1D Sod (64 cells, outflow), orders 0 and 2; a periodic smooth case at order 0
    nx = 64
    x = (np.arange(nx) + 0.5) / nx
    pr = np.zeros((5, nx))
    pr[0] = np.where(x <= 0.5, 1.0, 0.125)
    pr[4] = np.where(x <= 0.5, 1.0, 0.1)
    cases = [('sod_o0', pr, 0, 0, 0.2, 40), ('sod_o2', pr, 0, 2, 0.2, 40),
             ('per1_o0', smooth_prim(rng, (48,)), 1, 0, 10.0, 15)]
    for tag, p0, bc, order, tend, ncyc in cases:
        gr = Grid()
        gr.primVars, gr.consVars = p0.copy(), pc.primToCons_hydro(p0, g)
        gr.dx, gr.BcX = 1.0 / p0.shape[1], bc
        gr.numXCells = p0.shape[1]
        with quiet:
            r = evo.evolve_1D_hydro(gr, 0.0, tend, 0, ncyc, g, 0.8, 1e-12, 1e-12, 1e-12, order, order == 2, 0.5)
        d[tag + '_in'] = p0
        for i, name in enumerate(names):
            d[tag + '_' + name] = np.asarray(r[i])
    # 2D
    for tag, (ny, nx), (bcx, bcy), order, ncyc in (('p2o2', (16, 16), (1, 1), 2, 6),
                                                    ('p2o0', (16, 16), (1, 1), 0, 6),
                                                    ('m2o2', (12, 14), (0, 1), 2, 5),
                                                    ('o2o2', (12, 14), (0, 0), 2, 5)):
        p0 = smooth_prim(rng, (ny, nx))
        gr = Grid()
        gr.primVars, gr.consVars = p0.copy(), pc.primToCons_hydro(p0, g)
        gr.dx, gr.dy, gr.BcX, gr.BcY = 1.0 / nx, 1.0 / ny, bcx, bcy
        gr.numXCells, gr.numYCells = nx, ny
        with quiet:
            r = evo.evolve_2D_hydro(gr, 0.0, 10.0, 0, ncyc, g, 0.8, 1e-12, 1e-12, 1e-12, order, False, 0.5)
        d[tag + '_in'] = p0
        for i, name in enumerate(names):
            d[tag + '_' + name] = np.asarray(r[i])
    # 3D (outflow x raises in the reference, so x stays periodic)
    for tag, shape, (bcx, bcy, bcz), order, ncyc in (('p3o2', (8, 8, 8), (1, 1, 1), 2, 3),
                                                      ('p3o0', (8, 8, 8), (1, 1, 1), 0, 3),
                                                      ('m3o2', (6, 8, 10), (1, 0, 0), 2, 2)):
        p0 = smooth_prim(rng, shape)
        gr = Grid()
        gr.primVars, gr.consVars = p0.copy(), pc.primToCons_hydro(p0, g)
        gr.dx, gr.dy, gr.dz = 1.0 / shape[1], 1.0 / shape[0], 1.0 / shape[2]
        gr.BcX, gr.BcY, gr.BcZ = bcx, bcy, bcz
        gr.numYCells, gr.numXCells, gr.numZCells = shape
        with quiet:
            r = evo.evolve_3D_hydro(gr, 0.0, 10.0, 0, ncyc, g, 0.4, 1e-12, 1e-12, 1e-12, order, True, 0.5)
        d[tag + '_in'] = p0
        for i, name in enumerate(names):
            d[tag + '_' + name] = np.asarray(r[i])
    np.savez_compressed(os.path.join(OUT, 'hydro_evolve.npz'), **d)
    print('hydro golden vectors written to', OUT)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/PythonMHD')
