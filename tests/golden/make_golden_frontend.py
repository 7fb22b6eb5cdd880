"""Write golden vectors of the reference's initialisation path: SimulationGrid
(Source/SimulationGrid.py:47-169) built from init-file style inputs -- the
formatted primitive matrix, consVars from primVars (primToCons, :96-103), the face
fields from cell B (getFaceB*, Source/CT_HelperModule.py:29-103), widths and
coordinates (Source/SimulationGrid_HelperModule.py:414-482) -- and which inputs
the grid's checks reject (RuntimeError).

Run once, in the container that holds the reference checkout:

    python tests/golden/make_golden_frontend.py /root/reference/PythonMHD

Imports Source.SimulationGrid and Source.Initialization_HelperModule (NumPy
only); never Source.Simulation, whose file output needs modules that are absent.
Only the numeric results are committed (tests/golden/frontend.npz).
"""
import contextlib
import io
import os
import sys

sys.dont_write_bytecode = True

import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    sys.path.insert(0, ref_root)
    import Source.PythonMHD_Constants as c
    from Source.SimulationGrid import SimulationGrid
    import Source.Initialization_HelperModule as ih
    rng = np.random.default_rng(77)
    d = {}
    z1 = np.zeros(shape=(1))

    def grid_par(shape, bcs, lo=(0.0, -0.5, 0.25), hi=(1.0, 0.5, 1.75)):
        nd = len(shape)
        n = (shape[0],) if nd == 1 else ((shape[1], shape[0]) if nd == 2 else (shape[1], shape[0], shape[2]))
        gp = {}
        for a, (nk, mn, mx, bk) in enumerate(((c.NUM_X_CELLS, c.MIN_X, c.MAX_X, c.BC_X),
                                              (c.NUM_Y_CELLS, c.MIN_Y, c.MAX_Y, c.BC_Y),
                                              (c.NUM_Z_CELLS, c.MIN_Z, c.MAX_Z, c.BC_Z))[:nd]):
            gp[nk], gp[mn], gp[mx], gp[bk] = n[a], lo[a], hi[a], bcs[a]
        return gp

    cases = [('m1', (24,), (0,), True, 8), ('m2', (10, 12), (1, 0), True, 8), ('m3', (6, 8, 5), (0, 1, 1), True, 8),
             ('h1', (20,), (1,), False, 3), ('h1f', (20,), (0,), False, 5), ('h2', (8, 9), (0, 1), False, 5),
             ('h3', (5, 6, 7), (1, 1, 0), False, 5)]
    quiet = contextlib.redirect_stdout(io.StringIO())
    for tag, shape, bcs, mhd, nv in cases:
        p = rng.uniform(-1, 1, (nv,) + shape)
        p[0] = rng.uniform(0.5, 2.0, shape)
        p[nv - 1] = rng.uniform(0.5, 2.0, shape)
        gp = grid_par(shape, bcs)
        with quiet:
            g = SimulationGrid(p.copy(), gp, mhd, 1.4, 1e-12, 1e-12, 1e-12, z1, z1, z1, z1)
        d[tag + '_in'] = p
        d[tag + '_prim'] = g.primVars
        d[tag + '_cons'] = g.consVars
        d[tag + '_widths'] = np.array([g.dx] + ([g.dy] if g.nDim >= 2 else []) + ([g.dz] if g.nDim == 3 else []))
        d[tag + '_bcs'] = np.array([g.BcX, g.BcY, g.BcZ])
        d[tag + '_x'] = g.xCoords
        if g.nDim >= 2:
            d[tag + '_y'] = g.yCoords
        if g.nDim == 3:
            d[tag + '_z'] = g.zCoords
        if mhd and g.nDim >= 2:
            for nm in ('faceBx', 'faceBy', 'faceBz')[:g.nDim]:
                d[tag + '_' + nm] = getattr(g, nm)
    # coordinates of init files (Initialization_HelperModule), including a range whose
    # np.r_ construction would give one extra point if rounding went the other way
    for tag, gp in (('c1', {c.NUM_X_CELLS: 7, c.MIN_X: -0.3, c.MAX_X: 0.4}),
                    ('c2', {c.NUM_X_CELLS: 5, c.MIN_X: 0.0, c.MAX_X: 1.0, c.NUM_Y_CELLS: 3, c.MIN_Y: 1, c.MAX_Y: 2.5}),
                    ('c3', {c.NUM_X_CELLS: 3, c.MIN_X: 0.1, c.MAX_X: 0.7, c.NUM_Y_CELLS: 4, c.MIN_Y: 0.0,
                            c.MAX_Y: 1.0, c.NUM_Z_CELLS: 2, c.MIN_Z: -1.0, c.MAX_Z: 1.0})):
        nd = int(tag[1])
        fn = {1: ih.createCoordinates1D_Cartesian, 2: ih.createCoordinates2D_Cartesian,
              3: ih.createCoordinates3D_Cartesian}[nd]
        r = fn(gp)
        r = (r,) if nd == 1 else r
        for a, arr in zip('xyz', r):
            d['%s_%s' % (tag, a)] = arr
    # inputs the grid's checks reject: (tag, raises)
    bad = {}
    base = grid_par((8, 8), (1, 1))
    p8 = np.ones((8, 8, 8))
    tries = {
        'cells_not_int': (p8, dict(base, **{c.NUM_X_CELLS: 8.0}), True),
        'cells_mismatch': (p8, dict(base, **{c.NUM_X_CELLS: 9}), True),
        'one_cell': (np.ones((8, 8, 1)), dict(base, **{c.NUM_X_CELLS: 1}), True),
        'range_reversed': (p8, dict(base, **{c.MIN_Y: 1.0, c.MAX_Y: 0.0}), True),
        'range_type': (p8, dict(base, **{c.MAX_X: '1'}), True),
        'neg_density': (np.concatenate([-np.ones((1, 8, 8)), np.ones((7, 8, 8))]), base, True),
        'low_pressure': (np.concatenate([np.ones((7, 8, 8)), 1e-13 * np.ones((1, 8, 8))]), base, True),
        'mhd_seven_vars': (np.ones((7, 8, 8)), base, True),
        'ok': (p8, base, True),
    }
    for name, (p, gp, mhd) in tries.items():
        try:
            with quiet:
                SimulationGrid(p.copy(), gp, mhd, 1.4, 1e-12, 1e-12, 1e-12, z1, z1, z1, z1)
            bad[name] = 0
        except RuntimeError:
            bad[name] = 1
    d['reject_names'] = np.array(sorted(bad))
    d['reject_flags'] = np.array([bad[k] for k in sorted(bad)])
    np.savez_compressed(os.path.join(OUT, 'frontend.npz'), **d)
    print('front-end golden vectors written to', OUT, bad)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/PythonMHD')
