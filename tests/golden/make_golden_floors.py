"""Write golden trajectories of the reference's evolve functions in which the
density, pressure and energy floors fire (Source/PrimCons_HelperModule.py:499-554,
called at Source/EvolveFunctions_MHD.py:278-280, 513-519, 562-576, 911-922 and
Source/EvolveFunctions_Hydro.py:286-288, 437-443, 675-685).

Run once, in the container that holds the reference checkout:

    python tests/golden/make_golden_floors.py /root/reference/PythonMHD

It imports PythonMHD's NumPy modules (never ``Source.Simulation``), raises the
floors above parts of seeded states so that they bite, and saves inputs, outputs
and the number of values each applyFloors call floored (per cycle, per call site,
density and last variable) as ``tests/golden/floors.npz``.  The applyFloors
wrapper only counts; the reference's own function does the flooring.

3D calls that floor three or more values of one variable raise IndexError in the
reference (its message loop reuses ``i``, PrimCons_HelperModule.py:539-543); the
3D trajectories keep at most two per call, and one extra case records the raise.
"""
import contextlib
import io
import os
import sys

sys.dont_write_bytecode = True    # import the reference read-only: no __pycache__ in its tree

import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))

# call sites in the order the build's floor counters use: predictor (left, right)
# per direction x, y, z, then the 2D post-CTU sites (left, right) per direction
SITES_MHD = {
    'the x-direction reconstruction outputs (left)': 0,
    'the x-direction reconstruction outputs (right)': 1,
    'the y-direction reconstruction outputs (top)': 2,
    'the y-direction reconstruction outputs (bottom)': 3,
    'the z-direction reconstruction outputs (back)': 4,
    'the z-direction reconstruction outputs (forw)': 5,
    'the x-direction reconstruction outputs (left) after CTU update': 6,
    'the x-direction reconstruction outputs (right) after CTU update': 7,
    'the y-direction reconstruction outputs (top) after CTU update': 8,
    'the y-direction reconstruction outputs (bottom) after CTU update': 9,
}
NSITES = 10


def _site(msg):
    m = msg.strip()
    if m in SITES_MHD:
        return SITES_MHD[m]
    # the hydro modules word their messages differently: classify by direction/side
    low = m.lower()
    d = 0 if 'x-direction' in low else (1 if 'y-direction' in low else 2)
    right = any(w in low for w in ('right', 'bottom', 'forw'))
    return 2 * d + (1 if right else 0)


class Counter:
    """Per-cycle, per-site counts of floored values (density, last variable)."""

    def __init__(self, pc):
        self.pc = pc
        self.orig = pc.applyFloors
        self.reset()

    def reset(self):
        self.cycles = []
        self.called = set()

    def new_cycle(self):
        self.cycles.append(np.zeros((NSITES, 2), dtype=np.int64))
        self.called = set()

    def __call__(self, vars, isPrimitive, minDens, minPres, minEnergy, calledFromString):
        lo = minPres if isPrimitive else minEnergy
        n0 = int(np.count_nonzero(vars[0] < minDens))
        n1 = int(np.count_nonzero(vars[vars.shape[0] - 1] < lo))
        s = _site(calledFromString)
        # a cycle calls every site once: a site called again starts the next cycle
        if not self.cycles or s in self.called:
            self.new_cycle()
        self.called.add(s)
        self.cycles[-1][s, 0] += n0
        self.cycles[-1][s, 1] += n1
        return self.orig(vars, isPrimitive, minDens, minPres, minEnergy, calledFromString)


def _load(ref_root):
    sys.path.insert(0, ref_root)
    import Source.EvolveFunctions_MHD as evo
    import Source.EvolveFunctions_Hydro as hevo
    import Source.PrimCons_HelperModule as pc
    return evo, hevo, pc


class Grid:
    pass


def smooth_prim(rng, shape, nv=8, amp=0.25):
    grids = np.meshgrid(*[np.arange(n) / n for n in shape], indexing='ij')
    p = np.empty((nv,) + shape)
    base = [1.0, 0.0, 0.0, 0.0, 0.4, 0.3, 0.2, 1.0] if nv == 8 else [1.0, 0.0, 0.0, 0.0, 1.0]
    for v in range(nv):
        f = np.full(shape, base[v])
        for _ in range(3):
            k = rng.integers(1, 3, len(shape))
            ph = rng.uniform(0, 2 * np.pi)
            f = f + amp * rng.uniform(-1, 1) * np.sin(2 * np.pi * sum(ki * g for ki, g in zip(k, grids)) + ph)
        p[v] = f
    p[0] = np.abs(p[0]) + 0.3
    p[nv - 1] = np.abs(p[nv - 1]) + 0.3
    return p


def faces_from_cells(b, axis, bc):
    n = b.shape[axis]
    lo = np.take(b, [(n - 1 if bc == 1 else 0)] + list(range(n)), axis=axis)
    hi = np.take(b, list(range(n)) + [(0 if bc == 1 else n - 1)], axis=axis)
    return (lo + hi) / 2.0


def main(ref_root):
    evo, hevo, pc = _load(ref_root)
    cnt = Counter(pc)
    pc.applyFloors = cnt
    quiet = contextlib.redirect_stdout(io.StringIO())
    rng = np.random.default_rng(2024)
    d = {}

    def record(tag, r, names, floors):
        for i, name in enumerate(names):
            d[tag + '_' + name] = np.asarray(r[i])
        d[tag + '_counts'] = np.stack(cnt.cycles) if cnt.cycles else np.zeros((0, NSITES, 2), np.int64)
        d[tag + '_floors'] = np.asarray(floors, dtype=np.float64)
        tot = d[tag + '_counts'].sum()
        print('%-8s floored values: %d over %d cycles' % (tag, tot, len(cnt.cycles)))
        assert tot > 0, tag + ': no floor fired'

    names1 = ('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')
    names2 = ('prim', 'cons', 'fbx', 'fby', 't', 'cyc', 'prim2', 'cons2', 'fbx2', 'fby2', 't2')
    names3 = ('prim', 'cons', 'fbx', 'fby', 'fbz', 't', 'cyc', 'prim2', 'cons2', 'fbx2', 'fby2', 'fbz2', 't2')

    # ---- 1D MHD: Brio-Wu (64 cells, outflow) with the density and pressure floors
    # above the right state, orders 0 and 2; a periodic PPM case
    nx = 64
    x = (np.arange(nx) + 0.5) / nx
    bw = np.zeros((8, nx))
    bw[0] = np.where(x <= 0.5, 1.0, 0.125)
    bw[7] = np.where(x <= 0.5, 1.0, 0.1)
    bw[5] = np.where(x <= 0.5, 1.0, -1.0)
    bw[4] = 0.75
    for order in (0, 2):
        floors = (0.2, 0.3, 1e-12)
        gr = Grid()
        gr.primVars, gr.consVars = bw.copy(), pc.primToCons_mhd_allVars(bw, 2.0)
        gr.dx, gr.BcX = 1.0 / nx, 0
        cnt.reset()
        with quiet:
            r = evo.evolve_1D_mhd(gr, 0.0, 0.1, 0, 12, 2.0, 0.8, *floors, order, False, 0.5)
        d['m1o%d_in' % order] = bw
        record('m1o%d' % order, r, names1, floors)
    p1 = smooth_prim(rng, (48,))
    p1[4] = 0.6
    floors = (float(np.quantile(p1[0], 0.3)), float(np.quantile(p1[7], 0.3)), 1e-12)
    gr = Grid()
    gr.primVars, gr.consVars = p1.copy(), pc.primToCons_mhd_allVars(p1, 5.0 / 3.0)
    gr.dx, gr.BcX = 1.0 / 48, 1
    cnt.reset()
    with quiet:
        r = evo.evolve_1D_mhd(gr, 0.0, 10.0, 0, 8, 5.0 / 3.0, 0.8, *floors, 2, True, 0.5)
    d['m1p_in'] = p1
    record('m1p', r, names1, floors)

    # ---- 2D MHD: predictor floors (density, pressure) and the post-CTU density and
    # energy floors, PPM periodic and first-order outflow
    g = 5.0 / 3.0
    for tag, (ny, nx), bc, order, ncyc in (('m2o2', (16, 16), 1, 2, 4), ('m2o0', (12, 14), 0, 0, 4)):
        pr = smooth_prim(rng, (ny, nx))
        fbx = faces_from_cells(pr[4], 1, bc)
        fby = faces_from_cells(pr[5], 0, bc)
        pr[4] = (fbx[:, :-1] + fbx[:, 1:]) / 2.0
        pr[5] = (fby[:-1, :] + fby[1:, :]) / 2.0
        cons = pc.primToCons_mhd_allVars(pr, g)
        floors = (float(np.quantile(pr[0], 0.25)), float(np.quantile(pr[7], 0.25)),
                  float(np.quantile(cons[7], 0.35)))
        gr = Grid()
        gr.primVars, gr.consVars, gr.faceBx, gr.faceBy = pr.copy(), cons.copy(), fbx.copy(), fby.copy()
        gr.dx, gr.dy, gr.BcX, gr.BcY = 1.0 / nx, 1.0 / ny, bc, bc
        cnt.reset()
        with quiet:
            r = evo.evolve_2D_mhd(gr, 0.0, 10.0, 0, ncyc, g, 0.8, *floors, order, False, 0.5)
        d[tag + '_in_prim'], d[tag + '_in_cons'], d[tag + '_in_fbx'], d[tag + '_in_fby'] = pr, cons, fbx, fby
        record(tag, r, names2, floors)
        assert d[tag + '_counts'][:, 6:10, 1].sum() > 0, tag + ': the post-CTU energy floor did not fire'

    # ---- 3D MHD: single-cell dips of pressure and density below the floors, so
    # that each applyFloors call floors at most two values of each variable
    for tag, shape, bc, order, ncyc in (('m3o2', (8, 8, 8), 1, 2, 2), ('m3o0', (8, 8, 8), 1, 0, 2),
                                        ('m3o1', (6, 8, 10), 1, 0, 1)):
        pr = smooth_prim(rng, shape, amp=0.1)
        pr[7][3, 4, 5] = 0.01
        pr[0][5, 2, 2] = 0.02
        fbx = faces_from_cells(pr[4], 1, bc)
        fby = faces_from_cells(pr[5], 0, bc)
        fbz = faces_from_cells(pr[6], 2, bc)
        pr[4] = 0.5 * (fbx[:, :-1, :] + fbx[:, 1:, :])
        pr[5] = 0.5 * (fby[:-1, :, :] + fby[1:, :, :])
        pr[6] = 0.5 * (fbz[:, :, :-1] + fbz[:, :, 1:])
        cons = pc.primToCons_mhd_allVars(pr, g)
        floors = (0.05, 0.03, 1e-12)
        gr = Grid()
        gr.primVars, gr.consVars = pr.copy(), cons.copy()
        gr.faceBx, gr.faceBy, gr.faceBz = fbx.copy(), fby.copy(), fbz.copy()
        gr.dx, gr.dy, gr.dz = 1.0 / shape[1], 1.0 / shape[0], 1.0 / shape[2]
        gr.BcX = gr.BcY = gr.BcZ = bc
        cnt.reset()
        with quiet:
            r = evo.evolve_3D_mhd(gr, 0.0, 10.0, 0, ncyc, g, 0.4, *floors, order, True, 0.5)
        for nm, a in (('prim', pr), ('cons', cons), ('fbx', fbx), ('fby', fby), ('fbz', fbz)):
            d[tag + '_in_' + nm] = a
        record(tag, r, names3, floors)
        assert d[tag + '_counts'].max() <= 2

    # ---- 3D MHD with three floored values in one call: the reference raises
    pr = smooth_prim(rng, (8, 8, 8), amp=0.1)
    pr[7][3, 4, 2:5] = 0.01
    fbx = faces_from_cells(pr[4], 1, 1)
    fby = faces_from_cells(pr[5], 0, 1)
    fbz = faces_from_cells(pr[6], 2, 1)
    pr[4] = 0.5 * (fbx[:, :-1, :] + fbx[:, 1:, :])
    pr[5] = 0.5 * (fby[:-1, :, :] + fby[1:, :, :])
    pr[6] = 0.5 * (fbz[:, :, :-1] + fbz[:, :, 1:])
    cons = pc.primToCons_mhd_allVars(pr, g)
    gr = Grid()
    gr.primVars, gr.consVars = pr.copy(), cons.copy()
    gr.faceBx, gr.faceBy, gr.faceBz = fbx.copy(), fby.copy(), fbz.copy()
    gr.dx = gr.dy = gr.dz = 1.0 / 8
    gr.BcX = gr.BcY = gr.BcZ = 1
    raised = ''
    try:
        with quiet:
            evo.evolve_3D_mhd(gr, 0.0, 10.0, 0, 1, g, 0.4, 0.05, 0.03, 1e-12, 0, True, 0.5)
    except Exception as e:  # noqa: BLE001 -- the type is what is recorded
        raised = type(e).__name__
    print('m3raise: the reference raised', raised or 'nothing')
    assert raised == 'IndexError'
    for nm, a in (('prim', pr), ('cons', cons), ('fbx', fbx), ('fby', fby), ('fbz', fbz)):
        d['m3raise_in_' + nm] = a
    d['m3raise_floors'] = np.array([0.05, 0.03, 1e-12])

    # ---- hydro: 1D Sod (orders 0, 2), 2D periodic PPM, 3D periodic PPM with dips
    hnames = ('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')
    nx = 64
    x = (np.arange(nx) + 0.5) / nx
    sod = np.zeros((5, nx))
    sod[0] = np.where(x <= 0.5, 1.0, 0.125)
    sod[4] = np.where(x <= 0.5, 1.0, 0.1)
    for order in (0, 2):
        floors = (0.2, 0.3, 1e-12)
        gr = Grid()
        gr.primVars, gr.consVars = sod.copy(), pc.primToCons_hydro(sod, 1.4)
        gr.dx, gr.BcX, gr.numXCells = 1.0 / nx, 0, nx
        cnt.reset()
        with quiet:
            r = hevo.evolve_1D_hydro(gr, 0.0, 0.2, 0, 20, 1.4, 0.8, *floors, order, order == 2, 0.5)
        d['h1o%d_in' % order] = sod
        record('h1o%d' % order, r, hnames, floors)
    p2 = smooth_prim(rng, (16, 16), nv=5)
    floors = (float(np.quantile(p2[0], 0.3)), float(np.quantile(p2[4], 0.3)), 1e-12)
    gr = Grid()
    gr.primVars, gr.consVars = p2.copy(), pc.primToCons_hydro(p2, g)
    gr.dx, gr.dy, gr.BcX, gr.BcY = 1.0 / 16, 1.0 / 16, 1, 1
    gr.numXCells, gr.numYCells = 16, 16
    cnt.reset()
    with quiet:
        r = hevo.evolve_2D_hydro(gr, 0.0, 10.0, 0, 5, g, 0.8, *floors, 2, False, 0.5)
    d['h2o2_in'] = p2
    record('h2o2', r, hnames, floors)
    p3 = smooth_prim(rng, (8, 8, 8), nv=5, amp=0.1)
    p3[4][2, 3, 4] = 0.01
    p3[0][6, 1, 1] = 0.02
    floors = (0.05, 0.03, 1e-12)
    gr = Grid()
    gr.primVars, gr.consVars = p3.copy(), pc.primToCons_hydro(p3, g)
    gr.dx = gr.dy = gr.dz = 1.0 / 8
    gr.BcX = gr.BcY = gr.BcZ = 1
    gr.numYCells = gr.numXCells = gr.numZCells = 8
    cnt.reset()
    with quiet:
        r = hevo.evolve_3D_hydro(gr, 0.0, 10.0, 0, 2, g, 0.4, *floors, 2, True, 0.5)
    d['h3o2_in'] = p3
    record('h3o2', r, hnames, floors)
    assert d['h3o2_counts'].max() <= 2

    np.savez_compressed(os.path.join(OUT, 'floors.npz'), **d)
    print('floor golden vectors written to', OUT)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/PythonMHD')
