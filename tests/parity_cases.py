"""Shared inputs and runners of the parity tests (host build and GPU)."""
import os

import numpy as np

from oracle import mhd as M

G = os.path.join(os.path.dirname(__file__), 'golden')
GAMMA = 5.0 / 3.0

# (tag, shape, bc, order, ncyc) of the golden trajectories in tests/golden/evolve.npz
GOLD_2D = [('p2o2', (16, 16), 1, 2, 6), ('p2o0', (16, 16), 1, 0, 6), ('o2o2', (12, 14), 0, 2, 5)]
GOLD_3D = [('p3o2', (8, 8, 8), 1, 2, 3), ('p3o0', (8, 8, 8), 1, 0, 3),
           ('n3o2', (6, 8, 10), 1, 2, 2), ('o3o2', (6, 6, 6), 0, 2, 2)]


def golden():
    return np.load(os.path.join(G, 'evolve.npz'))


def run_solver(Solver, lib, shape, widths, bcs, prim, cons, faces, t_end, ncyc, precision=64, **kw):
    s = Solver(shape, widths, bcs, kw.pop('gamma', GAMMA), kw.pop('cfl'), precision=precision, lib=lib, **kw)
    s.set_state(prim, cons, faces)
    t, cyc, t2 = s.advance(0.0, t_end, 0, ncyc)
    p, c, f = s.get_state(0)
    p2, c2, f2 = s.get_state(1)
    s.close()
    return p, c, f, t, cyc, p2, c2, f2, t2


def smooth_state(rng, shape, bamp=0.3, vamp=0.3):
    """Smooth periodic prim state with face B from a discrete curl (div B = 0)."""
    nd = len(shape)
    g = np.meshgrid(*[np.arange(n) / n for n in shape], indexing='ij')
    prim = np.zeros((8,) + shape)
    ph = rng.uniform(0, 2 * np.pi, 16)
    prim[0] = 1.0 + 0.3 * np.sin(2 * np.pi * g[0] + ph[0]) * np.cos(2 * np.pi * g[-1] + ph[1])
    prim[7] = 1.0 + 0.2 * np.cos(2 * np.pi * (g[0] + g[-1]) + ph[2])
    for v in (1, 2, 3):
        prim[v] = vamp * np.sin(2 * np.pi * (v * g[0] + g[-1]) + ph[2 + v])
    if nd == 1:
        prim[4] = 0.75
        prim[5] = bamp * np.sin(2 * np.pi * g[0] + ph[7])
        prim[6] = bamp * np.cos(2 * np.pi * g[0] + ph[8])
        return prim, ()
    if nd == 2:
        ny, nx = shape
        # A_z at corners (j-1/2, i-1/2), periodic
        yc = (np.arange(ny + 1) / ny)[:, None]
        xc = (np.arange(nx + 1) / nx)[None, :]
        A = bamp / (2 * np.pi) * (np.cos(2 * np.pi * xc + ph[9]) + np.sin(2 * np.pi * yc + ph[10]))
        fbx = (A[1:, :] - A[:-1, :]) * ny + 0.4
        fby = -(A[:, 1:] - A[:, :-1]) * nx + 0.1
        prim[4] = (fbx[:, :-1] + fbx[:, 1:]) / 2.0
        prim[5] = (fby[:-1, :] + fby[1:, :]) / 2.0
        prim[6] = bamp * np.sin(2 * np.pi * g[1] + ph[11])
        return prim, (fbx, fby)
    ny, nx, nz = shape
    fbx = np.full((ny, nx + 1, nz), 0.5)
    fby = np.full((ny + 1, nx, nz), 0.2)
    fbz = np.full((ny, nx, nz + 1), -0.3)
    # add the curl of A = (0, 0, a(x, y)) for a divergence-free perturbation
    yc = (np.arange(ny + 1) / ny)[:, None, None]
    xc = (np.arange(nx + 1) / nx)[None, :, None]
    a = bamp / (2 * np.pi) * (np.cos(2 * np.pi * xc + ph[9]) * np.sin(2 * np.pi * yc + ph[10]))
    a = np.broadcast_to(a, (ny + 1, nx + 1, nz))
    fbx = fbx + (a[1:, :, :] - a[:-1, :, :]) * ny
    fby = fby - (a[:, 1:, :] - a[:, :-1, :]) * nx
    prim[4] = 0.5 * (fbx[:, :-1, :] + fbx[:, 1:, :])
    prim[5] = 0.5 * (fby[:-1, :, :] + fby[1:, :, :])
    prim[6] = 0.5 * (fbz[:, :, :-1] + fbz[:, :, 1:])
    return prim, (fbx, fby, fbz)


def widths_of(shape):
    """Cell widths (dx[, dy[, dz]]) of a unit box with the reference layout's shape."""
    if len(shape) == 1:
        return (1.0 / shape[0],)
    if len(shape) == 2:
        return (1.0 / shape[1], 1.0 / shape[0])
    return (1.0 / shape[1], 1.0 / shape[0], 1.0 / shape[2])


def oracle_run(shape, widths, bcs, prim, cons, faces, t_end, ncyc, P):
    nd = len(shape)
    if nd == 1:
        r = M.evolve_1d(prim, cons, widths[0], bcs[0], 0.0, t_end, 0, ncyc, P)
        return r[0], r[1], [], r[2], r[3], r[4], r[5], [], r[6]
    if nd == 2:
        r = M.evolve_2d(prim, cons, faces[0], faces[1], widths[0], widths[1], bcs[0], bcs[1],
                        0.0, t_end, 0, ncyc, P)
        return r[0], r[1], [r[2], r[3]], r[4], r[5], r[6], r[7], [r[8], r[9]], r[10]
    r = M.evolve_3d(prim, cons, faces[0], faces[1], faces[2], widths[0], widths[1], widths[2],
                    bcs[0], bcs[1], bcs[2], 0.0, t_end, 0, ncyc, P)
    return r[0], r[1], [r[2], r[3], r[4]], r[5], r[6], r[7], r[8], [r[9], r[10], r[11]], r[12]


def div_b(faces, widths):
    """max |div B| * dx / max |B| of the face field (the build's dimensionless metric)."""
    if len(faces) == 2:
        fbx, fby = faces
        dx, dy = widths[:2]
        d = (fbx[:, 1:] - fbx[:, :-1]) / dx + (fby[1:, :] - fby[:-1, :]) / dy
        bmax = max(np.max(np.abs(fbx)), np.max(np.abs(fby)))
        return float(np.max(np.abs(d)) * min(dx, dy) / bmax)
    fbx, fby, fbz = faces
    dx, dy, dz = widths
    d = ((fbx[:, 1:, :] - fbx[:, :-1, :]) / dx + (fby[1:, :, :] - fby[:-1, :, :]) / dy
         + (fbz[:, :, 1:] - fbz[:, :, :-1]) / dz)
    bmax = max(np.max(np.abs(fbx)), np.max(np.abs(fby)), np.max(np.abs(fbz)))
    return float(np.max(np.abs(d)) * min(dx, dy, dz) / bmax)
