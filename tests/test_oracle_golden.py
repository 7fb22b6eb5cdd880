"""The oracle against the golden vectors written from the reference's own helper
modules (tests/golden/make_golden.py).  Every comparison is bit-exact."""
import os

import numpy as np
import pytest

from oracle import mhd as M

G = os.path.join(os.path.dirname(__file__), 'golden')
GAMMA = 5.0 / 3.0


def load(name):
    return np.load(os.path.join(G, name))


def same(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape
    assert np.array_equal(a, b, equal_nan=True), float(np.nanmax(np.abs(a - b)))


def test_calc_dt():
    c = load('calc_dt.npz')
    same(M.calc_dt(c['cons1'], (), GAMMA, 0.8, (0.03,)), c['dt1'])
    same(M.calc_dt(c['cons2'], (c['fbx2'], c['fby2']), GAMMA, 0.8, (0.05, 0.07)), c['dt2'])
    same(M.calc_dt(c['cons3'], (c['fbx3'], c['fby3'], c['fbz3']), GAMMA, 0.4, (0.1, 0.11, 0.09)), c['dt3'])


@pytest.mark.parametrize('d', 'xyz')
@pytest.mark.parametrize('ef', [0, 1])
def test_roe(d, ef):
    r = load('roe.npz')
    f = M.roe_flux(r[d + '_wl'], r[d + '_wr'], r[d + '_ul'], r[d + '_ur'], r[d + '_bn'],
                   GAMMA, bool(ef), 0.5, d)
    same(f, r['%s_ef%d' % (d, ef)])


@pytest.mark.parametrize('bc', [0, 1])
def test_ppm(bc):
    p = load('ppm.npz')
    pr = p['p3_bc%d' % bc]
    for d, dn, ax in (('x', 0.1, 1), ('y', 0.12, 0), ('z', 0.09, 2)):
        w = pr[list(M.DIRVARS[d])]
        same(np.stack(M.recon_states(w, pr[M.BNVAR[d]], GAMMA, 0.01, dn, bc, d, ax, 2)),
             p['%s3_bc%d' % (d, bc)])
    pr = p['p2_bc%d' % bc]
    for d, ax in (('x', 1), ('y', 0)):
        w = pr[list(M.DIRVARS[d])]
        same(np.stack(M.recon_states(w, pr[M.BNVAR[d]], GAMMA, 0.02, 0.1, bc, d, ax, 2)),
             p['%s2_bc%d' % (d, bc)])
    pr = p['p1_bc%d' % bc]
    same(np.stack(M.recon_states(pr[list(M.DIRVARS['x'])], pr[4], 2.0, 0.003, 0.025, bc, 'x', 0, 2)),
         p['x1_bc%d' % bc])


@pytest.mark.parametrize('bc', [0, 1])
def test_helpers_3d(bc):
    h = load('helpers3d.npz')
    k = 'bc%d_' % bc
    pr, cons = h[k + 'prim'], h[k + 'cons']
    ux = cons[[0, 1, 2, 3, 5, 6, 7]]
    px = pr[[0, 1, 2, 3, 5, 6, 7]]
    py = pr[[0, 1, 2, 3, 6, 4, 7]]
    fb = (h[k + 'fbx'], h[k + 'fby'], h[k + 'fbz'])
    for d in 'xyz':
        same(M.recon_b_terms_3d(d, ux, px[0], *fb, 0.013, 0.1, 0.12, 0.09, bc), h[k + 'rec' + d])
    F, Gf, H, Ec = h[k + 'F'], h[k + 'G'], h[k + 'H'], h[k + 'Ec']
    same(M.corner_emf_3d_x(Gf, H, Ec[0], bc, bc, bc), h[k + 'Ex'])
    same(M.corner_emf_3d_y(F, H, Ec[1], bc, bc, bc), h[k + 'Ey'])
    same(M.corner_emf_3d_z(F, Gf, Ec[2], bc, bc, bc), h[k + 'Ez'])
    flux = {'x': F, 'y': Gf, 'z': H}
    for d in 'xyz':
        out = M.ctu_transverse_3d(d, h[k + 'UL' + d], h[k + 'UR' + d], px, py, flux, fb,
                                  h[k + 'E' + d], 0.013, 0.1, 0.12, 0.09, bc)
        same(np.stack(out), h[k + 'ctu' + d])


@pytest.mark.parametrize('bc', [0, 1])
def test_helpers_2d(bc):
    h = load('helpers2d.npz')
    k = 'bc%d_' % bc
    pr = h[k + 'prim']
    px = pr[[0, 1, 2, 3, 5, 6, 7]]
    py = pr[[0, 1, 2, 3, 6, 4, 7]]
    same(np.stack(M.recon_b_terms_2d(px[2], h[k + 'fbx'], 0.013, 0.1, bc, 1)), h[k + 'srcby'])
    same(np.stack(M.recon_b_terms_2d(px[1], h[k + 'fby'], 0.013, 0.12, bc, 0)), h[k + 'srcbx'])
    same(M.corner_emf_2d(h[k + 'F'], h[k + 'G'], h[k + 'Ec'], bc, bc), h[k + 'emf'])
    same(np.stack(M.ctu_transverse_2d_x(px, py, h[k + 'G'], h[k + 'fbx'], 0.013, 0.1, 0.12, bc)), h[k + 'ctux'])
    same(np.stack(M.ctu_transverse_2d_y(px, py, h[k + 'F'], h[k + 'fby'], 0.013, 0.1, 0.12, bc)), h[k + 'ctuy'])


@pytest.mark.parametrize('order', [0, 2])
def test_evolve_1d_briowu(order):
    e = load('evolve.npz')
    pin = e['bw_o%d_in' % order]
    P = M.Params(2.0, 0.8, order=order)
    r = M.evolve_1d(pin, M.prim_to_cons_all(pin, 2.0), 1 / 64, 0, 0.0, 0.1, 0, 25, P)
    for i, n in enumerate(('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')):
        same(r[i], e['bw_o%d_%s' % (order, n)])


def test_evolve_1d_periodic_entropy_fix():
    e = load('evolve.npz')
    P = M.Params(GAMMA, 0.8, order=2, entropy_fix=True)
    r = M.evolve_1d(e['p1_in'], M.prim_to_cons_all(e['p1_in'], GAMMA), 1 / 48, 1, 0.0, 10.0, 0, 15, P)
    for i, n in enumerate(('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')):
        same(r[i], e['p1_%s' % n])


@pytest.mark.parametrize('tag,shape,bc,order,ncyc', [
    ('p2o2', (16, 16), 1, 2, 6), ('p2o0', (16, 16), 1, 0, 6), ('o2o2', (12, 14), 0, 2, 5)])
def test_evolve_2d(tag, shape, bc, order, ncyc):
    e = load('evolve.npz')
    ny, nx = shape
    P = M.Params(GAMMA, 0.8, order=order)
    r = M.evolve_2d(e[tag + '_in_prim'], e[tag + '_in_cons'], e[tag + '_in_fbx'], e[tag + '_in_fby'],
                    1 / nx, 1 / ny, bc, bc, 0.0, 10.0, 0, ncyc, P)
    for i, n in enumerate(('prim', 'cons', 'fbx', 'fby', 't', 'cyc', 'prim2', 'cons2', 'fbx2', 'fby2', 't2')):
        same(r[i], e[tag + '_' + n])


@pytest.mark.parametrize('tag,shape,bc,order,ncyc', [
    ('p3o2', (8, 8, 8), 1, 2, 3), ('p3o0', (8, 8, 8), 1, 0, 3),
    ('n3o2', (6, 8, 10), 1, 2, 2), ('o3o2', (6, 6, 6), 0, 2, 2)])
def test_evolve_3d(tag, shape, bc, order, ncyc):
    e = load('evolve.npz')
    P = M.Params(GAMMA, 0.4, order=order, entropy_fix=True)
    r = M.evolve_3d(e[tag + '_in_prim'], e[tag + '_in_cons'], e[tag + '_in_fbx'], e[tag + '_in_fby'],
                    e[tag + '_in_fbz'], 1 / shape[1], 1 / shape[0], 1 / shape[2], bc, bc, bc,
                    0.0, 10.0, 0, ncyc, P)
    names = ('prim', 'cons', 'fbx', 'fby', 'fbz', 't', 'cyc',
             'prim2', 'cons2', 'fbx2', 'fby2', 'fbz2', 't2')
    for i, n in enumerate(names):
        same(r[i], e[tag + '_' + n])
