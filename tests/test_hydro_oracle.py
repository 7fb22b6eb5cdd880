"""The hydro oracle against the reference's golden vectors and invariants (CPU)."""
import os

import numpy as np
import pytest

from oracle import hydro as H

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
G = 1.4


def _load(name):
    return np.load(os.path.join(GOLD, name))


def test_calc_dt_golden():
    d = _load('hydro_calc_dt.npz')
    assert H.calc_dt(d['cons1'], G, 0.8, (0.03,)) == d['dt1']
    assert H.calc_dt(d['cons2'], G, 0.8, (0.05, 0.07)) == d['dt2']
    assert H.calc_dt(d['cons3'], G, 0.4, (0.1, 0.11, 0.09)) == d['dt3']


def test_calc_dt_uniform_closed_form():
    cons = H.prim_to_cons(np.array([1.3, 0.2, -0.4, 0.1, 0.9])[:, None, None] * np.ones((5, 3, 4)), G)
    cs = np.sqrt(G * 0.9 / 1.3)
    want = min(0.5 * 0.1 / (0.2 + cs), 0.5 * 0.2 / (0.4 + cs))
    assert abs(H.calc_dt(cons, G, 0.5, (0.1, 0.2)) - want) <= 1e-15 * want


@pytest.mark.parametrize('axis', ['x', 'y', 'z'])
@pytest.mark.parametrize('efix', [0, 1])
def test_roe_golden(axis, efix):
    d = _load('hydro_roe.npz')
    F = H.roe_flux(d['wl'], d['wr'], G, bool(efix), 0.5, 'xyz'.index(axis))
    assert np.array_equal(F, d['f%s_%d' % (axis, efix)])


@pytest.mark.parametrize('name,bc,d,axis,n', [
    ('x1_0', 0, 0, 0, 24), ('x2_0', 0, 0, 1, 14), ('x2_1', 1, 0, 1, 14), ('y2_0', 0, 1, 0, 12),
    ('y2_1', 1, 1, 0, 12), ('x3_0', 0, 0, 1, 7), ('x3_1', 1, 0, 1, 7), ('y3_0', 0, 1, 0, 6),
    ('y3_1', 1, 1, 0, 6), ('z3_0', 0, 2, 2, 8), ('z3_1', 1, 2, 2, 8)])
def test_ppm_golden(name, bc, d, axis, n):
    g = _load('hydro_ppm.npz')
    w = g['w%s' % name[1]]
    wl, wr = H.recon_states(w, G, 0.004, 1.0 / n, bc, d, axis, 2)
    assert np.array_equal(wl, g[name][0]) and np.array_equal(wr, g[name][1])


EVOLVE = [('sod_o0', 0, 0.2, 40, (0,)), ('sod_o2', 2, 0.2, 40, (0,)), ('per1_o0', 0, 10.0, 15, (1,)),
          ('p2o2', 2, 10.0, 6, (1, 1)), ('p2o0', 0, 10.0, 6, (1, 1)), ('m2o2', 2, 10.0, 5, (0, 1)),
          ('o2o2', 2, 10.0, 5, (0, 0)), ('p3o2', 2, 10.0, 3, (1, 1, 1)), ('p3o0', 0, 10.0, 3, (1, 1, 1)),
          ('m3o2', 2, 10.0, 2, (1, 0, 0))]


def evolve_case(tag):
    """(inputs, params, expected tuple) of one golden hydro trajectory."""
    d = _load('hydro_evolve.npz')
    _, order, tend, ncyc, bcs = next(c for c in EVOLVE if c[0] == tag)
    p0 = d[tag + '_in']
    nd = p0.ndim - 1
    shp = p0.shape[1:]
    widths = {1: lambda: (1.0 / shp[0],), 2: lambda: (1.0 / shp[1], 1.0 / shp[0]),
              3: lambda: (1.0 / shp[1], 1.0 / shp[0], 1.0 / shp[2])}[nd]()
    efix = (order == 2) if nd == 1 else (nd == 3)
    P = H.Params(G, 0.4 if nd == 3 else 0.8, order=order, entropy_fix=efix)
    want = [d[tag + '_' + k] for k in ('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')]
    return p0, widths, bcs, tend, ncyc, P, want


@pytest.mark.parametrize('tag', [c[0] for c in EVOLVE])
def test_evolve_golden(tag):
    p0, widths, bcs, tend, ncyc, P, want = evolve_case(tag)
    r = H.evolve(p0.copy(), H.prim_to_cons(p0, G, 0), widths, bcs, 0.0, tend, 0, ncyc, P)
    for a, b in zip(r, want):
        assert np.array_equal(np.asarray(a), b)


def _jac_prim(w, gamma, d):
    """Closed-form primitive flux Jacobian A (dW/dt + A dW/dn = 0) of direction d."""
    n = 1 + d
    rho, p = w[0], w[4]
    A = np.zeros((5, 5))
    for k in range(5):
        A[k, k] = w[n]
    A[0, n] = rho
    A[n, 4] = 1.0 / rho
    A[4, n] = gamma * p
    return A


@pytest.mark.parametrize('d', [0, 1, 2])
def test_eigen_identities(d):
    rng = np.random.default_rng(5 + d)
    w = np.stack([rng.uniform(0.3, 2, 20), *rng.uniform(-1, 1, (3, 20)), rng.uniform(0.2, 2, 20)])
    lam, L, R = H.eigen(w, G, d)
    for c in range(20):
        Lc, Rc = L[:, :, c], R[:, :, c]
        assert np.allclose(Lc @ Rc, np.eye(5), atol=1e-13)
        A = _jac_prim(w[:, c], G, d)
        assert np.allclose(A @ Rc, Rc @ np.diag(lam[:, c]), atol=1e-13)


@pytest.mark.parametrize('d', [0, 1, 2])
def test_roe_property_and_consistency(d):
    """For hydro the Roe linearisation is exact: F_R - F_L = sum_k lambda_k (L_k.dU) R_k."""
    rng = np.random.default_rng(11 + d)
    wl = np.stack([rng.uniform(0.3, 2, 50), *rng.uniform(-1, 1, (3, 50)), rng.uniform(0.2, 2, 50)])
    wr = np.stack([rng.uniform(0.3, 2, 50), *rng.uniform(-1, 1, (3, 50)), rng.uniform(0.2, 2, 50)])
    lam, Lm, R, FL, FR, UL, UR = H.roe_parts(wl, wr, G, False, 0.5, d)
    alpha = H._project(Lm, UR - UL)
    jump = sum(lam[k] * alpha[k] * R[:, k] for k in range(5))
    assert np.allclose(FR - FL, jump, rtol=1e-12, atol=1e-12)
    # consistency: equal states give the physical flux
    F = H.roe_flux(wl, wl, G, False, 0.5, d)
    assert np.allclose(F, FL, rtol=0, atol=1e-14)


def test_plm_constant_and_linear():
    w = np.ones((5, 12)) * np.array([1.0, 0.3, -0.2, 0.1, 0.8])[:, None]
    wl, wr = H.recon_states(w, G, 1e-3, 0.1, 1, 0, 0, 1)
    assert np.array_equal(wl, np.ones((5, 13)) * w[:, :1]) and np.array_equal(wr, wl)
    # linear density, dt -> 0: interior faces carry the linear interpolant
    w = np.ones((5, 16)) * np.array([1.0, 0.0, 0.0, 0.0, 1.0])[:, None]
    w[0] = 1.0 + 0.01 * np.arange(16)
    wl, wr = H.recon_states(w, G, 0.0, 0.1, 0, 0, 0, 1)
    mid = 1.0 + 0.01 * (np.arange(3, 14) - 0.5)
    assert np.allclose(wl[0, 3:14], mid, atol=1e-14) and np.allclose(wr[0, 3:14], mid, atol=1e-14)


def test_reduction_2d_to_1d():
    """y-uniform 2D data evolves as the 1D step (transverse terms vanish)."""
    rng = np.random.default_rng(3)
    p1 = np.stack([1 + 0.2 * np.sin(np.linspace(0, 2 * np.pi, 32, endpoint=False)),
                   0.1 * np.cos(np.linspace(0, 2 * np.pi, 32, endpoint=False)),
                   np.full(32, 0.05), np.zeros(32), np.full(32, 1.0)])
    p2 = np.repeat(p1[:, None, :], 6, axis=1)
    P = H.Params(G, 0.5, order=2)
    dt = 1e-3
    a, _ = H.step(p1, H.prim_to_cons(p1, G), dt, (1 / 32,), (1,), P)
    b, _ = H.step(p2, H.prim_to_cons(p2, G), dt, (1 / 32, 1 / 6), (1, 1), P)
    for j in range(6):
        assert np.allclose(b[:, j, :][[0, 1, 3, 4]], a[[0, 1, 3, 4]], rtol=0, atol=1e-14)
    del rng


@pytest.mark.parametrize('k', range(5))
def test_plm_tracing_single_wave_closed_form(k):
    """Hydro PLM with dt > 0 on linear data along one eigenvector R_k of a uniform
    state: the traced edge states follow the MUSCL-Hancock closed form of the
    characteristic tracing (wave k moving toward a face is corrected with its own
    speed, a wave moving away only with the edge wave's reference state), for the
    acoustic, entropy and shear waves."""
    w0 = np.array([1.1, 0.25, -0.3, 0.15, 0.9])
    lam, L, R = H.eigen(w0[:, None], G, 0)
    lam, r = lam[:, 0], R[:, k, 0]
    assert lam[0] < 0 < lam[4]
    n, eps, dt, dx = 16, 1e-7, 0.03, 0.1
    nu = dt / dx
    w = w0[:, None] + eps * np.arange(n)[None, :] * r[:, None]
    wl, wr = H.recon_states(w, G, dt, dx, 0, 0, 0, 1)
    a_right = 0.5 - 0.5 * nu * (lam[k] if lam[k] >= 0 else lam[4])
    a_left = -0.5 - 0.5 * nu * (lam[k] if lam[k] <= 0 else lam[0])
    tol = 1e-4 * eps
    for i in range(3, n - 3):
        assert np.max(np.abs(wl[:, i + 1] - (w[:, i] + eps * a_right * r))) < tol
        assert np.max(np.abs(wr[:, i] - (w[:, i] + eps * a_left * r))) < tol
