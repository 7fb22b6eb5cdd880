"""The drop-in boundary on the MI355X: the reference-signature evolve functions
(Source/EvolveFunctions_MHD.py:195, 393, 750; Source/EvolveFunctions_Hydro.py:232,
350, 559) through libpmhd.so, the device-resident state entry points with ROCm
tensors, and the torch.autograd wrapper of the adjoint on device tensors."""
import numpy as np
import pytest

from oracle import hydro as H
from oracle import mhd as M
from tests.parity_cases import GAMMA, golden, widths_of

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


class Grid:
    pass


@pytest.fixture
def ev():
    from pythonmhd_amd import evolve
    evolve.release_device_state()
    yield evolve
    evolve.release_device_state()


def test_evolve_mhd_tuples_on_gpu(ev):
    """evolve_{1,2,3}D_mhd return the reference's tuples bit for bit (golden)."""
    e = golden()
    g = Grid()
    g.primVars = e['bw_o2_in']
    g.consVars = M.prim_to_cons_all(g.primVars, 2.0)
    g.dx, g.BcX = 1 / 64, 0
    print('case: evolve_1D_mhd Brio-Wu PPM', flush=True)
    r = ev.evolve_1D_mhd(g, 0.0, 0.1, 0, 25, 2.0, 0.8, 1e-12, 1e-12, 1e-12, 2, False, 0.5)
    for i, n in enumerate(('prim', 'cons', 't', 'cyc', 'prim2', 'cons2', 't2')):
        assert np.array_equal(np.asarray(r[i]), e['bw_o2_' + n]), n
    tag = 'o2o2'
    g = Grid()
    g.primVars, g.consVars = e[tag + '_in_prim'], e[tag + '_in_cons']
    g.faceBx, g.faceBy = e[tag + '_in_fbx'], e[tag + '_in_fby']
    g.dx, g.dy, g.BcX, g.BcY = 1 / 14, 1 / 12, 0, 0
    print('case: evolve_2D_mhd outflow PPM', flush=True)
    r = ev.evolve_2D_mhd(g, 0.0, 10.0, 0, 5, 5 / 3, 0.8, 1e-12, 1e-12, 1e-12, 2, False, 0.5)
    for i, n in enumerate(('prim', 'cons', 'fbx', 'fby', 't', 'cyc', 'prim2', 'cons2', 'fbx2', 'fby2', 't2')):
        assert np.array_equal(np.asarray(r[i]), e[tag + '_' + n]), n
    for tag, shape in (('n3o2', (6, 8, 10)), ('o3o2', (6, 6, 6))):
        g = Grid()
        g.primVars, g.consVars = e[tag + '_in_prim'], e[tag + '_in_cons']
        g.faceBx, g.faceBy, g.faceBz = e[tag + '_in_fbx'], e[tag + '_in_fby'], e[tag + '_in_fbz']
        g.dx, g.dy, g.dz = 1 / shape[1], 1 / shape[0], 1 / shape[2]
        g.BcX = g.BcY = g.BcZ = 1 if tag == 'n3o2' else 0
        print('case: evolve_3D_mhd', tag, flush=True)
        r = ev.evolve_3D_mhd(g, 0.0, 10.0, 0, 2, 5 / 3, 0.4, 1e-12, 1e-12, 1e-12, 2, True, 0.5)
        names = ('prim', 'cons', 'fbx', 'fby', 'fbz', 't', 'cyc', 'prim2', 'cons2', 'fbx2', 'fby2', 'fbz2', 't2')
        for i, n in enumerate(names):
            assert np.array_equal(np.asarray(r[i]), e[tag + '_' + n]), (tag, n)


@pytest.mark.parametrize('tag', ['sod_o2', 'p2o2', 'p3o2'])
def test_evolve_hydro_tuples_on_gpu(ev, tag):
    from tests.test_hydro_oracle import evolve_case
    p0, widths, bcs, tend, ncyc, P, want = evolve_case(tag)
    g = Grid()
    g.primVars, g.consVars = p0, H.prim_to_cons(p0, P.gamma, 0)
    nd = p0.ndim - 1
    for name, w in zip(('dx', 'dy', 'dz'), widths):
        setattr(g, name, w)
    for name, b in zip(('BcX', 'BcY', 'BcZ'), bcs):
        setattr(g, name, b)
    shape = p0.shape[1:]
    g.numXCells = shape[0] if nd == 1 else shape[1]
    if nd >= 2:
        g.numYCells = shape[0]
    if nd == 3:
        g.numZCells = shape[2]
    fn = {1: ev.evolve_1D_hydro, 2: ev.evolve_2D_hydro, 3: ev.evolve_3D_hydro}[nd]
    print('case: evolve hydro', tag, flush=True)
    r = fn(g, 0.0, tend, 0, ncyc, P.gamma, P.cfl, 1e-12, 1e-12, 1e-12, P.order, P.entropy_fix, 0.5)
    for a, b in zip(r, want):
        assert np.array_equal(np.asarray(a), b)


def test_evolve_floors_and_stats_on_gpu(ev):
    """The 2D floor-firing golden trajectory through the drop-in front-end, and its
    floor counters (evolve.last_stats) against the reference's per-call counts."""
    from tests.test_oracle_floors import golden as floor_golden
    d = floor_golden()
    tag = 'm2o2'
    g = Grid()
    g.primVars, g.consVars = d[tag + '_in_prim'], d[tag + '_in_cons']
    g.faceBx, g.faceBy = d[tag + '_in_fbx'], d[tag + '_in_fby']
    g.dx, g.dy, g.BcX, g.BcY = 1 / 16, 1 / 16, 1, 1
    fl = [float(x) for x in d[tag + '_floors']]
    print('case: evolve_2D_mhd floors', flush=True)
    r = ev.evolve_2D_mhd(g, 0.0, 10.0, 0, 4, 5 / 3, 0.8, fl[0], fl[1], fl[2], 2, False, 0.5)
    for i, n in enumerate(('prim', 'cons', 'fbx', 'fby', 't', 'cyc', 'prim2', 'cons2', 'fbx2', 'fby2', 't2')):
        assert np.array_equal(np.asarray(r[i]), d[tag + '_' + n]), n
    st = ev.last_stats()
    assert np.array_equal(st['floored'], d[tag + '_counts'].sum(axis=0))


@pytest.mark.parametrize('precision', [64, 32])
def test_device_state_entry_points(precision):
    """set_state_device / get_state_device with ROCm float64 tensors give the same
    run, bit for bit, as the host-array entry points."""
    from pythonmhd_amd._lib import Solver
    from tests.parity_cases import smooth_state
    rng = np.random.default_rng(21)
    shape = (10, 12, 14)
    prim, faces = smooth_state(rng, shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    w = widths_of(shape)
    print('case: device state fp%d' % precision, flush=True)
    a = Solver(shape, w, (1, 1, 1), GAMMA, 0.4, order=1, precision=precision)
    a.set_state(prim, cons, faces)
    a.advance(0.0, 10.0, 0, 3)
    pa, ca, fa = a.get_state(0)
    p2a, _, _ = a.get_state(1)
    a.close()
    dev = torch.device('cuda')
    b = Solver(shape, w, (1, 1, 1), GAMMA, 0.4, order=1, precision=precision)
    b.set_stream(torch.cuda.current_stream().cuda_stream)
    b.set_state_device(torch.from_numpy(prim).to(dev), torch.from_numpy(cons).to(dev),
                       [torch.from_numpy(f).to(dev) for f in faces])
    b.advance(0.0, 10.0, 0, 3)
    pb, cb, fb = b.get_state_device(0)
    p2b, _, _ = b.get_state_device(1)
    b.set_stream(None)
    b.close()
    assert np.array_equal(pb.cpu().numpy(), pa) and np.array_equal(cb.cpu().numpy(), ca)
    assert np.array_equal(p2b.cpu().numpy(), p2a)
    for x, y in zip(fb, fa):
        assert np.array_equal(x.cpu().numpy(), y)


def test_device_state_rejects_host_tensors():
    """A host tensor handed to the HIP library's device entry points is refused on
    the host (it would fault the GPU)."""
    from pythonmhd_amd._lib import Solver
    s = Solver((8, 8), (1 / 8, 1 / 8), (1, 1), GAMMA, 0.8, order=1)
    with pytest.raises(ValueError):
        s.set_state_device(torch.zeros((8, 8, 8), dtype=torch.float64), torch.zeros((8, 8, 8), dtype=torch.float64),
                           [torch.zeros((8, 9), dtype=torch.float64), torch.zeros((9, 8), dtype=torch.float64)])
    s.close()


def _ke(c):
    return (0.5 * (c[1] ** 2 + c[2] ** 2 + c[3] ** 2) / c[0]).sum()


def test_fixed_dt_steps_on_device_tensors():
    """torch.autograd through the adjoint with ROCm tensors (no host copies): the
    gradient of the kinetic energy with respect to the initial density (configured
    adjoint state, pythonmhd_amd.problems.adjoint_state, with a smooth velocity)
    against central differences of the FP64 oracle."""
    from pythonmhd_amd.adjoint import AdjointRun, fixed_dt_steps
    from pythonmhd_amd.problems import adjoint_state
    from tests.parity_suite import _oracle_steps
    n = 24
    prim, cons, faces = adjoint_state(n, velocity=0.2)
    widths = (1.0 / n, 1.0 / n)
    dt, steps = 0.002, 6
    run = AdjointRun((n, n), widths, (1, 1), GAMMA, order=1, precision=64, checkpoint_every=2)
    dev = torch.device('cuda')
    rho = torch.tensor(prim[0], dtype=torch.float64, device=dev, requires_grad=True)
    rest = torch.tensor(prim[1:], dtype=torch.float64, device=dev)
    prim_t = torch.cat([rho[None], rest])
    v2 = rest[0] ** 2 + rest[1] ** 2 + rest[2] ** 2
    b2 = rest[3] ** 2 + rest[4] ** 2 + rest[5] ** 2
    cons_t = torch.stack([rho, rho * rest[0], rho * rest[1], rho * rest[2], rest[3], rest[4], rest[5],
                          rest[6] / (GAMMA - 1.0) + 0.5 * rho * v2 + 0.5 * b2])
    ft = [torch.tensor(f, dtype=torch.float64, device=dev) for f in faces]
    print('case: fixed_dt_steps on device tensors', flush=True)
    out = fixed_dt_steps(run, dt, steps, prim_t, cons_t, *ft)
    loss = _ke(out[1])
    loss.backward()
    grad = rho.grad.cpu().numpy()
    run.close()
    P = M.Params(GAMMA, 0.8, order=1)

    def oracle_loss(r):
        p = prim.copy()
        p[0] = r
        c = M.prim_to_cons_all(p, GAMMA)
        return float(_ke(_oracle_steps(2, p, c, list(faces), dt, steps, widths, (1, 1), P)[1]))

    # the step is piecewise smooth (limiters, upwind switches): central differences
    # at a handful of cells, judged by the median relative error, with every cell
    # within 1e-3 of the largest gradient entry
    rng = np.random.default_rng(4)
    eps = 1e-6
    gmax = float(np.max(np.abs(grad)))
    rel = []
    for _ in range(8):
        j, i = (int(x) for x in rng.integers(0, n, 2))
        r1 = prim[0].copy()
        r1[j, i] += eps
        r2 = prim[0].copy()
        r2[j, i] -= eps
        fd = (oracle_loss(r1) - oracle_loss(r2)) / (2 * eps)
        assert abs(grad[j, i] - fd) <= 1e-3 * gmax, (j, i, grad[j, i], fd)
        rel.append(abs(grad[j, i] - fd) / max(abs(fd), 1e-30))
    assert float(np.median(rel)) < 1e-5, rel


def test_fixed_dt_steps_refuses_a_reused_run():
    """Two fixed_dt_steps calls on one AdjointRun in one graph: the first call's
    backward raises instead of reversing the second call's trajectory."""
    from pythonmhd_amd.adjoint import AdjointRun, fixed_dt_steps
    from tests.parity_suite import _generic_state
    shape = (8, 8)
    prim, faces = _generic_state(np.random.default_rng(3), shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    dev = torch.device('cuda')
    run = AdjointRun(shape, widths_of(shape), (1, 1), GAMMA, order=1, precision=64, checkpoint_every=2)
    p = torch.tensor(prim, device=dev, requires_grad=True)
    c = torch.tensor(cons, device=dev)
    f = [torch.tensor(x, device=dev) for x in faces]
    o1 = fixed_dt_steps(run, 0.004, 2, p, c, *f)
    o2 = fixed_dt_steps(run, 0.004, 2, p * 1.0, c, *f)
    with pytest.raises(RuntimeError):
        (_ke(o1[1]) + _ke(o2[1])).backward()
    run.close()
