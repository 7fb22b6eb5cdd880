"""pmhd_cfl_dt_noalloc on the MI355X: the timestep of the allocation-free entry
point (static device scratch) against pmhd_compute_dt of an FP64 handle on the
same state, bit for bit, on two calls in a row (the second on a smaller grid), and
its rejections."""
import numpy as np
import pytest

from oracle import hydro as H
from oracle import mhd as M
from tests.parity_cases import GAMMA, smooth_state, widths_of

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _handle_dt(shape, widths, bcs, prim, cons, faces, cfl, advance=0, physics=0, gamma=GAMMA):
    from pythonmhd_amd._lib import Solver
    s = Solver(shape, widths, bcs, gamma, cfl, order=1, precision=64, physics=physics)
    s.set_state(prim, cons, faces)
    if advance:
        s.advance(0.0, 10.0, 0, advance)
    dt = s.compute_dt()
    p, c, f = s.get_state(0)
    s.close()
    return dt, c, f


def test_noalloc_dt_matches_handle_twice():
    from pythonmhd_amd._lib import cfl_dt_noalloc
    dev = torch.device('cuda')
    stream = torch.cuda.current_stream().cuda_stream
    # first call: a 3D box after two cycles (the handle's dt then comes from the
    # update's fused maxima)
    shape = (12, 10, 14)
    prim, faces = smooth_state(np.random.default_rng(5), shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    w = widths_of(shape)
    print('case: noalloc dt 3D 12x10x14 after 2 cycles', flush=True)
    want, c, f = _handle_dt(shape, w, (1, 1, 1), prim, cons, faces, 0.4, advance=2)
    out = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    cfl_dt_noalloc(shape, w, (1, 1, 1), GAMMA, 0.4, torch.from_numpy(c).to(dev),
                   [torch.from_numpy(x).to(dev) for x in f], dt_out=out, stream=stream)
    torch.cuda.synchronize()
    got = float(out.cpu().numpy()[0])
    assert got == want, (got, want)
    # second call, smaller input: a 2D grid with an outflow axis, into the same output
    shape = (6, 9)
    prim, faces = smooth_state(np.random.default_rng(6), shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    w = widths_of(shape)
    print('case: noalloc dt 2D 6x9', flush=True)
    want, c, f = _handle_dt(shape, w, (1, 0), prim, cons, faces, 0.8)
    cfl_dt_noalloc(shape, w, (1, 0), GAMMA, 0.8, torch.from_numpy(c).to(dev),
                   [torch.from_numpy(x).to(dev) for x in f], dt_out=out, stream=stream)
    torch.cuda.synchronize()
    got = float(out.cpu().numpy()[0])
    assert got == want, (got, want)


def test_noalloc_dt_hydro_and_1d():
    from pythonmhd_amd._lib import HYDRO, cfl_dt_noalloc
    dev = torch.device('cuda')
    n = 40
    x = (np.arange(n) + 0.5) / n
    prim = np.stack([1 + 0.2 * np.sin(2 * np.pi * x), 0.3 * np.cos(2 * np.pi * x), 0.1 * np.sin(4 * np.pi * x),
                     0.2 * np.cos(4 * np.pi * x), 1 + 0.1 * np.cos(2 * np.pi * x)])
    cons = H.prim_to_cons(prim, 1.4)
    print('case: noalloc dt hydro 1D', flush=True)
    want, c, _ = _handle_dt((n,), (1 / n,), (1,), prim, cons, (), 0.5, physics=HYDRO, gamma=1.4)
    got = cfl_dt_noalloc((n,), (1 / n,), (1,), 1.4, 0.5, torch.from_numpy(c).to(dev), physics=HYDRO)
    torch.cuda.synchronize()
    assert float(got.cpu().numpy()[0]) == want
    prim, _ = smooth_state(np.random.default_rng(9), (33,))
    cons = M.prim_to_cons_all(prim, GAMMA)
    print('case: noalloc dt MHD 1D', flush=True)
    want, c, _ = _handle_dt((33,), (1 / 33,), (0,), prim, cons, (), 0.8)
    got = cfl_dt_noalloc((33,), (1 / 33,), (0,), GAMMA, 0.8, torch.from_numpy(c).to(dev))
    torch.cuda.synchronize()
    assert float(got.cpu().numpy()[0]) == want


def test_noalloc_dt_rejections():
    """Rejected before any launch: FP32 parameters, a grid beyond the size limits,
    missing face arrays."""
    from pythonmhd_amd._lib import PmhdError, cfl_dt_noalloc, load, make_grid, make_params
    import ctypes
    lib = load()
    dev = torch.device('cuda')
    cons = torch.ones((8, 4, 4), dtype=torch.float64, device=dev)
    faces = [torch.ones((4, 5), dtype=torch.float64, device=dev), torch.ones((5, 4), dtype=torch.float64, device=dev)]
    with pytest.raises(PmhdError, match='face arrays'):
        cfl_dt_noalloc((4, 4), (0.25, 0.25), (1, 1), GAMMA, 0.8, cons, faces[:1])
    g = make_grid((4, 4), (0.25, 0.25), (1, 1))
    p = make_params(GAMMA, 0.8, precision=32)
    out = torch.zeros(1, dtype=torch.float64, device=dev)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    rc = lib.pmhd_cfl_dt_noalloc(ctypes.byref(g), ctypes.byref(p), ptr(cons), ptr(faces[0]), ptr(faces[1]), None,
                                 ptr(out), None)
    assert rc == -1 and b'precision must be 64' in lib.pmhd_last_error(None)
    g = make_grid((4, (1 << 24) + 1), (0.25, 0.25), (1, 1))
    p = make_params(GAMMA, 0.8, precision=64)
    rc = lib.pmhd_cfl_dt_noalloc(ctypes.byref(g), ctypes.byref(p), ptr(cons), ptr(faces[0]), ptr(faces[1]), None,
                                 ptr(out), None)
    assert rc == -1 and b'2^24' in lib.pmhd_last_error(None)
