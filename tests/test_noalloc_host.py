"""pmhd_cfl_dt_noalloc of the host test build (arrays in host memory): the same
timestep as pmhd_compute_dt of an FP64 handle, and the same rejections."""
import os

import numpy as np
import pytest

from oracle import mhd as M
from tests.parity_cases import GAMMA, smooth_state, widths_of

torch = pytest.importorskip('torch')
HOST_LIB = os.path.join(os.path.dirname(__file__), 'hostbuild', 'libpmhd_host.so')


@pytest.mark.parametrize('shape,bcs,cfl', [((12, 10, 14), (1, 1, 1), 0.4), ((6, 9), (1, 0), 0.8), ((33,), (0,), 0.8)])
def test_host_noalloc_dt_matches_handle(shape, bcs, cfl):
    from pythonmhd_amd._lib import Solver, cfl_dt_noalloc, load
    lib = load(HOST_LIB)
    prim, faces = smooth_state(np.random.default_rng(len(shape)), shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    w = widths_of(shape)
    s = Solver(shape, w, bcs, GAMMA, cfl, order=1, precision=64, lib=lib)
    s.set_state(prim, cons, faces)
    s.advance(0.0, 10.0, 0, 2)
    want = s.compute_dt()
    _, c, f = s.get_state(0)
    s.close()
    got = cfl_dt_noalloc(shape, w, bcs, GAMMA, cfl, torch.from_numpy(c), [torch.from_numpy(x) for x in f], lib=lib)
    assert float(got[0]) == want


def test_host_noalloc_rejects():
    from pythonmhd_amd._lib import PmhdError, cfl_dt_noalloc, load
    lib = load(HOST_LIB)
    cons = torch.ones((8, 4, 4), dtype=torch.float64)
    with pytest.raises(PmhdError, match='face arrays'):
        cfl_dt_noalloc((4, 4), (0.25, 0.25), (1, 1), GAMMA, 0.8, cons, [torch.ones((4, 5), dtype=torch.float64)],
                       lib=lib)
    with pytest.raises(PmhdError, match='2.24'):
        cfl_dt_noalloc((4, (1 << 24) + 1), (0.25, 0.25), (1, 1), GAMMA, 0.8, cons, [], lib=lib)
