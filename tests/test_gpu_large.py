"""A grid large enough that pmhd_create picks the FP32 unit with non-temporal
result stores (interface fluxes along x above 256 MiB): its step against the FP64
step on the same state, within the FP32 tolerance of the configurations."""
import numpy as np
import pytest

from oracle import mhd as M
from tests.parity_cases import GAMMA, div_b, smooth_state, widths_of

pytestmark = pytest.mark.gpu


def test_streaming_fp32_unit_against_fp64():
    from pythonmhd_amd._lib import Solver
    shape = (160, 256, 256)   # 7 x 160 x 257 x 256 x 4 B = 295 MB of x fluxes
    assert 7 * shape[0] * (shape[1] + 1) * shape[2] * 4 > 256 << 20
    rng = np.random.default_rng(11)
    prim, faces = smooth_state(rng, shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    widths = widths_of(shape)
    print('case: FP32 streaming unit', shape, flush=True)
    out = {}
    for prec in (32, 64):
        s = Solver(shape, widths, (1, 1, 1), GAMMA, 0.4, order=1, precision=prec)
        s.set_state(prim, cons, faces)
        t, cyc, _ = s.advance(0.0, 1e30, 0, 3)
        p, c, f = s.get_state(0)
        s.close()
        out[prec] = (t, cyc, c, f)
    t32, c32, u32, f32 = out[32]
    t64, c64, u64, f64 = out[64]
    assert c32 == c64 == 3
    assert abs(t32 - t64) <= 1e-6 * t64
    l1 = float(np.mean(np.abs(u32[0] - u64[0])))
    assert l1 < 1e-5, l1
    # div B: FP64 at double rounding; FP32 (faces as float high/low pairs, CT in
    # double from float EMFs) about 1e-11 on this state (4e-12 from the host build
    # of the same FP32 code), far below the 1e-6 bound of the FP32 step
    d0 = div_b(faces, widths)
    assert div_b(f64, widths) < 1e-13 + 2 * d0
    assert div_b(f32, widths) < 1e-10, div_b(f32, widths)
