"""torch.autograd wrapper of the adjoint (CPU, through the host build of the step)."""
import os

import numpy as np
import pytest

torch = pytest.importorskip('torch')

from oracle import mhd as M  # noqa: E402
from tests.parity_cases import GAMMA, widths_of  # noqa: E402
from tests.parity_suite import _generic_state, _kinetic_cotangent  # noqa: E402

HOST_LIB = os.path.join(os.path.dirname(__file__), 'hostbuild', 'libpmhd_host.so')


def test_autograd_function_matches_adjoint():
    from pythonmhd_amd._lib import load
    from pythonmhd_amd.adjoint import AdjointRun, fixed_dt_steps
    lib = load(HOST_LIB)
    shape = (8, 10)
    rng = np.random.default_rng(12)
    prim, faces = _generic_state(rng, shape)
    widths = widths_of(shape)
    run = AdjointRun(shape, widths, (1, 1), GAMMA, order=1, precision=64, checkpoint_every=2, lib=lib)
    # initial density as the leaf: cons follows from prim inside the torch graph
    rho = torch.tensor(prim[0], dtype=torch.float64, requires_grad=True)
    p_rest = torch.tensor(prim[1:], dtype=torch.float64)
    prim_t = torch.cat([rho[None], p_rest])
    v2 = (p_rest[0] ** 2 + p_rest[1] ** 2 + p_rest[2] ** 2)
    b2 = (p_rest[3] ** 2 + p_rest[4] ** 2 + p_rest[5] ** 2)
    cons_t = torch.stack([rho, rho * p_rest[0], rho * p_rest[1], rho * p_rest[2], p_rest[3], p_rest[4], p_rest[5],
                          p_rest[6] / (GAMMA - 1.0) + 0.5 * rho * v2 + 0.5 * b2])
    fx = torch.tensor(faces[0], dtype=torch.float64)
    fy = torch.tensor(faces[1], dtype=torch.float64)
    out = fixed_dt_steps(run, 0.004, 4, prim_t, cons_t, fx, fy)
    c = out[1]
    loss = (0.5 * (c[1] ** 2 + c[2] ** 2 + c[3] ** 2) / c[0]).sum()
    loss.backward()
    # the same gradient from the C-ABI calls and the chain rule of cons(prim)
    cons = M.prim_to_cons_all(prim, GAMMA)
    _, co, _ = run.forward(prim, cons, faces, 0.004, 4)
    gp, gc, _ = run.backward(None, _kinetic_cotangent(co), None)
    v = prim[1:4]
    want = gp[0] + gc[0] + v[0] * gc[1] + v[1] * gc[2] + v[2] * gc[3] + 0.5 * (v * v).sum(0) * gc[7]
    assert np.allclose(rho.grad.numpy(), want, rtol=1e-12, atol=1e-15)
    run.close()


def test_reused_run_raises_in_backward():
    """Two fixed_dt_steps calls on one AdjointRun in one graph: the backward of the
    first raises (its checkpoints were overwritten by the second forward)."""
    from pythonmhd_amd._lib import load
    from pythonmhd_amd.adjoint import AdjointRun, fixed_dt_steps
    lib = load(HOST_LIB)
    shape = (6, 8)
    prim, faces = _generic_state(np.random.default_rng(3), shape)
    cons = M.prim_to_cons_all(prim, GAMMA)
    run = AdjointRun(shape, widths_of(shape), (1, 1), GAMMA, order=1, precision=64, checkpoint_every=2, lib=lib)
    p = torch.tensor(prim, requires_grad=True)
    c = torch.tensor(cons)
    f = [torch.tensor(x) for x in faces]
    o1 = fixed_dt_steps(run, 0.004, 2, p, c, *f)
    o2 = fixed_dt_steps(run, 0.004, 2, p * 1.0, c, *f)
    ke = lambda u: (0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / u[0]).sum()  # noqa: E731
    with pytest.raises(RuntimeError):
        (ke(o1[1]) + ke(o2[1])).backward()
    run.close()


def test_adjoint_rejects_1d_mhd():
    from pythonmhd_amd._lib import PmhdError, load
    from pythonmhd_amd.adjoint import AdjointRun
    with pytest.raises(PmhdError):
        AdjointRun((16,), (1 / 16,), (1,), GAMMA, order=1, precision=64, lib=load(HOST_LIB))
