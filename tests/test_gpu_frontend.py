"""The Simulation front-end on the MI355X: the Orszag-Tang init file's
dictionaries (keys and values of
InitializationFiles/SampleSimulations/2D_MHD/OrszagTang_InitializationFile.py:36-75)
run through pythonmhd_amd.simulation.Simulation on libpmhd.so and match the
oracle's restatement of the reference's driver loop bit for bit (final grid, .mat
snapshots, cycle count), and a checkpoint restart reproduces the uninterrupted
run."""
import numpy as np
import pytest

from pythonmhd_amd import constants as C
from tests.test_frontend_oracle import _check_against_oracle, _ot, ot_dicts

pytestmark = pytest.mark.gpu


@pytest.fixture
def gpu_evolve():
    from pythonmhd_amd import evolve
    evolve.release_device_state()
    yield evolve
    evolve.release_device_state()


def test_orszag_tang_init_file_dicts_on_gpu(gpu_evolve, tmp_path):
    from pythonmhd_amd.simulation import Simulation
    n = 32
    prim, cons, faces = _ot(n)
    simPar, gridPar = ot_dicts(n, tmp_path, t_lim=0.05, save_dt=0.02)
    print('case: Simulation Orszag-Tang 32^2 to t=0.05, saves every 0.02', flush=True)
    sim = Simulation(simPar, gridPar, prim, {}, cons, faces[0], faces[1])
    grid = sim.run()
    _check_against_oracle(sim, grid, prim, cons, faces, n, simPar)


def test_checkpoint_restart_on_gpu(gpu_evolve, tmp_path):
    from pythonmhd_amd.simulation import Simulation
    n = 32
    prim, cons, faces = _ot(n)
    simPar, gridPar = ot_dicts(n, tmp_path, t_lim=0.05, save_dt=0.02, saveCheckpoints=True,
                               **{C.PRECISION: 32})
    print('case: Simulation checkpoint restart (FP32)', flush=True)
    full = Simulation(simPar, gridPar, prim, {}, cons, faces[0], faces[1])
    g_full = full.run()
    cks = [f for f in full.savedFiles if f.endswith('.npz')]
    res = Simulation.resume(cks[-1], simPar, gridPar)
    g_res = res.run()
    assert res.numCycles == full.numCycles and res.simTime == full.simTime
    assert np.array_equal(g_res.consVars, g_full.consVars)
    assert np.array_equal(g_res.faceBx, g_full.faceBx) and np.array_equal(g_res.faceBy, g_full.faceBy)
