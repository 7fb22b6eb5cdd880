"""FP64 restatement of the reference's simulation driver loop (the oracle of
pythonmhd_amd.simulation).  TEST INFRASTRUCTURE ONLY -- imported by tests/ alone.

``run_simulation`` follows Source/Simulation.py:326-413: the evolve function of
the grid advances from the current time to the next stop (the next save/plot time
or tLim, the first stop unrounded, later ones rounded to the output precision,
:338-346 and :395-402); at a stop the state is recorded; unless the run is over
the grid rolls back to the evolve call's second-last state and time, while the
cycle count keeps the cycle that was undone (:404-413).  The evolve functions are
the oracle's (oracle/mhd.py, oracle/hydro.py).

Parity: the reference's Simulation module cannot be imported here (its file
output needs pyevtk and cv2, which are absent), so this loop is pinned by
invariants instead of golden vectors (tests/test_frontend_oracle.py): with no
stops it is one evolve call; each recorded state equals an evolve call from the
previous rollback state to the stop; the cycle count exceeds the cycles of the
final trajectory by one per rollback.  "parity unpinned" beyond those.
"""
import numpy as np

from . import hydro as H
from . import mhd as M


def time_precision(step, t_lim, limit=10):
    """Source/FileSaver.py:94-118: decimals that show step and t_lim exactly."""
    def prec(x):
        p = 1
        while abs(10.0 ** p * x - int(10.0 ** p * x)) > 1e-14 and p < limit:
            p += 1
        return p
    return max(prec(step), prec(t_lim))


def _evolve(state, widths, bcs, t0, t_end, cyc0, max_cycles, P, mhd):
    prim, cons, faces = state
    nd = prim.ndim - 1
    if not mhd:
        r = H.evolve(prim, cons, widths, bcs, t0, t_end, cyc0, max_cycles, P)
        return (r[0], r[1], ()), r[2], r[3], (r[4], r[5], ()), r[6]
    if nd == 1:
        r = M.evolve_1d(prim, cons, widths[0], bcs[0], t0, t_end, cyc0, max_cycles, P)
        return (r[0], r[1], ()), r[2], r[3], (r[4], r[5], ()), r[6]
    if nd == 2:
        r = M.evolve_2d(prim, cons, faces[0], faces[1], widths[0], widths[1], bcs[0], bcs[1], t0, t_end, cyc0,
                        max_cycles, P)
        return (r[0], r[1], (r[2], r[3])), r[4], r[5], (r[6], r[7], (r[8], r[9])), r[10]
    r = M.evolve_3d(prim, cons, faces[0], faces[1], faces[2], widths[0], widths[1], widths[2], bcs[0], bcs[1],
                    bcs[2], t0, t_end, cyc0, max_cycles, P)
    return (r[0], r[1], (r[2], r[3], r[4])), r[5], r[6], (r[7], r[8], (r[9], r[10], r[11])), r[12]


def run_simulation(state, widths, bcs, t_lim, max_cycles, P, mhd=True, save_dt=None, plot_dt=None):
    """Returns (final state, final time, cycles, [(stop time, state at the stop)])."""
    sim_time, cycles = 0.0, 0
    records = []
    next_plot = next_save = t_lim
    last_plot = last_save = 0.0
    plot_prec = time_precision(plot_dt, t_lim) if plot_dt else -1
    save_prec = time_precision(save_dt, t_lim) if save_dt else -1
    if plot_dt:
        next_plot = np.minimum(next_plot, plot_dt)
    if save_dt:
        next_save = np.minimum(next_save, save_dt)
    next_stop = np.minimum(next_plot, next_save)
    new = state
    while sim_time < t_lim and cycles < max_cycles:
        new, t_new, cyc_new, second, t_second = _evolve(state, widths, bcs, sim_time, next_stop, cycles,
                                                        max_cycles, P, mhd)
        state = new
        sim_time, cycles = t_new, cyc_new
        if save_dt and next_stop == next_save:
            last_save = sim_time
            records.append((sim_time, new))
        if plot_dt and next_stop == next_plot:
            last_plot = sim_time
        next_plot = next_save = t_lim
        if plot_dt:
            next_plot = np.round(np.minimum(next_plot, last_plot + plot_dt), plot_prec)
        if save_dt:
            next_save = np.round(np.minimum(next_save, last_save + save_dt), save_prec)
        next_stop = np.minimum(next_plot, next_save)
        if sim_time < t_lim and cycles < max_cycles:
            state = second
            sim_time = t_second
    return new, sim_time, cycles, records
