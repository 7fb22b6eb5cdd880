"""Benchmark of the MI355X CTU+CT step on the 3D ideal-MHD turbulence box.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]

Workload: 256^3 periodic box, gamma = 5/3, PLM + Roe + CTU + CT, CFL 0.4, FP32
bulk with FP64 islands; rho = p = 1, uniform B_x = 0.5, random-phase solenoidal
velocity with per-mode power |k|^-2 over integer modes 1 <= |k| <= 8 (seeded),
rescaled to rms Mach 0.5.  A step is one full cycle (CFL reduction included).
Inputs are resident in HBM before the timed region.  The same K steps are run in
FP64 for the FP64 rate and the FP32-vs-FP64 density error.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from pythonmhd_amd.problems import adjoint_state, brio_wu, kh_boxes, orszag_tang, turbulence  # noqa: E402

METRIC = "cell-updates/s (FP32 vs FP64 GPU) at 3D 256^3 MHD; L1 err vs FP64; adjoint/fwd time"
GAMMA = 5.0 / 3.0
HBM_PEAK = 8.0e12

# Algorithmic (compulsory) words per element of each launch: distinct inputs read
# once plus outputs written once, in units of the storage word (4 B FP32, 8 B FP64).
ALG_WORDS = {
    'slopes': 8 + 7,        # rotated prim + normal B in, limited slope out (per extended cell)
    'edges': 8 + 7 + 14,    # prim + slope in, two interface states out (per cell)
    'copy': 7 + 14,
    'predict': 20 + 21,     # two states + face/cell B terms in; U_L, U_R, flux out (per face)
    'sweep': 13 + 21,       # fused slopes+edges+predictor: prim (8), faces (3), momenta (2) of one
                            # cell in; U_L, U_R, flux out (per face)
    'cell_e': 7 + 3,
    'emf': 12 + 3,          # three face-flux components per face array + cell E in (per edge point)
    'ctu_correct': 34 + 1 + 7,  # states, transverse fluxes, EMF, faces, cell B/v, half-step B in; flux out
    'emf_corr': 12 + 3,
    'ct_half': 6 + 3,       # faces of the three directions and three edge EMFs in, faces out (per edge point)
    'ct_full': 6 + 3,
    'half_e': 19 + 3,
    'correct': 15 + 7,
    'update': 31 + 16,      # U, fluxes, new faces in; cons + prim out (per cell)
    'cfl': 11,
}


def div_b(faces, h):
    fbx, fby, fbz = faces
    d = (fbx[:, 1:, :] - fbx[:, :-1, :]) + (fby[1:, :, :] - fby[:-1, :, :]) + (fbz[:, :, 1:] - fbz[:, :, :-1])
    bmax = max(np.max(np.abs(fbx)), np.max(np.abs(fby)), np.max(np.abs(fbz)))
    return float(np.max(np.abs(d)) / bmax)


def run_gpu(prim, cons, faces, n, precision, steps, warmup, profile_steps=3, sweep=None):
    from pythonmhd_amd._lib import Solver
    h = 1.0 / n
    s = Solver((n, n, n), (h, h, h), (1, 1, 1), GAMMA, 0.4, order=1, entropy_fix=False,
               precision=precision)
    if sweep is not None:
        s.set_sweep(sweep)
    s.set_state(prim, cons, faces)
    t, cyc, _ = s.advance(0.0, 1e30, 0, warmup)
    s.synchronize()
    t0 = time.perf_counter()
    t, cyc, _ = s.advance(t, 1e30, cyc, cyc + steps)
    s.synchronize()
    el = time.perf_counter() - t0
    prof = None
    if profile_steps:
        s.set_profiling(True)
        s.advance(t, 1e30, cyc, cyc + profile_steps)
        s.synchronize()
        prof = s.profile()
        s.set_profiling(False)
    return s, el, prof, cyc + profile_steps


def run_to_time(shape, widths, bcs, gamma, cfl, order, precision, prim, cons, faces, t_end, clock='auto'):
    """One handle run from t = 0 to t_end (state uploaded before the clock starts).
    clock 'host': the host reads dt each cycle, as the reference's loop does;
    'device': dt and the loop condition stay on the device."""
    from pythonmhd_amd._lib import Solver
    s = Solver(shape, widths, bcs, gamma, cfl, order=order, entropy_fix=False, precision=precision)
    s.set_clock(clock)
    s.set_state(prim, cons, faces)
    s.synchronize()
    t0 = time.perf_counter()
    t, cyc, _ = s.advance(0.0, t_end, 0, 10 ** 9)
    s.synchronize()
    el = time.perf_counter() - t0
    state = s.get_state(0)
    s.close()
    return t, cyc, el, state


def run_configs():
    """The 1D Brio-Wu and 2D Orszag-Tang configurations, each to its end time:
    wall time, cycles and cell-updates/s on the GPU (FP64, and FP32 for Orszag-Tang
    with its density L1 distance from the FP64 run)."""
    out = {}
    n = 512
    prim, cons, _ = brio_wu(n)
    # each configuration with the clock pmhd_set_clock's default picks for it (the
    # device for grids of at most 2^16 cells) and, beside it, with the other clock
    auto = lambda cells: 'device' if cells <= 1 << 16 else 'host'
    bw = {"grid": "512 cells, outflow, gamma 2, PLM, CFL 0.8, FP64", "clock": auto(n)}
    for clock in ('device', 'host'):
        run_to_time((n,), (1.0 / n,), (0,), 2.0, 0.8, 1, 64, prim, cons, (), 0.01, clock)   # warm-up
        t, cyc, el, _ = run_to_time((n,), (1.0 / n,), (0,), 2.0, 0.8, 1, 64, prim, cons, (), 0.1, clock)
        r = {"t_end": t, "cycles": cyc, "seconds": el, "cell_updates_per_s": n * cyc / el}
        if clock == bw["clock"]:
            bw.update(r)
        else:
            bw[clock + "_clock"] = r
    out["brio_wu"] = bw
    prim, cons, faces = orszag_tang(n)
    ot = {"grid": "512^2 periodic, gamma 5/3, PPM, CFL 0.8", "t_end": 0.5, "clock": auto(n * n)}
    rho = {}
    for prec in (32, 64):
        res = {}
        for clock in ('device', 'host'):
            run_to_time((n, n), (1.0 / n,) * 2, (1, 1), GAMMA, 0.8, 2, prec, prim, cons, faces, 0.005,
                        clock)  # warm-up
            t, cyc, el, st = run_to_time((n, n), (1.0 / n,) * 2, (1, 1), GAMMA, 0.8, 2, prec, prim, cons, faces,
                                         0.5, clock)
            res[clock] = {"cycles": cyc, "seconds": el, "cell_updates_per_s": n * n * cyc / el}
            if clock == ot["clock"]:
                rho[prec] = st[1][0]
        other = 'host' if ot["clock"] == 'device' else 'device'
        ot["fp%d" % prec] = dict(res[ot["clock"]], **{other + "_clock": res[other]})
    ot["l1_rho_fp32_vs_fp64"] = float(np.mean(np.abs(rho[32] - rho[64])))
    out["orszag_tang"] = ot
    return out


def run_kh(cycles=200, seq_cycles=200, nbox=256):
    """The batch of KH boxes: one batch launch per phase (tile-tuned) vs one
    handle per box advanced box after box (per-box launches and dt reads)."""
    from pythonmhd_amd._lib import Batch, Solver
    boxes = kh_boxes(nbox)
    shapes = [(n, n) for n, _, _, _ in boxes]
    widths = [(1.0 / n, 1.0 / n) for n, _, _, _ in boxes]
    cells = sum(n * n for n, _, _, _ in boxes)
    b = Batch(shapes, widths, (1, 1), GAMMA, 0.8, order=1, entropy_fix=False, precision=32)
    for k, (n, prim, cons, faces) in enumerate(boxes):
        b.set_state(k, prim, cons, faces)
    t0 = time.perf_counter()
    b.autotune(3)
    tune_s = time.perf_counter() - t0
    b.advance(0.0, 1e30, 0, 2)          # warm-up (state restarts below)
    for k, (n, prim, cons, faces) in enumerate(boxes):
        b.set_state(k, prim, cons, faces)
    b.synchronize()
    t0 = time.perf_counter()
    _, cyc, _ = b.advance(0.0, 1e30, 0, cycles)
    b.synchronize()
    el = time.perf_counter() - t0
    tiles = b.tiles()
    b.close()
    assert int(cyc.min()) == cycles
    # the same boxes one handle each, advanced box after box for the same cycle
    # count, each with its clock on the device (no dt read per cycle); handles are
    # made and loaded before the clock starts
    handles = []
    for n, prim, cons, faces in boxes:
        s = Solver((n, n), (1.0 / n, 1.0 / n), (1, 1), GAMMA, 0.8, order=1, entropy_fix=False, precision=32)
        s.set_clock('device')
        s.set_state(prim, cons, faces)
        handles.append(s)
    handles[0].advance(0.0, 1e30, 0, 2)   # warm-up
    handles[0].set_state(*boxes[0][1:])
    for s in handles:
        s.synchronize()
    t0 = time.perf_counter()
    for s in handles:
        s.advance(0.0, 1e30, 0, seq_cycles)
    for s in handles:
        s.synchronize()
    el_seq = time.perf_counter() - t0
    for s in handles:
        s.close()
    v = cells * cycles / el
    vs = cells * seq_cycles / el_seq
    return {"boxes": nbox, "sides": "U{64..256} seed 7", "cells": cells, "cycles": cycles,
            "value": v, "unit": "cell-updates/s", "seconds": el, "autotune_seconds": tune_s,
            "tiles": {k: list(t) for k, t in sorted(tiles.items())},
            "sequential": {"value": vs, "cycles": seq_cycles, "seconds": el_seq,
                           "how": "one handle per box, box after box, each with its clock on the device"},
            "speedup_vs_sequential": v / vs}


def _ke_cotangent(c, area):
    g = np.zeros_like(c)
    g[0] = -0.5 * (c[1] ** 2 + c[2] ** 2 + c[3] ** 2) / c[0] ** 2 * area
    g[1:4] = c[1:4] / c[0] * area
    return g


def _rho_gradient(gp, gc, prim):
    """d loss / d rho_0 with cons built from prim (v, p, B held fixed)."""
    v = prim[1:4]
    return gp[0] + gc[0] + v[0] * gc[1] + v[1] * gc[2] + v[2] * gc[3] + 0.5 * (v * v).sum(0) * gc[7]


def run_adjoint(n=256, steps=200, every=20, ncheck=32, check_steps=20):
    """Loss = total kinetic energy after `steps` fixed-dt steps; gradient with respect
    to the initial density through the checkpointed adjoint; timings against the
    plain FP32 and FP64 steps of the same run."""
    from pythonmhd_amd._lib import Solver
    from pythonmhd_amd.adjoint import AdjointRun
    h = 1.0 / n
    area = h * h
    prim, cons, faces = adjoint_state(n)
    s = Solver((n, n), (h, h), (1, 1), GAMMA, 0.8, order=1, precision=64)
    s.set_state(prim, cons, faces)
    dt = s.compute_dt()
    out = {"grid": "%d^2 periodic, PLM + Roe + CTU + CT" % n, "steps": steps, "checkpoint_every": every,
           "dt": dt, "loss": "sum 0.5 rho v^2 dx dy after the last step",
           "accumulation": "deterministic (each cotangent's contributions summed in a fixed order)"}
    for prec in (32, 64):   # the plain step, same fixed dt
        sp = Solver((n, n), (h, h), (1, 1), GAMMA, 0.8, order=1, precision=prec)
        sp.set_state(prim, cons, faces)
        sp.step(dt)
        sp.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sp.step(dt)
        sp.synchronize()
        out["plain_fp%d_seconds" % prec] = time.perf_counter() - t0
        sp.close()
    grads = {}
    for prec in (32, 64):
        run = AdjointRun((n, n), (h, h), (1, 1), GAMMA, order=1, precision=prec, checkpoint_every=every)
        t0 = time.perf_counter()
        _, co, _ = run.forward(prim, cons, faces, dt, steps)
        tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        gp, gc, _ = run.backward(None, _ke_cotangent(co, area), None)
        tb = time.perf_counter() - t0
        # a second backward of the same forward (its accumulation plans already built)
        t0 = time.perf_counter()
        gp2, gc2, _ = run.backward(None, _ke_cotangent(co, area), None)
        tb2 = time.perf_counter() - t0
        run.close()
        same = bool(np.array_equal(gp.view(np.uint64), gp2.view(np.uint64))
                    and np.array_equal(gc.view(np.uint64), gc2.view(np.uint64)))
        grads[prec] = _rho_gradient(gp, gc, prim)
        out["fp%d" % prec] = {"forward_seconds": tf, "backward_seconds": tb, "second_backward_seconds": tb2,
                              "repeat_bitwise_equal": same,
                              "loss": float(np.sum(0.5 * (co[1] ** 2 + co[2] ** 2 + co[3] ** 2) / co[0]) * area),
                              "max_abs_grad": float(np.max(np.abs(grads[prec])))}
        print('bench: adjoint fp%d forward %.2f s backward %.2f s' % (prec, tf, tb), file=sys.stderr, flush=True)
    # a handle's first backward also builds its accumulation plans (once per handle)
    for prec in (32, 64):
        out["backward_over_plain_forward_fp%d" % prec] = (out["fp%d" % prec]["second_backward_seconds"]
                                                          / out["plain_fp%d_seconds" % prec])
        out["first_backward_over_plain_forward_fp%d" % prec] = (out["fp%d" % prec]["backward_seconds"]
                                                                / out["plain_fp%d_seconds" % prec])
    out["grad_fp32_vs_fp64_max_abs_diff"] = float(np.max(np.abs(grads[32] - grads[64])))
    # central differences (eps = 1e-4) of the FP64 loss at seeded cells, on a smooth
    # flow with velocity (the configured state is a stationary entropy mode whose
    # gradient is zero up to rounding), over check_steps steps
    prim2, cons2, faces2 = adjoint_state(n, velocity=0.2)
    run32 = AdjointRun((n, n), (h, h), (1, 1), GAMMA, order=1, precision=32, checkpoint_every=every)
    _, co32, _ = run32.forward(prim2, cons2, faces2, dt, check_steps)
    gp, gc, _ = run32.backward(None, _ke_cotangent(co32, area), None)
    run32.close()
    g32 = _rho_gradient(gp, gc, prim2)
    run = AdjointRun((n, n), (h, h), (1, 1), GAMMA, order=1, precision=64, checkpoint_every=every)
    _, co, _ = run.forward(prim2, cons2, faces2, dt, check_steps)
    gp, gc, _ = run.backward(None, _ke_cotangent(co, area), None)
    g = _rho_gradient(gp, gc, prim2)
    rng = np.random.default_rng(5)
    cells = [tuple(int(x) for x in rng.integers(0, n, 2)) for _ in range(ncheck)]
    errs = []
    for (j, i) in cells:
        vals = []
        for sgn in (1, -1):
            p3, c3 = prim2.copy(), cons2.copy()
            p3[0, j, i] += sgn * 1e-4
            c3[:, j, i] = 0
            w = p3[:, j, i]
            c3[0, j, i] = w[0]
            c3[1:4, j, i] = w[1:4] * w[0]
            c3[4:7, j, i] = w[4:7]
            c3[7, j, i] = w[7] / (GAMMA - 1.0) + 0.5 * w[0] * (w[1:4] ** 2).sum() + 0.5 * (w[4:7] ** 2).sum()
            _, cq, _ = run.forward(p3, c3, faces2, dt, check_steps)
            vals.append(float(np.sum(0.5 * (cq[1] ** 2 + cq[2] ** 2 + cq[3] ** 2) / cq[0]) * area))
        fd = (vals[0] - vals[1]) / 2e-4
        errs.append((g[j, i], fd, g32[j, i]))
    run.close()
    gmax = float(np.max(np.abs(g)))
    out["fd_check"] = {"state": "same rho, plus a smooth velocity field of rms 0.2", "precision": 64,
                       "steps": check_steps, "cells": ncheck, "eps": 1e-4, "max_abs_grad": gmax,
                       "max_err_over_max_grad": max(abs(a - b) for a, b, _ in errs) / gmax,
                       "median_rel_err": float(np.median([abs(a - b) / max(abs(b), 1e-300) for a, b, _ in errs])),
                       "fp32_forward": {
                           "max_abs_diff_vs_fp64_over_max_grad": float(np.max(np.abs(g32 - g))) / gmax,
                           "max_err_vs_fd_over_max_grad": max(abs(c - b) for _, b, c in errs) / gmax,
                           "median_rel_err_vs_fd": float(np.median([abs(c - b) / max(abs(b), 1e-300)
                                                                   for _, b, c in errs]))}}
    return out


def cpu_baseline(n=64, cycles=8):
    """The FP64 NumPy oracle (single thread) on a bounded sample of the workload."""
    from oracle import mhd as M
    prim, cons, faces = turbulence(n)
    P = M.Params(GAMMA, 0.4, order=1, entropy_fix=False)
    h = 1.0 / n
    dt = M.calc_dt(cons, faces, GAMMA, 0.4, (h, h, h))
    M.step_3d(prim, cons, *faces, dt, h, h, h, 1, 1, 1, P)      # warm-up
    t0 = time.perf_counter()
    for _ in range(cycles):
        dt = M.calc_dt(cons, faces, GAMMA, 0.4, (h, h, h))
        prim, cons, f0, f1, f2 = M.step_3d(prim, cons, *faces, dt, h, h, h, 1, 1, 1, P)
        faces = (f0, f1, f2)
    el = time.perf_counter() - t0
    return {"value": n ** 3 * cycles / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "FP64 NumPy oracle, same workload at %d^3, %d cycles after 1 warm-up cycle, "
                      "1 thread (host: %s, os.cpu_count()=%d)" % (n, cycles, _cpu_model(), os.cpu_count())}


def cpu_openmp_baseline(n=128, cycles=3):
    """The same kernel bodies compiled for the host with g++ -O2 and OpenMP (the
    build the CPU tests run; FP64, bit for bit the oracle's arithmetic) on all the
    host threads OpenMP is given: a multi-core CPU figure beside the single-thread
    NumPy oracle."""
    from pythonmhd_amd._lib import Solver, load
    lib_path = os.path.join(ROOT, 'tests', 'hostbuild', 'libpmhd_host.so')
    if not os.path.exists(lib_path):
        return None
    lib = load(lib_path)
    prim, cons, faces = turbulence(n)
    h = 1.0 / n
    s = Solver((n, n, n), (h, h, h), (1, 1, 1), GAMMA, 0.4, order=1, entropy_fix=False, precision=64, lib=lib)
    s.set_clock('host')
    s.set_state(prim, cons, faces)
    s.advance(0.0, 1e30, 0, 1)   # warm-up
    t0 = time.perf_counter()
    s.advance(0.0, 1e30, 1, 1 + cycles)
    el = time.perf_counter() - t0
    s.close()
    threads = int(os.environ.get('OMP_NUM_THREADS', '0') or 0) or os.cpu_count()
    return {"value": n ** 3 * cycles / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": "the step's own kernel bodies built for the host (g++ -O2, OpenMP, FP64), same workload at "
                      "%d^3, %d cycles after 1 warm-up cycle, %d threads (host: %s)" % (n, cycles, threads,
                                                                                      _cpu_model())}


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


TRAFFIC = os.path.join(ROOT, 'profiles', 'hbm_traffic_fp32_turbulence256.json')


def measured_traffic(kernel, n, word_bytes):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (calibrated against a plain device copy), for the workload
    they were measured on (FP32, 256^3); None otherwise."""
    if n != 256 or word_bytes != 4 or not os.path.exists(TRAFFIC):
        return None
    with open(TRAFFIC) as f:
        t = json.load(f)
    row = t.get('launches', {}).get(kernel)
    return None if row is None else row['fetch_bytes'] + row['write_bytes']


VALU = os.path.join(ROOT, 'profiles', 'valu_fp32_turbulence256.json')
SIMDS, CLOCK_HZ, CYCLES_PER_VALU = 1024, 2.4e9, 4     # MI355X: 256 CUs x 4 SIMDs; wave64 on a SIMD16


def valu_roof(ms_per_step, n=256, sweep=None):
    """The step against the VALU issue rate: the counted VALU instructions of one
    cycle (rocprofv3 SQ_INSTS_VALU per launch, committed for the FP32 256^3 default
    launch plan) at one wave64 instruction per SIMD per 4 cycles, over the measured
    cycle time.  None for other workloads."""
    if n != 256 or sweep is not None or not os.path.exists(VALU):
        return None
    with open(VALU) as f:
        v = json.load(f)
    inst = sum(v['per_cycle'].values())
    issue_ms = inst * CYCLES_PER_VALU / SIMDS / CLOCK_HZ * 1e3
    cells = n ** 3
    return {"valu_instructions_per_cell_update": round(inst * 64 / cells, 1), "issue_ms_per_step": round(issue_ms, 4),
            "frac": round(issue_ms / ms_per_step, 4), "ceiling_cell_updates_per_s": cells / issue_ms * 1e3,
            "source": os.path.relpath(VALU, ROOT)}


def roofline(prof, word_bytes, n=256):
    if not prof:
        return None
    best = max(prof.items(), key=lambda kv: kv[1][0])
    name, (ms, nl, ne) = best
    base = name.rsplit('_', 1)[0] if name.endswith(('_x', '_y', '_z')) else name
    words = ALG_WORDS.get(base, ALG_WORDS.get(name, 0))
    per_launch_s = ms / nl / 1e3
    bytes_per_launch = words * word_bytes * ne / nl
    ach = bytes_per_launch / per_launch_s / 1e9
    return {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(ach / (HBM_PEAK / 1e9), 4), "traffic": measured_traffic(name, n, word_bytes),
            "kernel": name,
            "avg_launch_ms": round(ms / nl, 4), "alg_bytes_per_launch": int(bytes_per_launch)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--n', type=int, default=256)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-fp64', action='store_true')
    ap.add_argument('--err-cycles', type=int, default=100,
                    help='cycles after which the FP32 and FP64 states are compared')
    ap.add_argument('--sweep', default='', help='fused sweep run lengths x,y,z (0: separate launches)')
    ap.add_argument('--no-kh', action='store_true')
    ap.add_argument('--kh-cycles', type=int, default=200)
    ap.add_argument('--no-adjoint', action='store_true')
    ap.add_argument('--no-configs', action='store_true', help='skip the Brio-Wu and Orszag-Tang runs')
    ap.add_argument('--adj-steps', type=int, default=200)
    ap.add_argument('--adj-n', type=int, default=256)
    ap.add_argument('--copy-probe', action='store_true',
                    help='only run a plain 1 GiB device copy three times (HBM counter calibration)')
    a = ap.parse_args()
    if a.copy_probe:
        import torch
        x = torch.ones(2 ** 28, device='cuda', dtype=torch.float32)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print('copy probe: 3 copies of %d bytes' % (4 * 2 ** 28), file=sys.stderr)
        return
    if a.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU path (--gpus %d refused)' % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    n, K, W = a.n, a.steps, a.warmup
    prim, cons, faces = turbulence(n)
    print('bench: %d^3 turbulence IC ready' % n, file=sys.stderr, flush=True)
    sweep = [int(x) for x in a.sweep.split(',')] if a.sweep else None
    s32, el32, prof32, cyc32 = run_gpu(prim, cons, faces, n, 32, K, W, sweep=sweep)
    cells = n ** 3
    v32 = cells * K / el32
    print('bench: fp32 %.3f ms/step, %.3e cell-updates/s' % (el32 / K * 1e3, v32), file=sys.stderr, flush=True)
    out = {"metric": METRIC, "value": v32, "unit": "cell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
           "ms_per_step": el32 / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "3D ideal-MHD turbulence box %d^3 periodic, gamma=5/3, PLM+Roe+CTU+CT, "
                                  "CFL 0.4, FP32 bulk with FP64 islands" % n,
                      "cells": cells, "reconstruct_order": 1, "parallelism": "single GPU"},
           "roofline": roofline(prof32, 4, n),
           "hbm_roof_fraction_step": v32 * 64 / HBM_PEAK,
           "valu_roof_step": valu_roof(el32 / K * 1e3, n, sweep),
           "kernel_ms_per_step": {k: round(v[0] / 3, 4) for k, v in sorted(prof32.items())}}
    if sweep is not None:
        out["config"]["sweep_run_lengths"] = sweep
    if not a.no_fp64:
        # both precisions continue (untimed) to the configured cycle count before
        # their states are compared
        ncmp = max(a.err_cycles, cyc32)
        t32, _, _ = s32.advance(0.0, 1e30, cyc32, ncmp)
        p32, c32, f32 = s32.get_state(0)
        s32.close()
        s64, el64, prof64, cyc64 = run_gpu(prim, cons, faces, n, 64, K, W, profile_steps=3)
        v64 = cells * K / el64
        print('bench: fp64 %.3f ms/step, %.3e cell-updates/s' % (el64 / K * 1e3, v64), file=sys.stderr,
              flush=True)
        s64.advance(0.0, 1e30, cyc64, ncmp)
        p64, c64, f64 = s64.get_state(0)
        s64.close()
        out["fp64"] = {"value": v64, "ms_per_step": el64 / K * 1e3, "roofline": roofline(prof64, 8, n),
                       "kernel_ms_per_step": {k: round(v[0] / 3, 4) for k, v in sorted(prof64.items())}}
        out["fp32_over_fp64_throughput"] = v32 / v64
        out["l1_rho_fp32_vs_fp64"] = float(np.mean(np.abs(c32[0] - c64[0])))
        out["divb_fp32"] = div_b(f32, 1.0 / n)
        out["divb_fp64"] = div_b(f64, 1.0 / n)
        out["fp32_vs_fp64_cycles"] = ncmp
        print('bench: after %d cycles L1(rho) fp32 vs fp64 %.3e, div B fp32 %.3e, fp64 %.3e' % (
            ncmp, out["l1_rho_fp32_vs_fp64"], out["divb_fp32"], out["divb_fp64"]), file=sys.stderr, flush=True)
    else:
        s32.close()
    if not a.no_kh:
        out["kh_batch"] = run_kh(cycles=a.kh_cycles)
        print('bench: KH batch %.3e cell-updates/s (sequential %.3e)' % (
            out["kh_batch"]["value"], out["kh_batch"]["sequential"]["value"]), file=sys.stderr, flush=True)
    if not a.no_configs:
        out.update(run_configs())
        print('bench: Brio-Wu fp64 %.3f s (%d cycles); Orszag-Tang to t=0.5: fp32 %.3e, fp64 %.3e cell-updates/s' % (
            out["brio_wu"]["seconds"], out["brio_wu"]["cycles"], out["orszag_tang"]["fp32"]["cell_updates_per_s"],
            out["orszag_tang"]["fp64"]["cell_updates_per_s"]), file=sys.stderr, flush=True)
    if not a.no_adjoint:
        out["adjoint"] = run_adjoint(n=a.adj_n, steps=a.adj_steps)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        omp = cpu_openmp_baseline()
        if omp is not None:
            out["cpu_baseline_openmp"] = omp
            print('bench: cpu OpenMP baseline %.3e cell-updates/s on %d threads' % (omp["value"], omp["cores"]),
                  file=sys.stderr, flush=True)
        print('bench: cpu baseline %.3e cell-updates/s' % out["cpu_baseline"]["value"], file=sys.stderr,
              flush=True)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
