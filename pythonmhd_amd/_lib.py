"""ctypes binding of the C ABI in include/pmhd.h.

``load()`` opens the in-tree HIP library ``pythonmhd_amd/libpmhd.so`` (built by
``__graft_entry__.build()``).  There is no fallback: if the library is missing or
cannot be loaded, ``load()`` raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(_HERE, 'libpmhd.so')

EXPORTS = ('pmhd_create', 'pmhd_destroy', 'pmhd_set_state', 'pmhd_get_state', 'pmhd_set_state_device',
           'pmhd_get_state_device', 'pmhd_set_stream', 'pmhd_cfl_dt_noalloc', 'pmhd_compute_dt',
           'pmhd_step', 'pmhd_set_sweep', 'pmhd_set_clock', 'pmhd_advance', 'pmhd_synchronize', 'pmhd_diagnostics',
           'pmhd_set_profiling',
           'pmhd_get_profile', 'pmhd_get_stats', 'pmhd_riemann_flux', 'pmhd_last_error',
           'pmhd_batch_create', 'pmhd_batch_destroy', 'pmhd_batch_set_state', 'pmhd_batch_get_state',
           'pmhd_batch_advance', 'pmhd_batch_autotune', 'pmhd_batch_get_tiles', 'pmhd_batch_synchronize',
           'pmhd_batch_set_profiling', 'pmhd_batch_get_profile', 'pmhd_batch_get_stats', 'pmhd_batch_last_error',
           'pmhd_adjoint_create', 'pmhd_adjoint_destroy', 'pmhd_adjoint_forward', 'pmhd_adjoint_backward',
           'pmhd_adjoint_forward_device', 'pmhd_adjoint_backward_device', 'pmhd_adjoint_set_stream',
           'pmhd_adjoint_set_deterministic',
           'pmhd_adjoint_set_profiling', 'pmhd_adjoint_get_profile', 'pmhd_adjoint_last_error')


class PmhdGrid(ctypes.Structure):
    _fields_ = [('ndim', ctypes.c_int32), ('pad0', ctypes.c_int32),
                ('nx', ctypes.c_int64), ('ny', ctypes.c_int64), ('nz', ctypes.c_int64),
                ('dx', ctypes.c_double), ('dy', ctypes.c_double), ('dz', ctypes.c_double),
                ('bcx', ctypes.c_int32), ('bcy', ctypes.c_int32), ('bcz', ctypes.c_int32),
                ('pad1', ctypes.c_int32)]


class PmhdParams(ctypes.Structure):
    _fields_ = [('gamma', ctypes.c_double), ('cfl', ctypes.c_double),
                ('min_dens', ctypes.c_double), ('min_pres', ctypes.c_double),
                ('min_energy', ctypes.c_double), ('epsilon', ctypes.c_double),
                ('reconstruct_order', ctypes.c_int32), ('entropy_fix', ctypes.c_int32),
                ('precision', ctypes.c_int32), ('physics', ctypes.c_int32),
                ('riemann', ctypes.c_int32), ('fix_quirks', ctypes.c_int32)]


class PmhdStats(ctypes.Structure):
    _fields_ = [('cycles', ctypes.c_int64), ('floored', (ctypes.c_int64 * 2) * 10),
                ('floored_max', (ctypes.c_int64 * 2) * 10), ('floored_total', ctypes.c_int64),
                ('nonfinite', ctypes.c_int64)]


# return codes of include/pmhd.h
OK, EINVAL, EHIP, ENOMEM, ENONFINITE = 0, -1, -2, -3, -4

# floor call sites (pmhd_stats rows): predictor x/y/z left/right, 2D post-CTU x/y
FLOOR_SITES = ('x left', 'x right', 'y top', 'y bottom', 'z back', 'z forward',
               'x left after CTU', 'x right after CTU', 'y top after CTU', 'y bottom after CTU')


class PmhdError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def _stats_dict(st):
    return {'cycles': st.cycles,
            'floored': np.array([[st.floored[s][v] for v in range(2)] for s in range(10)], dtype=np.int64),
            'floored_max': np.array([[st.floored_max[s][v] for v in range(2)] for s in range(10)],
                                    dtype=np.int64),
            'floored_total': st.floored_total, 'nonfinite': st.nonfinite}


_cache = {}


def _declare(lib):
    H = ctypes.c_void_p
    P = ctypes.c_void_p
    lib.pmhd_create.argtypes = [ctypes.POINTER(PmhdGrid), ctypes.POINTER(PmhdParams), ctypes.POINTER(H)]
    lib.pmhd_destroy.argtypes = [H]
    lib.pmhd_set_state.argtypes = [H, P, P, P, P, P]
    lib.pmhd_get_state.argtypes = [H, ctypes.c_int32, P, P, P, P, P]
    lib.pmhd_set_state_device.argtypes = [H, P, P, P, P, P]
    lib.pmhd_get_state_device.argtypes = [H, ctypes.c_int32, P, P, P, P, P]
    lib.pmhd_set_stream.argtypes = [H, P]
    lib.pmhd_cfl_dt_noalloc.argtypes = [ctypes.POINTER(PmhdGrid), ctypes.POINTER(PmhdParams), P, P, P, P, P, P]
    lib.pmhd_compute_dt.argtypes = [H, ctypes.POINTER(ctypes.c_double)]
    lib.pmhd_step.argtypes = [H, ctypes.c_double]
    lib.pmhd_set_sweep.argtypes = [H, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    lib.pmhd_set_clock.argtypes = [H, ctypes.c_int32]
    lib.pmhd_advance.argtypes = [H, ctypes.c_double, ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_double)]
    lib.pmhd_synchronize.argtypes = [H]
    lib.pmhd_diagnostics.argtypes = [H, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
    lib.pmhd_set_profiling.argtypes = [H, ctypes.c_int32]
    lib.pmhd_get_profile.argtypes = [H, ctypes.c_int32, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int32)]
    lib.pmhd_get_stats.argtypes = [H, ctypes.POINTER(PmhdStats)]
    lib.pmhd_riemann_flux.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, P, P, P, P,
                                      P, ctypes.c_double, ctypes.c_int32, ctypes.c_double, P]
    lib.pmhd_batch_get_stats.argtypes = [H, ctypes.c_int32, ctypes.POINTER(PmhdStats)]
    lib.pmhd_last_error.argtypes = [H]
    lib.pmhd_last_error.restype = ctypes.c_char_p
    I32, I64, D = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    lib.pmhd_batch_create.argtypes = [I32, ctypes.POINTER(PmhdGrid), ctypes.POINTER(PmhdParams), ctypes.POINTER(H)]
    lib.pmhd_batch_destroy.argtypes = [H]
    lib.pmhd_batch_set_state.argtypes = [H, I32, P, P, P, P, P]
    lib.pmhd_batch_get_state.argtypes = [H, I32, I32, P, P, P, P, P]
    lib.pmhd_batch_advance.argtypes = [H, D, D, I64, I64, P, P, P]
    lib.pmhd_batch_autotune.argtypes = [H, I32]
    lib.pmhd_batch_get_tiles.argtypes = [H, I32, ctypes.c_char_p, P, P, P, P, ctypes.POINTER(I32)]
    lib.pmhd_batch_synchronize.argtypes = [H]
    lib.pmhd_batch_set_profiling.argtypes = [H, I32]
    lib.pmhd_batch_get_profile.argtypes = lib.pmhd_get_profile.argtypes
    lib.pmhd_batch_last_error.argtypes = [H]
    lib.pmhd_batch_last_error.restype = ctypes.c_char_p
    lib.pmhd_adjoint_create.argtypes = [ctypes.POINTER(PmhdGrid), ctypes.POINTER(PmhdParams), I32, ctypes.POINTER(H)]
    lib.pmhd_adjoint_destroy.argtypes = [H]
    lib.pmhd_adjoint_forward.argtypes = [H, P, P, P, P, P, D, I64, P, P, P, P, P]
    lib.pmhd_adjoint_backward.argtypes = [H] + [P] * 10
    lib.pmhd_adjoint_forward_device.argtypes = [H, P, P, P, P, P, D, I64, P, P, P, P, P]
    lib.pmhd_adjoint_backward_device.argtypes = [H] + [P] * 10
    lib.pmhd_adjoint_set_stream.argtypes = [H, P]
    lib.pmhd_adjoint_set_deterministic.argtypes = [H, I32]
    lib.pmhd_adjoint_set_profiling.argtypes = [H, I32]
    lib.pmhd_adjoint_get_profile.argtypes = lib.pmhd_get_profile.argtypes
    lib.pmhd_adjoint_last_error.argtypes = [H]
    lib.pmhd_adjoint_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        if not name.endswith('last_error'):
            getattr(lib, name).restype = ctypes.c_int
    return lib


def load(path=None):
    """Open the step library (default: the in-tree HIP build).  Raises OSError /
    PmhdError when it is absent -- the product has no CPU fallback.

    torch (when installed) is imported first: it brings its own HIP runtime, whose
    soname the library's dependency then resolves to, so that torch's device
    pointers and streams and the library's share one runtime."""
    path = os.path.abspath(path or DEFAULT_LIB)
    if path in _cache:
        return _cache[path]
    if not os.path.exists(path):
        raise PmhdError('libpmhd not built: %s is missing (run __graft_entry__.build())' % path)
    if path == DEFAULT_LIB:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    lib = _declare(ctypes.CDLL(path))
    lib._pmhd_gpu = path == DEFAULT_LIB or not os.path.basename(path).startswith('libpmhd_host')
    _cache[path] = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def on_gpu(lib):
    """True for the HIP library (device arrays are GPU memory); False for the host
    test build of the same kernel bodies, whose "device" arrays are host memory."""
    return getattr(lib, '_pmhd_gpu', True)


def _dptr(t, shape=None, what='array', gpu=True):
    """Device address of a contiguous float64 torch tensor (None for None).  The
    tensor must live where the library's device arrays do: GPU memory for the HIP
    library (a host address there would fault the GPU), host memory for the host
    test build."""
    if t is None:
        return None
    import torch
    if t.dtype != torch.float64 or not t.is_contiguous():
        raise ValueError('%s must be a contiguous float64 tensor' % what)
    if bool(t.is_cuda) != bool(gpu):
        raise ValueError('%s must be a %s tensor for this library' % (what, 'GPU' if gpu else 'CPU'))
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError('%s has shape %s, expected %s' % (what, tuple(t.shape), tuple(shape)))
    return ctypes.c_void_p(t.data_ptr())


def torch_stream(device=None):
    """The current torch stream's hipStream_t (as an integer), for set_stream."""
    import torch
    return torch.cuda.current_stream(device).cuda_stream


def make_grid(shape, widths, bcs):
    """PmhdGrid of a grid of cell shape (nx,) / (ny, nx) / (ny, nx, nz)."""
    shape = tuple(int(s) for s in shape)
    g = PmhdGrid()
    g.ndim = len(shape)
    if g.ndim == 1:
        g.nx, g.ny, g.nz = shape[0], 1, 1
    elif g.ndim == 2:
        g.ny, g.nx, g.nz = shape[0], shape[1], 1
    else:
        g.ny, g.nx, g.nz = shape
    w = list(widths) + [1.0] * (3 - len(widths))
    b = list(bcs) + [bcs[0]] * (3 - len(bcs))
    g.dx, g.dy, g.dz = float(w[0]), float(w[1]), float(w[2])
    g.bcx, g.bcy, g.bcz = int(b[0]), int(b[1]), int(b[2])
    return g


# Riemann solvers (PMHD_RIEMANN_*) and reference quirks a grid may correct (PMHD_FIX_*)
ROE, HLLE, ROE_HLLE = 0, 1, 2
FIX_Z_BC, FIX_OUTFLOW_GHOSTS, FIX_DTDY_80BIT = 1, 2, 4


def make_params(gamma, cfl, min_dens=1e-12, min_pres=1e-12, min_energy=1e-12, order=2,
                entropy_fix=False, epsilon=0.5, precision=64, physics=0, riemann=ROE, fix_quirks=0):
    p = PmhdParams()
    p.gamma, p.cfl = float(gamma), float(cfl)
    p.min_dens, p.min_pres, p.min_energy = float(min_dens), float(min_pres), float(min_energy)
    p.epsilon = float(epsilon)
    p.reconstruct_order = int(order)
    p.entropy_fix = 1 if entropy_fix else 0
    p.precision = int(precision)
    p.physics = int(physics)
    p.riemann = int(riemann)
    p.fix_quirks = int(fix_quirks)
    return p


MHD, HYDRO = 0, 1


def nvars(physics):
    return 5 if physics == HYDRO else 8


def face_shapes(shape, physics=MHD):
    if len(shape) == 1 or physics == HYDRO:
        return []
    if len(shape) == 2:
        ny, nx = shape
        return [(ny, nx + 1), (ny + 1, nx)]
    ny, nx, nz = shape
    return [(ny, nx + 1, nz), (ny + 1, nx, nz), (ny, nx, nz + 1)]


def _state_args(shape, prim, cons, faces, physics=MHD):
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    cons = np.ascontiguousarray(cons, dtype=np.float64)
    nv = nvars(physics)
    if prim.shape != (nv,) + shape or cons.shape != (nv,) + shape:
        raise ValueError('state shape %s does not match the grid (%d,) + %s' % (prim.shape, nv, shape))
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in faces]
    want = face_shapes(shape, physics)
    if len(fs) != len(want):
        raise ValueError('expected %d face arrays' % len(want))
    for f, s in zip(fs, want):
        if f.shape != s:
            raise ValueError('face shape %s, expected %s' % (f.shape, s))
    keep = [prim, cons] + fs
    return keep, [_ptr(prim), _ptr(cons)] + [_ptr(f) for f in fs] + [None] * (3 - len(fs))


def cfl_dt_noalloc(shape, widths, bcs, gamma, cfl, cons, faces=(), dt_out=None, physics=MHD, stream=None,
                   lib=None):
    """pmhd_cfl_dt_noalloc: the CFL timestep of a state in float64 tensors (GPU
    tensors for the HIP library, CPU tensors for the host test build) into the
    one-element float64 tensor dt_out (allocated here when None), without any
    allocation inside the library.  Queued on ``stream`` (a hipStream_t integer;
    None: the default stream); returns dt_out without synchronising."""
    import torch
    lib = lib if lib is not None else load()
    g = make_grid(tuple(int(s) for s in shape), widths, bcs)
    p = make_params(gamma, cfl, precision=64, physics=physics)
    gpu = on_gpu(lib)
    if dt_out is None:
        dt_out = torch.zeros(1, dtype=torch.float64, device='cuda' if gpu else 'cpu')
    fp = [_dptr(f, None, 'face array', gpu) for f in faces] + [None] * (3 - len(faces))
    rc = lib.pmhd_cfl_dt_noalloc(ctypes.byref(g), ctypes.byref(p), _dptr(cons, None, 'cons', gpu), *fp,
                                 _dptr(dt_out, (1,), 'dt_out', gpu),
                                 None if stream is None else ctypes.c_void_p(int(stream)))
    if rc != 0:
        raise PmhdError('pmhd_cfl_dt_noalloc failed (%d): %s' % (rc, lib.pmhd_last_error(None).decode()), rc)
    return dt_out


class Solver:
    """One device-resident grid state driven through the C ABI."""

    def __init__(self, shape, widths, bcs, gamma, cfl, min_dens=1e-12, min_pres=1e-12,
                 min_energy=1e-12, order=2, entropy_fix=False, epsilon=0.5, precision=64,
                 lib=None, physics=MHD, riemann=ROE, fix_quirks=0):
        self.lib = lib if lib is not None else load()
        self.ndim = len(shape)
        self.shape = tuple(int(s) for s in shape)   # (nx,) / (ny, nx) / (ny, nx, nz)
        self.physics = int(physics)
        g = make_grid(self.shape, widths, bcs)
        p = make_params(gamma, cfl, min_dens, min_pres, min_energy, order, entropy_fix, epsilon, precision,
                        physics, riemann, fix_quirks)
        self.grid, self.params = g, p
        h = ctypes.c_void_p()
        rc = self.lib.pmhd_create(ctypes.byref(g), ctypes.byref(p), ctypes.byref(h))
        if rc != 0:
            raise PmhdError('pmhd_create failed (%d): %s' % (rc, self.lib.pmhd_last_error(None).decode()))
        self.h = h

    def _check(self, rc, what):
        if rc != 0:
            raise PmhdError('%s failed (%d): %s' % (what, rc, self.lib.pmhd_last_error(self.h).decode()), rc)

    def face_shapes(self):
        return face_shapes(self.shape, self.physics)

    def set_state(self, prim, cons, faces=()):
        keep, args = _state_args(self.shape, prim, cons, faces, self.physics)
        self._check(self.lib.pmhd_set_state(self.h, *args), 'pmhd_set_state')
        del keep

    def get_state(self, which=0):
        prim = np.empty((nvars(self.physics),) + self.shape)
        cons = np.empty((nvars(self.physics),) + self.shape)
        fs = [np.empty(s) for s in self.face_shapes()]
        fp = [_ptr(f) for f in fs] + [None] * (3 - len(fs))
        self._check(self.lib.pmhd_get_state(self.h, which, _ptr(prim), _ptr(cons), *fp), 'pmhd_get_state')
        return prim, cons, fs

    def set_state_device(self, prim, cons, faces=()):
        """set_state from device tensors (contiguous float64, the reference layout),
        converted on the device; no host copy."""
        nv = nvars(self.physics)
        want = face_shapes(self.shape, self.physics)
        faces = list(faces)
        if len(faces) != len(want):
            raise ValueError('expected %d face arrays' % len(want))
        g = on_gpu(self.lib)
        args = [_dptr(prim, (nv,) + self.shape, 'prim', g), _dptr(cons, (nv,) + self.shape, 'cons', g)]
        args += [_dptr(f, s, 'face array', g) for f, s in zip(faces, want)] + [None] * (3 - len(want))
        self._check(self.lib.pmhd_set_state_device(self.h, *args), 'pmhd_set_state_device')

    def get_state_device(self, which=0, out=None):
        """The current (0) or second-last (1) state as new device tensors (or into
        ``out`` = (prim, cons, faces)) of float64 in the reference layout."""
        import torch
        nv = nvars(self.physics)
        if out is None:
            dev = torch.device('cuda', torch.cuda.current_device()) if on_gpu(self.lib) else torch.device('cpu')
            prim = torch.empty((nv,) + self.shape, dtype=torch.float64, device=dev)
            cons = torch.empty_like(prim)
            faces = [torch.empty(s, dtype=torch.float64, device=dev) for s in self.face_shapes()]
        else:
            prim, cons, faces = out[0], out[1], list(out[2])
        want = self.face_shapes()
        g = on_gpu(self.lib)
        args = [_dptr(prim, (nv,) + self.shape, 'prim', g), _dptr(cons, (nv,) + self.shape, 'cons', g)]
        args += [_dptr(f, s, 'face array', g) for f, s in zip(faces, want)] + [None] * (3 - len(want))
        self._check(self.lib.pmhd_get_state_device(self.h, int(which), *args), 'pmhd_get_state_device')
        return prim, cons, faces

    def set_stream(self, stream):
        """Launch on the given hipStream_t (an integer, e.g. torch_stream()); None:
        the handle's own stream."""
        self._check(self.lib.pmhd_set_stream(self.h, None if stream is None else ctypes.c_void_p(int(stream))),
                    'pmhd_set_stream')

    def compute_dt(self):
        dt = ctypes.c_double()
        self._check(self.lib.pmhd_compute_dt(self.h, ctypes.byref(dt)), 'pmhd_compute_dt')
        return dt.value

    def step(self, dt):
        self._check(self.lib.pmhd_step(self.h, float(dt)), 'pmhd_step')

    def set_sweep(self, seg):
        """Faces per thread of the fused predictor sweep along (x, y, z); 0 selects
        separate slope, edge and predictor launches (same results)."""
        seg = list(seg) + [0] * (3 - len(seg))
        self._check(self.lib.pmhd_set_sweep(self.h, int(seg[0]), int(seg[1]), int(seg[2])), 'pmhd_set_sweep')

    CLOCKS = {'auto': 0, 'host': 1, 'device': 2}

    def set_clock(self, mode):
        """Where advance() keeps the clock: 'host' (dt read back every cycle),
        'device' (dt, clipping and the loop condition on the device; same results) or
        'auto' (the device for grids of at most 2^16 cells)."""
        self._check(self.lib.pmhd_set_clock(self.h, int(self.CLOCKS[mode])), 'pmhd_set_clock')

    def advance(self, t0, t_end, cycle0, max_cycles, allow_nonfinite=False):
        """The evolve loop; returns (t, cycle, t of the second-last state).  A state
        that becomes non-finite raises PmhdError (code ENONFINITE) unless
        allow_nonfinite, which returns the reference loop's NaN clock instead."""
        t = ctypes.c_double()
        cyc = ctypes.c_int64()
        t2 = ctypes.c_double()
        rc = self.lib.pmhd_advance(self.h, float(t0), float(t_end), int(cycle0), int(max_cycles),
                                   ctypes.byref(t), ctypes.byref(cyc), ctypes.byref(t2))
        if not (rc == ENONFINITE and allow_nonfinite):
            self._check(rc, 'pmhd_advance')
        return t.value, cyc.value, t2.value

    def stats(self):
        """Floor counters per call site (FLOOR_SITES rows; columns density and
        pressure/energy), their per-cycle maxima, cycles since set_state and the
        non-finite count of the current state."""
        st = PmhdStats()
        self._check(self.lib.pmhd_get_stats(self.h, ctypes.byref(st)), 'pmhd_get_stats')
        return _stats_dict(st)

    def diagnostics(self):
        """FP64 conserved totals (8), dimensionless max |div B|, max |B_face|, non-finite count."""
        tot = (ctypes.c_double * 8)()
        dvb = ctypes.c_double()
        bmx = ctypes.c_double()
        nnf = ctypes.c_int64()
        self._check(self.lib.pmhd_diagnostics(self.h, tot, ctypes.byref(dvb), ctypes.byref(bmx),
                                              ctypes.byref(nnf)), 'pmhd_diagnostics')
        return {'totals': np.array(tot[:]), 'max_divb': dvb.value, 'max_b': bmx.value,
                'nonfinite': nnf.value}

    def set_profiling(self, on):
        self._check(self.lib.pmhd_set_profiling(self.h, 1 if on else 0), 'pmhd_set_profiling')

    def profile(self):
        """{kernel name: (total ms, launches, elements)} recorded with HIP events."""
        m = 64
        names = ctypes.create_string_buffer(32 * m)
        ms = (ctypes.c_double * m)()
        nl = (ctypes.c_int64 * m)()
        ne = (ctypes.c_int64 * m)()
        n = ctypes.c_int32()
        self._check(self.lib.pmhd_get_profile(self.h, m, names, ms, nl, ne, ctypes.byref(n)), 'pmhd_get_profile')
        out = {}
        for k in range(n.value):
            nm = names.raw[32 * k:32 * k + 32].split(b'\0', 1)[0].decode()
            out[nm] = (ms[k], nl[k], ne[k])
        return out

    def synchronize(self):
        self._check(self.lib.pmhd_synchronize(self.h), 'pmhd_synchronize')

    def close(self):
        if getattr(self, 'h', None) is not None and self.h.value:
            self.lib.pmhd_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent boxes advanced together: one launch per phase over all boxes,
    per-box CFL timestep and clock on the device (pmhd_batch_* in include/pmhd.h)."""

    def __init__(self, shapes, widths, bcs, gamma, cfl, min_dens=1e-12, min_pres=1e-12,
                 min_energy=1e-12, order=2, entropy_fix=False, epsilon=0.5, precision=64, lib=None,
                 physics=MHD, riemann=ROE, fix_quirks=0):
        self.lib = lib if lib is not None else load()
        self.physics = int(physics)
        self.shapes = [tuple(int(x) for x in s) for s in shapes]
        n = len(self.shapes)
        if n < 1:
            raise ValueError('a batch needs at least one box')
        widths = list(widths) if len(widths) == n and hasattr(widths[0], '__len__') else [widths] * n
        bcs = list(bcs) if len(bcs) == n and hasattr(bcs[0], '__len__') else [bcs] * n
        grids = (PmhdGrid * n)(*[make_grid(s, w, b) for s, w, b in zip(self.shapes, widths, bcs)])
        p = make_params(gamma, cfl, min_dens, min_pres, min_energy, order, entropy_fix, epsilon, precision,
                        physics, riemann, fix_quirks)
        self.grids, self.params = grids, p
        h = ctypes.c_void_p()
        rc = self.lib.pmhd_batch_create(n, grids, ctypes.byref(p), ctypes.byref(h))
        if rc != 0:
            raise PmhdError('pmhd_batch_create failed (%d): %s' % (rc, self.lib.pmhd_batch_last_error(None).decode()))
        self.h = h
        self.nbox = n

    def _check(self, rc, what):
        if rc != 0:
            raise PmhdError('%s failed (%d): %s' % (what, rc, self.lib.pmhd_batch_last_error(self.h).decode()))

    def _shape(self, box):
        if not 0 <= int(box) < self.nbox:
            raise PmhdError('box index %d out of range (batch of %d boxes)' % (box, self.nbox))
        return self.shapes[int(box)]

    def set_state(self, box, prim, cons, faces=()):
        keep, args = _state_args(self._shape(box), prim, cons, faces, self.physics)
        self._check(self.lib.pmhd_batch_set_state(self.h, int(box), *args), 'pmhd_batch_set_state')
        del keep

    def get_state(self, box, which=0):
        shape = self._shape(box)
        prim = np.empty((nvars(self.physics),) + shape)
        cons = np.empty((nvars(self.physics),) + shape)
        fs = [np.empty(s) for s in face_shapes(shape, self.physics)]
        fp = [_ptr(f) for f in fs] + [None] * (3 - len(fs))
        self._check(self.lib.pmhd_batch_get_state(self.h, int(box), int(which), _ptr(prim), _ptr(cons), *fp),
                    'pmhd_batch_get_state')
        return prim, cons, fs

    def advance(self, t0, t_end, cycle0, max_cycles):
        """Returns per-box arrays (t, cycle, t_second_last)."""
        t = np.zeros(self.nbox)
        cyc = np.zeros(self.nbox, dtype=np.int64)
        t2 = np.zeros(self.nbox)
        self._check(self.lib.pmhd_batch_advance(self.h, float(t0), float(t_end), int(cycle0), int(max_cycles),
                                                _ptr(t), _ptr(cyc), _ptr(t2)), 'pmhd_batch_advance')
        return t, cyc, t2

    def stats(self, box):
        """Solver.stats of one box."""
        self._shape(box)
        st = PmhdStats()
        self._check(self.lib.pmhd_batch_get_stats(self.h, int(box), ctypes.byref(st)), 'pmhd_batch_get_stats')
        return _stats_dict(st)

    def autotune(self, reps=3):
        self._check(self.lib.pmhd_batch_autotune(self.h, int(reps)), 'pmhd_batch_autotune')

    def tiles(self):
        """{launch name: (threads along z, threads along x, grid, tuned ms)}."""
        m = 64
        names = ctypes.create_string_buffer(32 * m)
        bk = np.zeros(m, dtype=np.int32)
        bi = np.zeros(m, dtype=np.int32)
        gr = np.zeros(m, dtype=np.int32)
        ms = np.zeros(m)
        n = ctypes.c_int32()
        self._check(self.lib.pmhd_batch_get_tiles(self.h, m, names, _ptr(bk), _ptr(bi), _ptr(gr), _ptr(ms),
                                                  ctypes.byref(n)), 'pmhd_batch_get_tiles')
        out = {}
        for k in range(n.value):
            nm = names.raw[32 * k:32 * k + 32].split(b'\0', 1)[0].decode()
            out[nm] = (int(bk[k]), int(bi[k]), int(gr[k]), float(ms[k]))
        return out

    def set_profiling(self, on):
        self._check(self.lib.pmhd_batch_set_profiling(self.h, 1 if on else 0), 'pmhd_batch_set_profiling')

    def profile(self):
        m = 64
        names = ctypes.create_string_buffer(32 * m)
        ms = (ctypes.c_double * m)()
        nl = (ctypes.c_int64 * m)()
        ne = (ctypes.c_int64 * m)()
        n = ctypes.c_int32()
        self._check(self.lib.pmhd_batch_get_profile(self.h, m, names, ms, nl, ne, ctypes.byref(n)),
                    'pmhd_batch_get_profile')
        out = {}
        for k in range(n.value):
            nm = names.raw[32 * k:32 * k + 32].split(b'\0', 1)[0].decode()
            out[nm] = (ms[k], nl[k], ne[k])
        return out

    def synchronize(self):
        self._check(self.lib.pmhd_batch_synchronize(self.h), 'pmhd_batch_synchronize')

    def close(self):
        if getattr(self, 'h', None) is not None and self.h.value:
            self.lib.pmhd_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
