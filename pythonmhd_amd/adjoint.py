"""Gradients through runs of fixed-dt steps (pmhd_adjoint_* in include/pmhd.h).

``AdjointRun`` owns a device context for one grid: ``forward`` runs ``nsteps``
cycles of a fixed ``dt`` keeping a checkpoint every ``checkpoint_every`` steps,
``backward`` maps the cotangent of the final state to the cotangent of the
initial state of the last ``forward`` (recomputing each checkpoint segment, then
each step's intermediates, and running the adjoint of every phase).

``fixed_dt_steps`` exposes the same run to torch autograd: the initial state
(prim, cons, face arrays) are tensors, the final state is returned as tensors, and
``loss.backward()`` calls the adjoint.  Torch is only the tape here; the step and
its adjoint run on the GPU through libpmhd.  ROCm tensors are passed to the
library by device address (pmhd_adjoint_forward_device / _backward_device), on
torch's current stream: no host copies.
"""
import ctypes

import numpy as np

from ._lib import (MHD, PmhdError, _dptr, _ptr, face_shapes, load, make_grid, make_params, nvars, on_gpu,
                   torch_stream)


class AdjointRun:
    def __init__(self, shape, widths, bcs, gamma, order=1, entropy_fix=False, epsilon=0.5, precision=32,
                 physics=MHD, checkpoint_every=20, min_dens=1e-12, min_pres=1e-12, min_energy=1e-12, lib=None,
                 fix_quirks=0, deterministic=True):
        self.lib = lib if lib is not None else load()
        self.shape = tuple(int(s) for s in shape)
        self.physics = int(physics)
        g = make_grid(self.shape, widths, bcs)
        p = make_params(gamma, 1.0, min_dens, min_pres, min_energy, order, entropy_fix, epsilon, precision,
                        physics, fix_quirks=fix_quirks)
        self.grid, self.params = g, p
        self.forward_count = 0   # forward() calls so far (backward reverses the last one)
        h = ctypes.c_void_p()
        rc = self.lib.pmhd_adjoint_create(ctypes.byref(g), ctypes.byref(p), int(checkpoint_every), ctypes.byref(h))
        if rc != 0:
            raise PmhdError('pmhd_adjoint_create failed (%d): %s' %
                            (rc, self.lib.pmhd_adjoint_last_error(None).decode()))
        self.h = h
        self.set_deterministic(deterministic)

    def _check(self, rc, what):
        if rc != 0:
            raise PmhdError('%s failed (%d): %s' % (what, rc, self.lib.pmhd_adjoint_last_error(self.h).decode()))

    def _state(self, prim, cons, faces):
        nv = nvars(self.physics)
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        cons = np.ascontiguousarray(cons, dtype=np.float64)
        if prim.shape != (nv,) + self.shape or cons.shape != (nv,) + self.shape:
            raise ValueError('state shape %s does not match (%d,) + %s' % (prim.shape, nv, self.shape))
        want = face_shapes(self.shape, self.physics)
        fs = [np.ascontiguousarray(f, dtype=np.float64) for f in faces]
        if len(fs) != len(want) or any(f.shape != w for f, w in zip(fs, want)):
            raise ValueError('face arrays must have shapes %s' % (want,))
        return prim, cons, fs

    def forward(self, prim, cons, faces, dt, nsteps):
        prim, cons, fs = self._state(prim, cons, faces)
        po = np.empty_like(prim)
        co = np.empty_like(cons)
        fo = [np.empty_like(f) for f in fs]
        fi = [_ptr(f) for f in fs] + [None] * (3 - len(fs))
        fop = [_ptr(f) for f in fo] + [None] * (3 - len(fo))
        self._check(self.lib.pmhd_adjoint_forward(self.h, _ptr(prim), _ptr(cons), *fi, float(dt), int(nsteps),
                                                  _ptr(po), _ptr(co), *fop), 'pmhd_adjoint_forward')
        self.forward_count += 1
        return po, co, fo

    # ---- tensors in the library's memory space (ROCm tensors for the HIP library)
    def _tensor_args(self, tensors, shapes, what):
        g = on_gpu(self.lib)
        return [_dptr(t, s, what, g) for t, s in zip(tensors, shapes)] + [None] * (5 - len(tensors))

    def _shapes(self):
        nv = nvars(self.physics)
        return [(nv,) + self.shape, (nv,) + self.shape] + face_shapes(self.shape, self.physics)

    def _on_stream(self):
        if on_gpu(self.lib):
            self._check(self.lib.pmhd_adjoint_set_stream(self.h, ctypes.c_void_p(torch_stream())),
                        'pmhd_adjoint_set_stream')

    def forward_tensors(self, prim, cons, faces, dt, nsteps):
        """forward() on float64 tensors (device tensors for the HIP library);
        returns the final (prim, cons, faces) as new tensors on the same device."""
        import torch
        shapes = self._shapes()
        ins = [prim, cons] + list(faces)
        outs = [torch.empty(s, dtype=torch.float64, device=prim.device) for s in shapes]
        self._on_stream()
        a = self._tensor_args(ins, shapes, 'state')
        o = self._tensor_args(outs, shapes, 'state')
        self._check(self.lib.pmhd_adjoint_forward_device(self.h, a[0], a[1], a[2], a[3], a[4], float(dt), int(nsteps),
                                                         o[0], o[1], o[2], o[3], o[4]), 'pmhd_adjoint_forward_device')
        self.forward_count += 1
        return outs[0], outs[1], outs[2:]

    def backward_tensors(self, gprim=None, gcons=None, gfaces=None, device=None):
        """backward() on float64 tensors; returns the initial-state cotangents as new
        tensors.  None cotangents count as zero."""
        import torch
        shapes = self._shapes()
        nf = len(shapes) - 2
        gfaces = [None] * nf if gfaces is None else list(gfaces)
        ins = [gprim, gcons] + gfaces
        dev = device if device is not None else next(t.device for t in ins if t is not None)
        outs = [torch.empty(s, dtype=torch.float64, device=dev) for s in shapes]
        self._on_stream()
        a = self._tensor_args(ins, shapes, 'cotangent')
        o = self._tensor_args(outs, shapes, 'cotangent')
        self._check(self.lib.pmhd_adjoint_backward_device(self.h, a[0], a[1], a[2], a[3], a[4], o[0], o[1], o[2], o[3],
                                                          o[4]), 'pmhd_adjoint_backward_device')
        return outs[0], outs[1], outs[2:]

    def backward(self, gprim=None, gcons=None, gfaces=None):
        nv = nvars(self.physics)
        want = face_shapes(self.shape, self.physics)
        gp = None if gprim is None else np.ascontiguousarray(gprim, dtype=np.float64)
        gc = None if gcons is None else np.ascontiguousarray(gcons, dtype=np.float64)
        gfs = [None] * len(want) if gfaces is None else \
            [None if f is None else np.ascontiguousarray(f, dtype=np.float64) for f in gfaces]
        p0 = np.empty((nv,) + self.shape)
        c0 = np.empty((nv,) + self.shape)
        f0 = [np.empty(w) for w in want]
        gfi = [_ptr(f) for f in gfs] + [None] * (3 - len(gfs))
        f0p = [_ptr(f) for f in f0] + [None] * (3 - len(f0))
        self._check(self.lib.pmhd_adjoint_backward(self.h, _ptr(gp), _ptr(gc), *gfi, _ptr(p0), _ptr(c0), *f0p),
                    'pmhd_adjoint_backward')
        return p0, c0, f0

    def set_deterministic(self, on):
        """True (the default): backward() sums each cotangent's contributions in a
        fixed order, so repeated runs give the same bits; False: FP64 atomics."""
        self._check(self.lib.pmhd_adjoint_set_deterministic(self.h, 1 if on else 0), 'pmhd_adjoint_set_deterministic')

    def set_profiling(self, on):
        self._check(self.lib.pmhd_adjoint_set_profiling(self.h, 1 if on else 0), 'pmhd_adjoint_set_profiling')

    def close(self):
        if getattr(self, 'h', None) is not None and self.h.value:
            self.lib.pmhd_adjoint_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fixed_dt_steps(run, dt, nsteps, prim, cons, *faces):
    """Differentiable run of ``nsteps`` fixed-dt cycles: torch tensors in (initial
    prim, cons, face arrays), torch tensors out (final state, same order).

    The tensors must live where the library's arrays do: on the GPU for the HIP
    library (they go to the device entry points by address, on torch's current
    stream), on the CPU for the host test build.  The backward pass reverses the
    forward run it belongs to: a later forward() on the same ``run`` (another
    fixed_dt_steps call in the same graph) makes it raise instead of returning the
    gradient of the wrong trajectory -- use one AdjointRun per call."""
    import torch

    def f64(t):
        return None if t is None else t.detach().to(torch.float64).contiguous()

    class _Steps(torch.autograd.Function):
        @staticmethod
        def forward(ctx, prim_t, cons_t, *face_t):
            p, c, f = run.forward_tensors(f64(prim_t), f64(cons_t), [f64(x) for x in face_t], dt, nsteps)
            ctx.count = run.forward_count
            ctx.dtype = prim_t.dtype
            ctx.device = prim_t.device
            return tuple(a.to(prim_t.dtype) for a in [p, c] + list(f))

        @staticmethod
        def backward(ctx, gp, gc, *gf):
            if run.forward_count != ctx.count:
                raise RuntimeError('fixed_dt_steps: the AdjointRun ran another forward() after this one; its '
                                   'checkpoints no longer belong to this call (use one AdjointRun per call)')
            p0, c0, f0 = run.backward_tensors(f64(gp), f64(gc), [f64(x) for x in gf], device=ctx.device)
            return tuple(a.to(ctx.dtype) for a in [p0, c0] + list(f0))

    return _Steps.apply(prim, cons, *faces)
