"""Sensor recordings, a squared-error loss against target recordings, and its
gradient with respect to per-cell conductivity factors and Mitchell-Schaeffer
tau_close factors (mono_sensor_gradient: discrete adjoint of the
operator-split step, FP64 adjoint solves, checkpointed forward).  No
counterpart in the reference.

The tissue is a Mitchell-Schaeffer sheet or box (mitchell_shaeffer_2003.c) on an
isotropic tensor scaled per cell by g (the conductivity of cell c is
sigma * g[c]; faces take the harmonic mean, assembly_common.c:911-1048), with
tau_close of cell c multiplied by s[c].

    loss, dg, ds, traces = sensor_loss_grad(shape, sensors, nsteps, every, target, g=g, s=s)

and, for torch, `sensor_loss(g, s, spec)` is differentiable in g and s.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import FP32_REFINE, MODEL_MS2003, check, iso_tensor, lib, make_desc, ptr


def ms_desc_kwargs(shape, g=None, s=None, precision=FP32_REFINE, sigma=0.00013, dt=0.02,
                   stim=(0.0, 2.0, 0.2, 500.0, 0.0), cg_rel_tol=1e-10, **kw):
    """make_desc arguments of a Mitchell-Schaeffer tissue with per-cell g and s."""
    n = int(np.prod(shape))
    out = dict(model=MODEL_MS2003, precision=precision, tensor=iso_tensor(sigma), dt=dt, stim=stim,
               cg_rel_tol=cg_rel_tol, cg_tol=0.0, cg_max_its=1000)
    if g is not None:
        out["g_cell"] = np.ascontiguousarray(g, np.float64).reshape(n)
    if s is not None:
        sc = np.ones((n, 4), np.float32)
        sc[:, 0] = np.asarray(s, np.float32).reshape(n)
        out["ionic_scale"] = sc
    out.update(kw)
    return out


def _call(shape, sensors, nsteps, record_every, target, checkpoint_every, want_grad, **kw):
    d, keep = make_desc(tuple(shape), **ms_desc_kwargs(shape, **kw))
    n = int(np.prod(shape))
    sens = np.ascontiguousarray(sensors, np.int64)
    nrec = nsteps // record_every + 1
    traces = np.zeros((nrec, len(sens)))
    tgt = None if target is None else np.ascontiguousarray(target, np.float64).reshape(nrec, len(sens))
    loss = C.c_double(0.0)
    gg = np.zeros(n) if want_grad else None
    gs = np.zeros(n) if want_grad else None
    check(lib().mono_sensor_gradient(C.byref(d), int(nsteps), ptr(sens), len(sens), int(record_every), ptr(tgt),
                                     int(checkpoint_every), ptr(traces), C.byref(loss), ptr(gg), ptr(gs)),
          "mono_sensor_gradient")
    return loss.value, gg, gs, traces


def sensor_traces(shape, sensors, nsteps, record_every, **kw) -> np.ndarray:
    """V at the sensor cells every record_every steps from step 0: (nrec, nsens)."""
    return _call(shape, sensors, nsteps, record_every, None, 1 << 30, False, **kw)[3]


def sensor_loss_grad(shape, sensors, nsteps, record_every, target, checkpoint_every=50, **kw):
    """(loss, dloss/dg, dloss/ds, traces) for sum (traces - target)^2."""
    return _call(shape, sensors, nsteps, record_every, target, checkpoint_every, True, **kw)


try:
    import torch

    class SensorLoss(torch.autograd.Function):
        """loss(g, s) with the adjoint gradient; spec holds shape, sensors,
        nsteps, record_every, target, checkpoint_every and make_desc options."""

        @staticmethod
        def forward(ctx, g, s, spec):
            kw = dict(spec)
            shape, sensors = kw.pop("shape"), kw.pop("sensors")
            nsteps, every, target = kw.pop("nsteps"), kw.pop("record_every"), kw.pop("target")
            ck = kw.pop("checkpoint_every", 50)
            loss, dg, ds, _ = sensor_loss_grad(shape, sensors, nsteps, every, target, checkpoint_every=ck,
                                               g=g.detach().cpu().double().numpy(), s=s.detach().cpu().double().numpy(), **kw)
            ctx.save_for_backward(torch.from_numpy(dg).to(g.dtype), torch.from_numpy(ds).to(s.dtype))
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            dg, ds = ctx.saved_tensors
            return grad_out.to(dg.dtype) * dg, grad_out.to(ds.dtype) * ds, None

    def sensor_loss(g, s, spec):
        return SensorLoss.apply(g, s, spec)
except ImportError:  # torch is optional for the numpy interface
    pass


def config5_fields(shape=(256, 256, 1), seed=1):
    """Seeded per-cell fields: g ~ log-normal(0, 0.3), s (tau_close factor) ~ U[0.8, 1.2]."""
    n = int(np.prod(shape))
    rng = np.random.default_rng(seed)
    g = np.exp(rng.normal(0.0, 0.3, n))
    s = rng.uniform(0.8, 1.2, n).astype(np.float32).astype(np.float64)
    return g, s


def config5_sensors(shape=(256, 256, 1), count=64, seed=5):
    n = int(np.prod(shape))
    return np.sort(np.random.default_rng(seed).choice(n, size=count, replace=False)).astype(np.int64)
