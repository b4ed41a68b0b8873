"""Adjoint gradient of a sensor loss (mono_sensor_gradient) on the GPU.

No reference counterpart: the gradient is checked against central finite
differences of the oracle's own forward run (the oracle is pinned to the
reference's tissue runs elsewhere), against the FP64-forward gradient when the
forward runs in FP32 + refinement, for independence from the checkpoint
interval, and through the torch.autograd.Function wrapper.
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

H = (100.0, 100.0, 100.0)
SHAPE = (32, 24, 1)
N = 32 * 24
NSTEPS, EVERY = 600, 25
STIM = (0.0, 2.0, 0.2, 500.0, 0.0)


def _case():
    from monoalg3d_c_amd.adjoint import config5_fields
    g, s = config5_fields(SHAPE, seed=1)
    g2, s2 = config5_fields(SHAPE, seed=2)
    sensors = np.sort(np.random.default_rng(3).choice(N, 8, replace=False)).astype(np.int64)
    return g, s, g2, s2, sensors


def _oracle_loss(oracle, g, s, sensors, target):
    """The oracle's forward run (FP64, tight CG) and the same loss."""
    n = N
    a = oracle.alpha(0.14, 1.0, 0.02, *H)
    coef = oracle.assemble_iso(SHAPE, H, a, np.full((n, 3), 0.00013) * g[:, None])
    sc = np.ones((n, 4), np.float32)
    sc[:, 0] = s
    t = oracle.Tissue(SHAPE, H, oracle.MODEL_IDS["ms2003"], coef, stim=STIM, ionic_scale=sc, cg_tol=1e-26, cg_max_its=2000)
    rec = [t.v[sensors].copy()]
    for _ in range(NSTEPS // EVERY):
        t.step(EVERY)
        rec.append(t.v[sensors].copy())
    return float(((np.array(rec) - target) ** 2).sum())


def test_gradient_matches_oracle_finite_differences(oracle):
    """dL/dg and dL/ds of 6 cells each against central differences of the
    oracle's loss: within 1e-4 relative (plus 1e-6 of the largest component
    for components near zero).  s is stored in float32, so the difference
    quotient divides by the float32 step actually taken."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.adjoint import sensor_loss_grad, sensor_traces
    announce("adjoint FP64 forward vs oracle finite differences, 32x24 MS 12 ms")
    g, s, g2, s2, sensors = _case()
    target = sensor_traces(SHAPE, sensors, NSTEPS, EVERY, g=g2, s=s2, precision=m.FP64, stim=STIM)
    loss, dg, ds, tr = sensor_loss_grad(SHAPE, sensors, NSTEPS, EVERY, target, checkpoint_every=40, g=g, s=s,
                                        precision=m.FP64, stim=STIM)
    assert loss > 0 and np.isfinite(dg).all() and np.isfinite(ds).all()
    assert abs(loss - ((tr - target) ** 2).sum()) <= 1e-12 * loss
    # forward CG to a relative residual of 1e-10 on the GPU, 1e-13 on the oracle
    assert abs(_oracle_loss(oracle, g, s, sensors, target) - loss) <= 1e-5 * loss
    rng = np.random.default_rng(4)
    # cells where the gradient is not negligible, plus random ones
    for name, grad in (("g", dg), ("s", ds)):
        big = np.argsort(-np.abs(grad))[:3]
        cells = np.concatenate([big, rng.choice(N, 3, replace=False)])
        for c in cells:
            if name == "g":
                eps = 1e-4 * g[c]
                gp, gm = g.copy(), g.copy()
                gp[c] += eps
                gm[c] -= eps
                lp, lm = _oracle_loss(oracle, gp, s, sensors, target), _oracle_loss(oracle, gm, s, sensors, target)
                step = gp[c] - gm[c]
            else:
                sp, sm = s.copy(), s.copy()
                sp[c] = float(np.float32(s[c] * (1 + 1e-3)))
                sm[c] = float(np.float32(s[c] * (1 - 1e-3)))
                lp, lm = _oracle_loss(oracle, g, sp, sensors, target), _oracle_loss(oracle, g, sm, sensors, target)
                step = sp[c] - sm[c]
            fd = (lp - lm) / step
            tol = 1e-4 * abs(fd) + 1e-6 * np.abs(grad).max()
            assert abs(grad[c] - fd) <= tol, (name, int(c), grad[c], fd)


def test_fp32_forward_gradient_close_to_fp64():
    """FP32 states and CG with FP64 refinement in the forward, FP64 adjoint:
    both gradients within 1% (relative L2) of the FP64-forward gradient."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.adjoint import sensor_loss_grad, sensor_traces
    announce("adjoint FP32+refine forward vs FP64 forward")
    g, s, g2, s2, sensors = _case()
    target = sensor_traces(SHAPE, sensors, NSTEPS, EVERY, g=g2, s=s2, precision=m.FP64, stim=STIM)
    l64, dg64, ds64, _ = sensor_loss_grad(SHAPE, sensors, NSTEPS, EVERY, target, g=g, s=s, precision=m.FP64, stim=STIM)
    l32, dg32, ds32, _ = sensor_loss_grad(SHAPE, sensors, NSTEPS, EVERY, target, g=g, s=s, precision=m.FP32_REFINE, stim=STIM)
    assert abs(l32 - l64) <= 1e-3 * l64
    for a, b in ((dg32, dg64), (ds32, ds64)):
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b)


@pytest.mark.parametrize("prec_name", ["FP64", "FP32_REFINE"])
def test_checkpoint_interval_does_not_change_gradient(prec_name):
    """Recomputing from checkpoints every 7 steps, keeping one segment, or
    keeping every step's states without recomputation gives the same gradient
    bit for bit (the forward kernels are deterministic, and a checkpoint keeps
    the mixed solve's sweep schedule)."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.adjoint import sensor_loss_grad, sensor_traces
    prec = getattr(m, prec_name)
    announce(f"adjoint checkpoint interval 7 vs 1000 vs all, {prec_name}")
    g, s, g2, s2, sensors = _case()
    target = sensor_traces(SHAPE, sensors, 200, EVERY, g=g2, s=s2, precision=m.FP64, stim=STIM)
    a = sensor_loss_grad(SHAPE, sensors, 200, EVERY, target, checkpoint_every=7, g=g, s=s, precision=prec, stim=STIM)
    b = sensor_loss_grad(SHAPE, sensors, 200, EVERY, target, checkpoint_every=1000, g=g, s=s, precision=prec, stim=STIM)
    c = sensor_loss_grad(SHAPE, sensors, 200, EVERY, target, checkpoint_every=0, g=g, s=s, precision=prec, stim=STIM)
    for other in (b, c):
        assert a[0] == other[0]
        np.testing.assert_array_equal(a[1], other[1])
        np.testing.assert_array_equal(a[2], other[2])


def test_torch_autograd_function():
    import torch

    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.adjoint import sensor_loss, sensor_loss_grad, sensor_traces
    announce("adjoint through torch.autograd.Function")
    g, s, g2, s2, sensors = _case()
    target = sensor_traces(SHAPE, sensors, 200, EVERY, g=g2, s=s2, precision=m.FP64, stim=STIM)
    spec = dict(shape=SHAPE, sensors=sensors, nsteps=200, record_every=EVERY, target=target, precision=m.FP64, stim=STIM)
    gt = torch.tensor(g, dtype=torch.float64, requires_grad=True)
    st = torch.tensor(s, dtype=torch.float64, requires_grad=True)
    loss = 3.0 * sensor_loss(gt, st, spec)
    loss.backward()
    l0, dg, ds, _ = sensor_loss_grad(SHAPE, sensors, 200, EVERY, target, g=g, s=s, precision=m.FP64, stim=STIM)
    assert float(loss) == pytest.approx(3.0 * l0, rel=1e-15)
    np.testing.assert_allclose(gt.grad.numpy(), 3.0 * dg, rtol=1e-15, atol=0)
    np.testing.assert_allclose(st.grad.numpy(), 3.0 * ds, rtol=1e-15, atol=0)
