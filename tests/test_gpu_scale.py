"""The benchmark configurations at their stated sizes on the GPU.

  - configs 2 and 3 (16M cells, ORd, rotating anisotropic fibres; config 3
    with 10% of the cells non-conducting): FP64 and FP32+refinement side by
    side over the full 100 ms: no NaN, the same cells activated (a cell may
    activate in one run only if it does so in the last 0.5 ms), activation
    times within 0.5 ms, and the largest Vm difference printed;
  - config 5 (256x256 Mitchell-Schaeffer, 64 sensors, 200 ms, checkpoints
    every 50 steps): the gradient is finite, independent of the checkpoint
    interval, the FP32-forward gradient within 1e-2 (relative L2) of the
    FP64-forward one, and central differences of the FP64 forward agree with
    the adjoint on 4 cells.
These are size-independent properties: the reference's CPU path cannot run
16M cells for 100 ms in a test, and config 5 has no reference counterpart;
the smaller versions of each are compared with the reference's own runs in
test_gpu_reference.py and with the oracle in test_gpu_adjoint.py.
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

DT = 0.02


def _desc(shape, precision, active=None):
    import monoalg3d_c_amd as m
    d = dict(model=m.MODEL_ORD2011, precision=precision, aniso=True,
             tensor=m.layered_fibre_tensors(shape[2], -60.0, 60.0, 0.00024, 0.000045), tensor_source=m.TENSOR_PER_LAYER,
             cg_rel_tol=1e-10, cg_tol=0.0, cg_max_its=500, stim=(0.0, 2.0, -50.0, 500.0, 0.0))
    if active is not None:
        d["active"] = active
    return d


@pytest.mark.parametrize("config", ["config2", "config3"])
def test_16M_fp32_refine_against_fp64_100ms(config):
    import monoalg3d_c_amd as m
    shape = (400, 400, 100)
    active = m.fibrosis_mask(shape, 0.1, 3) if config == "config3" else None
    live = np.ones(int(np.prod(shape)), bool) if active is None else active.astype(bool)
    announce(f"{config}: 400x400x100 ORd, FP64 and FP32+refine side by side, 100 ms")
    t64 = m.Tissue(shape, **_desc(shape, m.FP64, active))
    t32 = m.Tissue(shape, **_desc(shape, m.FP32_REFINE, active))
    for t in (t64, t32):
        t.apd_setup(-30.0)
    vmax = 0.0
    nsteps = int(round(100.0 / DT))
    for k in range(1, nsteps + 1):
        assert t64.step(1) == 0, f"FP64 NaN at step {k}"
        assert t32.step(1) == 0, f"FP32+refine NaN at step {k}"
        t64.apd_sample()
        t32.apd_sample()
        if k % 250 == 0:
            vmax = max(vmax, float(np.abs(t64.v - t32.v)[live].max()))
            print(f"[{config}] {k * DT:.0f} ms, max |dV| so far {vmax:.3f} mV", flush=True)
    l64, l32 = t64.apd()["lat"][live], t32.apd()["lat"][live]
    t64.close()
    t32.close()
    a64, a32 = np.isfinite(l64), np.isfinite(l32)
    one = a64 ^ a32
    assert a64.sum() > live.sum() // 4, "the front crossed a quarter of the tissue at least"
    assert (np.where(a64, l64, l32)[one] >= 100.0 - 0.5).all(), "a cell activated in one run only, early"
    both = a64 & a32
    dlat = float(np.abs(l64[both] - l32[both]).max())
    print(f"[{config}] activated {a64.sum()} / {a32.sum()} cells, in one run only {one.sum()}, max |dLAT| {dlat:.4f} ms, "
          f"max |dVm| {vmax:.3f} mV", flush=True)
    assert dlat <= 0.5, dlat


def test_config5_adjoint_full_size():
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.adjoint import config5_fields, config5_sensors, sensor_loss_grad, sensor_traces
    shape = (256, 256, 1)
    nsteps, every = 10000, 50
    announce("config 5: 256x256 MS adjoint, 64 sensors, 200 ms, checkpoints every 50")
    g, s = config5_fields(shape, seed=1)
    g2, s2 = config5_fields(shape, seed=2)
    sens = config5_sensors(shape, 64, seed=5)
    target = sensor_traces(shape, sens, nsteps, every, g=g2, s=s2)
    loss, dg, ds, _ = sensor_loss_grad(shape, sens, nsteps, every, target, checkpoint_every=50, g=g, s=s)
    assert loss > 0 and np.isfinite(dg).all() and np.isfinite(ds).all()
    l2, dg2, ds2, _ = sensor_loss_grad(shape, sens, nsteps, every, target, checkpoint_every=0, g=g, s=s)
    assert l2 == loss
    np.testing.assert_array_equal(dg2, dg)
    np.testing.assert_array_equal(ds2, ds)
    l64, dg64, ds64, _ = sensor_loss_grad(shape, sens, nsteps, every, target, checkpoint_every=50, g=g, s=s, precision=m.FP64)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"[config5] loss {loss:.6e} (FP64 forward {l64:.6e}); gradient rel. L2 FP32-forward vs FP64: "
          f"g {rel(dg, dg64):.2e}, tau_close {rel(ds, ds64):.2e}", flush=True)
    assert rel(dg, dg64) <= 1e-2 and rel(ds, ds64) <= 1e-2

    def cdiff(c, h):
        gp, gm = g.copy(), g.copy()
        gp[c] += h * g[c]
        gm[c] -= h * g[c]
        tp = sensor_traces(shape, sens, nsteps, every, g=gp, s=s, precision=m.FP64)
        tm = sensor_traces(shape, sens, nsteps, every, g=gm, s=s, precision=m.FP64)
        return (((tp - target) ** 2).sum() - ((tm - target) ** 2).sum()) / (gp[c] - gm[c])

    # the MS gate switch makes the loss piecewise smooth: each cell is
    # differenced with two steps and judged where the two agree
    agreed = 0
    for c in np.argsort(-np.abs(dg64))[:6]:
        d1, d2 = cdiff(c, 1e-4), cdiff(c, 1e-5)
        if abs(d1 - d2) > 1e-4 * abs(d1):
            continue
        agreed += 1
        print(f"[config5] cell {c}: adjoint {dg64[c]:.8e}, central difference {d1:.8e}", flush=True)
        assert abs(dg64[c] - d1) <= 1e-4 * abs(d1), (int(c), dg64[c], d1)
        if agreed == 4:
            break
    assert agreed >= 4, agreed
