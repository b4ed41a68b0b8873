"""Ionic step on the GPU (through the C ABI) against the oracle (FP64, glibc).

The FP64 kernel evaluates the reference's expressions with the device's libm
(OCML) and FMA contraction, so agreement is to rounding, not bits: per-step
relative error <= 1e-12 per state and <= 1e-9 relative on V over a 2000-step
action potential.  The FP32 kernel is checked on the action-potential shape.
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

DT = 0.02


def _stim_schedule(nsteps, n, amp):
    stim = np.zeros((nsteps, n))
    t = 0.0
    for k in range(nsteps):
        if 0.0 <= t <= 2.0:
            stim[k, :] = amp * np.linspace(1.0, 0.2, n)
        t += DT
    return stim


@pytest.mark.parametrize("name,amp", [("tt2006", -50.0), ("ord2011", -50.0), ("ms2003", 0.2)])
def test_fp64_trace_matches_oracle(oracle, name, amp):
    import monoalg3d_c_amd as m
    announce(f"ionic fp64 trace {name}")
    mid = m.MODELS[name]
    n, nsteps = 4, 2000
    stim = _stim_schedule(nsteps, n, amp)
    ref = oracle.model_init(mid, n)
    gpu = np.tile(m.model_initial_state(mid), (n, 1))
    np.testing.assert_array_equal(gpu, ref)
    worst = 0.0
    for k in range(nsteps):
        oracle.model_solve(mid, ref, stim[k], DT, 1)
        gpu = m.ionic_run(mid, gpu, stim[k], DT, 1)
        scale = np.maximum(np.abs(ref).max(axis=0), 1e-30)
        worst = max(worst, float((np.abs(gpu - ref) / scale).max()))
    assert worst < 1e-9, worst


@pytest.mark.parametrize("name,amp", [("tt2006", -50.0), ("ord2011", -50.0), ("ms2003", 0.2)])
def test_fp64_one_step_pairs(oracle, name, amp):
    """One step from states sampled along an action potential."""
    import monoalg3d_c_amd as m
    announce(f"ionic fp64 one-step {name}")
    mid = m.MODELS[name]
    sv = oracle.model_init(mid, 1)
    samples, stims = [], []
    stim = _stim_schedule(6000, 1, amp)
    for k in range(6000):
        oracle.model_solve(mid, sv, stim[k], DT, 1)
        if k % 37 == 0:
            samples.append(sv[0].copy())
            stims.append(amp if k % 2 else 0.0)
    X = np.array(samples)
    S = np.array(stims)
    ref = X.copy()
    oracle.model_solve(mid, ref, S, DT, 1)
    gpu = m.ionic_run(mid, X, S, DT, 1)
    rel = np.abs(gpu - ref) / np.maximum(np.abs(ref), 1e-300)
    rel[np.abs(ref - gpu) < 1e-300] = 0.0
    assert rel.max() < 1e-12, (rel.max(), np.unravel_index(rel.argmax(), rel.shape))


@pytest.mark.parametrize("name,amp", [("tt2006", -50.0), ("ord2011", -50.0), ("ms2003", 0.2)])
def test_fp32_action_potential_shape(oracle, name, amp):
    import monoalg3d_c_amd as m
    announce(f"ionic fp32 AP {name}")
    mid = m.MODELS[name]
    nsteps = 20000
    stim = _stim_schedule(nsteps, 1, amp)
    ref = oracle.model_init(mid, 1)
    gpu = ref.copy()
    vr, vg = [], []
    for k in range(nsteps):
        oracle.model_solve(mid, ref, stim[k], DT, 1)
        vr.append(ref[0, 0])
    # the FP32 kernel over the same schedule, 1000 steps per call
    for k0 in range(0, nsteps, 1000):
        for k in range(k0, k0 + 1000):
            gpu = m.ionic_run(mid, gpu, stim[k], DT, 1, precision=m.FP32_REFINE)
            vg.append(gpu[0, 0])
    vr, vg = np.array(vr), np.array(vg)
    span = vr.max() - vr.min()
    assert np.abs(vg - vr).max() < 0.05 * span, np.abs(vg - vr).max()


@pytest.mark.parametrize("name", ["tt2006", "ord2011"])
def test_fp32_finite_through_zero_mV(oracle, name):
    """Cells whose potential sits at (or within rounding of) 0 mV: the FP32
    kernel must stay finite where the reference's GHK factors
    vF^2/RT * (...) / (exp(k vF/RT) - 1) are 0/0 in float; FP64 and FP32
    one-step results agree to FP32 accuracy there."""
    import monoalg3d_c_amd as m
    announce(f"ionic fp32 near 0 mV {name}")
    mid = m.MODELS[name]
    sv = oracle.model_init(mid, 1)
    stim = _stim_schedule(4000, 1, -50.0)
    plateau = None
    for k in range(4000):
        oracle.model_solve(mid, sv, stim[k], DT, 1)
        if k > 150 and plateau is None and sv[0, 0] > 5.0:
            plateau = sv[0].copy()
    assert plateau is not None
    vs = np.array([0.0, 1e-9, -1e-9, 1e-7, -1e-7, 3e-6, -3e-6, 1e-3])
    X = np.tile(plateau, (len(vs), 1))
    X[:, 0] = vs
    g32 = m.ionic_run(mid, X, None, DT, 1, precision=m.FP32_REFINE)
    g64 = m.ionic_run(mid, X, None, DT, 1)
    assert np.all(np.isfinite(g32)), g32[:, 0]
    # FP64 keeps the reference's form, which is 0/0 at exactly 0 mV (as in the
    # reference); compare where it is finite
    ok = np.all(np.isfinite(g64), axis=1)
    assert ok.sum() >= len(vs) - 1
    rel = np.abs(g32[ok] - g64[ok]) / np.maximum(np.abs(g64[ok]), 1e-6)
    assert rel.max() < 1e-3, rel.max()
