"""First-action-potential tracker (activation time, peak, APD90) on the GPU.

No reference counterpart (the reference's save_with_activation_times records
APD at a fixed threshold, save_mesh.c:754-913).  Pinned by closed forms: the
tracker is fed piecewise-linear action potentials through the C ABI
(mono_tissue_set_v before each sample), for which the threshold crossings of
the linear segments are known exactly and linear interpolation between samples
reproduces them; then run on a real TT2006 cable, where APD90 must land in the
model's physiological range and agree with the sampled trace.
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

DT = 0.02


def _trapezoid(t, t1, rise, hold, fall, vrest=-85.0, vpeak=20.0):
    """rest until t1, linear rise over `rise` ms, plateau for `hold`, linear fall over `fall`."""
    if t <= t1:
        return vrest
    if t <= t1 + rise:
        return vrest + (vpeak - vrest) * (t - t1) / rise
    if t <= t1 + rise + hold:
        return vpeak
    if t <= t1 + rise + hold + fall:
        return vpeak - (vpeak - vrest) * (t - t1 - rise - hold) / fall
    return vrest


def test_apd90_closed_form():
    import monoalg3d_c_amd as m
    announce("APD90 tracker, piecewise-linear action potentials")
    n = 6
    t1 = np.array([0.5, 1.03, 2.0, 3.71, 5.5, 200.0])  # the last cell never activates
    rise, hold, fall = 1.0, 40.0, 60.0
    t = m.Tissue((n, 1, 1), model=m.MODEL_TT2006, tensor=m.iso_tensor(0.00013), stim=(0.0, 0.0, 0.0, 0.0, 0.0))
    t.v = np.full(n, -85.0)
    t.apd_setup(-30.0)
    time = 0.0
    for _ in range(int(round(120.0 / DT))):
        t.step(1)
        time += DT
        t.v = np.array([_trapezoid(time, a, rise, hold, fall) for a in t1])
        t.apd_sample()
    out = t.apd()
    t.close()
    # closed forms on the linear segments
    lat = t1 + rise * (-30.0 + 85.0) / 105.0
    v90 = 20.0 - 0.9 * 105.0
    t_rep = t1 + rise + hold + fall * (20.0 - v90) / 105.0
    act = np.arange(n) < n - 1
    np.testing.assert_allclose(out["lat"][act], lat[act], atol=1e-9)
    np.testing.assert_allclose(out["apd90"][act], (t_rep - lat)[act], atol=1e-9)
    np.testing.assert_array_equal(out["vmax"][act], 20.0)
    assert np.isnan(out["lat"][~act]).all() and np.isnan(out["apd90"][~act]).all()


def test_apd90_tt2006_cable():
    """A 64-cell TT2006 cable (stimulus at x < 500 um) over 450 ms: every cell
    activates once, activation times increase along the cable, APD90 of this
    first beat from the model's resting state is 300-400 ms, and it matches
    the samples: V at lat + apd90 is V90 to interpolation accuracy."""
    import monoalg3d_c_amd as m
    announce("APD90 tracker, TT2006 cable 64 cells 450 ms")
    n = 64
    t = m.Tissue((n, 1, 1), model=m.MODEL_TT2006, tensor=m.iso_tensor(0.00013))
    t.apd_setup(-30.0)
    trace = [t.v]
    for _ in range(int(round(450.0 / DT))):
        t.step(1)
        t.apd_sample()
        trace.append(t.v)
    out = t.apd()
    t.close()
    trace = np.array(trace)
    assert np.isfinite(out["lat"]).all() and np.isfinite(out["apd90"]).all()
    assert np.all(np.diff(out["lat"][8:]) > 0)
    assert 300.0 < out["apd90"].min() and out["apd90"].max() < 400.0
    vrest = trace[0]
    v90 = out["vmax"] - 0.9 * (out["vmax"] - vrest)
    t_rep = out["lat"] + out["apd90"]
    times = np.arange(len(trace)) * DT
    for c in range(n):
        assert abs(np.interp(t_rep[c], times, trace[:, c]) - v90[c]) < 1e-9
        assert out["vmax"][c] == trace[:, c].max()
