"""The ionic kernel's launch-shape tuner (the first steps of a run time each
candidate block size, then the fastest is kept): the block size does not enter
any cell's arithmetic, so a tuned run and an untuned one (MONO_TUNE=0 when the
tissue is created) give the same states bit for bit, in both precisions, and
the tuner settles on one of its candidates."""
import os

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("model", ["tt2006", "ord2011"])
def test_tuned_run_equals_untuned_run(precision, model, monkeypatch):
    import monoalg3d_c_amd as m
    announce(f"ionic launch tuner {model} {precision}")
    prec = m.FP64 if precision == "fp64" else m.FP32_REFINE
    shape = (40, 36, 6)

    def run(tune):
        monkeypatch.setenv("MONO_TUNE", "1" if tune else "0")
        t = m.Tissue(shape, model=m.MODELS[model], precision=prec, tensor=m.iso_tensor(0.00013), cg_tol=1e-16)
        t.step(40)
        out = (t.v, t.state, t.stats)
        t.close()
        return out

    v1, s1, st1 = run(True)
    v0, s0, st0 = run(False)
    assert st1["ionic_block"] in (64, 128, 256), st1["ionic_block"]
    assert st0["ionic_block"] == 0
    np.testing.assert_array_equal(v1, v0)
    np.testing.assert_array_equal(s1, s0)
    assert st1["cg_iterations_total"] == st0["cg_iterations_total"]
