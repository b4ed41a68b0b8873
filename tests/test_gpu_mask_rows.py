"""Mask mode: rows formed from a 4-byte neighbourhood mask per cell.

On a fibrotic box with layered anisotropic fibres (the 16M-cell workload's
operator at 48 x 40 x 12) the operator builder enables mask mode after the
FP64 mask rows reproduced every assembled row bit for bit; the solves then
form their rows from the masks (the FP64 CG and the mixed solve's outer
residuals in FP64, its inner sweeps in FP32).  Against the same run with
stored rows (MONO_MASK_ROWS=0 when the tissue is created): the same CG
iteration counts and Vm within 1e-4 mV over 200 steps for FP32+refine, within
1e-7 mV over 100 steps for FP64 (the two differ by the rounding of the formed
rows, ~1e-7 of a row's diagonal in FP32, ~1e-16 in FP64).
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

SHAPE = (48, 40, 12)


def _desc(m, precision, active, jacobi=False):
    nz = SHAPE[2]
    t6 = m.layered_fibre_tensors(nz, -60.0, 60.0, 0.00024, 0.000045)
    return dict(model=m.MODEL_ORD2011, precision=precision, aniso=True, tensor=t6, tensor_source=m.TENSOR_PER_LAYER,
                cg_rel_tol=1e-10, cg_tol=0.0, cg_max_its=500, stim=(0.0, 2.0, -50.0, 500.0, 0.0), active=active,
                jacobi=jacobi)


def _run(m, monkeypatch, precision, mask, nsteps=200, jacobi=False):
    monkeypatch.setenv("MONO_MASK_ROWS", "1" if mask else "0")
    t = m.Tissue(SHAPE, **_desc(m, precision, m.fibrosis_mask(SHAPE, 0.1, 3), jacobi))
    t.step(nsteps)
    out = (t.v, t.state, t.stats)
    t.close()
    return out


def test_mask_rows_match_stored_rows_fp32_refine(monkeypatch):
    import monoalg3d_c_amd as m
    announce(f"mask rows vs stored rows, FP32+refine, fibrotic anisotropic {SHAPE}")
    v1, _, st1 = _run(m, monkeypatch, m.FP32_REFINE, True)
    v0, _, st0 = _run(m, monkeypatch, m.FP32_REFINE, False)
    assert st1["class_operator"] == 3 and st0["class_operator"] == 0
    assert st1["inner_iterations_total"] == st0["inner_iterations_total"]
    assert st1["outer_sweeps_total"] == st0["outer_sweeps_total"]
    d = float(np.abs(v1 - v0).max())
    print(f"[mask rows] max |dV| mask vs stored rows over 200 steps: {d:.3e} mV", flush=True)
    assert d < 1e-4, d
    # the run is not trivial: the front is inside the box
    assert (v1 > -30.0).any() and (v1 < -70.0).any()


def test_fp64_mask_rows_match_stored_rows(monkeypatch):
    import monoalg3d_c_amd as m
    announce(f"mask rows vs stored rows, FP64, fibrotic anisotropic {SHAPE}")
    v1, s1, st1 = _run(m, monkeypatch, m.FP64, True, nsteps=100)
    v0, s0, st0 = _run(m, monkeypatch, m.FP64, False, nsteps=100)
    assert st1["class_operator"] == 3 and st0["class_operator"] == 0
    assert st1["cg_iterations_total"] == st0["cg_iterations_total"]
    d = float(np.abs(v1 - v0).max())
    ds = float(np.abs(s1 - s0).max())
    print(f"[mask rows] FP64 max |dV| mask vs stored rows over 100 steps: {d:.3e} mV, "
          f"max |d state| {ds:.3e}", flush=True)
    assert d < 1e-7, d
    assert (v1 > -30.0).any() and (v1 < -70.0).any()


def test_fp64_jacobi_in_mask_mode(monkeypatch):
    """The Jacobi update reads row diagonals from the stored rows, so a
    preconditioned FP64 solve on a mask-mode operator multiplies by the stored
    rows: against the unpreconditioned solve over mask rows, Vm within 1e-4 mV
    over 100 steps (both stop at a relative residual of 1e-10, ~1e-8 mV per
    step, which the upstroke amplifies)."""
    import monoalg3d_c_amd as m
    announce(f"FP64 Jacobi solve on a mask-mode operator, fibrotic anisotropic {SHAPE}")
    vj, _, stj = _run(m, monkeypatch, m.FP64, True, nsteps=100, jacobi=True)
    v1, _, st1 = _run(m, monkeypatch, m.FP64, True, nsteps=100)
    assert stj["class_operator"] == 3 and st1["class_operator"] == 3
    assert np.isfinite(vj).all()
    d = float(np.abs(vj - v1).max())
    print(f"[mask rows] FP64 Jacobi vs unpreconditioned over 100 steps: max |dV| {d:.3e} mV, "
          f"CG iterations {stj['cg_iterations_total']} vs {st1['cg_iterations_total']}", flush=True)
    assert d < 1e-4, d
