"""Operator-split tissue run on the GPU against the oracle's run (FP64).

Both run the same loop (monodomain_solver.c:616-925) on the same inputs; the
oracle is pinned to the reference separately.  The reference's own CG stops on
||r||^2 <= 1e-16, which leaves ~1e-9 mV of solver noise, so Vm is compared to
1e-8 relative to the Vm range (100 mV), activation times to one step.
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

H = 100.0


def _run_pair(oracle, shape, model, steps, aniso=False, active=None, g_cell=None, precision=None):
    import monoalg3d_c_amd as m
    nx, ny, nz = shape
    n = nx * ny * nz
    a = oracle.alpha(0.14, 1.0, 0.02, H, H, H)
    g = np.ones(n) if g_cell is None else g_cell
    if aniso:
        t6 = m.layered_fibre_tensors(nz, -60.0, 60.0, 0.00024, 0.000045)
        coef = oracle.assemble_aniso(shape, (H, H, H), a, np.repeat(t6, nx * ny, axis=0) * g[:, None], active)
        kw = dict(aniso=True, tensor=t6, tensor_source=m.TENSOR_PER_LAYER)
    else:
        coef = oracle.assemble_iso(shape, (H, H, H), a, np.full((n, 3), 0.00013) * g[:, None], active)
        kw = dict(tensor=m.iso_tensor(0.00013))
    if active is not None:
        kw["active"] = active
    if g_cell is not None:
        kw["g_cell"] = g_cell
    if precision is not None:
        kw["precision"] = precision
    mid = oracle.MODEL_IDS[model]
    orc = oracle.Tissue(shape, (H, H, H), mid, coef, active=active)
    gpu = m.Tissue(shape, model=mid, **kw)
    worst, lat_r, lat_g = 0.0, np.full(n, np.inf), np.full(n, np.inf)
    live = np.ones(n, bool) if active is None else active.astype(bool)
    for k in range(steps):
        orc.step()
        gpu.step()
        vr, vg = orc.v, gpu.v
        worst = max(worst, float(np.abs(vr - vg)[live].max()))
        t = (k + 1) * 0.02
        lat_r[(vr > -30.0) & np.isinf(lat_r)] = t
        lat_g[(vg > -30.0) & np.isinf(lat_g)] = t
    return worst, lat_r, lat_g, gpu.stats, orc


@pytest.mark.parametrize("model", ["tt2006", "ord2011"])
def test_iso_tissue_matches_oracle(oracle, model):
    announce(f"tissue iso 16x16x4 {model} 10 ms")
    worst, lr, lg, st, _ = _run_pair(oracle, (16, 16, 4), model, 500)
    assert st["class_operator"] == 1 and st["stencil_points"] == 7
    assert worst < 1e-6, worst
    act = np.isfinite(lr)
    assert act.sum() > 0
    assert np.array_equal(act, np.isfinite(lg))
    assert np.abs(lr[act] - lg[act]).max() <= 0.02 + 1e-12


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("storage", ["codes", "per_cell"])
def test_iso_tissue_storage_modes_match_oracle(oracle, storage, fp32):
    """Isotropic rows with 15% inactive cells are stored as one-byte neighbour
    codes into a 128-row table; with a per-cell conductivity factor as well the
    rows differ cell by cell and are stored per cell.  Both against the oracle
    (FP64: Vm within 1e-6 mV; FP32 + refinement: activation times within one
    step of 0.02 ms, since both solves meet the same tolerance)."""
    import monoalg3d_c_amd as m
    shape = (20, 14, 4)
    n = 20 * 14 * 4
    rng = np.random.default_rng(21)
    act = (rng.random(n) >= 0.15).astype(np.uint8)
    act[:20 * 14] = 1  # keep the stimulated face connected
    g = np.exp(rng.normal(0.0, 0.3, n)) if storage == "per_cell" else None
    announce(f"tissue iso holes {shape} storage={storage} fp32={fp32}")
    worst, lr, lg, st, _ = _run_pair(oracle, shape, "tt2006", 400, active=act, g_cell=g,
                                     precision=m.FP32_REFINE if fp32 else None)
    assert st["class_operator"] == (2 if storage == "codes" else 0) and st["stencil_points"] == 7
    live = act.astype(bool)
    fin = np.isfinite(lr) & live
    assert fin.sum() > 50
    assert np.array_equal(fin, np.isfinite(lg) & live)
    if fp32:
        assert np.abs(lr[fin] - lg[fin]).max() <= 0.02 + 1e-12
    else:
        assert worst < 1e-6, worst
        assert np.abs(lr[fin] - lg[fin]).max() <= 0.02 + 1e-12


def test_aniso_tissue_matches_oracle(oracle):
    announce("tissue aniso 12x12x6 ord2011 5 ms")
    worst, lr, lg, st, _ = _run_pair(oracle, (12, 12, 6), "ord2011", 250, aniso=True)
    # interior rows hold 19 non-zeros; boundary rows use other corner slots, so the
    # stored slot union is 27
    assert st["class_operator"] == 1 and st["stencil_points"] == 27
    assert worst < 1e-6, worst


def test_aniso_fibrotic_tissue_matches_oracle(oracle):
    """Layered fibres with 10% inactive cells: the rows differ cell by cell, so
    the FP64 solve forms them from 4-byte neighbourhood masks (mask mode,
    class_operator 3); the oracle multiplies by its assembled rows.  Vm within
    1e-6 mV, activation times within one step."""
    shape = (16, 14, 6)
    n = 16 * 14 * 6
    rng = np.random.default_rng(33)
    act = (rng.random(n) >= 0.1).astype(np.uint8)
    act[:16 * 14] = 1  # keep the stimulated face connected
    announce(f"tissue aniso fibrotic {shape} tt2006 6 ms (mask rows)")
    worst, lr, lg, st, _ = _run_pair(oracle, shape, "tt2006", 300, aniso=True, active=act)
    assert st["class_operator"] == 3 and st["stencil_points"] == 27
    assert worst < 1e-6, worst
    live = act.astype(bool)
    fin = np.isfinite(lr) & live
    assert fin.sum() > 50
    assert np.array_equal(fin, np.isfinite(lg) & live)
    assert np.abs(lr[fin] - lg[fin]).max() <= 0.02 + 1e-12
