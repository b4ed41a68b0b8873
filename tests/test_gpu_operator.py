"""Device assembly and CG through the C ABI against the oracle.

Assembly is compared bit for bit (the device kernel forms each coefficient with
the same operations, no FMA contraction).  CG results are compared to the
oracle's CG to 1e-10 relative (different summation order).
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

H = (100.0, 100.0, 100.0)


def _fibres(shape):
    import monoalg3d_c_amd as m
    nx, ny, nz = shape
    return m.layered_fibre_tensors(nz, -60.0, 60.0, 0.00024, 0.000045)


@pytest.mark.parametrize("holes", [False, True])
def test_iso_assembly_bit_exact(oracle, holes):
    import monoalg3d_c_amd as m
    announce(f"assembly iso holes={holes}")
    shape = (9, 7, 5)
    n = 9 * 7 * 5
    act = None
    if holes:
        act = (np.random.default_rng(3).random(n) > 0.2).astype(np.uint8)
    coef, a = m.assemble(shape, tensor=m.iso_tensor(0.00013), active=act)
    ref = oracle.assemble_iso(shape, H, oracle.alpha(0.14, 1.0, 0.02, *H), np.full((n, 3), 0.00013), act)
    assert a == oracle.alpha(0.14, 1.0, 0.02, *H)
    np.testing.assert_array_equal(coef, ref)


@pytest.mark.parametrize("holes", [False, True])
def test_aniso_assembly_bit_exact(oracle, holes):
    import monoalg3d_c_amd as m
    announce(f"assembly aniso holes={holes}")
    shape = (8, 6, 5)
    n = 8 * 6 * 5
    t6 = _fibres(shape)
    act = None
    if holes:
        act = (np.random.default_rng(4).random(n) > 0.15).astype(np.uint8)
    coef, a = m.assemble(shape, aniso=True, tensor=t6, tensor_source=m.TENSOR_PER_LAYER, active=act)
    ref = oracle.assemble_aniso(shape, H, a, np.repeat(t6, 8 * 6, axis=0), act)
    np.testing.assert_array_equal(coef, ref)


def test_conductivity_factor_scales_tensor(oracle):
    """Per-cell factor g: D_i = g_i D (a field the reference has no input for);
    checked against the oracle fed the scaled per-cell tensors."""
    import monoalg3d_c_amd as m
    announce("assembly aniso with per-cell conductivity factor")
    shape = (6, 6, 4)
    n = 6 * 6 * 4
    g = np.exp(np.random.default_rng(5).normal(0.0, 0.3, n))
    t6 = _fibres(shape)
    coef, a = m.assemble(shape, aniso=True, tensor=t6, tensor_source=m.TENSOR_PER_LAYER, g_cell=g)
    per_cell = np.repeat(t6, 36, axis=0) * g[:, None]
    ref = oracle.assemble_aniso(shape, H, a, per_cell)
    np.testing.assert_array_equal(coef, ref)


@pytest.mark.parametrize("aniso,jacobi", [(False, False), (False, True), (True, False)])
def test_cg_stencil_matches_oracle(oracle, aniso, jacobi):
    import monoalg3d_c_amd as m
    announce(f"cg stencil aniso={aniso} jacobi={jacobi}")
    shape = (10, 9, 6)
    n = 10 * 9 * 6
    if aniso:
        coef, a = m.assemble(shape, aniso=True, tensor=_fibres(shape), tensor_source=m.TENSOR_PER_LAYER)
    else:
        coef, a = m.assemble(shape, tensor=m.iso_tensor(0.00013))
    rng = np.random.default_rng(6)
    b = rng.normal(size=n) * a * 80.0
    x0 = np.full(n, -80.0)
    xr, itr, er = oracle.cg_stencil(shape, coef, b, x0, 1e-20, 500, jacobi)
    xg, itg, eg = m.cg_stencil(shape, coef, b, x0, tol=1e-20, max_its=500, jacobi=jacobi)
    assert abs(itr - itg) <= 1
    assert np.abs(xg - xr).max() <= 1e-10 * np.abs(xr).max()
    # the answer satisfies the system
    r = b - oracle.stencil_apply(shape, coef, xg)
    assert (r @ r) <= 1e-18


@pytest.mark.parametrize("max_its", [1, 2, 5])
def test_cg_stencil_iteration_cap_matches_oracle(oracle, max_its):
    """A solve stopped by max_its (linear_system_solver.c:159-163) returns the
    iterate after exactly that many steps; on boxes of several planes the last
    x += alpha p is applied by the kernel that ends the solve."""
    import monoalg3d_c_amd as m
    announce(f"cg stencil max_its={max_its}")
    shape = (10, 9, 6)
    n = 10 * 9 * 6
    coef, a = m.assemble(shape, aniso=True, tensor=_fibres(shape), tensor_source=m.TENSOR_PER_LAYER)
    rng = np.random.default_rng(8)
    b = rng.normal(size=n) * a * 80.0
    x0 = np.full(n, -80.0)
    xr, itr, er = oracle.cg_stencil(shape, coef, b, x0, 1e-30, max_its, False)
    xg, itg, eg = m.cg_stencil(shape, coef, b, x0, tol=1e-30, max_its=max_its)
    assert itg == itr == max_its
    assert np.abs(xg - xr).max() <= 1e-10 * np.abs(xr).max()
    assert abs(eg - er) <= 1e-8 * er


def test_cg_stencil_converged_start(oracle):
    """x0 already solves the system: no iteration, x unchanged."""
    import monoalg3d_c_amd as m
    announce("cg stencil converged start")
    shape = (10, 9, 6)
    n = 10 * 9 * 6
    coef, a = m.assemble(shape, tensor=m.iso_tensor(0.00013))
    x0 = np.random.default_rng(9).normal(size=n) - 80.0
    b = oracle.stencil_apply(shape, coef, x0)
    xg, itg, eg = m.cg_stencil(shape, coef, b, x0, tol=1e-6, max_its=50)
    assert itg == 1
    np.testing.assert_array_equal(xg, x0)


def test_cg_mixed_reaches_fp64_tolerance(oracle):
    import monoalg3d_c_amd as m
    announce("cg stencil mixed precision")
    shape = (12, 10, 8)
    n = 12 * 10 * 8
    coef, a = m.assemble(shape, aniso=True, tensor=_fibres(shape), tensor_source=m.TENSOR_PER_LAYER)
    rng = np.random.default_rng(7)
    b = rng.normal(size=n) * a * 80.0
    x0 = np.zeros(n)
    xg, its, err = m.cg_stencil(shape, coef, b, x0, rel_tol=1e-10, max_its=500, precision=m.FP32_REFINE)
    r = b - oracle.stencil_apply(shape, coef, xg)
    assert np.sqrt(r @ r) <= 1e-10 * np.sqrt(b @ b) * 1.0001
    assert err <= 1e-20 * (b @ b)
