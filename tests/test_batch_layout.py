"""Batch packing (host logic) and the batched run on the oracle, without a GPU.

A batch must be block diagonal: every instance advanced inside the atlas must
equal its own stand-alone run.  The oracle runs both on the CPU (the atlas with
the per-cell stimulus mask and model scalings, each instance alone with its
scalings everywhere), which checks the packing, the masks and the scale fields
that the GPU batch is built from.
"""
import numpy as np
import pytest

H = (100.0, 100.0, 100.0)


def _insts():
    from monoalg3d_c_amd.batch import Instance
    return [Instance(30, 1, (1.2, 0.8, 1.1, 1.0)), Instance(12, 9, (0.7, 1.3, 0.9, 1.0)),
            Instance(25, 1), Instance(8, 8, (1.0, 1.0, 1.5, 1.0), stim_x_limit=300.0), Instance(5, 14)]


def test_layout_separates_instances():
    from monoalg3d_c_amd.batch import Layout, sweep_instances
    for insts in (_insts(), sweep_instances(64, 3, cable_cells=(10, 200), sheet_side=(4, 40))):
        L = Layout(insts)
        W, Hh, nz = L.shape
        assert nz == 1 and W % 64 == 0
        owner = np.full((Hh + 2, W + 2), -1)  # one-cell border all round
        for i, inst in enumerate(insts):
            x0, y0 = L.origin[i]
            assert 0 <= x0 and x0 + inst.nx <= W and 0 <= y0 and y0 + inst.ny <= Hh
            blk = owner[y0 + 1:y0 + 1 + inst.ny, x0 + 1:x0 + 1 + inst.nx]
            assert (blk == -1).all(), "overlap"
            blk[:] = i
        # no two instances touch, also diagonally
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                sh = np.roll(np.roll(owner, dy, 0), dx, 1)
                both = (owner >= 0) & (sh >= 0)
                assert (owner[both] == sh[both]).all()
        act = L.active()
        assert act.sum() == sum(i.n for i in insts)
        cells = np.concatenate([L.cells(i) for i in range(len(insts))])
        assert len(np.unique(cells)) == len(cells)
        assert 0 < L.efficiency() <= 1


def test_layout_fields():
    from monoalg3d_c_amd.batch import Layout
    insts = _insts()
    L = Layout(insts)
    m, s = L.stim_mask(100.0), L.ionic_scale()
    ids, tab = L.instance_ids(), L.scale_table()
    assert ids.dtype == np.uint16 and tab.shape == (len(insts), 4) and tab.dtype == np.float32
    for i, inst in enumerate(insts):
        c = L.cells(i).reshape(inst.ny, inst.nx)
        lx = np.arange(inst.nx)
        want = ((lx + 0.5) * 100.0 < inst.stim_x_limit).astype(np.uint8)
        assert (m[c] == want[None, :]).all()
        assert (s[c.reshape(-1)] == np.asarray(inst.scale, np.float32)).all()
        assert (ids[c.reshape(-1)] == i).all()
        assert (tab[i, :len(inst.scale)] == np.asarray(inst.scale, np.float32)).all()
    gaps = L.active() == 0
    assert (m[gaps] == 0).all()
    v = np.arange(L.n, dtype=float)
    for i, part in enumerate(L.split(v)):
        assert (part == L.cells(i)).all()


def test_layout_rejects_bad_input():
    from monoalg3d_c_amd.batch import Instance, Layout
    with pytest.raises(ValueError):
        Layout([])
    with pytest.raises(ValueError):
        Layout([Instance(0, 3)])
    with pytest.raises(ValueError):
        Layout([Instance(100, 1)], width=50)
    with pytest.raises(ValueError):
        Layout([Instance(10, 1)], gap=0)


def test_sweep_is_seeded():
    from monoalg3d_c_amd.batch import sweep_instances
    a, b = sweep_instances(256, 7), sweep_instances(256, 7)
    assert a == b and sweep_instances(256, 8) != a
    cab = [i for i in a if i.ny == 1]
    sh = [i for i in a if i.ny > 1]
    assert all(100 <= i.nx <= 2000 for i in cab) and all(32 <= i.nx == i.ny <= 256 for i in sh)
    f = np.array([i.scale[:3] for i in a])
    assert f.min() >= 0.5 and f.max() <= 1.5


def _oracle_run(oracle, shape, active, stim_mask, scale, steps):
    n = shape[0] * shape[1] * shape[2]
    a = oracle.alpha(0.14, 1.0, 0.02, *H)
    coef = oracle.assemble_iso(shape, H, a, np.full((n, 3), 0.00013), active)
    t = oracle.Tissue(shape, H, oracle.MODEL_IDS["tt2006"], coef, active=active, stim_mask=stim_mask,
                      ionic_scale=scale)
    t.step(steps)
    return t.v.copy()


def test_atlas_equals_standalone_runs_on_oracle(oracle):
    """Every instance of the packed atlas equals its own run (30 ms, TT2006,
    per-instance scalings and stimulus regions), to the CG tolerance: each
    solve may differ by |r|/alpha <= 1e-8/0.07 mV, over 1500 steps < 1e-5 mV."""
    from monoalg3d_c_amd.batch import Layout
    insts = _insts()
    L = Layout(insts)
    steps = 1500
    va = _oracle_run(oracle, L.shape, L.active(), L.stim_mask(100.0), L.ionic_scale(), steps)
    for i, inst in enumerate(insts):
        shape = (inst.nx, inst.ny, 1)
        sm = ((np.arange(inst.nx) + 0.5) * 100.0 < inst.stim_x_limit).astype(np.uint8)
        vs = _oracle_run(oracle, shape, None, np.tile(sm, inst.ny), np.tile(np.asarray(inst.scale, np.float32),
                                                                               (inst.n, 1)), steps)
        np.testing.assert_allclose(va[L.cells(i)], vs, rtol=0, atol=1e-5, err_msg=f"instance {i}")
    # the scalings matter: instance 0 with reference parameters is a different run
    v0 = _oracle_run(oracle, (30, 1, 1), None, None, None, steps)
    assert np.abs(v0 - va[L.cells(0)]).max() > 0.1
