"""Batched runs on the GPU: many instances packed into one atlas.

Each instance of a batch must equal its own run: the reference's own cable and
sheet runs (golden fixtures written by MonoAlg3D built from its sources) for
instances with the reference parameters, the oracle's packed run for scaled
instances, and the GPU's own stand-alone runs.  A batch's CG stops on the
atlas residual, a stand-alone run on its own, so each step's solves may differ
by |r|/alpha <= 1e-8/0.07 = 1.4e-7 mV; over the 1,000-3,000 steps of these runs
that allows up to 4e-4 mV, and FP64 runs agree within VM_TOL = 1e-4 mV (the
wave front of a 60 ms cable run shows 2e-5 mV) with activation times within
1e-3 ms.  FP32 + refinement keeps activation times within 0.5 ms of the FP64
batch.
"""
import os
import sys

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_golden import lat_from  # noqa: E402

H = (100.0, 100.0, 100.0)
VM_TOL = 1e-4


def _mixed():
    from monoalg3d_c_amd.batch import Instance
    return [Instance(300, 1), Instance(48, 32), Instance(150, 1, (1.3, 0.7, 1.2, 1.0)),
            Instance(20, 20, (0.6, 1.4, 0.8, 1.0), stim_x_limit=300.0), Instance(64, 1, (0.5, 1.5, 1.5, 1.0)),
            Instance(9, 40)]


def _snapshots(b, every, nsamp):
    out = [b.v()]
    for _ in range(nsamp - 1):
        b.step(every)
        out.append(b.v())
    return [np.array([s[i] for s in out]) for i in range(len(out[0]))]


def test_batch_instances_match_reference_runs():
    """A 300-cell cable and a 48x32 sheet with the reference parameters, packed
    with four scaled instances, against MonoAlg3D's stand-alone runs of each."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.batch import Batch
    announce("batch fp64 vs reference cable_tt2006_300 + sheet_tt2006_48x32, 60 ms")
    b = Batch(_mixed(), precision=m.FP64, cg_tol=1e-16, tensor=m.iso_tensor(0.00013))
    assert b.layout.shape[2] == 1
    snaps = _snapshots(b, 50, 61)
    st = b.stats
    b.close()
    assert st["nan_detected"] == 0
    for i, name in ((0, "tissue_cable_tt2006_300"), (1, "tissue_sheet_tt2006_48x32")):
        g = np.load(os.path.join(GOLD, name + ".npz"))
        ns = len(g["probe_vm"])
        s = snaps[i][:ns]
        dv = np.abs(s[:, g["probe_cells"]] - g["probe_vm"]).max()
        assert dv < VM_TOL, (name, dv)
        for key in g.files:
            if key.startswith("vm_t"):
                k = int(round(int(key[4:]) / (50 * 0.02)))
                assert np.abs(s[k] - g[key]).max() < VM_TOL, (name, key)
        lat = lat_from(s, 50)
        assert np.array_equal(np.isfinite(lat), np.isfinite(g["lat"])), name
        assert np.abs(lat - g["lat"]).max() <= 1e-3, name


def test_batch_matches_oracle_atlas(oracle):
    """Scaled instances (per-instance GKr/GCaL/GKs and stimulus region): the
    GPU atlas against the oracle's run of the same atlas, every cell, 30 ms."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.batch import Batch, Layout
    insts = _mixed()[2:]
    L = Layout(insts)
    announce(f"batch fp64 vs oracle atlas {L.shape}, 30 ms")
    n = L.n
    a = oracle.alpha(0.14, 1.0, 0.02, *H)
    coef = oracle.assemble_iso(L.shape, H, a, np.full((n, 3), 0.00013), L.active())
    t = oracle.Tissue(L.shape, H, oracle.MODEL_IDS["tt2006"], coef, active=L.active(), stim_mask=L.stim_mask(100.0),
                      ionic_scale=L.ionic_scale())
    b = Batch(insts, precision=m.FP64, cg_tol=1e-16, tensor=m.iso_tensor(0.00013))
    for _ in range(3):
        t.step(500)
        b.step(500)
        got = b.v()
        for i, want in enumerate(L.split(t.v)):
            np.testing.assert_allclose(got[i], want, rtol=0, atol=VM_TOL, err_msg=f"instance {i}")
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_instance_table_equals_per_cell_scalings(precision):
    """The batch's LDS-staged instance table gives the same states, bit for
    bit, as the same atlas run with the scalings stored per cell."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.batch import Batch
    insts = _mixed()
    prec = m.FP64 if precision == "fp64" else m.FP32_REFINE
    announce(f"batch instance table vs per-cell scalings, {precision}, 10 ms")
    b = Batch(insts, precision=prec, cg_tol=1e-16, tensor=m.iso_tensor(0.00013))
    L = b.layout
    t = m.Tissue(L.shape, h=H, precision=prec, cg_tol=1e-16, tensor=m.iso_tensor(0.00013), active=L.active(),
                 stim_mask=L.stim_mask(100.0), ionic_scale=L.ionic_scale())
    b.step(500)
    t.step(500)
    act = L.active() == 1
    assert np.array_equal(b.tissue.state[1:, act], t.state[1:, act])
    assert np.array_equal(b.tissue.v[act], t.v[act])
    b.close()
    t.close()


def test_batch_instance_equals_standalone_gpu_run():
    """Each scaled instance of the batch against a stand-alone GPU tissue of
    its own shape with its scalings on every cell (FP64, 20 ms)."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.batch import Batch
    insts = _mixed()
    announce("batch fp64 vs standalone gpu runs, 20 ms")
    b = Batch(insts, precision=m.FP64, cg_tol=1e-16, tensor=m.iso_tensor(0.00013))
    b.step(1000)
    got = b.v()
    b.close()
    for i, inst in enumerate(insts):
        sm = ((np.arange(inst.nx) + 0.5) * 100.0 < inst.stim_x_limit).astype(np.uint8)
        t = m.Tissue((inst.nx, inst.ny, 1), precision=m.FP64, cg_tol=1e-16, tensor=m.iso_tensor(0.00013),
                     stim_mask=np.tile(sm, inst.ny), ionic_scale=np.tile(np.asarray(inst.scale, np.float32), (inst.n, 1)))
        t.step(1000)
        np.testing.assert_allclose(got[i], t.v, rtol=0, atol=VM_TOL, err_msg=f"instance {i}")
        t.close()


def test_batch_fp32_refine_activation_times():
    """FP32 states + FP32 CG with FP64 defect correction: every instance's
    activation times within 0.5 ms of the FP64 batch, same cells activated."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.batch import Batch
    announce("batch fp32+refine vs fp64 activation times, 60 ms")
    res = {}
    for prec in (m.FP64, m.FP32_REFINE):
        b = Batch(_mixed(), precision=prec, cg_tol=1e-16, tensor=m.iso_tensor(0.00013))
        res[prec] = _snapshots(b, 10, 301)
        assert b.stats["nan_detected"] == 0
        b.close()
    for i in range(len(_mixed())):
        l64, l32 = lat_from(res[m.FP64][i], 10), lat_from(res[m.FP32_REFINE][i], 10)
        assert np.isfinite(l64).any()
        assert np.array_equal(np.isfinite(l64), np.isfinite(l32)), i
        fin = np.isfinite(l64)
        assert np.abs(l64[fin] - l32[fin]).max() <= 0.5, i


def test_sweep_file_runs_match_standalone_runs(tmp_path):
    """A sweep file in the reference's [batch]/[modify] format over cable
    length and the GKr factor: the four runs go into one atlas, and each run's
    V_it files equal a stand-alone run of that cable (FP64, 5 ms)."""
    import monoalg3d_c_amd as m
    from make_golden import case_files, read_snapshots
    from monoalg3d_c_amd.sweep import plan, run_sweep
    announce("sweep file: 2 cable lengths x 2 GKr factors, one atlas")
    ini, _, _ = case_files(str(tmp_path / "base"), (300, 1, 1), "tt2006", 5.0, 50, cable=True)
    b = tmp_path / "sweep.ini"
    b.write_text("[batch]\ninitial_config=" + ini + "\noutput_folder=" + str(tmp_path / "out") +
                 "\nnum_simulations_per_parameter_change=1\n[modify]\n"
                 "domain|cable_length=list|10000|15000\nextra_data|GKr_multiplicator=list|0.5|1.5\n")
    assert len(plan(str(b))) == 1
    done = run_sweep(str(b))
    assert [d[0] for d in done] == [1, 2, 3, 4]
    for (num_run, out), (length, gkr) in zip(done, [(100, 0.5), (100, 1.5), (150, 0.5), (150, 1.5)]):
        snaps = read_snapshots(out, (length, 1, 1), 50, 251)
        t = m.Tissue((length, 1, 1), precision=m.FP64, cg_tol=1e-16, tensor=m.iso_tensor(0.00013),
                     ionic_scale=np.tile(np.array([gkr, 1, 1, 1], np.float32), (length, 1)))
        want = [t.v]
        for _ in range(len(snaps) - 1):
            t.step(50)
            want.append(t.v)
        t.close()
        np.testing.assert_allclose(snaps, np.array(want), rtol=0, atol=VM_TOL, err_msg=f"run {num_run}")


def test_sweep_matches_reference_runs(tmp_path):
    """The reference's own batch sweep: a [batch] file over cable length x
    stimulus x_limit (tests/golden/make_golden.py sweep_batch_files), run by
    run_sweep as one atlas, against MonoAlg3D run once per expanded
    configuration (what each MPI rank of main_batch runs): every run's
    V_it files within VM_TOL of that run's reference Vm, activation times
    within 1e-3 ms, run numbers and output directories as main_batch names them."""
    from make_golden import SWEEP_LENGTHS, SWEEP_XLIMITS, read_snapshots, sweep_batch_files
    from monoalg3d_c_amd.sweep import plan, run_sweep
    announce("sweep file vs the reference's per-run outputs: 3 cable lengths x 2 stimulus x limits")
    g = np.load(os.path.join(GOLD, "sweep_cable_tt2006.npz"))
    batch, ini = sweep_batch_files(str(tmp_path))
    groups = plan(batch)
    assert len(groups) == 1 and len(groups[0]) == 6  # one atlas for the whole sweep
    done = run_sweep(batch)
    assert [d[0] for d in done] == [1, 2, 3, 4, 5, 6]
    k = 0
    for length in SWEEP_LENGTHS:
        for xl in SWEEP_XLIMITS:
            num, out = done[k]
            k += 1
            assert out.endswith(f"_run_{num}_cable_length_{length}_x_limit_{xl}"), out
            n = int(round(float(length) / 100.0))
            snaps = read_snapshots(out, (n, 1, 1), 50, int(round(40.0 / 0.02)) + 1)
            ref = g[f"run{num}_vm"]
            assert snaps.shape == ref.shape, (num, snaps.shape, ref.shape)
            dv = np.abs(snaps - ref).max()
            assert dv < VM_TOL, (num, dv)
            lat = lat_from(snaps, 50)
            assert np.array_equal(np.isfinite(lat), np.isfinite(g[f"run{num}_lat"])), num
            assert np.abs(lat - g[f"run{num}_lat"]).max() <= 1e-3, num


def test_batch_relative_tolerance_bounds_every_instance():
    """cg_rel_tol on a batch: the atlas stops on r.r <= rel^2 min_i ||b_i||^2
    (each instance's ||b_i||^2 taken on the device after every ionic step), so
    a 100-cell cable packed beside a 192x192 sheet is solved at least as
    tightly as alone with the relative test.  After every step of 10 ms the
    atlas residual r.r is below rel^2 ||b_cable||^2; with b = A v and
    ||A|| <= alpha + 4 * 0.013 on the cable (face coefficient sigma * area /
    distance = 0.00013 * 100), ||b_cable||^2 <= 3.04 ||alpha v_cable||^2, the
    bound checked here.  Both instances match their stand-alone runs with the
    relative test within VM_TOL."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.batch import Batch, Instance
    announce("batch with cg_rel_tol: small cable beside a large sheet, 10 ms")
    insts = [Instance(100, 1), Instance(192, 192)]
    rel = 1e-10
    a = m.alpha()
    b = Batch(insts, precision=m.FP64, cg_rel_tol=rel, tensor=m.iso_tensor(0.00013))
    for _ in range(500):
        b.step(1)
        cable = b.v()[0]
        assert b.stats["last_cg_error"] <= rel * rel * 3.04 * float(((a * cable) ** 2).sum())
    got = b.v()
    b.close()
    for i, inst in enumerate(insts):
        t = m.Tissue((inst.nx, inst.ny, 1), precision=m.FP64, cg_rel_tol=rel, cg_tol=0.0, tensor=m.iso_tensor(0.00013))
        t.step(500)
        dv = np.abs(got[i] - t.v).max()
        t.close()
        assert dv < VM_TOL, (i, dv)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config4_256_instances_500ms(precision):
    """Config 4 at its stated size: the seed-4 sweep of 256 instances (cables
    of 100-2000 cells, square sheets of 32-256 cells a side, TT2006 with GKr,
    GCaL, GKs x U[0.5, 1.5]), 1 Hz pacing, 500 ms, CG relative residual 1e-10
    per instance.  No NaN; every instance activates; per-instance activation
    times of FP32+refine within 0.5 ms of the FP64 batch (the FP64 batch is the
    reference for the scaled instances; scale-1 instances are pinned to the
    reference's runs above)."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.batch import Batch, sweep_instances
    insts = sweep_instances(256, 4)
    announce(f"config 4: 256 instances, 500 ms, {precision} (+ fp64 for the comparison)")
    out = {}
    for key, prec in (("fp64", m.FP64), ("fp32", m.FP32_REFINE)):
        if key != "fp64" and precision == "fp64":
            continue
        b = Batch(insts, precision=prec, cg_rel_tol=1e-10, cg_max_its=500, tensor=m.iso_tensor(0.00013),
                  stim=(0.0, 2.0, -50.0, 500.0, 1000.0))
        b.observe_setup(activation_threshold=-30.0, apd_threshold=-83.0, time_threshold=10.0, max_events=2)
        for k in range(25001):
            if k % 5 == 0:
                b.observe()
            if b.step(1) == 1:
                break
        st = b.stats
        acts = b.activations()
        b.close()
        assert st["nan_detected"] == 0, key
        out[key] = acts
    for i, a in enumerate(out["fp64"]):
        assert (a["n_act"] >= 1).all(), f"instance {i} did not activate everywhere"
    if precision == "fp32":
        worst = 0.0
        for a64, a32 in zip(out["fp64"], out["fp32"]):
            np.testing.assert_array_equal(a64["n_act"], a32["n_act"])
            worst = max(worst, float(np.abs(a64["act"][:, 0] - a32["act"][:, 0]).max()))
        print(f"[config4] max activation-time difference fp32+refine vs fp64: {worst:.4f} ms", flush=True)
        assert worst <= 0.5, worst
