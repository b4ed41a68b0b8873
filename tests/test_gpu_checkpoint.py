"""Checkpoint and restart on the GPU (mono_tissue_save_state / mono_tissue_restore_state).

Against the reference: the three files save_simulation_state wrote at the end
of a 6 ms run (10x8x2 TT2006 box, stimulus every 1.5 ms) restart a tissue
here, which must then repeat the reference's own restart to 12 ms; and the
files written here after the same 302 steps must hold what the reference's
hold (time and iteration exactly, potentials and states within the
GPU-vs-reference agreement of these runs).  The reference's restart and its
uninterrupted run differ by up to 2.3 mV (the stimulus start is not part of its
checkpoint), so the first comparison pins that behaviour too.

Against itself: a run saved and restarted in a new tissue repeats the
uninterrupted run bit for bit, in FP64 and in FP32 + refinement, with a
periodic stimulus, the activation observer, and the adaptive ODE solver.
"""
import os
import sys

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
H = (100.0, 100.0, 100.0)
# GPU vs reference over 300 steps of these runs: ~1e-8 mV where the wave passes
VM_TOL = 1e-5


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLD, "checkpoint_tt2006_10x8x2.npz"))


def _write_files(fx, d):
    os.makedirs(d, exist_ok=True)
    for name in ("grid", "monodomain_solver", "ode_solver"):
        fx[f"file_{name}"].tofile(os.path.join(d, f"{name}_checkpoint.dat"))


def _tissue(fx, **kw):
    import monoalg3d_c_amd as m
    shape = tuple(int(x) for x in fx["shape"])
    args = dict(precision=m.FP64, cg_tol=1e-16, tensor=m.iso_tensor(0.00013), stim=tuple(fx["stim"]))
    args.update(kw)
    return m.Tissue(shape, **args)


def _run_to(t, k, iters):
    out = []
    for target in iters:
        t.step(target - k)
        k = target
        out.append(t.v)
    return np.array(out)


def test_restart_from_reference_checkpoint(fx, tmp_path):
    announce("restore the reference's checkpoint, run to 12 ms, FP64")
    d = str(tmp_path / "ck")
    _write_files(fx, d)
    t = _tissue(fx)
    t.restore_state(d)
    st = t.stats
    assert st["steps"] == 302
    got = _run_to(t, 302, fx["iters"])
    t.close()
    assert np.abs(got - fx["restart_vm"]).max() < VM_TOL
    assert np.abs(got - fx["continuous_vm"]).max() > 1.0


def test_saved_files_match_reference_checkpoint(fx, tmp_path, oracle):
    """Our files after 302 steps, read by the oracle's reader of the
    reference's layout, against the reference's files."""
    announce("save after 302 steps vs the reference's checkpoint files")
    shape = tuple(int(x) for x in fx["shape"])
    t = _tissue(fx)
    t.step(302)
    d = str(tmp_path / "ours")
    t.save_state(d)
    t.close()
    read = lambda files: oracle.read_reference_checkpoint(*files, shape, H, 19)
    ref = read([fx[f"file_{k}"].tobytes() for k in ("grid", "monodomain_solver", "ode_solver")])
    ours = read([open(os.path.join(d, f"{k}_checkpoint.dat"), "rb").read()
                 for k in ("grid", "monodomain_solver", "ode_solver")])
    assert ours["time"] == ref["time"] and ours["iteration"] == ref["iteration"] == 302
    assert ours["gpu"] and not ours["adaptive"] and ours["aux"] is None
    assert np.abs(ours["v"] - ref["v"]).max() < VM_TOL
    # state 0 (Vm) is carried in the potential here; the gates and concentrations
    # within 1e-6 of each state's range over the cells
    s_ref, s_ours = ref["sv"][:, 1:], ours["sv"][:, 1:]
    scale = np.maximum(np.abs(s_ref).max(axis=0), 1e-12)
    assert (np.abs(s_ours - s_ref).max(axis=0) / scale).max() < 1e-6
    # the monodomain solver header: beta, cm and dt where the reference puts them
    sol = open(os.path.join(d, "monodomain_solver_checkpoint.dat"), "rb").read()
    assert np.frombuffer(sol[16:32], np.float64).tolist() == [0.14, 1.0]
    assert np.frombuffer(sol[80:88], np.float64)[0] == 0.02
    assert sol[104:112] == fx["file_monodomain_solver"].tobytes()[104:112]  # time_info.dt


@pytest.mark.parametrize("prec_name", ["FP64", "FP32_REFINE"])
def test_self_restart_is_bitwise(fx, tmp_path, prec_name):
    """Save at step 302, restore into a new tissue, run 298 more steps: Vm,
    states, step count, time and activation times equal the uninterrupted run
    bit for bit (the stimulus start after its shifts, the solver's sweep
    schedule and the observer travel in hip_solver_checkpoint.dat)."""
    import monoalg3d_c_amd as m
    prec = getattr(m, prec_name)
    announce(f"self restart bitwise, {prec_name}")
    d = str(tmp_path / "self")

    def observed(t, k0, k1):
        for k in range(k0, k1):
            if k % 10 == 0:
                t.observe()
            t.step(1)

    a = _tissue(fx, precision=prec)
    a.observe_setup(max_events=4)
    observed(a, 0, 302)
    a.save_state(d)
    observed(a, 302, 600)
    b = _tissue(fx, precision=prec)
    b.restore_state(d)
    observed(b, 302, 600)
    sa, sb = a.stats, b.stats
    assert sa["steps"] == sb["steps"] == 600 and sa["time"] == sb["time"]
    np.testing.assert_array_equal(a.v, b.v)
    np.testing.assert_array_equal(a.state, b.state)
    aa, ab = a.activations(), b.activations()
    assert aa["n_act"].max() >= 1
    for k in aa:
        np.testing.assert_array_equal(aa[k], ab[k])
    a.close()
    b.close()


def test_adaptive_self_restart_is_bitwise(tmp_path):
    """CRN 1998 with the adaptive Euler solver on a 40-cell cable: the per-cell
    step sizes and times (the reference's three extra state rows) travel with
    the states."""
    import monoalg3d_c_amd as m
    announce("adaptive CRN self restart bitwise")
    kw = dict(h=(100.0, 100.0, 100.0), model=m.MODEL_CRN1998, precision=m.FP64, dt=0.05, dt_ode=0.001,
              tensor=m.iso_tensor(0.0000176), cg_tol=1e-16, stim=(0.0, 2.0, -2000.0, 500.0, 0.0), adaptive=(1e-5, 1e-5))
    d = str(tmp_path / "crn")
    a = m.Tissue((40, 1, 1), **kw)
    a.step(200)
    a.save_state(d)
    a.step(300)
    b = m.Tissue((40, 1, 1), **kw)
    b.restore_state(d)
    b.step(300)
    np.testing.assert_array_equal(a.v, b.v)
    np.testing.assert_array_equal(a.state, b.state)
    assert a.stats["time"] == b.stats["time"]
    a.close()
    b.close()


def test_restore_rejects_mismatch(fx, tmp_path):
    import monoalg3d_c_amd as m
    announce("restore into mismatched tissues")
    d = str(tmp_path / "ck")
    _write_files(fx, d)
    for kw, msg in ((dict(model=m.MODEL_MS2003), "states per cell"), (dict(dt=0.01), "dt_pde")):
        t = _tissue(fx, **kw)
        with pytest.raises(RuntimeError, match=msg):
            t.restore_state(d)
        t.close()
    t = m.Tissue((10, 8, 3), precision=m.FP64, tensor=m.iso_tensor(0.00013))
    with pytest.raises(RuntimeError, match="not in the checkpoint"):
        t.restore_state(d)
    t.close()
    with pytest.raises(RuntimeError, match="cannot open"):
        t = _tissue(fx)
        try:
            t.restore_state(str(tmp_path / "missing"))
        finally:
            t.close()


def test_ini_restart_from_reference_checkpoint(fx, tmp_path):
    """The reference's own configuration with [restore_state] input_dir
    pointing at its checkpoint: the V_it files of the run equal the reference's
    restart."""
    from monoalg3d_c_amd.inputs import run
    sys.path.insert(0, GOLD)
    from make_golden import checkpoint_case, read_iters
    announce("ini [restore_state] from the reference's checkpoint")
    d = str(tmp_path / "ck")
    _write_files(fx, d)
    out = str(tmp_path / "run")
    ini = checkpoint_case(out, 12.0, f"[restore_state]\nmain_function=restore_simulation_state\ninput_dir={d}\n"
                                     f"[save_state]\nmain_function=save_simulation_state\noutput_dir={out}/ck2\n")
    run(ini)
    got = read_iters(out, tuple(int(x) for x in fx["shape"]), fx["iters"])
    assert np.abs(got - fx["restart_vm"]).max() < VM_TOL
    assert not os.path.exists(os.path.join(out, "V_it_300.bin"))
    assert os.path.exists(os.path.join(out, "ck2", "hip_solver_checkpoint.dat"))


_CROSS_PROCESS = r"""
import sys
import numpy as np
sys.path.insert(0, {root!r})
import monoalg3d_c_amd as m
shape = (96, 80, 12)
t = m.Tissue(shape, model=m.MODEL_ORD2011, precision=getattr(m, {prec!r}), aniso=True,
             tensor=m.layered_fibre_tensors(12, -60.0, 60.0, 0.00024, 0.000045), tensor_source=m.TENSOR_PER_LAYER,
             cg_rel_tol=1e-10, cg_tol=0.0, cg_max_its=500)
mode = {mode!r}
if mode == "restore":
    t.restore_state({ck!r})
t.step({steps})
if mode == "save":
    t.save_state({ck!r})
np.save({out!r}, t.v)
t.close()
"""


@pytest.mark.parametrize("prec", ["FP64", "FP32_REFINE"])
def test_restart_in_another_process_is_bitwise(prec, tmp_path):
    """Three processes: one runs 400 steps uninterrupted; one runs 200 and
    saves; a third restores and runs 200 more.  The plane-loop kernels' work
    split and partial-sum order are a function of the box and the CU count
    only (not of how many blocks a process measured resident), so the
    restarted run repeats the uninterrupted one bit for bit across processes.
    The box (96x80x12, ORd, rotating fibres) uses the 3-D tiled kernels."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    announce(f"cross-process restart, 96x80x12 ORd {prec}")
    ck = str(tmp_path / "ck")
    os.makedirs(ck)
    runs = {"full": ("none", 400), "first": ("save", 200), "second": ("restore", 200)}
    for name, (mode, steps) in runs.items():
        code = _CROSS_PROCESS.format(root=root, prec=prec, mode=mode, ck=ck, steps=steps, out=str(tmp_path / f"{name}.npy"))
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (name, r.stderr[-2000:])
    full, second = np.load(str(tmp_path / "full.npy")), np.load(str(tmp_path / "second.npy"))
    np.testing.assert_array_equal(second, full)
