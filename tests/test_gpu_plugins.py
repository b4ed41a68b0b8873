"""The plugins over libmonoalg_hip.so inside the reference's own MonoAlg3D.

The reference binary (built from its sources by oracle/ref.mk) loads every
plugin kind of the monodomain step through its dlopen ABI: the ionic model
([ode_solver]), ASSEMBLY_MATRIX ([assembly_matrix]), UPDATE_MONODOMAIN
([update_monodomain]) and SOLVE_LINEAR_SYSTEM ([linear_system_solver]).

  - Assembly and update: with one OpenMP thread (the reference's assembly
    inserts elements in thread order) a run whose assembly and update come
    from the plugins writes the same Vm files as the all-reference run (both
    with the reference's CPU CG and ionic model): bit for bit for isotropic
    rows, to the rounding of the CG's row sums for anisotropic rows, whose
    elements are listed in another order.
  - All four kinds at config 1 (64x64x16 TT2006, 50 ms) against the
    reference's own run, to the tolerances of the FP64 tissue tests.
  - The GPU half of the ODE ABI (set_model_initial_conditions_gpu /
    solve_model_odes_gpu, driven by plugins/gpu_abi_driver as the core's
    use_gpu path calls it) gives the same states, bit for bit, as the CPU half
    of the same plugin.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLD)
from make_golden import DT, case_files, lat_from, read_snapshots  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")
REF_EXE = os.path.join(REF_DIR, "bin", "MonoAlg3D")
HARNESS = os.path.join(REF_DIR, "bin", "ref_harness")
PLUG = os.path.join(ROOT, "monoalg3d_c_amd", "plugins")
MONO = os.path.join(PLUG, "libmono_hip_monodomain.so")
DRIVER = os.path.join(PLUG, "gpu_abi_driver")


def _need(*paths):
    missing = [p for p in paths if not os.path.exists(p)]
    if missing:
        pytest.skip(f"not built: {missing}")


def _run_ref(ini, timeout=600):
    r = subprocess.run([REF_EXE, "-c", ini], cwd=REF_DIR, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def _resolution_seconds(log):
    m = re.search(r"Resolution Time:\s*([0-9.eE+-]+)\s*μs", log)
    return float(m.group(1)) * 1e-6 if m else None


def _swap(ini, ode=None, assembly=None, update=False, cg=False, threads=None):
    """Select plugins over libmonoalg_hip.so in a reference .ini."""
    text = open(ini).read()
    if threads is not None:
        text = re.sub(r"num_threads = \d+", f"num_threads = {threads}", text)
    if ode:
        text = re.sub(r"library_file=shared_libs/\S+", f"library_file={os.path.join(PLUG, ode)}", text, count=1)
    if assembly:
        text = text.replace("init_function=set_initial_conditions_fvm",
                            f"library_file={MONO}\ninit_function=mono_set_initial_conditions_fvm")
        text = text.replace(f"main_function={assembly}_assembly_matrix", f"main_function=mono_{assembly}_assembly_matrix")
    if update:
        text = text.replace("main_function=update_monodomain_default",
                            f"library_file={MONO}\nmain_function=mono_update_monodomain")
    if cg:
        text = text.replace("main_function=conjugate_gradient", f"library_file={MONO}\nmain_function=mono_conjugate_gradient")
    out = ini.replace(".ini", "_plug.ini")
    with open(out, "w") as fh:
        fh.write(text)
    return out


def _outputs(out_dir, every, nsteps):
    return [np.fromfile(os.path.join(out_dir, f"V_it_{k}.bin")) for k in range(0, nsteps, every)]


@pytest.mark.parametrize("case", ["iso", "fibrotic"])
def test_assembly_and_update_plugins_are_bit_exact(case, tmp_path):
    """One thread, reference ionic model and CPU CG on both sides: the run
    whose matrix comes from mono_{homogeneous,anisotropic}_sigma_assembly_matrix
    (rows assembled on the device, written back as element lists in the
    reference's insertion order) and whose b comes from mono_update_monodomain
    (alpha * Vm on the device) writes the same V_it_<k>.bin files as the
    all-reference run: bit for bit for the isotropic rows, within 1e-9 mV for
    the anisotropic ones (see below)."""
    _need(REF_EXE, MONO)
    if case == "iso":
        shape, model, T, every, kw, asm = (16, 16, 4), "tt2006", 10.0, 50, {}, "homogeneous_sigma"
    else:
        shape, model, T, every, kw, asm = (24, 24, 6), "ord2011", 6.0, 50, dict(aniso=True, fibrosis=0.1), "anisotropic_sigma"
    announce(f"reference core, assembly+update plugins bit-exact, {case} {shape}")
    nsteps = int(round(T / DT)) + 1
    ref_dir, plug_dir = str(tmp_path / "ref"), str(tmp_path / "plug")
    ini_ref, act, _ = case_files(ref_dir, shape, model, T, every, **kw)
    ini_ref = _swap(ini_ref, threads=1)
    ini_plug, _, _ = case_files(plug_dir, shape, model, T, every, **kw)
    ini_plug = _swap(ini_plug, assembly=asm, update=True, threads=1)
    _run_ref(ini_ref)
    _run_ref(ini_plug)
    a, b = _outputs(ref_dir, every, nsteps), _outputs(plug_dir, every, nsteps)
    assert len(a) == len(b) and len(a) == (nsteps - 1) // every + 1
    worst = 0.0
    for k, (x, y) in enumerate(zip(a, b)):
        assert x.shape == y.shape
        if case == "iso":
            assert np.array_equal(x, y), f"snapshot {k}: max |dV| = {np.abs(x - y).max()}"
        else:
            # element values are the reference's bit for bit (test_gpu_operator,
            # oracle pinned on the reference's rows with holes), but an
            # anisotropic row lists its elements in direction order, while the
            # reference's list follows the first non-zero update of each
            # neighbour in its direction loop (assembly_common.c:541-825); the
            # CPU CG sums a row's products in list order, so the runs agree to
            # rounding of those sums
            np.testing.assert_allclose(y, x, rtol=0, atol=1e-9, err_msg=f"snapshot {k}")
        worst = max(worst, float(np.abs(x - y).max()))
    print(f"[plugins] {case}: max |dV| over {len(a)} snapshots = {worst:.3e} mV", flush=True)
    # the run is not trivial: at some snapshot the front is inside the tissue
    vs = [s.reshape(-1, 7)[:, 6] for s in a]
    assert any((v > -30.0).any() and (v < -70.0).any() for v in vs)


def test_reference_core_with_all_plugin_kinds_config1(tmp_path):
    """Config 1 in full (64x64x16 TT2006, isotropic 0.13 S/m, 50 ms) in the
    reference's MonoAlg3D with the ionic model, assembly, update and CG all
    from the plugins: Vm of the probe cells within 1e-6 mV (1e-8 of the 100 mV
    scale of Vm) and activation times
    within 1e-3 ms of the reference's own run; the drop-in rate is printed."""
    _need(REF_EXE, MONO, os.path.join(PLUG, "libmono_hip_tt2006.so"))
    announce("reference core + all four plugin kinds, config 1 64x64x16 tt2006 50 ms")
    g = np.load(os.path.join(GOLD, "tissue_config1_tt2006_64x64x16.npz"))
    shape = tuple(int(s) for s in g["shape"])
    every = int(g["every"])
    out = str(tmp_path / "run")
    ini, _, _ = case_files(out, shape, "tt2006", 50.0, every)
    ini = _swap(ini, ode="libmono_hip_tt2006.so", assembly="homogeneous_sigma", update=True, cg=True)
    log = _run_ref(ini, timeout=900)
    nsteps = int(round(50.0 / DT)) + 1
    snaps = read_snapshots(out, shape, every, nsteps)
    probes = g["probe_cells"]
    dv = np.abs(snaps[:, probes] - g["probe_vm"]).max()
    lat = lat_from(snaps, every)
    both = np.isfinite(g["lat"])
    assert np.array_equal(np.isfinite(lat), both)
    dlat = np.abs(lat[both] - g["lat"][both]).max()
    res = _resolution_seconds(log)
    n = int(np.prod(shape))
    print(f"[drop-in] config 1 through the reference core with all four plugin kinds: max |dV| = {dv:.3e} mV, "
          f"max |dLAT| = {dlat:.3e} ms, resolution time {res} s, "
          f"{(n * 2501 / res) if res else float('nan'):.3e} cell-steps/s", flush=True)
    assert dv < 1e-6, dv
    assert dlat <= 1e-3, dlat


@pytest.mark.parametrize("model,lib,amp", [("tt2006", "libmono_hip_tt2006.so", -50.0),
                                           ("ord2011", "libmono_hip_ord2011.so", -50.0),
                                           ("ms2003", "libmono_hip_ms2003.so", 0.2)])
def test_gpu_ode_abi_equals_cpu_abi(model, lib, amp, tmp_path):
    """set_model_initial_conditions_gpu + solve_model_odes_gpu (device SoA
    state with pitch, V in row 0) against set_model_initial_conditions_cpu +
    solve_model_odes_cpu of the same plugin over 600 steps of 64 cells with a
    per-cell stimulus for the first 100: the same states bit for bit, and the
    reference's own model within 1e-9 of each state's range (the FP64 trace
    bound of the ionic tests)."""
    _need(DRIVER, HARNESS, os.path.join(PLUG, lib))
    announce(f"GPU ODE ABI vs CPU ODE ABI, {model}")
    n, nsteps, rec = 64, 600, 100
    stim = np.zeros((nsteps, n))
    stim[:100, ::2] = amp
    stim[:100, 1::2] = 0.5 * amp
    sp = str(tmp_path / "stim.bin")
    stim.tofile(sp)
    gpu_out, cpu_out, ref_out = (str(tmp_path / f"{k}.bin") for k in ("gpu", "cpu", "ref"))
    r = subprocess.run([DRIVER, os.path.join(PLUG, lib), str(n), str(nsteps), str(DT), sp, gpu_out, str(rec)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([HARNESS, "ionic", os.path.join(PLUG, lib), str(n), str(nsteps), str(DT), "-", sp, cpu_out, str(rec)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    from make_golden import MODELS
    r = subprocess.run([HARNESS, "ionic", os.path.join(REF_DIR, "shared_libs", MODELS[model][0]), str(n), str(nsteps), str(DT),
                        "-", sp, ref_out, str(rec)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a, b, c = np.fromfile(gpu_out), np.fromfile(cpu_out), np.fromfile(ref_out)
    assert a.size == b.size == c.size == (nsteps // rec) * n * MODELS[model][2]
    assert np.array_equal(a, b), np.abs(a - b).max()
    neq = MODELS[model][2]
    a, c = a.reshape(-1, neq), c.reshape(-1, neq)
    scale = np.maximum(np.abs(c).max(axis=0), 1e-30)
    assert (np.abs(a - c) / scale).max() < 1e-9
