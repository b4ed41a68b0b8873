#!/usr/bin/env python3
"""Throughput of the operator-split monodomain step on one MI355X.

Workload (the metric's configuration): a 400 x 400 x 100 = 16,000,000-cell
cuboid at 100 um, O'Hara-Rudy 2011 endocardial cells (41 states), anisotropic
conductivity sigma_l / sigma_t = 0.24 / 0.045 S/m with fibres in the xy plane
rotating linearly with depth from -60 to +60 degrees (19-point rows),
planar stimulus -50 uA/uF on x < 500 um for 2 ms from t = 0, dt = 0.02 ms,
implicit Euler diffusion solved by CG to a relative residual of 1e-10.
It runs twice: FP64 plain, then FP32 states + FP32 CG inside FP64 defect
correction.  One step = ionic step + RHS update + converged diffusion solve for
every active cell; the metric is cell-timesteps per second.

The same line carries:
  - the fibrotic precision variant (10% of the cells non-conducting, drawn by
    SplitMix64 from seed 3 and left out as inactive cells), timed in both
    precisions;
  - the error report of FP32+refine against FP64 over the full 100 ms on both
    workloads (max |Vm| difference, activation times) and, since no ORd cell
    repolarises within 100 ms, APD90 on a 128 x 128 x 16 box of each workload
    over 400 ms;
  - the reference's own CPU path (oracle/_ref, OpenMP) on a bounded sample of
    the configuration's shape (custom mesh + layered fibre file);
  - the drop-in rate: config 1 (64x64x16 TT2006) through the reference's core
    with all four plugin kinds over libmonoalg_hip.so, beside the same
    configuration on the device-resident path.

Inputs are synthetic (generated here) and stay resident in HBM during the
timed region.  `value` is the FP32+refine rate on the 16M-cell workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cell-timesteps/s at 16M cells: FP64 plain vs FP32+refine at <=0.5ms LAT err"
FIBROSIS_SEED = 3
DT = 0.02


def workload_desc(shape, precision, active=None):
    import monoalg3d_c_amd as m
    nz = shape[2]
    t6 = m.layered_fibre_tensors(nz, -60.0, 60.0, 0.00024, 0.000045)
    d = dict(model=m.MODEL_ORD2011, precision=precision, aniso=True, tensor=t6,
             tensor_source=m.TENSOR_PER_LAYER, cg_rel_tol=1e-10, cg_tol=0.0, cg_max_its=500,
             stim=(0.0, 2.0, -50.0, 500.0, 0.0))
    if active is not None:
        d["active"] = active
    return d


def bytes_per_cell_step(neq, w_state, k_inner, mixed, outer, coef_slots=0, mask=False):
    """Algorithmic HBM bytes per active cell-step (compulsory traffic of the
    algorithm, perfect reuse of stencil neighbours).

    ionic: read neq-1 states + V (FP64), write neq-1 states + b (FP64).
    FP64: residual 4 words (x, b in; r, p out) + 11 words per CG iteration
    (the reference's textbook order: q = A p; x += a p; r -= a q; p = r + b p;
    dots fused), k_inner = the CG iterations run.  FP32+refine: per outer
    sweep the FP64 residual (x, b in; r out in FP32: 20 B) and, after the
    first, the FP64 correction (d FP32 in, x FP64 in/out: 20 B); the inner
    iterations 11 FP32 words each.  Operators stored per cell add coef_slots
    coefficients (FP64 in FP64 sweeps, FP32 in the inner ones) per stencil
    application; per-layer class tables add none; rows formed from
    neighbourhood masks add the 4-byte mask per stencil application."""
    ionic = 2 * (neq - 1) * w_state + 8 + 8
    mb = 4 if mask else 0
    if not mixed:
        return ionic + 4 * 8 + k_inner * 11 * 8 + (1 + k_inner) * (coef_slots * 8 + mb)
    return (ionic + outer * (8 + 8 + 4 + coef_slots * 8 + mb) + max(outer - 1, 0) * (4 + 8 + 8)
            + k_inner * (11 * 4 + coef_slots * 4 + mb))


def run_precision(shape, precision, steps, warmup, active=None):
    """Warm up, time `steps` steps (wall clock around the whole step, device
    resident), then time up to 10 more steps per phase with HIP events."""
    import monoalg3d_c_amd as m
    n = int(np.prod(shape)) if active is None else int(active.sum())
    t0 = time.time()
    tis = m.Tissue(shape, **workload_desc(shape, precision, active))
    setup_s = time.time() - t0
    tis.step(warmup)
    tis.sync()
    s0 = tis.stats
    t0 = time.perf_counter()
    tis.step(steps)
    tis.sync()
    wall = time.perf_counter() - t0
    s1 = tis.stats
    k2 = min(steps, 10)
    tis.set_timing(True)
    tis.step(k2)
    tis.sync()
    s2 = tis.stats
    tis.close()
    its = (s1["cg_iterations_total"] - s0["cg_iterations_total"]) / steps
    inner = (s1["inner_iterations_total"] - s0["inner_iterations_total"]) / steps
    outer = (s1["outer_sweeps_total"] - s0["outer_sweeps_total"]) / steps
    ionic_ms = (s2["ionic_ms_total"] - s1["ionic_ms_total"]) / k2
    solve_ms = (s2["solve_ms_total"] - s1["solve_ms_total"]) / k2
    return dict(rate=n * steps / wall, ms_per_step=1e3 * wall / steps, cells=n, cg_its=its, inner_its=inner, outer=outer,
                ionic_ms=ionic_ms, solve_ms=solve_ms, setup_s=setup_s, class_operator=s1["class_operator"],
                stencil_points=s1["stencil_points"])


def phase_roofline(r, neq, mixed, coef_slots, mask=False):
    """Per-phase algorithmic bandwidth of a run_precision result (GB/s)."""
    w = 4 if mixed else 8
    step_b = bytes_per_cell_step(neq, w, r["inner_its"], mixed, r["outer"], coef_slots, mask)
    ionic_b = (2 * (neq - 1) * w + 16) * r["cells"]
    solve_b = (step_b - (2 * (neq - 1) * w + 16)) * r["cells"]
    return dict(ionic=dict(achieved=ionic_b / (r["ionic_ms"] * 1e-3) / 1e9, bytes=ionic_b, ms=r["ionic_ms"]),
                solve=dict(achieved=solve_b / (r["solve_ms"] * 1e-3) / 1e9, bytes=solve_b, ms=r["solve_ms"]),
                step_bytes_per_cell=step_b,
                step_achieved=step_b * r["rate"] / 1e9, step_frac=step_b * r["rate"] / 1e9 / HBM_PEAK_GBS)


def error_run(shape, T, active=None, compare_every=50, progress=""):
    """FP64 and FP32+refine run side by side (untimed) from t = 0 to T ms on
    the same inputs: the largest |Vm| difference over active cells, sampled
    every compare_every steps, and per-cell activation time / APD90 from the
    on-device tracker (first upward crossing of -30 mV; APD90 to the 90%
    repolarisation level; both interpolated between steps)."""
    import monoalg3d_c_amd as m
    live = np.ones(int(np.prod(shape)), bool) if active is None else active.astype(bool)
    t64 = m.Tissue(shape, **workload_desc(shape, m.FP64, active))
    t32 = m.Tissue(shape, **workload_desc(shape, m.FP32_REFINE, active))
    for t in (t64, t32):
        t.apd_setup(-30.0)
    nsteps = int(round(T / DT))
    vmax_err, nan64, nan32 = 0.0, None, None
    t_start = time.time()
    for k in range(1, nsteps + 1):
        if t64.step(1) > 0 and nan64 is None:
            nan64 = k
        if t32.step(1) > 0 and nan32 is None:
            nan32 = k
        t64.apd_sample()
        t32.apd_sample()
        if k % compare_every == 0 or k == nsteps:
            d = np.abs(t64.v - t32.v)[live]
            vmax_err = max(vmax_err, float(d.max()))
        if progress and k % 1000 == 0:
            print(f"[bench] {progress}: {k}/{nsteps} steps, {time.time() - t_start:.0f} s", file=sys.stderr, flush=True)
    a64, a32 = t64.apd(), t32.apd()
    t64.close()
    t32.close()
    l64, l32 = a64["lat"][live], a32["lat"][live]
    act64, act32 = np.isfinite(l64), np.isfinite(l32)
    both = act64 & act32
    one = act64 ^ act32
    lat_err = float(np.abs(l64[both] - l32[both]).max()) if both.any() else None
    # a cell activated in one run only is a front cell at the end of the window:
    # it is charged the time left to the window's end
    if one.any():
        t_one = np.where(act64, l64, l32)[one]
        lat_err = max(lat_err or 0.0, float((T - t_one).max()))
    p64, p32 = a64["apd90"][live], a32["apd90"][live]
    rep = np.isfinite(p64) & np.isfinite(p32)
    return dict(shape=list(shape), simulated_ms=T, cells=int(live.sum()), nan_step_fp64=nan64, nan_step_fp32=nan32,
                max_abs_vm_mv=vmax_err, vm_compared_every_ms=compare_every * DT,
                activated_fp64=int(act64.sum()), activated_fp32=int(act32.sum()), activated_in_one_only=int(one.sum()),
                max_abs_lat_ms=lat_err,
                repolarised_fp64=int(np.isfinite(p64).sum()), repolarised_fp32=int(np.isfinite(p32).sum()),
                max_abs_apd90_ms=float(np.abs(p64[rep] - p32[rep]).max()) if rep.any() else None,
                mean_apd90_fp64_ms=float(p64[rep].mean()) if rep.any() else None,
                within_0p5_ms=bool(lat_err is not None and lat_err <= 0.5 and nan64 is None and nan32 is None))


def measured_traffic(kernel, shape):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, each
    scaled by the same counters on a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        e = d["kernels"][kernel]
        if tuple(d.get("shape", ())) != tuple(shape):
            return None
        return int(e["bytes_per_launch"])
    except (KeyError, ValueError, TypeError):
        return None


def copy_calibration(mb):
    """Three device-to-device copies of `mb` MiB through the HIP runtime (with
    HSA_ENABLE_SDMA=0 in the environment they run as the runtime's copy kernel):
    the known-size reference for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = C.c_size_t(mb << 20)
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), n) or hip.hipMalloc(C.byref(b), n):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 0, n)
    for _ in range(3):
        if hip.hipMemcpy(b, a, n, 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


REF_DIR = os.path.join(ROOT, "oracle", "_ref")
REF_EXE = os.path.join(REF_DIR, "bin", "MonoAlg3D")


def _ref_ini(threads, T, domain, assembly, ode_lib, tol, extra=""):
    return f"""[main]
num_threads = {threads}
dt_pde = {DT}
simulation_time = {T:.6f}
abort_on_no_activity = false
use_adaptivity = false
[update_monodomain]
main_function=update_monodomain_default
[assembly_matrix]
init_function=set_initial_conditions_fvm
{assembly}
[linear_system_solver]
tolerance={tol:.17g}
use_preconditioner=no
max_iterations=500
main_function=conjugate_gradient
[domain]
{domain}
[ode_solver]
dt={DT}
use_gpu=no
library_file={ode_lib}
[stim_plain]
start = 0.0
duration = 2.0
current = -50.0
x_limit = 500.0
main_function=stim_if_x_less_than
{extra}"""


def _iteration_times(log):
    """Per-iteration wall times (s) the reference logs ("Total Iteration time")."""
    return [int(m) * 1e-6 for m in re.findall(r"Total Iteration time: (\d+) us", log)]


def cpu_baseline(max_threads=16, shape=(256, 256, 32), steps=12, skip=2):
    """The reference's own CPU path (oracle/_ref MonoAlg3D, OpenMP) on a
    bounded sample of the configuration's shape: a custom mesh of a
    256 x 256 x 32 box with the layered fibre file (-60..+60 degrees over
    depth), ORd 2011, anisotropic 0.24/0.045 S/m, the planar stimulus, CG to
    the absolute r.r that a relative residual of 1e-10 gives at rest.  Rate =
    cells / mean per-iteration time of iterations skip..steps (the reference's
    own per-iteration log), i.e. a steady per-step rate after initialisation."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.inputs import write_custom_mesh, write_layered_fibres
    if not os.path.exists(REF_EXE):
        return None
    threads = max(1, min(os.cpu_count() or 1, max_threads))
    out_dir = tempfile.mkdtemp(prefix="mono_cpu_baseline_")
    mesh, fib = os.path.join(out_dir, "mesh.alg"), os.path.join(out_dir, "fibres.txt")
    t0 = time.time()
    n = write_custom_mesh(mesh, shape, 100.0)
    write_layered_fibres(fib, shape, -60.0, 60.0)
    alpha = m.alpha()
    tol = 1e-20 * n * (alpha * m.model_initial_v(m.MODEL_ORD2011)) ** 2
    domain = (f"name=Custom Mesh\nmain_function=initialize_grid_with_custom_mesh\nmesh_file={mesh}\n"
              f"number_of_points={n}\nstart_discretization=100.0")
    asm = f"sigma_l=0.00024\nsigma_t=0.000045\nsigma_n=0.000045\nfibers_file={fib}\nmain_function=anisotropic_sigma_assembly_matrix"
    ini = os.path.join(out_dir, "cpu_baseline.ini")
    with open(ini, "w") as fh:
        fh.write(_ref_ini(threads, DT * (steps - 1), domain, asm, "shared_libs/libohara_rudy_endo_2011.so", tol))
    r = subprocess.run([REF_EXE, "-c", ini], cwd=REF_DIR, capture_output=True, text=True, timeout=900)
    wall = time.time() - t0
    shutil.rmtree(out_dir, ignore_errors=True)
    if r.returncode != 0:
        return None
    its = _iteration_times(r.stdout)
    if len(its) <= skip:
        return None
    per_step = float(np.mean(its[skip:]))
    try:
        cpu = [ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name")][0]
    except (OSError, IndexError):
        cpu = "?"
    return dict(value=n / per_step, unit="cell-timesteps/s", cores=threads, kind="reference",
                sample=f"reference MonoAlg3D CPU path (OpenMP {threads} threads, {cpu}), custom mesh {shape[0]}x{shape[1]}x"
                       f"{shape[2]} = {n} cells + layered fibre file (-60..+60 deg over depth), ORd 2011, anisotropic "
                       f"19-point, CG to r.r <= {tol:.3g} (rel. 1e-10 at rest); rate = cells / mean iteration time of "
                       f"iterations {skip}..{len(its) - 1} (the reference's own per-iteration log); run wall {wall:.1f} s")


def dropin_rates(T=10.0):
    """Config 1 (64x64x16 TT2006, isotropic 0.13 S/m, CG r.r <= 1e-16) for T ms
    through the reference's own core with the ionic model, assembly, update and
    CG all from the plugins over libmonoalg_hip.so, against the same
    configuration on the device-resident path (mono_tissue_step, FP64)."""
    import monoalg3d_c_amd as m
    plug = os.path.join(ROOT, "monoalg3d_c_amd", "plugins")
    mono, ode = os.path.join(plug, "libmono_hip_monodomain.so"), os.path.join(plug, "libmono_hip_tt2006.so")
    shape = (64, 64, 16)
    n = int(np.prod(shape))
    out = {"workload": "config 1: 64x64x16 TT2006, isotropic 0.13 S/m, CG r.r <= 1e-16, dt 0.02 ms", "simulated_ms": T}
    if os.path.exists(REF_EXE) and os.path.exists(mono) and os.path.exists(ode):
        out_dir = tempfile.mkdtemp(prefix="mono_dropin_")
        domain = ("name=Cuboid Mesh\nstart_dx=100.0\nstart_dy=100.0\nstart_dz=100.0\nside_length_x=6400.0\n"
                  "side_length_y=6400.0\nside_length_z=1600.0\nmain_function=initialize_grid_with_cuboid_mesh")
        asm = f"library_file={mono}\nsigma_x=0.00013\nsigma_y=0.00013\nsigma_z=0.00013\nmain_function=mono_homogeneous_sigma_assembly_matrix"
        ini = _ref_ini(16, T, domain, asm, ode, 1e-16)
        ini = ini.replace("init_function=set_initial_conditions_fvm", "init_function=mono_set_initial_conditions_fvm")
        ini = ini.replace("main_function=update_monodomain_default", f"library_file={mono}\nmain_function=mono_update_monodomain")
        ini = ini.replace("main_function=conjugate_gradient", f"library_file={mono}\nmain_function=mono_conjugate_gradient")
        path = os.path.join(out_dir, "dropin.ini")
        with open(path, "w") as fh:
            fh.write(ini)
        r = subprocess.run([REF_EXE, "-c", path], cwd=REF_DIR, capture_output=True, text=True, timeout=600)
        shutil.rmtree(out_dir, ignore_errors=True)
        its = _iteration_times(r.stdout) if r.returncode == 0 else []
        if len(its) > 10:
            out["reference_core_with_plugins"] = dict(rate=n / float(np.mean(its[10:])), ms_per_step=1e3 * float(np.mean(its[10:])),
                                                      steps=len(its))
        else:
            out["reference_core_with_plugins"] = dict(error=(r.stderr or r.stdout)[-300:])
    t = m.Tissue(shape, model=m.MODEL_TT2006, tensor=m.iso_tensor(0.00013), cg_tol=1e-16)
    t.step(10)
    t.sync()
    k = int(round(T / DT))
    t0 = time.perf_counter()
    t.step(k)
    t.sync()
    wall = time.perf_counter() - t0
    t.close()
    out["device_resident"] = dict(rate=n * k / wall, ms_per_step=1e3 * wall / k, steps=k)
    return out


def run_batch(args) -> dict:
    """The batch sweep: 256 seeded instances (cables of 100-2000 cells, square
    sheets of 32-256 cells a side, TT2006 with GKr/GCaL/GKs scaled by U[0.5,1.5]
    per instance), packed into one atlas, 1 Hz pacing, dt 0.02 ms, CG relative
    residual 1e-10 per instance (as the atlas's conservative absolute bound).
    The rate counts instance cells only (not the atlas gaps)."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.batch import Batch, sweep_instances
    insts = sweep_instances(256, args.batch_seed)
    out = {}
    for key, prec in (("fp64", m.FP64), ("fp32", m.FP32_REFINE)):
        if args.only not in (key, "both"):
            continue
        t0 = time.time()
        b = Batch(insts, model=m.MODEL_TT2006, precision=prec, tensor=m.iso_tensor(0.00013), cg_rel_tol=1e-10,
                  cg_max_its=500, stim=(0.0, 2.0, -50.0, 500.0, 1000.0))
        setup_s = time.time() - t0
        n_cells = sum(i.n for i in insts)
        b.step(args.warmup)
        b.sync()
        s0 = b.stats
        t0 = time.perf_counter()
        b.step(args.steps)
        b.sync()
        wall = time.perf_counter() - t0
        s1 = b.stats
        k2 = min(args.steps, 10)
        b.tissue.set_timing(True)
        b.step(k2)
        b.sync()
        s2 = b.stats
        out[key] = dict(rate=n_cells * args.steps / wall, ms_per_step=1e3 * wall / args.steps,
                        cg_its=(s1["cg_iterations_total"] - s0["cg_iterations_total"]) / args.steps,
                        inner_its=(s1["inner_iterations_total"] - s0["inner_iterations_total"]) / args.steps,
                        ionic_ms=(s2["ionic_ms_total"] - s1["ionic_ms_total"]) / k2,
                        solve_ms=(s2["solve_ms_total"] - s1["solve_ms_total"]) / k2, setup_s=setup_s,
                        stencil_points=s1["stencil_points"], class_operator=s1["class_operator"])
        b.close()
        print(f"[bench] batch {key} {out[key]}", file=sys.stderr, flush=True)
    from monoalg3d_c_amd.batch import Layout
    lay = Layout(insts)
    head = out.get("fp32") or out["fp64"]
    return {
        "metric": "cell-timesteps/s, 256-instance batch (TT2006 cables and sheets, FP32+refine)",
        "value": head["rate"], "unit": "cell-timesteps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if "fp32" in out else "f64", "data": "synthetic",
        "config": {"workload": "256 instances from seed %d: cables of 100-2000 cells and square sheets of 32-256 "
                               "cells a side at 100 um, TT2006 with GKr/GCaL/GKs x U[0.5,1.5] per instance, 1 Hz "
                               "pacing, CG rel. residual 1e-10, dt 0.02 ms" % args.batch_seed,
                   "instance_cells": sum(i.n for i in insts), "atlas": list(lay.shape),
                   "atlas_efficiency": round(lay.efficiency(), 4),
                   "ms_for_500ms_simulated": head["ms_per_step"] * 25000},
        "fp64": out.get("fp64"), "fp32_refine": out.get("fp32"),
        "speedup_fp32_refine_vs_fp64": (out["fp32"]["rate"] / out["fp64"]["rate"]) if len(out) == 2 else None,
    }


def run_adjoint(args) -> dict:
    """The sensitivity workload: a 256 x 256 Mitchell-Schaeffer sheet at 100 um,
    per-cell conductivity factors ~ log-normal(0, 0.3) and tau_close factors ~
    U[0.8, 1.2] from seed 1, 64 sensors recording V every 1 ms over 200 ms,
    squared-error loss against the recordings of the seed-2 fields; FP32 +
    refinement forward, FP64 adjoint, checkpoints every 50 steps.  Reports the
    forward time, the loss + gradient time and their ratio, and checks the
    gradient against central differences (FP64 forward) on a few cells and
    against the FP64-forward gradient."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.adjoint import config5_fields, config5_sensors, sensor_loss_grad, sensor_traces
    shape = (256, 256, 1)
    nsteps, every = 10000, 50
    g, s = config5_fields(shape, seed=1)
    g2, s2 = config5_fields(shape, seed=2)
    sens = config5_sensors(shape, 64, seed=5)
    target = sensor_traces(shape, sens, nsteps, every, g=g2, s=s2)
    sensor_traces(shape, sens, 200, every, g=g, s=s)  # warm-up (kernel calibration)
    t0 = time.perf_counter()
    sensor_traces(shape, sens, nsteps, every, g=g, s=s)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss, dg, ds, _ = sensor_loss_grad(shape, sens, nsteps, every, target, checkpoint_every=50, g=g, s=s)
    t_grad = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss_all, dg_all, ds_all, _ = sensor_loss_grad(shape, sens, nsteps, every, target, checkpoint_every=0, g=g, s=s)
    t_grad_all = time.perf_counter() - t0
    same = float(max(np.linalg.norm(dg_all - dg) / np.linalg.norm(dg), np.linalg.norm(ds_all - ds) / np.linalg.norm(ds)))
    print(f"[bench] adjoint forward {t_fwd:.2f} s, loss+gradient {t_grad:.2f} s (checkpoints every 50), "
          f"{t_grad_all:.2f} s (all states kept, gradient rel. diff {same:.1e})", file=sys.stderr, flush=True)
    # FP64 reference gradient and finite differences (FP64 forward)
    l64, dg64, ds64, _ = sensor_loss_grad(shape, sens, nsteps, every, target, checkpoint_every=50, g=g, s=s,
                                          precision=m.FP64)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    fd = []
    # The MS gate switch (V < V_gate) makes the loss jump where a perturbation
    # moves some cell's V across V_gate at a step; a central difference that
    # straddles such a jump is meaningless, so each cell is differenced with
    # two step sizes and compared only where the two agree to 1e-4.
    def cdiff(c, rel):
        gp, gm = g.copy(), g.copy()
        gp[c] += rel * g[c]
        gm[c] -= rel * g[c]
        tp = sensor_traces(shape, sens, nsteps, every, g=gp, s=s, precision=m.FP64)
        tm = sensor_traces(shape, sens, nsteps, every, g=gm, s=s, precision=m.FP64)
        return (((tp - target) ** 2).sum() - ((tm - target) ** 2).sum()) / (gp[c] - gm[c])
    for c in np.argsort(-np.abs(dg64))[:args.adjoint_fd_cells]:
        d1, d2 = cdiff(c, 1e-4), cdiff(c, 1e-5)
        ok = abs(d1 - d2) <= 1e-4 * abs(d1)
        fd.append(dict(cell=int(c), adjoint_fp64=float(dg64[c]), adjoint_fp32_forward=float(dg[c]), central_difference=float(d1),
                       central_difference_small_step=float(d2), differences_agree=bool(ok),
                       rel_err_fp64=float(abs(dg64[c] - d1) / abs(d1))))
        print(f"[bench] fd {fd[-1]}", file=sys.stderr, flush=True)
    ratio = t_grad / t_fwd
    return {
        "metric": "adjoint gradient cost / forward cost, 256x256 MS sheet, 200 ms, 64 sensors",
        "value": ratio, "unit": "x forward", "n_gpus": 1, "steps": nsteps, "warmup": 200,
        "ms_per_step": 1e3 * t_grad / nsteps, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 forward / f64 adjoint", "data": "synthetic",
        "config": {"workload": "256x256 Mitchell-Schaeffer sheet at 100 um, g ~ logN(0,0.3), tau_close x U[0.8,1.2] "
                               "(seed 1), target from seed 2, 64 sensors every 1 ms over 200 ms, checkpoints every 50 steps"},
        "forward_s": t_fwd, "loss_and_gradient_s": t_grad, "loss": loss,
        "all_states_kept": {"loss_and_gradient_s": t_grad_all, "ratio": t_grad_all / t_fwd,
                            "state_bytes": 16 * 256 * 256 * (nsteps + 1), "gradient_rel_diff_vs_checkpointed": same},
        "grad_rel_l2_fp32_forward_vs_fp64": {"g": rel(dg, dg64), "tau_close": rel(ds, ds64)},
        "finite_differences": fd,
        "fd_max_rel_err": max((f["rel_err_fp64"] for f in fd if f["differences_agree"]), default=None),
    }


def make_line(args, shape, neq, res, res3, errors, cpu, dropin) -> dict:
    """The JSON line of the 16M-cell workload from the run results."""
    n = int(np.prod(shape))
    head = res.get("fp32") or res["fp64"]
    mixed = "fp32" in res
    rf = {k: phase_roofline(r, neq, k == "fp32", 0) for k, r in res.items()}
    rf3 = {k: phase_roofline(r, neq, k == "fp32", 19 if r["class_operator"] == 0 else 0, r["class_operator"] == 3)
           for k, r in res3.items()}

    # The dominant kernel is the ionic step: one launch per step, more device
    # time than any one kernel of the solve.  Its launch time is the ionic
    # phase, timed with HIP events on the library's stream.
    hk = "fp32" if mixed else "fp64"
    d = rf[hk]["ionic"]
    kname = "k_ionic_step32<1, false>" if hk == "fp32" else "k_ionic_step64<1, true, false>"  # ORd 2011 = model 1, no mask
    roofline = {"bound": "hbm", "achieved": round(d["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(d["achieved"] / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kname, shape),
                "kernel": kname, "launch_ms": round(res[hk]["ionic_ms"], 4), "bytes_per_launch": int(d["bytes"])}
    return {
        "metric": METRIC,
        "value": head["rate"],
        "unit": "cell-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"cuboid {shape[0]}x{shape[1]}x{shape[2]} cells at 100 um, ORd 2011 endo, "
                               "anisotropic 0.24/0.045 S/m fibres -60..+60 deg over depth, planar stimulus, "
                               "CG rel. residual 1e-10, dt 0.02 ms",
                   "cells": n, "precision_value": "fp32+refine" if mixed else "fp64"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64": res.get("fp64"),
        "fp32_refine": res.get("fp32"),
        "speedup_fp32_refine_vs_fp64": (res["fp32"]["rate"] / res["fp64"]["rate"]) if len(res) == 2 else None,
        "phase_roofline": rf,
        "fibrotic": None if not res3 else {
            "workload": "the same cuboid with 10%% of the cells non-conducting (SplitMix64 seed %d, left out as inactive "
                        "cells; each row formed from the cell's 4-byte neighbourhood mask and its layer's tensor when "
                        "class_operator is 3, else stored per cell)" % FIBROSIS_SEED,
            "active_cells": res3[next(iter(res3))]["cells"], "fp64": res3.get("fp64"), "fp32_refine": res3.get("fp32"),
            "speedup_fp32_refine_vs_fp64": (res3["fp32"]["rate"] / res3["fp64"]["rate"]) if len(res3) == 2 else None,
            "phase_roofline": rf3},
        "errors": errors or None,
        "dropin_config1": dropin,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--shape", type=str, default="400,400,100")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32", "both"], default="both")
    ap.add_argument("--lat-ms", type=float, default=100.0,
                    help="simulated ms of the untimed FP64 vs FP32+refine error runs at full size (0: skip)")
    ap.add_argument("--apd-ms", type=float, default=400.0,
                    help="simulated ms of the APD90 error runs on the 128x128x16 box (0: skip)")
    ap.add_argument("--no-fibrosis", action="store_true", help="skip the fibrotic variant")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in (reference core + plugins) rate")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run a device copy of this many MiB (counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--workload", choices=["cuboid", "batch", "adjoint"], default="cuboid",
                    help="cuboid: the 16M-cell metric; batch: the 256-instance sweep; adjoint: gradient cost")
    ap.add_argument("--adjoint-fd-cells", type=int, default=4)
    ap.add_argument("--batch-seed", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every GPU box of this build has one "
              f"GPU; multi-GPU sharding is a design note only)", file=sys.stderr)
        return 2
    import monoalg3d_c_amd as m
    if m.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 1
    if args.workload == "batch":
        print(json.dumps(run_batch(args)))
        return 0
    if args.workload == "adjoint":
        print(json.dumps(run_adjoint(args)))
        return 0
    shape = tuple(int(s) for s in args.shape.split(","))
    n = int(np.prod(shape))
    neq = m.model_neq(m.MODEL_ORD2011)
    fib = None if args.no_fibrosis else m.fibrosis_mask(shape, 0.1, FIBROSIS_SEED)
    res, res3 = {}, {}
    for key, prec in (("fp64", m.FP64), ("fp32", m.FP32_REFINE)):
        if args.only not in (key, "both"):
            continue
        res[key] = run_precision(shape, prec, args.steps, args.warmup)
        print(f"[bench] {key} {res[key]}", file=sys.stderr, flush=True)
        if fib is not None:
            res3[key] = run_precision(shape, prec, args.steps, args.warmup, active=fib)
            print(f"[bench] fibrotic {key} {res3[key]}", file=sys.stderr, flush=True)

    errors = {}
    if len(res) == 2:
        if args.lat_ms > 0:
            errors["config2"] = error_run(shape, args.lat_ms, progress="error run 16M")
            print(f"[bench] errors config2 {errors['config2']}", file=sys.stderr, flush=True)
            if fib is not None:
                errors["config3"] = error_run(shape, args.lat_ms, active=fib, progress="error run 16M fibrotic")
                print(f"[bench] errors config3 {errors['config3']}", file=sys.stderr, flush=True)
        if args.apd_ms > 0:
            small = (128, 128, 16)
            errors["apd90_config2_small"] = error_run(small, args.apd_ms, progress="APD90 run")
            if fib is not None:
                errors["apd90_config3_small"] = error_run(small, args.apd_ms, active=m.fibrosis_mask(small, 0.1, FIBROSIS_SEED),
                                                          progress="APD90 run fibrotic")
            print(f"[bench] errors apd90 {errors.get('apd90_config2_small')} {errors.get('apd90_config3_small')}",
                  file=sys.stderr, flush=True)
    if args.copy_calib_mb > 0:
        copy_calibration(args.copy_calib_mb)

    cpu = None if args.no_cpu_baseline else cpu_baseline()
    if cpu:
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)
    dropin = None if args.no_dropin else dropin_rates()
    if dropin:
        print(f"[bench] dropin {dropin}", file=sys.stderr, flush=True)
    out = make_line(args, shape, neq, res, res3, errors, cpu, dropin)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
