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
every cell; the metric is cell-timesteps per second.

Inputs are synthetic (generated here) and stay resident in HBM during the
timed region.  `value` is the FP32+refine rate; the FP64 rate, the speed-up
and the roofline of the dominant kernel are in the same JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cell-timesteps/s at 16M cells: FP64 plain vs FP32+refine at <=0.5ms LAT err"


def workload_desc(shape, precision):
    import monoalg3d_c_amd as m
    nz = shape[2]
    t6 = m.layered_fibre_tensors(nz, -60.0, 60.0, 0.00024, 0.000045)
    return dict(model=m.MODEL_ORD2011, precision=precision, aniso=True, tensor=t6,
                tensor_source=m.TENSOR_PER_LAYER, cg_rel_tol=1e-10, cg_tol=0.0, cg_max_its=500,
                stim=(0.0, 2.0, -50.0, 500.0, 0.0))


def bytes_per_cell_step(neq, w_state, k_iter, mixed, outer):
    """Algorithmic HBM bytes per cell-step (compulsory traffic, perfect stencil reuse).

    ionic: read neq-1 states + V (FP64), write neq-1 states + b (FP64)
    FP64 CG: residual 4 words (x, b in; r, p out) + 11 words per iteration
    (q = A p: p in, q out; x += a p: x, p in, x out; r -= a q: r, q in, r out;
    p = r + b p: r, p in, p out).  Class-table operator: 0 coefficient bytes.
    FP32+refine: per outer sweep residual (x, b FP64 in; r, p, d FP32 out) and
    correction (d FP32 in, x FP64 in/out); inner iterations 11 FP32 words.
    """
    ionic = 2 * (neq - 1) * w_state + 8 + 8
    if not mixed:
        return ionic + 4 * 8 + k_iter * 11 * 8
    return ionic + outer * (8 + 8 + 3 * 4) + max(outer - 1, 0) * (4 + 8 + 8) + k_iter * 11 * 4


def run_precision(shape, precision, steps, warmup):
    """Warm up, time `steps` steps (wall clock around the whole step, device
    resident), then time up to 10 more steps per phase with HIP events."""
    import monoalg3d_c_amd as m
    n = shape[0] * shape[1] * shape[2]
    t0 = time.time()
    tis = m.Tissue(shape, **workload_desc(shape, precision))
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
    return dict(rate=n * steps / wall, ms_per_step=1e3 * wall / steps, cg_its=its, inner_its=inner, outer=outer,
                ionic_ms=ionic_ms, solve_ms=solve_ms, setup_s=setup_s, class_operator=s1["class_operator"],
                stencil_points=s1["stencil_points"])


def lat_run(shape, precision, lat_ms):
    """Untimed: a fresh run to lat_ms with the activation observer on every step
    (save_with_activation_times with print_rate 1, previous Vm kept in FP64);
    returns first activation times and the first step that reported a NaN."""
    import monoalg3d_c_amd as m
    tis = m.Tissue(shape, **workload_desc(shape, precision))
    tis.observe_setup(-30.0, -83.0, 10.0, 1, vm_double=True)
    nan_step = None
    for k in range(int(round(lat_ms / 0.02)) + 1):
        tis.observe()
        if tis.step(1) > 0 and nan_step is None:
            nan_step = k
    a = tis.activations()
    tis.close()
    lat = a["act"][:, 0].astype(np.float64)
    lat[a["n_act"] == 0] = np.inf
    return lat, nan_step


def lat_error(r64, r32, lat_ms):
    """Activation-time error of FP32+refine against FP64 (ms).  A cell activated
    in one run only is a front cell at the end of the window: its error is at
    least the time left to the window's end, which is what it is charged."""
    (lat64, nan64), (lat32, nan32) = r64, r32
    act64, act32 = np.isfinite(lat64), np.isfinite(lat32)
    both = act64 & act32
    err = np.abs(lat64[both] - lat32[both]).max() if both.any() else 0.0
    one = act64 ^ act32
    if one.any():
        t_one = np.where(act64, lat64, lat32)[one]
        err = max(err, float((lat_ms + 0.02 - t_one).max()))
    return dict(nan_step_fp64=nan64, nan_step_fp32=nan32, activated_fp64=int(act64.sum()),
                activated_fp32=int(act32.sum()), activated_in_one_only=int(one.sum()),
                max_abs_err_ms=float(err) if (both.any() or one.any()) else None)


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


def cpu_baseline(max_threads=16):
    """Reference CPU path (oracle/_ref MonoAlg3D, OpenMP) on a bounded sample:
    a 128 x 128 x 32 cuboid with ORd 2011 and anisotropic 0.24/0.045 S/m
    conductivity (uniform in-plane fibre at 45 degrees: same 19-point rows),
    30 steps; rate = cells * steps / Resolution Time."""
    exe = os.path.join(ROOT, "oracle", "_ref", "bin", "MonoAlg3D")
    if not os.path.exists(exe):
        return None
    threads = max(1, min(os.cpu_count() or 1, max_threads))
    nx, ny, nz, steps = 128, 128, 32, 30
    out_dir = os.path.join(ROOT, "gpurun_out", "cpu_baseline")
    os.makedirs(out_dir, exist_ok=True)
    f = np.cos(np.pi / 4)
    ini = f"""[main]
num_threads = {threads}
dt_pde = 0.02
simulation_time = {0.02 * (steps - 1):.6f}
abort_on_no_activity = false
use_adaptivity = false
[update_monodomain]
main_function=update_monodomain_default
[assembly_matrix]
init_function=set_initial_conditions_fvm
sigma_l=0.00024
sigma_t=0.000045
sigma_n=0.000045
f=[{f},{f},0]
main_function=anisotropic_sigma_assembly_matrix
[linear_system_solver]
tolerance=1e-16
use_preconditioner=no
max_iterations=500
main_function=conjugate_gradient
[domain]
name=Cuboid Mesh
start_dx=100.0
start_dy=100.0
start_dz=100.0
side_length_x={nx * 100.0}
side_length_y={ny * 100.0}
side_length_z={nz * 100.0}
main_function=initialize_grid_with_cuboid_mesh
[ode_solver]
dt=0.02
use_gpu=no
library_file=shared_libs/libohara_rudy_endo_2011.so
[stim_plain]
start = 0.0
duration = 2.0
current = -50.0
x_limit = 500.0
main_function=stim_if_x_less_than
"""
    path = os.path.join(out_dir, "cpu_baseline.ini")
    with open(path, "w") as fh:
        fh.write(ini)
    r = subprocess.run([exe, "-c", path], cwd=os.path.join(ROOT, "oracle", "_ref"), capture_output=True, text=True,
                       timeout=600)
    if r.returncode != 0:
        return None
    res_us, loops = None, 0
    for line in r.stdout.splitlines():
        if line.startswith("Resolution Time:"):
            res_us = float(line.split(":")[1].split("μs")[0])
        if line.startswith("t = "):
            loops += 1
    if not res_us:
        return None
    iters = int(round(0.02 * (steps - 1) / 0.02)) + 1
    n = nx * ny * nz
    return dict(value=n * iters / (res_us * 1e-6), unit="cell-timesteps/s", cores=threads, kind="reference",
                sample=f"reference MonoAlg3D CPU (OpenMP {threads} threads), {nx}x{ny}x{nz} cuboid, ORd 2011, "
                       f"anisotropic 19-point, {iters} steps from t=0 incl. stimulus; rate = cells*steps/Resolution Time")


def run_batch(args) -> dict:
    """The batch sweep: 256 seeded instances (cables of 100-2000 cells, square
    sheets of 32-256 cells a side, TT2006 with GKr/GCaL/GKs scaled by U[0.5,1.5]
    per instance), packed into one atlas, 1 Hz pacing, dt 0.02 ms, CG relative
    residual 1e-10.  The rate counts instance cells only (not the atlas gaps)."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.batch import Batch, sweep_instances
    insts = sweep_instances(256, args.batch_seed)
    out = {}
    for key, prec in (("fp64", m.FP64), ("fp32", m.FP32_REFINE)):
        if args.only not in (key, "both"):
            continue
        t0 = time.time()
        b = Batch(insts, model=m.MODEL_TT2006, precision=prec, tensor=m.iso_tensor(0.00013), cg_rel_tol=1e-10,
                  cg_tol=0.0, cg_max_its=500, stim=(0.0, 2.0, -50.0, 500.0, 1000.0))
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--shape", type=str, default="400,400,100")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32", "both"], default="both")
    ap.add_argument("--lat-ms", type=float, default=20.0,
                    help="simulated ms of the untimed FP64 vs FP32+refine activation-time check (0: skip)")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run a device copy of this many MiB (counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--workload", choices=["cuboid", "batch", "adjoint"], default="cuboid",
                    help="cuboid: the 16M-cell metric; batch: the 256-instance sweep; adjoint: gradient cost")
    ap.add_argument("--adjoint-fd-cells", type=int, default=4)
    ap.add_argument("--batch-seed", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (multi-GPU sharding is a design "
              f"note only, see DESIGN.md)", file=sys.stderr)
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
    n = shape[0] * shape[1] * shape[2]
    neq = m.model_neq(m.MODEL_ORD2011)
    res = {}
    if args.only in ("fp64", "both"):
        res["fp64"] = run_precision(shape, m.FP64, args.steps, args.warmup)
        print(f"[bench] fp64 {res['fp64']}", file=sys.stderr, flush=True)
    if args.only in ("fp32", "both"):
        res["fp32"] = run_precision(shape, m.FP32_REFINE, args.steps, args.warmup)
        print(f"[bench] fp32+refine {res['fp32']}", file=sys.stderr, flush=True)
    head = res.get("fp32") or res["fp64"]
    mixed = "fp32" in res
    lat_check = None
    if args.lat_ms > 0 and len(res) == 2:
        lat64 = lat_run(shape, m.FP64, args.lat_ms)
        lat32 = lat_run(shape, m.FP32_REFINE, args.lat_ms)
        lat_check = dict(simulated_ms=args.lat_ms, **lat_error(lat64, lat32, args.lat_ms))
        lat_check["within_0p5_ms"] = (lat_check["max_abs_err_ms"] is not None and lat_check["max_abs_err_ms"] <= 0.5
                                      and lat_check["nan_step_fp32"] is None and lat_check["nan_step_fp64"] is None)
        print(f"[bench] lat check {lat_check}", file=sys.stderr, flush=True)
    if args.copy_calib_mb > 0:
        copy_calibration(args.copy_calib_mb)

    # per-phase rooflines (algorithmic bytes / device time of the phase)
    rf = {}
    for key, r in res.items():
        w = 8 if key == "fp64" else 4
        k_it = r["cg_its"] if key == "fp64" else r["inner_its"]
        step_b = bytes_per_cell_step(neq, w, k_it, key == "fp32", r["outer"])
        ionic_bytes = (2 * (neq - 1) * w + 16) * n
        solve_bytes = (step_b - (2 * (neq - 1) * w + 16)) * n
        rf[key] = dict(
            ionic=dict(achieved=ionic_bytes / (r["ionic_ms"] * 1e-3) / 1e9, bytes=ionic_bytes, ms=r["ionic_ms"]),
            solve=dict(achieved=solve_bytes / (r["solve_ms"] * 1e-3) / 1e9, bytes=solve_bytes, ms=r["solve_ms"]),
            step_bytes_per_cell=step_b)

    # The dominant kernel is the ionic step: one launch per step, more device
    # time than any one kernel of the solve (whose phase time spreads over the
    # stencil, update and residual kernels; see profiles/).  Its launch time is
    # the ionic phase, timed with HIP events on the library's stream.
    hk = "fp32" if mixed else "fp64"
    d = rf[hk]["ionic"]
    kname = "k_ionic_step32<1>" if hk == "fp32" else "k_ionic_step64<1, true>"  # ORd 2011 = model 1; FP64: one-step kernel
    roofline = {"bound": "hbm", "achieved": round(d["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(d["achieved"] / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kname, shape),
                "kernel": kname, "launch_ms": round(res[hk]["ionic_ms"], 4), "bytes_per_launch": int(d["bytes"])}
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    out = {
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
        "lat_check": lat_check,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
