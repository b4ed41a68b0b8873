#!/usr/bin/env python3
"""Generate the golden fixtures in tests/golden/ from the reference itself.

Runs the reference's CPU build made by oracle/ref.mk (oracle/_ref: the
MonoAlg3D binary and its plugins, built from /root/reference sources) and the
harness oracle/ref_harness.c, which drives the reference plugins through their
own dlopen ABI.  Writes small .npz fixtures plus provenance.json.  Needs the
reference sources (this container); the fixtures themselves are data and
travel with the repository.

    python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py ionic      # one group
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
DT = 0.02
H = 100.0
PROV = {}

MODELS = {
    "tt2006": ("libten_tusscher_2006.so", -50.0, 19),
    "ord2011": ("libohara_rudy_endo_2011.so", -50.0, 41),
    "ms2003": ("libmitchell_shaeffer_2003.so", 0.2, 2),
}


def run(cmd, cwd=REF):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{cmd} failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return r.stdout


def stim_window(nsteps):
    """Steps whose accumulated time lies in [0, 2] ms (ode_solver.c:319-327)."""
    on, t = np.zeros(nsteps, bool), 0.0
    for k in range(nsteps):
        on[k] = 0.0 <= t <= 2.0
        t += DT
    return on


def ionic(tmp):
    """G1: action-potential traces; G2: one-step input/output pairs."""
    for name, (lib, amp, neq) in MODELS.items():
        nsteps, rec = 25000, 50
        amps = np.array([amp, 0.6 * amp])
        n = len(amps)
        stim = np.where(stim_window(nsteps)[:, None], amps[None, :], 0.0)
        sp, op = os.path.join(tmp, "stim.bin"), os.path.join(tmp, "out.bin")
        stim.tofile(sp)
        run([os.path.join(REF, "bin", "ref_harness"), "ionic", "./shared_libs/" + lib, str(n), str(nsteps), str(DT), "-", sp, op,
             str(rec)])
        trace = np.fromfile(op).reshape(nsteps // rec, n, neq)
        # one-step pairs from states along a trace recorded every 97 steps
        run([os.path.join(REF, "bin", "ref_harness"), "ionic", "./shared_libs/" + lib, str(n), str(nsteps), str(DT), "-", sp, op,
             "97"])
        states = np.fromfile(op).reshape(-1, n, neq).reshape(-1, neq)[:200]
        rng = np.random.default_rng(11)
        s1 = np.where(rng.random(len(states)) < 0.5, amp, 0.0)
        st_in = os.path.join(tmp, "states.bin")
        states.tofile(st_in)
        s1.tofile(sp)
        run([os.path.join(REF, "bin", "ref_harness"), "ionic", "./shared_libs/" + lib, str(len(states)), "1", str(DT), st_in, sp, op])
        after = np.fromfile(op).reshape(len(states), neq)
        np.savez_compressed(os.path.join(HERE, f"ionic_{name}.npz"), trace=trace, trace_every=rec, trace_amps=amps,
                            pair_in=states, pair_stim=s1, pair_out=after, dt=DT)
        PROV[f"ionic_{name}.npz"] = (f"ref_harness ionic on {lib}: {n} cells, {nsteps} steps of {DT} ms, stimulus amps "
                                     f"{amps.tolist()} for t in [0,2] ms, states every {rec} steps; 200 one-step pairs")


def ionic_crn(tmp):
    """CRN 1998 (a model of the reference's default solvers): G1/G2 for the
    fixed-step Euler (dt 0.01 ms) and G3 for the adaptive Euler (dt_pde 0.02,
    min_dt 0.001, tolerances 1e-5): state and (dt, time_new, previous_dt) per
    cell every 100 calls over 400 ms."""
    lib = "./shared_libs/libcourtemanche_ramirez_nattel_1998.so"
    neq = 21
    amps = np.array([-2000.0, -1200.0])
    n = len(amps)
    H_ = os.path.join(REF, "bin", "ref_harness")
    sp, op = os.path.join(tmp, "stim.bin"), os.path.join(tmp, "out.bin")
    fx = {"trace_amps": amps}
    # fixed step
    dt, nsteps, rec = 0.01, 40000, 100
    on, t = np.zeros(nsteps, bool), 0.0
    for k in range(nsteps):
        on[k] = 0.0 <= t <= 2.0
        t += dt
    np.where(on[:, None], amps[None, :], 0.0).tofile(sp)
    run([H_, "ionic", lib, str(n), str(nsteps), str(dt), "-", sp, op, str(rec)])
    fx.update(euler_dt=dt, euler_every=rec, euler_trace=np.fromfile(op).reshape(-1, n, neq))
    run([H_, "ionic", lib, str(n), str(nsteps), str(dt), "-", sp, op, "97"])
    states = np.fromfile(op).reshape(-1, neq)[:200]
    s1 = np.where(np.random.default_rng(12).random(len(states)) < 0.5, amps[0], 0.0)
    st_in = os.path.join(tmp, "states.bin")
    states.tofile(st_in)
    s1.tofile(sp)
    run([H_, "ionic", lib, str(len(states)), "1", str(dt), st_in, sp, op])
    fx.update(euler_pair_in=states, euler_pair_stim=s1, euler_pair_out=np.fromfile(op).reshape(len(states), neq))
    # adaptive
    dtp, min_dt, tol, nsteps, rec = 0.02, 0.001, 1e-5, 20000, 100
    on = stim_window(nsteps)
    np.where(on[:, None], amps[None, :], 0.0).tofile(sp)
    args = [str(dtp), str(min_dt), str(tol), str(tol)]
    run([H_, "ionic_adpt", lib, str(n), str(nsteps)] + args + ["-", sp, op, str(rec)])
    recs = np.fromfile(op).reshape(-1, n * (neq + 3))
    fx.update(adpt_dt_pde=dtp, adpt_min_dt=min_dt, adpt_tol=tol, adpt_every=rec,
              adpt_trace=recs[:, :n * neq].reshape(-1, n, neq), adpt_aux=recs[:, n * neq:].reshape(-1, 3, n))
    np.savez_compressed(os.path.join(HERE, "ionic_crn1998.npz"), **fx)
    PROV["ionic_crn1998.npz"] = (f"ref_harness ionic / ionic_adpt on libcourtemanche_ramirez_nattel_1998.so: "
                                 f"{n} cells, stimulus amps {amps.tolist()} for t in [0,2] ms; forward Euler dt 0.01 "
                                 f"(state every 100 steps, 400 ms) and 200 one-step pairs; adaptive Euler dt_pde 0.02, "
                                 f"min_dt 0.001, abs/rel tol 1e-5 (state + dt/time_new/previous_dt every 100 calls, 400 ms)")


def fibres_layered(shape, th0=-60.0, th1=60.0):
    nx, ny, nz = shape
    f = np.zeros((nx * ny * nz, 3))
    for c in range(nx * ny * nz):
        z = c // (nx * ny)
        th = np.deg2rad(th0 + (th1 - th0) * z / (nz - 1))
        v = np.array([np.cos(th), np.sin(th), 0.0])
        m = np.sqrt((v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]))
        f[c] = v / m
    return f


def read_rows(path, n):
    raw = open(path, "rb").read()
    A = np.zeros((n, 27))
    off = 0
    import struct
    nx_, ny_ = read_rows.shape[0], read_rows.shape[1]
    while off < len(raw):
        row, ne = struct.unpack_from("<ii", raw, off)
        off += 8
        rx, ry, rz = row % nx_, (row // nx_) % ny_, row // (nx_ * ny_)
        for _ in range(ne):
            col, v = struct.unpack_from("<id", raw, off)
            off += 12
            cx, cy, cz = col % nx_, (col // nx_) % ny_, col // (nx_ * ny_)
            slot = (cz - rz + 1) * 9 + (cy - ry + 1) * 3 + (cx - rx + 1)
            A[row, slot] += v
    return A


def assembly(tmp):
    """G4: assembled rows of the reference's assembly plugins."""
    shape = (7, 6, 5)
    nx, ny, nz = shape
    n = nx * ny * nz
    read_rows.shape = shape
    rng = np.random.default_rng(12)
    out = {}
    for holes in (False, True):
        act = np.ones(n, np.uint8)
        if holes:
            act[rng.random(n) < 0.15] = 0
        fib = fibres_layered(shape)
        cells = os.path.join(tmp, "cells.bin")
        with open(cells, "wb") as fh:
            for c in range(n):
                fh.write(bytes([act[c]]))
                fh.write(fib[c].tobytes())
        tag = "holes" if holes else "box"
        ip = os.path.join(tmp, "iso.bin")
        run([os.path.join(REF, "bin", "ref_harness"), "assemble", "iso", str(nx), str(ny), str(nz), str(H), cells, ip, "0.00013"])
        out[f"iso_{tag}"] = read_rows(ip, n)
        ap = os.path.join(tmp, "an.bin")
        run([os.path.join(REF, "bin", "ref_harness"), "assemble", "aniso", str(nx), str(ny), str(nz), str(H), cells, ap, "0.00024",
             "0.000045", "0.000045"])
        out[f"aniso_{tag}"] = read_rows(ap, n)
        out[f"active_{tag}"] = act
    out["fibres"] = fibres_layered(shape)
    np.savez_compressed(os.path.join(HERE, "assembly.npz"), shape=np.array(shape), **out)
    PROV["assembly.npz"] = ("ref_harness assemble (homogeneous_sigma_assembly_matrix sigma 0.00013; "
                            "anisotropic_sigma_assembly_matrix 0.00024/0.000045 with per-cell fibres rotating -60..60 deg "
                            "over z) on a 7x6x5 cuboid, with and without 15% inactive cells; rows as 27-point stencils")


def ini_text(out_dir, T, model_lib, domain, assembly_sec, print_rate):
    return f"""[main]
num_threads = 8
dt_pde = {DT}
simulation_time = {T}
abort_on_no_activity = false
use_adaptivity = false
[update_monodomain]
main_function=update_monodomain_default
[save_result]
print_rate={print_rate}
output_dir={out_dir}
main_function=save_as_text_or_binary
file_prefix=V
binary=true
[assembly_matrix]
init_function=set_initial_conditions_fvm
{assembly_sec}
[linear_system_solver]
tolerance=1e-16
use_preconditioner=no
max_iterations=200
main_function=conjugate_gradient
[domain]
{domain}
[ode_solver]
dt={DT}
use_gpu=no
library_file=shared_libs/{model_lib}
[stim_plain]
start = 0.0
duration = 2.0
current = -50.0
x_limit = 500.0
main_function=stim_if_x_less_than
"""


def read_snapshots(out_dir, shape, every, nsteps):
    nx, ny, nz = shape
    n = nx * ny * nz
    snaps = []
    for k in range(0, nsteps, every):
        rec = np.fromfile(os.path.join(out_dir, f"V_it_{k}.bin")).reshape(-1, 7)
        lex = (np.floor(rec[:, 0] / H) + nx * (np.floor(rec[:, 1] / H) + ny * np.floor(rec[:, 2] / H))).astype(int)
        v = np.full(n, np.nan)
        v[lex] = rec[:, 6]
        snaps.append(v)
    return np.array(snaps)


def lat_from(snaps, every, thr=-30.0):
    """First upward crossing of thr, linearly interpolated between samples (ms)."""
    lat = np.full(snaps.shape[1], np.inf)
    for k in range(1, len(snaps)):
        a, b = snaps[k - 1], snaps[k]
        hit = np.isinf(lat) & (a < thr) & (b >= thr)
        tt = (k - 1 + (thr - a[hit]) / (b[hit] - a[hit])) * every * DT
        lat[hit] = tt
    return lat


def case_files(out_dir, shape, model, T, every, aniso=False, fibrosis=0.0, cable=False):
    """Write the reference inputs of a tissue case into out_dir: run.ini and,
    for the anisotropic cases, a custom mesh (fibrotic cells omitted, seed 13)
    and a fibre file.  Returns (ini path, active mask, per-cell fibres or None)."""
    nx, ny, nz = shape
    n = nx * ny * nz
    os.makedirs(out_dir, exist_ok=True)
    active = np.ones(n, np.uint8)
    if cable:
        assert ny == 1 and nz == 1 and not aniso
        domain = (f"name=Cable Mesh\nstart_dx={H}\nstart_dy={H}\nstart_dz={H}\ncable_length={nx * H}\n"
                  f"main_function=initialize_grid_with_cable_mesh")
        asm = ("sigma_x=0.00013\nsigma_y=0.00013\nsigma_z=0.00013\n"
               "main_function=homogeneous_sigma_assembly_matrix")
        fib = None
    elif not aniso:
        domain = (f"name=Cuboid Mesh\nstart_dx={H}\nstart_dy={H}\nstart_dz={H}\nside_length_x={nx * H}\n"
                  f"side_length_y={ny * H}\nside_length_z={nz * H}\nmain_function=initialize_grid_with_cuboid_mesh")
        asm = ("sigma_x=0.00013\nsigma_y=0.00013\nsigma_z=0.00013\n"
               "main_function=homogeneous_sigma_assembly_matrix")
        fib = None
    else:
        rng = np.random.default_rng(13)
        if fibrosis > 0:
            active[rng.random(n) < fibrosis] = 0
        fib = fibres_layered(shape)
        mesh = os.path.join(out_dir, "mesh.alg")
        fibf = os.path.join(out_dir, "fibres.txt")
        idx = np.nonzero(active)[0]
        with open(mesh, "w") as fm, open(fibf, "w") as ff:
            for c in idx:
                x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
                fm.write(f"{(x + 0.5) * H:.1f},{(y + 0.5) * H:.1f},{(z + 0.5) * H:.1f},{H / 2},{H / 2},{H / 2}\n")
                f0 = fib[c]
                ff.write("%.17g %.17g %.17g %.17g %.17g 0 0 0 1\n" % (f0[0], f0[1], f0[2], -f0[1], f0[0]))
        domain = (f"name=Custom Mesh\nmain_function=initialize_grid_with_custom_mesh\nmesh_file={mesh}\n"
                  f"number_of_points={len(idx)}\nstart_discretization={H}")
        asm = ("sigma_l=0.00024\nsigma_t=0.000045\nsigma_n=0.000045\n"
               f"fibers_file={fibf}\nmain_function=anisotropic_sigma_assembly_matrix")
    lib = MODELS[model][0]
    ini = os.path.join(out_dir, "run.ini")
    with open(ini, "w") as fh:
        fh.write(ini_text(out_dir, T, lib, domain, asm, every))
    return ini, active, fib


def tissue(tmp, name, shape, model, T, every, aniso=False, fibrosis=0.0, keep_times=(), cable=False):
    nx, ny, nz = shape
    n = nx * ny * nz
    out_dir = os.path.join(tmp, name)
    ini, active, fib = case_files(out_dir, shape, model, T, every, aniso=aniso, fibrosis=fibrosis, cable=cable)
    log = run([os.path.join(REF, "bin", "MonoAlg3D"), "-c", ini])
    nsteps = int(round(T / DT)) + 1
    snaps = read_snapshots(out_dir, shape, every, nsteps)
    its = [int(line.split("Iterations = ")[1].split(",")[0]) for line in log.splitlines() if line.startswith("t = ")]
    fx = {"shape": np.array(shape), "every": every, "active": active, "lat": lat_from(snaps, every),
          "cg_iterations_every_1ms": np.array(its)}
    for tms in keep_times:
        fx[f"vm_t{int(tms)}"] = snaps[int(round(tms / (every * DT)))]
    probes = np.linspace(0, n - 1, 12).astype(int)
    fx["probe_cells"] = probes
    fx["probe_vm"] = snaps[:, probes]
    if fib is not None:
        fx["fibres"] = fib
    np.savez_compressed(os.path.join(HERE, f"tissue_{name}.npz"), **fx)
    PROV[f"tissue_{name}.npz"] = (f"reference MonoAlg3D (oracle/_ref) {nx}x{ny}x{nz} cells at {H} um"
                                  f"{' (cable mesh)' if cable else ''}, {model}, "
                                  f"{'anisotropic 0.24/0.045 S/m, fibres -60..60 deg over z (custom mesh + fibres file)' if aniso else 'isotropic 0.13 S/m'}"
                                  f"{f', {fibrosis:.0%} cells omitted (inactive)' if fibrosis else ''}; stimulus -50 on x<500 um for "
                                  f"t in [0,2] ms; CG tol 1e-16; {T} ms; Vm saved every {every} steps; LAT = first crossing of "
                                  f"-30 mV interpolated between saved samples")


def tissue_crn_adpt(tmp, name="cable_crn1998_adpt_40", ncell=40, T=100.0, every=20):
    """Reference MonoAlg3D run of CRN 1998 with the adaptive Euler ODE solver on
    a cable (as example_configs/cable_mesh_with_CRN_atrial.ini: dt_pde 0.05,
    adaptive with min dt 0.001, sigma 0.0000176, stimulus -2000 for 2 ms)."""
    dtp = 0.05
    out_dir = os.path.join(tmp, name)
    os.makedirs(out_dir, exist_ok=True)
    domain = (f"name=Cable Mesh\nstart_dx={H}\nstart_dy={H}\nstart_dz={H}\ncable_length={ncell * H}\n"
              f"main_function=initialize_grid_with_cable_mesh")
    asm = "sigma_x=0.0000176\nsigma_y=0.0000176\nsigma_z=0.0000176\nmain_function=homogeneous_sigma_assembly_matrix"
    ini = ini_text(out_dir, T, "libcourtemanche_ramirez_nattel_1998.so", domain, asm, every)
    ini = ini.replace(f"dt_pde = {DT}", f"dt_pde = {dtp}").replace(f"[ode_solver]\ndt={DT}", "[ode_solver]\nadaptive=true\ndt=0.001")
    ini = ini.replace("current = -50.0", "current = -2000.0")
    path = os.path.join(out_dir, "run.ini")
    with open(path, "w") as fh:
        fh.write(ini)
    log = run([os.path.join(REF, "bin", "MonoAlg3D"), "-c", path])
    nsteps = int(round(T / dtp)) + 1
    snaps = read_snapshots(out_dir, (ncell, 1, 1), every, nsteps)
    lat = np.full(ncell, np.inf)
    for k in range(1, len(snaps)):
        a, b = snaps[k - 1], snaps[k]
        hit = np.isinf(lat) & (a < -30.0) & (b >= -30.0)
        lat[hit] = (k - 1 + (-30.0 - a[hit]) / (b[hit] - a[hit])) * every * dtp
    its = [int(line.split("Iterations = ")[1].split(",")[0]) for line in log.splitlines() if line.startswith("t = ")]
    np.savez_compressed(os.path.join(HERE, f"tissue_{name}.npz"), shape=np.array((ncell, 1, 1)), every=every, dt_pde=dtp,
                        vm=snaps, lat=lat, cg_iterations_every_1ms=np.array(its))
    PROV[f"tissue_{name}.npz"] = (f"reference MonoAlg3D (oracle/_ref) cable mesh of {ncell} cells at {H} um, CRN 1998 "
                                  f"with the adaptive Euler ODE solver (dt_pde {dtp}, min dt 0.001, abs/rel tol 1e-5), "
                                  f"sigma 0.0000176, stimulus -2000 on x<500 um for t in [0,2] ms, CG tol 1e-16, {T} ms; "
                                  f"Vm of every cell every {every} steps; LAT = first crossing of -30 mV interpolated")


CK_SHAPE, CK_T1, CK_T2, CK_STIM = (10, 8, 2), 6.0, 12.0, (0.0, 0.5, -50.0, 500.0, 1.5)


def checkpoint_case(out_dir, T, extra=""):
    """A 10x8x2 TT2006 box with a periodic stimulus (every 1.5 ms for 0.5 ms)
    and Vm saved every 50 steps; extra is appended to the .ini."""
    nx, ny, nz = CK_SHAPE
    os.makedirs(out_dir, exist_ok=True)
    domain = (f"name=Cuboid Mesh\nstart_dx={H}\nstart_dy={H}\nstart_dz={H}\nside_length_x={nx * H}\n"
              f"side_length_y={ny * H}\nside_length_z={nz * H}\nmain_function=initialize_grid_with_cuboid_mesh")
    asm = "sigma_x=0.00013\nsigma_y=0.00013\nsigma_z=0.00013\nmain_function=homogeneous_sigma_assembly_matrix"
    text = ini_text(out_dir, T, MODELS["tt2006"][0], domain, asm, 50)
    text = text.replace("duration = 2.0\n", f"duration = {CK_STIM[1]}\nperiod = {CK_STIM[4]}\n")
    ini = os.path.join(out_dir, "run.ini")
    with open(ini, "w") as fh:
        fh.write(text + extra)
    return ini


def read_iters(out_dir, shape, iters):
    nx, ny, nz = shape
    out = np.full((len(iters), nx * ny * nz), np.nan)
    for r, k in enumerate(iters):
        rec = np.fromfile(os.path.join(out_dir, f"V_it_{k}.bin")).reshape(-1, 7)
        lex = (np.floor(rec[:, 0] / H) + nx * (np.floor(rec[:, 1] / H) + ny * np.floor(rec[:, 2] / H))).astype(int)
        out[r, lex] = rec[:, 6]
    return out


def checkpoint(tmp):
    """The reference's checkpoint (save_simulation_state at the end of a 6 ms
    run), its restart to 12 ms (restore_simulation_state) and its uninterrupted
    12 ms run."""
    ck = os.path.join(tmp, "ck")
    os.makedirs(ck, exist_ok=True)
    bin_ = os.path.join(REF, "bin", "MonoAlg3D")
    run([bin_, "-c", checkpoint_case(os.path.join(tmp, "ck_a"), CK_T1,
                                      f"[save_state]\nmain_function=save_simulation_state\noutput_dir={ck}\n")])
    run([bin_, "-c", checkpoint_case(os.path.join(tmp, "ck_b"), CK_T2,
                                      f"[restore_state]\nmain_function=restore_simulation_state\ninput_dir={ck}\n")])
    run([bin_, "-c", checkpoint_case(os.path.join(tmp, "ck_c"), CK_T2)])
    iters = np.arange(350, 601, 50)
    fx = {"shape": np.array(CK_SHAPE), "stim": np.array(CK_STIM), "iters": iters,
          "restart_vm": read_iters(os.path.join(tmp, "ck_b"), CK_SHAPE, iters),
          "continuous_vm": read_iters(os.path.join(tmp, "ck_c"), CK_SHAPE, iters),
          "checkpoint_vm": read_iters(os.path.join(tmp, "ck_a"), CK_SHAPE, [300])[0]}
    for name in ("grid", "monodomain_solver", "ode_solver"):
        fx[f"file_{name}"] = np.fromfile(os.path.join(ck, f"{name}_checkpoint.dat"), np.uint8)
    np.savez_compressed(os.path.join(HERE, "checkpoint_tt2006_10x8x2.npz"), **fx)
    PROV["checkpoint_tt2006_10x8x2.npz"] = (
        f"reference MonoAlg3D (oracle/_ref) 10x8x2 cuboid at {H} um, tt2006, isotropic 0.13 S/m, stimulus -50 on "
        f"x<500 um for 0.5 ms every 1.5 ms; save_simulation_state at the end of a {CK_T1} ms run (file_* = the three "
        f"checkpoint files as bytes), restore_simulation_state from it to {CK_T2} ms (restart_vm), the same "
        f"{CK_T2} ms uninterrupted (continuous_vm); Vm at iterations `iters`, lexicographic cells; checkpoint_vm = "
        f"iteration 300 of the first run")


def read_acm(path, shape):
    """The reference's activation_info .acm file (save_mesh.c:754-913): one line
    per active cell, 'cx,cy,cz,hx,hy,hz n [ t1 .. tn ] [ apd1 .. ]'."""
    nx, ny, nz = shape
    n = nx * ny * nz
    lines = open(path).read().splitlines()
    last = int(lines[0])
    nact = np.zeros(n, np.int32)
    acts, apds = [[] for _ in range(n)], [[] for _ in range(n)]
    for ln in lines[1:]:
        geo, rest = ln.split(" ", 1)
        cx, cy, cz = (float(v) for v in geo.split(",")[:3])
        c = int(np.floor(cx / H)) + nx * (int(np.floor(cy / H)) + ny * int(np.floor(cz / H)))
        k = rest.index("[")
        nact[c] = int(rest[:k])
        a, b = rest[k:].split("]", 1)
        acts[c] = [float(v) for v in a.strip("[ ").split()]
        apds[c] = [float(v) for v in b.strip().strip("[]").split()]
    m = max(1, max(len(a) for a in acts), max(len(a) for a in apds))
    act = np.full((n, m), np.nan)
    apd = np.full((n, m), np.nan)
    for c in range(n):
        act[c, :len(acts[c])] = acts[c]
        apd[c, :len(apds[c])] = apds[c]
    return last, nact, act, apd


def activation(tmp):
    """G7: activation times and APDs from the reference's save_with_activation_times
    (save_mesh.c:730-913) on a paced TT2006 slab: 12x12x2 cells, isotropic,
    stimuli at t = 0 and 400 ms, 700 ms, sampled every 10 steps."""
    shape = (12, 12, 2)
    nx, ny, nz = shape
    out_dir = os.path.join(tmp, "act")
    os.makedirs(out_dir, exist_ok=True)
    domain = (f"name=Cuboid Mesh\nstart_dx={H}\nstart_dy={H}\nstart_dz={H}\nside_length_x={nx * H}\n"
              f"side_length_y={ny * H}\nside_length_z={nz * H}\nmain_function=initialize_grid_with_cuboid_mesh")
    asm = "sigma_x=0.00013\nsigma_y=0.00013\nsigma_z=0.00013\nmain_function=homogeneous_sigma_assembly_matrix"
    ini = ini_text(out_dir, 700.0, MODELS["tt2006"][0], domain, asm, 10)
    ini = ini.replace("main_function=save_as_text_or_binary\nfile_prefix=V\nbinary=true",
                      "main_function=save_with_activation_times\ninit_function=init_save_with_activation_times\n"
                      "end_function=end_save_with_activation_times\nactivation_threshold=-30.0\napd_threshold=-83.0\n"
                      "time_threshold=10.0")
    ini = ini.replace("x_limit = 500.0", "x_limit = 500.0\nperiod = 400.0")
    path = os.path.join(out_dir, "run.ini")
    with open(path, "w") as fh:
        fh.write(ini)
    run([os.path.join(REF, "bin", "MonoAlg3D"), "-c", path])
    last, nact, act, apd = read_acm(os.path.join(out_dir, "activation_info_it_0.acm"), shape)
    np.savez_compressed(os.path.join(HERE, "activation_tt2006_12x12x2.npz"), shape=np.array(shape), print_rate=10,
                        last_call=last, n_act=nact, act=act, apd=apd)
    PROV["activation_tt2006_12x12x2.npz"] = (
        "reference MonoAlg3D save_with_activation_times (activation -30 mV, APD -83 mV, time threshold 10 ms, "
        "print_rate 10) on a 12x12x2 TT2006 isotropic 0.13 S/m cuboid, stimulus -50 on x<500 um for 2 ms every "
        "400 ms, 700 ms; activation times and APDs per cell from activation_info_it_0.acm")


def read_octave(path):
    """Octave text matrix (sparse: 'row col value' 1-based; dense: one value per line)."""
    lines = open(path).read().splitlines()
    data = [ln for ln in lines if ln and not ln.startswith("#")]
    if any("sparse" in ln for ln in lines[:5]):
        a = np.array([[float(t) for t in ln.split()] for ln in data])
        return a[:, 0].astype(np.int32) - 1, a[:, 1].astype(np.int32) - 1, a[:, 2]
    return np.array([float(ln) for ln in data])


def solvers(tmp):
    """The reference's own CG test systems (tests_bin/A1-A3, B1-B3, X1-X3; test_solvers.c:28-111)."""
    src = "/root/reference/tests_bin"
    out = {}
    for k in (1, 2, 3):
        r, c, v = read_octave(os.path.join(src, f"A{k}.txt"))
        out[f"A{k}_row"], out[f"A{k}_col"], out[f"A{k}_val"] = r, c, v
        out[f"B{k}"] = read_octave(os.path.join(src, f"B{k}.txt"))
        out[f"X{k}"] = read_octave(os.path.join(src, f"X{k}.txt"))
    np.savez_compressed(os.path.join(HERE, "solver_systems.npz"), **out)
    PROV["solver_systems.npz"] = ("tests_bin/A1-A3, B1-B3, X1-X3 of the reference (Octave SPD systems used by "
                                  "test_solvers.c with a 1e-3 absolute tolerance), converted to COO arrays")


SWEEP_LENGTHS = ["2000", "3500", "5000"]
SWEEP_XLIMITS = ["500", "1000"]


def sweep_batch_files(out_dir):
    """The base run (a TT2006 cable of 5000 um, 40 ms, Vm every 50 steps) and a
    [batch] file sweeping domain|cable_length and stim_plain|x_limit, as the
    reference's batch mode reads them (main_batch.c:17-212).  Returns (batch
    file, base .ini)."""
    ini, _, _ = case_files(os.path.join(out_dir, "base"), (50, 1, 1), "tt2006", 40.0, 50, cable=True)
    batch = os.path.join(out_dir, "sweep.ini")
    with open(batch, "w") as fh:
        fh.write(f"[batch]\ninitial_config={ini}\noutput_folder={os.path.join(out_dir, 'out')}\n"
                 "num_simulations_per_parameter_change=1\n[modify]\n"
                 f"domain|cable_length=list|{'|'.join(SWEEP_LENGTHS)}\nstim_plain|x_limit=list|{'|'.join(SWEEP_XLIMITS)}\n")
    return batch, ini


def sweep(tmp):
    """Every run of the sweep above on the reference binary, one MonoAlg3D run
    per expanded configuration (what each MPI rank of main_batch runs): runs
    numbered from 1 over the product of the lists, the last varying fastest
    (batch_utils.c:215-296)."""
    batch, ini = sweep_batch_files(os.path.join(tmp, "sweep"))
    base = open(ini).read()
    fx, k = {}, 1
    runs = []
    for length in SWEEP_LENGTHS:
        for xl in SWEEP_XLIMITS:
            d = os.path.join(tmp, "sweep", f"ref_run_{k}")
            text = base.replace("cable_length=5000.0", f"cable_length={length}").replace("x_limit = 500.0", f"x_limit = {xl}")
            text = text.replace(f"output_dir={os.path.join(tmp, 'sweep', 'base')}", f"output_dir={d}")
            assert f"cable_length={length}" in text and f"x_limit = {xl}" in text and f"output_dir={d}" in text
            os.makedirs(d, exist_ok=True)
            path = os.path.join(d, "run.ini")
            with open(path, "w") as fh:
                fh.write(text)
            run([os.path.join(REF, "bin", "MonoAlg3D"), "-c", path])
            n = int(round(float(length) / H))
            snaps = read_snapshots(d, (n, 1, 1), 50, int(round(40.0 / DT)) + 1)
            fx[f"run{k}_vm"] = snaps
            fx[f"run{k}_lat"] = lat_from(snaps, 50)
            runs.append((k, float(length), float(xl)))
            k += 1
    fx["runs"] = np.array(runs)
    np.savez_compressed(os.path.join(HERE, "sweep_cable_tt2006.npz"), **fx)
    PROV["sweep_cable_tt2006.npz"] = ("reference MonoAlg3D (oracle/_ref) run once per run of a [batch] sweep of a "
                                      "TT2006 cable (5000 um base, isotropic 0.13 S/m, CG tol 1e-16, 40 ms, Vm every "
                                      "50 steps): domain|cable_length in {2000,3500,5000} x stim_plain|x_limit in "
                                      "{500,1000}, runs numbered as main_batch does (last list fastest)")


def main():
    groups = sys.argv[1:] or ["ionic", "assembly", "tissue", "solvers"]
    if not os.path.exists(os.path.join(REF, "bin", "ref_harness")):
        raise SystemExit("build the reference first: make -f oracle/ref.mk")
    prov_path = os.path.join(HERE, "provenance.json")
    if os.path.exists(prov_path):
        PROV.update(json.load(open(prov_path)))
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        if "ionic" in groups:
            ionic(tmp)
        if "crn" in groups:
            ionic_crn(tmp)
            tissue_crn_adpt(tmp)
        if "assembly" in groups:
            assembly(tmp)
        if "tissue" in groups:
            tissue(tmp, "iso_tt2006_32x32x8", (32, 32, 8), "tt2006", 20.0, 5, keep_times=(10, 20))
            tissue(tmp, "aniso_ord_24x24x6", (24, 24, 6), "ord2011", 10.0, 5, aniso=True, keep_times=(5, 10))
            tissue(tmp, "fibrotic_ord_24x24x6", (24, 24, 6), "ord2011", 10.0, 5, aniso=True, fibrosis=0.1, keep_times=(5, 10))
        if "solvers" in groups:
            solvers(tmp)
        if "activation" in groups:
            activation(tmp)
        if "batch" in groups:
            # instances of a batch, each run on its own by the reference
            tissue(tmp, "cable_tt2006_300", (300, 1, 1), "tt2006", 60.0, 50, cable=True, keep_times=(20, 60))
            tissue(tmp, "sheet_tt2006_48x32", (48, 32, 1), "tt2006", 40.0, 50, keep_times=(20, 40))
        if "checkpoint" in groups:
            checkpoint(tmp)
        if "sweep" in groups:
            sweep(tmp)
        if "config1" in groups:
            tissue(tmp, "config1_tt2006_64x64x16", (64, 64, 16), "tt2006", 50.0, 10, keep_times=(20,))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    with open(prov_path, "w") as fh:
        json.dump(PROV, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
