"""Parameter sweeps from the reference's batch files, run in one process.

The reference's batch mode (main_batch.c:17-212) reads a file with

    [batch]
    initial_config = <.ini of the base run>
    output_folder = <dir>
    num_simulations_per_parameter_change = 1
    [modify]
    <section>|<name> = list|v1|v2|...      or   range|start|end|increment

builds the Cartesian product of the value lists, the last directive varying
fastest, repeated num_simulations_per_parameter_change times with run numbers
counting on (batch_utils.c:121-296), and gives each MPI rank a slice of the
runs.  Here the runs are grouped: runs that differ only in what an instance of
a batch atlas can carry (a cable's length, a one-cell-thick sheet's sides, the
TT2006 GKr/GCaL/GKs or MS tau_close factors, the stimulus x limit) run
together as one Batch on the GPU; runs that differ in anything else (dt, the
model, the conductivity, ...) form separate groups.  Each run writes its
outputs where the reference would: <output_folder>/<output_dir>_run_<n>
followed by _<name>_<value> for every changed parameter.

Per-instance factors use [extra_data] names: GKr_multiplicator,
GCaL_multiplicator, GKs_multiplicator (TT2006) and tau_close_multiplicator
(MS2003); the reference's TT2006 and MS plugins read no such parameters, so
these have no reference counterpart.
"""
from __future__ import annotations

import json
import os
import tempfile
from dataclasses import dataclass, field

import numpy as np

from . import FP64, MODEL_MS2003, MODEL_TT2006, TENSOR_UNIFORM, write_acm
from .batch import Batch, Instance
from .inputs import ConfigError, load_ini, num, read_ini, write_vm

# values that [main] and [ode_solver] parse with strtof (configure_new_parameters,
# batch_utils.c:9-112)
_FLOAT_PARSED = {("main", "dt_pde"), ("main", "simulation_time"), ("ode_solver", "dt")}
# parameters one instance of a batch atlas can carry
_PER_INSTANCE = {("domain", "cable_length"), ("domain", "side_length_x"), ("domain", "side_length_y"),
                 ("extra_data", "GKr_multiplicator"), ("extra_data", "GCaL_multiplicator"),
                 ("extra_data", "GKs_multiplicator"), ("extra_data", "tau_close_multiplicator")}


@dataclass
class BatchSpec:
    initial_config: str
    output_folder: str
    num_sims: int
    modify: list  # [(section, name, [values as strings])]


@dataclass
class Run:
    run_number: int
    params: list  # [(section, name, value string)]
    sections: dict = field(default_factory=dict)


def range_values(start: float, end: float, inc: float) -> list:
    """range|start|end|increment: start, start+inc, ... while <= end, each
    printed with %f (parse_range_or_list_values)."""
    out, v = [], start
    if inc <= 0:
        raise ConfigError("range increment must be positive")
    while v <= end:
        out.append("%f" % v)
        v += inc
    return out


def read_batch_ini(path: str) -> BatchSpec:
    cfg = read_ini(path)
    b = cfg.get("batch")
    if b is None or "initial_config" not in b:
        raise ConfigError(f"{path}: no [batch] initial_config")
    base = os.path.dirname(os.path.abspath(path))
    rel = lambda p: p if os.path.isabs(p) else os.path.join(base, p)
    modify = []
    for key, val in cfg.get("modify", {}).items():
        sec_name = key.split("|")
        if len(sec_name) != 2:
            raise ConfigError(f"bad [modify] directive {key}: section|name")
        parts = val.split("|")
        if parts[0] == "range":
            if len(parts) != 4:
                raise ConfigError(f"bad range {val}: range|start|end|increment")
            vals = range_values(float(parts[1]), float(parts[2]), float(parts[3]))
        elif parts[0] == "list":
            vals = parts[1:]
        else:
            raise ConfigError(f"bad [modify] value {val}: list|... or range|...")
        modify.append((sec_name[0], sec_name[1], vals))
    return BatchSpec(rel(b["initial_config"]), rel(b.get("output_folder", "batch_simulations")),
                     int(num(b.get("num_simulations_per_parameter_change", 1))), modify)


def generate_runs(modify: list, num_sims: int) -> list:
    """All combinations, last directive fastest, repeated num_sims times;
    run numbers from 1 (get_combinations / generate_all_simulations)."""
    if not modify:
        return []
    sets = [vals for _, _, vals in modify]
    combos = [[]]
    for vals in sets:
        combos = [c + [v] for c in combos for v in vals]
    runs, k = [], 1
    for _ in range(num_sims):
        for c in combos:
            runs.append(Run(k, [(sec, name, v) for (sec, name, _), v in zip(modify, c)]))
            k += 1
    return runs


def apply_run(sections: dict, run: Run) -> dict:
    out = {k: dict(v) for k, v in sections.items()}
    for sec, name, val in run.params:
        if sec.startswith("stim") and sec not in out:
            continue  # the reference changes existing stimulus sections only
        if (sec, name) in _FLOAT_PARSED:
            val = repr(float(np.float32(float(val))))
        out.setdefault(sec, {})[name] = val
    return out


def run_dir(output_folder: str, base_out_dir: str, run: Run) -> str:
    """"<output_folder>/<output_dir>_run_<n>_<name>_<value>..." as main_batch.c:117-130
    builds it ('/' in values becomes '_')."""
    return f"{output_folder}/{base_out_dir}_run_{run.run_number}" + "".join(
        f"_{n}_{v.replace('/', '_')}" for _, n, v in run.params)


def _write_ini(sections: dict, path: str) -> None:
    with open(path, "w") as fh:
        for sec, kv in sections.items():
            fh.write(f"[{sec}]\n")
            for k, v in kv.items():
                fh.write(f"{k} = {v}\n")


def _shared_key(sections: dict) -> str:
    s = {k: dict(v) for k, v in sections.items()}
    for sec, name in _PER_INSTANCE:
        s.get(sec, {}).pop(name, None)
    for k in s:
        if k.startswith("stim"):
            s[k].pop("x_limit", None)
    s.get("save_result", {}).pop("output_dir", None)
    return json.dumps(s, sort_keys=True)


def plan(batch_ini: str) -> list:
    """Groups of runs that share one Batch: [(shared sections, [Run, ...])]."""
    spec = read_batch_ini(batch_ini)
    base = read_ini(spec.initial_config)
    base_dir = os.path.dirname(os.path.abspath(spec.initial_config))
    for sec in base.values():  # make relative paths of the base config absolute
        for k in ("mesh_file", "fibers_file"):
            if k in sec and not os.path.isabs(sec[k]):
                sec[k] = os.path.join(base_dir, sec[k])
    base_out = base.get("save_result", {}).get("output_dir", "./")
    groups = {}
    runs = generate_runs(spec.modify, spec.num_sims) or [Run(1, [])]
    for r in runs:
        r.sections = apply_run(base, r)
        r.sections.setdefault("save_result", {})["output_dir"] = run_dir(spec.output_folder, base_out, r)
        groups.setdefault(_shared_key(r.sections), []).append(r)
    return list(groups.values())


def _instance(sim, sections) -> Instance:
    nx, ny, nz = sim.shape
    if nz != 1:
        raise ConfigError("a batch instance is a cable or a one-cell-thick sheet")
    if sim.active is not None and not np.all(sim.active):
        raise ConfigError("a batch instance has no inactive cells of its own")
    if sim.desc.get("tensor_source", TENSOR_UNIFORM) != TENSOR_UNIFORM:
        raise ConfigError("a batch shares one conductivity tensor across its instances")
    ex = sections.get("extra_data", {})
    m = sim.desc["model"]
    if m == MODEL_TT2006:
        scale = (num(ex.get("GKr_multiplicator", 1)), num(ex.get("GCaL_multiplicator", 1)),
                 num(ex.get("GKs_multiplicator", 1)), 1.0)
    elif m == MODEL_MS2003:
        scale = (num(ex.get("tau_close_multiplicator", 1)), 1.0, 1.0, 1.0)
    else:
        scale = (1.0, 1.0, 1.0, 1.0)
    return Instance(nx, ny, tuple(float(np.float32(x)) for x in scale), sim.desc["stim"][3])


def run_sweep(batch_ini: str, precision: int = FP64, max_events: int = 8) -> list:
    """Run every run of the sweep; returns [(run_number, output_dir)]."""
    done = []
    work = tempfile.TemporaryDirectory(prefix="mono_sweep_")
    for gi, runs in enumerate(plan(batch_ini)):
        sims = []
        for r in runs:
            tmp = os.path.join(work.name, f"group{gi}_run{r.run_number}.ini")
            _write_ini(r.sections, tmp)
            sims.append(load_ini(tmp))
        s0 = sims[0]
        insts = [_instance(s, r.sections) for s, r in zip(sims, runs)]
        kw = {k: v for k, v in s0.desc.items() if k not in ("model", "active")}
        b = Batch(insts, h=s0.h, model=s0.desc["model"], precision=precision, **kw)
        try:
            act = s0.save_function == "save_with_activation_times"
            for s in sims:
                if s.output_dir:
                    os.makedirs(s.output_dir, exist_ok=True)
            if act:
                b.observe_setup(activation_threshold=s0.activation["activation_threshold"],
                                apd_threshold=s0.activation["apd_threshold"],
                                time_threshold=s0.activation["time_threshold"], max_events=max_events)
            for k in range(s0.nsteps):
                if s0.print_rate > 0 and k % s0.print_rate == 0:
                    if act:
                        b.observe()
                    elif s0.save_function:
                        for s, v in zip(sims, b.v()):
                            write_vm(os.path.join(s.output_dir, s.file_prefix), k, s.shape, s.h, v, None, s.binary)
                if b.step(1) == 1:
                    break
            if act:
                for s, a in zip(sims, b.activations()):
                    write_acm(os.path.join(s.output_dir, "activation_info_it_0.acm"), s.shape, s.h, a["n_act"], a["act"],
                              a["apd"])
        finally:
            b.close()
        done += [(r.run_number, s.output_dir) for r, s in zip(runs, sims)]
    work.cleanup()
    return sorted(done)
