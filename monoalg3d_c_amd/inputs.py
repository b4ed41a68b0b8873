"""Reference inputs on the structured device layout: .ini files, cuboid / cable
/ custom meshes, fibre files, and the reference's output files.

A MonoAlg3D configuration (config_parser.c; sections [main], [domain],
[assembly_matrix], [ode_solver], [linear_system_solver], [stim_*],
[save_result]) becomes a Simulation that runs on one device-resident Tissue:

    sim = load_ini("cuboid_ohara.ini")
    run(sim)          # writes V_it_<k>.bin / .txt and activation_info_it_0.acm like the reference

Domains (domains_library/domain.c, domain_helpers.c):
  initialize_grid_with_cuboid_mesh  side_length_{x,y,z} / start_d{x,y,z} cells
  initialize_grid_with_cable_mesh   cable_length / start_dx x 1 x 1 (domain.c:87-102)
  initialize_grid_with_custom_mesh  mesh_file lines "cx,cy,cz,hx,hy,hz" at one
                                    spacing start_discretization; a cell is active
                                    when its centre is listed (domain_helpers.c:95-278)
The reference keeps cells in Hilbert order of its octree; here they are
lexicographic on the box, and every file is keyed by cell centres, so the two
orders are one permutation fixed by the centres (cell = floor(c / h) per axis).

Assembly (matrix_assembly.c): homogeneous_sigma_assembly_matrix (sigma_x/y/z)
and anisotropic_sigma_assembly_matrix (sigma_l/t/n with a fibers_file of one
"f s n" line per mesh-file line, normalised as read_fibers does,
assembly_common.c:1148-1188, or constant f/s/n vectors).  The tensor is
calc_tensor2 when sigma_t == sigma_n and calc_tensor otherwise
(assembly_common.c:1102-1139).
"""
from __future__ import annotations

import ast
import math
import operator
import os
from dataclasses import dataclass, field

import numpy as np

from . import (FP64, MODEL_CRN1998, MODEL_MS2003, MODEL_ORD2011, MODEL_TT2006, TENSOR_PER_CELL, TENSOR_UNIFORM, Tissue,
               fibre_tensor, iso_tensor, write_acm)

# [ode_solver] library_file basename -> model
MODEL_LIBS = {
    "libten_tusscher_2006": MODEL_TT2006,
    "libohara_rudy_endo_2011": MODEL_ORD2011,
    "libmitchell_shaeffer_2003": MODEL_MS2003,
    "libcourtemanche_ramirez_nattel_1998": MODEL_CRN1998,
}


class ConfigError(ValueError):
    pass


# ------------------------------------------------------------------ ini ----

def read_ini(path: str) -> dict:
    """Sections -> {key: value} (strings), in file order.  ';' and '#' start
    comment lines; `key = value` or `key=value`."""
    sections, cur = {}, None
    with open(path) as fh:
        for raw in fh:
            line = raw.strip()
            if not line or line[0] in ";#":
                continue
            if line.startswith("[") and line.endswith("]"):
                cur = line[1:-1].strip()
                sections.setdefault(cur, {})
                continue
            if "=" not in line or cur is None:
                raise ConfigError(f"{path}: cannot parse line {raw!r}")
            k, v = line.split("=", 1)
            sections[cur][k.strip()] = v.strip()
    return sections


_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
        ast.Pow: operator.pow, ast.USub: operator.neg, ast.UAdd: operator.pos}


def num(text) -> float:
    """A numeric value; arithmetic expressions as the reference's parser accepts
    them (config_parser.c evaluates values with tinyexpr): + - * / ^ and parentheses."""
    if isinstance(text, (int, float)):
        return float(text)
    s = str(text).strip().replace("^", "**")

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.BinOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](ev(node.left), ev(node.right))
        if isinstance(node, ast.UnaryOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](ev(node.operand))
        raise ConfigError(f"not a number: {text!r}")
    try:
        return ev(ast.parse(s, mode="eval"))
    except SyntaxError as e:
        raise ConfigError(f"not a number: {text!r}") from e


def flag(text) -> bool:
    return str(text).strip().lower() in ("yes", "true", "1")


def _get(sec: dict, key: str, default=None, where=""):
    if key in sec:
        return sec[key]
    if default is None:
        raise ConfigError(f"missing {where}{key}")
    return default


# --------------------------------------------------------------- domains ----

def box_cells(side: float, h: float) -> int:
    n = int(round(side / h))
    if n < 1 or abs(n * h - side) > 1e-6 * max(side, h):
        raise ConfigError(f"side {side} is not a whole number of cells of {h}")
    return n


def read_custom_mesh(path: str, npoints: int, h: float):
    """Cells listed in a custom mesh file on a box of spacing h.  Returns
    (shape, active[n] uint8, line_of_cell[n] int64, -1 where not listed)."""
    cent = np.zeros((npoints, 3))
    with open(path) as fh:
        for i in range(npoints):
            parts = fh.readline().split(",")
            if len(parts) < 3:
                raise ConfigError(f"{path}: line {i + 1} has fewer than 3 fields")
            cent[i] = [float(parts[0]), float(parts[1]), float(parts[2])]
    idx = np.floor(cent / h).astype(np.int64)
    if (idx < 0).any() or not np.allclose((idx + 0.5) * h, cent, rtol=0, atol=1e-6 * h):
        raise ConfigError(f"{path}: centres are not on a grid of spacing {h}")
    shape = tuple(int(v) for v in idx.max(axis=0) + 1)
    nx, ny, nz = shape
    lex = idx[:, 0] + nx * (idx[:, 1] + ny * idx[:, 2])
    if len(np.unique(lex)) != len(lex):
        raise ConfigError(f"{path}: a cell is listed twice")
    line = np.full(nx * ny * nz, -1, np.int64)
    line[lex] = np.arange(npoints)
    return shape, (line >= 0).astype(np.uint8), line


def _normalise(v):
    m = math.sqrt((v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]))
    return [v[0] / m, v[1] / m, v[2] / m]


def read_fibres(path: str) -> np.ndarray:
    """(m, 9) rows f, s, n, each normalised (read_fibers with normalize_vector)."""
    rows = []
    with open(path) as fh:
        for line in fh:
            p = line.split()
            if len(p) < 9:
                continue
            v = [float(x) for x in p[:9]]
            rows.append(_normalise(v[0:3]) + _normalise(v[3:6]) + _normalise(v[6:9]))
    return np.array(rows)


def write_custom_mesh(path: str, shape, h: float, active=None) -> int:
    """A custom mesh file (initialize_grid_with_custom_mesh, domain_helpers.c:95-278)
    listing the active cells of a box in lexicographic order, one
    "cx,cy,cz,h/2,h/2,h/2" line each; returns the number of lines."""
    nx, ny, nz = shape
    xs = [f"{(x + 0.5) * h:.1f}" for x in range(nx)]
    half = f"{h / 2}"
    count = 0
    with open(path, "w") as fh:
        for z in range(nz):
            cz = f"{(z + 0.5) * h:.1f}"
            for y in range(ny):
                cy = f"{(y + 0.5) * h:.1f}"
                tail = f",{cy},{cz},{half},{half},{half}\n"
                if active is None:
                    fh.write("".join(x + tail for x in xs))
                    count += nx
                else:
                    row = active[(z * ny + y) * nx:(z * ny + y + 1) * nx]
                    fh.write("".join(xs[x] + tail for x in np.nonzero(row)[0]))
                    count += int(row.sum())
    return count


def write_layered_fibres(path: str, shape, theta0_deg: float, theta1_deg: float, active=None) -> None:
    """The fibre file matching write_custom_mesh's lines: f = (cos t, sin t, 0),
    s = (-sin t, cos t, 0), n = e_z with t rising linearly over depth from
    theta0 to theta1 (endpoints included), one "f s n" line per listed cell."""
    nx, ny, nz = shape
    with open(path, "w") as fh:
        for z in range(nz):
            th = math.radians(theta0_deg + (theta1_deg - theta0_deg) * (z / (nz - 1) if nz > 1 else 0.0))
            c, s_ = math.cos(th), math.sin(th)
            line = "%.17g %.17g 0 %.17g %.17g 0 0 0 1\n" % (c, s_, -s_, c)
            if active is None:
                fh.write(line * (nx * ny))
            else:
                fh.write(line * int(active[z * nx * ny:(z + 1) * nx * ny].sum()))


def calc_tensor(f, s, n, sl, st, sn) -> np.ndarray:
    """D = sl f f^T + st s s^T + sn n n^T, summed as (ff + ss) + nn
    (assembly_common.c:1102-1119); returns (xx, yy, zz, xy, xz, yz)."""
    D = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            D[i, j] = ((sl * (f[i] * f[j])) + (st * (s[i] * s[j]))) + (sn * (n[i] * n[j]))
    return np.array([D[0, 0], D[1, 1], D[2, 2], D[0, 1], D[0, 2], D[1, 2]])


def tensor_of(f, s, n, sl, st, sn) -> np.ndarray:
    return fibre_tensor(f, sl, st) if st == sn else calc_tensor(f, s, n, sl, st, sn)


# ------------------------------------------------------------ simulation ----

@dataclass
class Simulation:
    shape: tuple
    h: tuple
    desc: dict                   # make_desc keyword arguments
    simulation_time: float
    dt_pde: float
    nsteps: int                  # loop iterations of solve_monodomain
    print_rate: int = 0
    output_dir: str | None = None
    save_function: str | None = None
    file_prefix: str = "V"
    binary: bool = False
    activation: dict = field(default_factory=dict)   # save_with_activation_times parameters
    active: np.ndarray | None = None
    checkpoint_dir: str | None = None   # [save_state] output_dir
    checkpoint_rate: int = 0            # [save_state] save_rate (0: at the end only)
    restore_dir: str | None = None      # [restore_state] input_dir

    @property
    def n(self) -> int:
        return int(np.prod(self.shape))


def loop_iterations(final_t: float, dt: float) -> int:
    """Iterations of `while (cur_time - final_t <= dt)` with cur_time += dt from
    0 (monodomain_solver.c:616)."""
    t, k = 0.0, 0
    while t - final_t <= dt:
        t += dt
        k += 1
    return k


def load_ini(path: str) -> Simulation:
    cfg = read_ini(path)
    base = os.path.dirname(os.path.abspath(path))
    rel = lambda p: p if os.path.isabs(p) else os.path.join(base, p)
    main = cfg.get("main", {})
    dt_pde = num(_get(main, "dt_pde", where="[main] "))
    T = num(_get(main, "simulation_time", where="[main] "))

    # domain
    dom = cfg.get("domain")
    if dom is None:
        raise ConfigError("no [domain] section")
    fn = _get(dom, "main_function", where="[domain] ")
    active = None
    line_of_cell = None
    if fn == "initialize_grid_with_cuboid_mesh":
        h = tuple(num(_get(dom, k, where="[domain] ")) for k in ("start_dx", "start_dy", "start_dz"))
        shape = tuple(box_cells(num(_get(dom, f"side_length_{a}", where="[domain] ")), hh) for a, hh in zip("xyz", h))
    elif fn == "initialize_grid_with_cable_mesh":
        h = tuple(num(_get(dom, k, where="[domain] ")) for k in ("start_dx", "start_dy", "start_dz"))
        shape = (box_cells(num(_get(dom, "cable_length", where="[domain] ")), h[0]), 1, 1)
    elif fn == "initialize_grid_with_custom_mesh":
        hh = num(_get(dom, "start_discretization", where="[domain] "))
        h = (hh, hh, hh)
        shape, active, line_of_cell = read_custom_mesh(rel(_get(dom, "mesh_file", where="[domain] ")),
                                                       int(num(_get(dom, "number_of_points", where="[domain] "))), hh)
        if active.all():
            active = None
    else:
        raise ConfigError(f"[domain] main_function {fn} is not a structured box")
    n = int(np.prod(shape))

    # assembly
    asm = cfg.get("assembly_matrix", {})
    afn = _get(asm, "main_function", where="[assembly_matrix] ")
    desc = {}
    if afn == "homogeneous_sigma_assembly_matrix":
        desc["tensor"] = iso_tensor(*(num(_get(asm, f"sigma_{a}", where="[assembly_matrix] ")) for a in "xyz"))
    elif afn == "anisotropic_sigma_assembly_matrix":
        sl, st, sn = (num(_get(asm, f"sigma_{a}", where="[assembly_matrix] ")) for a in "ltn")
        desc["aniso"] = True
        if "fibers_file" in asm:
            if line_of_cell is None:
                raise ConfigError("a fibers_file is indexed by the lines of a custom mesh file")
            fib = read_fibres(rel(asm["fibers_file"]))
            t6 = np.zeros((n, 6))
            for c in np.nonzero(line_of_cell >= 0)[0]:
                r = fib[line_of_cell[c]]
                t6[c] = tensor_of(r[0:3], r[3:6], r[6:9], sl, st, sn)
            desc["tensor"], desc["tensor_source"] = t6, TENSOR_PER_CELL
        else:
            vec = lambda key, dflt: [num(x) for x in asm[key].split(",")] if key in asm else dflt
            f, s, nn = vec("f", [1.0, 0.0, 0.0]), vec("s", [0.0, 1.0, 0.0]), vec("n", [0.0, 0.0, 1.0])
            desc["tensor"], desc["tensor_source"] = tensor_of(f, s, nn, sl, st, sn), TENSOR_UNIFORM
    else:
        raise ConfigError(f"[assembly_matrix] main_function {afn} is not supported")
    if active is not None:
        desc["active"] = active

    # ionic model
    ode = cfg.get("ode_solver", {})
    lib = os.path.splitext(os.path.basename(_get(ode, "library_file", where="[ode_solver] ")))[0]
    if lib not in MODEL_LIBS:
        raise ConfigError(f"ionic model {lib} is not available (models: {', '.join(MODEL_LIBS)})")
    desc["model"] = MODEL_LIBS[lib]
    # configure_ode_solver_from_options (ode_solver.c:306-336): [ode_solver] dt
    # defaults to 0 (config_parser.c:297); a fixed-step run then steps 0.01 ms,
    # an adaptive one (or auto_dt) starts from min dt 1e-10 with max dt = dt_pde;
    # tolerances are `abstol` / `reltol` (config_parser.c:1648-1651, defaults 1e-5)
    dt_ode = num(ode.get("dt", 0.0))
    if flag(ode.get("adaptive", "no")):
        if flag(ode.get("auto_dt", "no")) or dt_ode == 0.0:
            dt_ode = 1e-10
        tol = lambda key, alt: num(ode.get(key, ode.get(alt, 1e-5)))
        desc["adaptive"] = (tol("abstol", "abs_tol"), tol("reltol", "rel_tol"))
    else:
        if dt_ode == 0.0:
            dt_ode = 0.01
        if dt_pde < dt_ode:
            # the loop steps dt_ode, ALPHA keeps the configured dt_pde
            # (monodomain_solver.c:512-519,1326)
            desc["alpha_dt"] = dt_pde
            dt_pde = dt_ode
    desc["dt"] = dt_pde
    desc["dt_ode"] = dt_ode
    desc["beta"] = num(main.get("beta", 0.14))
    desc["cm"] = num(main.get("cm", 1.0))

    # linear solver
    ls = cfg.get("linear_system_solver", {})
    desc["cg_tol"] = num(ls.get("tolerance", 1e-16))
    desc["cg_max_its"] = int(num(ls.get("max_iterations", 200)))
    desc["jacobi"] = flag(ls.get("use_preconditioner", "no"))

    # stimulus
    stims = [(k, v) for k, v in cfg.items() if k.startswith("stim")]
    if len(stims) > 1:
        raise ConfigError("one stimulus section is supported")
    if stims:
        name, st = stims[0]
        if st.get("main_function", "stim_if_x_less_than") != "stim_if_x_less_than":
            raise ConfigError(f"[{name}] main_function {st['main_function']} is not supported")
        desc["stim"] = (num(st.get("start", 0.0)), num(st.get("duration", 0.0)), num(st.get("current", 0.0)),
                        num(_get(st, "x_limit", where=f"[{name}] ")), num(st.get("period", 0.0)))
    else:
        desc["stim"] = (0.0, 0.0, 0.0, 0.0, 0.0)

    sim = Simulation(shape=shape, h=h, desc=desc, simulation_time=T, dt_pde=dt_pde, nsteps=loop_iterations(T, dt_pde),
                     active=active)
    sv = cfg.get("save_result")
    if sv:
        sim.print_rate = int(num(sv.get("print_rate", 0)))
        sim.output_dir = rel(sv["output_dir"]) if "output_dir" in sv else None
        sim.save_function = sv.get("main_function")
        sim.file_prefix = sv.get("file_prefix", "V")
        sim.binary = flag(sv.get("binary", "no"))
        if sim.save_function == "save_with_activation_times":
            sim.activation = dict(time_threshold=num(sv.get("time_threshold", 10.0)),
                                  activation_threshold=num(sv.get("activation_threshold", -30.0)),
                                  apd_threshold=num(sv.get("apd_threshold", -83.0)))
    # checkpoints (monodomain_solver.c:231-260,555-562): both directories
    # default to the [save_result] output directory
    ss = cfg.get("save_state")
    if ss is not None:
        if ss.get("main_function", "save_simulation_state") != "save_simulation_state":
            raise ConfigError(f"[save_state] main_function {ss['main_function']} is not supported")
        sim.checkpoint_dir = rel(ss["output_dir"]) if "output_dir" in ss else sim.output_dir
        sim.checkpoint_rate = int(num(ss.get("save_rate", 0)))
        if not sim.checkpoint_dir:
            raise ConfigError("[save_state] needs an output_dir")
    rs = cfg.get("restore_state")
    if rs is not None:
        if rs.get("main_function", "restore_simulation_state") != "restore_simulation_state":
            raise ConfigError(f"[restore_state] main_function {rs['main_function']} is not supported")
        sim.restore_dir = rel(rs["input_dir"]) if "input_dir" in rs else sim.output_dir
        if not sim.restore_dir:
            raise ConfigError("[restore_state] needs an input_dir")
    return sim


def cell_records(shape, h, v, active=None) -> np.ndarray:
    """(m, 7) rows cx, cy, cz, hx/2, hy/2, hz/2, v of the active cells, x fastest."""
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rec = np.stack([(x.ravel() + 0.5) * h[0], (y.ravel() + 0.5) * h[1], (z.ravel() + 0.5) * h[2],
                    np.full(x.size, h[0] / 2.0), np.full(x.size, h[1] / 2.0), np.full(x.size, h[2] / 2.0), v], axis=1)
    return rec if active is None else rec[np.asarray(active, bool)]


def write_vm(path_prefix: str, iteration: int, shape, h, v, active=None, binary=True) -> str:
    """save_as_text_or_binary (save_mesh.c:248-420): <prefix>_it_<k>.bin, 7
    doubles per cell, or .txt, '%g,...' lines."""
    rec = cell_records(shape, h, v, active)
    path = f"{path_prefix}_it_{iteration}.{'bin' if binary else 'txt'}"
    if binary:
        rec.astype(np.float64).tofile(path)
    else:
        with open(path, "w") as fh:
            for r in rec:
                fh.write("%g,%g,%g,%g,%g,%g,%g\n" % tuple(r))
    return path


def run(sim: Simulation | str, precision: int = FP64, max_events: int = 8, progress=None) -> dict:
    """Run the loop of solve_monodomain for the configuration, saving as the
    [save_result] section asks; returns the tissue's statistics."""
    if isinstance(sim, str):
        sim = load_ini(sim)
    t = Tissue(sim.shape, h=sim.h, precision=precision, **sim.desc)
    try:
        saving = bool(sim.output_dir and sim.print_rate > 0 and sim.save_function)
        if saving:
            os.makedirs(sim.output_dir, exist_ok=True)
        act = sim.save_function == "save_with_activation_times"
        if saving and act:
            t.observe_setup(sim.activation["activation_threshold"], sim.activation["apd_threshold"],
                            sim.activation["time_threshold"], max_events)
        k, cur = 0, 0.0
        if sim.restore_dir:
            t.restore_state(sim.restore_dir)
            st = t.stats
            k, cur = st["steps"], st["time"]
        if sim.checkpoint_dir:
            os.makedirs(sim.checkpoint_dir, exist_ok=True)
        # while (cur_time - finalT <= dt_pde), count and time carried over a
        # restart (monodomain_solver.c:582-616,906-933)
        while cur - sim.simulation_time <= sim.dt_pde:
            if saving and k % sim.print_rate == 0:
                if act:
                    t.observe()
                else:
                    write_vm(os.path.join(sim.output_dir, sim.file_prefix), k, sim.shape, sim.h, t.v, sim.active, sim.binary)
            if t.step(1) == 1:
                break
            if progress:
                progress(k)
            k += 1
            cur += sim.dt_pde
            if sim.checkpoint_dir and sim.checkpoint_rate and k % sim.checkpoint_rate == 0:
                t.save_state(sim.checkpoint_dir)
        if sim.checkpoint_dir and sim.checkpoint_rate == 0:
            t.save_state(sim.checkpoint_dir)
        if saving and act:
            a = t.activations()
            write_acm(os.path.join(sim.output_dir, "activation_info_it_0.acm"), sim.shape, sim.h, a["n_act"], a["act"],
                      a["apd"], active=sim.active)
        return t.stats
    finally:
        t.close()
