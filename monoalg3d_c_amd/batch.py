"""Many independent cable and sheet instances in one device-resident run.

The reference runs a parameter sweep as one MPI process per run
(main_batch.c:17-212; the run list is the Cartesian product built by
generate_all_simulations, batch_utils.c:215-296).  Here the instances of a sweep
share one GPU and one set of launches: they are packed into a single 2D atlas
(nz = 1) with at least one inactive cell between any two of them.  An inactive
cell carries no flux (the finite-volume operator treats it like a missing mesh
cell), so the atlas operator is block diagonal, one block per instance, and each
step of the atlas advances every instance by one step of its own run:

  - the ionic launch covers every cell of the atlas; each cell carries its
    instance's index (2 bytes) into a table of the instances' model parameters
    (TT2006 GKr/GCaL/GKs, MS2003 tau_close), which the ionic kernel stages in
    LDS, and a per-cell stimulus mask gives each instance its own
    stimulated region (x < stim_x_limit in the instance's own frame, the rule of
    stim_if_x_less_than, stimuli_library/stimuli.c);
  - one CG solves all blocks at once.  It stops when the atlas residual r.r is
    below the tolerance, which bounds every instance's own r.r by the same
    tolerance, so with the reference's absolute test (`tolerance`) each
    instance is solved at least as tightly as its stand-alone run.  A relative
    test (cg_rel_tol, ||r|| <= rel ||b||) cannot be applied to the atlas as
    is: its ||b|| sums over all instances, so a small cable packed beside large
    sheets would be solved far less tightly than alone.  The library computes
    each instance's ||b_i||^2 on the device after every ionic step and stops
    on r.r <= rel^2 * min_i ||b_i||^2, which implies ||r_i|| <= rel ||b_i||
    for every instance i (the reference's relative test, per instance).

Instances share the model, dt, cell size, conductivity and the stimulus timing.
A cable of n cells is the reference's initialize_grid_with_cable_mesh
(domain.c:87-102: a cuboid n*dx x dy x dz), a sheet of nx x ny cells a cuboid one
cell deep.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import MAX_INSTANCES, MODEL_TT2006, Tissue


@dataclass(frozen=True)
class Instance:
    """One run of a sweep: an nx x ny sheet (ny = 1: a cable of nx cells).

    scale: the four per-cell model factors (TT2006: GKr, GCaL, GKs, unused;
    MS2003: tau_close, unused x3); all 1 is the reference model.
    stim_x_limit: cells with centre x < stim_x_limit (um, instance frame) are
    stimulated."""
    nx: int
    ny: int = 1
    scale: tuple = (1.0, 1.0, 1.0, 1.0)
    stim_x_limit: float = 500.0

    @property
    def n(self) -> int:
        return self.nx * self.ny


@dataclass
class Layout:
    """Shelf packing of instances into one W x H atlas (nz = 1).

    Instances are placed tallest first; each goes on the first shelf with room
    (shelves are as tall as their first instance), else on a new shelf.  `gap`
    inactive cells separate neighbours on a shelf and consecutive shelves."""
    instances: list
    width: int | None = None
    gap: int = 1
    origin: list = field(default_factory=list)
    shape: tuple = (0, 0, 1)

    def __post_init__(self):
        if not self.instances:
            raise ValueError("a batch needs at least one instance")
        if self.gap < 1:
            raise ValueError("instances must be separated by at least one inactive cell")
        for i, inst in enumerate(self.instances):
            if inst.nx < 1 or inst.ny < 1:
                raise ValueError(f"instance {i} has no cells ({inst.nx} x {inst.ny})")
        area = sum((inst.nx + self.gap) * (inst.ny + self.gap) for inst in self.instances)
        widest = max(inst.nx for inst in self.instances)
        W = self.width or max(widest, int(np.ceil(np.sqrt(area))))
        W = -(-W // 64) * 64 if self.width is None else W  # whole 64-cell tile columns
        if widest > W:
            raise ValueError(f"atlas width {W} is narrower than an instance ({widest} cells)")
        order = sorted(range(len(self.instances)), key=lambda i: (-self.instances[i].ny, -self.instances[i].nx, i))
        shelves = []  # [y0, height, next free x]
        origin = [None] * len(self.instances)
        top = 0
        for i in order:
            inst = self.instances[i]
            for sh in shelves:
                if inst.ny <= sh[1] and sh[2] + inst.nx <= W:
                    origin[i] = (sh[2], sh[0])
                    sh[2] += inst.nx + self.gap
                    break
            else:
                y0 = top
                shelves.append([y0, inst.ny, inst.nx + self.gap])
                origin[i] = (0, y0)
                top = y0 + inst.ny + self.gap
        self.origin = origin
        self.shape = (W, top - self.gap, 1)

    @property
    def n(self) -> int:
        return self.shape[0] * self.shape[1]

    def cells(self, i: int) -> np.ndarray:
        """Atlas cell index of instance i's cells, in the instance's own
        lexicographic order (x fastest)."""
        inst, (x0, y0) = self.instances[i], self.origin[i]
        ys, xs = np.meshgrid(np.arange(inst.ny), np.arange(inst.nx), indexing="ij")
        return ((y0 + ys) * self.shape[0] + (x0 + xs)).reshape(-1)

    def active(self) -> np.ndarray:
        a = np.zeros(self.n, np.uint8)
        for i in range(len(self.instances)):
            a[self.cells(i)] = 1
        return a

    def stim_mask(self, dx: float) -> np.ndarray:
        m = np.zeros(self.n, np.uint8)
        for i, inst in enumerate(self.instances):
            lx = np.arange(inst.nx)
            hit = (lx * dx + dx / 2.0) < inst.stim_x_limit
            m[self.cells(i)] = np.tile(hit, inst.ny).astype(np.uint8)
        return m

    def instance_ids(self) -> np.ndarray:
        """Instance of each atlas cell (uint16); gap cells name instance 0
        (they are inactive and never read it)."""
        ids = np.zeros(self.n, np.uint16)
        for i in range(len(self.instances)):
            ids[self.cells(i)] = i
        return ids

    def scale_table(self) -> np.ndarray:
        """The four model factors of each instance, one row per instance."""
        tab = np.ones((len(self.instances), 4), np.float32)
        for i, inst in enumerate(self.instances):
            tab[i, :len(inst.scale)] = inst.scale
        return tab

    def ionic_scale(self) -> np.ndarray:
        """The per-cell form of scale_table (n x 4)."""
        return self.scale_table()[self.instance_ids()]

    def split(self, values: np.ndarray) -> list:
        """Per-instance views of an atlas field (leading axis = cells)."""
        return [values[self.cells(i)] for i in range(len(self.instances))]

    def efficiency(self) -> float:
        """Fraction of atlas cells that belong to an instance."""
        return sum(inst.n for inst in self.instances) / self.n


class Batch:
    """A sweep of instances advanced together on one GPU.

    Keyword arguments other than width/gap go to Tissue (model, precision, dt,
    tensor, cg_tol, stim timing ...); stim's x limit is replaced by each
    instance's own stim_x_limit."""

    def __init__(self, instances, h=(100.0, 100.0, 100.0), model=MODEL_TT2006, width=None, gap=1, **kw):
        self.instances = list(instances)
        self.layout = Layout(self.instances, width=width, gap=gap)
        if len(self.instances) > MAX_INSTANCES:
            raise ValueError(f"{len(self.instances)} instances; one batch holds at most {MAX_INSTANCES}")
        for bad in ("active", "stim_mask", "ionic_scale", "instance", "instance_scale", "g_cell"):
            if bad in kw:
                raise ValueError(f"{bad} is set by the batch layout")
        self.tissue = Tissue(self.layout.shape, h=h, model=model, active=self.layout.active(),
                             stim_mask=self.layout.stim_mask(h[0]), instance=self.layout.instance_ids(),
                             instance_scale=self.layout.scale_table(), **kw)

    def close(self) -> None:
        self.tissue.close()

    def step(self, k: int = 1) -> int:
        return self.tissue.step(k)

    def sync(self) -> None:
        self.tissue.sync()

    @property
    def stats(self) -> dict:
        return self.tissue.stats

    def v(self) -> list:
        """Vm of every instance, each in its own lexicographic order."""
        return self.layout.split(self.tissue.v)

    def observe_setup(self, **kw) -> None:
        self.tissue.observe_setup(**kw)

    def observe(self) -> None:
        self.tissue.observe()

    def activations(self) -> list:
        a = self.tissue.activations()
        out = [dict() for _ in self.instances]
        for k, val in a.items():
            for i, part in enumerate(self.layout.split(val)):
                out[i][k] = part
        return out


def sweep_instances(n_instances: int, seed: int, cable_cells=(100, 2000), sheet_side=(32, 256),
                    scale_range=(0.5, 1.5), cable_fraction=0.5) -> list:
    """A seeded sweep of uneven instances: cables of U{cable_cells} cells and
    square sheets of U{sheet_side} cells a side, each with GKr, GCaL and GKs
    scaled by independent U[scale_range] factors."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_instances):
        f = rng.uniform(scale_range[0], scale_range[1], 3)
        sc = (float(np.float32(f[0])), float(np.float32(f[1])), float(np.float32(f[2])), 1.0)
        if rng.random() < cable_fraction:
            out.append(Instance(int(rng.integers(cable_cells[0], cable_cells[1] + 1)), 1, sc))
        else:
            s = int(rng.integers(sheet_side[0], sheet_side[1] + 1))
            out.append(Instance(s, s, sc))
    return out
