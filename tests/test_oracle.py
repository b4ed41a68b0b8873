"""The oracle against the reference's own outputs and against fixed mathematics.

Golden fixtures (tests/golden/, written by make_golden.py from the reference
built from its sources) pin the oracle bit for bit where the arithmetic is the
same (ionic steps, assembly), and to the reference's solver tolerance where it
is not (the tissue run's CG).  Mathematical checks cover what the fixtures leave
open: the reference's own CG test systems (test_solvers.c:28-111), a constant
field in the operator's null space up to the alpha term, symmetry of the
isotropic operator, CG against a direct solve, and closed forms of alpha and
the fibre tensor.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
H = (100.0, 100.0, 100.0)
DT = 0.02


def _load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def _stim_on(nsteps):
    """Steps whose accumulated time lies in [0, 2] ms (ode_solver.c:319-327)."""
    on, t = np.zeros(nsteps, bool), 0.0
    for k in range(nsteps):
        on[k] = 0.0 <= t <= 2.0
        t += DT
    return on


# ---------------------------------------------------------------- ionic ----

@pytest.mark.parametrize("name", ["tt2006", "ord2011", "ms2003"])
def test_ionic_trace_bit_exact(oracle, name):
    """25000 steps of two stimulated cells, every 50th state, bit for bit."""
    g = _load(f"ionic_{name}")
    mid = oracle.MODEL_IDS[name]
    trace, every, amps = g["trace"], int(g["trace_every"]), g["trace_amps"]
    n = len(amps)
    sv = oracle.model_init(mid, n)
    on = _stim_on(len(trace) * every)
    zero = np.zeros(n)
    for j in range(len(trace)):
        for k in range(j * every, (j + 1) * every):
            oracle.model_solve(mid, sv, amps if on[k] else zero, DT, 1)
        np.testing.assert_array_equal(sv, trace[j], err_msg=f"record {j}")


@pytest.mark.parametrize("name", ["tt2006", "ord2011", "ms2003"])
def test_ionic_one_step_pairs_bit_exact(oracle, name):
    g = _load(f"ionic_{name}")
    mid = oracle.MODEL_IDS[name]
    sv = np.ascontiguousarray(g["pair_in"])
    oracle.model_solve(mid, sv, g["pair_stim"], float(g["dt"]), 1)
    np.testing.assert_array_equal(sv, g["pair_out"])


def test_ionic_initial_states(oracle):
    """INITIAL_V of each model header (float literals: ten_tusscher_2006.h:7,
    Ohara_Rudy_2011.h:8; double: mitchell_shaeffer_2003.h:8) and NEQ."""
    for name, v0 in (("tt2006", float(np.float32(-85.23))), ("ord2011", -85.0), ("ms2003", 0.00000820413566106744)):
        mid = oracle.MODEL_IDS[name]
        sv = oracle.model_init(mid, 3)
        assert sv.shape == (3, oracle.neq(mid))
        assert oracle.initial_v(mid) == v0
        assert np.all(sv == sv[0])
    assert [oracle.neq(oracle.MODEL_IDS[k]) for k in ("tt2006", "ord2011", "ms2003")] == [19, 41, 2]


def test_ionic_rest_is_steady(oracle):
    """Ventricular models do not fire on their own: without stimulus V stays
    below -80 mV for 200 ms and settles (change over the last 100 ms < 0.5 mV)."""
    for name in ("tt2006", "ord2011"):
        mid = oracle.MODEL_IDS[name]
        sv = oracle.model_init(mid, 1)
        vs = []
        for _ in range(200):
            oracle.model_solve(mid, sv, np.zeros(1), DT, 50)
            vs.append(sv[0, 0])
        assert max(vs) < -80.0, (name, max(vs))
        assert abs(vs[-1] - vs[99]) < 0.5, (name, vs[99], vs[-1])


# ------------------------------------------ CRN 1998, forward / adaptive ----
# Courtemanche-Ramirez-Nattel 1998 runs on the reference's default solvers
# (default_solvers.c): plain forward Euler, or the adaptive Euler with per-cell
# step size.  Fixtures: ionic_crn1998.npz (ref_harness over the compiled
# reference plugin) and tissue_cable_crn1998_adpt_40.npz (MonoAlg3D run).

def _crn_on(nsteps, dt):
    on, t = np.zeros(nsteps, bool), 0.0
    for k in range(nsteps):
        on[k] = 0.0 <= t <= 2.0
        t += dt
    return on


def test_crn_euler_trace_and_pairs_bit_exact(oracle):
    g = _load("ionic_crn1998")
    mid = oracle.MODEL_IDS["crn1998"]
    amps, dt, every = g["trace_amps"], float(g["euler_dt"]), int(g["euler_every"])
    trace = g["euler_trace"]
    n = len(amps)
    sv = oracle.model_init(mid, n)
    on = _crn_on(len(trace) * every, dt)
    for j in range(len(trace)):
        for k in range(j * every, (j + 1) * every):
            oracle.model_solve(mid, sv, amps if on[k] else np.zeros(n), dt, 1)
        np.testing.assert_array_equal(sv, trace[j], err_msg=f"record {j}")
    x = np.ascontiguousarray(g["euler_pair_in"])
    oracle.model_solve(mid, x, g["euler_pair_stim"], dt, 1)
    np.testing.assert_array_equal(x, g["euler_pair_out"])


def _adaptive_run(oracle, g, n_records, min_dt=None, max_dt=None):
    mid = oracle.MODEL_IDS["crn1998"]
    amps = g["trace_amps"]
    n = len(amps)
    dtp, tol, every = float(g["adpt_dt_pde"]), float(g["adpt_tol"]), int(g["adpt_every"])
    min_dt = float(g["adpt_min_dt"]) if min_dt is None else min_dt
    max_dt = dtp if max_dt is None else max_dt
    sv = oracle.model_init(mid, n)
    aux = np.zeros((n, 3))
    aux[:, 0] = min_dt
    on = _crn_on(n_records * every, dtp)
    recs, t = [], 0.0
    for k in range(n_records * every):
        oracle.model_solve_adaptive(mid, sv, amps if on[k] else np.zeros(n), t + min_dt, aux, min_dt, max_dt, tol, tol)
        t += dtp
        if (k + 1) % every == 0:
            recs.append((sv.copy(), aux.copy()))
    return recs


def test_crn_adaptive_trace_bit_exact(oracle):
    """400 ms of two stimulated cells: state and the per-cell dt, time_new and
    previous_dt after every 100 calls, bit for bit (the step-size sequence is
    the reference's)."""
    g = _load("ionic_crn1998")
    recs = _adaptive_run(oracle, g, len(g["adpt_trace"]))
    for j, (sv, aux) in enumerate(recs):
        np.testing.assert_array_equal(sv, g["adpt_trace"][j], err_msg=f"record {j}")
        np.testing.assert_array_equal(aux.T, g["adpt_aux"][j], err_msg=f"record {j}")


def test_crn_adaptive_converges_with_tolerance(oracle):
    """The error control does what it is for: against forward Euler at a 25x
    smaller fixed step (dt 4e-4 ms, a stand-in for the exact solution) the
    median |V| error of two 40 ms action potentials shrinks as the tolerance
    tightens (1e-4 -> 1e-5 -> 1e-6) and is below 0.5 mV at the default 1e-5.
    (With min dt = max dt the reference's time bookkeeping takes zero-length
    steps, t + dt - dt rounding, so it is not plain Euler either.)"""
    mid = oracle.MODEL_IDS["crn1998"]
    amps = np.array([-2000.0, -1200.0])
    n = 2
    dtf, T = 4e-4, 40.0
    a = oracle.model_init(mid, n)
    ref, t = [], 0.0
    per_ms = int(round(1 / dtf))
    for k in range(int(round(T / dtf))):
        oracle.model_solve(mid, a, amps if t <= 2.0 else np.zeros(n), dtf, 1)
        t += dtf
        if (k + 1) % per_ms == 0:
            ref.append(a[:, 0].copy())
    med = []
    for tol in (1e-4, 1e-5, 1e-6):
        b = oracle.model_init(mid, n)
        aux = np.zeros((n, 3))
        aux[:, 0] = 1e-4
        got, t = [], 0.0
        for k in range(int(round(T / 0.02))):
            oracle.model_solve_adaptive(mid, b, amps if t <= 2.0 else np.zeros(n), t + 1e-4, aux, 1e-4, 0.02, tol, tol)
            t += 0.02
            if (k + 1) % 50 == 0:
                got.append(b[:, 0].copy())
        med.append(float(np.median(np.abs(np.array(got) - np.array(ref)))))
    assert med[0] > med[1] > med[2], med
    assert med[1] < 0.5, med


def test_crn_adaptive_tissue_matches_reference(oracle):
    """Cable of 40 CRN cells with the adaptive solver (dt_pde 0.05, min dt
    0.001) against MonoAlg3D's run.  The step-size control makes discrete
    accept/reject choices, so the CG's round-off (different summation order in
    the reference's Hilbert-ordered cells) shifts single steps near upstrokes:
    activation times agree within 0.05 ms, Vm to 0.05 mV in the median and
    within 2 mV at the steepest upstroke samples."""
    g = _load("tissue_cable_crn1998_adpt_40")
    n = 40
    shape = (n, 1, 1)
    a = oracle.alpha(0.14, 1.0, 0.05, *H)
    coef = oracle.assemble_iso(shape, H, a, np.full((n, 3), 0.0000176))
    t = oracle.Tissue(shape, H, oracle.MODEL_IDS["crn1998"], coef, dt=0.05, stim=(0.0, 2.0, -2000.0, 500.0, 0.0),
                      adaptive=(0.001, 1e-5, 1e-5))
    every = int(g["every"])
    vm = [t.v.copy()]
    for _ in range(len(g["vm"]) - 1):
        t.step(every)
        vm.append(t.v.copy())
    vm = np.array(vm)
    d = np.abs(vm - g["vm"])
    assert np.median(d) < 0.05 and d.max() < 2.0, (np.median(d), d.max())
    lat = np.full(n, np.inf)
    for k in range(1, len(vm)):
        lo, hi = vm[k - 1], vm[k]
        hit = np.isinf(lat) & (lo < -30.0) & (hi >= -30.0)
        lat[hit] = (k - 1 + (-30.0 - lo[hit]) / (hi[hit] - lo[hit])) * every * 0.05
    assert np.array_equal(np.isfinite(lat), np.isfinite(g["lat"]))
    assert np.abs(lat - g["lat"]).max() < 0.05


# ------------------------------------------------- ionic scalings --------
# Per-cell scalings have no reference counterpart (batch sweeps scale TT2006
# GKr/GCaL/GKs and MS2003 tau_close).  All ones must be the reference model
# bit for bit; otherwise each factor is pinned by what it must do to the
# model's equations: act only through its own current (zeroing that current's
# gate removes the factor's effect, zeroing another gate does not), enter the
# Euler step of V affinely with the physiological sign, and for tau_close
# divide the closing rate of h only above V_gate.

@pytest.mark.parametrize("name", ["tt2006", "ord2011", "ms2003"])
def test_scaling_ones_is_reference_model(oracle, name):
    g = _load(f"ionic_{name}")
    mid = oracle.MODEL_IDS[name]
    sv = np.ascontiguousarray(g["pair_in"])
    oracle.model_solve_scaled(mid, sv, g["pair_stim"], np.ones((len(sv), 4), np.float32), float(g["dt"]), 1)
    np.testing.assert_array_equal(sv, g["pair_out"])


def _tt_plateau(oracle):
    mid = oracle.MODEL_IDS["tt2006"]
    sv = oracle.model_init(mid, 1)
    on = _stim_on(101)
    for k in range(101):
        oracle.model_solve(mid, sv, np.array([-50.0 if on[k] else 0.0]), DT, 1)
    oracle.model_solve(mid, sv, np.zeros(1), DT, 2000)
    assert sv[0, 0] > 0.0, sv[0, 0]  # 40 ms after the stimulus: plateau, outward IKr/IKs, inward ICaL
    return sv


# scale slot -> (state index of its current's gate, sign of dV per unit factor)
TT_SLOTS = {0: (1, -1), 1: (7, +1), 2: (3, -1)}  # GKr via Xr1, GCaL via d, GKs via Xs


def test_tt2006_scaling_acts_on_its_own_current(oracle):
    mid = oracle.MODEL_IDS["tt2006"]
    base = _tt_plateau(oracle)
    for slot, (gate, sign) in TT_SLOTS.items():
        for zeroed, _ in TT_SLOTS.values():
            y = base.copy()
            y[0, zeroed] = 0.0
            outs = []
            for f in (0.5, 1.0, 1.5):
                sc = np.ones((1, 4), np.float32)
                sc[0, slot] = f
                z = y.copy()
                oracle.model_solve_scaled(mid, z, np.zeros(1), sc, DT, 1)
                outs.append(z[0])
            if zeroed == gate:
                np.testing.assert_array_equal(outs[0], outs[1])
                np.testing.assert_array_equal(outs[2], outs[1])
            else:
                lo, mid_v, hi = outs[0][0], outs[1][0], outs[2][0]
                assert np.sign(hi - mid_v) == sign and np.sign(mid_v - lo) == sign, (slot, lo, mid_v, hi)
                assert abs((hi - mid_v) - (mid_v - lo)) <= 1e-9 * abs(hi - mid_v), (slot, lo, mid_v, hi)


def test_ms2003_scaling_divides_closing_rate(oracle):
    mid = oracle.MODEL_IDS["ms2003"]
    for V, closes in ((0.5, True), (0.05, False)):
        steps = []
        for f in (0.5, 1.0, 2.0):
            y = np.array([[V, 0.8]])
            sc = np.array([[f, 1, 1, 1]], np.float32)
            oracle.model_solve_scaled(mid, y, np.zeros(1), sc, DT, 1)
            steps.append((y[0, 1] - 0.8, f, y[0, 0]))
        assert steps[0][2] == steps[1][2] == steps[2][2]  # V does not see tau_close
        if closes:
            # h' = -h / (150 s): the step times the factor is constant and equals -dt h / 150
            for dh, f, _ in steps:
                assert abs(dh * f - (-DT * 0.8 / 150.0)) < 1e-15
        else:
            assert steps[0][0] == steps[1][0] == steps[2][0] > 0.0


# ------------------------------------------------------------- assembly ----

def _alpha(oracle):
    return oracle.alpha(0.14, 1.0, DT, *H)


def test_alpha_closed_form(oracle):
    """ALPHA(beta, cm, dt, dx, dy, dz) = (beta*cm/dt) * UM2_TO_CM2 * dx*dy*dz
    (monodomain/constants.h:10-11), with UM2_TO_CM2 a float literal."""
    a = oracle.alpha(0.14, 1.0, 0.02, 100.0, 100.0, 100.0)
    assert a == (((0.14 * 1.0) / 0.02) * float(np.float32(1.0e-8))) * (100.0 * 100.0 * 100.0)
    assert oracle.alpha(0.14, 1.0, 0.01, 50.0, 50.0, 50.0) == pytest.approx(a / 4.0, rel=1e-15)


@pytest.mark.parametrize("tag", ["box", "holes"])
def test_iso_assembly_bit_exact(oracle, tag):
    g = _load("assembly")
    shape = tuple(int(s) for s in g["shape"])
    n = int(np.prod(shape))
    act = g[f"active_{tag}"]
    coef = oracle.assemble_iso(shape, H, _alpha(oracle), np.full((n, 3), 0.00013), act)
    np.testing.assert_array_equal(coef, g[f"iso_{tag}"])


@pytest.mark.parametrize("tag", ["box", "holes"])
def test_aniso_assembly_bit_exact(oracle, tag):
    g = _load("assembly")
    shape = tuple(int(s) for s in g["shape"])
    act = g[f"active_{tag}"]
    t6 = np.array([oracle.fibre_tensor(oracle.normalise(f), 0.00024, 0.000045) for f in g["fibres"]])
    coef = oracle.assemble_aniso(shape, H, _alpha(oracle), t6, act)
    np.testing.assert_array_equal(coef, g[f"aniso_{tag}"])


def _dense(shape, coef):
    nx, ny, nz = shape
    n = nx * ny * nz
    A = np.zeros((n, n))
    for c in range(n):
        x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
        for s in range(27):
            if coef[c, s] == 0.0:
                continue
            ox, oy, oz = s % 3 - 1, (s // 3) % 3 - 1, s // 9 - 1
            A[c, (x + ox) + nx * ((y + oy) + ny * (z + oz))] += coef[c, s]
    return A


@pytest.mark.parametrize("aniso", [False, True])
@pytest.mark.parametrize("holes", [False, True])
def test_constant_field_gives_alpha(oracle, aniso, holes):
    """A constant potential carries no flux: A 1 = alpha 1 on active cells
    (a dropped or mis-signed flux term breaks it)."""
    shape = (6, 5, 4)
    n = 120
    a = _alpha(oracle)
    act = (np.random.default_rng(5).random(n) > 0.2).astype(np.uint8) if holes else None
    if aniso:
        f = np.array([[np.cos(t), np.sin(t), 0.3] for t in np.linspace(-1.0, 1.0, n)])
        t6 = np.array([oracle.fibre_tensor(oracle.normalise(v), 0.00024, 0.000045) for v in f])
        coef = oracle.assemble_aniso(shape, H, a, t6, act)
    else:
        coef = oracle.assemble_iso(shape, H, a, np.full((n, 3), 0.00013), act)
    y = oracle.stencil_apply(shape, coef, np.ones(n), act)
    mask = np.ones(n, bool) if act is None else act.astype(bool)
    np.testing.assert_allclose(y[mask], a, rtol=1e-12)
    assert np.all(y[~mask] == 0.0)


def test_iso_operator_symmetric_positive(oracle):
    """The isotropic operator with per-cell conductivities is symmetric and
    A - alpha I is positive semi-definite (assembly_common.c:911-1048)."""
    shape = (5, 4, 3)
    n = 60
    sig = np.random.default_rng(6).uniform(0.0001, 0.0003, (n, 3))
    a = _alpha(oracle)
    coef = oracle.assemble_iso(shape, H, a, sig)
    A = _dense(shape, coef)
    np.testing.assert_allclose(A, A.T, rtol=1e-13, atol=0)
    assert np.linalg.eigvalsh(A).min() > 0.99 * a


def test_fibre_tensor_closed_form(oracle):
    """sigma = sigma_t I + (sigma_l - sigma_t) f f^T for a unit fibre f (calc_tensor2,
    assembly_common.c:1121-1139) and read_fibers' normalisation."""
    f = oracle.normalise([3.0, 4.0, 12.0])
    np.testing.assert_allclose(f, np.array([3.0, 4.0, 12.0]) / 13.0, rtol=1e-15)
    sl, st = 0.00024, 0.000045
    t = oracle.fibre_tensor(f, sl, st)
    M = st * np.eye(3) + (sl - st) * np.outer(f, f)
    np.testing.assert_allclose(t, [M[0, 0], M[1, 1], M[2, 2], M[0, 1], M[0, 2], M[1, 2]], rtol=1e-14)


# ---------------------------------------------------------------- solver ----

def _csr(g, k):
    r, c, v = g[f"A{k}_row"], g[f"A{k}_col"], g[f"A{k}_val"]
    n = len(g[f"B{k}"])
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    return np.cumsum(rp), c.astype(np.int32), v


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("jacobi", [False, True])
def test_cg_reference_systems(oracle, k, jacobi):
    """test_solvers.c:28-111: CG with tolerance 1e-16 reaches X to 1e-3."""
    g = _load("solver_systems")
    rp, col, val = _csr(g, k)
    b, X = g[f"B{k}"], g[f"X{k}"]
    x, its, err = oracle.cg_csr(rp, col, val, b, np.zeros_like(b), 1e-16, 1000, jacobi)
    assert np.abs(x - X).max() < 1e-3
    assert its > 1


def test_cg_matches_direct_solve(oracle):
    """CG on an assembled anisotropic operator reaches numpy's direct solve."""
    shape = (5, 4, 3)
    n = 60
    a = _alpha(oracle)
    f = np.array([[np.cos(t), np.sin(t), 0.0] for t in np.linspace(-1.0, 1.0, n)])
    t6 = np.array([oracle.fibre_tensor(oracle.normalise(v), 0.00024, 0.000045) for v in f])
    coef = oracle.assemble_aniso(shape, H, a, t6)
    A = _dense(shape, coef)
    b = np.random.default_rng(7).standard_normal(n)
    x, its, err = oracle.cg_stencil(shape, coef, b, np.zeros(n), 1e-26, 500)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)
    assert err <= 1e-26


def test_cg_counts_like_reference(oracle):
    """Iteration count starts at 1 (linear_system_solver.c:95-102): an exact warm
    start returns 1 without iterating; max_its caps the count."""
    shape = (4, 4, 2)
    n = 32
    coef = oracle.assemble_iso(shape, H, _alpha(oracle), np.full((n, 3), 0.00013))
    xs = np.random.default_rng(8).standard_normal(n)
    b = oracle.stencil_apply(shape, coef, xs)
    x, its, err = oracle.cg_stencil(shape, coef, b, xs, 1e-3, 50)
    assert its == 1 and np.array_equal(x, xs)
    x, its, err = oracle.cg_stencil(shape, coef, b, np.zeros(n), 1e-300, 3)
    assert its == 3


# ---------------------------------------------------------------- tissue ----

def _tissue_case(oracle, name, model, aniso):
    g = _load(name)
    shape = tuple(int(s) for s in g["shape"])
    n = int(np.prod(shape))
    act = g["active"]
    a = _alpha(oracle)
    if aniso:
        t6 = np.array([oracle.fibre_tensor(oracle.normalise(f), 0.00024, 0.000045) for f in g["fibres"]])
        coef = oracle.assemble_aniso(shape, H, a, t6, act)
    else:
        coef = oracle.assemble_iso(shape, H, a, np.full((n, 3), 0.00013), act)
    t = oracle.Tissue(shape, H, oracle.MODEL_IDS[model], coef, active=None if act.all() else act)
    every = int(g["every"])
    probes = g["probe_cells"]
    got = [t.v[probes].copy()]
    for _ in range(len(g["probe_vm"]) - 1):
        t.step(every)
        got.append(t.v[probes].copy())
    return g, np.array(got), act


@pytest.mark.parametrize("name,model,aniso", [
    ("tissue_iso_tt2006_32x32x8", "tt2006", False),
    ("tissue_fibrotic_ord_24x24x6", "ord2011", True),
    ("tissue_cable_tt2006_300", "tt2006", False),
    ("tissue_sheet_tt2006_48x32", "tt2006", False),
])
def test_tissue_run_matches_reference(oracle, name, model, aniso):
    """The oracle's operator-split loop against MonoAlg3D's own run: probe Vm
    within 1e-6 mV (the reference's CG stops at r.r <= 1e-16, so the two runs
    differ only by solver round-off)."""
    g, got, act = _tissue_case(oracle, name, model, aniso)
    ref = g["probe_vm"]
    keep = act[g["probe_cells"]].astype(bool)
    np.testing.assert_allclose(got[:, keep], ref[:, keep], rtol=0, atol=1e-6)


def test_activation_times_match_reference(oracle):
    """The activation/APD restatement on the oracle's own run against the
    reference's .acm output (12x12x2 TT2006, stimuli every 400 ms, 700 ms,
    save every 10 steps): identical counts, times and APDs (float32)."""
    g = _load("activation_tt2006_12x12x2")
    shape = tuple(int(s) for s in g["shape"])
    n = int(np.prod(shape))
    every = int(g["print_rate"])
    coef = oracle.assemble_iso(shape, H, _alpha(oracle), np.full((n, 3), 0.00013))
    t = oracle.Tissue(shape, H, oracle.TT2006, coef, stim=(0.0, 2.0, -50.0, 500.0, 400.0))
    obs = oracle.Observer(n)
    count = 0
    while t.time - 700.0 <= DT:                      # monodomain_solver.c:616
        if count % every == 0:
            obs.observe(t.v, t.time)
        t.step()
        count += 1
    for c in range(n):
        assert len(obs.acts[c]) == g["n_act"][c]
        ref_a = g["act"][c][np.isfinite(g["act"][c])]
        ref_p = g["apd"][c][np.isfinite(g["apd"][c])]
        np.testing.assert_array_equal(np.array(obs.acts[c], np.float32), ref_a.astype(np.float32), err_msg=f"cell {c}")
        np.testing.assert_array_equal(np.array(obs.apds[c], np.float32), ref_p.astype(np.float32), err_msg=f"cell {c}")
