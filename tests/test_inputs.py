"""Reference configuration files on the structured layout, without a GPU.

The .ini files, custom meshes and fibre files are the ones
tests/golden/make_golden.py hands to the reference binary for the tissue
fixtures; loading them must reproduce the inputs those fixtures were run with
(shape, inactive cells, per-cell tensors of the reference's own tensor rule,
model, stimulus, solver settings, loop length).
"""
import os
import sys

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_golden import case_files, read_snapshots  # noqa: E402


def test_numbers_and_flags():
    from monoalg3d_c_amd.inputs import ConfigError, flag, num
    assert num("0.02") == 0.02 and num("2*0.5") == 1.0 and num("-(1+2)^2") == -9.0 and num(" 1e-16 ") == 1e-16
    assert flag("yes") and flag("True") and not flag("no") and not flag("false")
    with pytest.raises(ConfigError):
        num("__import__('os')")


def test_loop_length():
    """while (cur_time - final <= dt) on the accumulated double time: the
    reference's config-1 run (50 ms at 0.02 ms) took 2,501 iterations; the
    accumulated time's rounding decides whether the loop takes one step past
    T / dt + 1 (it does for 20 ms)."""
    from monoalg3d_c_amd.inputs import loop_iterations
    assert loop_iterations(50.0, 0.02) == 2501
    t = 0.0
    for _ in range(1001):
        t += 0.02
    assert loop_iterations(20.0, 0.02) == (1002 if t - 20.0 <= 0.02 else 1001)


def test_cuboid_ini(tmp_path):
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.inputs import load_ini
    ini, act, fib = case_files(str(tmp_path), (32, 32, 8), "tt2006", 20.0, 5)
    sim = load_ini(ini)
    assert sim.shape == (32, 32, 8) and sim.h == (100.0, 100.0, 100.0)
    d = sim.desc
    assert d["model"] == m.MODEL_TT2006 and d["dt"] == 0.02 and d["dt_ode"] == 0.02
    np.testing.assert_array_equal(d["tensor"], m.iso_tensor(0.00013))
    assert d["stim"] == (0.0, 2.0, -50.0, 500.0, 0.0)
    assert d["cg_tol"] == 1e-16 and d["cg_max_its"] == 200 and not d["jacobi"]
    assert sim.nsteps in (1001, 1002) and sim.print_rate == 5 and sim.binary and sim.file_prefix == "V"
    assert sim.save_function == "save_as_text_or_binary" and sim.active is None


def test_custom_mesh_with_fibres(tmp_path, oracle):
    """Custom mesh without the fibrotic cells plus a fibre file: the active
    mask is the omitted-cell pattern, and every active cell's tensor is the
    reference's calc_tensor2 of its normalised fibre (sigma_t = sigma_n)."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.inputs import load_ini
    ini, act, fib = case_files(str(tmp_path), (24, 24, 6), "ord2011", 10.0, 5, aniso=True, fibrosis=0.1)
    sim = load_ini(ini)
    assert sim.shape == (24, 24, 6)
    np.testing.assert_array_equal(sim.active, act)
    d = sim.desc
    assert d["aniso"] and d["tensor_source"] == m.TENSOR_PER_CELL and d["model"] == m.MODEL_ORD2011
    for c in np.nonzero(act)[0][::37]:
        want = oracle.fibre_tensor(oracle.normalise(fib[c]), 0.00024, 0.000045)
        np.testing.assert_array_equal(d["tensor"][c], want)
    assert (d["tensor"][act == 0] == 0).all()


def test_full_tensor_rule():
    """sigma_t != sigma_n: D = sl f f^T + st s s^T + sn n n^T; for orthonormal
    f, s, n its eigenvectors are f, s, n with eigenvalues sl, st, sn."""
    from monoalg3d_c_amd.inputs import tensor_of
    th = 0.3
    f, s, n = [np.cos(th), np.sin(th), 0.0], [-np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]
    t = tensor_of(f, s, n, 3.0, 2.0, 1.0)
    D = np.array([[t[0], t[3], t[4]], [t[3], t[1], t[5]], [t[4], t[5], t[2]]])
    for vec, lam in ((f, 3.0), (s, 2.0), (n, 1.0)):
        np.testing.assert_allclose(D @ np.array(vec), lam * np.array(vec), atol=1e-15)


def test_cable_ini(tmp_path):
    from monoalg3d_c_amd.inputs import load_ini
    ini, act, fib = case_files(str(tmp_path), (300, 1, 1), "tt2006", 60.0, 50, cable=True)
    sim = load_ini(ini)
    assert sim.shape == (300, 1, 1) and sim.nsteps in (3001, 3002)


def test_written_records_read_back(tmp_path):
    """V_it_<k>.bin files as the reference writes them (7 doubles per active
    cell, keyed by centre) read back by the fixtures' reader."""
    from monoalg3d_c_amd.inputs import write_vm
    shape = (5, 4, 3)
    n = 60
    act = np.ones(n, np.uint8)
    act[[3, 17, 40]] = 0
    v = np.random.default_rng(0).normal(size=n)
    write_vm(str(tmp_path / "V"), 0, shape, (100.0, 100.0, 100.0), v, act, binary=True)
    back = read_snapshots(str(tmp_path), shape, 1, 1)[0]
    np.testing.assert_array_equal(back[act == 1], v[act == 1])
    assert np.isnan(back[act == 0]).all()
    p = write_vm(str(tmp_path / "T"), 7, shape, (100.0, 100.0, 100.0), v, None, binary=False)
    rows = np.loadtxt(p, delimiter=",")
    assert rows.shape == (60, 7) and np.allclose(rows[:, 6], v, rtol=1e-5, atol=1e-6)


def test_rejects_unsupported(tmp_path):
    from monoalg3d_c_amd.inputs import ConfigError, load_ini
    ini, act, fib = case_files(str(tmp_path), (8, 8, 2), "tt2006", 1.0, 5)
    text = open(ini).read()
    bad = tmp_path / "bad.ini"
    bad.write_text(text.replace("initialize_grid_with_cuboid_mesh", "initialize_grid_with_rabbit_mesh"))
    with pytest.raises(ConfigError, match="structured box"):
        load_ini(str(bad))
    bad.write_text(text.replace("libten_tusscher_2006", "libluo_rudy_1991"))
    with pytest.raises(ConfigError, match="not available"):
        load_ini(str(bad))


def _with_ode_section(tmp_path, body):
    """The cuboid test .ini with its [ode_solver] section replaced by `body`."""
    import re
    ini, act, fib = case_files(str(tmp_path), (8, 8, 2), "tt2006", 1.0, 5)
    text = open(ini).read()
    text = re.sub(r"\[ode_solver\][^\[]*", "[ode_solver]\n" + body + "\n", text)
    p = tmp_path / "ode.ini"
    p.write_text(text)
    return str(p)


def test_ode_solver_keys_and_defaults(tmp_path):
    """[ode_solver] as the reference reads it (config_parser.c:297-311,
    1645-1668; configure_ode_solver_from_options, ode_solver.c:306-336;
    monodomain_solver.c:510-519):
      - no dt: fixed-step runs step 0.01 ms (0.02 ms PDE step -> 2 ODE steps);
      - adaptive without dt (or auto_dt): min dt 1e-10, max dt = dt_pde;
      - adaptive tolerances are `abstol` / `reltol` (default 1e-5 each);
      - dt above dt_pde: the loop steps dt, ALPHA keeps the configured dt_pde."""
    from monoalg3d_c_amd.inputs import load_ini
    lib = "library_file=shared_libs/libten_tusscher_2006.so"
    d = load_ini(_with_ode_section(tmp_path, lib)).desc
    assert d["dt_ode"] == 0.01 and d["dt"] == 0.02 and "adaptive" not in d and "alpha_dt" not in d
    crn = "library_file=shared_libs/libcourtemanche_ramirez_nattel_1998.so"
    d = load_ini(_with_ode_section(tmp_path, crn + "\nadaptive=true")).desc
    assert d["dt_ode"] == 1e-10 and d["adaptive"] == (1e-5, 1e-5)
    d = load_ini(_with_ode_section(tmp_path, crn + "\nadaptive=true\ndt=0.001\nabstol=1e-7\nreltol=1e-6")).desc
    assert d["dt_ode"] == 0.001 and d["adaptive"] == (1e-7, 1e-6)
    d = load_ini(_with_ode_section(tmp_path, crn + "\nadaptive=true\ndt=0.001\nauto_dt=true")).desc
    assert d["dt_ode"] == 1e-10
    sim = load_ini(_with_ode_section(tmp_path, lib + "\ndt=0.05"))
    assert sim.desc["dt"] == 0.05 and sim.desc["dt_ode"] == 0.05 and sim.desc["alpha_dt"] == 0.02
    assert sim.dt_pde == 0.05


def test_splitmix64_known_answer():
    """SplitMix64 (Steele, Lea & Flood 2014): the first outputs from state 0 are
    0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F; the uniform draw
    is the top 53 bits."""
    import monoalg3d_c_amd as m
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = m.splitmix64_uniform(0, np.arange(3))
    np.testing.assert_array_equal(got, [(w >> 11) * 2.0 ** -53 for w in want])
    # stream position, not call order, decides a cell's draw
    np.testing.assert_array_equal(m.splitmix64_uniform(7, np.arange(10, 20)), m.splitmix64_uniform(7, np.arange(30))[10:20])
    mask = m.fibrosis_mask((40, 30, 20), 0.1, 3, chunk=1000)
    np.testing.assert_array_equal(mask, (m.splitmix64_uniform(3, np.arange(24000)) >= 0.1).astype(np.uint8))
    assert abs((1 - mask.mean()) - 0.1) < 0.01


def test_mesh_and_fibre_writers_read_back(tmp_path):
    """The reference-format inputs written for a box with holes read back to the
    same active cells and, per listed cell, the layered fibre of its depth."""
    import monoalg3d_c_amd as m
    from monoalg3d_c_amd.inputs import read_custom_mesh, read_fibres, write_custom_mesh, write_layered_fibres
    shape = (9, 7, 5)
    act = m.fibrosis_mask(shape, 0.2, 11)
    mesh, fib = str(tmp_path / "mesh.alg"), str(tmp_path / "fib.txt")
    count = write_custom_mesh(mesh, shape, 100.0, act)
    write_layered_fibres(fib, shape, -60.0, 60.0, act)
    assert count == act.sum()
    sh, active, line = read_custom_mesh(mesh, count, 100.0)
    f = read_fibres(fib)
    assert len(f) == count
    # the mesh's box is the bounding box of the listed cells
    nx, ny, nz = sh
    lex_back = np.nonzero(active)[0]
    xs, ys, zs = lex_back % nx, (lex_back // nx) % ny, lex_back // (nx * ny)
    want = np.nonzero(act)[0]
    np.testing.assert_array_equal(np.sort(xs + shape[0] * (ys + shape[1] * zs)), want)
    for k, c in enumerate(want):
        z = c // (shape[0] * shape[1])
        th = np.deg2rad(-60.0 + 120.0 * z / (shape[2] - 1))
        np.testing.assert_allclose(f[k][:3], [np.cos(th), np.sin(th), 0.0], atol=1e-15)
