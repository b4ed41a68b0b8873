"""Sweep files (the reference's [batch]/[modify] format) without a GPU: the run
list, its order and numbering, range values, output directory names, and the
grouping of runs into batch atlases."""
import os
import sys

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_golden import case_files  # noqa: E402


def test_range_values():
    """range|start|end|inc: values while <= end on the accumulated double, %f text."""
    from monoalg3d_c_amd.sweep import range_values
    assert range_values(0.5, 1.5, 0.25) == ["0.500000", "0.750000", "1.000000", "1.250000", "1.500000"]
    # 0.1 + 0.1 + 0.1 > 0.3 in double: the end value is not reached
    assert range_values(0.1, 0.3, 0.1) == ["0.100000", "0.200000"]


def test_run_order_and_numbers():
    """Cartesian product with the last directive varying fastest, repeated
    num_simulations_per_parameter_change times, run numbers from 1."""
    from monoalg3d_c_amd.sweep import generate_runs
    mod = [("domain", "cable_length", ["100", "200"]), ("extra_data", "GKr_multiplicator", ["0.5", "1.0", "1.5"])]
    runs = generate_runs(mod, 2)
    assert [r.run_number for r in runs] == list(range(1, 13))
    vals = [tuple(v for _, _, v in r.params) for r in runs[:6]]
    assert vals == [("100", "0.5"), ("100", "1.0"), ("100", "1.5"), ("200", "0.5"), ("200", "1.0"), ("200", "1.5")]
    assert [r.params for r in runs[6:]] == [r.params for r in runs[:6]]


def _sweep_files(tmp_path, modify):
    ini, _, _ = case_files(str(tmp_path / "base"), (300, 1, 1), "tt2006", 5.0, 50, cable=True)
    b = tmp_path / "sweep.ini"
    b.write_text("[batch]\ninitial_config=" + ini + "\noutput_folder=" + str(tmp_path / "out") +
                 "\nnum_simulations_per_parameter_change=1\n[modify]\n" + modify)
    return str(b), ini


def test_plan_groups_runs(tmp_path):
    from monoalg3d_c_amd.sweep import plan
    b, ini = _sweep_files(tmp_path, "domain|cable_length=list|10000|15000\nextra_data|GKr_multiplicator=list|0.5|1.5\n")
    groups = plan(b)
    assert len(groups) == 1 and len(groups[0]) == 4  # lengths and factors ride in one atlas
    r = groups[0][3]
    assert r.sections["domain"]["cable_length"] == "15000" and r.sections["extra_data"]["GKr_multiplicator"] == "1.5"
    base_out = str(tmp_path / "base")
    assert r.sections["save_result"]["output_dir"] == (str(tmp_path / "out") + "/" + base_out +
                                                       "_run_4_cable_length_15000_GKr_multiplicator_1.5")
    b2, _ = _sweep_files(tmp_path, "main|dt_pde=list|0.02|0.01\n")
    g2 = plan(b2)
    assert len(g2) == 2  # a different dt is a different launch
    # [main] values are parsed as float (strtof) by the reference
    assert float(g2[1][0].sections["main"]["dt_pde"]) == float(np.float32(0.01))


def test_bad_directives(tmp_path):
    from monoalg3d_c_amd.inputs import ConfigError
    from monoalg3d_c_amd.sweep import read_batch_ini
    b, _ = _sweep_files(tmp_path, "domain|cable_length=sequence|1|2\n")
    with pytest.raises(ConfigError):
        read_batch_ini(b)
