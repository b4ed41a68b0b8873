"""bench.py's JSON line is assembled from run results on the CPU, so a
formatting mistake shows here rather than at the end of a GPU run."""
import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _run(rate, ms, class_operator=1, cells=16000000):
    return dict(rate=rate, ms_per_step=ms, cells=cells, cg_its=9.0, inner_its=8.0, outer=3.0, ionic_ms=0.93,
                solve_ms=1.4, setup_s=0.1, class_operator=class_operator, stencil_points=27)


def test_full_line_is_json():
    args = argparse.Namespace(steps=40, warmup=10)
    res = {"fp64": _run(3.5e9, 4.5), "fp32": _run(6.0e9, 2.6)}
    res3 = {"fp64": _run(1.0e9, 14.0, 0, 14398217), "fp32": _run(2.0e9, 7.0, 0, 14398217)}
    errors = {"config2": {"max_abs_lat_ms": 0.001}}
    cpu = {"value": 6e6, "unit": "cell-timesteps/s", "cores": 16, "kind": "reference", "sample": "s"}
    line = bench.make_line(args, (400, 400, 100), 41, res, res3, errors, cpu, {"device_resident": {}})
    text = json.dumps(line)
    back = json.loads(text)
    assert back["value"] == 6.0e9 and back["dtype"] == "f32"
    assert "10% of the cells" in back["fibrotic"]["workload"]
    r = back["roofline"]
    assert r["bound"] == "hbm" and r["peak"] == 8000.0 and 0 < r["frac"] < 1
    assert abs(r["achieved"] - r["bytes_per_launch"] / (r["launch_ms"] * 1e-3) / 1e9) < 0.1
    for k in ("metric", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "data", "config", "cpu_baseline"):
        assert k in back


def test_partial_line_is_json():
    args = argparse.Namespace(steps=5, warmup=1)
    line = bench.make_line(args, (64, 64, 16), 41, {"fp64": _run(3.5e9, 4.5, cells=65536)}, {}, {}, None, None)
    back = json.loads(json.dumps(line))
    assert back["dtype"] == "f64" and back["fibrotic"] is None and back["speedup_fp32_refine_vs_fp64"] is None


def test_algorithmic_bytes_per_cell_step():
    """ORd (41 equations): ionic 2 x 40 states x 8 B + V in + b out = 656 B
    (FP64) and 2 x 40 x 4 + 16 = 336 B (FP32).  FP64 CG: residual 32 B, 88 B
    per iteration; per-cell rows add 19 x 8 B per stencil application, a
    neighbourhood mask 4 B.  FP32+refine: 20 B per outer residual, 20 B per
    correction after the first, 44 B per inner iteration."""
    assert bench.bytes_per_cell_step(41, 8, 8, False, 1) == 656 + 32 + 704
    assert bench.bytes_per_cell_step(41, 8, 8, False, 1, 19) == 656 + 32 + 704 + 9 * 152
    assert bench.bytes_per_cell_step(41, 8, 8, False, 1, 0, True) == 656 + 32 + 704 + 9 * 4
    assert bench.bytes_per_cell_step(41, 4, 8, True, 3) == 336 + 60 + 40 + 352
    assert bench.bytes_per_cell_step(41, 4, 8, True, 3, 0, True) == 336 + 60 + 40 + 352 + 11 * 4
    assert bench.bytes_per_cell_step(41, 4, 8, True, 3, 19) == 336 + 60 + 3 * 152 + 40 + 352 + 8 * 76


def test_mask_mode_counts_mask_bytes():
    args = argparse.Namespace(steps=40, warmup=10)
    res = {"fp64": _run(3.5e9, 4.5), "fp32": _run(6.0e9, 2.6)}
    res3 = {"fp64": _run(2.0e9, 7.0, 3, 14398217), "fp32": _run(2.9e9, 5.0, 3, 14398217)}
    line = bench.make_line(args, (400, 400, 100), 41, res, res3, {}, None, None)
    rf = line["fibrotic"]["phase_roofline"]
    # _run: 8 inner iterations, 3 outer sweeps
    assert rf["fp64"]["step_bytes_per_cell"] == 656 + 32 + 8 * 88 + 9 * 4
    assert rf["fp32"]["step_bytes_per_cell"] == 336 + 60 + 40 + 8 * 44 + 11 * 4
