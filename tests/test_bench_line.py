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
