"""Mask-formed anisotropic rows against the oracle (CPU).

The inner FP32 solve of the mixed-precision path forms 27-point anisotropic
rows from a 4-byte neighbourhood mask per cell (monoalg3d_c_amd/csrc/aniso_row.h)
instead of reading stored coefficients.  Built for the host, the same function
in FP64 must reproduce the oracle's assembled rows (oracle/assembly.c,
assembly_common.c:115-830) bit for bit; the operator builder repeats this
check on the device against the assembled rows before it enables the mask.
The fast form the kernels use (pair counts by popcount, one reciprocal per
cross term) must agree with those rows to 4e-6 of each row's diagonal in FP32
(the inner solve), keeping their zero pattern, and to 1e-13 in FP64 (the
outer residuals of the mixed-precision solve).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


def test_mask_rows_match_oracle_bitwise(tmp_path):
    if not os.path.exists(HIPCC) or shutil.which("gcc") is None:
        pytest.skip("hipcc / gcc not available")
    src = os.path.join(ROOT, "tests", "mask_rows_check.cpp")
    asm_o, chk_o, exe = (str(tmp_path / n) for n in ("assembly.o", "chk.o", "chk"))
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-c", os.path.join(ROOT, "oracle", "assembly.c"), "-o", asm_o,
                    "-I", os.path.join(ROOT, "oracle")], check=True)
    subprocess.run([HIPCC, "-O2", "-ffp-contract=off", "-std=c++17", "-I", os.path.join(ROOT, "monoalg3d_c_amd", "csrc"),
                    "-I", os.path.join(ROOT, "include"), "-c", src, "-o", chk_o], check=True, capture_output=True)
    subprocess.run([HIPCC, chk_o, asm_o, "-o", exe], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count(" 0 mismatching coefficients; fp32 form: 0 outside") == 6
    assert r.stdout.count("fast fp64 form: 0 outside") == 6
