"""The ionic models' device source compiled for the host, against the oracle.

monoalg3d_c_amd/csrc/ionic_models.h is host-and-device code; built here for
the host (tests/host/host_models.cpp, hipcc host pass) it runs the same FP64
arithmetic as the GPU kernels except for the host's log/pow/expm1.  The
oracle is pinned bit for bit to the reference's compiled models (test_oracle),
so this checks the device code's restatement before any GPU run:
  - its exp (Cody-Waite reduction + Taylor series) within 1 ulp of glibc's;
  - one step from states along an action potential within 1e-12 relative
    per state; a 2000-step trace of two stimulated cells within 1e-9 of each
    state's range.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host", "host_models.cpp")
DT = 0.02


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("host") / "host_models.so")
    subprocess.run([hipcc, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out, SRC],
                   check=True, capture_output=True)
    L = C.CDLL(out)
    L.host_step.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_double]
    L.host_exp.argtypes = [C.c_double]
    L.host_exp.restype = C.c_double
    return L


def test_exp_within_one_ulp(host):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-700, 700, 20000), rng.uniform(-1, 1, 20000), rng.uniform(-1e-6, 1e-6, 2000),
                         [0.0, 1.0, -1.0, 708.0, -708.0, 0.5 * np.log(2), -0.5 * np.log(2)]])
    got = np.array([host.host_exp(float(x)) for x in xs])
    ref = np.exp(xs)
    ulp = np.abs(got - ref) / np.spacing(np.abs(ref))
    assert ulp.max() <= 1.0, (ulp.max(), xs[ulp.argmax()])
    assert host.host_exp(0.0) == 1.0


def _stim(nsteps, n, amp):
    s = np.zeros((nsteps, n))
    t = 0.0
    for k in range(nsteps):
        if 0.0 <= t <= 2.0:
            s[k] = amp * np.linspace(1.0, 0.4, n)
        t += DT
    return s


MODELS = [("tt2006", 0, -50.0), ("ord2011", 1, -50.0), ("ms2003", 2, 0.2)]


@pytest.mark.parametrize("name,mid,amp", MODELS)
def test_one_step_pairs(host, oracle, name, mid, amp):
    om = oracle.MODEL_IDS[name]
    sv = oracle.model_init(om, 1)
    stim = _stim(6000, 1, amp)
    samples, stims = [], []
    for k in range(6000):
        oracle.model_solve(om, sv, stim[k], DT, 1)
        if k % 37 == 0:
            samples.append(sv[0].copy())
            stims.append(amp if k % 2 else 0.0)
    X = np.ascontiguousarray(np.array(samples))
    S = np.ascontiguousarray(np.array(stims))
    ref = X.copy()
    oracle.model_solve(om, ref, S, DT, 1)
    got = X.copy()
    assert host.host_step(mid, got.ctypes.data, S.ctypes.data, len(S), DT) == 0
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    rel[np.abs(ref - got) < 1e-300] = 0.0
    assert rel.max() < 1e-12, (rel.max(), np.unravel_index(rel.argmax(), rel.shape))


@pytest.mark.parametrize("name,mid,amp", MODELS)
def test_trace(host, oracle, name, mid, amp):
    om = oracle.MODEL_IDS[name]
    n, nsteps = 2, 2000
    stim = _stim(nsteps, n, amp)
    ref = oracle.model_init(om, n)
    got = np.ascontiguousarray(ref.copy())
    worst = 0.0
    for k in range(nsteps):
        oracle.model_solve(om, ref, stim[k], DT, 1)
        s = np.ascontiguousarray(stim[k])
        assert host.host_step(mid, got.ctypes.data, s.ctypes.data, n, DT) == 0
        scale = np.maximum(np.abs(ref).max(axis=0), 1e-30)
        worst = max(worst, float((np.abs(got - ref) / scale).max()))
    assert worst < 1e-9, worst
