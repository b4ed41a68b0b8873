"""mono_cg_csr_device_noalloc: the CSR solve on device operands with the
library's fixed device workspace, against mono_cg_csr (the same solve with
buffers it allocates) on the same operators, bit for bit: two calls in a row,
the second on a smaller operator (the workspace keeps the first call's
contents), with and without Jacobi.  Then the refusals: more rows than the
workspace holds (host check, no device work) and malformed operands (device
check, before the solve reads through them).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


def _operator(shape, seed):
    """Implicit-Euler diffusion rows on a box (7-point, varying conductances),
    diagonal alpha + sum of the face couplings: symmetric positive definite."""
    nx, ny, nz = shape
    n = nx * ny * nz
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.5, 1.5, (n, 3)) * 1.3e-4 * 100.0
    rows, cols, vals = [], [], []
    for c in range(n):
        x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
        diag = 0.07
        ent = []
        for axis, (dx, dy, dz) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
            for s in (-1, 1):
                xx, yy, zz = x + s * dx, y + s * dy, z + s * dz
                if 0 <= xx < nx and 0 <= yy < ny and 0 <= zz < nz:
                    o = xx + nx * (yy + ny * zz)
                    w = 2.0 * g[c, axis] * g[o, axis] / (g[c, axis] + g[o, axis])
                    ent.append((o, -w))
                    diag += w
        ent.append((c, diag))
        ent.sort()
        for o, v in ent:
            rows.append(c)
            cols.append(o)
            vals.append(v)
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, np.array(rows) + 1, 1)
    rp = np.cumsum(rp).astype(np.int64)
    b = rng.uniform(-6.0, 6.0, n)
    x0 = rng.uniform(-90.0, 20.0, n)
    return rp, np.array(cols, np.int32), np.array(vals), b, x0


def _hip():
    """The HIP runtime the library itself loaded (a second runtime in the same
    process, such as torch's, would not see the device)."""
    _lib()
    for line in open("/proc/self/maps"):
        if "libamdhip64.so" in line and "/" in line:
            H = C.CDLL(line.split()[-1])
            break
    else:
        raise RuntimeError("libamdhip64 not mapped after loading the library")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    return H


def _noalloc(L, rp, col, val, b, x0, tol, max_its, jacobi):
    H = _hip()
    arrs = [np.ascontiguousarray(a) for a in (rp, col, val, b, x0)]
    dev = []
    try:
        for a in arrs:
            p = C.c_void_p()
            assert H.hipMalloc(C.byref(p), a.nbytes) == 0
            dev.append(p)
            assert H.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        its, err = C.c_int(), C.c_double()
        rc = L.mono_cg_csr_device_noalloc(len(rp) - 1, dev[0], dev[1], dev[2], dev[3], dev[4], C.c_double(tol), max_its,
                                          int(jacobi), C.byref(its), C.byref(err), None)
        x = np.empty_like(arrs[4])
        assert H.hipMemcpy(x.ctypes.data, dev[4], x.nbytes, 2) == 0  # device to host
        return rc, x, its.value, err.value
    finally:
        for p in dev:
            H.hipFree(p)


def _lib():
    from monoalg3d_c_amd import _lib
    L = _lib.lib()
    L.mono_cg_csr_device_noalloc.argtypes = [C.c_int] + [C.c_void_p] * 5 + [C.c_double, C.c_int, C.c_int,
                                                                             C.POINTER(C.c_int), C.POINTER(C.c_double),
                                                                             C.c_void_p]
    L.mono_cg_csr_device_noalloc.restype = C.c_int
    return L


@pytest.mark.parametrize("jacobi", [False, True])
def test_noalloc_equals_allocating_solve(jacobi):
    import monoalg3d_c_amd as m
    L = _lib()
    for shape, seed in (((40, 30, 12), 1), ((17, 9, 5), 2)):
        announce(f"no-alloc CSR CG vs mono_cg_csr, {shape}, jacobi={jacobi}")
        rp, col, val, b, x0 = _operator(shape, seed)
        want, its_w, err_w = m.cg_csr(rp, col, val, b, x0, tol=1e-20, max_its=400, jacobi=jacobi)
        rc, got, its_g, err_g = _noalloc(L, rp, col, val, b, x0, 1e-20, 400, jacobi)
        assert rc == 0, L.mono_last_error()
        assert its_g == its_w and its_w > 3, (its_g, its_w)
        assert err_g == err_w, (err_g, err_w)
        np.testing.assert_array_equal(got, want)


def test_noalloc_refuses_too_many_rows():
    L = _lib()
    announce("no-alloc CSR CG refuses n > max rows")
    its, err = C.c_int(), C.c_double()
    # the row count is checked before any pointer is used
    one = C.c_void_p(1)
    rc = L.mono_cg_csr_device_noalloc(MAX_ROWS + 1, one, one, one, one, one, 1e-16, 10, 0, C.byref(its), C.byref(err), None)
    assert rc != 0
    assert b"exceed" in L.mono_last_error()


def test_noalloc_refuses_malformed_operands():
    L = _lib()
    announce("no-alloc CSR CG refuses malformed device operands")
    rp, col, val, b, x0 = _operator((6, 5, 4), 3)
    bad_rp = rp.copy()
    bad_rp[3], bad_rp[4] = bad_rp[4], bad_rp[3]  # decreasing, same total
    rc, _, _, _ = _noalloc(L, bad_rp, col, val, b, x0, 1e-16, 50, False)
    assert rc != 0 and b"row_ptr" in L.mono_last_error()
    bad_col = col.copy()
    bad_col[7] = len(rp) - 1  # == n
    rc, _, _, _ = _noalloc(L, rp, bad_col, val, b, x0, 1e-16, 50, False)
    assert rc != 0 and b"column" in L.mono_last_error()
    # a good call still works after the refusals
    rc, _, its, _ = _noalloc(L, rp, col, val, b, x0, 1e-16, 50, False)
    assert rc == 0 and its > 1
