"""mlq_fftfwdr_noalloc (device arrays, no device allocation: the intermediate
spectrum lives in the library's per-device static scratch) against
mlq_fftfwdr on the same field, two calls in a row (128^3, then 32^3 with the
first call's contents still in the scratch), and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libmaxlq.so is linked against (device buffers for the test)."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _noalloc(g, u):
    hip = _hip()
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros((g.n, g.n, g.n // 2 + 1), np.complex128)
    ud, od = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(ud), u.nbytes) == 0
    assert hip.hipMalloc(C.byref(od), out.nbytes) == 0
    try:
        assert hip.hipMemcpy(ud, u.ctypes.data, u.nbytes, H2D) == 0
        used0 = g.mem_info()[0]
        g._chk(g.lib.mlq_fftfwdr_noalloc(g.ctx, ud, od))
        assert g.mem_info()[0] == used0   # nothing allocated through the context either
        assert hip.hipMemcpy(out.ctypes.data, od, out.nbytes, D2H) == 0
    finally:
        hip.hipFree(ud)
        hip.hipFree(od)
    return out


def test_fftfwdr_noalloc_matches_fftfwdr(monkeypatch):
    from maxlq_amd.lq import MaxLq
    monkeypatch.setenv("MLQ_AUTOTUNE", "0")
    rng = np.random.default_rng(7)
    for n in (128, 32):                      # the second call is smaller
        g = MaxLq(n, q=4.0, precision="fp64")
        u = rng.standard_normal((n, n, n))
        ref = g.fftfwdr(u)                   # the existing entry point (host arrays)
        got = _noalloc(g, u)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"n={n}: max |noalloc - fftfwdr| / max |fftfwdr| = {err:.2e}")
        assert err < 1e-13
        # against numpy's transform (the reference's fftfwdr is the DFT)
        npref = np.fft.rfftn(u, axes=(0, 1, 2))
        assert np.max(np.abs(got - npref)) / np.max(np.abs(npref)) < 1e-13
        g.close()


def test_fftfwdr_noalloc_rejects_large_grids(monkeypatch):
    from maxlq_amd import capi
    from maxlq_amd.lq import MaxLq
    monkeypatch.setenv("MLQ_AUTOTUNE", "0")
    g = MaxLq(256, q=4.0, precision="fp64")
    rc = g.lib.mlq_fftfwdr_noalloc(g.ctx, C.c_void_p(1), C.c_void_p(1))   # rejected before use
    assert rc == capi.MLQ_EINVAL
    assert "n <= 128" in g.lib.mlq_last_error(g.ctx).decode()
    g.close()
