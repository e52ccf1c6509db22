"""mlfma_rcs_device_noalloc (the allocation-free RCS) against mlfma_rcs through the C ABI.

Two calls in a row on one GPU, the second on a smaller mesh with fewer angles (both use the
GPU's one scratch array), each compared with the allocating entry point on the same currents:
total dBsm within 1e-9 dB, the V and H powers within 1e-9 of the total power (the
cross-polarised part can sit hundreds of dB down, where dB values are ill-conditioned).
A request beyond the scratch size is rejected before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

from mlfma_pmchwt_amd import api, geometry

pytestmark = pytest.mark.gpu


_hip = None


def _hip_runtime():
    """The HIP runtime libmlfma.so itself links (device buffers of the caller, no torch)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipFree.argtypes = [C.c_void_p]
    return _hip


def _device(a):
    hip = _hip_runtime()
    a = np.ascontiguousarray(a)
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
    return p


def _noalloc_rcs(op, I, th, ph):
    lib = api.library()
    lib.mlfma_rcs_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    hip = _hip_runtime()
    dI, dth, dph = _device(I.astype(np.complex128)), _device(th.astype(np.float64)), _device(ph.astype(np.float64))
    dout = _device(np.zeros((th.size, 3)))
    rc = lib.mlfma_rcs_device_noalloc(op._h, dI, th.size, dth, dph, dout)
    out = np.zeros((th.size, 3))
    assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), dout, out.nbytes, 2) == 0   # device to host
    for p in (dI, dth, dph, dout):
        hip.hipFree(p)
    return rc, out


def _check(a, b):
    assert np.max(np.abs(a[:, 0] - b[:, 0])) < 1e-9
    tot = 10.0 ** (b[:, 0] / 10.0)
    for c in (1, 2):
        assert np.max(np.abs(10.0 ** (a[:, c] / 10.0) - 10.0 ** (b[:, c] / 10.0)) / tot) < 1e-9


@pytest.mark.parametrize("first,second", [((api.PMCHWT, 10, 181), (api.CFIE, 6, 37))])
def test_noalloc_rcs_matches_rcs_two_calls(first, second):
    rng = np.random.default_rng(7)
    for ie, freq, nang in (first, second):
        mesh = geometry.geodesic_sphere(freq, 0.8)
        print(f"[noalloc] ie={ie} geodesic {freq} angles {nang}", flush=True)
        op = api.Operator(mesh, integral_equation=ie, eps_r_re=3.0, eps_r_im=-0.2)
        I = rng.standard_normal(op.n) + 1j * rng.standard_normal(op.n)
        th = np.linspace(0.0, 180.0, nang)
        ph = np.full(nang, 30.0)
        rc, got = _noalloc_rcs(op, I, th, ph)
        assert rc == 0, api.library().mlfma_last_error(op._h)
        ref = op.rcs(I, th, ph, pol=0.0)
        _check(got, ref)
        op.close()


def test_noalloc_rcs_rejects_beyond_scratch():
    op = api.Operator(geometry.geodesic_sphere(10, 0.8), integral_equation=api.PMCHWT, eps_r_re=3.0)
    nang = 20000   # 20000 x (24 + 48 * 12) B = 12 MB > 8 MiB
    th = np.zeros(nang)
    rc, _ = _noalloc_rcs(op, np.zeros(op.n, dtype=np.complex128), th, th)
    assert rc == 1
    assert b"scratch" in api.library().mlfma_last_error(op._h)
    op.close()
