"""sac_getdt_courant_noalloc: getdt_courant on a state the caller keeps in
device memory, with its reduction slots in a static device array (no device
allocation in the call).  Checked through the C ABI against
sac_dev_getdt_courant on the same state, on two calls in a row, the second
with a smaller grid (the slots keep the first call's values, so they must be
cleared by the call), in FP64 and in mixed precision."""
import ctypes

import numpy as np
import pytest

from sac_amd import device, problems

pytestmark = pytest.mark.gpu


def _hip():
    """the HIP runtime the library uses (device buffers of the caller)"""
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def _to_device(hip, a):
    a = np.ascontiguousarray(a)
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
    return p


def _noalloc(prob, precision):
    L = device.load_library()
    hip = _hip()
    ne = problems.var_index(prob.ndir)["nevolve"]
    dtype = np.float64 if precision == "fp64" else np.float32
    bufs = [_to_device(hip, prob.w[:ne].astype(dtype)), _to_device(hip, prob.w[ne:]),
            _to_device(hip, device.grid_dx(prob))]
    nx = np.array(list(prob.nx) + [1] * (3 - prob.ndim), dtype=np.int32)
    dt, cm = ctypes.c_double(), ctypes.c_double()
    try:
        rc = L.sac_getdt_courant_noalloc(prob.ndim, prob.ndir, device._ip(nx),
                                         device.PREC[precision], bufs[0], bufs[1], bufs[2],
                                         float(prob.eqpar[0]), prob.courantpar, None,
                                         ctypes.byref(dt), ctypes.byref(cm))
        assert rc == 0, L.sac_dev_last_error()
    finally:
        for b in bufs:
            hip.hipFree(b)
    return dt.value, cm.value


def _handle(prob, precision):
    dev = device.Device(prob, precision=precision)
    dev.upload(prob.w)
    out = dev.getdt_courant(prob.courantpar)
    w = dev.download()
    dev.close()
    return out, w


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_getdt_equals_handle_getdt(precision):
    big = problems.stratified(n=24, nxyz=(24, 20, 18), amp=3e-2)
    small = problems.brio_wu(n=64)
    for prob in (big, small):
        (dt_h, cm_h), w = _handle(prob, precision)
        prob.w = w  # the state the handle held (FP32-rounded in mixed precision)
        dt_n, cm_n = _noalloc(prob, precision)
        assert dt_n == dt_h and cm_n == cm_h, (prob.name, dt_n, dt_h, cm_n, cm_h)


def test_noalloc_rejects_bad_sizes():
    L = device.load_library()
    nx = np.array([4, 1, 1], dtype=np.int32)
    dt = ctypes.c_double()
    rc = L.sac_getdt_courant_noalloc(1, 1, device._ip(nx), 0, ctypes.c_void_p(8),
                                     ctypes.c_void_p(8), ctypes.c_void_p(8), 2.0, 0.3, None,
                                     ctypes.byref(dt), None)
    assert rc == 1 and b"nx[0]=4" in L.sac_dev_last_error()
    nx = np.array([1 << 13, 1 << 13, 1 << 13], dtype=np.int32)
    rc = L.sac_getdt_courant_noalloc(3, 3, device._ip(nx), 0, ctypes.c_void_p(8),
                                     ctypes.c_void_p(8), ctypes.c_void_p(8), 2.0, 0.3, None,
                                     ctypes.byref(dt), None)
    assert rc == 1 and b"cells" in L.sac_dev_last_error()
