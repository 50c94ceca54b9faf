"""CPU checks of the drop-in boundary: the C-ABI library loads without a GPU
and exports every entry point declared in include/sac_amd.h; the Python
binding declares the same set; init fails loudly (no CPU fallback) when no
GPU is present."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sac_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sac_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "sac_dev_init" in syms and "sac_dev_advance" in syms
    assert len(syms) >= 15


def test_library_exports_every_declared_symbol():
    from sac_amd import batch, closure, device
    L = device.load_library()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(device.EXPORTS) | set(batch.EXPORTS) | set(closure.EXPORTS) == set(declared_symbols())


def test_init_fails_loudly_without_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    from sac_amd import device, problems
    with pytest.raises(device.SacError):
        device.Device(problems.brio_wu(n=50))


def test_init_rejects_unsupported_configurations():
    """argument validation happens before any device call"""
    import ctypes
    import numpy as np
    from sac_amd import device
    L = device.load_library()
    h = ctypes.c_void_p()
    nx = np.array([20], dtype=np.int32)
    iws = np.arange(1, 7, dtype=np.int32)
    eq = np.array([2.0, -0.5, 0, 0, 0, 1.0])  # eta < 0 (specialeta) -> not built
    dx = np.full(24, 0.05)
    tb = np.ones(20, dtype=np.int32)
    rc = L.sac_dev_init(1, 2, device._ip(nx), 10, device._ip(iws), 6, device._dp(eq), 6, 3, 6,
                        device._dp(dx), device._ip(tb), 0, 4, 0, ctypes.byref(h))
    assert rc == 2 and b"resistive" in L.sac_dev_last_error()
    eq[1] = 0.0
    rc = L.sac_dev_init(1, 2, device._ip(nx), 11, device._ip(iws), 6, device._dp(eq), 6, 3, 6,
                        device._dp(dx), device._ip(tb), 0, 4, 0, ctypes.byref(h))
    assert rc == 1 and b"nw" in L.sac_dev_last_error()
    # 'zero' fills run region by region without stored values: not together
    # with fixed, fixed1 or grad1; a code past 'zero' is not a boundary type
    tb[:] = 11
    tb[3] = 2
    rc = L.sac_dev_init(1, 2, device._ip(nx), 10, device._ip(iws), 6, device._dp(eq), 6, 3, 6,
                        device._dp(dx), device._ip(tb), 0, 4, 0, ctypes.byref(h))
    assert rc == 2 and b"'zero'" in L.sac_dev_last_error()
    tb[3] = 12
    rc = L.sac_dev_init(1, 2, device._ip(nx), 10, device._ip(iws), 6, device._dp(eq), 6, 3, 6,
                        device._dp(dx), device._ip(tb), 0, 4, 0, ctypes.byref(h))
    assert rc == 2 and b"boundary code 12" in L.sac_dev_last_error()


def test_batch_init_rejects_bad_ensembles():
    """The ensemble entry point validates its arguments before any device call."""
    import ctypes
    import numpy as np
    from sac_amd import batch
    L = batch._lib()
    h = ctypes.c_void_p()
    eq = np.array([5.0 / 3.0, 0, 0, 0, 0, 1.0])
    w = np.zeros(13 * 16 * 16)
    dx = np.full(32, 0.1)
    n = np.array([12, 12], dtype=np.int32)
    args = lambda ndim, nc: (ndim, 1, batch._ip(nc), 13, batch._dp(w), batch._dp(dx),
                             batch._dp(eq), 6, 6, 0.3, 0.1, 10, ctypes.byref(h))
    assert L.sac_batch_init_nd(*args(3, n)) == 2
    assert b"2-D patches" in L.sac_batch_last_error()
    assert L.sac_batch_init_nd(*args(2, np.array([12, 5], dtype=np.int32))) == 1
    assert b"axis 2" in L.sac_batch_last_error()
    assert not h.value
