"""ctypes binding of libsac_amd.so (include/sac_amd.h).

This is the Python side of the drop-in boundary, used by tests and by
bench.py.  It never falls back to CPU code: a missing library or a missing
GPU raises SacError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import problems

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsac_amd.so")

SLOT_W, SLOT_S1, SLOT_S2, SLOT_S3 = 0, 1, 2, 3
K_FINAL, K_AXIS, K_GHOST, K_CMAX, K_COPY = 0, 1, 2, 3, 4
K_AXIS1, K_AXIS2, K_AXIS3 = 5, 6, 7
K_CLOSURE, K_PLANE = 8, 9
ADV = {"onestep": 1, "twostep": 2, "threestep": 3, "fourstep": 4, "sterck": 5, "jameson": 6}
PREC = {"fp64": 0, "mixed": 1}

EXPORTS = [
    "sac_dev_init", "sac_dev_free", "sac_dev_upload", "sac_dev_download",
    "sac_dev_upload_slot", "sac_dev_download_slot", "sac_dev_getboundary",
    "sac_dev_getdt_courant", "sac_dev_getdt_special", "sac_dev_advance",
    "sac_dev_advect1", "sac_dev_set_divbfix", "sac_dev_set_plane_tiles", "sac_dev_set_resistive", "sac_dev_copy_slot", "sac_dev_get_maxviscoef", "sac_dev_set_maxviscoef",
    "sac_dev_timing", "sac_dev_timing_get", "sac_dev_synchronize",
    "sac_dev_memory", "sac_dev_save_begin", "sac_dev_save_ready", "sac_dev_save_end",
    "sac_getdt_courant_noalloc", "sac_dev_last_error",
]


class SacError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libsac_amd.so and declare its signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SacError(f"{path} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(path)
    I, D, V = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "sac_dev_init": (I, [I, I, ip, I, ip, I, dp, I, I, I, dp, ip, I, I, I,
                             ctypes.POINTER(V)]),
        "sac_dev_free": (I, [V]),
        "sac_dev_upload": (I, [V, dp]),
        "sac_dev_download": (I, [V, dp]),
        "sac_dev_upload_slot": (I, [V, I, dp]),
        "sac_dev_download_slot": (I, [V, I, dp]),
        "sac_dev_getboundary": (I, [V]),
        "sac_dev_getdt_courant": (I, [V, D, dp, dp]),
        "sac_dev_getdt_special": (I, [V, dp]),
        "sac_dev_advance": (I, [V, D]),
        "sac_dev_advect1": (I, [V, D, I, I]),
        "sac_dev_set_divbfix": (I, [V, I]),
        "sac_dev_set_plane_tiles": (I, [V, I]),
        "sac_dev_set_resistive": (I, [V, I, D]),
        "sac_dev_copy_slot": (I, [V, I, I]),
        "sac_dev_get_maxviscoef": (I, [V, dp]),
        "sac_dev_set_maxviscoef": (I, [V, D]),
        "sac_dev_timing": (I, [V, I]),
        "sac_dev_timing_get": (I, [V, I, dp, ctypes.POINTER(ctypes.c_longlong)]),
        "sac_dev_synchronize": (I, [V]),
        "sac_dev_memory": (I, [V, ctypes.POINTER(ctypes.c_longlong)]),
        "sac_dev_save_begin": (I, [V]),
        "sac_dev_save_ready": (I, [V, ip]),
        "sac_dev_save_end": (I, [V, dp]),
        "sac_getdt_courant_noalloc": (I, [I, I, ip, I, V, V, V, D, D, V, dp, dp]),
        "sac_dev_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _check(rc):
    if rc != 0:
        msg = load_library().sac_dev_last_error().decode()
        raise SacError(f"libsac_amd error {rc}: {msg}")


def grid_dx(prob) -> np.ndarray:
    """Per-axis cell widths after gridsetup1 (src/vacgrid.t:578-629) for a
    tensor grid: linear extrapolation of the mesh coordinates into the ghost
    layers, then dx_i=(x_{i+1}-x_{i-1})/2.  Returned concatenated over axes,
    each axis covering its whole grid extent."""
    out = []
    for d in range(prob.ndim):
        # coordinate d along its own axis (other indices at the first mesh cell)
        idx = [2 if ax < prob.ndim else 0 for ax in (2, 1, 0)]
        sl = list(idx)
        sl[2 - d] = slice(None)
        xs = np.array(prob.x[d][tuple(sl)], dtype=np.float64)
        G = xs.shape[0]
        q = np.zeros(G + 2)
        q[1:-1] = xs
        # upper layers: mesh end is grid index G-3 (0-based), q offset +1
        ixe = G - 3 + 1
        for ix in range(ixe + 1, G + 2):
            a = abs(ixe - ix)
            q[ix] = (1 + a) * q[ixe] - a * q[ixe - 1]
        ixe = 2 + 1
        for ix in range(0, ixe):
            a = abs(ixe - ix)
            q[ix] = (1 + a) * q[ixe] - a * q[ixe + 1]
        out.append(0.5 * (q[2:] - q[:-2]))
    return np.concatenate(out)


class Device:
    """One problem on the GPU behind the C ABI."""

    def __init__(self, prob, typeadvance="fourstep", precision="fp64", dx=None):
        L = load_library()
        self.prob = prob
        self.ndim, self.ndir = prob.ndim, prob.ndir
        self.nw = prob.nw
        self.shape = prob.w.shape
        nx = np.ascontiguousarray(prob.nx, dtype=np.int32)
        iws = np.ascontiguousarray(prob.iws(), dtype=np.int32)
        eq = np.ascontiguousarray(prob.eqpar, dtype=np.float64)
        self.dx = np.ascontiguousarray(grid_dx(prob) if dx is None else dx, dtype=np.float64)
        tb = np.ascontiguousarray(prob.typeB, dtype=np.int32).reshape(-1)
        h = ctypes.c_void_p()
        _check(L.sac_dev_init(prob.ndim, prob.ndir, _ip(nx), prob.nw, _ip(iws), len(iws),
                              _dp(eq), len(eq), problems.GRAV0, problems.NU, _dp(self.dx),
                              _ip(tb), 0, ADV[typeadvance], PREC[precision],
                              ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load_library().sac_dev_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def upload(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.shape == self.shape
        _check(load_library().sac_dev_upload(self._h, _dp(w)))

    def download(self):
        out = np.zeros(self.shape)
        _check(load_library().sac_dev_download(self._h, _dp(out)))
        return out

    def save_begin(self):
        """Start an asynchronous copy of the state to host memory."""
        _check(load_library().sac_dev_save_begin(self._h))

    def save_ready(self):
        r = ctypes.c_int()
        _check(load_library().sac_dev_save_ready(self._h, ctypes.byref(r)))
        return bool(r.value)

    def save_end(self):
        """Wait for the copy started by save_begin; the state at that time."""
        out = np.zeros(self.shape)
        _check(load_library().sac_dev_save_end(self._h, _dp(out)))
        return out

    def upload_slot(self, slot, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.shape == self.shape
        _check(load_library().sac_dev_upload_slot(self._h, slot, _dp(w)))

    def download_slot(self, slot):
        out = np.zeros(self.shape)
        _check(load_library().sac_dev_download_slot(self._h, slot, _dp(out)))
        return out

    def getboundary(self):
        _check(load_library().sac_dev_getboundary(self._h))

    def getdt_courant(self, courantpar):
        dt = ctypes.c_double()
        cm = ctypes.c_double()
        _check(load_library().sac_dev_getdt_courant(self._h, courantpar, ctypes.byref(dt),
                                                    ctypes.byref(cm)))
        return dt.value, cm.value

    def getdt_special(self, dt):
        v = ctypes.c_double(dt)
        _check(load_library().sac_dev_getdt_special(self._h, ctypes.byref(v)))
        return v.value

    def advance(self, dt):
        _check(load_library().sac_dev_advance(self._h, dt))

    def advect1(self, qdt, wct_slot, w_slot):
        _check(load_library().sac_dev_advect1(self._h, qdt, wct_slot, w_slot))

    def set_resistive(self, compactres=False, dtdiffpar=0.5):
        """Resistive source form and dtdiffpar (eta = eqpar[1] > 0)."""
        _check(load_library().sac_dev_set_resistive(self._h, 1 if compactres else 0, dtdiffpar))

    def set_plane_tiles(self, on=True):
        """3-D stages: x and y work in one pass over z-plane tiles (default)
        or one line-tile pass per axis."""
        _check(load_library().sac_dev_set_plane_tiles(self._h, 1 if on else 0))

    def set_divbfix(self, on=True):
        """Powell's div B sources in every advect1 (ndim > 1)."""
        _check(load_library().sac_dev_set_divbfix(self._h, 1 if on else 0))

    def copy_slot(self, dst, src):
        _check(load_library().sac_dev_copy_slot(self._h, dst, src))

    @property
    def maxviscoef(self):
        v = ctypes.c_double()
        _check(load_library().sac_dev_get_maxviscoef(self._h, ctypes.byref(v)))
        return v.value

    @maxviscoef.setter
    def maxviscoef(self, v):
        _check(load_library().sac_dev_set_maxviscoef(self._h, v))

    def step(self, courantpar):
        """One main-loop iteration (src/vac.t:60-104): dt, then advance."""
        dt, _ = self.getdt_courant(courantpar)
        dt = self.getdt_special(dt)
        self.advance(dt)
        return dt

    def timing(self, enable):
        _check(load_library().sac_dev_timing(self._h, int(enable)))

    def timing_get(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        _check(load_library().sac_dev_timing_get(self._h, kernel, ctypes.byref(ms),
                                                 ctypes.byref(n)))
        return ms.value, n.value

    def synchronize(self):
        _check(load_library().sac_dev_synchronize(self._h))

    def memory(self):
        b = ctypes.c_longlong()
        _check(load_library().sac_dev_memory(self._h, ctypes.byref(b)))
        return b.value


def run(prob, steps=None, tmax=None, typeadvance="fourstep", dtmin=1.0e-99):
    """Upload, getboundary, then the main loop (src/vac.t:55-106) until
    itmax, tmax or dt < dtmin (timetofinish, src/vac.t:585-605); returns the
    final state and the list of dt of the steps taken."""
    dev = Device(prob, typeadvance=typeadvance)
    dev.upload(prob.w)
    dev.getboundary()
    t, it, dts = 0.0, 0, []
    steps = prob.itmax if steps is None else steps
    tmax = prob.tmax if tmax is None else tmax
    while True:
        dt, _ = dev.getdt_courant(prob.courantpar)
        dt = dev.getdt_special(dt)
        if it >= steps or t >= tmax or dt < dtmin:
            break
        dev.advance(dt)
        dts.append(dt)
        t += dt
        it += 1
    w = dev.download()
    dev.close()
    return w, dts
