"""ctypes binding of the C oracle (oracle/sac_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke()
and by bench.py's cpu_baseline leg, always as the checker and never as the
thing measured or shipped.

Arrays use the reference layout: w has shape (nw, G3, G2, G1) in C order,
which is the memory of the Fortran array w(ixG1,ixG2,ixG3,nw).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SAC_ORACLE_LIB") or os.path.join(_HERE, "_build", "libsac_oracle.so")

BC = {"cont": 1, "fixed": 2, "periodic": 3, "symm": 4, "symm0": 5, "asymm": 6,
      "cont1": 7, "fixed1": 8, "grad1": 9, "contCD4": 10}
ADV = {"onestep": 1, "twostep": 2, "threestep": 3, "fourstep": 4,
       "sterck": 5, "jameson": 6}

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with gcc into oracle/_build (plain C, -O2, no
    fast-math, no FMA contraction so that the Fortran evaluation order is
    kept)."""
    if os.environ.get("SAC_ORACLE_LIB"):
        return LIB_PATH  # a prebuilt variant chosen by the caller
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    src = os.path.join(_HERE, "sac_oracle.c")
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(
            ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
             "-shared", "-o", LIB_PATH, src, "-lm"])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        D = ctypes.c_double
        I = ctypes.c_int
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        sig = {
            "oracle_new": (P, [I, I, ip, dp, I, I, I, ip, I]),
            "oracle_free": (None, [P]),
            "oracle_nw": (I, [P]),
            "oracle_ncell": (ctypes.c_long, [P]),
            "oracle_set_x": (I, [P, dp]),
            "oracle_get_dx": (I, [P, dp]),
            "oracle_set_w": (I, [P, dp]),
            "oracle_get_w": (I, [P, dp]),
            "oracle_set_iws": (I, [P, ip, I]),
            "oracle_set_divbfix": (None, [P, I]),
            "oracle_set_resistive": (None, [P, I, D]),
            "oracle_getdt_res": (D, [P, D]),
            "oracle_getboundary": (None, [P]),
            "oracle_getboundary_on": (None, [P, dp]),
            "oracle_getdt_courant": (D, [P, D, dp]),
            "oracle_getdt_special": (D, [P, D]),
            "oracle_maxviscoef": (D, [P]),
            "oracle_set_maxviscoef": (None, [P, D]),
            "oracle_advance": (I, [P, D, I]),
            "oracle_advect1_on": (None, [P, D, dp, dp]),
            "oracle_step": (D, [P, D, I]),
            "oracle_set_time": (None, [P, I, I, D]),
            "oracle_time": (D, [P]),
            "oracle_it": (I, [P]),
            "oracle_gradient4": (None, [P, dp, I, dp]),
            "oracle_gradient1": (None, [P, dp, I, I, dp]),
            "oracle_getflux": (None, [P, dp, I, I, dp, ip]),
            "oracle_getcmax": (None, [P, dp, I, dp]),
            "oracle_setnu": (None, [P, dp, I, I, dp, dp]),
            "oracle_addsource_visc": (None, [P, D, dp, dp]),
            "oracle_addsource_grav": (None, [P, D, dp, dp]),
            "oracle_centdiff4": (None, [P, D, dp, dp]),
            "oracle_poison_scratch": (None, [P, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Oracle:
    """One problem instance of the CPU restatement."""

    def __init__(self, ndim, ndir, nx, eqpar, typeB, grav0_=3, nu_=6,
                 fullgridini=False, iws=None):
        self.ndim, self.ndir = ndim, ndir
        self.nx = tuple(int(v) for v in nx)
        self.G = tuple(n + 4 for n in self.nx) + (1,) * (3 - ndim)
        self.shape = self.G[::-1]  # (G3, G2, G1)
        eq = np.ascontiguousarray(eqpar, dtype=np.float64)
        tb = np.ascontiguousarray(typeB, dtype=np.int32).reshape(-1)
        nxa = np.ascontiguousarray(self.nx, dtype=np.int32)
        self._h = lib().oracle_new(ndim, ndir, _ip(nxa), _dp(eq), len(eq),
                                   grav0_, nu_, _ip(tb), int(fullgridini))
        if not self._h:
            raise ValueError("oracle_new rejected the problem")
        self.nw = lib().oracle_nw(self._h)
        self.n = lib().oracle_ncell(self._h)
        if iws is not None:
            a = np.ascontiguousarray(iws, dtype=np.int32)
            if lib().oracle_set_iws(self._h, _ip(a), len(a)) != 0:
                raise ValueError("bad iws")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_free(h)
            self._h = None

    # -- state ---------------------------------------------------------
    def field_shape(self):
        return self.shape

    def set_x(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.ndim * self.n
        lib().oracle_set_x(self._h, _dp(x))

    def dx(self):
        out = np.zeros((self.ndim,) + self.shape)
        lib().oracle_get_dx(self._h, _dp(out))
        return out

    def set_w(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.nw * self.n
        lib().oracle_set_w(self._h, _dp(w))

    def get_w(self):
        out = np.zeros((self.nw,) + self.shape)
        lib().oracle_get_w(self._h, _dp(out))
        return out

    # -- algorithm -----------------------------------------------------
    def getboundary(self):
        lib().oracle_getboundary(self._h)

    def getboundary_on(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64).copy()
        lib().oracle_getboundary_on(self._h, _dp(w))
        return w

    def getdt_courant(self, courantpar):
        cm = ctypes.c_double(0.0)
        dt = lib().oracle_getdt_courant(self._h, courantpar, ctypes.byref(cm))
        return dt, cm.value

    def getdt_special(self, dt_in):
        return lib().oracle_getdt_special(self._h, dt_in)

    @property
    def maxviscoef(self):
        return lib().oracle_maxviscoef(self._h)

    @maxviscoef.setter
    def maxviscoef(self, v):
        lib().oracle_set_maxviscoef(self._h, v)

    def set_resistive(self, compactres=False, dtdiffpar=0.5):
        lib().oracle_set_resistive(self._h, 1 if compactres else 0, dtdiffpar)

    def getdt_res(self, dt):
        return lib().oracle_getdt_res(self._h, dt)

    def set_divbfix(self, on=True):
        lib().oracle_set_divbfix(self._h, 1 if on else 0)

    def advance(self, dt, typeadvance="fourstep"):
        r = lib().oracle_advance(self._h, dt, ADV[typeadvance])
        if r != 0:
            raise ValueError(typeadvance)

    def advect1(self, qdt, wCT, w):
        wCT = np.ascontiguousarray(wCT, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64).copy()
        lib().oracle_advect1_on(self._h, qdt, _dp(wCT), _dp(w))
        return w

    def step(self, courantpar, typeadvance="fourstep"):
        return lib().oracle_step(self._h, courantpar, ADV[typeadvance])

    def set_time(self, it, itmin, t):
        lib().oracle_set_time(self._h, it, itmin, t)

    @property
    def t(self):
        return lib().oracle_time(self._h)

    @property
    def it(self):
        return lib().oracle_it(self._h)

    # -- building blocks -------------------------------------------------
    def gradient4(self, q, idim):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros(self.shape)
        lib().oracle_gradient4(self._h, _dp(q), idim, _dp(out))
        return out

    def gradient1(self, q, idim, kind):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros(self.shape)
        lib().oracle_gradient1(self._h, _dp(q), idim, kind, _dp(out))
        return out

    def getflux(self, w, iw, idim):
        w = np.ascontiguousarray(w, dtype=np.float64)
        out = np.zeros(self.shape)
        tr = ctypes.c_int(0)
        lib().oracle_getflux(self._h, _dp(w), iw, idim, _dp(out), ctypes.byref(tr))
        return out, bool(tr.value)

    def getcmax(self, w, idim):
        w = np.ascontiguousarray(w, dtype=np.float64)
        out = np.zeros(self.shape)
        lib().oracle_getcmax(self._h, _dp(w), idim, _dp(out))
        return out

    def setnu(self, w, iw, idim):
        w = np.ascontiguousarray(w, dtype=np.float64)
        nuR = np.zeros(self.shape)
        nuL = np.zeros(self.shape)
        lib().oracle_setnu(self._h, _dp(w), iw, idim, _dp(nuR), _dp(nuL))
        return nuR, nuL

    def addsource_visc(self, qdt, w, wnew):
        w = np.ascontiguousarray(w, dtype=np.float64)
        wnew = np.ascontiguousarray(wnew, dtype=np.float64).copy()
        lib().oracle_addsource_visc(self._h, qdt, _dp(w), _dp(wnew))
        return wnew

    def addsource_grav(self, qdt, w, wnew):
        w = np.ascontiguousarray(w, dtype=np.float64)
        wnew = np.ascontiguousarray(wnew, dtype=np.float64).copy()
        lib().oracle_addsource_grav(self._h, qdt, _dp(w), _dp(wnew))
        return wnew

    def centdiff4(self, qdt, wCT, w):
        wCT = np.ascontiguousarray(wCT, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64).copy()
        lib().oracle_centdiff4(self._h, qdt, _dp(wCT), _dp(w))
        return w

    def poison_scratch(self, value):
        lib().oracle_poison_scratch(self._h, value)

    def mesh(self, a):
        """Mesh view (drop two ghost layers on active axes) of a field or
        stack of fields."""
        sl = []
        for d in range(3):  # shape order (G3, G2, G1)
            ax = 2 - d  # reference axis index 0..2
            sl.append(slice(2, -2) if ax < self.ndim else slice(None))
        return a[(Ellipsis,) + tuple(sl)]


def from_problem(prob):
    """An oracle set up with a sac_amd Problem (grid, eqpar, boundaries,
    state), with getboundary applied and the clock at it=itmin=0, t=0."""
    from sac_amd import problems
    o = Oracle(prob.ndim, prob.ndir, prob.nx, prob.eqpar, prob.typeB,
               grav0_=problems.GRAV0, nu_=problems.NU, iws=prob.iws())
    o.set_x(prob.x.reshape(-1))
    o.set_w(prob.w)
    o.getboundary()
    o.set_time(0, 0, 0.0)
    return o
