"""ctypes binding of the tube-ensemble entry points of libsac_amd.so
(sac_batch_*, include/sac_amd.h): many independent 1-D MHD tubes of uneven
length, or 2-D MHD patches of uneven size, advanced together on one GPU,
each with its own dt, t and it.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import problems
from .device import SacError, _dp, _ip, grid_dx, load_library

EXPORTS = [
    "sac_batch_init", "sac_batch_init_nd", "sac_batch_free", "sac_batch_tune", "sac_batch_shape",
    "sac_batch_set_shape",
    "sac_batch_run", "sac_batch_status", "sac_batch_download", "sac_batch_memory",
    "sac_batch_compactions", "sac_batch_last_error",
]

STATE = {0: "running", 1: "finished", 2: "failed", 3: "dt below dtmin"}

_declared = False


def _lib():
    global _declared
    L = load_library()
    if not _declared:
        I, D, V = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        llp = ctypes.POINTER(ctypes.c_longlong)
        sig = {
            "sac_batch_init": (I, [I, ip, I, dp, dp, dp, I, I, D, D, I, ctypes.POINTER(V)]),
            "sac_batch_init_nd": (I, [I, I, ip, I, dp, dp, dp, I, I, D, D, I, ctypes.POINTER(V)]),
            "sac_batch_shape": (I, [V, ip, ip, ip]),
            "sac_batch_set_shape": (I, [V, I]),
            "sac_batch_free": (I, [V]),
            "sac_batch_tune": (I, [V, I, ip, ip, dp]),
            "sac_batch_run": (I, [V, I, ctypes.c_longlong, llp]),
            "sac_batch_status": (I, [V, dp, ip, ip]),
            "sac_batch_download": (I, [V, dp]),
            "sac_batch_memory": (I, [V, llp]),
            "sac_batch_compactions": (I, [V, ip, ip]),
            "sac_batch_last_error": (ctypes.c_char_p, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _declared = True
    return L


def _check(rc):
    if rc != 0:
        raise SacError(f"libsac_amd batch error {rc}: {_lib().sac_batch_last_error().decode()}")


def riemann_ensemble(count=4096, nmin=200, nmax=2000, seed=4096, nu=1.0):
    """The seeded tube ensemble as Problems (gamma 5/3, cont boundaries)."""
    return [problems.riemann_tube(n, l, r, gamma=5.0 / 3.0, nu=nu)
            for n, l, r in problems.tube_ensemble(count, nmin, nmax, seed)]


def patch_ensemble(count=1024, nmin=32, nmax=160, seed=2048, nu=1.0):
    """The seeded 2-D patch ensemble as Problems (gamma 5/3, cont boundaries)."""
    return [problems.riemann_patch(nx, ny, l, r, gamma=5.0 / 3.0, nu=nu)
            for nx, ny, l, r in problems.patch_ensemble(count, nmin, nmax, seed)]


class Ensemble:
    """Independent 1-D tubes or 2-D patches (ndir=3) on one device."""

    def __init__(self, tubes, courantpar=0.3, tmax=0.1, itmax=1000000):
        L = _lib()
        self.tubes = tubes
        ndim = tubes[0].ndim
        nw = tubes[0].nw
        for t in tubes:
            if t.ndim != ndim or ndim not in (1, 2) or t.ndir != 3:
                raise SacError("an ensemble holds 1-D tubes or 2-D patches, all with ndir=3")
        self.ndim = ndim
        self.ncell = np.ascontiguousarray([t.nx[d] for t in tubes for d in range(ndim)],
                                          dtype=np.int32)
        # per tube w(G1[,G2], nw) column-major == prob.w in C order
        w = np.concatenate([np.ascontiguousarray(t.w.reshape(nw, -1)).ravel() for t in tubes])
        dx = np.concatenate([grid_dx(t) for t in tubes])
        eq = np.ascontiguousarray(tubes[0].eqpar, dtype=np.float64)
        h = ctypes.c_void_p()
        _check(L.sac_batch_init_nd(ndim, len(tubes), _ip(self.ncell), nw, _dp(w), _dp(dx),
                                   _dp(eq), len(eq), problems.NU, courantpar, tmax, itmax,
                                   ctypes.byref(h)))
        self._h = h
        self.nw = nw

    def close(self):
        if getattr(self, "_h", None):
            _lib().sac_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, trial_steps=2):
        ck, tb, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(_lib().sac_batch_tune(self._h, trial_steps, ctypes.byref(ck), ctypes.byref(tb),
                                     ctypes.byref(ms)))
        return ck.value, tb.value, ms.value

    def shape(self):
        """Current tile: (cells per thread, extent along x, extent along y)."""
        ck, tx, ty = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(_lib().sac_batch_shape(self._h, ctypes.byref(ck), ctypes.byref(tx),
                                      ctypes.byref(ty)))
        return ck.value, tx.value, ty.value

    def set_shape(self, index):
        """Use launch shape `index` of the autotuner's list without timing."""
        _check(_lib().sac_batch_set_shape(self._h, index))

    def run(self, steps_per_graph=32, max_steps=10 ** 7):
        n = ctypes.c_longlong()
        _check(_lib().sac_batch_run(self._h, steps_per_graph, max_steps, ctypes.byref(n)))
        return n.value

    def status(self):
        k = len(self.tubes)
        t = np.zeros(k)
        it = np.zeros(k, dtype=np.int32)
        st = np.zeros(k, dtype=np.int32)
        _check(_lib().sac_batch_status(self._h, _dp(t), _ip(it), _ip(st)))
        return t, it, st

    def download(self):
        """Final state of every tube, shaped like its Problem's w."""
        total = int(sum(t.w.size for t in self.tubes))
        w = np.zeros(total)
        _check(_lib().sac_batch_download(self._h, _dp(w)))
        out, off = [], 0
        for t in self.tubes:
            sz = t.w.size
            out.append(w[off:off + sz].reshape(t.w.shape))
            off += sz
        return out

    def memory(self):
        b = ctypes.c_longlong()
        _check(_lib().sac_batch_memory(self._h, ctypes.byref(b)))
        return b.value

    def compactions(self):
        """(tile rebuilds so far, tubes the tiles cover now)"""
        n, c = ctypes.c_int(), ctypes.c_int()
        _check(_lib().sac_batch_compactions(self._h, ctypes.byref(n), ctypes.byref(c)))
        return n.value, c.value


def _main():
    """Run the seeded tube and patch ensembles (default tiles) for a few
    steps with direct kernel launches (steps_per_graph = -1), e.g. under a
    kernel tracer."""
    import argparse
    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--tubes", type=int, default=4096)
    ap.add_argument("--patches", type=int, default=1024)
    args = ap.parse_args()
    for name, probs in (("tubes", riemann_ensemble(count=args.tubes)),
                        ("patches", patch_ensemble(count=args.patches))):
        e = Ensemble(probs, courantpar=0.3, tmax=0.1)
        n = e.run(steps_per_graph=-1, max_steps=args.steps)
        ck, tx, ty = e.shape()
        e.close()
        print(f"{name}: {len(probs)} problems, {n} steps, tile {tx}x{ty}", flush=True)


if __name__ == "__main__":
    _main()
