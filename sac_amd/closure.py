"""ctypes binding of the learned hyperdiffusion closure (sac_closure_*,
include/sac_amd.h) on a grid handle (sac_amd.device.Device).

Typical use: collect (stencil, exact setnu ratio) samples from states of the
plain solver, train the MLP on the GPU, then enable it so that advance()
uses the predicted ratios wherever the shock sensor stays below its
threshold.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from .device import SLOT_W, SacError, load_library

EXPORTS = [
    "sac_closure_nparam", "sac_closure_init", "sac_closure_set_params",
    "sac_closure_get_params", "sac_closure_collect", "sac_closure_collect_from",
    "sac_closure_clear",
    "sac_closure_train", "sac_closure_grad", "sac_closure_enable",
]

N_FEAT, N_TARG = 32, 8

_declared = False


def _lib():
    global _declared
    L = load_library()
    if not _declared:
        I, D, V = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        LL = ctypes.c_longlong
        fp = ctypes.POINTER(ctypes.c_float)
        sig = {
            "sac_closure_nparam": (I, []),
            "sac_closure_init": (I, [V, LL, I, ctypes.c_uint]),
            "sac_closure_set_params": (I, [V, fp, I]),
            "sac_closure_get_params": (I, [V, fp, I]),
            "sac_closure_collect": (I, [V, I, ctypes.POINTER(LL)]),
            "sac_closure_collect_from": (I, [V, V, I, ctypes.POINTER(LL)]),
            "sac_closure_clear": (I, [V]),
            "sac_closure_train": (I, [V, I, I, D, LL, LL, ctypes.POINTER(D)]),
            "sac_closure_grad": (I, [V, fp, fp, I, fp, ctypes.POINTER(D)]),
            "sac_closure_enable": (I, [V, I, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _declared = True
    return L


def _check(rc):
    if rc != 0:
        raise SacError(f"libsac_amd closure error {rc}: {load_library().sac_dev_last_error().decode()}")


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def nparam():
    return _lib().sac_closure_nparam()


class Closure:
    """The learned closure of one Device."""

    def __init__(self, dev, capacity=0, max_blocks=256, seed=1):
        self.dev = dev
        self.max_blocks = max_blocks
        _check(_lib().sac_closure_init(dev.handle, capacity, max_blocks, seed))
        self.n = 0

    @property
    def params(self):
        p = np.zeros(nparam(), dtype=np.float32)
        _check(_lib().sac_closure_get_params(self.dev.handle, _fp(p), p.size))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_lib().sac_closure_set_params(self.dev.handle, _fp(p), p.size))

    def collect(self, slot=SLOT_W):
        n = ctypes.c_longlong()
        _check(_lib().sac_closure_collect(self.dev.handle, slot, ctypes.byref(n)))
        self.n = n.value
        return self.n

    def collect_from(self, src, slot=SLOT_W):
        """Samples of another Device's state (FP64, same GPU) into this buffer."""
        n = ctypes.c_longlong()
        _check(_lib().sac_closure_collect_from(self.dev.handle, src.handle, slot, ctypes.byref(n)))
        self.n = n.value
        return self.n

    def clear(self):
        _check(_lib().sac_closure_clear(self.dev.handle))
        self.n = 0

    def train(self, steps, blocks=256, lr=1e-3, base=0, stride=None):
        """Adam steps over the collected samples; returns the last step's
        mean squared error."""
        if stride is None:
            stride = 1000003
            while math.gcd(stride, self.n) != 1:
                stride += 2
        loss = ctypes.c_double()
        _check(_lib().sac_closure_train(self.dev.handle, steps, blocks, lr, base, stride,
                                        ctypes.byref(loss)))
        return loss.value

    def grad(self, feat, targ):
        """(loss, gradient) of one pass over explicit samples."""
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        targ = np.ascontiguousarray(targ, dtype=np.float32)
        n = feat.shape[0]
        g = np.zeros(nparam(), dtype=np.float32)
        loss = ctypes.c_double()
        _check(_lib().sac_closure_grad(self.dev.handle, _fp(feat), _fp(targ), n, _fp(g),
                                       ctypes.byref(loss)))
        return loss.value, g

    def enable(self, on=True, sensor=0.1, exact_through_closure=False):
        """Use the network's ratios in advance (on); with
        exact_through_closure the exact ratios take the same path (a check
        of the learned path's storage and indexing)."""
        mode = (2 if exact_through_closure else 1) if on else 0
        _check(_lib().sac_closure_enable(self.dev.handle, mode, sensor))


def constant_params(value):
    """Parameters of the network that ignores its inputs and predicts
    `value` (ratio/6) for every output: the closure without a learned part."""
    p = np.zeros(nparam(), dtype=np.float32)
    off_b4 = 64 * 32 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 16 * 64
    p[off_b4:off_b4 + 7] = value
    return p
