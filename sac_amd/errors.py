"""Per-variable error measures between two states in the reference layout
(w of shape (nw, G3, G2, G1)), taken over the mesh cells: the figures every
result reports against the FP64 oracle or the FP64 GPU run."""
from __future__ import annotations

import numpy as np


def mesh(prob, a):
    """the mesh cells of a (nw, G3, G2, G1) array of problem prob"""
    return a[(Ellipsis,) + prob.mesh_slices()]


def rel_l2(a, b):
    """rel-L2 of a against b; absolute L2 when b is identically zero"""
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def per_var_errors(prob, a, b, nvar=None):
    """rel-L2 of each of the first nvar variables of a against b"""
    nvar = nvar or a.shape[0]
    ma, mb = mesh(prob, a), mesh(prob, b)
    return [rel_l2(ma[i], mb[i]) for i in range(nvar)]


def per_var_maxabs(prob, a, b, nvar=None):
    """max |a - b| over the mesh of each of the first nvar variables"""
    nvar = nvar or a.shape[0]
    ma, mb = mesh(prob, a), mesh(prob, b)
    return [float(np.abs(ma[i] - mb[i]).max()) for i in range(nvar)]
