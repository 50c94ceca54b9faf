"""Shared helpers: build an oracle instance from a Problem, run the main
loop on either side, and compare states per variable."""
from __future__ import annotations

import numpy as np

from sac_amd import problems


def make_oracle(prob, oracle_mod):
    o = oracle_mod.Oracle(prob.ndim, prob.ndir, prob.nx, prob.eqpar, prob.typeB,
                          grav0_=problems.GRAV0, nu_=problems.NU, iws=prob.iws())
    o.set_x(prob.x.reshape(-1))
    o.set_w(prob.w)
    o.getboundary()
    o.set_time(0, 0, 0.0)
    return o


def oracle_run(prob, oracle_mod, steps, typeadvance="fourstep"):
    o = make_oracle(prob, oracle_mod)
    dts = [o.step(prob.courantpar, typeadvance) for _ in range(steps)]
    return o.get_w(), dts, o


def mesh(prob, a):
    return a[(Ellipsis,) + prob.mesh_slices()]


def rel_l2(a, b):
    """rel-L2 of a against b; absolute L2 when b is identically zero"""
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def per_var_errors(prob, a, b, nvar=None):
    nvar = nvar or a.shape[0]
    ma, mb = mesh(prob, a), mesh(prob, b)
    return [rel_l2(ma[i], mb[i]) for i in range(nvar)]


def assert_close_state(prob, got, ref, tol, nvar=None, what=""):
    errs = per_var_errors(prob, got, ref, nvar)
    worst = max(errs)
    assert worst <= tol, f"{what} per-variable rel-L2 {['%.2e' % e for e in errs]} > {tol}"
    return errs


def perturbed(prob, seed=1, amp=1e-3):
    """Same problem with a smooth random perturbation on every evolving
    variable of the mesh (keeps rho_t and pressure positive)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w = prob.w.copy()
    ms = prob.mesh_slices()
    v = problems.var_index(prob.ndir)
    for iw in range(v["nevolve"]):
        base = np.abs(w[iw][ms]).max() or 1.0
        w[iw][ms] += amp * base * rng.standard_normal(w[iw][ms].shape)
    prob.w = w
    return prob
