"""Shared helpers: build an oracle instance from a Problem, run the main
loop on either side, and compare states per variable."""
from __future__ import annotations

import numpy as np

from sac_amd import problems
from sac_amd.errors import mesh, per_var_errors, per_var_maxabs, rel_l2  # noqa: F401


def make_oracle(prob, oracle_mod):
    return oracle_mod.from_problem(prob)


def oracle_run(prob, oracle_mod, steps, typeadvance="fourstep"):
    o = make_oracle(prob, oracle_mod)
    dts = [o.step(prob.courantpar, typeadvance) for _ in range(steps)]
    return o.get_w(), dts, o


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


def ulp_perturbed(prob, seed):
    """prob.w with every evolving mesh value moved by -1, 0 or +1 ulp (drawn
    per value from `seed`): a start that differs from prob.w by rounding alone"""
    ne = problems.var_index(prob.ndir)["nevolve"]
    w = prob.w.copy()
    ms = (slice(0, ne),) + prob.mesh_slices()
    m = w[ms]
    k = np.random.default_rng(seed).integers(-1, 2, size=m.shape)
    w[ms] = np.where(k > 0, np.nextafter(m, np.inf), np.where(k < 0, np.nextafter(m, -np.inf), m))
    return w
