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


def zero_boundary_problem(ndim):
    """A problem with 'zero' boundaries (src/vacgrid.t:294-345) in 1-D (a
    Riemann tube, zero on the left), 2-D (a Riemann patch, zero along x) or
    3-D (the stratified box, zero along x and z)."""
    if ndim == 1:
        prob = problems.riemann_tube(40, {"rho": 1.0, "p": 0.8, "v": [0.1, -0.2, 0.3],
                                          "b": [0.6, 0.5, -0.4]},
                                     {"rho": 0.3, "p": 0.2, "v": [-0.1, 0.2, 0.0],
                                      "b": [0.6, -0.5, 0.4]})
        prob.typeB[:, :] = [problems.BC["zero"], problems.BC["cont"]]
        prob.typeB[1, 1] = problems.BC["zero"]
    elif ndim == 2:
        prob = problems.riemann_patch(14, 12, {"rho": 1.0, "p": 0.8, "v": [0.1, -0.2, 0.3],
                                              "b": [0.6, 0.5, -0.4]},
                                      {"rho": 0.3, "p": 0.2, "v": [-0.1, 0.2, 0.0],
                                       "b": [0.6, 0.5, 0.4]})
        prob.typeB[:, :] = [problems.BC[c] for c in ("zero", "zero", "periodic", "periodic")]
    else:
        # the background's ghosts given with the state (as an ini file of the
        # whole grid gives them): with zero background ghosts the first
        # fill's conversions set the momentum of the ghosts filled before the
        # background's to zero, and the stratified run blows up on both sides
        prob = problems.stratified(n=10, nxyz=(10, 9, 8), amp=3e-2)
        ms = prob.mesh_slices()
        for iw in range(8, 13):
            prob.w[iw] = np.pad(prob.w[iw][ms], 2, mode="edge")
        prob.typeB[:, :] = [problems.BC[c] for c in ("zero", "cont", "periodic", "periodic",
                                                    "zero", "zero")]
    return prob
