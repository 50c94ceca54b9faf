"""Closed-form pins of the oracle's three time-step limits.

  getdt_courant  src/vac.t:641-699: dt = courantpar / max_{mesh,d}(cmax_d/dx_d)
                 with cmax_d the fast magnetosonic speed along d plus |v_d|;
  getdt_visc     src/vacusr.viscosity.t:509-541: dt = dtdiffpar/max(maxviscoef/dx^2)
                 with dtdiffpar forced to 0.25 (:530), maxviscoef reset to 0
                 afterwards (:539), so a step with no setnu call since the last
                 reset has no viscous limit;
  getdt_grav     src/vacusr.gravity.t:97-138: dtgrav = 1/sqrt(max_d |g|/dx_d),
                 evaluated only at it == itmin and kept for later steps.

Each expected value is computed here from the mathematics (a textbook fast
speed, the grid spacing, the gravity parameter), never from the oracle's
own intermediate arrays, so a changed constant (0.25, the 1 of 1/sqrt, a
factor on courantpar) or a dropped |v| fails one of them.
"""
import numpy as np
import pytest

from sac_amd import problems
from tests import mhd_textbook as tb
from tests.helpers import make_oracle, mesh
from tests.test_oracle_pins import _blank


def _random_state(prob, v, seed, bmag=0.8, vmax=0.4):
    """positive density and pressure, nonzero velocity and split field"""
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = prob.w.shape[1:]
    gamma = prob.eqpar[0]
    ndir = prob.ndir
    rho_b = 0.5 + rng.uniform(size=shape)
    rho_p = 0.3 * rng.uniform(size=shape)
    rho = rho_b + rho_p
    vel = [vmax * rng.uniform(-1, 1, size=shape) for _ in range(ndir)]
    p_b = 0.5 + rng.uniform(size=shape)
    p_p = 0.2 * rng.uniform(-1, 1, size=shape)
    B = [bmag * rng.uniform(-1, 1, size=shape) for _ in range(ndir)]
    bg = [0.5 * b for b in B]
    bp = [b - g for b, g in zip(B, bg)]
    w = prob.w
    w[0] = rho_p
    w[v["rhob"] - 1] = rho_b
    for c in range(ndir):
        w[c + 1] = rho * vel[c]
        w[v["b0"] + c] = bp[c]
        w[v["bg0"] + c] = bg[c]
    w[v["eb"] - 1] = p_b / (gamma - 1) + 0.5 * sum(g * g for g in bg)
    w[v["e"] - 1] = (p_p / (gamma - 1) +
                     0.5 * (rho * sum(u * u for u in vel) + sum(a * a for a in bp)) +
                     sum(a * b for a, b in zip(bp, bg)))
    return rho, vel, p_b + p_p, B


@pytest.mark.parametrize("ndim,ndir,nx", [(1, 2, (40,)), (2, 3, (12, 10)), (3, 3, (9, 7, 6))])
def test_getdt_courant_closed_form(oracle_mod, ndim, ndir, nx):
    prob, v = _blank(ndim, ndir, nx, bc="cont", lengths=tuple(1.0 + 0.5 * d for d in range(ndim)))
    rho, vel, p, B = _random_state(prob, v, seed=11 + ndim)
    gamma = prob.eqpar[0]
    o = make_oracle(prob, oracle_mod)
    cmax_over_dx = 0.0
    for d in range(ndim):
        dx = (1.0 + 0.5 * d) / nx[d]
        c = tb.fast_speed(rho, p, B, d, gamma) + np.abs(vel[d])
        cmax_over_dx = max(cmax_over_dx, float(mesh(prob, c).max()) / dx)
    for courantpar in (0.3, 0.8):
        dt, cm = o.getdt_courant(courantpar)
        assert cm == pytest.approx(cmax_over_dx, rel=1e-12)
        assert dt == pytest.approx(courantpar / cmax_over_dx, rel=1e-12)


def test_getdt_courant_uniform_tube(oracle_mod):
    """uniform 1-D state: dt = courantpar dx / (c_f + |v|) exactly as the
    textbook gives it (B along and across the tube, flow along it)"""
    prob, v = _blank(1, 2, (50,), bc="cont", gamma=2.0)
    rho, u, p, b1, b2 = 0.7, -0.35, 0.4, 0.6, -0.9
    w = prob.w
    w[0] = rho
    w[1] = rho * u
    w[v["b0"]] = b1
    w[v["b0"] + 1] = b2
    w[v["e"] - 1] = p / (2.0 - 1.0) + 0.5 * (rho * u * u + b1 * b1 + b2 * b2)
    o = make_oracle(prob, oracle_mod)
    cs2 = 2.0 * p / rho
    a2 = (b1 * b1 + b2 * b2) / rho
    cf = np.sqrt(0.5 * (a2 + cs2 + np.sqrt((a2 + cs2) ** 2 - 4.0 * b1 * b1 / rho * cs2)))
    dt, _ = o.getdt_courant(0.3)
    assert dt == pytest.approx(0.3 * (1.0 / 50) / (cf + abs(u)), rel=1e-13)


@pytest.mark.parametrize("ndim,nx", [(1, (40,)), (3, (9, 7, 6))])
def test_getdt_visc_closed_form_and_reset(oracle_mod, ndim, nx):
    """dt = 0.25 min(dx)^2 / maxviscoef, maxviscoef is 0 afterwards and the
    next call (no setnu since) leaves dt unlimited"""
    ndir = 2 if ndim == 1 else 3
    lengths = tuple(1.0 + 0.5 * d for d in range(ndim))
    prob, v = _blank(ndim, ndir, nx, bc="cont", nu=1.0, lengths=lengths)
    prob.w[v["rhob"] - 1] = 1.0
    prob.w[v["e"] - 1] = 1.5
    o = make_oracle(prob, oracle_mod)
    dxmin = min(lengths[d] / nx[d] for d in range(ndim))
    # in 3-D the switch slot nu_=6 is also grav3_ (src/vacusrpar.f:4-5): the
    # stored 1e-30 acts as g_3 and gives the only other limit
    free = 1.0 / np.sqrt(1e-30 / dxmin) if ndim == 3 else 1e30
    for mv in (3.7e-2, 1.25):
        o.maxviscoef = mv
        dt = o.getdt_special(1e30)
        assert dt == pytest.approx(0.25 * dxmin ** 2 / mv, rel=1e-14)
        assert o.maxviscoef == 0.0
        assert o.getdt_special(1e30) == pytest.approx(free, rel=1e-14)
    # a larger limit from outside is kept
    o.maxviscoef = 1e-12
    assert o.getdt_special(1e-3) == 1e-3


def test_maxviscoef_is_the_largest_setnu_coefficient(oracle_mod):
    """the accumulator getdt_visc reads is the largest nuR/nuL of the setnu
    calls since the reset (src/vacusr.viscosity.t:429,450); on a unit step in
    v_1 setnu's known answer is c_tot c_hyp 2 dx (ratio 2 at the jump)"""
    n = 40
    prob, v = _blank(1, 2, (n,), bc="cont", nu=1.0, gamma=5.0 / 3.0)
    w = prob.w
    w[0] = 1.0
    w[v["e"] - 1] = 1.0 / (5.0 / 3.0 - 1.0)  # p=1, rho=1, B=0
    ms = prob.mesh_slices()
    xm = np.arange(n)
    w[1][ms] = np.where(xm < n // 2, 0.0, 1.0)  # m_1 = v_1 step
    w[v["e"] - 1][ms] += 0.5 * w[1][ms] ** 2
    o = make_oracle(prob, oracle_mod)
    o.maxviscoef = 0.0
    w0 = o.get_w()
    nuR, nuL = o.setnu(w0, 2, 1)  # momentum m_1 along x
    cs = np.sqrt(5.0 / 3.0)
    c_tot = float(mesh(prob, cs + np.abs(w[1])).max())  # B=0: cmax = c_s + |v|
    expect = c_tot * 0.4 * 2.0 * (1.0 / n)
    assert float(max(nuR.max(), nuL.max())) == pytest.approx(expect, rel=1e-13)
    assert o.maxviscoef == pytest.approx(expect, rel=1e-13)
    dt = o.getdt_special(1e30)
    assert dt == pytest.approx(0.25 * (1.0 / n) ** 2 / expect, rel=1e-13)


def test_first_step_has_no_viscous_limit_then_lagged(oracle_mod):
    """step 1 runs on the courant limit alone (maxviscoef starts at 0); the
    advance fills maxviscoef, and step 2's dt is min(courant, 0.25 dx^2/maxviscoef)
    with the value collected during step 1 (src/vac.t:60-88)"""
    prob = problems.brio_wu(n=120)
    o = make_oracle(prob, oracle_mod)
    dtc, _ = o.getdt_courant(prob.courantpar)
    dt1 = o.step(prob.courantpar)
    assert dt1 == dtc
    mv = o.maxviscoef
    assert mv > 0.0
    dtc2, _ = o.getdt_courant(prob.courantpar)
    dt2 = o.step(prob.courantpar)
    assert dt2 == pytest.approx(min(dtc2, 0.25 * (1.0 / 120) ** 2 / mv), rel=1e-14)


@pytest.mark.parametrize("ndim,nx,axis,g", [(1, (30,), 0, -2.5), (2, (16, 10), 1, 4.0),
                                            (3, (9, 7, 6), 0, -1.0 / 0.2 * 0.6)])
def test_getdt_grav_closed_form_and_kept(oracle_mod, ndim, nx, axis, g):
    """dtgrav = 1/sqrt(|g|/min(dx)) from the first step's getdt_grav; later
    steps keep that value, and a run whose first getdt_special comes after
    itmin has no gravity limit"""
    ndir = 2 if ndim == 1 else 3
    lengths = tuple(1.0 + 0.25 * d for d in range(ndim))
    grav = [0.0, 0.0, 0.0]
    grav[axis] = g
    prob, v = _blank(ndim, ndir, nx, bc="cont", grav=tuple(grav), lengths=lengths)
    prob.w[v["rhob"] - 1] = 1.0
    prob.w[v["e"] - 1] = 1.5
    dxmin = min(lengths[d] / nx[d] for d in range(ndim))
    expect = 1.0 / np.sqrt(abs(g) / dxmin)
    o = make_oracle(prob, oracle_mod)
    o.set_time(0, 0, 0.0)
    assert o.getdt_special(1e30) == pytest.approx(expect, rel=1e-14)
    o.set_time(7, 0, 0.1)
    assert o.getdt_special(1e30) == pytest.approx(expect, rel=1e-14)
    assert o.getdt_special(0.5 * expect) == 0.5 * expect
    o2 = make_oracle(prob, oracle_mod)
    o2.set_time(3, 0, 0.0)
    assert o2.getdt_special(1e30) == 1e30
