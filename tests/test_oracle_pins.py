"""Pins of the C oracle against what the mathematics fixes.

The reference has no golden data and cannot be built here, so each oracle
routine is checked against an independent statement: closed forms, a
textbook ideal-MHD flux of the total state, hand-computed setnu values,
conservation and symmetry, and the textbook RK4 tableau.  Every check is
chosen so that a dropped term, a wrong sign or a shifted index fails it.
"""
import numpy as np
import pytest

from sac_amd import problems
from tests import mhd_textbook as tb
from tests.helpers import make_oracle, mesh

BC = problems.BC


def _blank(ndim, ndir, nx, gamma=5.0 / 3.0, nu=0.0, grav=(0.0, 0.0, 0.0), bc="periodic",
           lengths=None):
    """Problem with all fields zero and uniform coordinates.  nu>0 switches
    hyperdiffusion on; the stored value is 1e-30 because slot nu_=6 is also
    the axis-3 gravity of the 3-D user parameters (the reference's index
    collision)."""
    nu = 1e-30 if nu else 0.0
    v = problems.var_index(ndir)
    G = tuple(n + 4 for n in nx) + (1,) * (3 - ndim)
    shape = G[::-1]
    w = np.zeros((v["nw"],) + shape)
    x = np.zeros((ndim,) + shape)
    lengths = lengths or (1.0,) * ndim
    for d in range(ndim):
        c = (np.arange(G[d]) - 1.5) * (lengths[d] / nx[d])
        sh = [1, 1, 1]
        sh[2 - d] = G[d]
        x[d] = np.broadcast_to(c.reshape(sh), shape)
    eq = np.zeros(problems.NEQPAR)
    eq[0] = gamma
    eq[problems.NU - 1] = nu
    for d in range(ndim):
        if grav[d]:
            eq[problems.GRAV0 + d] = grav[d]
    t = np.full((v["nw"], 2 * ndim), BC[bc], dtype=np.int32)
    return problems.Problem("blank", ndim, ndir, tuple(nx), w, x, eq, t), v


def _zero_evolving(w, ne):
    """output array for a right-hand-side evaluation: the reference reads the
    background (getptotal_bg, bg) from the output array w (src/vaccd.t:57,63)"""
    z = w.copy()
    z[:ne] = 0.0
    return z


def _dx_axes(o, ndim):
    dx = o.dx()
    axes = []
    for d in range(ndim):
        idx = [2 if (2 - a) < ndim else 0 for a in range(3)]
        idx[2 - d] = slice(None)
        axes.append(dx[d][tuple(idx)])
    return axes


def _coords(prob, d):
    """cell-centre coordinate of axis d on the grid (shape G3,G2,G1)"""
    G = prob.G[d]
    L = 1.0
    c = (np.arange(G) - 1.5) * (L / prob.nx[d])
    sh = [1, 1, 1]
    sh[2 - d] = G
    return np.broadcast_to(c.reshape(sh), prob.w.shape[1:]).copy()


# ---------------------------------------------------------------------
# gradient4 / gradient1

@pytest.mark.parametrize("d", [0, 1, 2])
def test_gradient4_polynomial_exact(oracle_mod, d):
    prob, v = _blank(3, 3, (10, 9, 8))
    o = make_oracle(prob, oracle_mod)
    X = _coords(prob, d)
    for deg in range(5):
        q = 0.3 + X ** deg if deg else np.full_like(X, 0.7)
        dq = deg * X ** (deg - 1) if deg else np.zeros_like(X)
        g = mesh(prob, o.gradient4(q, d + 1))
        exp = mesh(prob, dq)
        a = 2 - d
        inner = [slice(None)] * 3
        inner[a] = slice(2, -2)
        np.testing.assert_allclose(g[tuple(inner)], exp[tuple(inner)], atol=1e-11, rtol=1e-11)
        edge = [slice(None)] * 3
        for sl in (slice(0, 2), slice(-2, None)):
            edge[a] = sl
            assert np.all(g[tuple(edge)] == 0.0), "two outermost mesh layers must be zero"
    # degree 5 is not exact: the stencil really is 4th order, not higher
    q = X ** 5
    g = mesh(prob, o.gradient4(q, d + 1))
    exp = mesh(prob, 5 * X ** 4)
    inner = [slice(None)] * 3
    inner[2 - d] = slice(2, -2)
    err = np.abs(g - exp)[tuple(inner)].max()
    h = 1.0 / prob.nx[d]
    assert 1e-3 * h ** 4 < err < 10 * h ** 4


@pytest.mark.parametrize("bc", ["periodic", "cont"])
def test_gradient1_family(oracle_mod, bc):
    prob, v = _blank(2, 2, (9, 7), bc=bc)
    o = make_oracle(prob, oracle_mod)
    for d in range(2):
        X = _coords(prob, d)
        q = 2.0 - 3.0 * X
        for kind in (0, 1, 2):
            g = o.gradient1(q, d + 1, kind)
            # range mesh+1 on every axis
            core = g[0, 1:-1, 1:-1]
            a = 1 - d  # axis index within the 2-D (G2,G1) block
            n = core.shape[a]
            first = np.take(core, [0, 1], axis=a)
            last = np.take(core, [n - 2, n - 1], axis=a)
            mid = np.take(core, np.arange(2, n - 2), axis=a)
            np.testing.assert_allclose(mid, -3.0, rtol=1e-12)
            if bc == "periodic":
                # no zeroing at periodic sides: the first mesh cell keeps its
                # derivative (taken across the periodic image)
                assert not np.all(np.take(core, [1], axis=a) == 0.0)
            else:
                assert np.all(first == 0.0) and np.all(last == 0.0)


# ---------------------------------------------------------------------
# fluxes: split perturbation form == textbook MHD of the total state

def _random_state(prob, v, rng, rhob_var=True, uniform_eb=True, bgvec=(0.3, -0.5, 0.4)):
    w = prob.w
    shape = w.shape[1:]
    ndir = prob.ndir
    w[v["rhob"] - 1] = 1.0 + (0.5 * rng.uniform(size=shape) if rhob_var else 0.0)
    w[0] = 0.2 * rng.uniform(-1, 1, size=shape)
    for c in range(ndir):
        w[v["m0"] + c] = 0.3 * rng.uniform(-1, 1, size=shape)
        w[v["b0"] + c] = 0.2 * rng.uniform(-1, 1, size=shape)
        w[v["bg0"] + c] = bgvec[c]
    bg2 = sum(b * b for b in bgvec[:ndir])
    pb = 1.0
    w[v["eb"] - 1] = pb / (prob.eqpar[0] - 1.0) + 0.5 * bg2 if uniform_eb else \
        (pb + 0.2 * rng.uniform(size=shape)) / (prob.eqpar[0] - 1.0) + 0.5 * bg2
    # perturbed energy with a positive total pressure
    rt = w[0] + w[v["rhob"] - 1]
    m2 = sum(w[v["m0"] + c] ** 2 for c in range(ndir))
    b2 = sum(w[v["b0"] + c] ** 2 for c in range(ndir))
    bbg = sum(w[v["b0"] + c] * w[v["bg0"] + c] for c in range(ndir))
    pprime = 0.1 * rng.uniform(-1, 1, size=shape)
    w[v["e"] - 1] = pprime / (prob.eqpar[0] - 1.0) + 0.5 * (m2 / rt + b2) + bbg
    return w


@pytest.mark.parametrize("ndim,ndir", [(1, 2), (1, 3), (2, 3), (3, 3)])
def test_centdiff4_equals_textbook_total_state(oracle_mod, ndim, ndir):
    """uniform background pressure and field, non-uniform background density,
    no gravity, no hyperdiffusion: d(perturbation)/dt from centdiff4 equals
    -sum_d D4(F_textbook(total)) on every mesh cell"""
    nx = {1: (24,), 2: (12, 10), 3: (10, 9, 8)}[ndim]
    prob, v = _blank(ndim, ndir, nx, bc="periodic")
    rng = np.random.Generator(np.random.PCG64(11 + ndim))
    _random_state(prob, v, rng)
    o = make_oracle(prob, oracle_mod)
    w = o.get_w()
    ne = v["nevolve"]
    R = (o.centdiff4(1.0, w, _zero_evolving(w, ne)))[:ne]
    dxa = _dx_axes(o, ndim)
    Rt = tb.rhs_flux(w, ndir, ndim, prob.eqpar[0], dxa)
    # Reference quirk: centdiff4 adds the background coupling
    # bg_idir*bg_idim*d_idir(v_idim) for spatial idim only (src/vaccd.t:61-63),
    # so with ndir > ndim the terms of velocity components beyond ndim that
    # the textbook energy equation carries are absent.
    rt = w[0] + w[v["rhob"] - 1]
    for d in range(ndim):
        for l in range(ndim, ndir):
            vl = w[v["m0"] + l] / rt
            Rt[v["e"] - 1] = Rt[v["e"] - 1] - w[v["bg0"] + d] * w[v["bg0"] + l] * tb.d4(vl, d, dxa[d])
    scale = max(np.abs(mesh(prob, r)).max() for r in Rt)
    for i in range(ne):
        np.testing.assert_allclose(mesh(prob, R[i]), mesh(prob, Rt[i]), atol=1e-12 * scale,
                                   rtol=0, err_msg=f"variable {i + 1}")


def test_centdiff4_hydrostatic_gravity(oracle_mod):
    """hydrostatic background p_b(z) cubic, rho_b = -p_b'/g, uniform B_b,
    gravity on axis 1, v_1 linear along axis 1: the split form plus gravity
    on the perturbed density equals textbook(total) - textbook(background)
    away from the axis-1 edge band (where the reference zeroes D4)."""
    ndim, ndir = 3, 3
    prob, v = _blank(ndim, ndir, (12, 8, 8), bc="periodic")
    g1 = -2.0
    prob.eqpar[problems.GRAV0] = g1
    rng = np.random.Generator(np.random.PCG64(5))
    _random_state(prob, v, rng, rhob_var=False)
    w = prob.w
    Z = _coords(prob, 0)
    pb = 2.0 - 0.8 * Z + 0.1 * Z ** 3
    rhob = (0.8 - 0.3 * Z ** 2) / (-g1)  # p_b' = rho_b g1
    assert np.all(rhob > 0)
    bgv = np.array([0.3, -0.5, 0.4])
    w[v["rhob"] - 1] = rhob
    w[v["eb"] - 1] = pb / (prob.eqpar[0] - 1.0) + 0.5 * (bgv ** 2).sum()
    # v_1 = a(y,z) + b(y,z) * x_1
    Y = _coords(prob, 1)
    X3 = _coords(prob, 2)
    vel1 = 0.1 * np.sin(2 * np.pi * Y) + 0.05 * np.cos(2 * np.pi * X3) * Z
    rt = w[0] + rhob
    w[v["m0"]] = rt * vel1
    o = make_oracle(prob, oracle_mod)
    w = o.get_w()
    ne = v["nevolve"]
    R = o.centdiff4(1.0, w, _zero_evolving(w, ne))[:ne]
    dxa = _dx_axes(o, ndim)
    grav = [g1, 0.0, 0.0]
    Rt = tb.rhs_flux(w, ndir, ndim, prob.eqpar[0], dxa, grav=grav)
    Rb = tb.rhs_flux(w, ndir, ndim, prob.eqpar[0], dxa, grav=grav, background=True)
    sl = (slice(None), slice(None), slice(4, -4))  # axis-1 edge band excluded
    scale = max(np.abs(mesh(prob, r)).max() for r in Rt)
    for i in range(ne):
        got = mesh(prob, R[i])[sl]
        exp = mesh(prob, Rt[i] - Rb[i])[sl]
        np.testing.assert_allclose(got, exp, atol=1e-11 * scale, rtol=0,
                                   err_msg=f"variable {i + 1}")


# ---------------------------------------------------------------------
# cmax

def test_getcmax_closed_forms(oracle_mod):
    ndim, ndir = 3, 3
    prob, v = _blank(ndim, ndir, (6, 6, 6))
    rng = np.random.Generator(np.random.PCG64(2))
    gamma = prob.eqpar[0]
    shape = prob.w.shape[1:]
    rho_b = 0.5 + rng.uniform(size=shape)
    rho_p = 0.3 * rng.uniform(size=shape)
    rho = rho_b + rho_p
    vel = [0.4 * rng.uniform(-1, 1, size=shape) for _ in range(3)]
    p_b = 0.5 + rng.uniform(size=shape)
    p_p = 0.2 * rng.uniform(-1, 1, size=shape)
    for case in ("zeroB", "parallel", "perp", "general"):
        if case == "zeroB":
            B = [np.zeros(shape)] * 3
        elif case == "parallel":
            B = [0.9 * np.ones(shape), np.zeros(shape), np.zeros(shape)]
        elif case == "perp":
            B = [np.zeros(shape), 0.7 * np.ones(shape), -0.2 * np.ones(shape)]
        else:
            B = [rng.uniform(-1, 1, size=shape) for _ in range(3)]
        bg = [0.5 * b for b in B]  # split field between background and perturbation
        bp = [b - g for b, g in zip(B, bg)]
        w = prob.w
        w[0] = rho_p
        w[v["rhob"] - 1] = rho_b
        for c in range(3):
            w[c + 1] = rho * vel[c]
            w[v["b0"] + c] = bp[c]
            w[v["bg0"] + c] = bg[c]
        bg2 = sum(g * g for g in bg)
        w[v["eb"] - 1] = p_b / (gamma - 1) + 0.5 * bg2
        bpbg = sum(a * b for a, b in zip(bp, bg))
        w[v["e"] - 1] = p_p / (gamma - 1) + 0.5 * (rho * sum(u * u for u in vel) +
                                                 sum(a * a for a in bp)) + bpbg
        o = make_oracle(prob, oracle_mod)
        p = p_b + p_p
        for d in range(3):
            cm = mesh(prob, o.getcmax(o.get_w(), d + 1))
            cs = np.sqrt(gamma * p / rho)
            if case == "zeroB":
                exp = cs + np.abs(vel[d])
            elif case == "parallel" and d == 0:
                exp = np.maximum(np.abs(B[0]) / np.sqrt(rho), cs) + np.abs(vel[d])
            elif case == "perp" and d == 0:
                exp = np.sqrt(cs ** 2 + (B[1] ** 2 + B[2] ** 2) / rho) + np.abs(vel[d])
            else:
                exp = tb.fast_speed(rho, p, B, d, gamma) + np.abs(vel[d])
            np.testing.assert_allclose(cm, mesh(prob, exp), rtol=1e-12, err_msg=f"{case} d={d}")


# ---------------------------------------------------------------------
# setnu known answers

def _state_for_setnu(n=40, amp=1e-3, ndir=3, bc="cont"):
    prob, v = _blank(1, ndir, (n,), bc=bc, nu=1.0)
    w = prob.w
    w[0] = 1.0
    gamma = prob.eqpar[0]
    w[v["e"] - 1] = 1.0 / (gamma - 1.0)  # p=1, rho=1, v=0, B=0
    return prob, v


def test_setnu_step_profile(oracle_mod):
    """unit step in v_1 (via m_1): ratios [1,2,2,2,1] on the five faces
    around the jump, zero elsewhere; nuL is nuR shifted by one cell"""
    prob, v = _state_for_setnu()
    w = prob.w
    i0 = 20  # 0-based grid index of the last cell before the jump
    amp = 1e-3
    w[1, 0, 0, i0 + 1:] = amp  # m_1 = rho*v with rho=1
    o = make_oracle(prob, oracle_mod)
    wf = o.get_w()
    nuR, nuL = o.setnu(wf, 2, 1)  # iw=m1
    gamma = prob.eqpar[0]
    ctot = np.sqrt(gamma * 1.0 / 1.0) + amp  # cs + max|v| (p,rho unchanged by m)
    # pressure changes by the kinetic energy taken from e: recompute exactly
    rho = 1.0
    p = (gamma - 1) * (wf[v["e"] - 1, 0, 0] - 0.5 * wf[1, 0, 0] ** 2 / rho)
    ctot = (np.sqrt(gamma * p / rho) + np.abs(wf[1, 0, 0]) / rho)[2:-2].max()
    dx = 1.0 / 40
    K = ctot * 0.4 * dx
    exp = np.zeros(prob.w.shape[-1])
    for off, r in zip(range(-2, 3), [1, 2, 2, 2, 1]):
        exp[i0 + off] = r
    got = nuR[0, 0] / K
    np.testing.assert_allclose(got[2:-2], exp[2:-2], rtol=1e-13, atol=1e-15)
    expL = np.zeros_like(exp)
    expL[1:] = exp[:-1]
    np.testing.assert_allclose(nuL[0, 0, 2:-2] / K, expL[2:-2], rtol=1e-13, atol=1e-15)


def test_setnu_density_and_energy_coefficients(oracle_mod):
    """unit step in rho' and in the thermal-energy code 173: same [1,2,2,2,1]
    ratio pattern, with c_hyp = 0.04 (momentum uses 0.4)"""
    prob, v = _state_for_setnu()
    gamma = prob.eqpar[0]
    i0 = 20
    w = prob.w
    w[0, 0, 0, i0 + 1:] = 1.5  # rho' step (rho_t = 1 -> 1.5), p = 1 kept
    o = make_oracle(prob, oracle_mod)
    wf = o.get_w()
    rho = wf[0, 0, 0]
    ctot = np.sqrt(gamma * 1.0 / rho)[2:-2].max()
    exp = np.zeros(prob.w.shape[-1])
    for off, r in zip(range(-2, 3), [1, 2, 2, 2, 1]):
        exp[i0 + off] = r
    dx = 1.0 / 40
    nuR, _ = o.setnu(wf, 1, 1)
    np.testing.assert_allclose(nuR[0, 0, 2:-2] / (ctot * 0.04 * dx), exp[2:-2], rtol=1e-13,
                               atol=1e-15)
    # thermal energy q = e - (b^2 + m^2/rho)/2: step of p at fixed rho
    prob2, v = _state_for_setnu()
    prob2.w[v["e"] - 1, 0, 0, i0 + 1:] = 2.0 / (gamma - 1.0)
    o2 = make_oracle(prob2, oracle_mod)
    wf2 = o2.get_w()
    p = (gamma - 1.0) * wf2[v["e"] - 1, 0, 0]
    ctot2 = np.sqrt(gamma * p)[2:-2].max()
    nuR2, _ = o2.setnu(wf2, 173, 1)
    np.testing.assert_allclose(nuR2[0, 0, 2:-2] / (ctot2 * 0.04 * dx), exp[2:-2], rtol=1e-13,
                               atol=1e-15)


def test_visc_cross_terms_known_answer(oracle_mod):
    """Shear v_2 = step(x_1) * (a + b x_2) with v_1 = 0: m_1 receives only the
    cross term d_2(rho_t * d_1 v_2 * (nuL+nuR)/4).  At the last cell before
    the step (ratios 2 and 2 on its faces) that is rho_t * 0.4 c_tot * b / 2
    per unit qdt.  The same shape in b_2 gives b_1 only its cross term
    -d_2(d_1 b_2 * (nuL+nuR)/2) = -0.04 c_tot' * b."""
    prob, v = _blank(2, 3, (24, 24), nu=1.0, bc="periodic")
    gamma = prob.eqpar[0]
    w = prob.w
    w[v["rhob"] - 1] = 1.0
    w[v["eb"] - 1] = 1.0 / (gamma - 1.0)  # p = 1
    i0 = 12  # grid index (axis 1) of the last cell before the step
    Y = _coords(prob, 1)
    a_, b_ = 0.3, 0.8
    h = a_ + b_ * Y
    step = np.zeros_like(Y)
    step[..., i0 + 1:] = 1.0
    amp = 1e-3
    v2 = amp * step * h
    w[v["m0"] + 1] = 1.0 * v2
    w[v["e"] - 1] = 0.5 * v2 * v2  # keeps p' = 0
    o = make_oracle(prob, oracle_mod)
    wf = o.get_w()
    dw = o.addsource_visc(1.0, wf, _zero_evolving(wf, v["nevolve"]))
    ctot = np.sqrt(gamma * 1.0)  # c_s; v_1 = 0 and B = 0
    j = 12  # an interior row of axis 2
    got = dw[v["m0"], 0, j, i0]
    exp = 1.0 * 0.4 * ctot * (amp * b_) / 2.0
    np.testing.assert_allclose(got, exp, rtol=1e-12)
    # magnetic analogue
    prob2, v = _blank(2, 3, (24, 24), nu=1.0, bc="periodic")
    w2 = prob2.w
    w2[v["rhob"] - 1] = 1.0
    w2[v["eb"] - 1] = 1.0 / (gamma - 1.0)
    b2 = amp * step * h
    w2[v["b0"] + 1] = b2
    w2[v["e"] - 1] = 0.5 * b2 * b2
    o2 = make_oracle(prob2, oracle_mod)
    wf2 = o2.get_w()
    dw2 = o2.addsource_visc(1.0, wf2, _zero_evolving(wf2, v["nevolve"]))
    ctot2 = np.sqrt(gamma * 1.0 + (b2[0, 2:-2, 2:-2] ** 2).max())  # perpendicular field
    got2 = dw2[v["b0"], 0, j, i0]
    np.testing.assert_allclose(got2, -0.04 * ctot2 * (amp * b_), rtol=1e-12)


def test_setnu_constant_and_linear_give_zero(oracle_mod):
    prob, v = _state_for_setnu()
    X = _coords(prob, 0)
    w = prob.w
    w[v["b0"] + 1] = 0.2 + 0.01 * X  # b_2: linear -> d3 = 0
    prob.ndim = 1
    o = make_oracle(prob, oracle_mod)
    # b_l for l != k exists only for ndim >= 2; use rho (constant) and e code
    nuR, nuL = o.setnu(o.get_w(), 1, 1)
    assert np.all(nuR[0, 0, 2:-2] == 0) and np.all(nuL[0, 0, 2:-2] == 0)
    w2 = o.get_w()
    w2[0] = 1.0 + 0.01 * X  # linear density: d3 = 0 exactly (up to rounding)
    nuR, _ = o.setnu(w2, 1, 1)
    assert np.abs(nuR[0, 0, 4:-4]).max() < 1e-9


def test_setnu_mirror_layer(oracle_mod):
    """a single bump on the fourth mesh cell is copied into the layer beyond
    the grid (mirror of the cell six inside), which makes nuL of the first
    mesh cell non-zero even though the bump is three cells away"""
    prob, v = _state_for_setnu(bc="periodic")
    w = prob.w
    i6 = 5  # 0-based grid index = 1-based 6 = Mlo+3
    w[0, 0, 0, i6] = 1.0 + 1e-3
    o = make_oracle(prob, oracle_mod)
    wf = o.get_w()
    nuR, nuL = o.setnu(wf, 1, 1)
    assert nuL[0, 0, 2] > 0.0, "mirror layer must reach the first mesh cell"
    assert nuR[0, 0, 2] > 0.0
    # same bump moved to the fifth mesh cell: the mirror copies a zero
    w2 = prob.w.copy()
    w2[0, 0, 0, i6] = 1.0
    w2[0, 0, 0, i6 + 1] = 1.0 + 1e-3
    o2 = make_oracle(type(prob)(**{**prob.__dict__, "w": w2}), oracle_mod)
    nuR2, nuL2 = o2.setnu(o2.get_w(), 1, 1)
    assert nuL2[0, 0, 2] == 0.0


def test_setnu_bound(oracle_mod):
    """md3 <= 6 md1 on every window, so nu <= 6 c_tot c_hyp dx"""
    prob, v = _state_for_setnu(n=64)
    rng = np.random.Generator(np.random.PCG64(9))
    prob.w[1, 0, 0, 2:-2] = 1e-2 * rng.uniform(-1, 1, size=64)
    o = make_oracle(prob, oracle_mod)
    wf = o.get_w()
    nuR, nuL = o.setnu(wf, 2, 1)
    gamma = prob.eqpar[0]
    p = (gamma - 1) * (wf[v["e"] - 1, 0, 0] - 0.5 * wf[1, 0, 0] ** 2)
    ctot = (np.sqrt(gamma * p) + np.abs(wf[1, 0, 0]))[2:-2].max()
    bound = 6 * ctot * 0.4 / 64
    assert nuR.max() <= bound * (1 + 1e-12) and nuL.max() <= bound * (1 + 1e-12)
    assert nuR.max() > 0.1 * bound


# ---------------------------------------------------------------------
# hyperdiffusion: conservation, dissipation, symmetry, scratch independence

def _visc_problem(ndim, seed=3, bc="periodic"):
    nx = {1: (40,), 2: (22, 20), 3: (22, 20, 20)}[ndim]
    prob, v = _blank(ndim, 3, nx, bc=bc, nu=1.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    w = prob.w
    shape = w.shape[1:]
    w[v["rhob"] - 1] = 1.0
    w[v["eb"] - 1] = 1.0 / (prob.eqpar[0] - 1.0) + 0.5 * 0.25
    w[v["bg0"]] = 0.5
    # perturbation confined away from the boundaries (7 cells margin)
    mask = np.ones(shape, dtype=bool)
    for d in range(ndim):
        a = 2 - d
        idx = np.arange(shape[a])
        inner = (idx >= 9) & (idx < shape[a] - 9)
        sh = [1, 1, 1]
        sh[a] = shape[a]
        mask &= inner.reshape(sh)
    for iw in range(v["nevolve"]):
        if iw == v["e"] - 1:
            continue
        w[iw] = np.where(mask, 1e-2 * rng.uniform(-1, 1, size=shape), 0.0)
    w[v["e"] - 1] = np.where(mask, 1e-2 * rng.uniform(-1, 1, size=shape), 0.0)
    return prob, v


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_visc_conserves_and_dissipates(oracle_mod, ndim):
    prob, v = _visc_problem(ndim)
    o = make_oracle(prob, oracle_mod)
    w = o.get_w()
    dw = o.addsource_visc(1.0, w, _zero_evolving(w, v["nevolve"]))
    for iw in range(v["nevolve"]):
        s = mesh(prob, dw[iw]).sum()
        scale = np.abs(mesh(prob, dw[iw])).sum() + 1e-300
        assert abs(s) <= 1e-12 * scale, f"variable {iw + 1} not conserved: {s} vs {scale}"
    # density diffusion is dissipative: sum(rho' * d rho') < 0
    assert (mesh(prob, w[0]) * mesh(prob, dw[0])).sum() < 0.0


def _reflect(prob, v, d):
    """mirror the whole state along reference axis d: index reversal and sign
    flip of the d-components of m, b and bg"""
    w = prob.w.copy()
    a = 2 - d
    w = np.flip(w, axis=1 + a).copy()
    for name in ("m0", "b0", "bg0"):
        w[v[name] + d] *= -1.0
    return w


@pytest.mark.parametrize("d", [0, 1, 2])
def test_stage_reflection_symmetry(oracle_mod, d):
    """one advect1 (fluxes + hyperdiffusion + gravity off) commutes with a
    reflection of any axis: catches left/right index slips"""
    prob, v = _visc_problem(3, seed=8, bc="periodic")
    rng = np.random.Generator(np.random.PCG64(12))
    prob.w[v["rhob"] - 1] = 1.0 + 0.3 * rng.uniform(size=prob.w.shape[1:])
    o = make_oracle(prob, oracle_mod)
    w = o.get_w()
    out = o.advect1(1e-3, w, w)
    probR = type(prob)(**{**prob.__dict__, "w": _reflect(type(prob)(**{**prob.__dict__, "w": w}), v, d)})
    oR = make_oracle(probR, oracle_mod)
    wR = oR.get_w()
    outR = oR.advect1(1e-3, wR, wR)
    back = _reflect(type(prob)(**{**prob.__dict__, "w": outR}), v, d)
    ne = v["nevolve"]
    for iw in range(ne):
        a_, b_ = mesh(prob, back[iw] - w[iw]), mesh(prob, out[iw] - w[iw])
        scale = np.abs(b_).max() + 1e-300
        np.testing.assert_allclose(a_, b_, atol=1e-10 * scale, rtol=0, err_msg=f"var {iw + 1}")


def _permute(prob, v, perm):
    """relabel axes: new axis k = old axis perm[k] (3-D, ndir=3)"""
    w = prob.w
    # array axes are (axis3, axis2, axis1) -> transpose accordingly
    old_to_arr = {0: 2, 1: 1, 2: 0}
    new_arr_order = [old_to_arr[perm[2]], old_to_arr[perm[1]], old_to_arr[perm[0]]]
    wt = np.stack([np.transpose(w[i], new_arr_order) for i in range(w.shape[0])])
    out = wt.copy()
    for name in ("m0", "b0", "bg0"):
        for k in range(3):
            out[v[name] + k] = wt[v[name] + perm[k]]
    nx = tuple(prob.nx[perm[k]] for k in range(3))
    return out, nx


def test_stage_axis_permutation(oracle_mod):
    """relabelling the axes (1,2,3)->(2,3,1) commutes with one advect1:
    catches per-axis index errors in the cross terms"""
    # a cube, so that dx is identical on every axis
    prob, v = _blank(3, 3, (12, 12, 12), nu=1.0)
    rng = np.random.Generator(np.random.PCG64(21))
    shape = prob.w.shape[1:]
    w = prob.w
    w[v["rhob"] - 1] = 1.0 + 0.2 * rng.uniform(size=shape)
    w[v["eb"] - 1] = 2.0
    for c in range(3):
        w[v["bg0"] + c] = 0.2 * (c + 1)
    for iw in range(v["nevolve"]):
        w[iw] = 1e-2 * rng.uniform(-1, 1, size=shape)
    w[v["e"] - 1] += 0.5
    o = make_oracle(prob, oracle_mod)
    w = o.get_w()
    out = o.advect1(1e-3, w, w)
    perm = (1, 2, 0)
    wp, nxp = _permute(type(prob)(**{**prob.__dict__, "w": w}), v, perm)
    probP = type(prob)(**{**prob.__dict__, "w": wp, "nx": nxp})
    oP = make_oracle(probP, oracle_mod)
    outP = oP.advect1(1e-3, oP.get_w(), oP.get_w())
    exp, _ = _permute(type(prob)(**{**prob.__dict__, "w": out}), v, perm)
    ne = v["nevolve"]
    for iw in range(ne):
        a_, b_ = mesh(probP, outP[iw] - wp[iw]), mesh(probP, exp[iw] - wp[iw])
        scale = np.abs(b_).max() + 1e-300
        np.testing.assert_allclose(a_, b_, atol=1e-10 * scale, rtol=0, err_msg=f"var {iw + 1}")


def test_results_do_not_depend_on_stale_scratch(oracle_mod):
    prob = problems.stratified(n=10, nxyz=(10, 8, 8), amp=5e-2)
    o1 = make_oracle(prob, oracle_mod)
    w = o1.get_w()
    a = o1.advect1(1e-3, w, w)
    o2 = make_oracle(prob, oracle_mod)
    o2.poison_scratch(1e30)
    b = o2.advect1(1e-3, w, w)
    np.testing.assert_array_equal(mesh(prob, a), mesh(prob, b))


# ---------------------------------------------------------------------
# gravity closed form, RK4 tableau, invariants

def test_gravity_uniform_state_closed_form(oracle_mod):
    """uniform state under gravity: m_1 = g rho' t, e' = e0 + g^2 rho'^2 t^2
    / (2 rho_t); RK4 integrates the polynomial exactly"""
    prob, v = _blank(3, 3, (8, 8, 8), nu=1.0, grav=(-1.5, 0.0, 0.0), bc="periodic")
    w = prob.w
    r0, rb, e0 = 0.2, 1.0, 0.7
    w[0] = r0
    w[v["rhob"] - 1] = rb
    w[v["e"] - 1] = e0
    w[v["eb"] - 1] = 2.0
    o = make_oracle(prob, oracle_mod)
    t = 0.0
    for _ in range(5):
        dt = 0.013
        o.advance(dt)
        t += dt
    wf = o.get_w()
    g = -1.5
    np.testing.assert_allclose(mesh(prob, wf[1]), g * r0 * t, rtol=1e-13)
    np.testing.assert_allclose(mesh(prob, wf[v["e"] - 1]), e0 + g * g * r0 * r0 * t * t / (2 * (r0 + rb)),
                               rtol=1e-13)
    np.testing.assert_array_equal(mesh(prob, wf[0]), r0)


def test_fourstep_is_classical_rk4(oracle_mod):
    """advance('fourstep') equals w + dt/6 (k1 + 2k2 + 2k3 + k4) built from
    the oracle's own right-hand side with getboundary on each stage input"""
    prob = problems.stratified(n=10, nxyz=(10, 8, 8), amp=5e-2)
    o = make_oracle(prob, oracle_mod)
    w = o.get_w()
    dt = 2e-3

    ne = 8

    def R(state):
        return o.advect1(1.0, state, _zero_evolving(state, ne))

    k1 = R(w)
    s = w.copy(); s[:ne] = w[:ne] + dt / 2 * k1[:ne]; s = o.getboundary_on(s)
    k2 = R(s)
    s = w.copy(); s[:ne] = w[:ne] + dt / 2 * k2[:ne]; s = o.getboundary_on(s)
    k3 = R(s)
    s = w.copy(); s[:ne] = w[:ne] + dt * k3[:ne]; s = o.getboundary_on(s)
    k4 = R(s)
    exp = w[:ne] + dt / 6 * (k1[:ne] + 2 * k2[:ne] + 2 * k3[:ne] + k4[:ne])
    o.advance(dt)
    got = o.get_w()[:ne]
    for iw in range(ne):
        scale = np.abs(mesh(prob, exp[iw] - w[iw])).max() + 1e-300
        np.testing.assert_allclose(mesh(prob, got[iw] - w[iw]), mesh(prob, exp[iw] - w[iw]),
                                   atol=1e-11 * scale, rtol=0, err_msg=f"var {iw + 1}")


def test_hydrostatic_zero_perturbation_stays_zero(oracle_mod):
    prob = problems.stratified(n=8, nxyz=(10, 8, 8), amp=0.0, noise=0.0)
    o = make_oracle(prob, oracle_mod)
    for _ in range(3):
        o.step(prob.courantpar)
    wf = o.get_w()
    assert np.all(wf[:8] == 0.0)


def test_uniform_state_stays_uniform(oracle_mod):
    prob, v = _blank(3, 3, (8, 8, 8), nu=1.0, bc="cont")
    w = prob.w
    w[0], w[1], w[2], w[3] = 0.1, 0.05, -0.02, 0.03
    w[v["e"] - 1] = 1.7
    for c in range(3):
        w[v["b0"] + c] = 0.1 * (c + 1)
        w[v["bg0"] + c] = 0.2
    w[v["rhob"] - 1] = 1.0
    w[v["eb"] - 1] = 1.5
    o = make_oracle(prob, oracle_mod)
    w0 = o.get_w()
    for _ in range(3):
        o.step(0.3)
    wf = o.get_w()
    for iw in range(8):
        np.testing.assert_allclose(mesh(prob, wf[iw]), mesh(prob, w0[iw]), rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------------
# getboundary semantics

def test_getboundary_semantics_1d(oracle_mod):
    prob, v = _blank(1, 2, (8,), bc="cont")
    prob.w[:, 0, 0, 2:-2] = np.arange(1, 9)[None, :] * (1 + np.arange(prob.w.shape[0]))[:, None]
    for codes, exp_lo, exp_hi in (((1, 1), (1, 1), (8, 8)), ((3, 3), (7, 8), (1, 2)),
                                  ((4, 6), (2, 1), (-8, -7)), ((2, 2), (1, 1), (8, 8))):
        prob.typeB[:, :] = codes
        o = make_oracle(prob, oracle_mod)
        w = o.get_w()
        np.testing.assert_array_equal(w[0, 0, 0, :2], exp_lo, err_msg=str(codes))
        np.testing.assert_array_equal(w[0, 0, 0, -2:], exp_hi, err_msg=str(codes))
        if codes == (2, 2):
            w2 = w.copy()
            w2[:, 0, 0, 2:-2] *= 3.0
            w3 = o.getboundary_on(w2)
            np.testing.assert_array_equal(w3[0, 0, 0, :2], exp_lo)  # restored, not re-copied


# stability polynomials sum_k c_k z^k of the integrators (src/vac.t:406-475)
# for a linear right-hand side: twostep is the midpoint rule, threestep is
# Shu's SSP-RK3, fourstep and jameson are exact to z^4 for linear problems,
# sterck's stage factors 0.12, 1/4, 1/2, 1 give 1/8 and 0.12/8 (0.12 is a
# default-real literal there)
_STAB = {
    "onestep": [1.0, 1.0],
    "twostep": [1.0, 1.0, 0.5],
    "threestep": [1.0, 1.0, 0.5, 1.0 / 6.0],
    "fourstep": [1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0],
    "jameson": [1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0],
    "sterck": [1.0, 1.0, 0.5, 1.0 / 8.0, float(np.float32(0.12)) / 8.0],
}


@pytest.mark.parametrize("adv", sorted(_STAB))
def test_integrator_stability_polynomial(oracle_mod, adv):
    """One step of a small perturbation of a uniform periodic MHD state
    (no hyperdiffusion, so the right-hand side is linear in the perturbation
    up to second-order terms) equals sum_k c_k (dt L)^k delta, with
    (dt L)^k delta built by repeated right-hand-side evaluations.  The
    polynomials differ at z^2..z^4 by 1e-3..1e-2 of the step at this dt, far
    above the 1e-5 left by the nonlinear terms."""
    prob, v = _blank(2, 3, (24, 20))
    ne = v["nevolve"]
    rng = np.random.default_rng(11)
    u0 = prob.w.copy()
    u0[0] = 1.0
    u0[v["e"] - 1] = 1.5
    u0[v["b0"]] = 0.6
    u0[v["b0"] + 1] = -0.4
    u0[v["m0"]] = 0.3
    d = np.zeros_like(u0)
    x, y = _coords(prob, 0), _coords(prob, 1)
    for iw in range(ne):
        a, b, ph = rng.uniform(-1, 1, 3)
        d[iw] = 1e-7 * (a * np.sin(2 * np.pi * x + ph) + b * np.cos(2 * np.pi * (x + 2 * y)))
    prob.w = u0 + d
    o = make_oracle(prob, oracle_mod)
    w = o.get_w()
    base = o.getboundary_on(u0)
    dt = 4e-3  # courant number ~0.3 on this grid

    def dtL(vec):
        s = base.copy()
        s[:ne] = base[:ne] + vec[:ne]
        s = o.getboundary_on(s)
        r = o.advect1(dt, s, _zero_evolving(s, ne))
        return r[:ne]

    terms = [w[:ne] - base[:ne]]
    for _ in range(len(_STAB[adv]) - 1):
        terms.append(dtL(terms[-1]))
    exp = sum(c * t for c, t in zip(_STAB[adv][1:], terms[1:]))
    o.advance(dt, adv)
    got = o.get_w()[:ne] - w[:ne]
    for iw in range(ne):
        scale = np.abs(mesh(prob, exp[iw])).max()
        if scale == 0.0:
            continue
        np.testing.assert_allclose(mesh(prob, got[iw]), mesh(prob, exp[iw]), atol=1e-5 * scale,
                                   rtol=0, err_msg=f"{adv} var {iw + 1}")


def test_threestep_is_ssp_rk3_butcher(oracle_mod):
    """advance('threestep') (Shu-Osher form) equals the Butcher form of SSP-RK3:
    k1=R(w), k2=R(w+dt k1), k3=R(w+dt/4 (k1+k2)), w+dt (k1+k2+4k3)/6"""
    prob = problems.stratified(n=10, nxyz=(10, 8, 8), amp=5e-2)
    o = make_oracle(prob, oracle_mod)
    w = o.get_w()
    dt = 2e-3
    ne = 8

    def R(state):
        return o.advect1(1.0, state, _zero_evolving(state, ne))

    k1 = R(w)
    s = w.copy(); s[:ne] = w[:ne] + dt * k1[:ne]; s = o.getboundary_on(s)
    k2 = R(s)
    s = w.copy(); s[:ne] = w[:ne] + dt / 4 * (k1[:ne] + k2[:ne]); s = o.getboundary_on(s)
    k3 = R(s)
    exp = w[:ne] + dt * (k1[:ne] + k2[:ne] + 4 * k3[:ne]) / 6
    o.advance(dt, "threestep")
    got = o.get_w()[:ne]
    for iw in range(ne):
        scale = np.abs(mesh(prob, exp[iw] - w[iw])).max() + 1e-300
        np.testing.assert_allclose(mesh(prob, got[iw] - w[iw]), mesh(prob, exp[iw] - w[iw]),
                                   atol=1e-11 * scale, rtol=0, err_msg=f"var {iw + 1}")


def test_extrapolating_and_mirror_boundaries(oracle_mod):
    """cont1, fixed1, grad1 and contCD4 (src/vacgrid.t:260-290,376-407,476-507)
    on a 1-D tube with hand-computed ghosts: mesh values q_i = 3 + 2 i
    (i = 1..N) plus a bump at the edge cells so that mirror and
    extrapolation differ"""
    from sac_amd import problems as P
    n = 8
    prob, v = _blank(1, 2, (n,), bc="cont")
    codes = {1: P.BC["cont1"], 2: P.BC["fixed1"], 3: P.BC["grad1"], 4: P.BC["contCD4"]}
    for iw, c in codes.items():
        prob.typeB[iw - 1, :] = c
    q = 3.0 + 2.0 * np.arange(1, n + 1)
    q[0] += 0.5     # first mesh cell
    q[-1] -= 0.25   # last mesh cell
    for iw in codes:
        prob.w[iw - 1, 0, 0, 2:-2] = q
    o = make_oracle(prob, oracle_mod)  # the first getboundary stores fixed1/grad1
    w = o.get_w()[:, 0, 0, :]
    # linear extrapolation from the edge and the inner edge
    lo = [q[0] + 2 * (q[0] - q[1]), q[0] + (q[0] - q[1])]
    hi = [q[-1] + (q[-1] - q[-2]), q[-1] + 2 * (q[-1] - q[-2])]
    for iw in (1, 2, 3):
        np.testing.assert_allclose(w[iw - 1, :2], lo, rtol=1e-15)
        np.testing.assert_allclose(w[iw - 1, -2:], hi, rtol=1e-15)
    # contCD4: ghost i-1 <- cell i+1, ghost i-2 <- cell i+2 (edge cell i)
    np.testing.assert_array_equal(w[3, :2], [q[2], q[1]])
    np.testing.assert_array_equal(w[3, -2:], [q[-2], q[-3]])
    # second fill after the mesh changed: cont1 extrapolates again, fixed1
    # restores the stored values, grad1 keeps the stored differences to the
    # new edge values
    w2 = o.get_w()
    q2 = q * 1.5 - 1.0
    for iw in codes:
        w2[iw - 1, 0, 0, 2:-2] = q2
    r = o.getboundary_on(w2)[:, 0, 0, :]
    lo2 = [q2[0] + 2 * (q2[0] - q2[1]), q2[0] + (q2[0] - q2[1])]
    hi2 = [q2[-1] + (q2[-1] - q2[-2]), q2[-1] + 2 * (q2[-1] - q2[-2])]
    np.testing.assert_allclose(r[0, :2], lo2, rtol=1e-15)
    np.testing.assert_allclose(r[0, -2:], hi2, rtol=1e-15)
    np.testing.assert_array_equal(r[1, :2], w[1, :2])
    np.testing.assert_array_equal(r[1, -2:], w[1, -2:])
    np.testing.assert_allclose(r[2, :2], [q2[0] + (lo[0] - q[0]), q2[0] + (lo[1] - q[0])], rtol=1e-15)
    np.testing.assert_allclose(r[2, -2:], [q2[-1] + (hi[0] - q[-1]), q2[-1] + (hi[1] - q[-1])],
                               rtol=1e-15)
    np.testing.assert_array_equal(r[3, :2], [q2[2], q2[1]])


def test_powell_divb_source_closed_form(oracle_mod):
    """addsource_divb (src/vacphys.mhd0.t:264-297): with b+bg linear in
    x and y, div B is exact for gradient4 away from the two zeroed edge
    cells per side, and one advect1 with divbfix differs from the same step
    without it by -qdt*divB*((b+bg) for m, m/rho_t for b, m.(b+bg)/rho_t for
    e), nothing else"""
    prob, v = _blank(2, 3, (12, 10), bc="cont")
    x, y = _coords(prob, 0), _coords(prob, 1)
    w = prob.w
    w[0] = 0.3 + 0.1 * x            # rho'
    w[v["rhob"] - 1] = 1.2
    w[v["m0"]] = 0.2 - 0.1 * y
    w[v["m0"] + 1] = 0.1 + 0.05 * x
    w[v["m0"] + 2] = -0.3
    w[v["e"] - 1] = 2.0
    s1, t1, s2 = 0.7, -0.4, 0.25
    w[v["b0"]] = 0.3 + s1 * x
    w[v["b0"] + 1] = -0.2 + s2 * y
    w[v["b0"] + 2] = 0.15 * x
    w[v["bg0"]] = 0.2 + t1 * x
    w[v["bg0"] + 1] = 0.05
    w[v["bg0"] + 2] = 0.1
    w[v["eb"] - 1] = 3.0
    o = make_oracle(prob, oracle_mod)
    wct = o.get_w()
    qdt = 1e-3
    off = o.advect1(qdt, wct, wct)
    o.set_divbfix(True)
    on = o.advect1(qdt, wct, wct)
    o.set_divbfix(False)
    d = on - off
    divb = s1 + t1 + s2
    rt = wct[0] + wct[v["rhob"] - 1]
    exp = np.zeros_like(d)
    for c in range(3):
        bt = wct[v["b0"] + c] + wct[v["bg0"] + c]
        exp[v["m0"] + c] = -qdt * bt * divb
        exp[v["b0"] + c] = -qdt * wct[v["m0"] + c] / rt * divb
        exp[v["e"] - 1] += -qdt * wct[v["m0"] + c] * bt / rt * divb
    inner = (0, slice(4, -4), slice(4, -4))
    for iw in range(v["nevolve"]):
        scale = np.abs(exp[iw][inner]).max() + 1e-300
        np.testing.assert_allclose(d[iw][inner], exp[iw][inner], atol=1e-10 * scale, rtol=0,
                                   err_msg=f"var {iw + 1}")
    assert np.abs(d[0][inner]).max() == 0.0  # rho' has no Powell term


@pytest.mark.parametrize("compact", [False, True])
def test_resistive_sources_closed_form(oracle_mod, compact):
    """Resistive MHD with constant eta (src/vacphys.mhdres.t): for
    B_y = c x^2 (centred differences of a quadratic are exact) one advect1
    differs from the same step with eta=0 by b_y += qdt eta 2c and, for the
    energy, by the discrete closed forms of the two source forms: res2
    (conservative) e += qdt eta 2c^2 (3x^2 + h^2), res1 (compact) e += qdt
    eta 6 c^2 x^2; nothing else changes"""
    prob, v = _blank(2, 3, (16, 12), bc="cont")
    x = _coords(prob, 0)
    h = 1.0 / prob.nx[0]
    c2, eta, qdt = 0.8, 2e-3, 1e-3
    w = prob.w
    w[0] = 0.1
    w[v["rhob"] - 1] = 1.0
    w[v["e"] - 1] = 2.0
    w[v["eb"] - 1] = 1.0
    w[v["b0"] + 1] = c2 * x * x
    prob.eqpar[1] = eta
    o = make_oracle(prob, oracle_mod)
    o.set_resistive(compact, 0.5)
    wct = o.get_w()
    on = o.advect1(qdt, wct, wct)
    prob0 = type(prob)(**{**prob.__dict__, "eqpar": prob.eqpar.copy()})
    prob0.eqpar[1] = 0.0
    off = make_oracle(prob0, oracle_mod).advect1(qdt, wct, wct)
    d = on - off
    inner = (0, slice(5, -5), slice(5, -5))
    xi = x[inner]
    np.testing.assert_allclose(d[v["b0"] + 1][inner], qdt * eta * 2 * c2, rtol=1e-9)
    if compact:
        exp_e = qdt * eta * 6 * c2 * c2 * xi * xi
    else:
        exp_e = qdt * eta * 2 * c2 * c2 * (3 * xi * xi + h * h)
    np.testing.assert_allclose(d[v["e"] - 1][inner], exp_e, rtol=1e-9, atol=1e-22)
    for iw in range(v["nevolve"]):
        if iw not in (v["b0"] + 1, v["e"] - 1):
            assert np.abs(d[iw][inner]).max() < 1e-20, iw


def test_resistive_dt_limit(oracle_mod):
    """getdt_res = dtdiffpar min(dx)^2 / eta with the par value of dtdiffpar
    until getdt_visc has assigned 0.25 to it"""
    prob, v = _blank(2, 3, (16, 12), bc="cont", nu=1.0)
    prob.w[v["rhob"] - 1] = 1.0
    prob.w[v["e"] - 1] = 2.0
    prob.eqpar[1] = 2e-3
    o = make_oracle(prob, oracle_mod)
    o.set_resistive(False, 0.6)
    dxmin = min(1.0 / 16, 1.0 / 12)
    assert o.getdt_res(1e30) == pytest.approx(0.6 * dxmin ** 2 / 2e-3, rel=1e-14)
    o.getdt_special(1e30)  # getdt_visc sets the common dtdiffpar to 0.25
    assert o.getdt_res(1e30) == pytest.approx(0.25 * dxmin ** 2 / 2e-3, rel=1e-14)


def _zero_state(ndim, nx, gamma, seed, moving):
    """random state for the 'zero' boundary pins: rho' and e of either sign
    around positive totals, the background positive in every cell (ghosts
    included, so that the whole-grid conversions never divide by zero);
    moving=False: no momentum and no perturbed field"""
    prob, v = _blank(ndim, 3, nx, gamma=gamma, bc="cont")
    rng = np.random.default_rng(seed)
    w = prob.w
    ms = prob.mesh_slices()
    w[v["rhob"] - 1] = rng.uniform(1.0, 2.0, w[0].shape)
    w[v["eb"] - 1] = rng.uniform(2.0, 3.0, w[0].shape)
    for c in range(3):
        w[v["bg0"] + c] = rng.uniform(-0.5, 0.5, w[0].shape)
    w[0][ms] = rng.uniform(-0.2, 0.2, w[0][ms].shape)
    w[v["e"] - 1][ms] = rng.uniform(1.0, 2.0, w[0][ms].shape)
    if moving:
        for c in range(3):
            w[v["m0"] + c][ms] = rng.uniform(-0.3, 0.3, w[0][ms].shape)
            w[v["b0"] + c][ms] = rng.uniform(-0.3, 0.3, w[0][ms].shape)
    return prob, v


def test_zero_boundary_is_contcd4_when_the_conversions_are_exact(oracle_mod):
    """'zero' (src/vacgrid.t:294-345) is contCD4's copy made in primitive
    variables with the whole grid converted before and after.  With gamma=2,
    no momentum and no perturbed field the conversions are exact (p = e,
    v = 0), so the fill must equal contCD4's everywhere, bit for bit."""
    for ndim, nx in ((1, (12,)), (2, (10, 9))):
        prob, v = _zero_state(ndim, nx, 2.0, seed=ndim, moving=False)
        prob.typeB[:, :] = BC["zero"]
        wz = make_oracle(prob, oracle_mod).get_w()
        prob.typeB[:, :] = BC["contCD4"]
        wc = make_oracle(prob, oracle_mod).get_w()
        np.testing.assert_array_equal(wz, wc)


def test_zero_boundary_copies_primitive_variables(oracle_mod):
    """On a general state the 'zero' ghosts hold the contCD4 copy of the
    primitive variables: rho' and b exactly, v = m/rho_t and the thermal
    pressure p = (gamma-1)(e - m^2/(2 rho_t) - b^2/2 - b.bg) to rounding; the
    mesh keeps rho' and b exactly and m, e to rounding (each conversion
    round trip moves them by a few ulps)."""
    gamma = 5.0 / 3.0
    prob, v = _zero_state(1, (12,), gamma, seed=7, moving=True)
    prob.typeB[:, 0] = BC["cont"]
    prob.typeB[:, 1] = BC["zero"]  # upper boundary
    w0 = prob.w.copy()
    w = make_oracle(prob, oracle_mod).get_w()
    ne = 8
    rt = w[0] + w[v["rhob"] - 1]
    vel = np.stack([w[v["m0"] + c] / rt for c in range(3)])
    b = np.stack([w[v["b0"] + c] for c in range(3)])
    bg = np.stack([w[v["bg0"] + c] for c in range(3)])
    p = (gamma - 1.0) * (w[v["e"] - 1] - 0.5 * rt * (vel ** 2).sum(0) - 0.5 * (b ** 2).sum(0)
                         - (b * bg).sum(0))
    sl = lambda i: (0, 0, i)  # noqa: E731
    mhi = 2 + 12 - 1
    for ghost, src in ((mhi + 1, mhi - 1), (mhi + 2, mhi - 2)):
        assert w[(0,) + sl(ghost)] == w[(0,) + sl(src)]
        for c in range(3):
            assert w[(v["b0"] + c,) + sl(ghost)] == w[(v["b0"] + c,) + sl(src)]
            assert vel[(c,) + sl(ghost)] == pytest.approx(vel[(c,) + sl(src)], rel=1e-14, abs=1e-16)
        assert p[sl(ghost)] == pytest.approx(p[sl(src)], rel=1e-13)
    m = (slice(None), 0, 0, slice(2, 2 + 12))
    np.testing.assert_array_equal(w[0][m[1:]], w0[0][m[1:]])
    for c in range(3):
        np.testing.assert_array_equal(w[v["b0"] + c][m[1:]], w0[v["b0"] + c][m[1:]])
        np.testing.assert_allclose(w[v["m0"] + c][m[1:]], w0[v["m0"] + c][m[1:]], rtol=1e-15,
                                   atol=1e-16)
    np.testing.assert_allclose(w[v["e"] - 1][m[1:]], w0[v["e"] - 1][m[1:]], rtol=1e-14)
    # the lower (cont) ghosts copy the edge cell of the converted state
    for c in range(ne):
        assert w[(c,) + sl(0)] == w[(c,) + sl(2)]
