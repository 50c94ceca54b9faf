"""Synthetic initial states for the SAC perturbation-MHD configurations.

Each builder returns a Problem: the reference-layout state w (shape
(nw, G3, G2, G1), the memory of the Fortran array w(ixG^T,nw)), mesh
coordinates x, the eqpar array, boundary codes and run parameters.  Ghost
cells of w are zero, as in the reference after readfileini with
fullgridini=F; getboundary fills them.

The eqpar layout is the reference's: gamma_=1, eta_=2
(src/vacpar.t.mhd:26) followed by the user parameters of
sac/src/vacusrpar.t.stuart:3-6: gravity component d at grav0_+d with
grav0_=3, and the hyperdiffusion switch at nu_=6.  eqpar always has six
entries so that nu_ is in bounds in 1-D as well.  In 3-D slot 6 is both the
axis-3 gravity and the switch (the reference's index collision); the
stratified configurations store 1e-30 there, which turns hyperdiffusion on
and makes the axis-3 gravity vanish below double-precision resolution.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BC = {"cont": 1, "fixed": 2, "periodic": 3, "symm": 4, "symm0": 5, "asymm": 6,
      "cont1": 7, "fixed1": 8, "grad1": 9, "contCD4": 10, "zero": 11}
BC_NAMES = {v: k for k, v in BC.items()}

GRAV0 = 3  # grav0_ (1-based)
NU = 6     # nu_ (1-based)
NEQPAR = 6


def var_index(ndir):
    """1-based variable indices of src/vacpar.t.mhd:9-13."""
    e = ndir + 2
    eb = e + ndir + 1
    return {"rho": 1, "m0": 1, "e": e, "b0": e, "eb": eb, "rhob": eb + 1,
            "bg0": eb + 1, "nw": eb + 1 + ndir, "nevolve": e + ndir}


@dataclass
class Problem:
    name: str
    ndim: int
    ndir: int
    nx: tuple
    w: np.ndarray           # (nw, G3, G2, G1)
    x: np.ndarray           # (ndim, G3, G2, G1), mesh entries filled
    eqpar: np.ndarray       # (6,)
    typeB: np.ndarray       # (nw, 2*ndim) codes
    courantpar: float = 0.3
    tmax: float = 1e99
    itmax: int = 10_000_000
    typeadvance: str = "fourstep"
    meta: dict = field(default_factory=dict)

    @property
    def nw(self):
        return self.w.shape[0]

    @property
    def G(self):
        return tuple(n + 4 for n in self.nx) + (1,) * (3 - self.ndim)

    def iws(self):
        v = var_index(self.ndir)
        return list(range(1, v["nevolve"] + 1))

    def mesh_slices(self):
        sl = []
        for ax in (2, 1, 0):
            sl.append(slice(2, -2) if ax < self.ndim else slice(None))
        return tuple(sl)


def _alloc(ndim, ndir, nx):
    v = var_index(ndir)
    G = tuple(n + 4 for n in nx) + (1,) * (3 - ndim)
    shape = G[::-1]
    return np.zeros((v["nw"],) + shape), np.zeros((ndim,) + shape), v


def _mesh_coords(ndim, nx, lengths=None):
    """cell centres x_i=(i-1/2)*L/N on each active axis, on the mesh"""
    lengths = lengths or (1.0,) * ndim
    return [(np.arange(nx[d]) + 0.5) * (lengths[d] / nx[d]) for d in range(ndim)]


def brio_wu(n=800, gamma=2.0, nu=1.0):
    """1-D Brio-Wu shock tube, d=12 (ndim=1, ndir=2, nw=10), zero
    background, cont boundaries (par/briowu-test12 states, inipar layout)."""
    ndim, ndir = 1, 2
    w, x, v = _alloc(ndim, ndir, (n,))
    xc = _mesh_coords(1, (n,))[0]
    x[0, 0, 0, 2:-2] = xc
    left = xc <= 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    b1 = np.full(n, 0.75)
    b2 = np.where(left, 1.0, -1.0)
    e = p / (gamma - 1.0) + 0.5 * (b1 * b1 + b2 * b2)
    w[v["rho"] - 1, 0, 0, 2:-2] = rho
    w[v["e"] - 1, 0, 0, 2:-2] = e
    w[v["b0"] + 1 - 1, 0, 0, 2:-2] = b1
    w[v["b0"] + 2 - 1, 0, 0, 2:-2] = b2
    eq = np.zeros(NEQPAR)
    eq[0] = gamma
    eq[NU - 1] = nu
    tb = np.full((v["nw"], 2), BC["cont"], dtype=np.int32)
    return Problem("brio_wu", ndim, ndir, (n,), w, x, eq, tb,
                   courantpar=0.3, tmax=0.1)


def stratified(n=128, gamma=5.0 / 3.0, H=0.2, amp=1e-3, noise=1e-6,
               width=0.05, seed=20261202, nu=1e-30, nxyz=None):
    """3-D stratified atmosphere on a static background (axis 1 vertical):
    rho_b=exp(-z/H), isothermal p_b=p0*rho_b with p0=1/gamma, uniform
    vertical B_b with beta=1 at the base, hydrostatic g1=-p0/H; Gaussian
    vertical-velocity pulse plus seeded uniform noise on v1..v3; axis 1
    'fixed', axes 2,3 'periodic'."""
    ndim, ndir = 3, 3
    nx = tuple(nxyz) if nxyz is not None else (n, n, n)
    w, x, v = _alloc(ndim, ndir, nx)
    c = _mesh_coords(3, nx)
    ms = (slice(2, -2),) * 3
    # meshgrid in (axis3, axis2, axis1) order == array order
    Z3, Z2, Z1 = np.meshgrid(c[2], c[1], c[0], indexing="ij")
    x[0][ms], x[1][ms], x[2][ms] = Z1, Z2, Z3
    p0 = 1.0 / gamma
    rhob = np.exp(-Z1 / H)
    pb = p0 * rhob
    bg1 = np.sqrt(2.0 * p0)
    eb = pb / (gamma - 1.0) + 0.5 * bg1 * bg1
    r2 = (Z1 - 0.5) ** 2 + (Z2 - 0.5) ** 2 + (Z3 - 0.5) ** 2
    vz = amp * np.exp(-r2 / (width * width))
    rng = np.random.Generator(np.random.PCG64(seed))
    nz = rng.uniform(-noise, noise, size=(3,) + Z1.shape)
    vel = [vz + nz[0], nz[1], nz[2]]
    rhot = rhob  # rho' = 0
    for d in range(3):
        w[v["m0"] + d + 1 - 1][ms] = rhot * vel[d]
    w[v["e"] - 1][ms] = 0.5 * rhot * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    w[v["eb"] - 1][ms] = eb
    w[v["rhob"] - 1][ms] = rhob
    w[v["bg0"] + 1 - 1][ms] = bg1
    eq = np.zeros(NEQPAR)
    eq[0] = gamma
    eq[GRAV0 + 1 - 1] = -p0 / H   # g1 at grav0_+1
    eq[NU - 1] = nu               # slot 6: nu switch (and g3)
    tb = np.zeros((v["nw"], 6), dtype=np.int32)
    tb[:, 0:2] = BC["fixed"]
    tb[:, 2:6] = BC["periodic"]
    return Problem("stratified", ndim, ndir, nx, w, x, eq, tb,
                   courantpar=0.3, itmax=200,
                   meta={"seed": seed, "amp": amp, "noise": noise, "H": H})


def riemann_tube(n, left, right, gamma=5.0 / 3.0, nu=1.0):
    """1-D MHD Riemann tube, d=13 (ndim=1, ndir=3, nw=13), zero background,
    cont boundaries, diaphragm at x=0.5 on [0,1].  left/right are dicts with
    rho, p, v (3), b (3)."""
    ndim, ndir = 1, 3
    w, x, v = _alloc(ndim, ndir, (n,))
    xc = _mesh_coords(1, (n,))[0]
    x[0, 0, 0, 2:-2] = xc
    L = xc <= 0.5
    rho = np.where(L, left["rho"], right["rho"])
    p = np.where(L, left["p"], right["p"])
    vel = [np.where(L, left["v"][d], right["v"][d]) for d in range(3)]
    b = [np.where(L, left["b"][d], right["b"][d]) for d in range(3)]
    e = p / (gamma - 1.0) + 0.5 * (rho * sum(vv * vv for vv in vel) + sum(bb * bb for bb in b))
    w[0, 0, 0, 2:-2] = rho
    for d in range(3):
        w[v["m0"] + d + 1 - 1, 0, 0, 2:-2] = rho * vel[d]
        w[v["b0"] + d + 1 - 1, 0, 0, 2:-2] = b[d]
    w[v["e"] - 1, 0, 0, 2:-2] = e
    eq = np.zeros(NEQPAR)
    eq[0] = gamma
    eq[NU - 1] = nu
    tb = np.full((v["nw"], 2), BC["cont"], dtype=np.int32)
    return Problem("riemann_tube", ndim, ndir, (n,), w, x, eq, tb,
                   courantpar=0.3, tmax=0.1)


def tube_ensemble(count=4096, nmin=200, nmax=2000, seed=4096):
    """Seeded parameters of the uneven 1-D tube ensemble: N_i ~ U{nmin..nmax};
    per side rho,p ~ U[0.1,1], v ~ U[-0.5,0.5]^3, b2,b3 ~ U[-1,1]; b1 ~ U[0,1]
    shared by both sides."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = rng.integers(nmin, nmax + 1, size=count)
    out = []
    for i in range(count):
        sides = []
        for _ in range(2):
            rho, p = rng.uniform(0.1, 1.0, size=2)
            vel = rng.uniform(-0.5, 0.5, size=3)
            b23 = rng.uniform(-1.0, 1.0, size=2)
            sides.append({"rho": rho, "p": p, "v": vel, "b23": b23})
        b1 = rng.uniform(0.0, 1.0)
        for s in sides:
            s["b"] = np.array([b1, s["b23"][0], s["b23"][1]])
        out.append((int(ns[i]), sides[0], sides[1]))
    return out


def riemann_patch(nx, ny, left, right, gamma=5.0 / 3.0, nu=1.0):
    """2-D MHD Riemann patch, d=23 (ndim=2, ndir=3, nw=13), zero background,
    cont boundaries on [0,1]^2 with nx x ny cells, the two states split by
    the diagonal x + y = 1 (left below it).  The in-plane field b1, b2 must
    be the same on both sides (div B = 0 across the oblique front)."""
    ndim, ndir = 2, 3
    w, x, v = _alloc(ndim, ndir, (nx, ny))
    xc, yc = _mesh_coords(2, (nx, ny))
    X, Y = np.meshgrid(xc, yc)          # (ny, nx), x fastest
    x[0, 0, 2:-2, 2:-2] = X
    x[1, 0, 2:-2, 2:-2] = Y
    L = X + Y <= 1.0
    rho = np.where(L, left["rho"], right["rho"])
    p = np.where(L, left["p"], right["p"])
    vel = [np.where(L, left["v"][d], right["v"][d]) for d in range(3)]
    b = [np.where(L, left["b"][d], right["b"][d]) for d in range(3)]
    e = p / (gamma - 1.0) + 0.5 * (rho * sum(vv * vv for vv in vel) + sum(bb * bb for bb in b))
    w[0, 0, 2:-2, 2:-2] = rho
    for d in range(3):
        w[v["m0"] + d, 0, 2:-2, 2:-2] = rho * vel[d]
        w[v["b0"] + d, 0, 2:-2, 2:-2] = b[d]
    w[v["e"] - 1, 0, 2:-2, 2:-2] = e
    eq = np.zeros(NEQPAR)
    eq[0] = gamma
    eq[NU - 1] = nu
    tb = np.full((v["nw"], 4), BC["cont"], dtype=np.int32)
    return Problem("riemann_patch", ndim, ndir, (nx, ny), w, x, eq, tb,
                   courantpar=0.3, tmax=0.05)


def patch_ensemble(count=1024, nmin=32, nmax=160, seed=2048):
    """Seeded parameters of an uneven 2-D patch ensemble: nx, ny ~
    U{nmin..nmax} each; per side rho,p ~ U[0.1,1], v ~ U[-0.5,0.5]^3,
    b3 ~ U[-1,1]; b1, b2 ~ U[0,1] shared by both sides."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = rng.integers(nmin, nmax + 1, size=(count, 2))
    out = []
    for i in range(count):
        sides = []
        for _ in range(2):
            rho, p = rng.uniform(0.1, 1.0, size=2)
            vel = rng.uniform(-0.5, 0.5, size=3)
            b3 = rng.uniform(-1.0, 1.0)
            sides.append({"rho": rho, "p": p, "v": vel, "b3": b3})
        b12 = rng.uniform(0.0, 1.0, size=2)
        for s in sides:
            s["b"] = np.array([b12[0], b12[1], s["b3"]])
        out.append((int(ns[i, 0]), int(ns[i, 1]), sides[0], sides[1]))
    return out
