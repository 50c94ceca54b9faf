"""Textbook ideal-MHD quantities on the total (background + perturbation)
state, written independently of the oracle, in conservative variables
U = (rho, m, E, B) with p = (gamma-1)(E - m^2/(2 rho) - B^2/2) (Heaviside-
Lorentz units with mu=1, as in VAC).  Used to pin the oracle's split
perturbation form: for a static background in equilibrium the time
derivative of the perturbation must equal the textbook right-hand side of
the total state minus that of the background.
"""
import numpy as np


def total_state(w, ndir):
    """(rho, m[ndir], E, B[ndir]) of the total state from reference-layout w"""
    e = ndir + 2
    eb = e + ndir + 1
    rho = w[0] + w[eb]  # rho' + rhob (rhob is variable eb+1, 1-based)
    m = [w[c] for c in range(1, ndir + 1)]
    E = w[e - 1] + w[eb - 1]
    B = [w[e + c - 1] + w[eb + 1 + c - 1] for c in range(1, ndir + 1)]
    return rho, m, E, B


def background_state(w, ndir):
    e = ndir + 2
    eb = e + ndir + 1
    rho = w[eb]
    m = [np.zeros_like(rho) for _ in range(ndir)]
    E = w[eb - 1].copy()
    B = [w[eb + 1 + c - 1].copy() for c in range(1, ndir + 1)]
    return rho, m, E, B


def flux(U, d, gamma):
    """ideal-MHD flux along axis d of conservative U; returns list in the
    order rho, m_1..m_ndir, E, B_1..B_ndir"""
    rho, m, E, B = U
    ndir = len(m)
    v = [mc / rho for mc in m]
    B2 = sum(b * b for b in B)
    p = (gamma - 1.0) * (E - 0.5 * sum(mc * vc for mc, vc in zip(m, v)) - 0.5 * B2)
    ptot = p + 0.5 * B2
    vB = sum(vc * b for vc, b in zip(v, B))
    out = [m[d]]
    for c in range(ndir):
        f = m[c] * v[d] - B[c] * B[d]
        if c == d:
            f = f + ptot
        out.append(f)
    out.append((E + ptot) * v[d] - B[d] * vB)
    for c in range(ndir):
        out.append(v[d] * B[c] - B[d] * v[c])
    return out


def d4(q, d, dx_axis):
    """4th-order central derivative along reference axis d (0-based) at the
    mesh cells, zero on the two outermost mesh layers of every side (the
    serial reference's edge treatment) and outside the mesh.  q has shape
    (G3,G2,G1); dx_axis holds dx along axis d for every grid index."""
    a = 2 - d
    G = q.shape[a]
    out = np.zeros_like(q)
    idx = np.arange(4, G - 4)  # mesh 2..G-3 without two layers per side
    def take(off):
        return np.take(q, idx + off, axis=a)
    shape = [1, 1, 1]
    shape[a] = len(idx)
    dxv = dx_axis[idx].reshape(shape)
    val = (-take(2) + 8.0 * take(1) - 8.0 * take(-1) + take(-2)) / (12.0 * dxv)
    sl = [slice(None)] * 3
    sl[a] = idx
    out[tuple(sl)] = val
    return out


def rhs_flux(w, ndir, ndim, gamma, dx_axes, grav=None, background=False):
    """-sum_d D4_d(F_d(U)) + gravity source for the total (or background)
    state; returns list ordered like the evolving variables."""
    U = background_state(w, ndir) if background else total_state(w, ndir)
    nv = 2 + 2 * ndir
    R = [np.zeros_like(w[0]) for _ in range(nv)]
    for d in range(ndim):
        F = flux(U, d, gamma)
        for i in range(nv):
            R[i] = R[i] - d4(F[i], d, dx_axes[d])
    if grav is not None:
        rho, m, E, B = U
        for d in range(ndim):
            if grav[d] != 0.0:
                R[1 + d] = R[1 + d] + rho * grav[d]
                R[1 + ndir] = R[1 + ndir] + m[d] * grav[d]
    return R


def fast_speed(rho, p, B, d, gamma):
    """fast magnetosonic speed along axis d (textbook closed form)"""
    cs2 = gamma * p / rho
    a2 = sum(b * b for b in B) / rho
    ad2 = B[d] * B[d] / rho
    return np.sqrt(0.5 * (a2 + cs2 + np.sqrt((a2 + cs2) ** 2 - 4.0 * ad2 * cs2)))
