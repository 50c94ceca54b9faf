"""Host-side checks of the ensemble generators (no GPU): the seeded draws
follow the stated distributions, and each initial state is a consistent
MHD state (positive pressure, div B = 0 across the discontinuity)."""
import numpy as np

from sac_amd import problems


def test_tube_ensemble_draws_are_seeded_and_in_range():
    a = problems.tube_ensemble(count=64, nmin=200, nmax=2000, seed=4096)
    b = problems.tube_ensemble(count=64, nmin=200, nmax=2000, seed=4096)
    assert [x[0] for x in a] == [x[0] for x in b]
    for n, left, right in a:
        assert 200 <= n <= 2000
        for s in (left, right):
            assert 0.1 <= s["rho"] <= 1.0 and 0.1 <= s["p"] <= 1.0
            assert np.all(np.abs(s["v"]) <= 0.5)
        assert left["b"][0] == right["b"][0]  # b1 shared: div B = 0 along x


def test_patch_ensemble_draws_and_states():
    draws = problems.patch_ensemble(count=32, nmin=32, nmax=160, seed=2048)
    again = problems.patch_ensemble(count=32, nmin=32, nmax=160, seed=2048)
    assert [(d[0], d[1]) for d in draws] == [(d[0], d[1]) for d in again]
    assert all(np.array_equal(d[2]["b"], e[2]["b"]) for d, e in zip(draws, again))
    v = problems.var_index(3)
    for nx, ny, left, right in draws:
        assert 32 <= nx <= 160 and 32 <= ny <= 160
        # in-plane field shared by both sides: the normal component across
        # the oblique front x + y = 1 is continuous
        assert np.array_equal(left["b"][:2], right["b"][:2])
    nx, ny, left, right = draws[0]
    p = problems.riemann_patch(nx, ny, left, right)
    assert p.w.shape == (13, 1, ny + 4, nx + 4) and p.ndim == 2 and p.ndir == 3
    m = p.w[:, 0, 2:-2, 2:-2]
    rho = m[0]
    mom = m[1:4]
    e = m[v["e"] - 1]
    b = m[v["b0"]:v["b0"] + 3]
    gamma = p.eqpar[0]
    pth = (gamma - 1.0) * (e - 0.5 * (np.sum(mom ** 2, axis=0) / rho + np.sum(b ** 2, axis=0)))
    assert np.all(pth > 0.0)
    # the two states sit on either side of the diagonal
    xc = (np.arange(nx) + 0.5) / nx
    yc = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(xc, yc)
    assert np.allclose(rho[X + Y <= 1.0], left["rho"]) and np.allclose(rho[X + Y > 1.0], right["rho"])
    # ghosts stay zero until getboundary fills them
    assert not p.w[:, :, :2, :].any() and not p.w[:, :, :, -2:].any()
