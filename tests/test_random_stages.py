"""One RK stage (advect1: fluxes, hyperdiffusion, gravity, the register
update and maxviscoef) on seeded random grids against the oracle: random
extents, random boundary pairs per axis, random perturbations, both 3-D
schedules (plane tiles and one line tile per axis), 2-D patches and 1-D
tubes; and three RK4 steps on 8 more.  Bar as for the fixed cases:
increments rel-L2 <= 1e-11 per variable, states 1e-12, maxviscoef 1e-12;
the 3-D cases in mixed precision at FP32 resolution (increments 1e-5,
states 1e-6, maxviscoef 1e-5)."""
import numpy as np
import pytest

from sac_amd import problems
from tests.helpers import assert_close_state, make_oracle, per_var_errors, perturbed

pytestmark = pytest.mark.gpu

# boundary pairs (lower, upper) of one axis; periodic only as a pair
PAIRS = [("periodic", "periodic"), ("cont", "cont"), ("fixed", "fixed"), ("cont1", "grad1"),
         ("symm", "asymm"), ("fixed1", "cont"), ("contCD4", "symm0"), ("asymm", "fixed")]
LEFT = {"rho": 1.0, "p": 0.8, "v": [0.1, -0.2, 0.3], "b": [0.6, 0.5, -0.4]}
RIGHT = {"rho": 0.3, "p": 0.2, "v": [-0.1, 0.2, 0.0], "b": [0.6, -0.5, 0.4]}


@pytest.fixture(scope="module")
def dev_mod():
    from sac_amd import device
    device.load_library()
    return device


def _problem(ndim, seed):
    rng = np.random.default_rng(seed)
    if ndim == 3:
        nxyz = tuple(int(x) for x in rng.integers(6, 21, size=3))
        prob = problems.stratified(n=nxyz[0], nxyz=nxyz, amp=3e-2)
    elif ndim == 2:
        prob = problems.riemann_patch(int(rng.integers(8, 40)), int(rng.integers(8, 40)),
                                      LEFT, RIGHT)
    else:
        prob = problems.riemann_tube(int(rng.integers(12, 120)), LEFT, RIGHT)
    codes = []
    for _ in range(ndim):
        codes += PAIRS[int(rng.integers(len(PAIRS)))]
    prob.typeB[:, :] = np.array([problems.BC[c] for c in codes], dtype=np.int32)[None, :]
    return prob, codes


def _stage(dev_mod, oracle_mod, prob, seed, plane, precision="fp64"):
    o = make_oracle(prob, oracle_mod)
    w0 = o.get_w()
    wct = perturbed(type(prob)(**{**prob.__dict__, "w": w0.copy()}), seed=seed, amp=1e-3).w
    wct = o.getboundary_on(wct)
    ref = o.advect1(1e-4, wct, w0)
    dev = dev_mod.Device(prob, precision=precision)
    if plane is not None:
        dev.set_plane_tiles(plane)
    dev.upload(prob.w)
    dev.getboundary()
    dev.upload_slot(dev_mod.SLOT_S1, wct)
    dev.upload_slot(dev_mod.SLOT_S2, w0)
    dev.advect1(1e-4, dev_mod.SLOT_S1, dev_mod.SLOT_S2)
    got = dev.download_slot(dev_mod.SLOT_S2)
    mv = dev.maxviscoef
    dev.close()
    return got, ref, w0, mv, o.maxviscoef


@pytest.mark.parametrize("ndim,seed", [(3, s) for s in range(12)] + [(2, s) for s in range(6)] +
                         [(1, s) for s in range(6)])
def test_random_stage_matches_oracle(dev_mod, oracle_mod, ndim, seed):
    prob, codes = _problem(ndim, 1000 * ndim + seed)
    plane = bool(seed % 2) if ndim == 3 else None
    got, ref, w0, mv, mv_ref = _stage(dev_mod, oracle_mod, prob, seed, plane)
    ne = problems.var_index(prob.ndir)["nevolve"]
    what = f"{ndim}-D {prob.nx} {codes} plane={plane}"
    errs = per_var_errors(prob, got[:ne] - w0[:ne], ref[:ne] - w0[:ne])
    assert max(errs) <= 1e-11, f"{what}: increment rel-L2 {errs}"
    assert_close_state(prob, got, ref, 1e-12, ne, what)
    assert mv == pytest.approx(mv_ref, rel=1e-12), what


@pytest.mark.parametrize("seed", range(8))
def test_random_stage_mixed_matches_oracle(dev_mod, oracle_mod, seed):
    prob, codes = _problem(3, 5000 + seed)
    plane = bool(seed % 2)
    got, ref, w0, mv, mv_ref = _stage(dev_mod, oracle_mod, prob, seed, plane, "mixed")
    what = f"mixed {prob.nx} {codes} plane={plane}"
    errs = per_var_errors(prob, got[:8] - w0[:8], ref[:8] - w0[:8])
    print(what, "increments", " ".join(f"{e:.1e}" for e in errs),
          "states", " ".join(f"{e:.1e}" for e in per_var_errors(prob, got, ref, 8)))
    assert max(errs) <= 1e-5, f"{what}: increment rel-L2 {errs}"
    assert_close_state(prob, got, ref, 1e-6, 8, what)
    assert mv == pytest.approx(mv_ref, rel=1e-5), what


@pytest.mark.parametrize("ndim,seed", [(3, s) for s in range(4)] + [(2, s) for s in range(2)] +
                         [(1, s) for s in range(2)])
def test_random_steps_match_oracle(dev_mod, oracle_mod, ndim, seed):
    """three RK4 steps (the last stage's (w1 + 2 w2 + w3 - 4 w)/3 and every
    ghost fill of the random boundary pairs) under the oracle's dt sequence"""
    from tests.helpers import oracle_run
    prob, codes = _problem(ndim, 7000 + 100 * ndim + seed)
    ref, dts, _ = oracle_run(prob, oracle_mod, 3)
    dev = dev_mod.Device(prob)
    if ndim == 3:
        dev.set_plane_tiles(bool(seed % 2))
    dev.upload(prob.w)
    dev.getboundary()
    for dt in dts:
        dev.advance(dt)
    w = dev.download()
    dev.close()
    ne = problems.var_index(prob.ndir)["nevolve"]
    assert_close_state(prob, w, ref, 1e-12, ne, f"{ndim}-D {prob.nx} {codes}")
