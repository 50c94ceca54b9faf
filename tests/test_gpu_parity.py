"""GPU parity: libsac_amd.so (through its C ABI) against the CPU oracle.

Bar: per-variable relative L2 <= 1e-12 on the mesh for FP64 (the tolerance
of the FP64 path), identical dt sequences to 1e-12.  Sizes are small so the
oracle finishes in seconds.
"""
import numpy as np
import pytest

from sac_amd import problems
from tests.helpers import (assert_close_state, make_oracle, mesh, oracle_run,
                           perturbed, per_var_errors, zero_boundary_problem)

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def dev_mod():
    from sac_amd import device
    device.load_library()
    return device


def _gpu_run(dev_mod, prob, steps, typeadvance="fourstep"):
    dev = dev_mod.Device(prob, typeadvance=typeadvance)
    dev.upload(prob.w)
    dev.getboundary()
    dts = [dev.step(prob.courantpar) for _ in range(steps)]
    w = dev.download()
    dev.close()
    return w, dts


def test_getboundary_all_types(dev_mod, oracle_mod):
    """ghost fill of every boundary type against the sequential reference
    fill, on index-coded data (each cell value unique)"""
    for codes in ([1, 1, 3, 3, 2, 2], [3, 3, 1, 1, 3, 3], [4, 6, 2, 1, 3, 3],
                  [6, 4, 3, 3, 5, 2], [7, 8, 9, 10, 3, 3], [10, 9, 8, 7, 7, 8],
                  [9, 9, 10, 10, 8, 7], [3, 3, 9, 8, 10, 4]):
        prob = problems.stratified(n=8, nxyz=(8, 7, 6))
        prob.typeB[:, :] = np.array(codes, dtype=np.int32)[None, :]
        # asymm on a different variable set than symm
        prob.typeB[3, 0] = 1
        rng = np.random.Generator(np.random.PCG64(7))
        w = prob.w.copy()
        ms = prob.mesh_slices()
        for iw in range(w.shape[0]):
            w[iw][ms] = rng.uniform(0.5, 1.5, size=w[iw][ms].shape)
        prob.w = w
        o = make_oracle(prob, oracle_mod)
        ref = o.get_w()
        dev = dev_mod.Device(prob)
        dev.upload(prob.w)
        dev.getboundary()
        got = dev.download()
        np.testing.assert_array_equal(got, ref, err_msg=f"codes {codes}")
        # second fill (fixed values restored from storage)
        w2 = got.copy()
        w2[:8][ms] *= 1.25
        dev.upload(w2)
        dev.getboundary()
        got2 = dev.download()
        ref2 = o.getboundary_on(w2)
        np.testing.assert_array_equal(got2[:8], ref2[:8], err_msg=f"refill {codes}")
        dev.close()


def test_upload_after_first_fill_refills_background_ghosts(dev_mod):
    """An upload after the first getboundary brings the caller's background
    ghost cells (zero in a fresh problem array); the next getboundary fills
    them again, so the run equals one on a fresh handle bit for bit."""
    prob = problems.stratified(n=24, nxyz=(24, 20, 12), amp=0.1)
    dts = [2e-4, 1e-4, 1e-4]
    dev = dev_mod.Device(prob)
    dev.upload(prob.w)
    dev.getboundary()
    for dt in dts:
        dev.advance(dt)
    ref = dev.download()
    dev.close()
    dev = dev_mod.Device(prob)
    dev.upload(prob.w)
    dev.getboundary()
    dev.advance(dts[0])
    dev.advance(dts[1])
    dev.upload(prob.w)
    dev.getboundary()
    for dt in dts:
        dev.advance(dt)
    got = dev.download()
    dev.close()
    assert np.isfinite(got).all()
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("case", ["briowu", "riemann3", "strat3d", "strat2d"])
def test_advect1_stage(dev_mod, oracle_mod, case):
    """one advect1 ('cd4' + sources + getboundary) from an arbitrary wCT"""
    if case == "briowu":
        prob = problems.brio_wu(n=64)
    elif case == "riemann3":
        prob = problems.riemann_tube(96, {"rho": 1.0, "p": 0.8, "v": [0.1, -0.2, 0.3],
                                          "b": [0.6, 0.5, -0.4]},
                                     {"rho": 0.3, "p": 0.2, "v": [-0.1, 0.2, 0.0],
                                      "b": [0.6, -0.7, 0.2]})
    elif case == "strat3d":
        prob = problems.stratified(n=12, nxyz=(12, 10, 8), amp=3e-2)
    else:
        prob = problems.stratified(n=12, nxyz=(14, 12, 1), amp=3e-2)
        prob = _to_2d(prob)
    o = make_oracle(prob, oracle_mod)
    w0 = o.get_w()
    wct = perturbed(type(prob)(**{**prob.__dict__, "w": w0.copy()}), seed=3, amp=1e-3).w
    wct = o.getboundary_on(wct)
    qdt = 1e-3
    ref = o.advect1(qdt, wct, w0)
    dev = dev_mod.Device(prob)
    dev.upload(prob.w)
    dev.getboundary()
    dev.upload_slot(dev_mod.SLOT_S1, wct)
    dev.upload_slot(dev_mod.SLOT_S2, w0)
    dev.advect1(qdt, dev_mod.SLOT_S1, dev_mod.SLOT_S2)
    got = dev.download_slot(dev_mod.SLOT_S2)
    dev.close()
    ne = problems.var_index(prob.ndir)["nevolve"]
    # compare the increments, which is where the stage's arithmetic shows
    errs = per_var_errors(prob, got[:ne] - w0[:ne], ref[:ne] - w0[:ne])
    assert max(errs) <= 1e-11, f"increment rel-L2 {errs}"
    assert_close_state(prob, got, ref, TOL, ne, case)


def _to_2d(prob):
    """drop axis 3 of a stratified problem built with nz=1 -> ndim=2, ndir=3"""
    ndim, ndir = 2, 3
    nx = prob.nx[:2]
    v = problems.var_index(ndir)
    G = (nx[0] + 4, nx[1] + 4)
    w = np.zeros((v["nw"], 1, G[1], G[0]))
    x = np.zeros((2, 1, G[1], G[0]))
    w[:, 0, :, :] = prob.w[:, 2, :, :]
    x[0, 0] = prob.x[0, 2]
    x[1, 0] = prob.x[1, 2]
    tb = np.zeros((v["nw"], 4), dtype=np.int32)
    tb[:, 0:2] = problems.BC["fixed"]
    tb[:, 2:4] = problems.BC["periodic"]
    return problems.Problem("strat2d", ndim, ndir, nx, w, x, prob.eqpar.copy(), tb,
                            courantpar=0.3)


def _gpu_run_prescribed(dev_mod, prob, dts):
    """advance with a prescribed dt sequence (the oracle's)"""
    dev = dev_mod.Device(prob)
    dev.upload(prob.w)
    dev.getboundary()
    for dt in dts:
        dev.advance(dt)
    w = dev.download()
    dev.close()
    return w


def test_briowu_steps(dev_mod, oracle_mod):
    """Shock tubes: the dt of the reference is set by maxviscoef, which in
    the exactly-constant regions of a shock tube is a ratio of rounding-level
    differences; any change of rounding changes it (two builds of the oracle
    itself diverge, tests/test_dt_conditioning.py).  The state update is
    therefore compared under the oracle's own dt sequence."""
    prob = problems.brio_wu(n=800)
    steps = 200
    ref, rdts, _ = oracle_run(prob, oracle_mod, steps)
    got = _gpu_run_prescribed(dev_mod, prob, rdts)
    assert_close_state(prob, got, ref, TOL, 6, "brio-wu")


@pytest.mark.parametrize("case", ["briowu", "strat3d", "tube_constant_regions", "patch2d"])
def test_stage_maxviscoef_and_dt(dev_mod, oracle_mod, case):
    """maxviscoef gathered by one stage, and the courant dt, on identical
    inputs.  The tube case has exactly constant regions with nonzero
    velocity: setnu's ratio there must come out as exact zeros (no rounding
    residue from a product fused into the q differences).  The 2-D patch
    (an oblique front between two uniform states) after one step has both;
    on it the oracle's and the GPU's maxviscoef of the later RK stages of
    the next step part by up to 2e-5 on identical inputs (faces whose
    md3/md1 is a ratio of rounding-level differences), so the check is
    made on the state after a full step."""
    if case == "patch2d":
        nx, ny, left, right = problems.patch_ensemble(count=3, nmin=12, nmax=40, seed=9)[1]
        prob = problems.riemann_patch(nx, ny, left, right)
        nsteps = 1
    elif case == "briowu":
        prob = problems.brio_wu(n=400)
        nsteps = 3
    elif case == "tube_constant_regions":
        n, left, right = problems.tube_ensemble(count=1, nmin=117, nmax=117, seed=5)[0]
        prob = problems.riemann_tube(n, left, right)
        nsteps = 0
    else:
        prob = problems.stratified(n=12, nxyz=(12, 10, 8))
        nsteps = 2
    o = make_oracle(prob, oracle_mod)
    for _ in range(nsteps):
        o.step(prob.courantpar)
    w = o.get_w()
    dev = dev_mod.Device(prob)
    dev.upload(prob.w)
    dev.getboundary()
    dev.upload(w)
    dtc_o, cm_o = o.getdt_courant(prob.courantpar)
    dtc_g, cm_g = dev.getdt_courant(prob.courantpar)
    assert cm_g == pytest.approx(cm_o, rel=1e-14)
    o.maxviscoef = 0.0
    ref = o.advect1(1e-5, w, w)
    dev.upload_slot(dev_mod.SLOT_S1, w)
    dev.upload_slot(dev_mod.SLOT_S2, w)
    dev.maxviscoef = 0.0
    dev.advect1(1e-5, dev_mod.SLOT_S1, dev_mod.SLOT_S2)
    assert dev.maxviscoef == pytest.approx(o.maxviscoef, rel=1e-13)
    dt_o = o.getdt_special(dtc_o)
    dt_g = dev.getdt_special(dtc_g)
    assert dt_g == pytest.approx(dt_o, rel=1e-13)
    dev.close()


def test_stratified_steps(dev_mod, oracle_mod):
    prob = problems.stratified(n=16, nxyz=(16, 12, 10))
    steps = 8
    ref, rdts, _ = oracle_run(prob, oracle_mod, steps)
    got, gdts = _gpu_run(dev_mod, prob, steps)
    np.testing.assert_allclose(gdts, rdts, rtol=1e-12)
    assert_close_state(prob, got, ref, TOL, 8, "stratified")


def test_onestep(dev_mod, oracle_mod):
    prob = problems.brio_wu(n=200)
    prob.courantpar = 0.05  # courant-limited: dt is well conditioned
    ref, rdts, _ = oracle_run(prob, oracle_mod, 20, typeadvance="onestep")
    got, gdts = _gpu_run(dev_mod, prob, 20, typeadvance="onestep")
    np.testing.assert_allclose(gdts, rdts, rtol=1e-12)
    assert_close_state(prob, got, ref, TOL, 6, "onestep")


@pytest.mark.parametrize("adv", ["twostep", "threestep", "sterck", "jameson"])
def test_integrators(dev_mod, oracle_mod, adv):
    """the remaining Runge-Kutta integrators of src/vac.t:406-475 against the
    oracle: 3-D stratified (free-running, well-conditioned dt) and a
    courant-limited Brio-Wu tube"""
    prob = problems.stratified(n=16, nxyz=(16, 12, 10))
    ref, rdts, _ = oracle_run(prob, oracle_mod, 6, typeadvance=adv)
    got, gdts = _gpu_run(dev_mod, prob, 6, typeadvance=adv)
    np.testing.assert_allclose(gdts, rdts, rtol=1e-12)
    assert_close_state(prob, got, ref, TOL, 8, adv)
    prob = problems.brio_wu(n=200)
    prob.courantpar = 0.05
    ref, rdts, _ = oracle_run(prob, oracle_mod, 15, typeadvance=adv)
    got, gdts = _gpu_run(dev_mod, prob, 15, typeadvance=adv)
    np.testing.assert_allclose(gdts, rdts, rtol=1e-12)
    assert_close_state(prob, got, ref, TOL, 6, adv)


@pytest.mark.parametrize("case", ["strat3d", "strat2d"])
def test_powell_divb_stage(dev_mod, oracle_mod, case):
    """one advect1 stage with Powell's div B sources (divbfix) against the
    oracle, on a state whose perturbed field has a nonzero divergence"""
    if case == "strat3d":
        prob = problems.stratified(n=12, nxyz=(12, 10, 8), amp=3e-2)
    else:
        prob = _to_2d(problems.stratified(n=12, nxyz=(14, 12, 1), amp=3e-2))
    rng = np.random.default_rng(5)
    v = problems.var_index(prob.ndir)
    w = prob.w.copy()
    ms = prob.mesh_slices()
    for c in range(prob.ndir):
        w[v["b0"] + c][ms] += 1e-2 * rng.uniform(-1, 1, size=w[v["b0"] + c][ms].shape)
    prob.w = w
    o = make_oracle(prob, oracle_mod)
    o.set_divbfix(True)
    w0 = o.get_w()
    wct = perturbed(type(prob)(**{**prob.__dict__, "w": w0.copy()}), seed=3, amp=1e-3).w
    wct = o.getboundary_on(wct)
    ref = o.advect1(1e-3, wct, w0)
    dev = dev_mod.Device(prob)
    dev.set_divbfix(True)
    dev.upload(prob.w)
    dev.getboundary()
    dev.upload_slot(dev_mod.SLOT_S1, wct)
    dev.upload_slot(dev_mod.SLOT_S2, w0)
    dev.advect1(1e-3, dev_mod.SLOT_S1, dev_mod.SLOT_S2)
    got = dev.download_slot(dev_mod.SLOT_S2)
    dev.close()
    ne = v["nevolve"]
    errs = per_var_errors(prob, got[:ne] - w0[:ne], ref[:ne] - w0[:ne])
    assert max(errs) <= 1e-11, errs
    # the sources are really on: the same stage without them differs
    o.set_divbfix(False)
    ref_off = o.advect1(1e-3, wct, w0)
    assert max(per_var_errors(prob, ref_off[:ne] - w0[:ne], ref[:ne] - w0[:ne])) > 1e-6


@pytest.mark.parametrize("case,compact", [("strat3d", False), ("strat3d", True),
                                          ("strat2d", False), ("riemann3", True)])
def test_resistive_stage(dev_mod, oracle_mod, case, compact):
    """one advect1 stage with resistive sources (constant eta) against the
    oracle; and the resistive dt limit of getdt_special"""
    if case == "strat3d":
        prob = problems.stratified(n=12, nxyz=(12, 10, 8), amp=3e-2)
    elif case == "strat2d":
        prob = _to_2d(problems.stratified(n=12, nxyz=(14, 12, 1), amp=3e-2))
    else:
        prob = problems.riemann_tube(96, {"rho": 1.0, "p": 0.8, "v": [0.1, -0.2, 0.3],
                                          "b": [0.6, 0.5, -0.4]},
                                     {"rho": 0.3, "p": 0.2, "v": [-0.1, 0.2, 0.0],
                                      "b": [0.6, -0.7, 0.2]})
    prob.eqpar[1] = 3e-3
    o = make_oracle(prob, oracle_mod)
    o.set_resistive(compact, 0.4)
    w0 = o.get_w()
    wct = perturbed(type(prob)(**{**prob.__dict__, "w": w0.copy()}), seed=3, amp=1e-3).w
    wct = o.getboundary_on(wct)
    ref = o.advect1(1e-3, wct, w0)
    dev = dev_mod.Device(prob)
    dev.set_resistive(compact, 0.4)
    dev.upload(prob.w)
    dev.getboundary()
    dev.upload_slot(dev_mod.SLOT_S1, wct)
    dev.upload_slot(dev_mod.SLOT_S2, w0)
    dev.advect1(1e-3, dev_mod.SLOT_S1, dev_mod.SLOT_S2)
    got = dev.download_slot(dev_mod.SLOT_S2)
    ne = problems.var_index(prob.ndir)["nevolve"]
    errs = per_var_errors(prob, got[:ne] - w0[:ne], ref[:ne] - w0[:ne])
    assert max(errs) <= 1e-11, errs
    # the dt limit before and after getdt_visc (which sets dtdiffpar=0.25)
    assert dev.getdt_special(1e30) == pytest.approx(o.getdt_special(o.getdt_res(1e30)), rel=1e-13)
    assert dev.getdt_special(1e30) == pytest.approx(o.getdt_special(o.getdt_res(1e30)), rel=1e-13)
    dev.close()
    # the sources are really on
    prob0 = type(prob)(**{**prob.__dict__, "eqpar": prob.eqpar.copy()})
    prob0.eqpar[1] = 0.0
    ref0 = make_oracle(prob0, oracle_mod).advect1(1e-3, wct, w0)
    assert max(per_var_errors(prob, ref0[:ne] - w0[:ne], ref[:ne] - w0[:ne])) > 1e-8


def test_resistive_free_steps(dev_mod, oracle_mod):
    """free-running steps with eta > 0 (the resistive dt limit binds) against
    the oracle"""
    prob = problems.stratified(n=12, nxyz=(12, 10, 8), amp=3e-2)
    prob.eqpar[1] = 1.0  # 0.5 (1/12)^2 / 1 = 3.5e-3, below the courant dt
    o = make_oracle(prob, oracle_mod)
    o.set_resistive(False, 0.5)
    rdts = [o.step(prob.courantpar, "fourstep") for _ in range(5)]
    dev = dev_mod.Device(prob)
    dev.set_resistive(False, 0.5)
    dev.upload(prob.w)
    dev.getboundary()
    gdts = [dev.step(prob.courantpar) for _ in range(5)]
    got = dev.download()
    dev.close()
    np.testing.assert_allclose(gdts, rdts, rtol=1e-12)
    assert_close_state(prob, got, o.get_w(), TOL, 8, "resistive")


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_zero_boundary_matches_oracle(dev_mod, oracle_mod, ndim):
    """'zero' boundaries (src/vacgrid.t:294-345): every such fill converts the
    whole grid to primitive variables and back, so the library fills region
    by region in the reference's order.  The first fill (ghosts of a fresh
    problem array, background ghosts zero, so the conversions divide by zero
    in the ghosts not yet filled, as the reference's do) and a refill match
    the oracle bit for bit, NaN included; then a few RK4 steps under the
    oracle's dt sequence match to 1e-12."""
    prob = zero_boundary_problem(ndim)
    o = make_oracle(prob, oracle_mod)
    ref = o.get_w()
    dev = dev_mod.Device(prob)
    dev.upload(prob.w)
    dev.getboundary()
    got = dev.download()
    np.testing.assert_array_equal(got, ref)
    w2 = got.copy()
    ms = prob.mesh_slices()
    w2[:8][(slice(None),) + ms] *= 1.001
    dev.upload(w2)
    dev.getboundary()
    got2 = dev.download()
    ref2 = o.getboundary_on(w2)
    np.testing.assert_array_equal(got2[:8], ref2[:8])
    dev.close()
    assert np.isfinite(mesh(prob, ref2[:8])).all()
    steps = 3
    w_ref, dts, _ = oracle_run(prob, oracle_mod, steps)
    w_gpu = _gpu_run_prescribed(dev_mod, prob, dts)
    assert_close_state(prob, w_gpu, w_ref, TOL, 8, f"zero {ndim}-D")
