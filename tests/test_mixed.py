"""Mixed precision (SAC_PREC_MIXED): evolving fields, fluxes and RK registers
in FP32 with compensated flux differences; background fields, pressure
recovery, register update, cmax/dt and maxviscoef in FP64.

Bar: relative L2 per variable <= 1e-5 against the FP64 oracle (the bar of
the reduced-precision form), except where stated.  The stratified
atmosphere amplifies rounding differences: two FP64 builds of the oracle
that differ only in FMA contraction drift apart by ~1e4 x the unit
roundoff over 200 steps at 32^3, so the horizon of the free-running check
is kept where FP32 rounding (6e-8) stays under the bar.
"""
import numpy as np
import pytest

from sac_amd import problems
from tests.helpers import (make_oracle, oracle_run, per_var_errors, perturbed,
                           zero_boundary_problem)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev_mod():
    from sac_amd import device
    device.load_library()
    return device


def _stage(dev_mod, oracle_mod, prob):
    o = make_oracle(prob, oracle_mod)
    w0 = o.get_w()
    wct = perturbed(type(prob)(**{**prob.__dict__, "w": w0.copy()}), seed=3, amp=1e-3).w
    wct = o.getboundary_on(wct)
    ref = o.advect1(1e-3, wct, w0)
    dev = dev_mod.Device(prob, precision="mixed")
    dev.upload(prob.w)
    dev.getboundary()
    dev.upload_slot(dev_mod.SLOT_S1, wct)
    dev.upload_slot(dev_mod.SLOT_S2, w0)
    dev.advect1(1e-3, dev_mod.SLOT_S1, dev_mod.SLOT_S2)
    got = dev.download_slot(dev_mod.SLOT_S2)
    dev.close()
    ne = problems.var_index(prob.ndir)["nevolve"]
    return (per_var_errors(prob, got[:ne] - w0[:ne], ref[:ne] - w0[:ne]),
            per_var_errors(prob, got, ref, ne))


@pytest.mark.parametrize("nxyz", [(12, 10, 8), (13, 10, 8)])
def test_mixed_stage_stratified(dev_mod, oracle_mod, nxyz):
    """one stage from identical inputs: increments at FP32 resolution (an odd
    x extent leaves the combining update's last cell pair with one cell)"""
    inc, st = _stage(dev_mod, oracle_mod,
                     problems.stratified(n=nxyz[0], nxyz=nxyz, amp=3e-2))
    assert max(inc) <= 1e-6, inc
    assert max(st) <= 1e-6, st


@pytest.mark.parametrize("n", [64, 63])
def test_mixed_stage_shock_tube(dev_mod, oracle_mod, n):
    """one stage across discontinuities: states at FP32 resolution; the
    increments of the smallest fields (e, b2 near the jump) to 1e-4"""
    inc, st = _stage(dev_mod, oracle_mod, problems.brio_wu(n=n))
    assert max(st) <= 1e-5, st
    assert max(inc) <= 1e-4, inc


def test_mixed_stratified_free_running(dev_mod, oracle_mod):
    prob = problems.stratified(n=24)
    steps = 100
    ref, rdts, _ = oracle_run(prob, oracle_mod, steps)
    dev = dev_mod.Device(prob, precision="mixed")
    dev.upload(prob.w)
    dev.getboundary()
    dts = [dev.step(prob.courantpar) for _ in range(steps)]
    w = dev.download()
    dev.close()
    errs = per_var_errors(prob, w, ref, 8)
    assert max(errs) <= 1e-5, errs
    np.testing.assert_allclose(dts, rdts, rtol=1e-3)


def test_mixed_odd_extent_free_running(dev_mod, oracle_mod):
    """odd x extent: the cmax reductions of the combining update count the
    last cell of every row once and nothing past the mesh, so the
    free-running dt follows the oracle's"""
    prob = problems.stratified(n=15, nxyz=(15, 12, 9), amp=3e-2)
    steps = 30
    ref, rdts, _ = oracle_run(prob, oracle_mod, steps)
    dev = dev_mod.Device(prob, precision="mixed")
    dev.upload(prob.w)
    dev.getboundary()
    dts = [dev.step(prob.courantpar) for _ in range(steps)]
    w = dev.download()
    dev.close()
    errs = per_var_errors(prob, w, ref, 8)
    assert max(errs) <= 1e-5, errs
    np.testing.assert_allclose(dts, rdts, rtol=1e-4)


def test_mixed_brio_wu_prescribed_dt(dev_mod, oracle_mod):
    """300 steps across shocks under the oracle's dt sequence (FP32 setnu
    ratios with denormal-safe reciprocals stay finite and accurate)"""
    prob = problems.brio_wu(n=400)
    ref, rdts, _ = oracle_run(prob, oracle_mod, 300)
    dev = dev_mod.Device(prob, precision="mixed")
    dev.upload(prob.w)
    dev.getboundary()
    for dt in rdts:
        dev.advance(dt)
    w = dev.download()
    dev.close()
    assert np.isfinite(w).all()
    errs = per_var_errors(prob, w, ref, 6)
    assert max(errs) <= 1e-5, errs


@pytest.mark.parametrize("ndim", [1, 3])
def test_mixed_zero_boundary(dev_mod, oracle_mod, ndim):
    """'zero' boundaries in mixed precision (the primitive/conserve round
    trips of each fill on the FP32 register, in FP64 arithmetic): three RK4
    steps under the oracle's dt sequence at FP32 resolution"""
    prob = zero_boundary_problem(ndim)
    ref, rdts, _ = oracle_run(prob, oracle_mod, 3)
    dev = dev_mod.Device(prob, precision="mixed")
    dev.upload(prob.w)
    dev.getboundary()
    for dt in rdts:
        dev.advance(dt)
    w = dev.download()
    dev.close()
    errs = per_var_errors(prob, w, ref, 8)
    assert max(errs) <= 1e-5, errs


def test_mixed_memory_is_smaller(dev_mod):
    prob = problems.stratified(n=16)
    d64 = dev_mod.Device(prob, precision="fp64")
    d32 = dev_mod.Device(prob, precision="mixed")
    m64, m32 = d64.memory(), d32.memory()
    d64.close()
    d32.close()
    assert m32 < 0.65 * m64, (m32, m64)
