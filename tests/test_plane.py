"""3-D stages with the x and y work in one pass over z-plane tiles
(sac_dev_set_plane_tiles, on by default): one stage against the oracle with
tiles that end inside the mesh and for every boundary type along x and y
(mirror layers, edge zeroing, periodic neighbours), and whole runs against
the per-axis line-tile schedule."""
import numpy as np
import pytest

from sac_amd import problems
from tests.helpers import assert_close_state, make_oracle, per_var_errors, perturbed

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev_mod():
    from sac_amd import device
    device.load_library()
    return device


def _stage(dev_mod, oracle_mod, prob, plane):
    o = make_oracle(prob, oracle_mod)
    w0 = o.get_w()
    wct = perturbed(type(prob)(**{**prob.__dict__, "w": w0.copy()}), seed=5, amp=1e-3).w
    wct = o.getboundary_on(wct)
    ref = o.advect1(1e-3, wct, w0)
    dev = dev_mod.Device(prob)
    dev.set_plane_tiles(plane)
    dev.upload(prob.w)
    dev.getboundary()
    dev.upload_slot(dev_mod.SLOT_S1, wct)
    dev.upload_slot(dev_mod.SLOT_S2, w0)
    dev.advect1(1e-3, dev_mod.SLOT_S1, dev_mod.SLOT_S2)
    got = dev.download_slot(dev_mod.SLOT_S2)
    mv = dev.maxviscoef
    dev.close()
    return got, ref, w0, mv, o.maxviscoef


@pytest.mark.parametrize("nxyz,codes", [
    ((20, 18, 9), ["periodic", "periodic", "fixed", "fixed", "cont", "cont"]),
    ((35, 17, 6), ["cont", "fixed", "periodic", "periodic", "periodic", "periodic"]),
    ((16, 16, 7), ["fixed", "fixed", "periodic", "periodic", "periodic", "periodic"]),
    ((9, 33, 8), ["cont1", "grad1", "symm", "asymm", "fixed", "cont"]),
])
def test_plane_stage_matches_oracle(dev_mod, oracle_mod, nxyz, codes):
    prob = problems.stratified(n=nxyz[0], nxyz=nxyz, amp=3e-2)
    prob.typeB[:, :] = np.array([problems.BC[c] for c in codes], dtype=np.int32)[None, :]
    got, ref, w0, mv, mv_ref = _stage(dev_mod, oracle_mod, prob, True)
    errs = per_var_errors(prob, got[:8] - w0[:8], ref[:8] - w0[:8])
    assert max(errs) <= 1e-11, f"increment rel-L2 {errs}"
    assert_close_state(prob, got, ref, 1e-12, 8, str(codes))
    assert mv == pytest.approx(mv_ref, rel=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_plane_run_equals_line_run(dev_mod, precision):
    """The two schedules evaluate the same terms; only the order in which the
    x-y cross terms join the rate differs, so free runs agree to rounding."""
    prob = problems.stratified(n=24, nxyz=(24, 20, 12), amp=3e-2)
    out = []
    for plane in (True, False):
        dev = dev_mod.Device(prob, precision=precision)
        dev.set_plane_tiles(plane)
        dev.upload(prob.w)
        dev.getboundary()
        dts = [dev.step(prob.courantpar) for _ in range(10)]
        out.append((dev.download(), np.array(dts)))
        dev.close()
    tol = 1e-12 if precision == "fp64" else 1e-5
    errs = per_var_errors(prob, out[0][0], out[1][0], 8)
    assert max(errs) <= tol, errs
    np.testing.assert_allclose(out[0][1], out[1][1], rtol=tol)
