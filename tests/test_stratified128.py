"""The FP64 path at the stratified 128^3 configuration, 200 RK4 steps,
against the oracle's run of the same configuration
(tests/golden/stratified_128_200steps.npz, written by make_stratified128.py).

The oracle's full final state is too large to commit, so the GPU run is
compared on what the fixture holds: every 8th cell along each axis, the
plane at axis-3 index 64, 8 probe cells, and per-variable L2 norms, sums and
max |w| over the whole mesh.  The GPU run takes the oracle's dt sequence
(the state comparison) and also runs free (its own dt sequence is compared
with the oracle's).  Bar: rel-L2 <= 1e-12 per variable on the sampled cells
and on the plane, or 4x the oracle's own FMA/no-FMA spread recorded in the
fixture where that is larger."""
import os

import numpy as np
import pytest

from sac_amd import problems
from sac_amd.errors import mesh

pytestmark = pytest.mark.gpu

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "stratified_128_200steps.npz")
NE = 8


def _summary(prob, w, probe_idx):
    m = mesh(prob, w[:NE])
    norms = np.stack([np.sqrt((m ** 2).reshape(NE, -1).sum(1)), m.reshape(NE, -1).sum(1),
                      np.abs(m).reshape(NE, -1).max(1)], axis=1)
    probes = np.stack([m[:, k3, k2, k1] for k1, k2, k3 in probe_idx], axis=1)
    n = prob.nx[0]
    return {"norms": norms, "sub": m[:, ::8, ::8, ::8], "slice_k": m[:, n // 2, :, :],
            "probes": probes}


def _rel(a, b):
    return [float(np.linalg.norm(a[i] - b[i]) / max(np.linalg.norm(b[i]), 1e-300))
            for i in range(NE)]


@pytest.fixture(scope="module")
def fixture():
    if not os.path.exists(PATH):
        pytest.skip("the 128^3 fixture has not been generated")
    return np.load(PATH)


def test_stratified128_under_oracle_dt(fixture):
    from sac_amd import device
    prob = problems.stratified(n=128)
    dev = device.Device(prob)
    dev.upload(prob.w)
    dev.getboundary()
    for dt in fixture["dts"]:
        dev.advance(float(dt))
    w = dev.download()
    dev.close()
    s = _summary(prob, w, fixture["probe_idx"])
    bars = [max(1e-12, 4.0 * f) for f in fixture["floor"]]
    for key in ("sub", "slice_k", "probes"):
        errs = _rel(s[key], fixture[key])
        print(f"{key}: rel-L2 " + " ".join(f"{e:.1e}" for e in errs))
        assert all(e <= b for e, b in zip(errs, bars)), (key, errs, bars)
    rn = np.abs(s["norms"][:, 0] - fixture["norms"][:, 0]) / fixture["norms"][:, 0]
    print("L2 norms: rel. difference " + " ".join(f"{e:.1e}" for e in rn))
    assert all(e <= b for e, b in zip(rn, bars)), (rn, bars)
    # the max |w| is attained at one cell: equal to rounding
    np.testing.assert_allclose(s["norms"][:, 2], fixture["norms"][:, 2], rtol=1e-10)


def test_stratified128_free_running_dt(fixture):
    from sac_amd import device
    prob = problems.stratified(n=128)
    dev = device.Device(prob)
    dev.upload(prob.w)
    dev.getboundary()
    dts = [dev.step(prob.courantpar) for _ in range(len(fixture["dts"]))]
    w = dev.download()
    dev.close()
    rel = np.abs(np.array(dts) - fixture["dts"]) / fixture["dts"]
    print(f"dt sequence: max rel. difference {rel.max():.1e}")
    assert rel.max() <= 1e-10, rel.max()
    s = _summary(prob, w, fixture["probe_idx"])
    errs = _rel(s["sub"], fixture["sub"])
    print("free-running, sampled cells: rel-L2 " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) <= 1e-9, errs
