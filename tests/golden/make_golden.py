"""Write the golden fixtures of tests/golden from the C oracle.

The reference itself cannot be built in this image (its sources need the
reference's own preprocessor/configure scripts and an absent user module),
so the fixtures hold the oracle's final states for the configurations the
GPU tests replay; they pin the GPU path on the box without running the
oracle there.  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as om  # noqa: E402
from sac_amd import problems  # noqa: E402
from tests.helpers import make_oracle, mesh  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def run(prob, steps=None, tmax=None):
    o = make_oracle(prob, om)
    dts = []
    t = 0.0
    while (steps is None or len(dts) < steps) and (tmax is None or t < tmax):
        dt = o.step(prob.courantpar)
        dts.append(dt)
        t += dt
    return o.get_w(), np.array(dts), t


def floor(prob, dts, w):
    """per-variable rel-L2 between this oracle build and the same oracle
    built with FMA contraction, both advanced with the same dt sequence:
    the reproducibility floor of the reference algorithm for this run"""
    from tests.test_dt_conditioning import _fma_build, _run_problem
    wf = _run_problem(_fma_build(), prob, list(dts))
    ne = problems.var_index(prob.ndir)["nevolve"]
    return np.array([np.linalg.norm(mesh(prob, wf)[i] - mesh(prob, w)[i]) /
                     max(np.linalg.norm(mesh(prob, w)[i]), 1e-300) for i in range(ne)])


def save(name, prob, w, dts, t, **meta):
    ne = problems.var_index(prob.ndir)["nevolve"]
    fl = floor(prob, dts, w)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), w=mesh(prob, w[:ne]), dts=dts,
                        t=t, floor=fl, **meta)
    print("  reproducibility floor (FMA vs no-FMA oracle, same dt):",
          " ".join(f"{x:.1e}" for x in fl))
    print(name, "steps", len(dts), "t", t, "bytes", os.path.getsize(os.path.join(HERE, name + ".npz")))


def main():
    om.build()
    p = problems.brio_wu(n=800)
    w, dts, t = run(p, tmax=p.tmax)
    save("briowu_800_t0.1", p, w, dts, t, config="brio_wu n=800 gamma=2 courant=0.3 tmax=0.1")
    p = problems.stratified(n=16, nxyz=(16, 14, 12))
    w, dts, t = run(p, steps=40)
    save("stratified_16x14x12_40steps", p, w, dts, t,
         config="stratified nxyz=(16,14,12) seed=20261202 40 RK4 steps")
    p = problems.riemann_tube(300, {"rho": 1.0, "p": 1.0, "v": [0.0, 0.2, -0.1], "b": [0.7, 1.0, 0.3]},
                              {"rho": 0.2, "p": 0.15, "v": [0.1, -0.3, 0.0], "b": [0.7, -0.8, 0.5]})
    w, dts, t = run(p, tmax=0.1)
    save("riemann3_300_t0.1", p, w, dts, t, config="riemann_tube n=300 ndir=3 tmax=0.1")


if __name__ == "__main__":
    main()
