"""Write the 128^3 stratified fixture (tests/golden/stratified_128_200steps.npz)
from the C oracle.

Workload: the stratified 3-D perturbation on a static background at 128^3
(problems.stratified(n=128): rho_b = exp(-z/H), isothermal, uniform vertical
B with beta = 1 at the base, axis-1 gravity, Gaussian velocity pulse plus
seeded noise, fixed/periodic/periodic), 200 free-running RK4 steps with
hyperdiffusion.  The full final state (8 evolving fields x 128^3 doubles,
134 MB) is too large to commit, so the fixture holds what pins a GPU run of
the same configuration without the oracle:

  dts        the 200 time steps the oracle chose (getdt_courant and the
             lagged viscous limit)
  norms      per-variable L2 norm, sum and max |w| over the mesh
  sub        every 8th mesh cell along each axis (8 x 16^3 values)
  slice_k    the mesh plane at axis-3 index 64 (8 x 128 x 128 values)
  probes     8 fixed cells (index triples in `probe_idx`) x 8 variables
  floor      per-variable rel-L2 of the same oracle built with FMA
             contraction and advanced with the same dt sequence, on the
             full state: the reproducibility floor of the algorithm here

Takes about 2.5 hours on one core (two 200-step runs at 128^3, ~20 s a step).  Run from
the repository root:  python tests/golden/make_stratified128.py

`--ulp` adds to an existing fixture the spread of the oracle itself under a
one-ulp change of its initial state (every evolving mesh value moved by -1,
0 or +1 ulp), for five seeds, in runs of about 75 minutes each, side by
side:

  ulp_sub, ulp_slice_k, ulp_probes
             per-variable rel-L2 between a perturbed run (the same dt
             sequence) and `sub`, `slice_k`, `probes`, the maximum over the
             seeds (ulp_*_all: each seed; ulp_seeds: the seeds): how far
             apart two FP64 runs of this configuration end when their
             inputs differ by rounding alone.  The response to rounding is
             dominated by a few switches of setnu's maxima and differs
             several-fold between seeds.
  ulp_dts    the dt sequence of the first seed's run when it chooses its own
             time steps
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as om  # noqa: E402
from sac_amd import problems  # noqa: E402
from tests.helpers import make_oracle, mesh, ulp_perturbed  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
NAME = "stratified_128_200steps"
N = 128
STEPS = 200
PROBES = np.array([[64, 64, 64], [10, 64, 64], [118, 64, 64], [64, 3, 120],
                   [32, 96, 17], [100, 5, 5], [0, 0, 0], [127, 127, 127]], dtype=np.int64)


def summarize(prob, w):
    """fixture quantities of a state w (nw, G3, G2, G1) -> dict"""
    ne = problems.var_index(prob.ndir)["nevolve"]
    m = mesh(prob, w[:ne])  # (ne, n3, n2, n1)
    norms = np.stack([np.sqrt((m ** 2).reshape(ne, -1).sum(1)), m.reshape(ne, -1).sum(1),
                      np.abs(m).reshape(ne, -1).max(1)], axis=1)
    # probe index triples are (i1, i2, i3), 0-based mesh indices
    probes = np.stack([m[:, k3, k2, k1] for k1, k2, k3 in PROBES], axis=1)
    return {"norms": norms, "sub": np.ascontiguousarray(m[:, ::8, ::8, ::8]),
            "slice_k": np.ascontiguousarray(m[:, N // 2, :, :]), "probes": probes}


def main():
    om.build()
    prob = problems.stratified(n=N)
    o = make_oracle(prob, om)
    t0 = time.time()
    dts = []
    for i in range(STEPS):
        dts.append(o.step(prob.courantpar))
        if i % 20 == 19:
            print(f"  oracle step {i + 1}/{STEPS} dt={dts[-1]:.6e} ({time.time() - t0:.0f} s)",
                  flush=True)
    w = o.get_w()
    del o
    s = summarize(prob, w)
    from tests.test_dt_conditioning import _fma_build, _run_problem
    print("  FMA build, same dt sequence ...", flush=True)
    wf = _run_problem(_fma_build(), problems.stratified(n=N), dts)
    ne = problems.var_index(prob.ndir)["nevolve"]
    a, b = mesh(prob, wf[:ne]), mesh(prob, w[:ne])
    floor = np.array([np.linalg.norm(a[i] - b[i]) / max(np.linalg.norm(b[i]), 1e-300)
                      for i in range(ne)])
    path = os.path.join(HERE, NAME + ".npz")
    np.savez_compressed(path, dts=np.array(dts), floor=floor, probe_idx=PROBES,
                        config=f"stratified n={N} seed=20261202, {STEPS} free-running RK4 steps",
                        **s)
    print("  reproducibility floor:", " ".join(f"{x:.1e}" for x in floor))
    print(NAME, "bytes", os.path.getsize(path), f"{time.time() - t0:.0f} s")


def _ulp_leg(args):
    """one oracle run from a one-ulp start (seed): its summary under the
    fixture's dt sequence (free=False), or its own dt sequence (free=True)"""
    seed, free, dts = args
    om.build()
    prob = problems.stratified(n=N)
    prob.w = ulp_perturbed(prob, seed)
    o = make_oracle(prob, om)
    out = []
    for i in range(STEPS):
        if free:
            out.append(o.step(prob.courantpar))
        else:
            o.advance(float(dts[i]))
    return (summarize(prob, o.get_w()) if not free else None), np.array(out)


KEYS = ("sub", "slice_k", "probes")


def add_ulp_spread(seeds, free_seed=20261203):
    """Add the one-ulp starts of `seeds` (fixed dt) that the fixture does not
    hold yet, and the free-running dt sequence of free_seed if missing; each
    run in a process of its own.  ulp_<key> is the per-variable maximum over
    the seeds held, ulp_<key>_all the value of each seed."""
    from multiprocessing import Pool
    path = os.path.join(HERE, NAME + ".npz")
    fx = dict(np.load(path))
    have = [int(x) for x in fx.get("ulp_seeds", [])]
    if "ulp_seeds" not in fx and "ulp_sub" in fx:  # one seed, stored before seeds were listed
        have = [free_seed]
        for key in KEYS:
            fx["ulp_" + key + "_all"] = fx["ulp_" + key][None, :]
    new = [sd for sd in seeds if sd not in have]
    jobs = [(sd, False, fx["dts"]) for sd in new]
    if "ulp_dts" not in fx:
        jobs.append((free_seed, True, None))
    t0 = time.time()
    with Pool(max(1, len(jobs))) as pool:
        res = pool.map(_ulp_leg, jobs)

    def rel(a, b):
        return np.array([np.linalg.norm(a[i] - b[i]) / max(np.linalg.norm(b[i]), 1e-300)
                         for i in range(a.shape[0])])
    for (sd, free, _), (s, dts_free) in zip(jobs, res):
        if free:
            fx["ulp_dts"] = dts_free
            continue
        for key in KEYS:
            v = rel(s[key], fx[key])[None, :]
            allk = "ulp_" + key + "_all"
            fx[allk] = np.concatenate([fx[allk], v]) if allk in fx else v
        have.append(sd)
    fx["ulp_seeds"] = np.array(have, dtype=np.int64)
    for key in KEYS:
        fx["ulp_" + key] = fx["ulp_" + key + "_all"].max(axis=0)
        print(f"  one-ulp spread, {key}, max over {len(have)} starts: " +
              " ".join(f"{x:.1e}" for x in fx["ulp_" + key]))
    r = np.abs(fx["ulp_dts"] - fx["dts"]) / fx["dts"]
    print(f"  one-ulp spread of the free-running dt: max {r.max():.1e} at step {r.argmax()}")
    np.savez_compressed(path, **fx)
    print(NAME, "bytes", os.path.getsize(path), f"{time.time() - t0:.0f} s")


if __name__ == "__main__":
    if "--ulp" in sys.argv[1:]:
        add_ulp_spread([20261203, 1, 2, 3, 4])
    else:
        main()
