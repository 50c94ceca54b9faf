"""How the oracle's own rounding variants end the ensemble members whose
fate is in question (tests/golden/ensemble_{which}_marginal.npz).

A member that comes close to blowing up (its viscous dt falling towards
dtmin) can finish at t=0.1 in one FP64 run and stop at dt < dtmin in
another that differs from it by rounding alone.  For each member given,
the reference's main loop (make_ensemble_oracle.main_loop, the same loop as
the ensemble fixture) runs under three variants of the oracle:

  fma    the oracle built with FMA contraction (tests/test_dt_conditioning.py)
  ulp1   the plain build, the member's initial state moved by one ulp
  ulp2   (every evolving mesh value -1, 0 or +1 ulp; seeds 1 and 2)

Recorded per member and variant: the end state (1 finished, 2 failed,
3 stopped) and the steps taken.  A member whose variants end differently
from each other or from the plain run (ensemble_{which}_oracle.npz) is
rounding-marginal: no FP64 implementation can be held to either ending.

From the repository root (one process per member and variant):

    python tests/golden/make_ensemble_marginal.py patches 631
"""
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
VARIANTS = ("fma", "ulp1", "ulp2")


def _run(args):
    which, i, variant = args
    if variant == "fma":
        from tests.test_dt_conditioning import _fma_build
        os.environ["SAC_ORACLE_LIB"] = _fma_build()
    sys.path.insert(0, HERE)
    from make_ensemble_oracle import _members, main_loop
    from tests.helpers import ulp_perturbed
    prob = _members(which)[i]
    if variant.startswith("ulp"):
        prob.w = ulp_perturbed(prob, int(variant[3:]))
    it, _, state = main_loop(prob)
    return i, variant, it, state


def main():
    which = sys.argv[1]
    members = [int(x) for x in sys.argv[2:]]
    from tests.test_dt_conditioning import _fma_build
    _fma_build()  # once, before the workers
    jobs = [(which, i, v) for i in members for v in VARIANTS]
    t0 = time.time()
    # a fresh process per run: the oracle library is chosen at import
    ctx = mp.get_context("spawn")
    with ctx.Pool(len(jobs), maxtasksperchild=1) as pool:
        res = pool.map(_run, jobs, chunksize=1)
    state = np.zeros((len(members), len(VARIANTS)), dtype=np.int32)
    it = np.zeros((len(members), len(VARIANTS)), dtype=np.int64)
    for i, v, n, s in res:
        state[members.index(i), VARIANTS.index(v)] = s
        it[members.index(i), VARIANTS.index(v)] = n
    path = os.path.join(HERE, f"ensemble_{which}_marginal.npz")
    np.savez_compressed(path, members=np.array(members, dtype=np.int64),
                        variants=np.array(VARIANTS), state=state, it=it)
    for k, i in enumerate(members):
        print(f"{which} member {i}: " + ", ".join(
            f"{v} state {state[k, j]} after {it[k, j]} steps" for j, v in enumerate(VARIANTS)))
    print(path, f"{time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
