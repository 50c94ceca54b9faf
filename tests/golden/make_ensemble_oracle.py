"""Run the C oracle on every member of the seeded ensembles and record how
each run ends (tests/golden/ensemble_{tubes,patches}_oracle.npz).

Each member runs the reference's main loop (src/vac.t:55-106) on its own:
dt = getdt_courant (src/vac.t:641-699) limited by getdt_visc
(src/vacusr.viscosity.t:509-541), then timetofinish (src/vac.t:585-605):
finished when t >= tmax (0.1) or it >= itmax, stopped when dt < dtmin
(smalldouble, 1e-99, as the ensemble kernels use), failed when dt is not a
positive finite number; otherwise one RK4 step.  Recorded per member: the
steps taken (it), t, and the end state (1 finished, 2 failed, 3 stopped).

The ensembles are the ones the tube and patch ensemble runs use:
batch.riemann_ensemble(4096) (1-D tubes of 200-2000 cells, seed 4096) and
batch.patch_ensemble(1024) (2-D patches of 32-160 cells per axis, seed
2048).  Runs in worker processes.  From the repository root:

    python tests/golden/make_ensemble_oracle.py tubes   [workers]
    python tests/golden/make_ensemble_oracle.py patches [workers]
"""
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

HERE = os.path.dirname(os.path.abspath(__file__))
TMAX = 0.1
DTMIN = 1.0e-99
ITMAX = 1_000_000
FINISHED, FAILED, STOPPED = 1, 2, 3


def _members(which):
    from sac_amd import batch
    return batch.riemann_ensemble(count=4096) if which == "tubes" else \
        batch.patch_ensemble(count=1024)


def run_member(args):
    which, i = args
    return (i,) + main_loop(_members_cache[which][i])


def main_loop(prob):
    """the reference's main loop on one member -> (steps, t, end state)"""
    from oracle import oracle as om
    from tests.helpers import make_oracle
    o = make_oracle(prob, om)
    t, it = 0.0, 0
    while True:
        dt, _ = o.getdt_courant(prob.courantpar)
        dt = o.getdt_special(dt)
        if it >= ITMAX or t >= TMAX:
            state = FINISHED
            break
        if dt < DTMIN:
            state = STOPPED
            break
        if not (dt > 0.0) or not np.isfinite(dt):
            state = FAILED
            break
        o.advance(dt)
        it += 1
        t += dt
        o.set_time(it, 0, t)
    return it, t, state


_members_cache = {}


def _init(which):
    _members_cache[which] = _members(which)


def main():
    which = sys.argv[1] if len(sys.argv) > 1 else "tubes"
    workers = int(sys.argv[2]) if len(sys.argv) > 2 else 6
    from oracle import oracle as om
    om.build()
    n = len(_members(which))
    it = np.zeros(n, dtype=np.int64)
    t = np.zeros(n)
    st = np.zeros(n, dtype=np.int32)
    t0 = time.time()
    # longest members first so that the pool drains evenly
    sizes = [np.prod(p.nx) for p in _members(which)]
    order = list(np.argsort(sizes)[::-1])
    with mp.Pool(workers, initializer=_init, initargs=(which,)) as pool:
        for k, (i, n_it, tt, s) in enumerate(pool.imap_unordered(
                run_member, [(which, int(j)) for j in order], chunksize=1)):
            it[i], t[i], st[i] = n_it, tt, s
            if k % 256 == 255:
                print(f"  {k + 1}/{n} members ({time.time() - t0:.0f} s)", flush=True)
    path = os.path.join(HERE, f"ensemble_{which}_oracle.npz")
    np.savez_compressed(path, it=it, t=t, state=st, tmax=TMAX, dtmin=DTMIN,
                        config=f"{which} ensemble, oracle main loop per member to t={TMAX}")
    stopped = np.nonzero(st == STOPPED)[0]
    print(f"{which}: {n} members, finished {(st == FINISHED).sum()}, stopped "
          f"{len(stopped)} {list(zip(stopped.tolist(), it[stopped].tolist()))}, "
          f"failed {(st == FAILED).sum()}; {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
