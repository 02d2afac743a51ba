#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ with the FP64 CPU oracle.

The reference repository holds no test data, so these fixtures are the
oracle's own outputs (oracle/ref_cpu.c is pinned by tests/test_oracle_pins.py)
for runs too long to repeat in every test session:
  config1.npz  -- 64x64 elements, p=6, point force, Ricker 10 Hz, 0.2 CFL,
                  1000 steps: times and the receiver rows (ux uy vx vy ax ay curl)
  demo_readme.npz -- the README example (-mx 64 -my 64 -bdegree 6 -tmax 0.4,
                  demo dt factor 0.5): times and rows of the three stations
  small_*.npz  -- short heterogeneous runs: rows and final u, v, a
                  (small_legacy_p4: cubic-spline, P0 and cubic-spline with a
                  set width moment sources, se2wave.c:1925-2198)
Run:  python tests/golden/make_golden.py [name ...]   (all: about a minute on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import se2dr_fork_amd as S  # noqa: E402  (host setup only: dt and scenario data)
from se2dr_fork_amd import scenarios as SC  # noqa: E402
from oracle import oracle as O  # noqa: E402


def save(name, **arrays):
    np.savez_compressed(os.path.join(HERE, name), **arrays)
    print("wrote", name, {k: v.shape for k, v in arrays.items()})


def run(sc, nt=None, state=False):
    b = SC.build(sc, with_solver=False)
    pr = O.Problem(sc, dt=b.dt, nt=nt or b.nsteps, nthreads=os.cpu_count())
    t0 = time.time()
    times, trace, u, v, a = pr.run(want_state=state)
    print(f"  {sc.name}: {len(times)} steps in {time.time() - t0:.1f} s")
    return b.dt, times, trace, u, v, a


def legacy_p4():
    """9x8 elements, p=4: cubic-spline, P0 and cubic-spline (h = 600) moment sources."""
    sc = SC.small(mx=9, my=8, p=4, nt=80, nsrc=3, nrec=6, seed=17)
    sc.src_impl = [S.IMPL_SPLINE, S.IMPL_P0, S.IMPL_SPLINE]
    sc.src_h = [0.0, 0.0, 600.0]
    sc.name = "legacy"
    return sc


def main(only=()):
    want = lambda n: not only or n in only  # noqa: E731
    if want("config1"):
        sc = SC.config1()
        dt, times, trace, _, _, _ = run(sc)
        save("config1.npz", dt=np.array(dt), times=times, trace=trace)
    if want("demo_readme"):
        sc = SC.demo(mx=64, p=6, nt=100000, tmax=0.4, cfl_factor=0.5)
        dt, times, trace, _, _, _ = run(sc)
        save("demo_readme.npz", dt=np.array(dt), times=times, trace=trace)
    for name, sc in [("small_p3", SC.small(mx=6, my=5, p=3, nt=120)),
                     ("small_p4_force", SC.small(mx=7, my=3, p=4, kind=S.SOURCE_FORCE, nt=100)),
                     ("small_p8", SC.small(mx=4, my=3, p=8, seed=5, nt=80)),
                     ("small_legacy_p4", legacy_p4())]:
        if want(name):
            dt, times, trace, u, v, a = run(sc, state=True)
            save(f"{name}.npz", dt=np.array(dt), times=times, trace=trace, u=u, v=v, a=a)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
