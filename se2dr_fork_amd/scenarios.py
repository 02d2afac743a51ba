"""Workload definitions: the reference demo and the benchmark configurations.

A Scenario is plain data (mesh, domain, per-cell velocities, sources with
Ricker STFs, receivers, time stepping); `build` turns it into the product's
host objects and GPU solver.  Random draws use the library's splitmix64
stream (se2w_rng_uniform) so that every consumer regenerates the same
arrays from the seed.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import (FP64, IMPL_NEAREST_QPOINT, IMPL_SPLINE, SCHED_FUSED, SOURCE_FORCE, SOURCE_MOMENT, STF, Receivers, Solver, Source,
               SpecFE, _ptr, libraries)


@dataclass
class Scenario:
    name: str
    mx: int
    my: int
    p: int
    scale: Tuple[float, float]
    shift: Tuple[float, float]
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    src_xy: np.ndarray                      # (ns, 2)
    src_type: List[int]                     # SOURCE_FORCE | SOURCE_MOMENT
    src_values: np.ndarray                  # (ns, 4): (Fx, Fy, 0, 0) or (Mxx, Mxy, Myx, Myy)
    stf: List[Tuple[str, float, float, float]]  # ("ricker", t0, freq, amp)
    rec_xy: np.ndarray                      # (nr, 2)
    cfl_factor: float = 0.2
    nt: int = 100
    tmax: float = 1.0e30
    use_nearest_internal_gll: bool = True
    dt: Optional[float] = None              # explicit dt overrides cfl_factor * CFL
    src_impl: Optional[List[int]] = None    # per source IMPL_NEAREST_QPOINT (default), IMPL_POINTWISE, IMPL_SPLINE, IMPL_P0
    src_h: Optional[List[float]] = None     # per source cubic-spline kernel width (-sm_h); <= 0 or None: the default

    @property
    def nnodes(self) -> int:
        return (self.p * self.mx + 1) * (self.p * self.my + 1)

    def lame(self):
        """mu = Vs^2 rho, lambda = Vp^2 rho - 2 mu (se2wave.c:1166-1167)."""
        mu = self.vs * self.vs * self.rho
        lam = self.vp * self.vp * self.rho - 2.0 * mu
        return lam, mu


class Rng:
    """splitmix64 stream of the native library (53-bit uniforms)."""

    def __init__(self, seed: int):
        self.state = C.c_uint64(seed)
        _, self._host = libraries()

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * self._host.se2w_rng_uniform(C.byref(self.state))

    def normal(self) -> float:
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def random_cells(mx, my, seed, vp_lo, vp_hi, rho_lo, rho_hi):
    _, host = libraries()
    ne = mx * my
    vp, vs, rho = np.zeros(ne), np.zeros(ne), np.zeros(ne)
    rc = host.se2w_material_random_cells(mx, my, seed, vp_lo, vp_hi, rho_lo, rho_hi, _ptr(vp), _ptr(vs), _ptr(rho))
    if rc:
        raise RuntimeError("se2w_material_random_cells failed")
    return vp, vs, rho


def layered_cells(mx, my, nlayers, seed, vp_lo, vp_hi, rho_lo, rho_hi):
    _, host = libraries()
    ne = mx * my
    vp, vs, rho = np.zeros(ne), np.zeros(ne), np.zeros(ne)
    rc = host.se2w_material_layered(mx, my, nlayers, seed, vp_lo, vp_hi, rho_lo, rho_hi, _ptr(vp), _ptr(vs),
                                    _ptr(rho))
    if rc:
        raise RuntimeError("se2w_material_layered failed")
    return vp, vs, rho


def demo(mx=8, my=None, p=2, nt=100, tmax=0.4, cfl_factor=0.5) -> Scenario:
    """se2wave_demo (se2wave.c:4541-4720): [-1000,1000]^2, constant material,
    two moment sources with Ricker STFs, three receivers."""
    my = mx if my is None else my
    ne = mx * my
    return Scenario(
        name="demo", mx=mx, my=my, p=p, scale=(2.0e3, 2.0e3), shift=(-1.0e3, -1.0e3),
        vp=np.full(ne, 4746.3670317412243), vs=np.full(ne, 2740.2554625435928), rho=np.full(ne, 1000.0),
        src_xy=np.array([[0.1, 100.1], [400.1, 400.1]]), src_type=[SOURCE_MOMENT, SOURCE_MOMENT],
        src_values=np.array([[1000.0, 0.0, 0.0, 1000.0], [0.0, -1000.0, 1000.0, 0.0]]),
        stf=[("ricker", 0.15, 12.0, 1.0), ("ricker", 0.25, 4.0, 1.0)],
        rec_xy=np.array([[100.0, 200.0], [150.0, 200.0], [180.0, 200.0]]),
        cfl_factor=cfl_factor, nt=nt, tmax=tmax)


def config1(nt=1000, mx=64, p=6) -> Scenario:
    """Single FP64 instance: [0,4000]^2, 64^2 elements, p=6, homogeneous
    Vp 6000 / Vs 3464 / rho 2670, point force in y at the centre, Ricker
    f0=10 Hz t0=0.12 s, dt = 0.2 CFL, receiver 500 m from the source."""
    ne = mx * mx
    return Scenario(
        name="config1", mx=mx, my=mx, p=p, scale=(4000.0, 4000.0), shift=(0.0, 0.0),
        vp=np.full(ne, 6000.0), vs=np.full(ne, 3464.0), rho=np.full(ne, 2670.0),
        src_xy=np.array([[2000.0, 2000.0]]), src_type=[SOURCE_FORCE], src_values=np.array([[0.0, 1.0e10, 0.0, 0.0]]),
        stf=[("ricker", 0.12, 10.0, 1.0)], rec_xy=np.array([[2500.0, 2000.0]]), cfl_factor=0.2, nt=nt)


def _sources_and_line(rng: Rng, ns: int, nr: int, f0: float, L=4000.0, kind=SOURCE_MOMENT):
    """ns sources in the inner 80% of [0,L]^2 and nr receivers on a horizontal
    line 150 m above the first source, spanning +-300 m around it."""
    xy = np.array([[rng.uniform(0.1 * L, 0.9 * L), rng.uniform(0.1 * L, 0.9 * L)] for _ in range(ns)])
    vals = []
    for _ in range(ns):
        if kind == SOURCE_MOMENT:
            vals.append([1.0e9, 0.0, 0.0, 1.0e9])   # explosive
        else:
            vals.append([0.0, 1.0e10, 0.0, 0.0])
    x0, y0 = xy[0]
    xs = np.linspace(x0 - 300.0, x0 + 300.0, nr) if nr > 1 else np.array([x0 + 300.0])
    rec = np.stack([np.clip(xs, 0.0, L), np.full(nr, min(y0 + 150.0, L))], axis=1)
    return xy, [kind] * ns, np.array(vals), [("ricker", 1.2 / f0, f0, 1.0)] * ns, rec


def config2(nt=2000, mx=512, p=4, seed=2) -> Scenario:
    """512^2, p=4, per-cell Vp ~ U[4000,6000], Vs = Vp/sqrt(3),
    rho ~ U[2500,3000] (seed 2, element order), two Ricker sources f0=15 Hz at
    seeded positions, 16 receivers on a line, 2000 steps at 0.2 CFL."""
    vp, vs, rho = random_cells(mx, mx, seed, 4000.0, 6000.0, 2500.0, 3000.0)
    xy, ty, vals, stf, rec = _sources_and_line(Rng(1000 + seed), 2, 16, 15.0)
    return Scenario(name="config2", mx=mx, my=mx, p=p, scale=(4000.0, 4000.0), shift=(0.0, 0.0), vp=vp, vs=vs,
                    rho=rho, src_xy=xy, src_type=ty, src_values=vals, stf=stf, rec_xy=rec, cfl_factor=0.2, nt=nt)


def config3_instance(i: int, nt=1500, p=6) -> Scenario:
    """Instance i of the uneven ensemble: mx, my ~ U{64..192}, p=6, per-cell
    materials as config2, one Ricker f0 ~ U[5,20] Hz, 8 receivers."""
    rng = Rng(3 + i)
    mx = 64 + int(rng.uniform() * 129)
    my = 64 + int(rng.uniform() * 129)
    f0 = rng.uniform(5.0, 20.0)
    vp, vs, rho = random_cells(mx, my, 3 + i, 4000.0, 6000.0, 2500.0, 3000.0)
    xy, ty, vals, stf, rec = _sources_and_line(rng, 1, 8, f0)
    return Scenario(name=f"config3[{i}]", mx=mx, my=my, p=p, scale=(4000.0, 4000.0), shift=(0.0, 0.0), vp=vp, vs=vs,
                    rho=rho, src_xy=xy, src_type=ty, src_values=vals, stf=stf, rec_xy=rec, cfl_factor=0.2, nt=nt)


def config5(nt=500, mx=2048, p=8, seed=5) -> Scenario:
    """2048^2, p=8, 20 horizontal layers Vp ~ U[3000,7000], rho ~
    U[2200,3200], four Ricker sources f0=8 Hz, 64 receivers, 500 steps."""
    vp, vs, rho = layered_cells(mx, mx, 20, seed, 3000.0, 7000.0, 2200.0, 3200.0)
    xy, ty, vals, stf, rec = _sources_and_line(Rng(1000 + seed), 4, 64, 8.0)
    return Scenario(name="config5", mx=mx, my=mx, p=p, scale=(4000.0, 4000.0), shift=(0.0, 0.0), vp=vp, vs=vs,
                    rho=rho, src_xy=xy, src_type=ty, src_values=vals, stf=stf, rec_xy=rec, cfl_factor=0.2, nt=nt)


def with_near_source_receivers(sc: Scenario, offsets=((0.7, 0.0), (0.0, 1.9), (-2.6, 1.1), (2.4, -2.5))) -> Scenario:
    """The scenario with its Ricker sources peaking at t = 0 and one extra
    receiver a few metres from each source (appended after its own
    receivers), so that a short run carries signal to those receivers (config
    5's 500 steps simulate 0.7 ms, in which no wave reaches its receiver
    line)."""
    ns = len(sc.src_xy)
    sc.stf = [("ricker", 0.0, f, a) for (_, _, f, a) in sc.stf]
    off = np.array([offsets[s % len(offsets)] for s in range(ns)])
    sc.rec_xy = np.concatenate([np.asarray(sc.rec_xy).reshape(-1, 2), sc.src_xy + off])
    return sc


def small(mx=6, my=5, p=3, seed=11, nt=40, kind=SOURCE_MOMENT, nsrc=2, nrec=4) -> Scenario:
    """A small heterogeneous case for parity tests (uneven mx, my)."""
    vp, vs, rho = random_cells(mx, my, seed, 4000.0, 6000.0, 2500.0, 3000.0)
    rng = Rng(seed + 7)
    xy = np.array([[rng.uniform(600.0, 3400.0), rng.uniform(600.0, 3400.0)] for _ in range(nsrc)])
    vals = np.array([[1.0e9, 2.0e8, -3.0e8, 0.7e9] if kind == SOURCE_MOMENT else [2.0e9, 1.0e10, 0, 0]
                     for _ in range(nsrc)])
    rec = np.array([[rng.uniform(0.0, 4000.0), rng.uniform(0.0, 4000.0)] for _ in range(nrec)])
    stf = [("ricker", 0.02 + 0.01 * s, 25.0 + 5 * s, 1.0) for s in range(nsrc)]
    return Scenario(name="small", mx=mx, my=my, p=p, scale=(4000.0, 4000.0), shift=(0.0, 0.0), vp=vp, vs=vs, rho=rho,
                    src_xy=xy, src_type=[kind] * nsrc, src_values=vals, stf=stf, rec_xy=rec, cfl_factor=0.2, nt=nt)


@dataclass
class Built:
    spec: SpecFE
    sources: list
    stfs: list
    receivers: Receivers
    dt: float
    nsteps: int
    solver: Optional[Solver] = None


def build(sc: Scenario, precision=FP64, schedule=SCHED_FUSED, with_solver=True) -> Built:
    """Host objects (and the GPU solver) for a scenario, in se2wave_demo's order."""
    spec = SpecFE(sc.mx, sc.my, sc.p)
    spec.scale_coords(sc.scale, sc.shift)
    spec.set_cell_material_velocity(sc.vp, sc.vs, sc.rho)
    sources, stfs = [], []
    for s in range(len(sc.src_type)):
        impl = IMPL_NEAREST_QPOINT if sc.src_impl is None else sc.src_impl[s]
        kh = None if (sc.src_h is None or impl != IMPL_SPLINE) else sc.src_h[s]
        sources.append(Source(spec, sc.src_type[s], sc.src_xy[s], sc.src_values[s], impl=impl,
                              use_nearest_internal_gll=sc.use_nearest_internal_gll, kernel_h=kh))
        kind, a, b, c = sc.stf[s]
        stfs.append(STF(kind, a, b, c))
    rec = Receivers(spec, sc.rec_xy)
    dt = sc.dt if sc.dt is not None else spec.timestep() * sc.cfl_factor
    # the reference stops after the step at which time >= tmax (se2wave.c:4772)
    nsteps, t = sc.nt, 0.0
    for k in range(1, sc.nt + 1):
        t = t + dt
        if t >= sc.tmax:
            nsteps = k
            break
    b = Built(spec=spec, sources=sources, stfs=stfs, receivers=rec, dt=dt, nsteps=nsteps)
    if with_solver:
        b.solver = Solver(spec, sources, stfs, rec, precision=precision, schedule=schedule)
    return b
