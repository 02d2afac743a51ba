"""Seismogram surrogate: an MLP that maps a source position and a coarse grid
of P-velocity perturbations to receiver velocity seismograms.

The solver is both the teacher and the fallback.  Training data come from the
batched ensemble step (Ensemble): each instance is a [0, 4000 m]^2 mesh of
96 x 96 elements of degree 4 whose Vp is 5000 m/s times (1 + d) with d drawn
N(0, 5%) on an 8 x 8 grid of element blocks (Vs = Vp / sqrt(3), rho = 2700),
one vertical point force with a 1 Hz Ricker source-time function at a
position drawn uniformly over the inner 80% of the domain, and four receivers
outside that region.  Every instance uses the same dt (0.5 dt_CFL at the
largest Vp of a 4-sigma perturbation, the reference demo's factor), and the
receivers' vx, vy are sampled every STRIDE steps: 4 receivers x 2 components
x 512 samples = 4096 outputs.

Inputs (66): the source position in units of 100 m from the domain centre
and the 64 fractional perturbations d.  Outputs are divided by one global
scale (the training set's RMS).  Network: 66 -> 1024 -> 1024 -> 1024 -> 4096,
GELU, BF16 MFMA products with FP32 accumulation, Adam (se2w_mlp.h).  The
training set is extended with the problem's two exact mirror symmetries
(mirror_x, mirror_y), which quadruples it at no solver cost.

`Surrogate.predict_or_solve` answers inputs inside the training distribution
with the network and runs the plain solver for the others.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import FP32, SOURCE_FORCE, Ensemble, Mlp
from . import scenarios as SC

L = 4000.0
VP0, RHO, SIGMA = 5000.0, 2700.0, 0.05
GRID = 8
F0 = 1.0  # Ricker peak frequency (Hz); the delay is 1.2 / f0
DT, STRIDE, NSAMP = 5.0e-4, 14, 512
POS_UNIT = 100.0
RECEIVERS = np.array([[200.0, 2000.0], [3800.0, 2000.0], [2000.0, 200.0], [2000.0, 3800.0]])
WIDTHS = [2 + GRID * GRID, 1024, 1024, 1024, len(RECEIVERS) * 2 * NSAMP]


def instance(seed: int, mx=96, p=4, f0=F0):
    """(scenario, input vector) of the training/held-out instance `seed`."""
    rng = SC.Rng(40000 + seed)
    xs, ys = rng.uniform(0.1 * L, 0.9 * L), rng.uniform(0.1 * L, 0.9 * L)
    d = np.array([SIGMA * rng.normal() for _ in range(GRID * GRID)]).reshape(GRID, GRID)  # [block y][block x]
    return scenario(xs, ys, d, mx, p, f0), encode(xs, ys, d)


def encode(xs, ys, d) -> np.ndarray:
    return np.concatenate([[(xs - 0.5 * L) / POS_UNIT, (ys - 0.5 * L) / POS_UNIT], np.ravel(d)])


def decode(x: np.ndarray):
    xs = 0.5 * L + POS_UNIT * x[0]
    ys = 0.5 * L + POS_UNIT * x[1]
    return xs, ys, np.asarray(x[2:]).reshape(GRID, GRID)


def scenario(xs, ys, d, mx=96, p=4, f0=F0) -> SC.Scenario:
    ey, ex = np.divmod(np.arange(mx * mx), mx)
    vp = VP0 * (1.0 + d[ey * GRID // mx, ex * GRID // mx])
    return SC.Scenario(name="surrogate", mx=mx, my=mx, p=p, scale=(L, L), shift=(0.0, 0.0), vp=vp,
                       vs=vp / np.sqrt(3.0), rho=np.full(mx * mx, RHO), src_xy=np.array([[xs, ys]]),
                       src_type=[SOURCE_FORCE], src_values=np.array([[0.0, 1.0e10, 0.0, 0.0]]),
                       stf=[("ricker", 1.2 / f0, f0, 1.0)], rec_xy=RECEIVERS.copy(), nt=STRIDE * NSAMP, dt=DT)


def seismograms(trace: np.ndarray) -> np.ndarray:
    """Receiver rows [nsteps][nr][7] -> outputs [nr][vx, vy][NSAMP] (flattened)."""
    v = trace[STRIDE - 1::STRIDE, :, 2:4][:NSAMP]  # [time][rec][comp]
    return np.ascontiguousarray(v.transpose(1, 2, 0)).ravel()


def solve(scs: Sequence[SC.Scenario], group=1152, log=None) -> np.ndarray:
    """Plain-solver outputs of the scenarios, advanced in ensembles of `group`
    (log(message), if given, is called after each ensemble)."""
    out = np.zeros((len(scs), WIDTHS[-1]))
    for g0 in range(0, len(scs), group):
        builts = [SC.build(sc, precision=FP32) for sc in scs[g0:g0 + group]]
        ens = Ensemble([b.solver for b in builts])
        if len(builts) >= 64:
            ens.autotune(steps=5, reps=2)
        res = ens.advance(STRIDE * NSAMP, [b.dt for b in builts])
        for i, (_, tr) in enumerate(res):
            out[g0 + i] = seismograms(tr)
        ens.close()
        for b in builts:
            b.solver.close()
        if log:
            log(f"solved {min(g0 + group, len(scs))} of {len(scs)} instances")
    return out


@dataclass
class Dataset:
    x: np.ndarray        # [n][66]
    y: np.ndarray        # [n][4096]
    seconds: float       # wall time of the solves


def generate(seeds: Sequence[int], mx=96, p=4, group=1152, log=None, f0=F0) -> Dataset:
    pairs = [instance(s, mx, p, f0) for s in seeds]
    t0 = time.perf_counter()
    y = solve([sc for sc, _ in pairs], group, log)
    return Dataset(x=np.array([x for _, x in pairs]), y=y, seconds=time.perf_counter() - t0)


class Surrogate:
    def __init__(self, widths: Optional[List[int]] = None, batch=256, seed=0, mx=96, p=4, f0=F0):
        self.widths = list(widths or WIDTHS)
        self.mlp = Mlp(self.widths, batch=batch, seed=seed)
        self.scale = 1.0
        self.mx, self.p, self.f0 = mx, p, f0
        self.losses = None

    def fit(self, data: Dataset, epochs=200, lr=1e-3, seed=0, mirror=True, cosine=True) -> np.ndarray:
        """Adam from lr (cosine-decayed to 0 over the run unless cosine=False)."""
        self.scale = float(np.sqrt(np.mean(data.y ** 2))) or 1.0
        if mirror:
            data = augment(data)
        self.mlp.set_lr_schedule(cosine)
        self.mlp.set_data(data.x, data.y / self.scale)
        self.losses = self.mlp.train(epochs, lr=lr, seed=seed)
        return self.losses

    def predict(self, x) -> np.ndarray:
        return self.mlp.forward(np.atleast_2d(x)).astype(np.float64) * self.scale

    @staticmethod
    def in_domain(x) -> np.ndarray:
        """Inputs inside the training distribution: source in the inner 80%
        and every perturbation within 4 standard deviations."""
        x = np.atleast_2d(x)
        inner = 0.4 * L / POS_UNIT
        return (np.abs(x[:, :2]) <= inner).all(axis=1) & (np.abs(x[:, 2:]) <= 4.0 * SIGMA).all(axis=1)

    def predict_or_solve(self, x):
        """Network output for in-domain rows, the plain solver's for the rest;
        returns (outputs, mask of rows answered by the solver)."""
        x = np.atleast_2d(x)
        inside = self.in_domain(x)
        y = np.zeros((x.shape[0], self.widths[-1]))
        if inside.any():
            y[inside] = self.predict(x[inside])
        if (~inside).any():
            scs = [scenario(*decode(r), self.mx, self.p, self.f0) for r in x[~inside]]
            y[~inside] = solve(scs)
        return y, ~inside

    def close(self):
        self.mlp.close()


def mirror_x(x: np.ndarray, y: np.ndarray):
    """The same data mirrored about x = L/2 (exact symmetry of the problem:
    the mesh, the vertical force and the receiver set are symmetric).  The
    source moves to L - xs, the perturbation grid flips in x, receivers 0 and 1
    swap, and vx changes sign."""
    x2 = x.copy()
    x2[:, 0] = -x[:, 0]
    x2[:, 2:] = x[:, 2:].reshape(-1, GRID, GRID)[:, :, ::-1].reshape(-1, GRID * GRID)
    yr = y.reshape(-1, len(RECEIVERS), 2, NSAMP)
    y2 = yr[:, [1, 0, 2, 3]].copy()
    y2[:, :, 0] *= -1.0
    return x2, y2.reshape(y.shape)


def mirror_y(x: np.ndarray, y: np.ndarray):
    """The same data mirrored about y = L/2: the mirrored force points down,
    so by linearity the fields change sign; net effect: the source moves to
    L - ys, the perturbation grid flips in y, receivers 2 and 3 swap, and vx
    changes sign."""
    x2 = x.copy()
    x2[:, 1] = -x[:, 1]
    x2[:, 2:] = x[:, 2:].reshape(-1, GRID, GRID)[:, ::-1, :].reshape(-1, GRID * GRID)
    yr = y.reshape(-1, len(RECEIVERS), 2, NSAMP)
    y2 = yr[:, [0, 1, 3, 2]].copy()
    y2[:, :, 0] *= -1.0
    return x2, y2.reshape(y.shape)


def augment(data: Dataset) -> Dataset:
    """Training set extended by the problem's two mirror symmetries (4x)."""
    xs, ys = [data.x], [data.y]
    x1, y1 = mirror_x(data.x, data.y)
    x2, y2 = mirror_y(data.x, data.y)
    x3, y3 = mirror_y(x1, y1)
    return Dataset(x=np.concatenate([data.x, x1, x2, x3]), y=np.concatenate([data.y, y1, y2, y3]),
                   seconds=data.seconds)


def rel_l2(a, b) -> float:
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def per_sample_rel_l2(a, b) -> np.ndarray:
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)
