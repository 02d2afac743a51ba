"""GPU parity of the batched ensemble step.

Every member of an ensemble runs the same tile bodies as a single instance,
only launched over all members' tiles at once, so each member
must be bit-identical to its own single-instance run with the same kernel
variant -- and, in FP64, to the CPU oracle.
"""
import numpy as np
import pytest

import se2dr_fork_amd as S
from se2dr_fork_amd import scenarios as SC

pytestmark = pytest.mark.gpu


def uneven(p, n=5, seed=100):
    """n members of degree p with different sizes, materials, sources and receivers."""
    shapes = [(6, 5), (11, 3), (4, 9), (13, 7), (3, 3), (9, 12), (7, 2)]
    return [SC.small(mx=shapes[i % len(shapes)][0], my=shapes[i % len(shapes)][1], p=p, seed=seed + i,
                     nsrc=1 + i % 3, nrec=2 + i % 4) for i in range(n)]


def single_runs(scs, precision, nsteps, kernel=None, chunks=(None,)):
    out = []
    for sc in scs:
        b = SC.build(sc, precision=precision)
        if kernel is not None:
            b.solver.ctx.set_kernel(kernel)
        rows = []
        for k in chunks:
            rows.append(b.solver.advance(k or nsteps, b.dt)[1])
        out.append((np.concatenate(rows), *b.solver.ctx.get_state()))
        b.solver.close()
    return out


def ensemble(scs, precision, kernel=None, graph=True):
    builts = [SC.build(sc, precision=precision) for sc in scs]
    ens = S.Ensemble([b.solver for b in builts])
    if kernel is not None:
        ens.set_kernel(kernel)
    ens.set_graph(graph)
    return builts, ens


def ens_runs(builts, ens, nsteps, chunks=(None,)):
    dts = [b.dt for b in builts]
    rows = [[] for _ in builts]
    for k in chunks:
        res = ens.advance(k or nsteps, dts)
        for i, (_, tr) in enumerate(res):
            rows[i].append(tr)
    return [(np.concatenate(rows[i]), *b.solver.ctx.get_state()) for i, b in enumerate(builts)]


def assert_same(got, want):
    for (tr, u, v, a), (tr0, u0, v0, a0) in zip(got, want):
        np.testing.assert_array_equal(tr, tr0)
        np.testing.assert_array_equal(u, u0)
        np.testing.assert_array_equal(v, v0)
        np.testing.assert_array_equal(a, a0)


@pytest.mark.parametrize("p", [3, 6])
def test_fp64_ensemble_bit_identical_to_oracle(oracle, p):
    scs = uneven(p)
    builts, ens = ensemble(scs, S.FP64)
    got = ens_runs(builts, ens, 30, chunks=(17, 13))
    for sc, b, (tr, u, v, a) in zip(scs, builts, got):
        pr = oracle.Problem(sc, dt=b.dt, nt=30)
        _, tr0, u0, v0, a0 = pr.run(max_steps=30)
        np.testing.assert_array_equal(tr, tr0)
        np.testing.assert_array_equal(u, u0)
        np.testing.assert_array_equal(v, v0)
        np.testing.assert_array_equal(a, a0)
    ens.close()


@pytest.mark.parametrize("p", [4, 6, 8])
def test_fp64_ensemble_every_variant_bit_identical_to_oracle(oracle, p):
    """Every FP64 batched kernel (the LDS tile kernels and the register
    kernel schedules built for p) advances each member exactly as the
    oracle does."""
    scs = uneven(p, n=3, seed=600 + p)
    want = []
    for sc in scs:
        b = SC.build(sc, precision=S.FP64, with_solver=False)
        _, tr0, u0, v0, a0 = oracle.Problem(sc, dt=b.dt, nt=20).run(max_steps=20)
        want.append((tr0, u0, v0, a0))
    tried = 0
    for var in [S.KERNEL_LDS, S.KERNEL_LDS_WIDE] + [S.KERNEL_REG64 + k for k in range(9)]:
        builts, ens = ensemble(scs, S.FP64)
        try:
            ens.set_kernel(var)
        except S.Se2wError:
            ens.close()
            # one lane per element: p <= 4; schedules 7, 8 only (and only they) at p = 8
            s = var - S.KERNEL_REG64
            assert (s >= 4 and p > 4) or ((s >= 7) != (p == 8))
            continue
        tried += 1
        assert_same(ens_runs(builts, ens, 20, chunks=(7, 13)), want)
        ens.close()
    assert tried == {4: 9, 6: 6, 8: 4}[p]


def test_fp64_ensemble_with_cubic_spline_and_p0_sources(oracle):
    """Members whose sources span several elements (the cubic-spline form)
    or sit on one cell (P0): each member's table of source-time values is
    widened to its element sources, bit-identical to the oracle."""
    scs = uneven(4, n=3, seed=500)
    impls = ([S.IMPL_SPLINE], [S.IMPL_P0, S.IMPL_SPLINE], [S.IMPL_NEAREST_QPOINT, S.IMPL_SPLINE, S.IMPL_P0])
    for sc, im in zip(scs, impls):
        sc.src_type = [S.SOURCE_MOMENT] * len(sc.src_type)
        sc.src_values = np.tile([1.3, -0.4, 0.7, 2.1], (len(sc.src_type), 1))
        sc.src_impl = im
    builts, ens = ensemble(scs, S.FP64)
    got = ens_runs(builts, ens, 24, chunks=(11, 13))
    for sc, b, (tr, u, v, a) in zip(scs, builts, got):
        pr = oracle.Problem(sc, dt=b.dt, nt=24)
        _, tr0, u0, v0, a0 = pr.run(max_steps=24)
        np.testing.assert_array_equal(tr, tr0)
        np.testing.assert_array_equal(u, u0)
        np.testing.assert_array_equal(v, v0)
        np.testing.assert_array_equal(a, a0)
    ens.close()


@pytest.mark.parametrize("prec", [S.FP32, S.FP32_KAHAN])
@pytest.mark.parametrize("p", [4, 6, 8])
def test_fp32_ensemble_equals_single_runs_for_every_variant(prec, p):
    scs = uneven(p, n=4, seed=200 + p)
    variants = [S.KERNEL_LDS] + [S.KERNEL_REG + k for k in range(8)] + [S.KERNEL_MFMA, S.KERNEL_LINE, S.KERNEL_LDS_WIDE]
    tried = 0
    for var in variants:
        builts, ens = ensemble(scs, prec)
        try:
            ens.set_kernel(var)
        except S.Se2wError:
            ens.close()
            continue
        tried += 1
        got = ens_runs(builts, ens, 24)
        ens.close()
        want = single_runs(scs, prec, 24, kernel=var)
        assert_same(got, want)
    assert tried >= 2


def test_graph_and_direct_launches_agree():
    scs = uneven(4, n=3, seed=300)
    b1, e1 = ensemble(scs, S.FP32_KAHAN, graph=True)
    b2, e2 = ensemble(scs, S.FP32_KAHAN, graph=False)
    g1 = ens_runs(b1, e1, 20, chunks=(10, 10, 5))  # the second chunk replays the cached graph
    g2 = ens_runs(b2, e2, 20, chunks=(10, 10, 5))
    assert_same(g1, g2)
    assert_same(g1, single_runs(scs, S.FP32_KAHAN, 20, kernel=e1.kernel(), chunks=(10, 10, 5)))
    e1.close()
    e2.close()


def test_ensemble_autotune_keeps_the_state():
    scs = uneven(6, n=3, seed=400)
    builts, ens = ensemble(scs, S.FP32)
    ens.set_kernel(S.KERNEL_LDS)
    ens_runs(builts, ens, 12)
    best, ms = ens.autotune(steps=5, reps=2)
    assert ens.kernel() == best and ms > 0
    got = ens_runs(builts, ens, 12)
    want = []
    for sc in scs:  # 12 steps with the LDS kernel, then 12 with the tuned one
        b = SC.build(sc, precision=S.FP32)
        b.solver.ctx.set_kernel(S.KERNEL_LDS)
        b.solver.advance(12, b.dt)
        b.solver.ctx.set_kernel(best)
        tr = b.solver.advance(12, b.dt)[1]
        want.append((tr, *b.solver.ctx.get_state()))
    assert_same(got, want)
    tiles, grid = ens.launch()
    assert tiles >= len(scs) and grid == max(8, (tiles + 7) // 8 * 8)  # one workgroup per tile
    ens.close()


def test_ensemble_argument_errors():
    a = SC.build(SC.small(mx=4, my=4, p=3), precision=S.FP32)
    b = SC.build(SC.small(mx=5, my=4, p=4), precision=S.FP32)
    with pytest.raises(S.Se2wError, match="degree and precision must match"):
        S.Ensemble([a.solver, b.solver])
    c = SC.build(SC.small(mx=5, my=4, p=3), precision=S.FP32, schedule=S.SCHED_SPLIT)
    with pytest.raises(S.Se2wError, match="fused schedule"):
        S.Ensemble([a.solver, c.solver])
    with pytest.raises(S.Se2wError, match="appears twice"):
        S.Ensemble([a.solver, a.solver])
    e = S.Ensemble([a.solver])
    with pytest.raises(S.Se2wError):
        e.set_kernel(99)
    e.close()
