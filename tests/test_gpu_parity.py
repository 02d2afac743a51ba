"""GPU parity: the HIP step through the C ABI against the FP64 CPU oracle.

FP64 mode follows the reference's operation order with FMA contraction off,
so it must be bit-identical to the oracle: receiver rows, u, v and a.  FP32
modes are checked against FP64 within stated tolerances.
"""
import os
import subprocess

import numpy as np
import pytest

import se2dr_fork_amd as S
from se2dr_fork_amd import scenarios as SC
from conftest import GOLDEN, ROOT, read_vts, rel_l2

pytestmark = pytest.mark.gpu


def gpu_run(sc, precision=S.FP64, schedule=S.SCHED_FUSED, nsteps=None, kernel=None, chunks=1):
    b = SC.build(sc, precision=precision, schedule=schedule)
    if kernel is not None:
        b.solver.ctx.set_kernel(kernel)
    n = nsteps or b.nsteps
    times, rows = [], []
    base, rem = divmod(n, chunks)
    for c in range(chunks):
        k = base + (1 if c < rem else 0)
        t, r = b.solver.advance(k, b.dt)
        times.append(t)
        rows.append(r)
    u, v, a = b.solver.ctx.get_state()
    return b, np.concatenate(times), np.concatenate(rows), u, v, a


def oracle_run(oracle, sc, b):
    pr = oracle.Problem(sc, dt=b.dt, nt=b.nsteps)
    return pr.run()


def small_cases():
    yield SC.small(mx=6, my=5, p=3, nt=40)
    yield SC.small(mx=7, my=3, p=4, kind=S.SOURCE_FORCE, nt=35)
    yield SC.demo(mx=8, p=2, nt=50, cfl_factor=0.5)
    for p in [1, 2, 5, 6, 7, 8]:
        sc = SC.small(mx=5, my=4, p=p, nt=25, seed=20 + p)
        sc.use_nearest_internal_gll = p > 1
        yield sc
    yield SC.small(mx=33, my=17, p=4, nt=30, nsrc=3, nrec=9)  # several tiles, partial edge tiles
    for kind, p in ((S.SOURCE_MOMENT, 4), (S.SOURCE_FORCE, 6)):  # SOURCE_IMPL_POINTWISE
        sc = SC.small(mx=11, my=9, p=p, nt=30, nsrc=3, kind=kind, seed=50 + p)
        sc.src_impl = [S.IMPL_POINTWISE] * 3
        sc.name = "pointwise"
        yield sc
    yield from legacy_cases()


def legacy_cases():
    """Cubic-spline (se2wave.c:1925-2057) and P0 (2059-2198) moment sources
    beside a nearest-node one; the wide kernel (-sm_h) spans tile borders."""
    I = S
    for p, mx, my, impls, h in ((3, 13, 11, (I.IMPL_SPLINE, I.IMPL_P0, I.IMPL_NEAREST_QPOINT), None),
                                (6, 9, 8, (I.IMPL_SPLINE, I.IMPL_SPLINE, I.IMPL_P0), None),
                                (4, 37, 19, (I.IMPL_P0, I.IMPL_SPLINE, I.IMPL_SPLINE), [0.0, 0.0, 390.0]),
                                (2, 8, 8, (I.IMPL_SPLINE, I.IMPL_SPLINE), [700.0, 0.0])):
        sc = SC.small(mx=mx, my=my, p=p, nt=30, nsrc=len(impls), nrec=6, seed=80 + p)
        sc.src_impl = list(impls)
        sc.src_h = h
        sc.name = "legacy"
        yield sc


@pytest.mark.parametrize("sc", list(small_cases()), ids=lambda s: f"{s.name}-{s.mx}x{s.my}-p{s.p}")
def test_fp64_bit_identical_to_oracle(oracle, sc):
    b, times, rows, u, v, a = gpu_run(sc)
    to, ro, uo, vo, ao = oracle_run(oracle, sc, b)
    np.testing.assert_array_equal(times, to)
    np.testing.assert_array_equal(rows, ro)
    np.testing.assert_array_equal(u, uo)
    np.testing.assert_array_equal(v, vo)
    np.testing.assert_array_equal(a, ao)


FP64_VARIANTS = [S.KERNEL_LDS, S.KERNEL_LDS_WIDE] + [S.KERNEL_REG64 + k for k in range(9)]


def fp64_variant_built(k, p):
    """Register schedules: 0..6 up to p = 7 (one lane per element, 4..6, up to
    p = 4), 7 and 8 (one wave per SIMD) at p = 8; the LDS kernels always."""
    if k < S.KERNEL_REG64:
        return True
    s = k - S.KERNEL_REG64
    return (s >= 7) == (p == 8) and (s < 4 or s >= 7 or p <= 4)


@pytest.mark.parametrize("p, mx, my", [(2, 9, 7), (3, 33, 5), (4, 41, 9), (4, 3, 2), (4, 70, 13), (4, 40, 1),
                                        (4, 1, 17), (5, 23, 11), (6, 33, 9), (6, 2, 17), (7, 19, 7), (8, 35, 9),
                                        (8, 2, 5)])
def test_fp64_kernel_variants_bit_identical_to_oracle(oracle, p, mx, my):
    """Every FP64 step kernel (LDS tile kernels, register kernel schedules;
    above p = 4 the lane-pair schedules only, at p = 8 the one-wave ones) runs
    the reference's operation order: receiver rows and the whole state
    bit-identical to the oracle, on meshes of one tile, several tiles with
    partial edge tiles, one element row and one element column."""
    sc = SC.small(mx=mx, my=my, p=p, nt=30, seed=100 + p + mx, nsrc=2, nrec=5)
    b0 = SC.build(sc)
    pr = oracle.Problem(sc, dt=b0.dt, nt=b0.nsteps)
    to, ro, uo, vo, ao = pr.run()
    b0.solver.close()
    ran = 0
    for k in FP64_VARIANTS:
        try:
            b, times, rows, u, v, a = gpu_run(sc, kernel=k, chunks=2)
        except S.Se2wError:
            assert not fp64_variant_built(k, p)
            continue
        assert fp64_variant_built(k, p)
        ran += 1
        np.testing.assert_array_equal(rows, ro, err_msg=f"variant {k}")
        np.testing.assert_array_equal(u, uo, err_msg=f"variant {k}")
        np.testing.assert_array_equal(v, vo, err_msg=f"variant {k}")
        np.testing.assert_array_equal(a, ao, err_msg=f"variant {k}")
        b.solver.close()
    assert ran == sum(fp64_variant_built(k, p) for k in FP64_VARIANTS)


def legacy_p4():
    """The scenario of tests/golden/small_legacy_p4.npz: cubic-spline, P0 and
    cubic-spline (h = 600) moment sources on 9x8 elements of degree 4."""
    sc = SC.small(mx=9, my=8, p=4, nt=80, nsrc=3, nrec=6, seed=17)
    sc.src_impl = [S.IMPL_SPLINE, S.IMPL_P0, S.IMPL_SPLINE]
    sc.src_h = [0.0, 0.0, 600.0]
    return sc


@pytest.mark.parametrize("name", ["small_p3", "small_p4_force", "small_p8", "small_legacy_p4"])
def test_fp64_golden_fixtures(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    sc = {"small_p3": lambda: SC.small(mx=6, my=5, p=3, nt=120),
          "small_p4_force": lambda: SC.small(mx=7, my=3, p=4, kind=S.SOURCE_FORCE, nt=100),
          "small_p8": lambda: SC.small(mx=4, my=3, p=8, seed=5, nt=80),
          "small_legacy_p4": legacy_p4}[name]()
    b, times, rows, u, v, a = gpu_run(sc)
    assert b.dt == float(z["dt"])
    np.testing.assert_array_equal(rows, z["trace"])
    np.testing.assert_array_equal(u, z["u"])
    np.testing.assert_array_equal(v, z["v"])
    np.testing.assert_array_equal(a, z["a"])


def test_legacy_sources_fp32_error():
    """FP32 modes with cubic-spline and P0 sources against the FP64 run."""
    sc = list(legacy_cases())[2]
    sc.nt = 200
    _, _, rows, _, _, _ = gpu_run(sc)
    for prec, tol_uv in [(S.FP32_KAHAN, 1e-5), (S.FP32, 3e-5)]:
        _, _, r32, _, _, _ = gpu_run(sc, precision=prec)
        assert rel_l2(r32[..., 0:2], rows[..., 0:2]) < tol_uv  # displacement
        assert rel_l2(r32[..., 2:4], rows[..., 2:4]) < tol_uv  # velocity


def test_config1_fp64_matches_golden_and_fp32_error():
    """64x64 elements, p=6, point force, 1000 steps at 0.2 CFL."""
    z = np.load(os.path.join(GOLDEN, "config1.npz"))
    b, times, rows, _, _, _ = gpu_run(SC.config1())
    assert b.nsteps == 1000
    np.testing.assert_array_equal(times, z["times"])
    np.testing.assert_array_equal(rows, z["trace"])
    for prec, tol_uv in [(S.FP32_KAHAN, 1e-5), (S.FP32, 3e-5)]:
        _, _, r32, _, _, _ = gpu_run(SC.config1(), precision=prec)
        assert rel_l2(r32[..., 0:2], rows[..., 0:2]) < tol_uv  # displacement
        assert rel_l2(r32[..., 2:4], rows[..., 2:4]) < tol_uv  # velocity


def test_readme_demo_figure():
    """The README example (-mx 64 -my 64 -bdegree 6 -tmax 0.4, demo dt) against
    the oracle fixture, and the features of the reference's vy figure
    (docs/figs/se2wave-demo-vy.png): onset ~0.1 s, +5.2e-9 near 0.16 s,
    -6e-9 near 0.187 s."""
    z = np.load(os.path.join(GOLDEN, "demo_readme.npz"))
    b, times, rows, _, _, _ = gpu_run(SC.demo(mx=64, p=6, nt=100000, tmax=0.4, cfl_factor=0.5), chunks=7)
    np.testing.assert_array_equal(rows, z["trace"])
    vy = rows[:, 0, 3]
    i, j = np.argmax(vy), np.argmin(vy)
    assert 5.0e-9 < vy[i] < 5.5e-9 and 0.155 < times[i] < 0.166
    assert -6.3e-9 < vy[j] < -5.7e-9 and 0.182 < times[j] < 0.192


@pytest.mark.parametrize("p", [4, 6, 7])
def test_split_schedule_equals_fused(oracle, p):
    """The split schedule (element kernel storing f, then the update kernel;
    in FP64 the register kernel's element-only form for 2 <= p <= 7) gives
    the fused step's bits, and the oracle's."""
    sc = SC.small(mx=9 + p, my=7, p=p, nt=30, seed=70 + p)
    b, _, r1, u1, v1, a1 = gpu_run(sc, schedule=S.SCHED_FUSED)
    _, _, r2, u2, v2, a2 = gpu_run(sc, schedule=S.SCHED_SPLIT)
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(u1, u2)
    np.testing.assert_array_equal(a1, a2)
    _, ro, uo, _, _ = oracle_run(oracle, sc, b)
    np.testing.assert_array_equal(r2, ro)
    np.testing.assert_array_equal(u2, uo)
    for prec in [S.FP32, S.FP32_KAHAN]:
        _, _, r1, u1, _, _ = gpu_run(sc, precision=prec, schedule=S.SCHED_FUSED)
        _, _, r2, u2, _, _ = gpu_run(sc, precision=prec, schedule=S.SCHED_SPLIT)
        assert rel_l2(u2, u1) < 1e-5


@pytest.mark.parametrize("prec, schedule", [(S.FP64, S.SCHED_FUSED), (S.FP32, S.SCHED_FUSED),
                                             (S.FP32_KAHAN, S.SCHED_FUSED), (S.FP64, S.SCHED_SPLIT),
                                             (S.FP32_KAHAN, S.SCHED_SPLIT)])
@pytest.mark.parametrize("mesh", ["one tile", "many tiles"])
def test_chunked_advance_equals_single_advance(prec, schedule, mesh):
    """A run's last step writes the plain state (u, v, a) and, for a next run
    with the same dt, that run's predicted state with the operations of the
    predictor kernel: chunking changes no bits in any mode or schedule."""
    sc = SC.small(mx=6, my=5, p=3, nt=60) if mesh == "one tile" else SC.small(mx=41, my=37, p=4, nt=45, seed=13)
    _, t1, r1, u1, v1, a1 = gpu_run(sc, precision=prec, schedule=schedule)
    _, t2, r2, u2, v2, a2 = gpu_run(sc, precision=prec, schedule=schedule, chunks=7)
    np.testing.assert_array_equal(t1, t2)
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(u1, u2)
    np.testing.assert_array_equal(v1, v2)
    np.testing.assert_array_equal(a1, a2)


@pytest.mark.parametrize("prec", [S.FP64, S.FP32, S.FP32_KAHAN])
def test_run_after_a_dt_change_or_a_state_write_predicts_from_the_plain_state(prec):
    """The predicted state a run leaves behind serves only a next run with
    the same dt on an unwritten state.  A run with another dt, or after
    set_state, must predict from the plain state: it equals a fresh context
    given that plain state (bit for bit; after the dt change FP32+Kahan's
    carried compensation, which a fresh context starts at zero, allows
    round-off, far below the O(1) error of a stale prediction)."""
    sc = SC.small(mx=20, my=11, p=3, nt=40, seed=21)
    rng = np.random.default_rng(5)

    def fresh(state, dt, psi, n):
        b = SC.build(sc, precision=prec)
        b.solver.ctx.set_state(*state)
        b.solver.ctx.run(n, dt, psi)
        return (*b.solver.ctx.get_state(), b.solver.ctx.receivers())

    b = SC.build(sc, precision=prec)
    ctx = b.solver.ctx
    ns = ctx.counts()[0]
    assert ns > 0
    psi1, psi2 = rng.standard_normal(15 * ns), rng.standard_normal(10 * ns)
    ctx.run(15, b.dt, psi1)
    plain = ctx.get_state()
    ctx.run(10, 0.5 * b.dt, psi2)  # another dt
    got = (*ctx.get_state(), ctx.receivers())
    want = fresh(plain, 0.5 * b.dt, psi2, 10)
    for g, w in zip(got, want):
        if prec == S.FP32_KAHAN:
            assert rel_l2(g, w) < 1e-6
        else:
            np.testing.assert_array_equal(g, w)
    new = tuple(x + 1e-4 * rng.standard_normal(x.size) for x in got[:3])
    ctx.set_state(*new)  # a state written between two runs with the same dt
    ctx.run(10, 0.5 * b.dt, psi2)
    got = (*ctx.get_state(), ctx.receivers())
    want = fresh(new, 0.5 * b.dt, psi2, 10)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)


def test_state_roundtrip_and_restart(oracle):
    sc = SC.small(mx=6, my=5, p=3, nt=30)
    rng = np.random.default_rng(3)
    u0, v0, a0 = (rng.standard_normal(2 * sc.nnodes) * 1e-3 for _ in range(3))
    b = SC.build(sc)
    b.solver.ctx.set_state(u0, v0, a0)
    u, v, a = b.solver.ctx.get_state()
    np.testing.assert_array_equal(u, u0)
    np.testing.assert_array_equal(v, v0)
    np.testing.assert_array_equal(a, a0)
    _, rows = b.solver.advance(30, b.dt)
    pr = oracle.Problem(sc, dt=b.dt, nt=30)
    _, ro, uo, vo, ao = pr.run(u0=u0, v0=v0, a0=a0)
    np.testing.assert_array_equal(rows, ro)
    np.testing.assert_array_equal(b.solver.ctx.get_state()[0], uo)


@pytest.mark.parametrize("prec", [S.FP64, S.FP32_KAHAN])
def test_edge_cases_against_the_oracle(oracle, prec):
    """Free vibration with no sources and no receivers (a nonzero initial
    state, run in single steps and then a chunk); one element with sources
    on the domain's corner and left edge and receivers on two corners
    (element location clamped at the upper bound, 3035-3048, 2736-2760);
    one element row and one element column.  FP64 bit-identical to the
    oracle, FP32+Kahan within 1e-5."""
    def check(got, want, tol):
        for g, w in zip(got, want):
            if prec == S.FP64:
                np.testing.assert_array_equal(g, w)
            elif w.size:
                assert rel_l2(g, w) < tol

    sc = SC.small(mx=5, my=3, p=3, nsrc=0, nrec=0, nt=20)
    sc.src_xy, sc.src_values, sc.rec_xy = np.zeros((0, 2)), np.zeros((0, 4)), np.zeros((0, 2))
    rng = np.random.default_rng(11)
    u0, v0 = rng.standard_normal(2 * sc.nnodes) * 1e-3, rng.standard_normal(2 * sc.nnodes) * 1e-2
    a0 = np.zeros_like(u0)
    b = SC.build(sc, precision=prec)
    b.solver.ctx.set_state(u0, v0, a0)
    rows = [b.solver.advance(1, b.dt)[1] for _ in range(5)] + [b.solver.advance(15, b.dt)[1]]
    assert all(r.size == 0 for r in rows)
    _, _, uo, vo, ao = oracle.Problem(sc, dt=b.dt, nt=20).run(u0=u0, v0=v0, a0=a0)
    check(b.solver.ctx.get_state(), (uo, vo, ao), 1e-5)
    b.solver.close()

    corner = SC.small(mx=1, my=1, p=4, nsrc=2, nrec=2, nt=40)
    corner.src_xy = np.array([[4000.0, 4000.0], [0.0, 2000.0]])
    corner.rec_xy = np.array([[0.0, 0.0], [4000.0, 4000.0]])
    row = SC.small(mx=19, my=1, p=2, nsrc=1, nrec=3, nt=40, seed=12)
    col = SC.small(mx=1, my=23, p=5, nsrc=1, nrec=3, nt=40, seed=13)
    for sc in (corner, row, col):
        b, _, rows, u, v, a = gpu_run(sc, precision=prec, chunks=3)
        _, ro, uo, vo, ao = oracle.Problem(sc, dt=b.dt, nt=b.nsteps).run()
        check((u, v), (uo, vo), 1e-5)
        if prec == S.FP64:
            np.testing.assert_array_equal(rows, ro)
            np.testing.assert_array_equal(a, ao)
        b.solver.close()


def test_assemble_forces(oracle):
    for sc in [SC.small(mx=9, my=7, p=4), SC.small(mx=5, my=6, p=7, seed=9)]:
        u = np.random.default_rng(sc.p).standard_normal(2 * sc.nnodes)
        fo = oracle.Problem(sc, dt=1e-5).assemble(u)
        for prec in [S.FP64, S.FP32, S.FP32_KAHAN]:
            built = SC.build(sc, precision=prec)
            f = built.solver.ctx.assemble_forces(u)
            if prec == S.FP64:
                np.testing.assert_array_equal(f, fo)
            else:
                assert rel_l2(f, fo) < 2e-6


@pytest.mark.parametrize("p, mx, my", [(4, 41, 37), (6, 23, 19), (8, 13, 11)])
def test_assemble_forces_with_every_kernel_variant(oracle, p, mx, my):
    """f = K(u) after set_kernel of each variant the context accepts: the call
    runs an element-only kernel with that kernel's own tiling (fused-only
    variants fall back to the LDS tile kernel), so every node of a
    multi-tile mesh is assembled.  FP64 bit-identical to the oracle; FP32
    within rounding."""
    sc = SC.small(mx=mx, my=my, p=p, seed=90 + p)
    u = np.random.default_rng(p).standard_normal(2 * sc.nnodes)
    fo = oracle.Problem(sc, dt=1e-5).assemble(u)
    hip, _ = S.libraries()
    for prec in (S.FP64, S.FP32, S.FP32_KAHAN):
        built = SC.build(sc, precision=prec)
        ctx = built.solver.ctx
        tried = 0
        for k in range(hip.se2w_kernel_count()):
            try:
                ctx.set_kernel(k)
            except S.Se2wError:
                continue
            tried += 1
            f = ctx.assemble_forces(u)
            if prec == S.FP64:
                np.testing.assert_array_equal(f, fo)
            else:
                assert rel_l2(f, fo) < 2e-6, (prec, k)
        assert tried >= (2 if prec == S.FP64 else 4)
        built.solver.close()


@pytest.mark.parametrize("p, mx, my", [(4, 37, 33), (6, 23, 19), (4, 3, 2), (4, 40, 1), (2, 9, 7), (3, 8, 5),
                                        (5, 11, 6), (7, 12, 9), (8, 17, 10), (3, 70, 20), (6, 35, 34)])
def test_fp32_kernel_variants_agree(p, mx, my):
    """Register-blocked schedules share the arithmetic (one lane or a lane pair
    per element, full or top-edge-only halo rows): identical bits.  The
    LDS-staged kernel sums in another order: agreement to FP32 rounding."""
    hip, _ = S.libraries()
    sc = SC.small(mx=mx, my=my, p=p, nt=60)
    for prec in [S.FP32, S.FP32_KAHAN]:
        _, _, r0, u0, _, _ = gpu_run(sc, precision=prec, kernel=S.KERNEL_LDS)
        ref = None
        for k in range(S.KERNEL_REG, S.KERNEL_MFMA):
            try:
                _, _, r, u, v, a = gpu_run(sc, precision=prec, kernel=k)
            except S.Se2wError:
                continue
            if ref is None:
                ref = (r, u)
            np.testing.assert_array_equal(r, ref[0])
            np.testing.assert_array_equal(u, ref[1])
        assert rel_l2(ref[1], u0) < 2e-6


def test_autotune_keeps_the_state():
    sc = SC.small(mx=40, my=36, p=4, nt=50)
    b1 = SC.build(sc, precision=S.FP32_KAHAN)
    first = b1.solver.ctx.kernel()
    b1.solver.advance(20, b1.dt)
    best, ms = b1.solver.ctx.autotune(steps=5, reps=2)
    assert best >= 0
    _, r1 = b1.solver.advance(30, b1.dt)
    u1 = b1.solver.ctx.get_state()[0]
    # the same sequence without the tuning trials in between
    b2 = SC.build(sc, precision=S.FP32_KAHAN)
    b2.solver.ctx.set_kernel(first)
    b2.solver.advance(20, b2.dt)
    b2.solver.ctx.set_kernel(best)
    _, r2 = b2.solver.advance(30, b2.dt)
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(u1, b2.solver.ctx.get_state()[0])


def test_reserve_makes_advances_allocation_free():
    """After reserve(n), advances of up to n steps do not grow any run
    buffer, and their results equal those of an unreserved solver."""
    sc = SC.small(mx=9, my=7, p=4, nt=90)
    rows = []
    for reserve in (True, False):
        b = SC.build(sc, precision=S.FP32_KAHAN)
        b.solver.set_uv_receivers([S.ReceiverUV(b.spec, [1500.3, 2200.7], interpolate=True)])
        if reserve:
            b.solver.reserve(40)
            caps = b.solver.ctx.reserved()
            ns, nr, ner, _ = b.solver.ctx.counts()
            assert caps[0] >= 40 * ns and caps[1] >= 40 * nr * (6 + 2 * 5) and caps[2] >= 40 * ner * 25 * 4
        out = [b.solver.advance(n, b.dt)[1] for n in (40, 5, 40)]
        if reserve:
            assert b.solver.ctx.reserved() == caps
        rows.append(np.concatenate(out))
        b.solver.close()
    np.testing.assert_array_equal(rows[0], rows[1])


def test_norms():
    sc = SC.small(mx=6, my=5, p=3, nt=40)
    b, _, _, u, v, _ = gpu_run(sc)
    n = b.solver.ctx.norms()
    np.testing.assert_allclose(n, [u.max(), u.min(), np.linalg.norm(u), v.max(), v.min(), np.linalg.norm(v)],
                               rtol=1e-13)


def test_receiver_files_byte_identical(oracle, tmp_path):
    sc = SC.small(mx=6, my=5, p=3, nt=40)
    b, times, rows, _, _, _ = gpu_run(sc)
    (tmp_path / "gpu").mkdir()
    (tmp_path / "orc").mkdir()
    b.receivers.write_header(str(tmp_path / "gpu"))
    b.receivers.append_rows(times[:17], rows[:17])
    b.receivers.append_rows(times[17:], rows[17:])
    to, ro, _, _, _ = oracle_run(oracle, sc, b)
    oracle.Problem(sc, dt=b.dt).write_receiver_files(str(tmp_path / "orc"), to, ro)
    for f in ["closestqpsource-receiverCP-uva-6x5-p3-rank0.dat", "closestqpsource-receiverCP-uva-6x5-p3.mpimeta"]:
        assert (tmp_path / "gpu" / f).read_bytes() == (tmp_path / "orc" / f).read_bytes()


def test_program_demo_run_matches_oracle(oracle, tmp_path):
    """The se2wave executable with the reference's demo options writes the
    same receiver file as the oracle run of the same problem."""
    exe = os.path.join(ROOT, "se2dr_fork_amd", "bin", "se2wave")
    args = ["-mx", "16", "-bdegree", "3", "-tmax", "0.05", "-cfl_factor", "0.5", "-of", "4", "-output_dir", str(tmp_path)]
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "[se2wave] Using time step size:" in r.stdout and "[step" in r.stdout
    sc = SC.demo(mx=16, p=3, nt=100000, tmax=0.05, cfl_factor=0.5)
    b = SC.build(sc, with_solver=False)
    pr = oracle.Problem(sc, dt=b.dt, nt=b.nsteps)
    to, ro, _, _, _ = pr.run()
    (tmp_path / "orc").mkdir()
    pr.write_receiver_files(str(tmp_path / "orc"), to, ro)
    name = "closestqpsource-receiverCP-uva-16x16-p3-rank0.dat"
    assert (tmp_path / name).read_bytes() == (tmp_path / "orc" / name).read_bytes()
    # snapshots every 4 steps and at the end, written while the next steps
    # run: the FP64 state of the oracle at those steps
    steps = list(range(4, len(to), 4)) + [len(to)]
    assert len(steps) >= 3
    for k in steps:
        _, _, uo, vo, _ = oracle.Problem(sc, dt=b.dt, nt=k).run()
        a = read_vts(tmp_path / f"step-{k:04d}.vts")
        for f, ref in (("disp", uo), ("velo", vo)):
            np.testing.assert_array_equal(a[f + "_x"], ref[0::2])
            np.testing.assert_array_equal(a[f + "_y"], ref[1::2])


@pytest.mark.parametrize("source_impl, sm_h", [(2, None), (2, 90.0), (3, None)])
def test_program_legacy_source_forms_match_oracle(oracle, tmp_path, source_impl, sm_h):
    """-source_impl 2 (cubic spline, -sm_h its width) and 3 (P0), numbered as
    ElastoDynamicsSetSourceImplementation numbers them (se2wave.c:2202-2229),
    applied to the demo's sources: the receiver file equals the oracle's and
    the printed kernel lines carry the oracle's width and \\int W."""
    exe = os.path.join(ROOT, "se2dr_fork_amd", "bin", "se2wave")
    args = ["-mx", "12", "-bdegree", "4", "-tmax", "0.03", "-cfl_factor", "0.5", "-of", "100000",
            "-source_impl", str(source_impl), "-output_dir", str(tmp_path)]
    if sm_h is not None:
        args += ["-sm_h", str(sm_h)]
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    sc = SC.demo(mx=12, p=4, nt=100000, tmax=0.03, cfl_factor=0.5)
    impl = S.IMPL_SPLINE if source_impl == 2 else S.IMPL_P0
    sc.src_impl = [impl] * len(sc.src_type)
    sc.src_h = None if sm_h is None else [sm_h] * len(sc.src_type)
    b = SC.build(sc, with_solver=False)
    pr = oracle.Problem(sc, dt=b.dt, nt=b.nsteps)
    if impl == S.IMPL_SPLINE:
        for s in range(len(sc.src_type)):
            _, h, int_w = pr.assembled_source(s)
            assert ("\\int W = %1.14e \n" % int_w) in r.stdout
            if sm_h is not None:
                assert h == sm_h and ("kernel_h = %1.4e <from options>" % h) in r.stdout
            else:
                assert ("kernel_h = %1.4e <with pfac>" % h) in r.stdout
    to, ro, _, _, _ = pr.run()
    (tmp_path / "orc").mkdir()
    pr.write_receiver_files(str(tmp_path / "orc"), to, ro)
    name = "closestqpsource-receiverCP-uva-12x12-p4-rank0.dat"
    assert (tmp_path / name).read_bytes() == (tmp_path / "orc" / name).read_bytes()


@pytest.mark.parametrize("prec", [S.FP64, S.FP32_KAHAN])
def test_snapshot_taken_while_steps_run(prec):
    """se2w_snapshot_begin/wait return the state at the step they were begun,
    with later steps enqueued behind them; a second begin reuses the buffers
    and a begin while one is in flight waits for it."""
    sc = SC.small(mx=9, my=7, p=4, nt=60)
    b = SC.build(sc, precision=prec)
    ctx = b.solver.ctx
    with pytest.raises(S.Se2wError, match="no snapshot"):
        ctx.snapshot_wait()
    b.solver.advance(15, b.dt)
    u, v, _ = ctx.get_state()
    ctx.snapshot_begin()
    b.solver.advance(10, b.dt)
    su, sv = ctx.snapshot_wait()
    np.testing.assert_array_equal(su, u)
    np.testing.assert_array_equal(sv, v)
    u2, v2, _ = ctx.get_state()
    assert not np.array_equal(u2, u)
    ctx.snapshot_begin()
    ctx.snapshot_begin()
    b.solver.advance(5, b.dt)
    su, sv = ctx.snapshot_wait()
    np.testing.assert_array_equal(su, u2)
    np.testing.assert_array_equal(sv, v2)
    b.solver.close()


def test_config2_full_size(oracle):
    """512x512 elements, p=4 (4.2M nodes): FP64 GPU bit-identical to the oracle
    over a truncated run; FP32 modes against FP64 over the full 2000 steps."""
    sc = SC.config2()
    b, _, _, u, v, a = gpu_run(sc, nsteps=12)
    pr = oracle.Problem(sc, dt=b.dt, nt=12)
    _, _, uo, vo, ao = pr.run()
    np.testing.assert_array_equal(u, uo)
    np.testing.assert_array_equal(v, vo)
    np.testing.assert_array_equal(a, ao)
    _, _, r64, _, _, _ = gpu_run(sc)
    _, _, r32k, _, _, _ = gpu_run(sc, precision=S.FP32_KAHAN)
    assert rel_l2(r32k[..., 0:2], r64[..., 0:2]) < 1e-5
    assert rel_l2(r32k[..., 2:4], r64[..., 2:4]) < 1e-5


def _config5_with_signal():
    """The 2048x2048, p=8 layered workload (268M nodes) with its four Ricker
    sources peaking at t = 0 and only the receivers 0.7-3.5 m from each
    source, so that 500 steps (0.7 ms simulated) carry signal to every
    receiver."""
    sc = SC.with_near_source_receivers(SC.config5())
    sc.rec_xy = sc.rec_xy[-len(sc.src_xy):]
    return sc


@pytest.mark.timeout(900)
def test_config5_full_size(oracle):
    """2048x2048 elements, p=8 (268,468,225 nodes, FP64 state 4.3 GB per
    field).  FP64: two steps bit-identical to the oracle (state and receiver
    rows).  FP32 and FP32+Kahan: the whole u, v state and the receiver
    displacement/velocity against FP64 after the workload's 500 steps."""
    sc = _config5_with_signal()
    b, _, r2, u, v, a = gpu_run(sc, nsteps=2)
    print(f"config5: GPU FP64 2 steps, dt = {b.dt:.6e}", flush=True)
    pr = oracle.Problem(sc, dt=b.dt, nt=2)
    _, ro, uo, vo, ao = pr.run()
    print("config5: oracle 2 steps", flush=True)
    assert np.abs(uo).max() > 0.0 and np.abs(ro[..., :4]).max() > 0.0
    np.testing.assert_array_equal(r2, ro)
    for x, y in ((u, uo), (v, vo), (a, ao)):
        assert np.array_equal(x, y)
    del u, v, a, uo, vo, ao, pr
    b.solver.close()
    b64, _, r64, u64, v64, _ = gpu_run(sc, nsteps=500)
    b64.solver.close()
    print("config5: GPU FP64 500 steps", flush=True)
    for prec, tol in ((S.FP32, 2e-5), (S.FP32_KAHAN, 2e-5)):
        b32, _, r32, u32, v32, _ = gpu_run(sc, precision=prec, nsteps=500)
        b32.solver.close()
        eu, ev = rel_l2(u32, u64), rel_l2(v32, v64)
        er = rel_l2(r32[..., :4], r64[..., :4])
        print(f"config5: precision {prec}: state u {eu:.2e} v {ev:.2e}, receivers {er:.2e}", flush=True)
        assert eu < tol and ev < tol and er < tol
        del u32, v32


def test_gpu_error_paths():
    ctx = S.Context(4, 4, 3, precision=S.FP32)
    with pytest.raises(S.Se2wError, match="set tables, cells and mass"):
        ctx.run(5, 1e-4)
    with pytest.raises(S.Se2wError, match="out of range"):
        ctx.set_sources([16], np.zeros(2 * 16))
    with pytest.raises(S.Se2wError, match="gll"):
        ctx.set_receivers([0], [16])
    with pytest.raises(S.Se2wError, match="register-blocked"):
        S.Context(4, 4, 3, precision=S.FP64).set_kernel(S.KERNEL_REG)
    with pytest.raises(S.Se2wError, match="p = 9"):
        S.Context(4, 4, 9)


def test_fp64_default_kernel_per_degree():
    """A new FP64 context runs the FP64 register kernel wherever it is built:
    its first schedule for 2 <= p <= 7, the first one-wave schedule at p = 8;
    the LDS tile kernel at p = 1.  The p = 8 error names the schedules."""
    for p in range(1, 9):
        k = S.Context(6, 5, p, precision=S.FP64).kernel()
        assert k == (S.KERNEL_LDS if p == 1 else S.KERNEL_REG64 + 7 if p == 8 else S.KERNEL_REG64), (p, k)
    with pytest.raises(S.Se2wError, match="schedules 19 and 20 at p = 8"):
        S.Context(6, 5, 8, precision=S.FP64).set_kernel(S.KERNEL_REG64)


@pytest.mark.parametrize("p", [6, 7, 8])
def test_mfma_kernel_matches_the_vector_kernels(p):
    """Matrix-core contractions (FP32 in, FP32 accumulate) sum in another
    order than the vector kernels: agreement to FP32 rounding, and the same
    distance from the FP64 result."""
    sc = SC.small(mx=23, my=17, p=p, nt=50, nsrc=2, nrec=5, seed=60 + p)  # partial edge tiles
    _, _, r64, u64, _, _ = gpu_run(sc, precision=S.FP64)
    for prec in [S.FP32, S.FP32_KAHAN]:
        _, _, r0, u0, v0, _ = gpu_run(sc, precision=prec, kernel=S.KERNEL_LDS)
        for kern in (S.KERNEL_MFMA, S.KERNEL_LINE):
            _, _, r1, u1, v1, _ = gpu_run(sc, precision=prec, kernel=kern)
            assert rel_l2(u1, u0) < 2e-6 and rel_l2(v1, v0) < 2e-6
            assert rel_l2(r1[..., :4], r64[..., :4]) < 5e-6
            assert rel_l2(u1, u64) < 3 * max(rel_l2(u0, u64), 1e-7)


def test_mfma_kernel_is_rejected_where_not_built():
    b = SC.build(SC.small(mx=5, my=4, p=4), precision=S.FP32)
    with pytest.raises(S.Se2wError, match="matrix-core"):
        b.solver.ctx.set_kernel(S.KERNEL_MFMA)
    b64 = SC.build(SC.small(mx=5, my=4, p=6), precision=S.FP64)
    with pytest.raises(S.Se2wError, match="matrix-core"):
        b64.solver.ctx.set_kernel(S.KERNEL_MFMA)


@pytest.mark.parametrize("p", [3, 5])
def test_single_station_receivers_match_the_oracle(oracle, p):
    """RecordUV / RecordUV_interp on the GPU against the oracle's restatement
    (se2wave.c:2351-2581): FP64 rows bit-identical at every step, over runs
    advanced in chunks; FP32+Kahan rows within its displacement/velocity
    tolerance of the FP64 rows."""
    sc = SC.small(mx=9, my=7, p=p, nt=24, seed=44 + p)
    xy = [[1500.3, 2200.7], [1500.3, 2200.7], [3110.0, 640.2], [0.0, 3999.0], [2222.2, 1777.7]]
    interp = [0, 1, 1, 0, 1]
    rows = {}
    for prec in (S.FP64, S.FP32_KAHAN):
        b = SC.build(sc, precision=prec)
        b.solver.set_uv_receivers([S.ReceiverUV(b.spec, xr, interpolate=bool(i)) for xr, i in zip(xy, interp)])
        parts = []
        for n in (1, 7, 16):
            b.solver.advance(n, b.dt)
            parts.append(b.solver.uv_trace(n))
        rows[prec] = np.concatenate(parts)
        b.solver.close()
    pr = oracle.Problem(sc, dt=b.dt, nt=24, uv_xy=xy, uv_interp=interp)
    _, _, _, _, _, uvo = pr.run(want_uv=True)
    assert rows[S.FP64].shape == uvo.shape == (24, 5, 4)
    np.testing.assert_array_equal(rows[S.FP64], uvo)
    assert rel_l2(rows[S.FP32_KAHAN], uvo) < 1e-5


@pytest.mark.parametrize("p", [3, 8])
def test_wide_tile_kernel_equals_the_tile_kernel(p):
    """512 threads per tile instead of 256: the same per-point and per-node
    arithmetic, so identical bits in every mode (FP64 = the oracle's)."""
    sc = SC.small(mx=13, my=11, p=p, nt=30, seed=70 + p)
    for prec in [S.FP64, S.FP32, S.FP32_KAHAN]:
        _, _, r0, u0, v0, a0 = gpu_run(sc, precision=prec, kernel=S.KERNEL_LDS)
        for kern in (S.KERNEL_LDS_WIDE,):
            _, _, r1, u1, v1, a1 = gpu_run(sc, precision=prec, kernel=kern)
            np.testing.assert_array_equal(r1, r0)
            np.testing.assert_array_equal(u1, u0)
            np.testing.assert_array_equal(v1, v0)
            np.testing.assert_array_equal(a1, a0)
