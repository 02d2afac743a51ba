"""se2w_assemble_forces_noalloc, the entry point that allocates nothing,
against se2w_assemble_forces (the same kernels): identical results on two
calls in a row, the second on a smaller mesh, in FP64 (and against the
oracle) and FP32; a mesh beyond its scratch limit is refused with an error.
Collected last (the zz prefix): it must pass alone and after the others."""
import numpy as np
import pytest

import se2dr_fork_amd as S
from se2dr_fork_amd import scenarios as SC

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [S.FP64, S.FP32, S.FP32_KAHAN])
def test_noalloc_forces_equal_the_allocating_entry_point(oracle, prec):
    results = []
    for sc in (SC.small(mx=41, my=37, p=4, seed=3), SC.small(mx=9, my=7, p=6, seed=4)):  # second call smaller
        b = SC.build(sc, precision=prec)
        ctx = b.solver.ctx
        u = np.random.default_rng(sc.mx).standard_normal(2 * sc.nnodes)
        f_ref = ctx.assemble_forces(u)
        f_na = ctx.assemble_forces_noalloc(u)
        np.testing.assert_array_equal(f_na, f_ref)
        if prec == S.FP64:
            np.testing.assert_array_equal(f_na, oracle.Problem(sc, dt=1e-5).assemble(u))
        results.append(f_na)
        b.solver.close()
    assert results[0].size > results[1].size


def test_noalloc_forces_after_a_context_switch_and_reuse():
    """Two contexts share the device's scratch: calls alternate between them
    and repeat with new inputs; each result equals the allocating call's."""
    a = SC.build(SC.small(mx=20, my=11, p=3, seed=5), precision=S.FP64)
    b = SC.build(SC.small(mx=7, my=5, p=5, seed=6), precision=S.FP32)
    rng = np.random.default_rng(7)
    for _ in range(2):
        for built in (a, b):
            ctx = built.solver.ctx
            u = rng.standard_normal(2 * ctx.nnodes)
            np.testing.assert_array_equal(ctx.assemble_forces_noalloc(u), ctx.assemble_forces(u))


def test_noalloc_forces_refuse_meshes_beyond_the_limit():
    ctx = S.Context(512, 512, 4, precision=S.FP64)  # 4.2 M nodes: above the 64 MiB scratch
    D, w1d = S.deriv_matrix(4), S.gll(4)[1]
    ctx.set_tables(D, w1d, np.full(512, 4000.0 / 512), np.full(512, 4000.0 / 512))
    ctx.set_cells(np.full(512 * 512, 2.0e10), np.full(512 * 512, 1.0e10))
    with pytest.raises(S.Se2wError, match="scratch"):
        ctx.assemble_forces_noalloc(np.zeros(2 * ctx.nnodes))
    ctx.close()
