"""The learned surrogate end to end on a reduced mesh (24 x 24 elements,
p = 4): data from the batched solver, training on the GPU kernels, held-out
error against the solver, the gain over going without it (time per answer
against the solver's time per solve, at an error well below the trivial
predictors'), and the solver fallback outside the training distribution."""
import time

import numpy as np
import pytest

import se2dr_fork_amd as S
from se2dr_fork_amd import scenarios as SC
from se2dr_fork_amd import surrogate as SG

pytestmark = pytest.mark.gpu

MX = 24


@pytest.fixture(scope="module")
def data():
    tr = SG.generate(range(1024), mx=MX)
    te = SG.generate(range(100000, 100128), mx=MX)
    return tr, te


def test_mirror_symmetries_are_exact(data):
    _, te = data
    for mirror in (SG.mirror_x, SG.mirror_y):
        xm, ym = mirror(te.x[:3], te.y[:3])
        direct = SG.solve([SG.scenario(*SG.decode(r), mx=MX) for r in xm])
        assert SG.rel_l2(direct, ym) < 2e-3  # FP32 round-off only


def test_surrogate_learns_and_is_faster_than_the_solver(data):
    tr, te = data
    sg = SG.Surrogate(widths=[66, 256, 256, 4096], batch=128, mx=MX)
    losses = sg.fit(tr, epochs=60)
    assert losses[-1] < 0.2 * losses[0]
    sg.predict(te.x)
    t0 = time.perf_counter()
    yp = sg.predict(te.x)
    per_answer = (time.perf_counter() - t0) / len(te.x)
    err = SG.rel_l2(yp, te.y)
    mean_err = SG.rel_l2(np.broadcast_to(tr.y.mean(axis=0), te.y.shape), te.y)
    d = ((te.x[:, None, :2] - tr.x[None, :, :2]) ** 2).sum(-1)
    nn_err = SG.rel_l2(tr.y[np.argmin(d, axis=1)], te.y)
    assert err < 0.5 * mean_err and err < 0.8 * nn_err, (err, mean_err, nn_err)
    # the plain solver, one instance at a time
    sc, _ = SG.instance(100000, mx=MX)
    b = SC.build(sc, precision=S.FP32)
    t0 = time.perf_counter()
    b.solver.advance(SG.STRIDE * SG.NSAMP, b.dt)
    per_solve = time.perf_counter() - t0
    b.solver.close()
    assert per_solve > 1000 * per_answer, (per_solve, per_answer)
    sg.close()


def test_fallback_to_the_solver_outside_the_training_distribution(data):
    tr, te = data
    sg = SG.Surrogate(widths=[66, 128, 4096], batch=64, mx=MX)
    sg.fit(tr, epochs=2)
    x = te.x[:3].copy()
    x[1, 0] = 0.45 * SG.L / SG.POS_UNIT  # source outside the inner 80%
    x[2, 5] = 0.3                        # a 6-sigma perturbation
    y, solved = sg.predict_or_solve(x)
    assert list(solved) == [False, True, True]
    np.testing.assert_array_equal(y[1:], SG.solve([SG.scenario(*SG.decode(r), mx=MX) for r in x[1:]]))
    np.testing.assert_allclose(y[0], sg.predict(x[:1])[0])
    sg.close()
