"""Pins of the oracle's MRI operator layer and solvers (CPU, small sizes)."""

import numpy as np
import pytest

from helpers import gauss_smaps, rand_cplx, rand_traj, rel_l2, truncnorm_traj
from oracle import mri, solvers


def small_op(exact, sense=True, density=None, B=1):
    k = truncnorm_traj(300, 2, 1)
    shape = (12, 10)
    smaps = gauss_smaps(3, shape, 2) if sense else None
    return mri.Operator(k, shape, n_coils=3, n_batchs=B, smaps=smaps, density=density,
                        exact=exact)


@pytest.mark.parametrize("sense", [True, False])
def test_nufft_operator_matches_exact(sense):
    a, b = small_op(True, sense), small_op(False, sense)
    x = rand_cplx((1, 1 if sense else 3, 12, 10), 3)
    y = rand_cplx((1, 3, 300), 4)
    assert rel_l2(b.op(x), a.op(x)) < 5e-6
    assert rel_l2(b.adj_op(y), a.adj_op(y)) < 5e-6
    # adjointness of the normalised operator
    assert abs(np.vdot(a.op(x), y) - np.vdot(x, a.adj_op(y))) < 1e-10 * abs(np.vdot(x, a.adj_op(y)))


def test_dc_is_adj_of_residual_with_density_once():
    d = np.random.default_rng(5).uniform(0.5, 2, 300)
    op = small_op(True, True, density=d)
    x = rand_cplx((1, 1, 12, 10), 6)
    y = rand_cplx((1, 3, 300), 7)
    ref = op.adj_op(op.op(x) - y)
    np.testing.assert_allclose(op.data_consistency(x, y), ref, rtol=1e-12)
    # density applied exactly once: compare with the explicit matrix form
    op2 = small_op(True, True)
    r = op2.op(x) - y
    np.testing.assert_allclose(ref, op2.adj_op(r * d), rtol=1e-10)


def test_power_method_dense_norm():
    op = mri.Operator(rand_traj(200, 2, 8), (8, 6), exact=True)
    A = np.stack([op.op(e.reshape(8, 6)).ravel() for e in np.eye(48)], 1)
    smax = np.linalg.norm(A, 2) ** 2
    L = solvers.power_method(200, op, rand_cplx((8, 6), 9))
    assert L == pytest.approx(smax, rel=1e-4)


def test_cg_first_step_and_convergence():
    op = mri.Operator(rand_traj(400, 2, 10), (8, 8), exact=True)
    xt = rand_cplx((8, 8), 11).astype(np.complex128)
    y = op.op(xt)
    L = solvers.power_method(100, op, rand_cplx((8, 8), 12))
    one = solvers.cg(op, y, L, max_iter=0)
    g = op.data_consistency(np.zeros((1, 1, 8, 8)), y)
    np.testing.assert_allclose(one, -g / L, rtol=1e-12)
    # residual decreases monotonically (reference tests/operators/test_optim.py:60-68)
    res = [np.linalg.norm(op.op(solvers.cg(op, y, L, max_iter=k)) - y) for k in range(0, 12, 3)]
    assert all(b <= a * (1 + 1e-9) for a, b in zip(res, res[1:]))
    x = solvers.cg(op, y, L, max_iter=60, tol=1e-12)
    A = np.stack([op.op(e.reshape(8, 8)).ravel() for e in np.eye(64)], 1)
    xls = np.linalg.lstsq(A, y.ravel(), rcond=None)[0]
    assert rel_l2(x.ravel(), xls) < 1e-3


def test_pipe_uniform_on_cartesian_grid():
    f = np.fft.fftfreq(16)
    k = np.stack(np.meshgrid(f, f, indexing="ij"), -1).reshape(-1, 2)
    d = solvers.pipe(k, (16, 16), max_iter=10)
    assert np.std(d) / np.mean(d) < 1e-5
