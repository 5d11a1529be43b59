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


def _nyquist_radial(case):
    """The reference's Pipe cases (``tests/case_trajectories.py:40-56``):
    2D radial 128 x 16 on 32^2, and its 3D counterpart ``rotate`` with the
    default zero tilts (the 2D pattern repeated in the kz = 0 plane), here
    at 64 x 8 x 64 on 16^3 so that the numpy oracle finishes in seconds."""
    from hipnufft.trajectories import initialize_2D_radial

    if case == "radial2D":
        return initialize_2D_radial(128, 16).reshape(-1, 2).astype(np.float32), (32, 32)
    base = initialize_2D_radial(64, 8).reshape(-1, 2)
    b3 = np.concatenate([base, np.zeros((len(base), 1))], -1)
    return np.concatenate([b3] * 64).astype(np.float32), (16, 16, 16)


def _spread_interp_only(traj, shape, osf):
    """A_s A_s^H with the spread / interp-only operators on the image grid
    (the kernel of ``upsampfac = osf``), normalised like the operator
    (``sqrt(prod(N) 2^d)`` on each side), from the oracle's spreading and
    interpolation, which are pinned on their own (adjoint pair, kernel sums,
    the published kernel table)."""
    from oracle import nufft

    om = (traj * np.float32(2 * np.pi)).astype(np.float32)
    w, beta = nufft.kernel_params(1e-6, osf)
    norm2 = float(np.prod(shape) * 2 ** len(shape))
    return lambda d: nufft.interp(om, nufft.spread(om, d, shape, w, beta), w, beta) / norm2


@pytest.mark.parametrize("osf", [2.0, 1.5])
def test_pipe_fixed_point_on_radial(osf):
    """Pipe's update ``d <- d / |A_s A_s^H d|`` drives |A_s A_s^H d| to 1
    (``operators/interfaces/cufinufft.py:1311-1324``): the mean residual
    falls at least 5x from the first to the tenth iteration and ends below
    1%.  An inverted update, the full NUFFT in place of A_s, or a grid of
    the wrong size leaves it near 1."""
    traj, shape = _nyquist_radial("radial2D")
    AA = _spread_interp_only(traj, shape, osf)
    res = []
    for k in (1, 10):
        d = solvers.pipe(traj, shape, max_iter=k, osf=osf, normalize=False)
        res.append(float(np.mean(np.abs(np.abs(AA(d)) - 1))))
    assert res[1] < 0.01 and res[1] < res[0] / 5, res


@pytest.mark.parametrize("osf", [2.0, 1.5])
def test_pipe_normalised_psf_energy(osf):
    """With ``normalize`` the weights give ``mean|A^H(d * A 1)| = 1`` (the
    reference's "energy of the PSF = 1", ``finufft.py:240-244``), checked
    with the exact FP64 NDFT operator rather than the NUFFT the oracle uses."""
    traj, shape = _nyquist_radial("radial2D")
    d = solvers.pipe(traj, shape, max_iter=10, osf=osf)
    ex = mri.Operator(traj, shape, exact=True)
    rec = ex.adj_op(d * ex.op(np.ones(shape)).ravel())
    assert abs(np.mean(np.abs(rec)) - 1) < 1e-5
    assert np.all(d > 0)


@pytest.mark.parametrize("case", ["radial2D", "radial3D"])
@pytest.mark.parametrize("osf", [2.0, 1.5])
def test_pipe_correlates_with_radius(case, osf):
    """The reference's bar for the finufft family
    (``tests/operators/test_density_for_op.py:20-57``): normalised weights
    against normalised |k|, r within 0.3 and slope within 0.4 at osf 2,
    r within 0.6 at osf 1.5."""
    from helpers import assert_correlate

    traj, shape = _nyquist_radial(case)
    d = solvers.pipe(traj, shape, max_iter=10, osf=osf)
    dist = np.linalg.norm(traj, axis=-1)
    a, b = d / d.mean(), dist / dist.mean()
    if osf == 2.0:
        assert_correlate(a, b, slope=1, slope_err=0.4, r_value_err=0.3)
    else:
        assert_correlate(a, b, slope=1, slope_err=None, r_value_err=0.6)


@pytest.mark.parametrize("sense", [True, False])
def test_sense_linop_matches_operator_matrix(sense):
    """``sense_linop`` equals the oracle operator (exact transform) applied
    column by column, and its rmatvec is the conjugate transpose."""
    k = truncnorm_traj(120, 2, 31)
    shape = (6, 5)
    C = 3
    sm = gauss_smaps(C, shape, 32) if sense else None
    op = mri.Operator(k, shape, n_coils=C, smaps=sm, exact=True)
    A = solvers.sense_linop(op.omega, shape, smaps=sm, n_coils=C)  # the same float32 radians
    n_in = 30 if sense else C * 30
    x = rand_cplx(n_in, 33).astype(np.complex128)
    np.testing.assert_allclose(A.matvec(x), op.op(x.reshape(1, -1, *shape)).ravel(), rtol=1e-12)
    y = rand_cplx(C * 120, 34).astype(np.complex128)
    np.testing.assert_allclose(A.rmatvec(y), op.adj_op(y.reshape(1, C, -1)).ravel(), rtol=1e-12)
    assert abs(np.vdot(A.matvec(x), y) - np.vdot(x, A.rmatvec(y))) < 1e-12 * abs(np.vdot(x, A.rmatvec(y)))


@pytest.mark.parametrize("damp", [0.0, 0.3])
def test_lsmr_converges_to_damped_lstsq(damp):
    """LSMR iterates tend to ``(A^H A + damp^2 I)^-1 A^H b``, and its
    residual ``||A^H r||`` decreases (Fong & Saunders' monotonicity)."""
    k = truncnorm_traj(200, 2, 35)
    shape = (6, 6)
    A = solvers.sense_linop(k, shape, smaps=gauss_smaps(2, shape, 36))
    b = rand_cplx(400, 37).astype(np.complex128)
    Ad = np.stack([A.matvec(e) for e in np.eye(36)], 1)
    want = np.linalg.solve(Ad.conj().T @ Ad + damp ** 2 * np.eye(36), Ad.conj().T @ b)
    x = solvers.lsmr(A, b, 200, damp=damp, atol=0, btol=0, conlim=0)
    assert rel_l2(x, want) < 1e-8
    ar = [np.linalg.norm(Ad.conj().T @ (b - Ad @ solvers.lsmr(A, b, it, damp=damp))
                         - damp ** 2 * solvers.lsmr(A, b, it, damp=damp)) for it in (1, 3, 6, 10)]
    assert all(q <= p * (1 + 1e-9) for p, q in zip(ar, ar[1:])), ar
