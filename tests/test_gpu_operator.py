"""The hipnufft operator against the oracle, mirroring the reference's tests.

References (behaviour mirrored, tolerances from the reference or tighter):
* tests/operators/test_operator_ref.py:49-74   backend vs NDFT
* tests/operators/test_interfaces.py:138-208   adjointness, Lipschitz, shape errors
* tests/operators/test_batch.py:106-272         batches, DC, density once, no mutation, solvers
* tests/operators/test_update.py:131-248        update_samples == fresh operator
* tests/operators/test_density_for_op.py:20-57  Pipe density vs radial distance
* tests/operators/test_autodiff.py:185-374      autograd vs NDFT-matrix autograd
"""

import numpy as np
import pytest
import torch

from helpers import gauss_smaps, radial2d, rand_cplx, rand_traj, rel_l2, truncnorm_traj
from oracle import mri, ndft, solvers

pytestmark = pytest.mark.gpu
DEV = "cuda"


def get_op(*args, **kw):
    import hipnufft

    return hipnufft.get_operator("hipnufft")(*args, **kw)


def grid_case(N, d):
    f = np.fft.fftfreq(N)
    return np.stack(np.meshgrid(*([f] * d)), axis=-1).reshape(-1, d), (N,) * d


CASES = {
    "random2D": lambda: (truncnorm_traj(1000, 2, 0), (64, 128)),
    "grid2D": lambda: grid_case(16, 2),
    "grid3D": lambda: grid_case(16, 3),
    "radial2D": lambda: (radial2d(10, 500), (64, 64)),
    "random3D": lambda: (rand_traj(20000, 3, 1), (24, 32, 20)),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("sense", [False, True])
def test_op_adj_vs_ndft(case, sense):
    k, shape = CASES[case]()
    C = 3 if sense else 2
    smaps = gauss_smaps(C, shape, 5) if sense else None
    op = get_op(k, shape, n_coils=C, smaps=smaps)
    ref = mri.Operator(k, shape, n_coils=C, smaps=smaps, exact=True)
    x = rand_cplx((1, 1 if sense else C, *shape), 6)
    y = rand_cplx((1, C, len(k)), 7)
    assert rel_l2(op.op(x), ref.op(x)) < 2e-5
    assert rel_l2(op.adj_op(y), ref.adj_op(y)) < 2e-5
    # reference-style elementwise check (atol = rtol = 1e-4)
    np.testing.assert_allclose(op.op(x), ref.op(x), atol=1e-4, rtol=1e-4)


def test_autoadjoint():
    k, shape = CASES["random2D"]()
    op = get_op(k, shape, n_coils=2)
    rel = []
    for i in range(5):
        x = torch.from_numpy(rand_cplx((1, 2, *shape), 10 + i)).to(DEV)
        y = torch.from_numpy(rand_cplx((1, 2, len(k)), 20 + i)).to(DEV)
        lhs = torch.vdot(op.op(x).reshape(-1).to(torch.complex128), y.reshape(-1).to(torch.complex128))
        rhs = torch.vdot(x.reshape(-1).to(torch.complex128), op.adj_op(y).reshape(-1).to(torch.complex128))
        rel.append(float(abs(lhs - rhs) / abs(rhs)))
    assert np.mean(rel) < 5e-5


def test_shapes_squeeze_and_errors():
    k, shape = CASES["random2D"]()
    for B, C, sq in [(1, 1, True), (2, 3, False), (1, 4, True), (3, 1, True)]:
        op = get_op(k, shape, n_coils=C, n_batchs=B, squeeze_dims=sq)
        x = torch.zeros((B, C, *shape), dtype=torch.complex64, device=DEV)
        y = op.op(x)
        z = op.adj_op(y)
        exp_k = tuple(s for i, s in enumerate((B, C, len(k))) if not (sq and i < 2 and s == 1))
        exp_i = tuple(s for i, s in enumerate((B, C, *shape)) if not (sq and i < 2 and s == 1))
        assert tuple(y.shape) == exp_k and tuple(z.shape) == exp_i
        assert op.img_full_shape == (B, C, *shape) and op.ksp_full_shape == (B, C, len(k))
    with pytest.raises(ValueError):
        op.check_shape(image=np.zeros((shape[0] + 10, shape[1] + 1)))
    with pytest.raises(ValueError):
        op.check_shape(ksp=np.zeros((C, len(k) + 10)))
    with pytest.raises(ValueError):
        op.op(np.zeros((shape[0] + 1, shape[1]), np.complex64))
    with pytest.raises(ValueError):
        get_op(k, shape, n_coils=4, n_trans=3)


def test_array_interfaces():
    k, shape = CASES["random2D"]()
    op = get_op(k, shape)
    x = rand_cplx(shape, 30)
    y_np = op.op(x)
    assert isinstance(y_np, np.ndarray)
    y_cpu = op.op(torch.from_numpy(x))
    assert isinstance(y_cpu, torch.Tensor) and not y_cpu.is_cuda
    y_gpu = op.op(torch.from_numpy(x).to(DEV))
    assert y_gpu.is_cuda
    np.testing.assert_allclose(y_np, y_cpu.numpy(), rtol=0, atol=0)
    np.testing.assert_allclose(y_np, y_gpu.cpu().numpy(), rtol=0, atol=0)
    z = op.adj_op(y_np)
    assert isinstance(z, np.ndarray) and z.shape == (1, 1, *shape)


@pytest.mark.parametrize("sense", [False, True])
def test_batch_equals_per_batch(sense):
    k, shape = CASES["radial2D"]()
    B, C = 3, 2
    smaps = gauss_smaps(C, shape, 40) if sense else None
    op = get_op(k, shape, n_coils=C, n_batchs=B, smaps=smaps)
    one = get_op(k, shape, n_coils=C, n_batchs=1, smaps=smaps)
    x = rand_cplx((B, 1 if sense else C, *shape), 41)
    y = rand_cplx((B, C, len(k)), 42)
    yb = op.op(x)
    zb = op.adj_op(y)
    for b in range(B):
        np.testing.assert_allclose(yb[b], one.op(x[b:b + 1])[0], rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(zb[b], one.adj_op(y[b:b + 1])[0], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("sense", [False, True])
@pytest.mark.parametrize("density", [False, True])
def test_data_consistency_and_density_once(sense, density):
    k, shape = CASES["random2D"]()
    C = 3
    smaps = gauss_smaps(C, shape, 50) if sense else None
    d = np.random.default_rng(51).uniform(0.5, 2, len(k)).astype(np.float32) if density else None
    op = get_op(k, shape, n_coils=C, smaps=smaps, density=d)
    x = rand_cplx((1, 1 if sense else C, *shape), 52)
    y = rand_cplx((1, C, len(k)), 53)
    dc = op.data_consistency(x, y)
    naive = op.adj_op(op.op(x) - y)
    assert rel_l2(dc, naive) < 1e-5
    ref = mri.Operator(k, shape, n_coils=C, smaps=smaps, density=d, exact=True)
    assert rel_l2(dc, ref.data_consistency(x, y)) < 2e-5


def test_inputs_not_mutated():
    k, shape = CASES["random2D"]()
    smaps = gauss_smaps(2, shape, 60)
    d = np.ones(len(k), np.float32) * 0.5
    op = get_op(k, shape, n_coils=2, smaps=smaps, density=d)
    x = torch.from_numpy(rand_cplx((1, 1, *shape), 61)).to(DEV)
    y = torch.from_numpy(rand_cplx((1, 2, len(k)), 62)).to(DEV)
    x0, y0 = x.clone(), y.clone()
    op.op(x)
    op.adj_op(y)
    op.data_consistency(x, y)
    assert torch.equal(x, x0) and torch.equal(y, y0)


def test_update_samples_equals_fresh():
    k1, shape = CASES["random2D"]()
    k2 = truncnorm_traj(1000, 2, 99)
    op = get_op(k1, shape, n_coils=2)
    op.update_samples(k2)
    fresh = get_op(k2, shape, n_coils=2)
    x = rand_cplx((1, 2, *shape), 70)
    np.testing.assert_array_equal(op.op(x), fresh.op(x))
    y = rand_cplx((1, 2, len(k2)), 71)
    np.testing.assert_array_equal(op.adj_op(y), fresh.adj_op(y))


@pytest.mark.parametrize("name", ["random2D", "random3D"])
def test_lipschitz_toeplitz_form(name):
    """The power iterations on the Toeplitz form give the NUFFT form's value
    from the same start vector (the two operators agree to the NUFFT
    tolerance)."""
    k, shape = CASES[name]()
    op = get_op(k, shape, n_coils=2, density=np.linspace(0.5, 1.5, len(k)).astype(np.float32))
    L = op.get_lipschitz_cst(10, seed=5)
    Lt = op.get_lipschitz_cst(10, seed=5, toeplitz=True)
    assert abs(Lt - L) <= 1e-4 * L
    assert op.n_coils == 2 and op._smaps is None


def test_lipschitz_upper_bound():
    from hipnufft import power_method

    k, shape = CASES["random2D"]()
    op = get_op(k, shape)
    L = op.get_lipschitz_cst(20, seed=3)
    vals = []
    for i in range(5):
        a = rand_cplx(shape, 80 + i)
        b = rand_cplx(shape, 90 + i)
        num = np.linalg.norm(op.adj_op(op.op(a)) - op.adj_op(op.op(b)))
        vals.append(num / np.linalg.norm(a - b))
    assert np.mean(vals) < 1.1 * L
    # the power iteration itself, from the oracle's start vector
    ref = mri.Operator(k, shape, exact=False)
    rng = np.random.default_rng(3)
    x0 = 1j * rng.random(shape) + rng.random(shape)
    Lr = solvers.power_method(20, ref, x0)
    one = get_op(k, shape, squeeze_dims=True)
    Lg, _ = power_method(20, one, x=torch.from_numpy(x0.astype(np.complex64)).to(DEV),
                         norm_func=lambda v: torch.linalg.vector_norm(v.reshape(-1)))
    assert float(Lg) == pytest.approx(Lr, rel=1e-3)
    assert L == pytest.approx(Lr, rel=2e-2)  # another start vector, 20 iterations


@pytest.mark.parametrize("optim", ["cg", "lsqr", "lsmr"])
def test_pinv_solver_shapes(optim):
    k, shape = CASES["radial2D"]()
    smaps = gauss_smaps(2, shape, 100)
    op = get_op(k, shape, n_coils=2, smaps=smaps)
    y = op.op(rand_cplx(shape, 101))
    x = op.pinv_solver(y, optim=optim, max_iter=5)
    assert tuple(x.shape) == (1, 1, *shape)
    assert np.isfinite(np.asarray(x)).all()


def test_cg_matches_oracle_cg():
    """The power method from the same start vector gives the oracle's
    Lipschitz constant; CG with that constant gives the oracle's iterates.
    (The operator draws its own start vector on the device, unseeded like
    the reference's numpy draw, so the constant is handed over.)"""
    from hipnufft import power_method

    k, shape = CASES["random2D"]()
    smaps = gauss_smaps(2, shape, 110)
    op = get_op(k, shape, n_coils=2, smaps=smaps)
    xt = rand_cplx(shape, 111)
    y = op.op(xt)
    rng = np.random.default_rng(7)
    x0 = 1j * rng.random(shape) + rng.random(shape)
    single = mri.Operator(k, shape)
    L = solvers.power_method(10, single, x0)
    one = get_op(k, shape, squeeze_dims=True)
    Lg, _ = power_method(10, one, x=torch.from_numpy(x0.astype(np.complex64)).to(DEV),
                         norm_func=lambda v: torch.linalg.vector_norm(v.reshape(-1)))
    assert abs(float(Lg) - L) < 1e-5 * L
    op.get_lipschitz_cst = lambda **_: L
    x = op.pinv_solver(y, optim="cg", max_iter=8)
    ref = mri.Operator(k, shape, n_coils=2, smaps=smaps)
    xr = solvers.cg(ref, y.astype(np.complex128), L, max_iter=8)
    assert rel_l2(x, xr) < 1e-4


def test_cg_graph_matches_eager_over_replays():
    k, shape = CASES["radial2D"]()
    smaps = gauss_smaps(4, shape, 120)
    op = get_op(k, shape, n_coils=4, smaps=smaps)
    y = torch.from_numpy(op.op(rand_cplx(shape, 121))).to(DEV)
    L = op.get_lipschitz_cst(seed=11)
    op.get_lipschitz_cst = lambda **_: L
    eager = op.pinv_solver(y, optim="cg", max_iter=12, graph=False)
    for rep in range(3):
        # eager work and synchronisation between graph runs
        _ = op.adj_op(op.op(torch.ones(shape, dtype=torch.complex64, device=DEV)))
        torch.cuda.synchronize()
        g = op.pinv_solver(y, optim="cg", max_iter=12, graph=True)
        torch.cuda.synchronize()
        assert float(torch.linalg.vector_norm(g - eager) / torch.linalg.vector_norm(eager)) < 1e-6


def test_cg_early_stop_predicated():
    # with a huge tolerance the reference stops at the first check: one step
    k, shape = CASES["radial2D"]()
    op = get_op(k, shape)
    y = op.op(rand_cplx(shape, 130))
    L = op.get_lipschitz_cst(seed=1)
    op.get_lipschitz_cst = lambda **_: L
    a = op.pinv_solver(y, optim="cg", max_iter=6, tol=1e9, graph=False)
    b = op.pinv_solver(y, optim="cg", max_iter=6, tol=1e9, graph=True)
    np.testing.assert_allclose(np.asarray(a), np.asarray(b.cpu() if hasattr(b, "cpu") else b),
                               rtol=1e-6, atol=1e-7)


def _correlate(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    r = np.corrcoef(a, b)[0, 1]
    slope = np.cov(a, b)[0, 1] / np.var(a)
    return r, abs(slope)


def _pipe_case(case):
    """The reference's Pipe cases (``tests/case_trajectories.py:40-56``):
    2D radial 128 x 16 on 32^2, and ``case_nyquist_radial3D`` -- ``rotate()``
    with its default zero tilts, i.e. the 2D radial repeated 128 times in
    the kz = 0 plane, on 32^3; ``radial3D_small`` is the same construction
    at 64 x 8 x 64 on 16^3, small enough for the numpy oracle."""
    from hipnufft.trajectories import initialize_2D_radial

    if case == "radial2D_nyq":
        traj, shape = initialize_2D_radial(32 * 4, 16), (32, 32)
    else:
        nc, ns, nr, n = (128, 16, 128, 32) if case == "radial3D_nyq" else (64, 8, 64, 16)
        base = initialize_2D_radial(nc, ns)
        b3 = np.concatenate([base, np.zeros((*base.shape[:2], 1))], -1)
        traj, shape = np.concatenate([b3] * nr), (n, n, n)
    return traj.reshape(-1, traj.shape[-1]).astype(np.float32), shape


@pytest.mark.parametrize("osf", [2, 1.5])
@pytest.mark.parametrize("case", ["radial2D_nyq", "radial3D_nyq", "radial3D_small"])
def test_pipe_density(case, osf):
    """Pipe on the device: the reference's bar for the finufft family
    (``tests/operators/test_density_for_op.py:20-57``: r within 0.3 and slope
    within 0.4 at osf 2, r within 0.6 at osf 1.5) and, where the numpy oracle
    finishes in seconds, the oracle's Pipe (same fixed-point iterations on
    spread / interp-only operators, pinned by its fixed point and PSF energy)."""
    import hipnufft

    traj, shape = _pipe_case(case)
    d = hipnufft.get_density("pipe")(traj, shape, backend="hipnufft", osf=osf, max_iter=10)
    d = np.asarray(d)
    dist = np.linalg.norm(traj, axis=-1)
    a, b = d / d.mean(), dist / dist.mean()
    if osf == 2:
        r, slope = _correlate(a, b)
        assert abs(r - 1) < 0.3 and abs(slope - 1) < 0.4
    else:
        r, _ = _correlate(a, b)
        assert abs(r - 1) < 0.6
    if case != "radial3D_nyq":
        ref = solvers.pipe(traj, shape, max_iter=10, osf=float(osf))
        err = rel_l2(d, ref)
        print(f"pipe {case} osf {osf}: rel L2 vs oracle {err:.2e}")
        assert err < 1e-3


@pytest.mark.parametrize("sense", [False, True])
def test_autograd_vs_ndft_matrix(sense):
    k, shape = (truncnorm_traj(300, 2, 140), (16, 12))
    C = 2
    smaps = gauss_smaps(C, shape, 141) if sense else None
    op = get_op(k, shape, n_coils=C, smaps=smaps)
    ad = op.make_autograd(wrt_data=True, wrt_traj=True)
    xin = torch.from_numpy(rand_cplx((1, 1 if sense else C, *shape), 142)).to(DEV)
    x = xin.clone().requires_grad_(True)
    target = torch.from_numpy(rand_cplx((1, C, len(k)), 143)).to(DEV)
    loss = torch.sum(torch.abs(ad.op(x) - target) ** 2)
    gx, gt = torch.autograd.grad(loss, [x, ad.samples])

    # FP64 NDFT matrix on the same radians, differentiated by torch
    w = ad.samples.detach().double().cpu().clone().requires_grad_(True)
    r = [torch.linspace(-s / 2, s / 2 - 1, s, dtype=torch.float64) for s in shape]
    grid = torch.stack(torch.meshgrid(*r, indexing="ij")).reshape(2, -1)
    A = torch.exp(-1j * (w @ grid)) / np.sqrt(np.prod(shape) * 4)
    xd = xin.detach().cpu().to(torch.complex128).clone().requires_grad_(True)
    if sense:
        S = torch.from_numpy(smaps).to(torch.complex128).reshape(C, -1)
        y = (A @ (S * xd.reshape(1, -1)).T).T[None]
    else:
        y = (A @ xd.reshape(C, -1).T).T[None]
    lref = torch.sum(torch.abs(y - target.cpu().to(torch.complex128)) ** 2)
    gxr, gtr = torch.autograd.grad(lref, [xd, w])
    assert rel_l2(gx.cpu().numpy(), gxr.numpy()) < 1e-4
    assert rel_l2(gt.cpu().numpy(), gtr.numpy()) < 1e-3


def test_autograd_adjoint_traj_vs_ndft_matrix():
    k, shape = (truncnorm_traj(250, 2, 150), (12, 10))
    op = get_op(k, shape, n_coils=1)
    ad = op.make_autograd(wrt_data=True, wrt_traj=True)
    yin = torch.from_numpy(rand_cplx((1, 1, len(k)), 151)).to(DEV)
    y = yin.clone().requires_grad_(True)
    tgt = torch.from_numpy(rand_cplx((1, 1, *shape), 152)).to(DEV)
    loss = torch.sum(torch.abs(ad.adj_op(y) - tgt) ** 2)
    gy, gt = torch.autograd.grad(loss, [y, ad.samples])
    w = ad.samples.detach().double().cpu().clone().requires_grad_(True)
    r = [torch.linspace(-s / 2, s / 2 - 1, s, dtype=torch.float64) for s in shape]
    grid = torch.stack(torch.meshgrid(*r, indexing="ij")).reshape(2, -1)
    A = torch.exp(-1j * (w @ grid)) / np.sqrt(np.prod(shape) * 4)
    yd = yin.detach().cpu().to(torch.complex128).clone().requires_grad_(True)
    x = (A.conj().T @ yd.reshape(-1)).reshape(1, 1, *shape)
    lref = torch.sum(torch.abs(x - tgt.cpu().to(torch.complex128)) ** 2)
    gyr, gtr = torch.autograd.grad(lref, [yd, w])
    assert rel_l2(gy.cpu().numpy(), gyr.numpy()) < 1e-4
    assert rel_l2(gt.cpu().numpy(), gtr.numpy()) < 1e-3


def test_autograd_traj_central_fd():
    k, shape = (truncnorm_traj(200, 2, 160), (16, 16))
    op = get_op(k, shape, n_coils=1)
    ad = op.make_autograd(wrt_data=False, wrt_traj=True)
    x = torch.from_numpy(rand_cplx((1, 1, *shape), 161)).to(DEV)
    tgt = torch.from_numpy(rand_cplx((1, 1, len(k)), 162)).to(DEV)
    loss = torch.sum(torch.abs(ad.op(x) - tgt) ** 2)
    (gt,) = torch.autograd.grad(loss, [ad.samples])
    base = op.samples.detach().clone()
    h = 1e-3
    rng = np.random.default_rng(163)
    for m, d in zip(rng.integers(0, len(k), 6), rng.integers(0, 2, 6)):
        vals = []
        for s in (h, -h):
            w = base.clone()
            w[m, d] += s
            op2 = get_op(w, shape, n_coils=1)
            vals.append(float(torch.sum(torch.abs(op2.op(x) - tgt) ** 2)))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - float(gt[m, d])) <= 2e-2 * max(1.0, abs(fd))


@pytest.mark.parametrize("dim,sense,density", [(2, True, False), (2, False, True),
                                               (3, True, True), (2, True, True)])
def test_gram_toeplitz_matches_adj_op(dim, sense, density):
    """gram_op (Toeplitz kernel) == adj_op(op(x)) (reference test_batch.py:275-293:
    rtol 2e-4, atol 1e-4)."""
    from hipnufft import MRIHipNUFFT

    shape = (48, 40) if dim == 2 else (16, 20, 12)
    traj = truncnorm_traj(3000, dim, 90)
    C = 3
    smaps = gauss_smaps(C, shape, seed=91) if sense else None
    dens = np.random.default_rng(92).uniform(0.2, 2.0, traj.shape[0]).astype(np.float32) \
        if density else False
    op = MRIHipNUFFT(traj, shape, n_coils=C, n_batchs=2, smaps=smaps, density=dens)
    ic = 1 if sense else C
    x = torch.from_numpy(rand_cplx((2, ic, *shape), seed=93)).cuda()
    ref = op.adj_op(op.op(x)).cpu().numpy()
    g = op.gram_op(x).cpu().numpy()
    assert g.shape == ref.shape
    np.testing.assert_allclose(g, ref, rtol=2e-4, atol=1e-4 * np.abs(ref).max())
    assert rel_l2(g, ref) < 2e-5
    # the direct path and a recomputed kernel after a density change
    np.testing.assert_allclose(op.gram_op(x, toeplitz=False).cpu().numpy(), ref, rtol=1e-6, atol=1e-6)
    op.density = None
    ref2 = op.adj_op(op.op(x)).cpu().numpy()
    assert rel_l2(op.gram_op(x).cpu().numpy(), ref2) < 2e-5


@pytest.mark.parametrize("graph", [False, True])
def test_cg_toeplitz_matches_nufft_cg(graph):
    """CG with Toeplitz gradients equals CG with NUFFT data consistency
    (same L, same recurrences) to the 1e-4 reconstruction bar."""
    from hipnufft import MRIHipNUFFT, cg

    shape = (40, 48)
    traj = radial2d(40, 160)
    smaps = gauss_smaps(4, shape, seed=95)
    op = MRIHipNUFFT(traj, shape, n_coils=4, smaps=smaps)
    x = torch.from_numpy(rand_cplx(shape, seed=96)).cuda()
    y = op.op(x)
    L = op.get_lipschitz_cst(seed=0)
    op.get_lipschitz_cst = lambda **_: L
    a = cg(op, y, max_iter=12, graph=graph).cpu().numpy()
    b = cg(op, y, max_iter=12, graph=graph, toeplitz=True).cpu().numpy()
    assert rel_l2(b, a) < 1e-4


@pytest.mark.parametrize("dim,sense,density", [(2, True, True), (3, False, True),
                                               (2, False, False), (3, True, False)])
def test_gram_toeplitz_vs_exact_normal_operator(dim, sense, density):
    """The Toeplitz ``gram_op`` against the exact FP64 normal operator
    A^H D A of the oracle (NDFT, density applied once), not against the
    NUFFT ``adj_op(op(.))`` it replaces."""
    shape = (24, 20) if dim == 2 else (12, 10, 8)
    traj = truncnorm_traj(2000, dim, 97)
    C = 2
    smaps = gauss_smaps(C, shape, 98) if sense else None
    dens = (np.random.default_rng(99).uniform(0.5, 2.0, len(traj)).astype(np.float32)
            if density else None)
    op = get_op(traj, shape, n_coils=C, smaps=smaps, density=dens if density else False)
    ref = mri.Operator(op.samples.cpu().numpy(), shape, n_coils=C, smaps=smaps, density=dens,
                       exact=True)
    x = rand_cplx((1, 1 if sense else C, *shape), 100)
    g = op.gram_op(torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert rel_l2(g, ref.adj_op(ref.op(x))) < 2e-5
