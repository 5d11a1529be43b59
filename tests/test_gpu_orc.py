"""Off-resonance corrected operator on the GPU against the FP64 oracle
(oracle/orc.py): the batched L-interpolator chain through the C ABI (coils,
shots, chunks that split images), the operator wrapper, the device
factorizations, the conjugate-phase bar of the reference's test_orc.py,
Lipschitz constant and autograd."""

import math

import numpy as np
import pytest
import torch

from helpers import assert_correlate, gauss_smaps, rand_cplx, rel_l2, truncnorm_traj
from oracle import orc

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _radians(traj):
    return torch.from_numpy(np.asarray(traj, np.float32) * np.float32(2 * np.pi)).to(DEV)


@pytest.mark.parametrize("shape", [(20, 24), (10, 12, 8)])
@pytest.mark.parametrize("sense", [True, False])
@pytest.mark.parametrize("n_shots", [1, 3])
@pytest.mark.parametrize("n_trans", [5, 64])
def test_orc_chain_vs_oracle(shape, sense, n_shots, n_trans):
    """B*C*L = 2*3*4 transforms; n_trans = 5 splits images across chunks."""
    from hipnufft import _capi

    d, NK, L, nb, nc = len(shape), 70, 4, 2, 3
    M = NK * n_shots
    traj = truncnorm_traj(M, d, 1)
    B = rand_cplx((NK, L), 2)
    C = rand_cplx((L, *shape), 3)
    smaps = gauss_smaps(nc, shape, 4) if sense else None
    x = rand_cplx((nb, 1 if sense else nc, *shape), 5)
    y = rand_cplx((nb, nc, M), 6)
    dens = np.random.default_rng(7).uniform(0.5, 1.5, M).astype(np.float32)
    scale = 1.0 / math.sqrt(np.prod(shape) * 2 ** d)

    p = _capi.RawPlan(shape, n_trans=n_trans)
    p.setpts(_radians(traj))
    td = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    out = torch.empty((nb, nc, M), dtype=torch.complex64, device=DEV)
    p.orc_op(td(x), td(smaps), td(C), td(B), out, nb, nc, scale)
    ref = orc.op(traj, shape, x, B, C, smaps)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5
    img = torch.empty((nb, 1 if sense else nc, *shape), dtype=torch.complex64, device=DEV)
    p.orc_adj(td(y), td(smaps), td(C), td(B), td(dens), img, nb, nc, scale)
    ref = orc.adj(traj, shape, y, B, C, smaps, dens)
    assert rel_l2(img.cpu().numpy(), ref) < 1e-5


def test_orc_chain_rejects_bad_arguments():
    from hipnufft import _capi

    p = _capi.RawPlan((16, 16), n_trans=4)
    p.setpts(_radians(truncnorm_traj(100, 2, 8)))
    C = torch.ones((2, 16, 16), dtype=torch.complex64, device=DEV)
    x = torch.ones((1, 1, 16, 16), dtype=torch.complex64, device=DEV)
    out = torch.empty((1, 1, 100), dtype=torch.complex64, device=DEV)
    with pytest.raises(ValueError):  # 30 samples per shot do not divide 100
        p.orc_op(x, None, C, torch.ones((30, 2), dtype=torch.complex64, device=DEV), out, 1, 1,
                 1.0)
    with pytest.raises(ValueError):  # B has 3 columns, C 2 interpolators
        p.orc_op(x, None, C, torch.ones((50, 3), dtype=torch.complex64, device=DEV), out, 1, 1,
                 1.0)


@pytest.mark.parametrize("kind", ["numpy", "torch-cuda"])
def test_operator_with_given_interpolators(kind):
    from hipnufft import MRIFourierCorrected, MRIHipNUFFT

    shape, NK, n_shots, L, nc = (24, 20), 90, 2, 5, 4
    traj = truncnorm_traj(NK * n_shots, 2, 9)
    smaps = gauss_smaps(nc, shape, 10)
    dens = np.random.default_rng(11).uniform(0.5, 1.5, NK * n_shots).astype(np.float32)
    fo = MRIHipNUFFT(traj, shape, n_coils=nc, smaps=smaps, density=dens)
    B = rand_cplx((NK, L), 12)
    C = rand_cplx((L, *shape), 13)
    op = fo.with_off_resonance_correction(None, interpolator=(B, C))
    assert isinstance(op, MRIFourierCorrected)
    assert op.n_interpolators == L and op.n_samples_per_shot == NK and op.n_shots == n_shots
    x = rand_cplx(shape, 14)
    y = rand_cplx((nc, NK * n_shots), 15)
    if kind == "torch-cuda":
        x, y = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)
    k = op.op(x)
    im = op.adj_op(y)
    assert type(k) is type(x) and tuple(k.shape) == (1, nc, NK * n_shots)
    assert tuple(im.shape) == (1, 1, *shape)
    cpu = lambda a: a.cpu().numpy() if isinstance(a, torch.Tensor) else a
    assert rel_l2(cpu(k), orc.op(traj, shape, cpu(x)[None, None], B, C, smaps)) < 1e-5
    assert rel_l2(cpu(im), orc.adj(traj, shape, cpu(y)[None], B, C, smaps, dens)) < 1e-5
    # data consistency is adj_op(op(x) - y)
    dc = cpu(op.data_consistency(x, y))
    assert rel_l2(dc, cpu(op.adj_op(cpu(op.op(x)) - cpu(y)))) < 1e-5


def test_zero_field_is_the_plain_operator():
    from hipnufft import MRIHipNUFFT

    shape = (16, 20)
    traj = truncnorm_traj(400, 2, 16)
    fo = MRIHipNUFFT(traj, shape, n_coils=2, squeeze_dims=True)
    op = fo.with_off_resonance_correction(np.linspace(0, 5e-3, 400), b0_map=np.zeros(shape))
    assert op.n_interpolators == 1
    x = rand_cplx((2, *shape), 17)
    assert rel_l2(op.op(x), fo.op(x)) < 1e-6
    y = rand_cplx((2, 400), 18)
    assert rel_l2(op.adj_op(y), fo.adj_op(y)) < 1e-6


@pytest.mark.parametrize("method", ["svd", "mti"])
@pytest.mark.parametrize("complex_", [False, True])
def test_device_factorizations_vs_oracle(method, complex_):
    from hipnufft import get_orc_factorization, make_b0map, make_t2smap

    b0, mask = make_b0map((48, 48))
    r2s = None
    if complex_:
        t2s, _ = make_t2smap((48, 48), 15.0)
        r2s = mask * (1.0 / (t2s * 1e-3 + 1e-9))
    f = orc.complex_fieldmap(b0, r2s).astype(np.complex64)
    t = np.linspace(0.0, 5e-3, 400).astype(np.float32)
    B, C, _ = get_orc_factorization(method)(torch.from_numpy(f).to(DEV),
                                            torch.from_numpy(t).to(DEV),
                                            torch.from_numpy(mask).to(DEV), L=40, n_bins=4096)
    assert B.is_cuda and C.is_cuda and tuple(C.shape) == (40, 48, 48)
    B, C = B.cpu().numpy(), C.cpu().numpy()
    E = np.exp(np.outer(t, f[mask]))
    np.testing.assert_allclose(B @ C[:, mask], E, atol=5e-3, rtol=5e-3)
    rB, rC, _ = getattr(orc, method)(f, t, mask, L=40, n_bins=4096)
    assert rel_l2(B @ C.reshape(40, -1), rB @ rC.reshape(40, -1)) < 1e-4


def test_device_mfi():
    from hipnufft import get_orc_factorization, make_b0map

    b0, mask = make_b0map((48, 48))
    f = orc.complex_fieldmap(b0).astype(np.complex64)
    t = np.linspace(0.0, 5e-3, 400).astype(np.float32)
    B, C, _ = get_orc_factorization("mfi")(torch.from_numpy(f).to(DEV), torch.from_numpy(t).to(DEV),
                                           torch.from_numpy(mask).to(DEV), L=40, n_bins=4096)
    C = torch.stack([C[l] for l in range(len(C))])
    np.testing.assert_allclose((B @ C.reshape(40, -1)).cpu().numpy()[:, mask.ravel()],
                               np.exp(np.outer(t, f[mask])), atol=5e-3, rtol=5e-3)


@pytest.fixture(scope="module")
def cp_case():
    """The reference's test_orc.py case: 1000 truncated-normal samples,
    (64, 128) image, radial B0 map in [-300, 300] Hz, readout 0 .. 50 ms,
    SVD interpolators with the automatic L."""
    from hipnufft import MRIHipNUFFT, make_b0map

    rng = np.random.default_rng(0)
    s = rng.normal(0, 0.16, 6000)
    traj = s[np.abs(s) < 0.48][:2000].reshape(1000, 2).astype(np.float32)
    shape = (64, 128)
    fo = MRIHipNUFFT(traj, shape, n_coils=1, n_batchs=1, density=False, squeeze_dims=True)
    b0, mask = make_b0map(shape, b0range=(-300, 300))
    t = np.linspace(0, 5e-2, 1000, dtype=np.float32)
    op = fo.with_off_resonance_correction(b0_map=b0, mask=mask, readout_time=t)
    E = orc.extended_fourier_matrix(traj, shape, orc.complex_fieldmap(b0), t)
    return op, E, shape


def test_orc_forward_conjugate_phase(cp_case):
    op, E, shape = cp_case
    x = rand_cplx(shape, 19)
    got = np.asarray(op.op(x))
    ideal = E @ x.ravel()
    assert_correlate(got.squeeze(), ideal, slope_err=0.005, r_value_err=0.002)
    print(f"L = {op.n_interpolators}, forward rel err vs conjugate phase {rel_l2(got, ideal):.2e}")


def test_orc_adjoint_conjugate_phase(cp_case):
    op, E, shape = cp_case
    y = rand_cplx(1000, 20)
    got = np.asarray(op.adj_op(y))
    ideal = (E.conj().T @ y).reshape(shape)
    assert_correlate(got.squeeze(), ideal, slope_err=0.005, r_value_err=0.002)
    print(f"adjoint rel err vs conjugate phase {rel_l2(got, ideal):.2e}")


def test_lipschitz_vs_oracle_norm():
    from hipnufft import MRIHipNUFFT

    shape, NK, L = (16, 16), 200, 3
    traj = truncnorm_traj(NK, 2, 21)
    B = rand_cplx((NK, L), 22)
    C = rand_cplx((L, *shape), 23)
    fo = MRIHipNUFFT(traj, shape, n_coils=4, smaps=gauss_smaps(4, shape, 24))
    op = fo.with_off_resonance_correction(None, interpolator=(B, C))
    lip = op.get_lipschitz_cst(max_iter=200)
    A = np.stack([orc.op(traj, shape, e.reshape(1, 1, *shape), B, C).ravel()
                  for e in np.eye(int(np.prod(shape)))], axis=1)
    ref = np.linalg.norm(A, 2) ** 2
    assert abs(lip - ref) < 1e-2 * ref


def test_autograd_data_and_field_map():
    from hipnufft import MRIHipNUFFT

    shape, NK, L = (16, 12), 150, 3
    traj = truncnorm_traj(NK, 2, 25)
    fo = MRIHipNUFFT(traj, shape, n_coils=1)
    B = rand_cplx((NK, L), 26)
    C = rand_cplx((L, *shape), 27)
    t = np.linspace(0, 4e-3, NK).astype(np.float32)
    op = fo.with_off_resonance_correction(t, b0_map=np.zeros(shape, np.float32),
                                          interpolator=(B, C))
    ad = op.make_autograd(wrt_data=True, wrt_field_map=True)
    x = torch.from_numpy(rand_cplx((1, 1, *shape), 28)).to(DEV).requires_grad_(True)
    v = torch.from_numpy(rand_cplx((1, 1, NK), 29)).to(DEV)
    y = ad.op(x)
    y.backward(v)
    xn, vn = x.detach().cpu().numpy(), v.cpu().numpy()
    assert rel_l2(x.grad.cpu().numpy(), orc.adj(traj, shape, vn, B, C)) < 1e-5
    fm_grad = ad.field_map.grad.cpu().numpy()
    ref = xn * orc.adj(traj, shape, vn * t, B, C)  # autodiff.py:44-55
    assert rel_l2(fm_grad.reshape(ref.shape), ref) < 1e-5
    with pytest.raises(ValueError):
        op.make_autograd(wrt_traj=True)


def test_stacked_operator_wrapped_by_interpolator_loop():
    """A stacked operator is wrapped the reference's way (L calls of its op /
    adj_op): against the stacked oracle composed with the B / C weights."""
    from hipnufft import MRIStackedHipNUFFT
    from oracle import stacked as ost

    shape, NS, L, nc = (12, 10, 8), 60, 3, 2
    s2 = truncnorm_traj(NS, 2, 30)
    zi = np.array([0, 2, 3, 5, 7])
    M = NS * len(zi)
    smaps = gauss_smaps(nc, shape, 31)
    st = MRIStackedHipNUFFT(s2, shape, n_coils=nc, smaps=smaps, z_index=zi)
    B = rand_cplx((NS, L), 32)  # one row per sample of a plane: NK = NS, a shot per plane
    C = rand_cplx((L, *shape), 33)
    op = st.with_off_resonance_correction(None, interpolator=(B, C))
    assert not op._fused and op.n_shots == len(zi)
    x = rand_cplx(shape, 34)
    Bf = B[np.arange(M) % NS]
    ref = sum(Bf[:, l] * ost.op(s2, zi, (C[l] * x)[None], shape, smaps) for l in range(L))
    got = np.asarray(op.op(x))
    assert rel_l2(got.reshape(ref.shape), ref) < 1e-5
    y = rand_cplx((nc, M), 35)
    ref = sum(np.conj(C[l]) * ost.adj(s2, zi, (np.conj(Bf[:, l]) * y)[None], shape, smaps)
              for l in range(L))
    got = np.asarray(op.adj_op(y))
    assert rel_l2(got.reshape(ref.shape), ref) < 1e-5
    assert op.get_lipschitz_cst(max_iter=20) > 0
