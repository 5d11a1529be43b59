"""Parity at the headline configuration and on 192- and 384-point image axes.

Headline: 3D golden-angle phyllotaxis radial, 4096 spokes x 512 samples
(2,097,152 points), image 192^3 (ellipsoid phantom), 16 Gaussian-bump SENSE
coils, eps 1e-6 -- the 384-point fine-grid axes run the compile-time FFT
instances of the row and strided passes.  Mirrors the reference's backend
check against the NDFT (``tests/operators/test_operator_ref.py:49-74``,
atol = rtol = 1e-4) with a subsampled exact oracle:

* op at 4096 seeded random samples (all 16 coils) and adj_op at 2048 seeded
  random voxels against the exact FP64 NDFT (``tests/ndft_gpu.py``);
* adjointness ``<A x, y> = <x, A^H y>`` (``test_interfaces.py:138-158``);
* the 20-iteration graph-captured CG equals the eager CG at the same L.

The CG accuracy bar of the metric (rel err <= 1e-4) on a reduced 3D problem
of the same kind (phyllotaxis 1024 x 128 = 131,072 points, 48^3, 4 coils):
the HIP CG against the oracle's CG (``oracle/solvers.py:cg``, restating
``extras/optim.py:801-902``) run on the exact FP64 NDFT operator, same L,
with and without density compensation.
"""

import numpy as np
import pytest
import torch

import ndft_gpu
from helpers import gauss_smaps, rand_cplx, rel_l2, truncnorm_traj
from oracle import solvers

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _rel(a, b):
    a = a.to(torch.complex128).reshape(-1)
    b = b.to(torch.complex128).reshape(-1)
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.fixture(scope="module")
def headline():
    from hipnufft import MRIHipNUFFT
    from hipnufft.trajectories import (ellipsoid_phantom, gaussian_smaps,
                                       initialize_3D_phyllotaxis_radial)

    shape, C = (192, 192, 192), 16
    traj = initialize_3D_phyllotaxis_radial(4096, 512).reshape(-1, 3).astype(np.float32)
    smaps = gaussian_smaps(C, shape, seed=1)
    op = MRIHipNUFFT(traj, shape, n_coils=C, smaps=smaps)
    assert tuple(op.raw_op.plan.info()["nf"]) == (384, 384, 384)
    x = torch.from_numpy(ellipsoid_phantom(shape, seed=0)).to(DEV)
    return op, x


def test_headline_op_adj_vs_ndft(headline):
    op, x = headline
    C, M, shape = op.n_coils, op.n_samples, op.shape
    y = op.op(x).reshape(C, M)
    z = op.adj_op(y.reshape(1, C, M)).reshape(shape)
    om = op.samples
    S = op.smaps
    gen = torch.Generator().manual_seed(0)
    idx = torch.randperm(M, generator=gen)[:4096].to(DEV)
    yref = ndft_gpu.type2_at(om, S * x, shape, idx) * op.inv_norm_factor      # (C, 4096)
    e_op = _rel(y[:, idx], yref)
    vox = torch.randperm(x.numel(), generator=gen)[:2048].to(DEV)
    acc = ndft_gpu.type1_at(om, y, shape, vox)                                   # (C, 2048)
    zref = torch.sum(torch.conj(S.reshape(C, -1)[:, vox]).to(torch.complex128) * acc, 0)
    zref = zref * op.inv_norm_factor
    e_adj = _rel(z.reshape(-1)[vox], zref)
    print(f"headline op rel L2 {e_op:.2e}, adj rel L2 {e_adj:.2e}")
    assert e_op < 1e-5 and e_adj < 1e-5
    # the reference's elementwise bar (atol = rtol = 1e-4, relative to the data scale)
    yr = yref.cpu().numpy()
    np.testing.assert_allclose(y[:, idx].cpu().numpy(), yr, rtol=1e-4,
                               atol=1e-4 * np.abs(yr).max())


def test_headline_adjointness(headline):
    op, x = headline
    C, M = op.n_coils, op.n_samples
    gen = torch.Generator(device=DEV).manual_seed(3)
    yy = torch.complex(torch.randn((1, C, M), generator=gen, device=DEV),
                       torch.randn((1, C, M), generator=gen, device=DEV))
    xr = torch.complex(torch.randn(op.shape, generator=gen, device=DEV),
                       torch.randn(op.shape, generator=gen, device=DEV))
    lhs = torch.vdot(op.op(xr).reshape(-1).to(torch.complex128), yy.reshape(-1).to(torch.complex128))
    rhs = torch.vdot(xr.reshape(-1).to(torch.complex128),
                     op.adj_op(yy).reshape(-1).to(torch.complex128))
    err = float(abs(lhs - rhs) / abs(lhs))
    print(f"headline adjointness {err:.2e}")
    assert err < 1e-5


def test_headline_cg_graph_equals_eager(headline):
    op, x = headline
    y = op.op(x)
    L = op.get_lipschitz_cst(seed=0)
    op.get_lipschitz_cst = lambda **_: L
    try:
        eager = op.pinv_solver(y, optim="cg", max_iter=20, graph=False)
        for _ in range(2):
            g = op.pinv_solver(y, optim="cg", max_iter=20, graph=True)
            torch.cuda.synchronize()
            assert _rel(g, eager) < 1e-6
        # the reconstruction explains the data: the normal-equation residual fell
        r0 = torch.linalg.vector_norm(op.adj_op(y))
        r20 = torch.linalg.vector_norm(op.data_consistency(g, y))
        assert float(r20) < 0.2 * float(r0)
    finally:
        del op.get_lipschitz_cst


@pytest.mark.parametrize("density", [False, True])
def test_reduced_3d_cg_vs_oracle_cg(density):
    """20-iteration CG on a 48^3 phyllotaxis problem: HIP (graph-captured,
    FP64 dots) vs the oracle CG on the exact FP64 NDFT operator, same L."""
    from hipnufft import MRIHipNUFFT
    from hipnufft.trajectories import (ellipsoid_phantom, gaussian_smaps,
                                       initialize_3D_phyllotaxis_radial)

    shape, C = (48, 48, 48), 4
    traj = initialize_3D_phyllotaxis_radial(1024, 128).reshape(-1, 3).astype(np.float32)
    smaps = gaussian_smaps(C, shape, seed=2)
    dens = None
    if density:
        # ramp-like weights 1 + 4|k|^2, at least 1: the reference takes its
        # step 1/L from the density-weighted operator and then iterates
        # without density, which converges only when D >= I makes L bound
        # the unweighted operator's constant
        r = np.linalg.norm(traj, axis=-1)
        dens = (1 + 4 * r ** 2).astype(np.float32)
    op = MRIHipNUFFT(traj, shape, n_coils=C, smaps=smaps,
                     density=dens if dens is not None else False)
    xt = torch.from_numpy(ellipsoid_phantom(shape, seed=3)).to(DEV)
    om = op.samples
    ref = ndft_gpu.SenseNdft(om, shape, smaps=torch.from_numpy(smaps).to(DEV), density=dens)
    y = ref.op(xt.cpu().numpy())                                      # exact data (1, C, M)
    L = op.get_lipschitz_cst(seed=4)
    op.get_lipschitz_cst = lambda **_: L
    got = op.pinv_solver(torch.from_numpy(y.astype(np.complex64)).to(DEV), optim="cg",
                         max_iter=20, graph=True)
    want = solvers.cg(ref, y, L, max_iter=20)
    err = rel_l2(got.cpu().numpy(), want)
    print(f"reduced 3D CG (density={density}) rel L2 vs oracle CG {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("shape,C", [((192, 192), 2), ((8, 192, 192), 3), ((384, 384), 2),
                                     ((4, 384, 384), 2)])
def test_192_point_axes_vs_ndft(shape, C):
    """384- and 768-point fine-grid axes (compile-time row and strided FFT
    instances) at small size: op / adj_op against the exact FP64 NDFT and
    adjointness."""
    from hipnufft import MRIHipNUFFT

    d = len(shape)
    traj = truncnorm_traj(40000, d, 21)
    smaps = gauss_smaps(C, shape, seed=22)
    op = MRIHipNUFFT(traj, shape, n_coils=C, smaps=smaps)
    nf = tuple(op.raw_op.plan.info()["nf"])[-d:]
    assert all(n == 2 * s for n, s in zip(nf, shape) if s in (192, 384)), nf
    x = torch.from_numpy(rand_cplx(shape, 23)).to(DEV)
    y = op.op(x).reshape(C, -1)
    yy = torch.from_numpy(rand_cplx((1, C, len(traj)), 24)).to(DEV)
    z = op.adj_op(yy).reshape(shape)
    om = op.samples
    S = op.smaps
    yref = ndft_gpu.type2(om, S * x, shape) * op.inv_norm_factor
    zref = torch.sum(torch.conj(S).to(torch.complex128)
                     * ndft_gpu.type1(om, yy.reshape(C, -1), shape), 0) * op.inv_norm_factor
    assert _rel(y, yref) < 1e-5
    assert _rel(z, zref) < 1e-5
    lhs = torch.vdot(y.reshape(-1).to(torch.complex128), yy.reshape(-1).to(torch.complex128))
    rhs = torch.vdot(x.reshape(-1).to(torch.complex128), z.reshape(-1).to(torch.complex128))
    assert float(abs(lhs - rhs) / abs(lhs)) < 1e-5
