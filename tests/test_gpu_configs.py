"""Parity at the exact sizes of the benchmark's secondary configurations.

* PR1 2D radial (100 spokes x 500 samples, 256^2, one coil, eps 1e-6):
  op and adj_op against the exact FP64 NDFT over every sample / pixel.
* 2D spiral (64 x 2048, 320^2, 32 coils, Pipe density, autotuned plan, all
  coils in one launch per stage): op / adj_op against the FP64 NDFT on
  random subsets, adjointness, and one launch per stage.
* Differentiable 2D radial (128 x 1024 + N(0, 1e-3) jitter, 256^2, 8 coils):
  image and trajectory gradients of an L2 loss against central finite
  differences (h = 1e-3) and the image gradient against the FP64 adjoint.

The FP64 NDFT is evaluated on the GPU in separable form (per-axis phase
matrices), which is exact and fast at these sizes.
"""

import numpy as np
import pytest
import torch

from helpers import gauss_smaps, rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _phases(om, shape, sign):
    """Per-axis exp(sign * i * omega_d * r_d) (complex128), r = i - N/2."""
    return [torch.exp(sign * 1j * om[:, d:d + 1].double()
                      * (torch.arange(n, device=om.device, dtype=torch.float64) - n // 2)[None, :])
            for d, n in enumerate(shape)]


def ndft2_type2(om, x, chunk=8192):
    """y_m = sum_n x_n exp(-i omega_m . r_n), 2D, x (B, N0, N1) or (N0, N1) -> (B, M) / (M,).
    Separable: (x @ E1^T) then a row-wise dot with E0, in chunks of samples."""
    xb = x.to(torch.complex128).reshape(-1, *x.shape[-2:])
    out = []
    for s in range(0, om.shape[0], chunk):
        E0, E1 = _phases(om[s:s + chunk], x.shape[-2:], -1)   # (m, N0), (m, N1)
        t = torch.matmul(xb, E1.T.contiguous())                # (B, N0, m)
        out.append(torch.sum(t * E0.T.contiguous()[None], 1))  # (B, m)
    y = torch.cat(out, -1)
    return y if x.dim() > 2 else y[0]


def ndft2_type1(om, y, shape):
    """x_n = sum_m y_m exp(+i omega_m . r_n), 2D, y (B, M) or (M,) -> (B, N0, N1) / (N0, N1).
    Separable: (E0 * y)^T @ E1 per batch row (no (M, N0, N1) intermediate)."""
    E0, E1 = _phases(om, shape, 1)                            # (M, N0), (M, N1)
    yb = y.to(torch.complex128).reshape(-1, om.shape[0])
    out = torch.stack([torch.matmul((E0 * yb[b, :, None]).T.contiguous(), E1)
                       for b in range(yb.shape[0])])
    return out if y.dim() > 1 else out[0]


def test_config_pr1_radial_vs_ndft():
    from hipnufft import MRIHipNUFFT
    from hipnufft.trajectories import initialize_2D_radial

    traj = initialize_2D_radial(100, 500).reshape(-1, 2).astype(np.float32)
    shape = (256, 256)
    rng = np.random.default_rng(0)
    img = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    op = MRIHipNUFFT(traj, shape, n_coils=1, eps=1e-6, squeeze_dims=True)
    x = torch.from_numpy(img).to(DEV)
    y = op.op(x)
    z = op.adj_op(y)
    om = op.samples
    yref = ndft2_type2(om, x) * op.inv_norm_factor
    zref = ndft2_type1(om, y, shape) * op.inv_norm_factor
    assert rel_l2(y.cpu().numpy(), yref.cpu().numpy()) < 1e-5
    assert rel_l2(z.cpu().numpy(), zref.cpu().numpy()) < 1e-5


def test_config_spiral_32coils_pipe_autotuned():
    from hipnufft import MRIHipNUFFT
    from hipnufft.trajectories import initialize_2D_spiral

    traj = initialize_2D_spiral(64, 2048).reshape(-1, 2).astype(np.float32)
    shape = (320, 320)
    C = 32
    smaps = gauss_smaps(C, shape, seed=5)
    op = MRIHipNUFFT(traj, shape, n_coils=C, smaps=smaps, density="pipe", autotune=True)
    assert op.tuning is not None and op.n_trans == C
    d = op.density
    assert torch.all(torch.isfinite(d)) and torch.all(d > 0)
    rng = np.random.default_rng(6)
    img = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    x = torch.from_numpy(img).to(DEV)
    plan = op.raw_op.plan
    plan.set_timing(True)
    y = op.op(x).reshape(C, -1)
    z = op.adj_op(y.reshape(1, C, -1)).reshape(shape)
    t = plan.get_timing()
    plan.set_timing(False)
    # every stage ran once for all 32 coils (one launch per stage and direction)
    assert t["interp"][1] == 1 and t["spread"][1] == 1 and t["rows_adj"][1] == 1
    om = op.samples
    S = torch.from_numpy(smaps).to(DEV)
    idx = torch.from_numpy(rng.choice(om.shape[0], 512, replace=False)).to(DEV)
    yref = ndft2_type2(om[idx], S * x) * op.inv_norm_factor          # (C, 512)
    assert rel_l2(y[:, idx].cpu().numpy(), yref.cpu().numpy()) < 1e-5
    # adjoint with density, at every pixel, from the FP64 type 1 per coil
    zref = torch.zeros(shape, dtype=torch.complex128, device=DEV)
    for c in range(0, C, 8):
        zc = ndft2_type1(om, y[c:c + 8] * d[None, :], shape)
        zref += torch.sum(torch.conj(S[c:c + 8]).to(torch.complex128) * zc, 0)
    zref *= op.inv_norm_factor
    assert rel_l2(z.cpu().numpy(), zref.cpu().numpy()) < 1e-5
    # adjointness <A x, y> = <x, A^H y> (density off)
    op.density = None
    yy = torch.from_numpy((rng.standard_normal((C, om.shape[0]))
                           + 1j * rng.standard_normal((C, om.shape[0]))).astype(np.complex64)).to(DEV)
    lhs = torch.vdot(op.op(x).reshape(-1).to(torch.complex128), yy.reshape(-1).to(torch.complex128))
    rhs = torch.vdot(x.reshape(-1).to(torch.complex128),
                     op.adj_op(yy.reshape(1, C, -1)).reshape(-1).to(torch.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_config_differentiable_radial_fd():
    from hipnufft import MRIHipNUFFT
    from hipnufft.trajectories import initialize_2D_radial

    rng = np.random.default_rng(7)
    traj = initialize_2D_radial(128, 1024).reshape(-1, 2)
    traj = (traj + rng.normal(0, 1e-3, traj.shape)).astype(np.float32)
    traj = np.clip(traj, -0.5, 0.5 - 1e-6).astype(np.float32)
    shape = (256, 256)
    C = 8
    smaps = gauss_smaps(C, shape, seed=8)
    op = MRIHipNUFFT(traj, shape, n_coils=C, smaps=smaps)
    ad = op.make_autograd(wrt_data=True, wrt_traj=True)
    img = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    x = torch.from_numpy(img).to(DEV).reshape(1, 1, *shape).requires_grad_(True)
    tgt = torch.from_numpy((rng.standard_normal((1, C, traj.shape[0]))
                            + 1j * rng.standard_normal((1, C, traj.shape[0]))).astype(np.complex64)).to(DEV)
    y = ad.op(x)
    loss = torch.sum(torch.abs(y - tgt) ** 2)
    gx, gt = torch.autograd.grad(loss, [x, ad.samples])
    # image gradient: 2 A^H (A x - t) (no density), in FP64 from the same residual
    resid = (y - tgt).detach().reshape(C, -1)
    om = op.samples.detach()
    S = torch.from_numpy(smaps).to(DEV).to(torch.complex128)
    gref = 2 * torch.sum(torch.conj(S) * ndft2_type1(om, resid, shape), 0) * op.inv_norm_factor
    assert rel_l2(gx.reshape(shape).cpu().numpy(), gref.cpu().numpy()) < 1e-4
    # trajectory gradient (radians): against the exact FP64 derivative of the
    # loss for each sampled coordinate, and against central differences
    # (h = 1e-3; only the perturbed sample's C values change, so the
    # difference is taken there).  Central differences carry an O(h^2 r^3)
    # truncation error (r up to 128 pixels here), about 1e-2 relative, so
    # they are held to 3e-2 while the gradient is held to the exact value.
    base = om.clone()
    h = 1e-3
    ms = rng.integers(0, traj.shape[0], 8)
    ds = rng.integers(0, 2, 8)
    xd = x.detach()
    xs = (S * xd.reshape(1, *shape).to(torch.complex128))            # (C, N0, N1)
    r = [torch.arange(n, device=DEV, dtype=torch.float64) - n // 2 for n in shape]
    gmax = float(torch.abs(gt).max())
    for m, d in zip(ms, ds):
        E0, E1 = _phases(base[m:m + 1], shape, -1)                      # (1, N0), (1, N1)
        ph = (E0[0][:, None] * E1[0][None, :])                           # (N0, N1)
        rd = r[0][:, None] if d == 0 else r[1][None, :]
        ym = torch.sum(xs * ph, (1, 2)) * op.inv_norm_factor
        dym = torch.sum(xs * ph * (-1j * rd), (1, 2)) * op.inv_norm_factor
        exact = float(torch.sum(2 * torch.real(torch.conj(ym - tgt[0, :, m].to(torch.complex128))
                                               * dym)))
        g = float(gt[m, d])
        assert abs(g - exact) <= 1e-3 * max(abs(exact), 1e-2 * gmax), (m, d, g, exact)
        vals = []
        for sgn in (h, -h):
            w = base.clone()
            w[m, d] += sgn
            op2 = MRIHipNUFFT(w, shape, n_coils=C, smaps=smaps)
            yv = op2.op(xd).reshape(C, -1)[:, m].to(torch.complex128)
            vals.append(float(torch.sum(torch.abs(yv - tgt[0, :, m].to(torch.complex128)) ** 2)))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g) <= 3e-2 * max(abs(fd), 1e-2 * gmax), (m, d, fd, g)
