"""Coil-map estimation on the GPU against the FP64 oracle (oracle/smaps.py):
the window, normalisation and blur kernels, LSQR against scipy's LSQR on the
exact NDFT, and ``low_frequency`` end to end (ellipse / Hann / rect windows,
blur, mask array)."""

import numpy as np
import pytest
import torch

from helpers import gauss_smaps, rand_cplx, rand_traj, rel_l2, truncnorm_traj
from oracle import smaps as osm

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("kind", ["ellipse", "rect", "hann", "hamming"])
@pytest.mark.parametrize("d", [2, 3])
def test_window_kernel(kind, d):
    from hipnufft import _capi

    traj = rand_traj(5000, d, 1) * 0.5
    thr = (0.1, 0.15, 0.12)[:d]
    w, out = _capi.kspace_window(torch.from_numpy(traj).to(DEV), thr, kind,
                                 ksp=torch.from_numpy(rand_cplx((3, 5000), 2)).to(DEV))
    ref = osm.window(traj, thr, kind)
    w = w.cpu().numpy()
    if kind in ("ellipse", "rect"):
        # away from the boundary (FP32 against FP64 evaluation) the weights are equal
        if kind == "ellipse":
            s = np.sum(traj.astype(np.float64) ** 2 / np.array(thr) ** 2, axis=1)
            far = np.abs(s - 1) > 1e-5
        else:
            far = np.all(np.abs(np.abs(traj) - np.array(thr)) > 1e-6, axis=1)
        np.testing.assert_array_equal(w[far], ref[far])
    else:
        np.testing.assert_allclose(w, ref, atol=2e-6)
    k = rand_cplx((3, 5000), 2)
    np.testing.assert_allclose(out.cpu().numpy(), k * w[None], rtol=1e-6, atol=1e-7)


def test_rss_normalize_kernel():
    from hipnufft import _capi

    m = rand_cplx((5, 12, 10, 6), 3)
    t = torch.from_numpy(m).to(DEV)
    _capi.rss_normalize(t, 1e-10)
    ref = osm.rss_normalize(m, 1e-10)
    assert rel_l2(t.cpu().numpy(), ref) < 1e-6


@pytest.mark.parametrize("shape,sigma", [((20, 24), (1.5, 0.7)), ((10, 12, 8), (1.0, 0.0, 2.0))])
def test_blur_kernel(shape, sigma):
    from hipnufft import _capi

    m = gauss_smaps(3, shape, 4) * (1 + 0.3 * rand_cplx((3, *shape), 5).real)
    t = torch.from_numpy(m.astype(np.complex64)).to(DEV)
    _capi.blur_magnitude(t, sigma)
    assert rel_l2(t.cpu().numpy(), osm.blur(m, sigma)) < 1e-5


def test_lsqr_vs_scipy_on_exact_ndft():
    """10 LSQR iterations of the calibrationless operator (2 coils) from the
    same data: the GPU NUFFT operator vs scipy's LSQR on the exact NDFT."""
    from hipnufft import MRIHipNUFFT

    shape, C = (16, 16), 2
    traj = truncnorm_traj(300, 2, 6)
    y = rand_cplx((C, 300), 7)
    op = MRIHipNUFFT(traj, shape, n_coils=C, squeeze_dims=True)
    x = op.pinv_solver(y, optim="lsqr", max_iter=10)
    ref = osm.lsqr(osm.calibless_operator(traj, shape, C), y.reshape(-1).astype(np.complex128), 10)
    assert rel_l2(np.asarray(x).reshape(-1), ref) < 1e-4


@pytest.mark.parametrize("window,blurr,masked", [("ellipse", 0.0, False), ("hann", 1.0, False),
                                                 ("rect", 0.0, True), ("ellipse", (1.0, 0.5), True)])
def test_low_frequency_vs_oracle(window, blurr, masked):
    from hipnufft import get_smaps

    shape, C = (20, 20), 4
    traj = truncnorm_traj(1500, 2, 8)
    true = gauss_smaps(C, shape, 9)
    from oracle import ndft

    img = np.ones(shape)
    ksp = np.stack([ndft.type2(traj, true[c] * img, shape) / ndft.norm_factor(shape)
                    for c in range(C)]).astype(np.complex64)
    mask = None
    if masked:
        mask = np.zeros(shape)
        mask[2:-2, 3:-3] = 1.0
    got = get_smaps("low_frequency")(traj, shape, ksp, "hipnufft", threshold=0.2,
                                     window_fun=window, blurr_factor=blurr,
                                     mask=mask if masked else False)
    ref = osm.low_frequency(traj, shape, ksp, threshold=0.2, window_fun=window,
                            blurr_factor=blurr, mask=mask)
    got = np.asarray(got)
    assert got.shape == (C, *shape)
    # magnitudes everywhere; complex values where the maps were not masked to
    # zero (there the blur brings back magnitude, and the reference's phase
    # np.angle of a signed zero is 0 or pi: this build takes 0)
    assert rel_l2(np.abs(got), np.abs(ref)) < 1e-4
    keep = mask > 0 if masked else np.ones(shape, bool)
    assert rel_l2(got[:, keep], ref[:, keep]) < 1e-4
    np.testing.assert_allclose(np.linalg.norm(got[:, keep], axis=0), 1.0, rtol=1e-5)


def test_low_frequency_device_tensors_and_mask_true():
    from hipnufft import low_frequency

    shape, C = (16, 16), 2
    traj = truncnorm_traj(800, 2, 10)
    ksp = torch.from_numpy(rand_cplx((C, 800), 11)).to(DEV)
    out = low_frequency(traj, shape, ksp, "hipnufft")
    assert isinstance(out, torch.Tensor) and out.is_cuda and out.shape == (C, *shape)
    try:
        import skimage  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError):
            low_frequency(traj, shape, ksp, "hipnufft", mask=True)
