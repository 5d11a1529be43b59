"""Sensitivity-map estimation on the device (``extras/smaps.py``).

``low_frequency`` (``smaps.py:219-306``): window the k-space centre
(``_extract_kspace_center``, ``:83-170``), solve for the coil images with a
calibrationless operator and ``pinv_solver`` (LSQR, ``max_iter``
iterations), optionally mask and blur the magnitudes, and divide by the
root sum of squares over coils.  The windows, the blur and the
normalisation are ``libhipnufft`` kernels; the solve is the hipnufft
operator's.  Masking with ``mask=True`` (Otsu threshold + convex hull) uses
scikit-image, as the reference does, and needs it installed; a mask array
needs nothing else.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _capi
from ._array import from_device, kind_of, to_device
from ._registry import MethodRegister

register_smaps = MethodRegister("smaps")
get_smaps = register_smaps.make_getter()


def _extract_kspace_center(kspace_data, kspace_loc, threshold=None, density=None,
                           window_fun="ellipse"):
    """(windowed data, locations, density) of the k-space centre (``smaps.py:83-170``).

    ``"rect"`` keeps only the samples inside the box (fewer samples);
    ``"ellipse"`` (default), ``"hann"`` / ``"hanning"``, ``"hamming"`` and a
    callable ``f(locations) -> weights`` weigh every sample.  Device tensors
    in and out.
    """
    d = kspace_loc.shape[1]
    if isinstance(threshold, (float, int)):
        threshold = (float(threshold),) * d
    threshold = tuple(float(t) for t in np.asarray(threshold).reshape(-1))
    if window_fun == "rect":
        w, _ = _capi.kspace_window(kspace_loc, threshold, "rect")
        sel = w > 0
        dc = density[sel] if density is not None else None
        return kspace_data[..., sel].contiguous(), kspace_loc[sel].contiguous(), dc
    if callable(window_fun):
        window = to_device(window_fun(kspace_loc), torch.float32, kspace_data.device)
        return kspace_data * window, kspace_loc, density
    if window_fun not in ("hann", "hanning", "hamming", "ellipse"):
        raise ValueError("Unsupported window function.")
    _, out = _capi.kspace_window(kspace_loc, threshold, window_fun, ksp=kspace_data,
                                 weights=False)
    return out.reshape(kspace_data.shape), kspace_loc, density


@register_smaps
def low_frequency(traj, shape, kspace_data, backend="hipnufft", threshold=0.1, density=None,
                  max_iter=10, window_fun="ellipse", blurr_factor=0.0, mask=False):
    """Low-frequency coil maps (``smaps.py:219-306``): (C, *shape), unit root
    sum of squares over coils.

    traj (..., d) in [-0.5, 0.5) (flattened); kspace_data (C, M); threshold
    per axis or scalar; density is passed through the window selection, as
    in the reference (the solve itself runs without it); blurr_factor: Gaussian
    sigma in cells per axis (0: none); mask: False, True (Otsu + convex hull,
    needs scikit-image) or an array multiplied into the maps.
    """
    from .base import get_operator

    kind = kind_of(kspace_data)
    shape = tuple(int(s) for s in shape)
    dev = torch.device("cuda", 0)
    if isinstance(kspace_data, torch.Tensor) and kspace_data.is_cuda:
        dev = kspace_data.device
    loc = to_device(np.asarray(traj, dtype=np.float32) if not isinstance(traj, torch.Tensor)
                    else traj, torch.float32, dev)
    loc = loc.reshape(-1, loc.shape[-1])
    ksp = to_device(kspace_data, torch.complex64, dev)
    ksp = ksp.reshape(-1, loc.shape[0])
    dens = None if density is None else to_device(density, torch.float32, dev)
    with torch.cuda.device(dev):
        k_space, samples, _ = _extract_kspace_center(ksp, loc, threshold, dens, window_fun)
        C = k_space.shape[-2]
        op = get_operator(backend)(samples, shape, n_coils=C, squeeze_dims=True,
                                   gpu_device_id=dev.index or 0)
        maps = to_device(op.pinv_solver(k_space, max_iter=max_iter), torch.complex64, dev)
        maps = maps.reshape(C, *shape).contiguous()
        masked = False
        if isinstance(mask, (np.ndarray, torch.Tensor)):
            m = to_device(mask, torch.complex64, dev)
            maps = (maps * m).contiguous()
            masked = bool(torch.any(m != 0))
        elif isinstance(mask, bool) and mask:
            try:
                from skimage.filters import threshold_otsu
                from skimage.morphology import convex_hull_image
            except ImportError as err:
                raise ImportError(
                    "The scikit-image module is not available. Please install it along with "
                    "the [extra] dependencies or using `pip install scikit-image`.") from err
            sos = torch.linalg.vector_norm(maps, dim=0).cpu().numpy()
            hull = convex_hull_image(sos > threshold_otsu(sos))
            maps = (maps * torch.from_numpy(hull).to(dev)).contiguous()
            masked = True
        blur = blurr_factor
        if isinstance(blur, (int, float)):
            blur = (float(blur),) * len(shape)
        blur = tuple(float(b) for b in blur)
        blurred = sum(blur) > 0
        if blurred:
            _capi.blur_magnitude(maps, blur)
        # the reference adds 1e-10 to the root sum of squares only after masking / blurring
        _capi.rss_normalize(maps, 1e-10 if (masked or blurred) else 0.0)
    return from_device(maps, kind)
