"""Exact FP64 non-uniform DFT -- test infrastructure (the oracle).

Restates ``src/mrinufft/operators/interfaces/nudft_numpy.py``:

* ``fourier_matrix``      <- ``get_fourier_matrix`` (:12-55)
* ``type2`` / ``type1``   <- ``implicit_type2_ndft`` / ``implicit_type1_ndft`` (:58-81)

Conventions (nudft_numpy.py:32-45): samples are in cycles, normalised into
[-0.5, 0.5) exactly like ``proper_trajectory(..., "unit")``
(``_utils.py:58-100``); image axis ``d`` pairs with sample coordinate ``d``;
the pixel grid along an axis of size ``N`` is ``linspace(-N/2, N/2-1, N)``;
type 2 uses ``exp(-2 i pi k.r)`` and type 1 the conjugate.

Pinned by ``tests/test_oracle_ndft.py``: the known answer
"Cartesian fftfreq samples -> fftn(fftshift(img))" of
``tests/test_ndft.py:58-79``; explicit == implicit (``test_ndft.py:19-55``);
adjointness; and brute-force scalar loops on tiny inputs.
"""

from __future__ import annotations

import numpy as np


def unit_trajectory(samples):
    """Flatten to (M, d) and normalise to [-0.5, 0.5) (``_utils.py:58-100``, "unit")."""
    k = np.asarray(samples, dtype=np.float64).reshape(-1, np.shape(samples)[-1])
    mx = np.max(np.abs(k)) if k.size else 0.0
    if mx - 1e-4 > 0.5:
        k = k / (2 * np.pi)
    if mx >= 0.5:
        k = (k + 0.5) % 1 - 0.5
    return k


def pixel_grid(shape):
    """(d, prod(shape)) pixel coordinates ``linspace(-N/2, N/2-1, N)`` per axis."""
    r = [np.linspace(-s / 2, s / 2 - 1, s) for s in shape]
    g = np.meshgrid(*r, indexing="ij")
    return np.stack(g).reshape(len(shape), -1)


def norm_factor(shape):
    """``sqrt(prod(N) * 2**d)`` (``operators/base.py:837-840``)."""
    return float(np.sqrt(np.prod(shape) * (2 ** len(shape))))


def fourier_matrix(samples, shape, normalize=False):
    """Explicit (M, prod(shape)) complex128 matrix ``exp(-2 i pi k.r)``."""
    k = unit_trajectory(samples)
    mat = np.exp(-2j * np.pi * (k @ pixel_grid(shape)))
    if normalize:
        mat /= norm_factor(shape)
    return mat


def type2(samples, image, shape, normalize=False, chunk=4096):
    """Image -> k-space, ``sum_r x_r exp(-2 i pi k.r)``; chunked over samples."""
    k = unit_trajectory(samples)
    r = pixel_grid(shape)
    x = np.asarray(image, dtype=np.complex128).reshape(-1)
    out = np.empty(len(k), dtype=np.complex128)
    for s in range(0, len(k), chunk):
        out[s : s + chunk] = np.exp(-2j * np.pi * (k[s : s + chunk] @ r)) @ x
    if normalize:
        out /= norm_factor(shape)
    return out


def type1(samples, coeffs, shape, normalize=False, chunk=4096):
    """K-space -> image, ``sum_m c_m exp(+2 i pi k_m.r)``; chunked over samples."""
    k = unit_trajectory(samples)
    r = pixel_grid(shape)
    c = np.asarray(coeffs, dtype=np.complex128).reshape(-1)
    out = np.zeros(r.shape[1], dtype=np.complex128)
    for s in range(0, len(k), chunk):
        out += c[s : s + chunk] @ np.exp(2j * np.pi * (k[s : s + chunk] @ r))
    if normalize:
        out /= norm_factor(shape)
    return out.reshape(shape)


def type2_at(samples, image, shape, rows):
    """Type 2 evaluated only at the sample indices ``rows`` (subsampled oracle)."""
    k = unit_trajectory(samples)[rows]
    return type2(k, image, shape)


def type1_at(samples, coeffs, shape, voxels, chunk=2048):
    """Type 1 evaluated only at the flat voxel indices ``voxels``."""
    k = unit_trajectory(samples)
    r = pixel_grid(shape)[:, voxels]
    c = np.asarray(coeffs, dtype=np.complex128).reshape(-1)
    out = np.zeros(r.shape[1], dtype=np.complex128)
    for s in range(0, len(k), chunk):
        out += c[s : s + chunk] @ np.exp(2j * np.pi * (k[s : s + chunk] @ r))
    return out
