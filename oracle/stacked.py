"""Stacked 2.5D NUFFT in FP64 numpy -- test infrastructure (the oracle).

Restates ``MRIStackedNUFFTGPU`` (``src/mrinufft/operators/stacked.py``):

* ``fftz`` / ``ifftz``  <- ``_fftz`` / ``_ifftz`` (:467-489): centred unitary
  DFT along the last axis, ``fftshift(fft(ifftshift(x), norm="ortho"))``
* ``op``   <- ``_op_sense_device`` / ``_op_calibless_device`` (:527-558,
  :639-662): z transform, the z_index planes, a 2D type-2 NDFT per plane,
  times ``1 / norm`` with ``norm = sqrt(4 X Y) * sqrt(2)`` (:447-450)
* ``adj``  <- ``_adj_op_sense_device`` / ``_adj_op_calibless_device``
  (:711-740, :767-790): density, 2D type-1 NDFT per plane times ``1 / norm``,
  planes assigned into a zero image at z_index (a repeated index keeps the
  last plane), inverse z transform, conjugate coil combine

The 2D transforms are the exact NDFT of ``oracle/ndft.py``.  Pinned by
``tests/test_oracle_stacked.py``: for even Z the stacked operator equals
the 3D NDFT of the equivalent 3D trajectory (planes at kz = (z - Z//2) / Z),
forward and adjoint; adjointness; the trajectory conversions round-trip.
"""

from __future__ import annotations

import numpy as np

from . import ndft


def fftz(x):
    return np.fft.fftshift(np.fft.fft(np.fft.ifftshift(x, axes=-1), axis=-1, norm="ortho"),
                           axes=-1)


def ifftz(x):
    return np.fft.fftshift(np.fft.ifft(np.fft.ifftshift(x, axes=-1), axis=-1, norm="ortho"),
                           axes=-1)


def norm_factor(shape):
    """``sqrt(4 X Y) * sqrt(2)``: the 2D operator's factor times sqrt(2)."""
    return ndft.norm_factor(shape[:2]) * np.sqrt(2)


def op(samples2d, z_index, image, shape, smaps=None):
    """(B, C, NZ * NS) k-space of image (B, X, Y, Z) with smaps (C, X, Y, Z),
    or image (B, C, X, Y, Z) without."""
    X, Y, Z = shape
    zi = np.asarray(z_index).reshape(-1)
    img = np.asarray(image, dtype=np.complex128)
    if smaps is not None:
        coil = img.reshape(-1, 1, X, Y, Z) * np.asarray(smaps, np.complex128)[None]
    else:
        coil = img.reshape(img.shape[0], -1, X, Y, Z)
    B, C = coil.shape[:2]
    kz = fftz(coil)[..., zi]                       # (B, C, X, Y, NZ)
    NS = len(np.asarray(samples2d).reshape(-1, 2))
    out = np.zeros((B, C, len(zi), NS), np.complex128)
    for b in range(B):
        for c in range(C):
            for i in range(len(zi)):
                out[b, c, i] = ndft.type2(samples2d, kz[b, c, :, :, i], (X, Y))
    return out.reshape(B, C, -1) / norm_factor(shape)


def adj(samples2d, z_index, ksp, shape, smaps=None, density=None, n_batchs=1):
    """Adjoint of ``op`` with the density applied first: (B, 1 or C, X, Y, Z)."""
    X, Y, Z = shape
    zi = np.asarray(z_index).reshape(-1)
    NS = len(np.asarray(samples2d).reshape(-1, 2))
    y = np.asarray(ksp, np.complex128).reshape(n_batchs, -1, len(zi), NS)
    if density is not None:
        y = y * np.asarray(density, np.float64)[None, None, None, :]
    B, C = y.shape[:2]
    imgz = np.zeros((B, C, X, Y, Z), np.complex128)
    for b in range(B):
        for c in range(C):
            for i in range(len(zi)):  # plain assignment: a repeated index keeps the last
                imgz[b, c, :, :, zi[i]] = ndft.type1(samples2d, y[b, c, i], (X, Y))
    imgz /= norm_factor(shape)
    img = ifftz(imgz)
    if smaps is not None:
        return np.sum(img * np.conj(np.asarray(smaps, np.complex128))[None], axis=1,
                      keepdims=True)
    return img
