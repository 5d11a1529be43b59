"""Low-frequency coil-map estimation in FP64 -- test infrastructure (the oracle).

Restates ``src/mrinufft/extras/smaps.py``:

* ``window``  <- ``_extract_kspace_center`` (:83-170): ellipse / rect /
  Hann / Hamming weights of the samples
* ``blur``    <- ``skimage.filters.gaussian(|maps|, sigma=(0,) + s) *
  exp(i angle(maps))`` (:293-298).  scikit-image is not installed here; its
  ``gaussian`` is ``scipy.ndimage.gaussian_filter`` with ``mode="nearest"``,
  ``truncate=4.0`` for float input, which is called directly
* ``lsqr``    <- ``scipy.sparse.linalg.lsqr``, the published implementation the
  reference's ``lsqr`` (``extras/optim.py:249-495``) states it is adapted from
* ``low_frequency`` <- :219-306 with the exact NDFT operator of
  ``oracle/ndft.py`` (calibrationless, no density, normalised)

Pinned by ``tests/test_oracle_smaps.py``: window values at hand-picked
points (closed forms), the blur of constants and of a delta (the truncated,
normalised Gaussian), LSQR converging to ``numpy.linalg.lstsq`` and
reproducing the first iterate's closed form, unit root sum of squares.
"""

from __future__ import annotations

import numpy as np
import scipy.ndimage as ndi
from scipy.sparse.linalg import LinearOperator
from scipy.sparse.linalg import lsqr as _lsqr

from . import ndft


def window(traj, threshold, kind):
    """Weights of the samples traj (M, d) (cycles) for the window ``kind``."""
    k = np.asarray(traj, dtype=np.float64).reshape(-1, np.shape(traj)[-1])
    d = k.shape[1]
    thr = np.broadcast_to(np.asarray(threshold, dtype=np.float64), (d,))
    if kind == "rect":
        return (np.sum(np.abs(k) - thr <= 0, axis=1) == d).astype(np.float64)
    if kind == "ellipse":
        return (np.sum(k ** 2 / thr ** 2, axis=1) <= 1).astype(np.float64)
    a0 = 0.5 if kind in ("hann", "hanning") else 0.53836
    return a0 + (1 - a0) * np.cos(np.pi * np.linalg.norm(k, axis=1) / thr[0])


def blur(maps, sigma):
    m = np.asarray(maps, dtype=np.complex128)
    mag = ndi.gaussian_filter(np.abs(m), sigma=(0,) + tuple(sigma), mode="nearest", truncate=4.0)
    return mag * np.exp(1j * np.angle(m))


def rss_normalize(maps, eps=0.0):
    m = np.asarray(maps, dtype=np.complex128)
    return m / (np.linalg.norm(m, axis=0) + eps)


def lsqr(A, b, max_iter):
    """LSQR iterates of ``min ||A x - b||`` from x = 0 (damp 0, atol = btol = 1e-6,
    conlim 1e8: the reference's defaults)."""
    return _lsqr(A, b, damp=0.0, atol=1e-6, btol=1e-6, conlim=1e8, iter_lim=max_iter)[0]


def calibless_operator(samples, shape, C):
    """The normalised type-2 NDFT applied to each of C coil images, as a
    block-diagonal LinearOperator on flattened (C * N) vectors."""
    F = ndft.fourier_matrix(samples, shape, normalize=True)
    M, N = F.shape
    return LinearOperator(
        (C * M, C * N), dtype=np.complex128,
        matvec=lambda x: (F @ x.reshape(C, N).T).T.reshape(-1),
        rmatvec=lambda y: (F.conj().T @ y.reshape(C, M).T).T.reshape(-1))


def low_frequency(traj, shape, kspace, threshold=0.1, max_iter=10, window_fun="ellipse",
                  blurr_factor=0.0, mask=None):
    k = np.asarray(traj, dtype=np.float64).reshape(-1, len(shape))
    y = np.asarray(kspace, dtype=np.complex128).reshape(-1, k.shape[0])
    C = y.shape[0]
    w = window(k, threshold, window_fun)
    if window_fun == "rect":
        sel = w > 0
        k, y = k[sel], y[:, sel]
    else:
        y = y * w[None]
    x = lsqr(calibless_operator(k, shape, C), y.reshape(-1), max_iter).reshape(C, *shape)
    eps = 0.0
    if mask is not None:
        x = x * mask
        eps = 1e-10 if np.any(mask) else 0.0
    bf = (blurr_factor,) * len(shape) if np.isscalar(blurr_factor) else tuple(blurr_factor)
    if sum(bf) > 0:
        x = blur(x, bf)
        eps = 1e-10
    return rss_normalize(x, eps)
