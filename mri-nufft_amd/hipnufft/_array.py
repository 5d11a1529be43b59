"""Array boundary: numpy / torch-cpu / torch-cuda in, the same kind out.

Restates the caller-type round trip of ``with_numpy_cupy``
(``src/mrinufft/_array_compat.py:183-222``) for a torch-native backend, and
``proper_trajectory`` (``src/mrinufft/_utils.py:58-100``).
"""

from __future__ import annotations

import logging
import math

import numpy as np
import torch

logger = logging.getLogger("hipnufft")

NP2TORCH = {
    np.dtype("float32"): torch.float32,
    np.dtype("float64"): torch.float64,
    np.dtype("complex64"): torch.complex64,
    np.dtype("complex128"): torch.complex128,
}


def kind_of(x):
    """'numpy', 'torch-cpu' or 'torch-cuda' (anything else array-like: 'numpy')."""
    if isinstance(x, torch.Tensor):
        return "torch-cuda" if x.is_cuda else "torch-cpu"
    return "numpy"


def to_device(x, dtype, device):
    """Contiguous tensor of ``dtype`` on ``device`` (zero copy when possible)."""
    if isinstance(x, torch.Tensor):
        # lazy conj / neg views keep the original memory: materialise them,
        # the C ABI reads raw device memory
        t = x.detach().resolve_conj().resolve_neg()
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x)))
    if t.dtype != dtype and not t.is_complex() and dtype in (torch.complex64, torch.complex128):
        t = t.to(dtype)
    return t.to(device=device, dtype=dtype, non_blocking=False).contiguous()


def from_device(t, kind):
    """Return ``t`` as the caller's array kind."""
    if kind == "torch-cuda":
        return t
    if kind == "torch-cpu":
        return t.cpu()
    return t.cpu().numpy()


def proper_trajectory(trajectory, normalize="pi"):
    """Flatten to (M, d) and rescale like ``_utils.proper_trajectory``.

    ``"pi"``: multiply by 2 pi when max|k| - 1e-4 < 0.5 (samples in
    [-0.5, 0.5) become radians).  ``"unit"``: divide by 2 pi when
    max|k| - 1e-4 > 0.5, then wrap into [-0.5, 0.5) when max|k| >= 0.5.
    Works on numpy arrays and torch tensors (kept on their device).
    """
    if isinstance(trajectory, torch.Tensor):
        new = trajectory.reshape(-1, trajectory.shape[-1])
        mx = float(new.abs().max()) if new.numel() else 0.0
    else:
        new = np.asarray(trajectory)
        new = new.reshape(-1, new.shape[-1])
        mx = float(np.max(np.abs(new))) if new.size else 0.0
    if normalize == "pi" and mx - 1e-4 < 0.5:
        logger.warning("Samples will be rescaled to [-pi, pi), assuming they were in [-0.5, 0.5)")
        new = new * 2 * math.pi
    elif normalize == "unit" and mx - 1e-4 > 0.5:
        logger.warning("Samples will be rescaled to [-0.5, 0.5), assuming they were in [-pi, pi)")
        new = new * 1 / (2 * math.pi)
    if normalize == "unit" and mx >= 0.5:
        new = (new + 0.5) % 1 - 0.5
    return new
