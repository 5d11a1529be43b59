"""Solvers restated in numpy -- test infrastructure.

* ``power_method`` <- ``operators/base.py:1155-1221`` (norm of A^H A x,
  stop when the norm changes by less than 1e-6, at most max_iter steps);
* ``cg``           <- ``extras/optim.py:801-902`` (Lipschitz step, density
  warm start ``_scaled_dcp`` :151-165, first step ``tol*v + g/L``, Polak-
  Ribiere beta with the non-conjugated ``numpy.dot`` and the lexicographic
  complex ``max(0, beta)``, stop at ``||g|| <= tol``);
* ``pipe``         <- ``operators/interfaces/cufinufft.py:1281-1331`` (fixed
  point ``d <- d / |A_s A_s^H d|`` with spread/interp-only operators on the
  image grid, then ``d /= mean|A^H(d A 1)|``).

Pinned by ``tests/test_oracle_solvers.py``: power method against the dense
spectral norm; CG against scipy's least-squares solution on a tiny exact
problem (monotone residual, convergence) and against a hand-unrolled first
step; Pipe against its fixed-point property on a Cartesian grid, where the
exact density is uniform.
"""

from __future__ import annotations

import numpy as np

from . import nufft
from .mri import Operator


def power_method(max_iter, op, x):
    """Largest eigenvalue of A^H A from the start vector ``x`` (single coil)."""
    x = np.array(x, dtype=np.complex128)
    x_norm = np.linalg.norm(x)
    x = x / x_norm
    for _ in range(max_iter):
        x_new = op.adj_op(op.op(x)).reshape(x.shape)
        x_new_norm = np.linalg.norm(x_new)
        x_new = x_new / x_new_norm
        if abs(x_norm - x_new_norm) < 1e-6:
            break
        x_norm = x_new_norm
        x = x_new
    return x_new_norm


def cg(op, y, L, max_iter=10, tol=1e-4, x_init=None):
    """Reference CG iterations with a given Lipschitz constant ``L``."""
    shape = (op.n_batchs, 1 if op.uses_sense else op.n_coils) + op.shape
    density = op.density
    if density is not None:
        if x_init is None:
            x0 = op.adj_op(y)
            x_init = x0 * np.linalg.norm(y) / np.linalg.norm(op.op(x0))
        op.density = None
    image = np.zeros(shape, dtype=np.complex128) if x_init is None else np.array(x_init).reshape(shape)
    velocity = np.zeros_like(image)
    grad = op.data_consistency(image, y).reshape(shape)
    velocity = tol * velocity + grad / L
    image = image - velocity
    try:
        for _ in range(max_iter):
            grad_new = op.data_consistency(image, y).reshape(shape)
            if np.linalg.norm(grad_new) <= tol:
                break
            beta = np.dot(grad_new.ravel(), (grad_new - grad).ravel()) / np.dot(grad.ravel(),
                                                                              grad.ravel())
            beta = max(0, beta)
            velocity = grad_new + beta * velocity
            image = image - velocity / L
            grad = grad_new
    finally:
        op.density = density
    return image


def pipe(samples, shape, max_iter=10, osf=2.0, normalize=True, eps=1e-6):
    """Pipe density with spread/interp-only operators (oracle transforms)."""
    shape = tuple(shape)
    s = np.asarray(samples).reshape(-1, len(shape))
    if np.max(np.abs(s)) - 1e-4 < 0.5:
        s = (s.astype(np.float32) * np.float32(2) * np.float32(np.pi)).astype(np.float32)
    grid = nufft.Plan(s, shape, eps=eps, upsampfac=osf, spreadinterponly=True)
    norm = float(np.sqrt(np.prod(shape) * 2 ** len(shape)))
    d = np.ones(len(s), dtype=np.complex128)
    for _ in range(max_iter):
        d = d / np.abs(grid.execute(grid.execute_adjoint(d) / norm) / norm)
    if normalize:
        op = Operator(s, shape, eps=eps)
        rec = op.adj_op(d * op.op(np.ones(shape)).ravel())
        d = d / np.mean(np.abs(rec))
    return np.abs(d)
