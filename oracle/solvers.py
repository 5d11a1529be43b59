"""Solvers restated in numpy -- test infrastructure.

* ``power_method`` <- ``operators/base.py:1155-1221`` (norm of A^H A x,
  stop when the norm changes by less than 1e-6, at most max_iter steps);
* ``cg``           <- ``extras/optim.py:801-902`` (Lipschitz step, density
  warm start ``_scaled_dcp`` :151-165, first step ``tol*v + g/L``, Polak-
  Ribiere beta with the non-conjugated ``numpy.dot`` and the lexicographic
  complex ``max(0, beta)``, stop at ``||g|| <= tol``);
* ``pipe``         <- ``operators/interfaces/cufinufft.py:1281-1331`` (fixed
  point ``d <- d / |A_s A_s^H d|`` with spread/interp-only operators on the
  image grid, then ``d /= mean|A^H(d A 1)|``);
* ``lsmr``         <- ``scipy.sparse.linalg.lsmr`` (Fong & Saunders), the
  published implementation the reference's ``lsmr`` (``extras/optim.py:498-798``)
  is adapted from, called directly, with the reference's defaults;
* ``sense_linop``  the exact FP64 SENSE operator (``oracle/ndft.py`` matrix,
  normalised, density-free) of one batch as a scipy ``LinearOperator``, for
  the scipy solvers.

Pinned by ``tests/test_oracle_solvers.py``: power method against the dense
spectral norm; CG against the least-squares solution on a tiny exact problem
(monotone residual, convergence) and against a hand-unrolled first step;
Pipe on a Cartesian grid (uniform), on radial trajectories by its fixed
point (|A_s A_s^H d| -> 1 with independently built spread / interp-only
operators), by the normalised PSF energy (mean|A^H(d A 1)| = 1 with the
exact NDFT operator) and by the reference's correlation bars against |k|;
LSMR converging to the (damped) least-squares solution; ``sense_linop``
against the dense SENSE matrix.
"""

from __future__ import annotations

import numpy as np
from scipy.sparse.linalg import LinearOperator
from scipy.sparse.linalg import lsmr as _lsmr

from . import ndft, nufft
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


def lsmr(A, b, max_iter, damp=0.0, atol=1e-6, btol=1e-6, conlim=1e8):
    """LSMR iterates of ``min ||A x - b||^2 + damp^2 ||x||^2`` from x = 0
    (the reference's defaults: atol = btol = 1e-6, conlim 1e8)."""
    return _lsmr(A, b, damp=damp, atol=atol, btol=btol, conlim=conlim, maxiter=max_iter)[0]


def sense_linop(samples, shape, smaps=None, n_coils=1):
    """Exact normalised SENSE (or calibrationless, ``smaps=None``) operator of
    one batch on flattened vectors: image (N,) or (C N,) -> k-space (C M,)."""
    F = ndft.fourier_matrix(samples, shape, normalize=True)
    M, N = F.shape
    if smaps is None:
        C = n_coils
        return LinearOperator(
            (C * M, C * N), dtype=np.complex128,
            matvec=lambda x: (F @ x.reshape(C, N).T).T.reshape(-1),
            rmatvec=lambda y: (F.conj().T @ y.reshape(C, M).T).T.reshape(-1))
    S = np.asarray(smaps, dtype=np.complex128).reshape(-1, N)
    C = S.shape[0]
    return LinearOperator(
        (C * M, N), dtype=np.complex128,
        matvec=lambda x: (F @ (S * x.reshape(1, N)).T).T.reshape(-1),
        rmatvec=lambda y: np.sum(np.conj(S) * (F.conj().T @ y.reshape(C, M).T).T, 0))
