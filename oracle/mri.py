"""MRI operator semantics over the oracle transforms -- test infrastructure.

Restates ``FourierOperatorSimple`` (``src/mrinufft/operators/base.py:880-1139``)
and the cufinufft device paths (``operators/interfaces/cufinufft.py:478-1178``):

* ``op``: ``y[b, c] = F(S_c x_b) / norm`` (SENSE) or ``F(x[b, c]) / norm``;
* ``adj_op``: ``x_b = sum_c conj(S_c) F^H(d * y[b, c]) / norm`` (density
  applied once, before the type 1, ``base.py:1068-1073``);
* ``data_consistency``: ``A^H D (A x - y)`` per coil (``base.py:1094-1139``);
* ``norm = sqrt(prod(N) 2^d)`` (``base.py:837-840``).

The raw transform is either the exact FP64 NDFT (``exact=True``) or the FP64
ES-NUFFT of ``oracle.nufft``.  Pinned by ``tests/test_oracle_solvers.py``
(``test_nufft_operator_matches_exact``: NDFT vs NUFFT agreement and
adjointness; ``test_dc_is_adj_of_residual_with_density_once``: DC ==
adj(op(x) - y) with the density applied once, against the explicit form).
"""

from __future__ import annotations

import numpy as np

from . import ndft, nufft


class Operator:
    def __init__(self, samples, shape, n_coils=1, n_batchs=1, smaps=None, density=None,
                 exact=False, eps=1e-6):
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(self.shape)
        s = np.asarray(samples).reshape(-1, self.ndim)
        if np.max(np.abs(s)) - 1e-4 < 0.5:  # unit -> radians like proper_trajectory("pi")
            s = (s.astype(np.float32) * np.float32(2) * np.float32(np.pi)).astype(np.float32)
        self.omega = s.astype(np.float32)
        self.unit = self.omega.astype(np.float64) / (2 * np.pi)
        self.n_samples = len(self.omega)
        self.smaps = None if smaps is None else np.asarray(smaps, dtype=np.complex128)
        self.n_coils = n_coils if smaps is None else self.smaps.shape[0]
        self.n_batchs = n_batchs
        self.density = None if density is None else np.asarray(density, dtype=np.float64)
        self.exact = exact
        if not exact:
            self.plan = nufft.Plan(self.omega, self.shape, eps=eps)
        self.norm_factor = float(np.sqrt(np.prod(self.shape) * 2 ** self.ndim))

    @property
    def uses_sense(self):
        return self.smaps is not None

    def _t2(self, x):
        if self.exact:
            return ndft.type2(self.unit, x, self.shape)
        return self.plan.execute(x)

    def _t1(self, y):
        if self.exact:
            return ndft.type1(self.unit, y, self.shape)
        return self.plan.execute_adjoint(y)

    def op(self, x):
        B, C = self.n_batchs, self.n_coils
        if self.uses_sense:
            x = np.asarray(x).reshape(B, *self.shape)
            out = [[self._t2(x[b] * self.smaps[c]) for c in range(C)] for b in range(B)]
        else:
            x = np.asarray(x).reshape(B, C, *self.shape)
            out = [[self._t2(x[b, c]) for c in range(C)] for b in range(B)]
        return np.asarray(out) / self.norm_factor

    def adj_op(self, y):
        B, C = self.n_batchs, self.n_coils
        y = np.asarray(y).reshape(B, C, self.n_samples)
        d = 1.0 if self.density is None else self.density
        if self.uses_sense:
            out = [sum(np.conj(self.smaps[c]) * self._t1(d * y[b, c]) for c in range(C))[None]
                   for b in range(B)]
        else:
            out = [[self._t1(d * y[b, c]) for c in range(C)] for b in range(B)]
        return np.asarray(out) / self.norm_factor

    def data_consistency(self, x, y):
        return self.adj_op(self.op(x) - np.asarray(y).reshape(self.n_batchs, self.n_coils, -1))
