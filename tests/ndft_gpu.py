"""Exact FP64 non-uniform DFT evaluated on the GPU in separable form -- a
test-side checker for problems too large for the numpy oracle.

The transform is the one of ``oracle/ndft.py`` (restating
``src/mrinufft/operators/interfaces/nudft_numpy.py:12-81``): coordinate d
(radians) pairs with image axis d, ``r = i - N/2``,

    type 2:  y_m = sum_r x_r exp(-i omega_m . r)
    type 1:  x_r = sum_m y_m exp(+i omega_m . r)

Because ``exp(-i omega . r)`` factors over the axes, both are evaluated as
per-axis phase matrices (m x N_d) and batched products in complex128, in
chunks of samples; no (M x N) matrix is formed.  ``test_ndft_gpu_matches_oracle``
pins these against the oracle's explicit matrix at small sizes.
"""

from __future__ import annotations

import numpy as np
import torch


def phases(om, shape, sign):
    """Per-axis exp(sign * i * omega_d * r_d), complex128, r = i - N_d/2."""
    return [torch.exp(sign * 1j * om[:, d:d + 1].double()
                      * (torch.arange(n, device=om.device, dtype=torch.float64) - n // 2)[None, :])
            for d, n in enumerate(shape)]


def type2(om, x, shape, chunk=4096):
    """x (T, *shape) complex -> (T, M) complex128 at the samples ``om`` (M, d) radians."""
    d = len(shape)
    xb = x.to(torch.complex128).reshape(-1, *shape)
    T = xb.shape[0]
    out = []
    for s in range(0, om.shape[0], chunk):
        E = phases(om[s:s + chunk], shape, -1)
        m = E[0].shape[0]
        # contract the last axis by a GEMM, the others by broadcast sums
        t = torch.matmul(xb.reshape(T, -1, shape[-1]), E[-1].T.contiguous())   # (T, P, m)
        t = t.reshape(T, *shape[:-1], m)
        for ax in range(d - 2, -1, -1):
            t = torch.sum(t * E[ax].T, dim=-2)          # (T, N0..N_ax, m) -> (.., N_{ax-1}, m)
        out.append(t.reshape(T, m))
    return torch.cat(out, -1)


def type1(om, y, shape, chunk=8192):
    """y (T, M) complex -> (T, *shape) complex128."""
    d = len(shape)
    yb = y.to(torch.complex128).reshape(-1, om.shape[0])
    T = yb.shape[0]
    acc = torch.zeros((T, *shape), dtype=torch.complex128, device=om.device)
    for s in range(0, om.shape[0], chunk):
        E = phases(om[s:s + chunk], shape, 1)
        m = E[0].shape[0]
        for t in range(T):
            w = yb[t, s:s + m]                                   # (m,)
            lead = w[:, None] * E[0]                             # (m, N0)
            if d == 1:
                acc[t] += torch.sum(lead, 0)
                continue
            for ax in range(1, d - 1):
                lead = (lead[..., None] * E[ax].reshape(m, *([1] * ax), shape[ax]))
            lead = lead.reshape(m, -1)                           # (m, P)
            acc[t] += torch.matmul(lead.T.contiguous(), E[-1]).reshape(shape)
    return acc


def type2_at(om, x, shape, idx, chunk=512):
    """type 2 at the sample indices ``idx`` only."""
    return type2(om[idx], x, shape, chunk)


def type1_at(om, y, shape, vox, chunk=32768):
    """type 1 at the flat voxel indices ``vox`` only: (T, len(vox)) complex128."""
    dev = om.device
    coords = torch.stack(torch.unravel_index(vox, tuple(shape)), -1).double()
    coords = coords - torch.tensor([n // 2 for n in shape], dtype=torch.float64, device=dev)
    yb = y.to(torch.complex128).reshape(-1, om.shape[0])
    acc = torch.zeros((yb.shape[0], vox.numel()), dtype=torch.complex128, device=dev)
    for s in range(0, om.shape[0], chunk):
        ph = torch.exp(1j * (om[s:s + chunk].double() @ coords.T))   # (m, nvox)
        acc += yb[:, s:s + chunk] @ ph
    return acc


class SenseNdft:
    """Exact FP64 SENSE operator on the GPU with the oracle operator's
    interface (``oracle.mri.Operator``: numpy in and out, ``n_batchs`` = 1,
    ``density`` attribute, ``data_consistency``), so that the oracle's
    solvers (``oracle.solvers.cg``) can run on problems the numpy transforms
    cannot finish in seconds.  ``om`` are radians; ``norm = sqrt(prod(N) 2^d)``
    (``operators/base.py:837-840``)."""

    def __init__(self, om, shape, smaps=None, n_coils=1, density=None):
        self.om = om.double()
        self.shape = tuple(shape)
        self.ndim = len(shape)
        self.smaps = None if smaps is None else smaps.to(torch.complex128)
        self.n_coils = n_coils if smaps is None else smaps.shape[0]
        self.n_batchs = 1
        self.uses_sense = smaps is not None
        self.density = density
        self.norm = float(np.sqrt(np.prod(shape) * 2 ** len(shape)))

    def _dev(self, a):
        return torch.as_tensor(np.asarray(a)).to(self.om.device).to(torch.complex128)

    def op(self, x):
        x = self._dev(x).reshape(-1, *self.shape)
        v = x * self.smaps if self.uses_sense else x
        y = type2(self.om, v, self.shape) / self.norm
        return y.reshape(1, self.n_coils, -1).cpu().numpy()

    def adj_op(self, y):
        y = self._dev(y).reshape(self.n_coils, -1)
        if self.density is not None:
            y = y * self._dev(self.density).reshape(1, -1)
        x = type1(self.om, y, self.shape) / self.norm
        if self.uses_sense:
            x = torch.sum(torch.conj(self.smaps) * x, 0, keepdim=True)
        return x.reshape(1, -1, *self.shape).cpu().numpy()

    def data_consistency(self, x, y):
        return self.adj_op(self.op(x) - np.asarray(y).reshape(1, self.n_coils, -1))
