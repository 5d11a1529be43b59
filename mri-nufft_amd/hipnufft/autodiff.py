"""Torch autograd for the hipnufft operator (``src/mrinufft/operators/autodiff.py``).

Backward rules follow the reference (Wang & Fessler, IEEE TCI 2023):

* op:      dL/dx = adj_op(dy) (density included, as ``autodiff.py:14-19``);
           dL/dw_{m,d} = Re sum_{b,c} conj(dy) * d y / d w_{m,d}   (``:22-41``)
* adj_op:  dL/dy = op(dx)                                           (``:63-66``);
           dL/dw_{m,d} = Re sum_{b,c} y * d/dw_{m,d} T+(conj(S_c) conj(dx))  (``:69-86``)

The reference evaluates the trajectory terms with ``ndim`` extra NUFFTs of
``x * r_d``.  Here they come from one type-2 chain per coil followed by the
``k_grad_traj`` kernel, which interpolates the fine grid with the analytic
derivative of the ES kernel (``hipnufft_sense_grad_traj``).  Units are
radians, like the reference for finufft / cufinufft (factor 1,
``autodiff.py:27-29``).  As in the reference, the adjoint's trajectory term
does not include the density weights.

For the off-resonance operator the field-map terms follow ``autodiff.py:44-55,
89-100``: op: ``x * adj_op(dy * t)``; adj_op: ``conj(dx) * adj_op(y * t)``
with t the readout time of every sample.
"""

from __future__ import annotations

import numpy as np
import torch

from ._array import to_device


def _grad_traj(nufft, img, dy, iflag, conj_smaps):
    B, C, M = nufft.n_batchs, nufft.n_coils, nufft.n_samples
    ic = 1 if nufft.uses_sense else C
    x = to_device(img, torch.complex64, nufft.device).reshape(B, ic, *nufft.shape)
    d = to_device(dy, torch.complex64, nufft.device).reshape(B, C, M)
    g = torch.empty((M, nufft.ndim), dtype=torch.float32, device=nufft.device)
    nufft.raw_op.plan.sense_grad_traj(x, nufft.smaps, d, B, C, nufft.inv_norm_factor, iflag,
                                      conj_smaps, g)
    return g


def _readout_times(nufft, device):
    return to_device(nufft.full_readout_time, torch.float32, device)


def _field_map_grad_op(nufft, x, dy):
    if not getattr(nufft, "_grad_wrt_field_map", False):
        return None
    return x * nufft.adj_op(dy * _readout_times(nufft, dy.device))


def _field_map_grad_adj(nufft, y, dx):
    if not getattr(nufft, "_grad_wrt_field_map", False):
        return None
    return dx.conj() * nufft.adj_op(y * _readout_times(nufft, y.device))


class _NUFFT_OP(torch.autograd.Function):
    """Autograd of ``op`` (image -> k-space)."""

    @staticmethod
    def forward(ctx, x, traj, field_map, nufft_op):
        ctx.save_for_backward(x)
        ctx.nufft = nufft_op
        return nufft_op.op(x)

    @staticmethod
    def backward(ctx, dy):
        (x,) = ctx.saved_tensors
        nufft = ctx.nufft
        gx = nufft.adj_op(dy) if nufft._grad_wrt_data else None
        gt = _grad_traj(nufft, x, dy, -1, False) if nufft._grad_wrt_traj else None
        if gx is not None:
            gx = gx.reshape(x.shape)
        return gx, gt, _field_map_grad_op(nufft, x, dy), None


class _NUFFT_ADJOP(torch.autograd.Function):
    """Autograd of ``adj_op`` (k-space -> image)."""

    @staticmethod
    def forward(ctx, y, traj, field_map, nufft_op):
        ctx.save_for_backward(y)
        ctx.nufft = nufft_op
        return nufft_op.adj_op(y)

    @staticmethod
    def backward(ctx, dx):
        (y,) = ctx.saved_tensors
        nufft = ctx.nufft
        gy = nufft.op(dx) if nufft._grad_wrt_data else None
        gt = None
        if nufft._grad_wrt_traj:
            # Re sum y * D T+(conj(S) conj(dx)) == Re sum conj(conj(y)) * D(...)
            gt = _grad_traj(nufft, dx.conj(), y.conj(), 1, True)
        if gy is not None:
            gy = gy.reshape(y.shape)
        return gy, gt, _field_map_grad_adj(nufft, y, dx), None


class MRINufftAutoGrad(torch.nn.Module):
    """Wrap a hipnufft operator for torch autodiff (``autodiff.py:157-456``)."""

    def __init__(self, nufft_op, wrt_data=True, wrt_traj=False, paired_batch=False,
                 wrt_field_map=False):
        if (wrt_data or wrt_traj or wrt_field_map) and nufft_op.squeeze_dims:
            raise ValueError("Squeezing dimensions is not supported for autodiff.")
        super().__init__()
        self.nufft_op = nufft_op
        self.nufft_op._grad_wrt_traj = wrt_traj
        self.nufft_op._grad_wrt_data = wrt_data
        self.nufft_op._grad_wrt_field_map = wrt_field_map
        if wrt_traj:
            self._samples_torch = self.nufft_op.samples.detach().clone()
            self._samples_torch.requires_grad = True
        self._field_map_torch = None
        if wrt_field_map:
            if not hasattr(nufft_op, "update_field_map"):
                raise ValueError("Underlying nufft operator does not support field map.")
            fm = nufft_op.field_map
            fm = fm.detach().clone() if isinstance(fm, torch.Tensor) else torch.from_numpy(
                np.ascontiguousarray(fm))
            self._field_map_torch = fm.to(nufft_op.device)
            self._field_map_torch.requires_grad = True
        self.paired_batch = paired_batch

    def _field_map_arg(self, field_map):
        if field_map is not None and not hasattr(self.nufft_op, "update_field_map"):
            raise ValueError("Underlying nufft operator does not support field map.")
        if field_map is None and self._field_map_torch is not None:
            return self._field_map_torch
        return field_map

    def op(self, x, smaps=None, samples=None, field_map=None):
        if self.paired_batch:
            return self._op_batched(x, smaps, samples)
        return _NUFFT_OP.apply(x, self.samples, self._field_map_arg(field_map), self.nufft_op)

    def adj_op(self, kspace, smaps=None, samples=None, field_map=None):
        if self.paired_batch:
            return self._adj_op_batched(kspace, smaps, samples)
        return _NUFFT_ADJOP.apply(kspace, self.samples, self._field_map_arg(field_map),
                                  self.nufft_op)

    def _op_batched(self, imgs, smaps=None, samples=None):
        out = []
        for i in range(len(imgs)):
            try:
                if smaps is not None:
                    self.nufft_op.smaps = smaps[i]
                if samples is not None:
                    self.samples = samples[i]
                out.append(_NUFFT_OP.apply(imgs[i], self.samples, None, self.nufft_op))
            except Exception as e:
                raise RuntimeError(f"Failed at batch index {i}") from e
        return torch.stack(out, dim=0)

    def _adj_op_batched(self, kspace, smaps=None, samples=None):
        out = []
        for i in range(len(kspace)):
            try:
                if smaps is not None:
                    self.nufft_op.smaps = smaps[i]
                if samples is not None:
                    self.samples = samples[i]
                out.append(_NUFFT_ADJOP.apply(kspace[i], self.samples, None, self.nufft_op))
            except Exception as e:
                raise RuntimeError(f"Failed at batch index {i + 1}") from e
        return torch.stack(out, dim=0)

    @property
    def samples(self):
        try:
            return self._samples_torch
        except AttributeError:
            return self.nufft_op.samples

    @samples.setter
    def samples(self, value):
        self.update_samples(value, unsafe=False)

    def update_samples(self, new_samples, *, unsafe=False):
        self._samples_torch = new_samples
        self.nufft_op.update_samples(new_samples.detach(), unsafe=unsafe)

    @property
    def field_map(self):
        if not hasattr(self.nufft_op, "update_field_map"):
            raise ValueError("Underlying nufft operator does not support field map.")
        if self._field_map_torch is not None:
            return self._field_map_torch
        return self.nufft_op.field_map

    @field_map.setter
    def field_map(self, value):
        self.update_field_map(value)

    def update_field_map(self, new_field_map):
        """Set the field map and recompute the operator's interpolators."""
        if not hasattr(self.nufft_op, "update_field_map"):
            raise ValueError("Underlying nufft operator does not support field map.")
        self._field_map_torch = new_field_map
        self.nufft_op.update_field_map(new_field_map.detach())

    def __getattr__(self, name):
        try:
            return super().__getattr__(name)
        except AttributeError:
            return getattr(self.nufft_op, name)
