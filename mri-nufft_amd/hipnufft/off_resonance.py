"""Off-resonance corrected operator (``src/mrinufft/operators/off_resonance.py``).

``MRIFourierCorrected`` wraps a hipnufft operator A and applies

    op:     y[b, c] = sum_l B[:, l] * A(C_l x_b)[c]              (``:230-262``)
    adj_op: x_b = sum_l conj(C_l) A^H(conj(B[:, l]) * y[b])      (``:264-302``)

with B (NK, L) temporal interpolators (sample m uses row m % NK, NK samples
per shot) and C (L, *shape) spatial interpolators from a field-map
factorization (``field_map.get_orc_factorization``) or given directly.  The
reference runs L separate NUFFTs; for a hipnufft operator the L
interpolators are L extra transforms per coil of ONE batched SENSE chain
(``hipnufft_orc_op`` / ``hipnufft_orc_adj``), the C_l products fused into
the image-side FFT pass and the B weights fused into the k-space passes.
Any other operator of this package (the stacked one, spread/interp-only
plans) is wrapped the reference's way: L calls of its ``op`` / ``adj_op`` on
device tensors, the weights applied with torch.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _capi
from ._array import from_device, kind_of, to_device
from .base import FourierOperatorBase, power_method, random_start
from .field_map import _CLazy, get_complex_fieldmap_rad, get_orc_factorization
from .operator import DEFAULT_WORKSPACE_BYTES


class MRIFourierCorrected(FourierOperatorBase):
    """Fourier operator with B0 (and R2*) inhomogeneity compensation.

    Parameters follow the reference (``off_resonance.py:24-105``):
    fourier_op (a hipnufft operator), b0_map (Hz, or a complex field map in
    rad/s), readout_time (s; (n_shots, n_pts) or flat), r2star_map (Hz),
    mask (support used by the histogram), interpolator ("svd", "mti",
    "mfi", {"name": ..., **kwargs} or a (B, C) tuple).
    """

    def __init__(self, fourier_op, b0_map=None, readout_time=None, r2star_map=None, mask=None,
                 interpolator="svd"):
        if not hasattr(fourier_op, "device"):
            raise ValueError("the off-resonance operator wraps an operator of this package")
        self._fourier_op = fourier_op
        # the batched chain needs a full NUFFT plan of MRIHipNUFFT
        self._fused = (hasattr(fourier_op, "raw_op")
                       and not fourier_op._plan_opts.get("spreadinterponly"))
        self.squeeze_dims = fourier_op.squeeze_dims
        self._grad_wrt_data = False
        self._grad_wrt_traj = False
        self._grad_wrt_field_map = False
        self._orc_plan = None
        self._orc_plan_key = None
        self.interpolator_method = None
        if b0_map is None and readout_time is None and not isinstance(interpolator, tuple):
            raise ValueError("b0_map  and readout_time required for off-resonance correction.")
        field_map = None if b0_map is None else get_complex_fieldmap_rad(b0_map, r2star_map)
        self.compute_interpolator(interpolator, field_map, readout_time, mask)

    # -------------------------------------------------------------- set-up
    def compute_interpolator(self, interpolators, field_map, readout_time, mask):
        """Set the B and C interpolators (``off_resonance.py:107-194``)."""
        fo = self._fourier_op
        dev = fo.device
        if field_map is not None and tuple(field_map.shape) != tuple(fo.shape):
            self.log.warning("field_map and mask will be interpolated to match image shape.")
            from scipy.ndimage import zoom

            fm = field_map.cpu().numpy() if isinstance(field_map, torch.Tensor) else field_map
            field_map = zoom(fm, zoom=tuple(np.array(fo.shape) / np.array(fm.shape)), order=1)
            if mask is not None:
                mk = mask.cpu().numpy() if isinstance(mask, torch.Tensor) else np.asarray(mask)
                mask = zoom(mk, zoom=tuple(np.array(fo.shape) / np.array(mk.shape)), order=0)
        self._field_map = field_map
        self.mask = mask
        self.readout_time = readout_time
        n_samples = fo.n_samples
        self.n_shots = 1
        if readout_time is not None:
            size = int(np.prod(np.shape(readout_time)))
            if size != n_samples:
                n_shot, r = divmod(n_samples, size)
                if r != 0:
                    raise ValueError("readout_time should divide or equal the size of the samples.")
                self.n_shots = n_shot

        if isinstance(interpolators, tuple):
            B, C = interpolators
            if not all(hasattr(a, "shape") for a in (B, C)):
                raise ValueError("Provide a tuple of 2 array_like data"
                                 " for space and time interpolators")
            if int(np.prod(B.shape)) != n_samples:
                n_shot, r = divmod(n_samples, B.shape[0])
                if r != 0:
                    raise ValueError("Time interpolator should divide or equal size of the samples.")
                self.n_shots = n_shot
            self._set_BC(B, C)
            return

        fm_zero = _all_zero(field_map)
        t_zero = _all_zero(readout_time)
        if fm_zero or t_zero:  # no off-resonance effect
            self.B = torch.ones((n_samples, 1), dtype=torch.complex64, device=dev)
            self.C = torch.ones((1, *fo.shape), dtype=torch.complex64, device=dev)
            return

        kwargs = {}
        if isinstance(interpolators, dict):
            kwargs = interpolators.copy()
            interpolators = kwargs.pop("name")
        if isinstance(interpolators, str):
            interpolators = get_orc_factorization(interpolators)
        if not callable(interpolators):
            raise ValueError(f"Unknown off-resonance interpolator ``{interpolators}``")
        self.interpolator_method = interpolators
        if field_map is None or readout_time is None:
            raise ValueError("b0_map  and readout_time required for off-resonance correction.")
        fm = to_device(get_complex_fieldmap_rad(field_map), torch.complex64, dev)
        t = to_device(readout_time, torch.float32, dev).reshape(-1)
        mk = None if mask is None else to_device(mask, torch.bool, dev)
        B, C, _ = interpolators(field_map=fm, readout_time=t, mask=mk, **kwargs)
        self._set_BC(B, C)

    def _set_BC(self, B, C):
        dev = self._fourier_op.device
        if isinstance(C, _CLazy):
            C = C.full()
        elif not isinstance(C, (np.ndarray, torch.Tensor)) and hasattr(C, "__getitem__"):
            C = torch.stack([to_device(C[i], torch.complex64, dev) for i in range(len(C))])
        B = to_device(B, torch.complex64, dev)
        C = to_device(C, torch.complex64, dev)
        if B.ndim == 1:
            B = B[:, None]
        C = C.reshape(-1, *self._fourier_op.shape)
        if B.shape[1] != C.shape[0]:
            raise ValueError("B (NK, L) and C (L, *shape) disagree on L")
        self.B, self.C = B, C

    def update_field_map(self, new_field_map):
        """Recompute the interpolators for a new field map."""
        self.compute_interpolator(self.interpolator_method, new_field_map, self.readout_time,
                                  self.mask)

    @property
    def field_map(self):
        return self._field_map

    @field_map.setter
    def field_map(self, new_field_map):
        self.update_field_map(new_field_map)

    @property
    def full_readout_time(self):
        """The readout times with each entry repeated n_shots times in a row
        (``numpy.repeat``, ``off_resonance.py:216-220``), used by the field-map
        gradient."""
        t = self.readout_time
        if isinstance(t, torch.Tensor):
            return t.reshape(-1).repeat_interleave(self.n_shots)
        return np.repeat(np.asarray(t).reshape(-1), self.n_shots)

    @property
    def n_interpolators(self):
        return int(self.B.shape[1])

    @property
    def n_samples_per_shot(self):
        return int(self.B.shape[0])

    @property
    def autograd_available(self):
        return self._fourier_op.autograd_available

    def __getattr__(self, name):
        if name.startswith("__") or name == "_fourier_op":
            raise AttributeError(name)
        return getattr(self._fourier_op, name)

    # ---------------------------------------------------------- the plan
    def _plan(self):
        """A plan whose workspace holds as many of the B*C*L transforms as the
        budget allows (the wrapped operator's when it is large enough)."""
        fo = self._fourier_op
        need = fo.n_batchs * fo.n_coils * self.n_interpolators
        nf = [2 * s for s in fo.shape]
        grid = 8 * int(np.prod(nf)) * max(1, int(fo._plan_opts["upsampfac"] / 2.0) ** 2)
        tmax = int(max(1, min(need, DEFAULT_WORKSPACE_BYTES // max(grid, 1))))
        T = -(-need // -(-need // tmax))  # equal chunks: no short tail launch
        if fo.raw_op.n_trans >= T:
            return fo.raw_op.plan
        key = (T, id(fo._samples))
        if self._orc_plan is None or self._orc_plan_key != key:
            plan = _capi.RawPlan(fo.shape, n_trans=T, eps=fo.eps, isign=-1, **fo._plan_opts)
            plan.setpts(fo._samples)
            self._orc_plan, self._orc_plan_key = plan, key
        return self._orc_plan

    # --------------------------------------------------------- transforms
    def _loop_op(self, x):
        """``sum_l B[:, l] * A(C_l x)`` by L calls of the wrapped operator."""
        fo = self._fourier_op
        B, C, M = fo.n_batchs, fo.n_coils, fo.n_samples
        NK = self.n_samples_per_shot
        sq, fo.squeeze_dims = fo.squeeze_dims, False
        try:
            y = torch.zeros((B, C, M // NK, NK), dtype=torch.complex64, device=fo.device)
            for l in range(self.n_interpolators):
                yl = to_device(fo.op(self.C[l] * x), torch.complex64, fo.device)
                y += yl.reshape(B, C, M // NK, NK) * self.B[:, l]
        finally:
            fo.squeeze_dims = sq
        return y.reshape(B, C, M)

    def _loop_adj(self, y, ic):
        """``sum_l conj(C_l) A^H(conj(B[:, l]) * y)`` by L calls of the wrapped operator."""
        fo = self._fourier_op
        B, C, M = fo.n_batchs, fo.n_coils, fo.n_samples
        NK = self.n_samples_per_shot
        y = y.reshape(B, C, M // NK, NK)
        sq, fo.squeeze_dims = fo.squeeze_dims, False
        try:
            img = torch.zeros((B, ic, *fo.shape), dtype=torch.complex64, device=fo.device)
            for l in range(self.n_interpolators):
                xl = fo.adj_op((y * self.B[:, l].conj()).reshape(B, C, M))
                img += to_device(xl, torch.complex64, fo.device).reshape(img.shape) * self.C[l].conj()
        finally:
            fo.squeeze_dims = sq
        return img

    def op(self, data, *args):
        """Forward with off-resonance -> (B, C, M) (squeezed if asked)."""
        fo = self._fourier_op
        fo.check_shape(image=data)
        kind = kind_of(data)
        B, C, M = fo.n_batchs, fo.n_coils, fo.n_samples
        if not self._fused:
            with torch.cuda.device(fo.device):
                x = to_device(data, torch.complex64, fo.device)
                x = x.reshape(B, 1 if fo.uses_sense else C, *fo.shape)
                return from_device(self._safe_squeeze(self._loop_op(x)), kind)
        with torch.cuda.device(fo.device):
            x = fo._img_in(data, 1 if fo.uses_sense else C)
            out = torch.empty((B, C, M), dtype=torch.complex64, device=fo.device)
            self._plan().orc_op(x, fo.smaps, self.C, self.B, out, B, C, fo.inv_norm_factor)
        return from_device(self._safe_squeeze(out), kind)

    def adj_op(self, coeffs, *args):
        """Adjoint with off-resonance -> (B, 1 or C, *shape) (squeezed if asked)."""
        fo = self._fourier_op
        fo.check_shape(ksp=coeffs)
        kind = kind_of(coeffs)
        B, C, M = fo.n_batchs, fo.n_coils, fo.n_samples
        with torch.cuda.device(fo.device):
            y = to_device(coeffs, torch.complex64, fo.device).reshape(B, C, M)
            ic = 1 if fo.uses_sense else C
            if not self._fused:
                return from_device(self._safe_squeeze(self._loop_adj(y, ic)), kind)
            out = torch.empty((B, ic, *fo.shape), dtype=torch.complex64, device=fo.device)
            self._plan().orc_adj(y, fo.smaps, self.C, self.B, fo.density, out, B, C,
                                 fo.inv_norm_factor)
        return from_device(self._safe_squeeze(out), kind)

    def data_consistency(self, image_data, obs_data):
        """``adj_op(op(x) - y)``."""
        return self.adj_op(self.op(image_data) - obs_data)

    def _single_coil_AHA(self, x):
        fo = self._fourier_op
        plan = self._plan()
        y = torch.empty((1, 1, fo.n_samples), dtype=torch.complex64, device=fo.device)
        out = torch.empty((1, 1, *fo.shape), dtype=torch.complex64, device=fo.device)
        plan.orc_op(x.reshape(1, 1, *fo.shape).contiguous(), None, self.C, self.B, y, 1, 1,
                    fo.inv_norm_factor)
        plan.orc_adj(y, None, self.C, self.B, fo.density, out, 1, 1, fo.inv_norm_factor)
        return out.reshape(fo.shape)

    def get_lipschitz_cst(self, max_iter=10, **kwargs):
        """Largest eigenvalue of A^H A for one coil without smaps (``off_resonance.py:312-365``)."""
        fo = self._fourier_op
        with torch.cuda.device(fo.device):
            x = random_start(fo.shape, fo.device, kwargs.get("seed"))
            if self._fused:
                lipschitz_cst, _ = power_method(max_iter, self._single_coil_AHA, x=x)
                return float(lipschitz_cst)
            saved = (fo._smaps, fo.n_coils, fo.n_batchs)
            fo._smaps, fo.n_coils, fo.n_batchs = None, 1, 1
            try:
                lipschitz_cst, _ = power_method(
                    max_iter, lambda v: self._loop_adj(self._loop_op(v), 1).reshape(fo.shape),
                    x=x)
            finally:
                fo._smaps, fo.n_coils, fo.n_batchs = saved
        return float(lipschitz_cst)

    def make_autograd(self, *, wrt_data=True, wrt_traj=False, wrt_field_map=False,
                      paired_batch=False):
        """Torch autograd wrapper (``off_resonance.py:401-444``): data and field-map
        gradients; the trajectory gradient is not available here."""
        if wrt_traj:
            raise ValueError("the trajectory gradient is not available for the "
                             "off-resonance operator")
        if not self._fourier_op.autograd_available:
            raise ValueError("Backend does not support auto-differentiation.")
        from .autodiff import MRINufftAutoGrad

        return MRINufftAutoGrad(self, wrt_data=wrt_data, wrt_traj=False,
                                wrt_field_map=wrt_field_map, paired_batch=paired_batch)


def _all_zero(a):
    if a is None:
        return False
    if isinstance(a, torch.Tensor):
        return bool(torch.all(a == 0))
    return bool(np.all(np.asarray(a) == 0))
