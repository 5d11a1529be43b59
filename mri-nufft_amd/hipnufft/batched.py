"""Many trajectories of uneven length in one plan: one launch per stage.

``MultiTrajectoryOperator`` holds K reconstructions that share an image
shape and a coil count but each have their own trajectory (any number of
samples), sensitivity maps and density.  All K point sets are binned and
sorted together (``hipnufft_setpts_sets``); every stage of ``op``,
``adj_op`` and ``data_consistency`` is then a single launch over all
trajectories and coils, with work items balanced over tiles whatever the
trajectory lengths.  ``cg`` solves the K problems together with per-problem
step sizes, inner products (``hipnufft_dot_c64_rows``) and stopping flags,
so each result equals the one ``MRIHipNUFFT.pinv_solver(optim="cg")`` gives
for that problem alone (same recurrences, ``extras/optim.py:801-902``).

The reference plans one trajectory per operator (``cufinufft.py:202-275``);
this replaces a Python loop over K operators.

k-space data of the K problems is one flat complex64 tensor laid out
``[k][c][M_k]`` back to back (``split`` / ``join`` convert to and from a
list of (C, M_k) tensors).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _capi
from ._array import from_device, kind_of, proper_trajectory, to_device
from .base import random_start
from .operator import DEFAULT_WORKSPACE_BYTES, MRIHipNUFFT


class MultiTrajectoryOperator:
    """K SENSE (or calibrationless) problems with uneven trajectories.

    Parameters
    ----------
    samples : sequence of K arrays (M_k, d) or (Nc, Ns, d), in [-0.5, 0.5) or radians.
    shape : image shape shared by all problems.
    n_coils : coils per problem.
    smaps : None (calibrationless: images (K, C, *shape)), (C, *shape) shared
        by all problems, or (K, C, *shape).
    density : None / False, True or "pipe" (Pipe per trajectory), or a
        sequence of K (M_k,) arrays.
    eps, n_trans, gpu_device_id, and plan tuning options (tile,
    spread_group, interp_group, max_subproblem, spread_max, upsampfac).
    """

    def __init__(self, samples, shape, n_coils=1, smaps=None, density=None, eps=1e-6,
                 n_trans=None, gpu_device_id=0, **plan_opts):
        if not MRIHipNUFFT.available:
            raise RuntimeError("hipnufft is not available (library or GPU missing)")
        self.device = torch.device("cuda", int(gpu_device_id))
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(self.shape)
        self.n_coils = int(n_coils)
        self.eps = float(eps)
        trajs = [proper_trajectory(s, normalize="pi") for s in samples]
        if not trajs:
            raise ValueError("need at least one trajectory")
        for t in trajs:
            if t.shape[1] != self.ndim:
                raise ValueError("trajectory dimension does not match the image shape")
        self.n_sets = len(trajs)
        self.sizes = [int(t.shape[0]) for t in trajs]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n_samples = int(self.offsets[-1])
        self.norm_factor = math.sqrt(float(np.prod(self.shape)) * (2 ** self.ndim))
        self.inv_norm_factor = 1.0 / self.norm_factor
        K, C = self.n_sets, self.n_coils
        with torch.cuda.device(self.device):
            self._coords = [to_device(t, torch.float32, self.device) for t in trajs]
            self.n_trans = self._choose_n_trans(n_trans, plan_opts)
            self.plan = _capi.RawPlan(self.shape, n_trans=self.n_trans, eps=self.eps, isign=-1,
                                      **plan_opts)
            self.plan.setpts_sets(self._coords)
            random_start((1,), self.device, 0)  # device RNG loaded before the first solve
            self._smaps = None
            if smaps is not None:
                s = to_device(smaps, torch.complex64, self.device)
                if s.shape == (C, *self.shape):
                    s = s.unsqueeze(0).expand(K, C, *self.shape)
                if tuple(s.shape) != (K, C, *self.shape):
                    raise ValueError(f"smaps must be (C, *shape) or (K, C, *shape), got {tuple(s.shape)}")
                self._smaps = s.contiguous()
            self._density = self._make_density(density)
        self._lipschitz = None

    # ----------------------------------------------------------------- setup
    def _choose_n_trans(self, n_trans, plan_opts):
        C, K = self.n_coils, self.n_sets
        if n_trans is not None:
            if int(n_trans) < C:
                raise ValueError("n_trans must be >= n_coils for several trajectories")
            return int(n_trans)
        up = float(plan_opts.get("upsampfac", 2.0))
        cells = float(np.prod([max(int(up * n), 16) for n in self.shape]))
        per_set = C * cells * 8.0
        spc = int(max(1, min(K, DEFAULT_WORKSPACE_BYTES // max(per_set, 1.0))))
        return spc * C

    def _make_density(self, density):
        if density is None or density is False:
            return None
        if density is True or density == "pipe":
            parts = [torch.as_tensor(MRIHipNUFFT.pipe(c, self.shape), device=self.device)
                     for c in self._coords]
        else:
            parts = [to_device(d, torch.float32, self.device).reshape(-1) for d in density]
        if [int(p.numel()) for p in parts] != self.sizes:
            raise ValueError("density sizes do not match the trajectories")
        return torch.cat([p.to(torch.float32) for p in parts]).contiguous()

    @property
    def uses_sense(self):
        return self._smaps is not None

    @property
    def uses_density(self):
        return self._density is not None

    @property
    def img_full_shape(self):
        return (self.n_sets, *(() if self.uses_sense else (self.n_coils,)), *self.shape)

    # --------------------------------------------------------- k-space layout
    def split(self, y):
        """Flat [k][c][M_k] tensor -> list of K (C, M_k) views."""
        C = self.n_coils
        out = []
        for k, m in enumerate(self.sizes):
            o = int(self.offsets[k]) * C
            out.append(y[o:o + C * m].reshape(C, m))
        return out

    def join(self, parts):
        """List of K (C, M_k) arrays -> flat [k][c][M_k] device tensor."""
        if len(parts) != self.n_sets:
            raise ValueError("need one k-space array per trajectory")
        flat = [to_device(p, torch.complex64, self.device).reshape(-1) for p in parts]
        for f, m in zip(flat, self.sizes):
            if f.numel() != self.n_coils * m:
                raise ValueError("k-space size does not match its trajectory")
        return torch.cat(flat)

    def _ksp_in(self, y):
        if isinstance(y, (list, tuple)):
            return self.join(y)
        y = to_device(y, torch.complex64, self.device).reshape(-1)
        if y.numel() != self.n_coils * self.n_samples:
            raise ValueError("k-space size does not match the trajectories")
        return y.contiguous()

    def _img_in(self, x):
        x = to_device(x, torch.complex64, self.device)
        return x.reshape(self.img_full_shape).contiguous()

    # ------------------------------------------------------------- transforms
    def op(self, x):
        """Forward of all K problems -> flat [k][c][M_k] (normalised like ``op``)."""
        with torch.cuda.device(self.device):
            xi = self._img_in(x)
            out = torch.empty(self.n_coils * self.n_samples, dtype=torch.complex64,
                              device=self.device)
            self.plan.sense_op(xi, self._smaps, out, self.n_sets, self.n_coils,
                               self.inv_norm_factor)
        return out

    def adj_op(self, y):
        """Adjoint (density applied) -> (K, *shape) (SENSE) or (K, C, *shape)."""
        with torch.cuda.device(self.device):
            yi = self._ksp_in(y)
            out = torch.empty(self.img_full_shape, dtype=torch.complex64, device=self.device)
            self.plan.sense_adj(yi, self._smaps, self._density, out, self.n_sets, self.n_coils,
                                self.inv_norm_factor)
        return out

    def data_consistency(self, x, y):
        """Fused ``A^H D (A x - y)`` of every problem -> image stack."""
        with torch.cuda.device(self.device):
            xi = self._img_in(x)
            yi = self._ksp_in(y)
            out = torch.empty(self.img_full_shape, dtype=torch.complex64, device=self.device)
            self.plan.sense_dc(xi, self._smaps, yi, self._density, out, self.n_sets,
                               self.n_coils, self.inv_norm_factor)
        return out

    # ------------------------------------------------------------------ solver
    def _aha_single_coil(self, x):
        """A_k^H D_k A_k x_k for every k without coils (the Lipschitz operator)."""
        K = self.n_sets
        y = torch.empty(self.n_samples, dtype=torch.complex64, device=self.device)
        self.plan.sense_op(x, None, y, K, 1, self.inv_norm_factor)
        out = torch.empty((K, *self.shape), dtype=torch.complex64, device=self.device)
        self.plan.sense_adj(y, None, self._density, out, K, 1, self.inv_norm_factor)
        return out

    def get_lipschitz_cst(self, max_iter=10, seed=None):
        """Per-problem power method on A^H D A without coils (``base.py:1155-1221``,
        ``cufinufft.py:1239-1279``); returns K floats.  Problems whose norm
        has converged keep their value while the others iterate."""
        K = self.n_sets
        with torch.cuda.device(self.device):
            x = random_start((K, *self.shape), self.device, seed)
            nrm = torch.linalg.vector_norm(x.reshape(K, -1), dim=1)
            x = x / nrm.reshape(K, *([1] * self.ndim))
            done = torch.zeros(K, dtype=torch.bool, device=self.device)
            val = nrm.clone()
            for _ in range(max_iter):
                xn = self._aha_single_coil(x)
                nn = torch.linalg.vector_norm(xn.reshape(K, -1), dim=1)
                conv = (nrm - nn).abs() < 1e-6
                val = torch.where(done, val, nn)
                xn = xn / nn.reshape(K, *([1] * self.ndim))
                keep = done.reshape(K, *([1] * self.ndim))
                x = torch.where(keep, x, xn)
                nrm = torch.where(done, nrm, nn)
                done = done | conv
                if bool(done.all()):
                    break
        return val.double().cpu().numpy()

    def cg(self, kspace_data, max_iter=10, tol=1e-4, graph=True, lipschitz=None):
        """CG on every problem at once, each with its own L, beta and stop
        test (the recurrences of ``extras/optim.py:801-902``); the inner
        products are FP64 row-wise dots.  ``graph=True`` captures one
        iteration in a HIP graph.  Returns the (K, ...) image stack."""
        K = self.n_sets
        dev = self.device
        with torch.cuda.device(dev):
            y = self._ksp_in(kspace_data)
            L = self.get_lipschitz_cst() if lipschitz is None else np.asarray(lipschitz, np.float64)
            L = np.broadcast_to(L, (K,)).astype(np.float64)
            shape = self.img_full_shape
            bshape = (K,) + (1,) * (len(shape) - 1)
            old_density = None
            if self.uses_density:
                x_init = self._scaled_dcp(y)
                old_density, self._density = self._density, None
            else:
                x_init = torch.zeros(shape, dtype=torch.complex64, device=dev)
            try:
                inv_l = torch.as_tensor(1.0 / L, device=dev, dtype=torch.float32).reshape(bshape)
                image = x_init.contiguous()
                grad = self.data_consistency(image, y)
                velocity = grad * inv_l  # tol * 0 + grad / L
                image = image - velocity
                image = self._cg_iterate(y, image, grad, velocity, inv_l, max_iter, tol, graph)
            finally:
                if old_density is not None:
                    self._density = old_density
        return image

    def _scaled_dcp(self, y):
        """Per-problem density-compensated warm start (``optim.py:151-165``)."""
        x = self.adj_op(y)
        yy = self.op(x)
        ny = torch.stack([torch.linalg.vector_norm(p) for p in self.split(y)])
        nyy = torch.stack([torch.linalg.vector_norm(p) for p in self.split(yy)])
        return x * (ny / nyy).reshape((self.n_sets,) + (1,) * (x.ndim - 1)).to(x.dtype)

    def _cg_iterate(self, y, image, grad, vel, inv_l, max_iter, tol, graph):
        K = self.n_sets
        dev = self.device
        shape = image.shape
        bshape = (K,) + (1,) * (len(shape) - 1)
        st = dict(image=image.contiguous(), grad=grad.contiguous(), vel=vel.contiguous(),
                  done=torch.zeros(bshape, dtype=torch.bool, device=dev))
        bufs = dict(nrm=torch.empty((K, 2), dtype=torch.float64, device=dev),
                    num=torch.empty((K, 2), dtype=torch.float64, device=dev),
                    den=torch.empty((K, 2), dtype=torch.float64, device=dev),
                    ws=torch.empty(K * _capi.DOT_WS, dtype=torch.float64, device=dev),
                    zero=torch.zeros(K, dtype=torch.float64, device=dev))
        tol2 = tol * tol

        def step():
            gn = self.data_consistency(st["image"], y)
            g2 = gn.reshape(K, -1)
            ws = bufs["ws"]
            _capi.dot_c64_rows(g2, g2, True, out=bufs["nrm"], ws=ws)
            diff = (gn - st["grad"]).reshape(K, -1)
            _capi.dot_c64_rows(g2, diff, False, out=bufs["num"], ws=ws)
            go = st["grad"].reshape(K, -1)
            _capi.dot_c64_rows(go, go, False, out=bufs["den"], ws=ws)
            stop = st["done"] | (bufs["nrm"][:, 0] <= tol2).reshape(bshape)
            n_re, n_im = bufs["num"][:, 0], bufs["num"][:, 1]
            d_re, d_im = bufs["den"][:, 0], bufs["den"][:, 1]
            dd = d_re * d_re + d_im * d_im
            b_re = (n_re * d_re + n_im * d_im) / dd
            b_im = (n_im * d_re - n_re * d_im) / dd
            pos = (b_re > 0) | ((b_re == 0) & (b_im > 0))
            b_re = torch.where(pos, b_re, bufs["zero"])
            b_im = torch.where(pos, b_im, bufs["zero"])
            beta = torch.complex(b_re, b_im).to(torch.complex64).reshape(bshape)
            vel_new = gn + beta * st["vel"]
            img_new = st["image"] - vel_new * inv_l
            st["image"].copy_(torch.where(stop, st["image"], img_new))
            st["vel"].copy_(torch.where(stop, st["vel"], vel_new))
            st["grad"].copy_(torch.where(stop, st["grad"], gn))
            st["done"].copy_(stop)

        if max_iter <= 0:
            return st["image"]
        step()
        remaining = max_iter - 1
        if remaining <= 0:
            return st["image"]
        if not graph:
            for _ in range(remaining):
                step()
            return st["image"]
        s = torch.cuda.Stream(device=dev)
        s.wait_stream(torch.cuda.current_stream(dev))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            step()
            remaining -= 1
            torch.cuda.current_stream(dev).synchronize()
            if remaining > 0:
                with torch.cuda.graph(g, stream=s):
                    step()
        torch.cuda.current_stream(dev).wait_stream(s)
        for _ in range(remaining):
            g.replay()
        torch.cuda.current_stream(dev).wait_stream(s)
        return st["image"]

    def to_kind(self, t, like):
        """Return ``t`` in the array kind of ``like`` (numpy / torch-cpu / cuda)."""
        return from_device(t, kind_of(like))

    def __repr__(self):
        return (f"MultiTrajectoryOperator(shape={self.shape}, n_sets={self.n_sets}, "
                f"n_coils={self.n_coils}, n_samples={self.n_samples}, n_trans={self.n_trans}, "
                f"sense={self.uses_sense}, density={self.uses_density})")
