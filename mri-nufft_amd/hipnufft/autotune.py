"""Tile / bin / grouping autotuner for libhipnufft plans.

For a given trajectory, image shape and transform count the tuner times one
SENSE-free type 2 + type 1 pair (the operator's inner op+adj) for candidate
plan options and keeps the fastest, by coordinate search:

1. tile shape (which sets both the bin sort and the spread / interp work
   granularity);
2. in 3D, the spreading kernel (``spread_method``: cell owners, block
   owners or the matrix-core kernel);
3. transforms accumulated per spreading workgroup (``spread_group``);
4. transforms gathered per interpolation workgroup (``interp_group``);
5. dense-bin splitting (``max_subproblem``, ``spread_max``).

Only candidates the kernels accept are tried (``tile_ok``).  Results are cached in-process by a
fingerprint of the trajectory (size, shape, coils, a strided subsample of the
coordinates), and optionally in a JSON file (``HIPNUFFT_TUNE_CACHE``).

The reference exposes these knobs as cufinufft options (``gpu_method``,
``gpu_sort``, ``gpu_binsize*``; ``cufinufft.py:38-46``) and leaves them at
library defaults.
"""

from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np
import torch

from . import _capi

_CACHE: dict = {}


def _kernel_width(eps, upsampfac=2.0):
    if upsampfac == 2.0:
        w = int(np.ceil(-np.log10(eps / 10.0)))
    else:
        w = int(np.ceil(-np.log(eps) / (np.pi * np.sqrt(1.0 - 1.0 / upsampfac))))
    return min(max(w, 2), 16)


def tile_ok(tile, w, interp_group=1, interp_span=None):
    """Whether the kernels accept ``tile`` (per-axis, C order) at width w:
    at most 256 (column, 4-cell segment) pairs per spreading tile; an
    interpolation region (interp_span tiles along the last axis, default 2
    in 3D, plus the width, with a 4-byte offset per row) within 64 KiB for
    several transforms per workgroup, within 160 KiB for one."""
    cells = int(np.prod(tile))
    zl = tile[-1]
    if (cells // zl) * -(-zl // 4) > 256:
        return False
    span = interp_span or (2 if len(tile) == 3 else 1)
    region = [t + w for t in tile[:-1]] + [span * tile[-1] + w]
    rows = int(np.prod(region[:-1])) if len(tile) > 1 else 0   # row-offset table
    g = max(interp_group, 1)
    return int(np.prod(region)) * 8 * g + 4 * rows <= (65536 if g > 1 else 163840)


def tile_candidates(ndim, w):
    if ndim == 1:
        cands = [(256,), (512,), (1024,)]
    elif ndim == 2:
        cands = [(32, 32), (16, 32), (32, 16), (16, 64), (64, 16), (16, 16)]
    else:
        cands = [(8, 8, 16), (8, 8, 8), (4, 8, 16), (4, 16, 16), (16, 4, 16), (4, 4, 32),
                 (8, 4, 32), (4, 8, 32)]
    return [c for c in cands if tile_ok(c, w)]


def fingerprint(coords, shape, n_trans, eps):
    c = coords.detach().reshape(-1, coords.shape[-1])
    step = max(1, c.shape[0] // 4096)
    sub = c[::step].to(torch.float32).cpu().numpy().tobytes()
    h = hashlib.sha1(sub).hexdigest()[:16]
    return f"{tuple(shape)}|{c.shape[0]}|{n_trans}|{eps:g}|{h}"


def _time_opts(coords, shape, n_trans, eps, opts, reps, x, y):
    plan = _capi.RawPlan(shape, n_trans=n_trans, eps=eps, **opts)
    plan.setpts(coords)
    yo = torch.empty((n_trans, coords.shape[0]), dtype=torch.complex64, device=coords.device)
    xo = torch.empty_like(x)
    plan.sense_op(x, None, yo, n_trans, 1, 1.0)       # warm-up (allocates workspaces)
    plan.sense_adj(y, None, None, xo, n_trans, 1, 1.0)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.sense_op(x, None, yo, n_trans, 1, 1.0)
        plan.sense_adj(y, None, None, xo, n_trans, 1, 1.0)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    plan.close()
    return best


def autotune(coords, shape, n_trans=1, eps=1e-6, upsampfac=2.0, reps=3, budget_s=60.0,
             verbose=False):
    """Best plan options for (M, d) radian ``coords`` on the device.

    Returns a dict of RawPlan keyword options (tile, spread_group,
    interp_group, max_subproblem, spread_max) plus ``_ms`` (op+adj time of the
    winner) and ``_tried`` (candidates timed).
    """
    shape = tuple(int(s) for s in shape)
    ndim = len(shape)
    key = fingerprint(coords, shape, n_trans, eps)
    if key in _CACHE:
        return dict(_CACHE[key])
    path = os.environ.get("HIPNUFFT_TUNE_CACHE")
    disk = {}
    if path and os.path.exists(path):
        with open(path) as f:
            disk = json.load(f)
        if key in disk:
            _CACHE[key] = disk[key]
            return dict(disk[key])
    w = _kernel_width(eps, upsampfac)
    dev = coords.device
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.randn((n_trans, *shape), dtype=torch.complex64, generator=g).to(dev)
    y = torch.randn((n_trans, coords.shape[0]), dtype=torch.complex64, generator=g).to(dev)
    t_start = time.perf_counter()
    tried = []

    def timed(opts):
        full = dict(upsampfac=upsampfac, **opts)
        try:
            ms = _time_opts(coords, shape, n_trans, eps, full, reps, x, y)
        except ValueError:  # a combination the kernels reject
            ms = float("inf")
        tried.append((dict(opts), ms))
        if verbose:
            print(f"[autotune] {opts} -> {ms:.3f} ms", flush=True)
        return ms

    best = dict(tile=None, spread_group=0, interp_group=0, max_subproblem=0, spread_max=0,
                spread_method=0)
    best_ms = timed(best)
    stages = [
        ("tile", tile_candidates(ndim, w)),
        ("spread_method", [1, 2, 4] if ndim == 3 else []),
        ("spread_group", [4, 8, 16]),
        # 2D / 3D: 0 interpolates every transform of a work item in one
        # workgroup (items of at most 256 points in 2D, 512 in 3D), g >= 1
        # stages g transforms side by side; 1D: 4 by default
        ("interp_group", {1: [1, 2], 2: [1, 4], 3: [1]}[ndim]),
        ("max_subproblem", {1: [512, 1024, 2048], 2: [128, 512, 1024], 3: [64, 256, 1024]}[ndim]),
        ("spread_max", [256, 512, 1024, 2048, 4096]),
    ]
    for name, values in stages:
        for v in values:
            if time.perf_counter() - t_start > budget_s:
                break
            cand = dict(best, **{name: v})
            if name == "interp_group" and cand["tile"] is not None and not tile_ok(
                    cand["tile"], w, v):
                continue
            if cand == best:
                continue
            ms = timed(cand)
            if ms < best_ms:
                best, best_ms = cand, ms
    result = dict(best, _ms=best_ms, _tried=len(tried))
    if result["tile"] is not None:
        result["tile"] = list(result["tile"])
    _CACHE[key] = result
    if path:
        disk[key] = result
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(disk, f, indent=1)
        os.replace(tmp, path)
    return dict(result)


def plan_options(result):
    """RawPlan keyword options of an ``autotune`` result (private keys dropped)."""
    return {k: v for k, v in result.items() if not k.startswith("_")}
