#!/usr/bin/env python
"""hipnufft benchmark: SENSE op+adj_op throughput and CG reconstruction latency.

Workload (the headline configuration): 3D golden-angle phyllotaxis radial
trajectory, 4096 spokes x 512 samples = 2,097,152 points, image 192^3
complex64 (seeded ellipsoid phantom + N(0, 0.01) noise), 16 Gaussian-bump
coil maps, eps = 1e-6, no density compensation.

* ``value``: SENSE op + adj_op pairs per second over all 16 coils, inputs
  resident in HBM (one step = ``op(x)`` then ``adj_op(y)``).
* ``cg_latency_ms``: p50 / p99 over repeated 20-iteration
  ``pinv_solver(y, optim="cg")`` reconstructions with the graph-captured
  iterations and FP64 inner products, after one untimed solve whose time
  (one-time initialisation included) is ``first_ms``; ``accuracy_reduced``
  holds the metric's CG accuracy bar checked on a reduced problem of the
  same kind (48^3, 131,072 phyllotaxis samples, 4 coils): the hipnufft CG
  against the same CG on an exact FP64 operator.
* ``accuracy``: relative L2 error of op (all coils, 256 random samples) and
  adj_op (256 random voxels) against an exact FP64 NDFT evaluated on the GPU.
* ``roofline``: the stage with the largest total time, its algorithmic HBM
  bytes per launch / its mean launch time (HIP events on the plan stream).
* ``cpu_baseline``: the numpy ES-NUFFT oracle on the host cores (one worker
  process per core, at most 16), one full 16-coil op + adj_op of the whole
  workload, not extrapolated.
* ``serving``: a bounded pass of the serving workload (mixed 2D / 3D CG
  requests), served vs plain per-call: requests/s and p50 / p99 latency.

``python bench.py --gpus N --steps K --warmup W`` (N must be 1).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mri-nufft_amd"))

PEAK_HBM_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def ndft_op_check(torch, traj_rad, img, smaps, y_gpu, inv_norm, nsamp, gen):
    """Exact FP64 SENSE forward at nsamp random samples via separable GEMMs."""
    dev = img.device
    M = traj_rad.shape[0]
    idx = torch.randperm(M, generator=gen)[:nsamp].to(dev)
    w = traj_rad[idx].double()  # (m, 3) radians
    N = img.shape[0]
    r = torch.arange(N, device=dev, dtype=torch.float64) - N // 2
    E = [torch.exp(-1j * w[:, d:d + 1] * r[None, :]) for d in range(3)]  # (m, N)
    C = smaps.shape[0]
    out = torch.empty((C, nsamp), dtype=torch.complex128, device=dev)
    for c in range(C):
        v = (smaps[c].to(torch.complex128) * img.to(torch.complex128))  # (z, y, x)
        t1 = torch.einsum("zyx,mx->mzy", v, E[2])
        t2 = torch.einsum("mzy,my->mz", t1, E[1])
        out[c] = torch.einsum("mz,mz->m", t2, E[0]) * inv_norm
    got = y_gpu[:, idx].to(torch.complex128)
    return float(torch.linalg.vector_norm(got - out) / torch.linalg.vector_norm(out))


def ndft_adj_check(torch, traj_rad, ksp, smaps, x_gpu, inv_norm, nvox, gen):
    """Exact FP64 SENSE adjoint at nvox random voxels."""
    dev = ksp.device
    shape = smaps.shape[1:]
    nvt = math.prod(shape)
    vox = torch.randperm(nvt, generator=gen)[:nvox].to(dev)
    coords = torch.stack(torch.unravel_index(vox, shape), -1).double() - shape[0] // 2
    C, M = ksp.shape
    acc = torch.zeros((C, nvox), dtype=torch.complex128, device=dev)
    chunk = 1 << 16
    for s in range(0, M, chunk):
        w = traj_rad[s:s + chunk].double()
        ph = torch.exp(1j * (w @ coords.T))  # (m, nvox)
        acc += ksp[:, s:s + chunk].to(torch.complex128) @ ph
    sm = smaps.reshape(C, -1)[:, vox].to(torch.complex128)
    ref = (torch.conj(sm) * acc).sum(0) * inv_norm
    got = x_gpu.reshape(-1)[vox].to(torch.complex128)
    return float(torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref))


class ExactSense:
    """Exact SENSE operator in FP64 (the NDFT evaluated separably with GEMMs
    on the device), with the interface hipnufft's CG uses.  Forward:
    y[c, m] = scale * sum_v S_c(v) x(v) exp(-i om_m . (v - N/2)); adjoint its
    conjugate transpose.  The reference solution of the CG accuracy figure."""

    def __init__(self, torch, om, shape, smaps, scale, chunk=8192):
        self.torch = torch
        self.shape = tuple(shape)
        self.img_full_shape = (1, *self.shape)
        self.squeeze_dims = False
        self.uses_density = False
        self.density = None
        self.device = om.device
        self.scale = float(scale)
        self.smaps = smaps.to(torch.complex128)
        self.chunk = chunk
        w = om.double()
        self.E = [torch.exp(-1j * w[:, d:d + 1] *
                            (torch.arange(n, device=om.device, dtype=torch.float64) - n // 2)[None])
                  for d, n in enumerate(self.shape)]  # (M, n_d) per axis
        self.L = None

    def _type2(self, v):
        torch = self.torch
        X, Y, Z = self.shape
        M = self.E[0].shape[0]
        out = torch.empty(M, dtype=torch.complex128, device=self.device)
        vz = v.reshape(X * Y, Z)
        for s in range(0, M, self.chunk):
            e = slice(s, s + self.chunk)
            t = vz @ self.E[2][e].transpose(0, 1).contiguous()                # (XY, m)
            t = (t.reshape(X, Y, -1) * self.E[1][e].transpose(0, 1).contiguous()[None]).sum(1)
            out[e] = (t * self.E[0][e].transpose(0, 1)).sum(0)
        return out

    def _type1(self, y):
        torch = self.torch
        X, Y, Z = self.shape
        acc = torch.zeros((X * Y, Z), dtype=torch.complex128, device=self.device)
        for s in range(0, y.shape[0], self.chunk):
            e = slice(s, s + self.chunk)
            wx = y[e, None] * torch.conj(self.E[0][e])                          # (m, X)
            wxy = (wx[:, :, None] * torch.conj(self.E[1][e])[:, None, :]).reshape(-1, X * Y)
            acc += wxy.transpose(0, 1).contiguous() @ torch.conj(self.E[2][e])
        return acc.reshape(X, Y, Z)

    def op(self, x):
        torch = self.torch
        x = x.reshape(self.shape).to(torch.complex128)
        out = torch.stack([self._type2(sc * x) for sc in self.smaps]) * self.scale
        return out.to(torch.complex64)[None]

    def adj_op(self, y):
        torch = self.torch
        y = y.reshape(self.smaps.shape[0], -1).to(torch.complex128)
        img = sum(torch.conj(sc) * self._type1(yc) for sc, yc in zip(self.smaps, y))
        return (img * self.scale).to(torch.complex64)[None]

    def data_consistency(self, x, y):
        return self.adj_op(self.op(x) - y.reshape(1, self.smaps.shape[0], -1))

    def get_lipschitz_cst(self, **_):
        return self.L


def reduced_cg_accuracy(torch, hipnufft):
    """The metric's CG accuracy bar on a reduced problem of the headline's
    kind (48^3, phyllotaxis 1024 x 128 = 131,072 samples, 4 coils): the
    20-iteration graph CG of the hipnufft operator against the same CG on
    the exact FP64 operator, same data (exact) and step size 1/L."""
    import numpy as np
    from hipnufft.optim import cg
    from hipnufft.trajectories import (ellipsoid_phantom, gaussian_smaps,
                                       initialize_3D_phyllotaxis_radial)

    shape, C = (48, 48, 48), 4
    traj = initialize_3D_phyllotaxis_radial(1024, 128).reshape(-1, 3).astype(np.float32)
    smaps = gaussian_smaps(C, shape, seed=2)
    op = hipnufft.get_operator("hipnufft")(traj, shape, n_coils=C, smaps=smaps)
    om = op.samples
    om = om if isinstance(om, torch.Tensor) else torch.from_numpy(np.asarray(om))
    om = om.to("cuda", torch.float32)
    ex = ExactSense(torch, om, shape, op.smaps, op.inv_norm_factor)
    x = torch.from_numpy(ellipsoid_phantom(shape, seed=3)).to("cuda")
    y = ex.op(x)
    L = op.get_lipschitz_cst(seed=4)
    ex.L = L
    op.get_lipschitz_cst = lambda **_: L
    got = cg(op, y, max_iter=20, graph=True)
    want = cg(ex, y, max_iter=20)
    rel = float(torch.linalg.vector_norm((got - want).reshape(-1).to(torch.complex128))
                / torch.linalg.vector_norm(want.reshape(-1).to(torch.complex128)))
    return {"problem": "48^3, phyllotaxis 1024 x 128 (131,072 samples), 4 coils, eps 1e-6, "
                       "20 iterations, exact data",
            "rel_l2_vs_exact_operator_cg": rel, "bar": 1e-4}


def stage_bytes(stage, op, T):
    """Algorithmic HBM bytes of one launch of a stage over T coils (complex64).

    Each array the stage must touch is counted once at its pruned size: the
    image (read once per launch, written once per batch), the T coil maps,
    the fine-grid lines each FFT pass reads and writes, the T x M k-space
    values and the sorted-point records (int32 index + int32 cell and float32
    offset per axis; spreading reads each record once per touched tile,
    about 2.6 times in 3D).
    """
    N = op.shape
    nf = op.raw_op.plan.info()["nf"]
    M = op.n_samples
    c8 = 8
    n_img = math.prod(N)
    pts = M * (4 + 8 * len(N))
    g_rows = N[0] * N[1] * nf[2]
    g_cols = N[0] * nf[1] * nf[2]
    g_full = nf[0] * nf[1] * nf[2]
    per_coil = {
        "rows_fwd": (n_img + g_rows) * c8,   # coil map read, grid rows written
        "cols_fwd": (g_rows + g_cols) * c8,
        "slabs_fwd": (g_cols + g_full) * c8,
        "interp": (g_full + M) * c8,         # grid read, samples written
        "spread": (M + g_full) * c8,         # samples read, grid written
        "slabs_adj": (g_full + g_cols) * c8,
        "cols_adj": (g_cols + g_rows) * c8,
        "rows_adj": (g_rows + n_img) * c8,   # grid rows read, coil map read
    }
    once = {"rows_fwd": n_img * c8, "rows_adj": n_img * c8, "interp": pts, "spread": 2.6 * pts}
    return per_coil[stage] * T + once.get(stage, 0)


def path_bytes(op, C):
    """Algorithmic bytes of one SENSE op+adj_op over C coils: image read and
    written, coil maps read twice, k-space written and read, coordinates read
    twice, and one write + one read of every coil's fine grid per direction."""
    N = math.prod(op.shape)
    G = math.prod(op.raw_op.plan.info()["nf"])
    M = op.n_samples
    d = len(op.shape)
    return 2 * N * 8 + 2 * C * N * 8 + 2 * C * M * 8 + 2 * M * d * 4 + 4 * C * G * 8


def measured_traffic(stage):
    """HBM bytes per launch of ``stage`` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE, WRITE_SIZE scaled by a plain copy
    kernel), or None."""
    path = os.path.join(ROOT, "profiles", "op_adj_3d_192_16coils_pmc_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    st = d.get("stages", {}).get(stage)
    if st is None:
        return None, None
    return float(st["bytes"]), os.path.relpath(path, ROOT)


def _cpu_coil_worker(path, coils):
    """One worker of the CPU baseline: SENSE op then adj_op of ``coils`` with
    the numpy ES-NUFFT oracle on the inputs saved (``.npy``, memory-mapped)
    in the folder ``path``; returns the
    summed coil-combined adjoint and the k-space checksum of its coils."""
    import numpy as np

    sys.path.insert(0, ROOT)
    from oracle import nufft as onufft

    om, img, smaps = (np.load(os.path.join(path, f"{k}.npy"), mmap_mode="r")
                      for k in ("om", "img", "smaps"))
    p = onufft.Plan(om, img.shape)
    inv = 1.0 / math.sqrt(math.prod(img.shape) * 2 ** img.ndim)
    acc = np.zeros(img.shape, dtype=np.complex128)
    ksum = 0.0
    for c in coils:
        y = p.execute(img * smaps[c]) * inv                         # op
        ksum += float(np.sum(np.abs(y)))
        acc += np.conj(smaps[c]) * p.execute_adjoint(y) * inv        # adj_op
    return acc, ksum


def cpu_baseline(traj32, img, smaps):
    """The numpy FP64 ES-NUFFT oracle (``oracle/nufft.py``) timed on the host
    cores: one full SENSE op + adj_op over every point and coil of the
    workload, the coils spread over worker processes (one per core, at most
    16: the CPU share of a one-GPU box), each single-threaded.  Nothing is
    extrapolated.  Process start and input loading are outside the timing."""
    import multiprocessing as mp
    import tempfile
    from concurrent.futures import ProcessPoolExecutor

    import numpy as np

    C = smaps.shape[0]
    cores = max(1, min(16, len(os.sched_getaffinity(0)), C))
    om = (traj32 * np.float32(2 * np.pi)).astype(np.float32)
    env_threads = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS")}
    os.environ.update(OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1")
    try:
        with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
            path = tmp
            for k, v in (("om", om), ("img", img), ("smaps", smaps)):
                np.save(os.path.join(tmp, f"{k}.npy"), v)
            groups = [list(range(C))[k::cores] for k in range(cores)]
            with ProcessPoolExecutor(cores, mp_context=mp.get_context("spawn")) as ex:
                list(ex.map(abs, range(cores)))                     # start the workers
                t0 = time.perf_counter()
                parts = list(ex.map(_cpu_coil_worker, [path] * cores, groups))
                elapsed = time.perf_counter() - t0
    finally:
        for k, v in env_threads.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    x = sum(a for a, _ in parts)
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next(line.split(":", 1)[1].strip() for line in f if line.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {
        "value": 1.0 / elapsed,
        "unit": "op+adj_op/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"numpy FP64 ES-NUFFT oracle (oracle/nufft.py): one full SENSE op + adj_op "
                   f"of the whole workload ({om.shape[0]} points, {C} coils, "
                   f"{tuple(img.shape)} image), coils spread over {cores} single-threaded "
                   f"worker processes, {elapsed:.1f} s wall; no extrapolation; CPU {cpu}; "
                   f"{len(os.sched_getaffinity(0))} cores visible"),
        "seconds": elapsed,
        "adj_checksum": float(np.abs(x).sum()),
    }


def stacked_benchmark(args, torch, np, hipnufft):
    """Stack-of-stars 128^3, 8 coils: 128 kz planes of a 2D radial pattern
    (256 spokes x 256 samples), 8,388,608 samples.  op+adj_op/s of the
    stacked operator (z FFT + one 2D plan launch of all planes) and of the
    full 3D operator on the equivalent 3D trajectory; rel. L2 difference of
    their forward outputs."""
    from hipnufft.stacked import stacked2traj3d
    from hipnufft.trajectories import gaussian_smaps

    shape, C, Z = (128, 128, 128), 8, 128
    nsp, ns = 256, 256
    r = 0.5 * np.linspace(-1, 1, ns, endpoint=False)
    th = np.pi * np.arange(nsp) / nsp
    s2 = np.stack([np.cos(th)[:, None] * r, np.sin(th)[:, None] * r], -1).reshape(-1, 2)
    s2 = s2.astype(np.float32)
    zi = np.arange(Z)
    smaps = gaussian_smaps(C, shape, seed=1)
    dev = torch.device("cuda", 0)
    gen = torch.Generator(device="cpu").manual_seed(2)
    x = (torch.randn(shape, generator=gen) + 1j * torch.randn(shape, generator=gen))
    x = x.to(torch.complex64).to(dev)
    res = {}
    t = time.perf_counter()
    st = hipnufft.get_operator("stacked-hipnufft")(s2, shape, smaps=smaps, n_coils=C, z_index=zi)
    res["stacked_setup_s"] = round(time.perf_counter() - t, 3)
    t = time.perf_counter()
    full = hipnufft.get_operator("hipnufft")(stacked2traj3d(s2, zi, Z), shape, n_coils=C,
                                             smaps=smaps)
    res["full_setup_s"] = round(time.perf_counter() - t, 3)

    def timed(op, steps, warmup):
        y = op.op(x)
        for _ in range(warmup):
            op.adj_op(op.op(x))
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            op.adj_op(op.op(x))
        e1.record()
        e1.synchronize()
        return y, e0.elapsed_time(e1) / steps

    ys, ms_s = timed(st, args.steps, args.warmup)
    yf, ms_f = timed(full, args.steps, args.warmup)
    diff = float(torch.linalg.vector_norm((ys.reshape(-1) - yf.reshape(-1)).to(torch.complex128))
                 / torch.linalg.vector_norm(yf.reshape(-1).to(torch.complex128)))
    log(f"stacked {ms_s:.2f} ms/op+adj, full 3D {ms_f:.2f} ms/op+adj, rel diff {diff:.2e}")
    return {"metric": "stacked op+adj_op/s (stack-of-stars 128³, 8 coils)",
            "value": round(1000.0 / ms_s, 3), "unit": "op+adj_op/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_s, 3),
            "higher_is_better": True, "dtype": "c64 (FP32)",
            "data": "synthetic (seeded image, Gaussian coil maps)",
            "config": {"workload": "stack-of-stars 128^3, 8 coils, 128 kz planes x 256 spokes "
                                   "x 256 samples (8,388,608 samples), eps 1e-6"},
            "full_3d": {"value": round(1000.0 / ms_f, 3), "ms_per_step": round(ms_f, 3)},
            "speedup_vs_full_3d": round(ms_f / ms_s, 2),
            "rel_l2_stacked_vs_full_3d": diff, **res}


def orc_benchmark(args, torch, np, hipnufft):
    """Off-resonance corrected SENSE on the 3D 192^3, 16-coil workload
    (phyllotaxis radial 4096 x 512): radial B0 map in [-300, 300] Hz, 10 ms
    readout per spoke, SVD interpolators with L = ``--orc``.  op+adj_op/s of
    the batched chain (all B*C*L transforms per launch) and of the
    reference's pattern on the same plain operator (L NUFFT calls on
    ``C_l x`` with the B weights applied outside); rel. L2 difference of
    their forward outputs."""
    from hipnufft.trajectories import (ellipsoid_phantom, gaussian_smaps,
                                       initialize_3D_phyllotaxis_radial)

    shape, C, L, NK = (192, 192, 192), 16, args.orc, 512
    traj = initialize_3D_phyllotaxis_radial(4096, NK).reshape(-1, 3).astype(np.float32)
    dev = torch.device("cuda", 0)
    x = torch.from_numpy(ellipsoid_phantom(shape, seed=0)).to(dev)
    smaps = gaussian_smaps(C, shape, seed=1)
    b0, mask = hipnufft.make_b0map(shape, b0range=(-300, 300))
    t_read = np.linspace(0, 10e-3, NK, dtype=np.float32)
    t0 = time.perf_counter()
    fo = hipnufft.get_operator("hipnufft")(traj, shape, n_coils=C, smaps=smaps)
    orc_op = fo.with_off_resonance_correction(t_read, b0_map=b0, mask=mask,
                                              interpolator={"name": "svd", "L": L})
    orc_op.op(x)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    Bt, Cm = orc_op.B, orc_op.C
    reps = fo.n_samples // NK

    def loop_op(v):
        y = None
        for l in range(L):
            yl = fo.op(Cm[l] * v).reshape(1, C, reps, NK) * Bt[:, l]
            y = yl if y is None else y + yl
        return y.reshape(1, C, -1)

    def loop_adj(k):
        k = k.reshape(1, C, reps, NK)
        img = None
        for l in range(L):
            il = fo.adj_op((k * Bt[:, l].conj()).reshape(1, C, -1)) * Cm[l].conj()
            img = il if img is None else img + il
        return img

    def timed(fwd, adj, steps, warmup):
        y = fwd(x)
        for _ in range(warmup):
            adj(fwd(x))
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            adj(fwd(x))
        e1.record()
        e1.synchronize()
        return y, e0.elapsed_time(e1) / steps

    yb, ms_b = timed(orc_op.op, orc_op.adj_op, args.steps, args.warmup)
    yl, ms_l = timed(loop_op, loop_adj, args.steps, args.warmup)
    diff = float(torch.linalg.vector_norm((yb.reshape(-1) - yl.reshape(-1)).to(torch.complex128))
                 / torch.linalg.vector_norm(yl.reshape(-1).to(torch.complex128)))
    log(f"orc batched {ms_b:.1f} ms/op+adj, L-loop {ms_l:.1f} ms/op+adj, rel diff {diff:.2e}")
    return {"metric": f"off-resonance op+adj_op/s (3D 192³, 16 coils, L={L})",
            "value": round(1000.0 / ms_b, 3), "unit": "op+adj_op/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_b, 3),
            "higher_is_better": True, "dtype": "c64 (FP32)",
            "data": "synthetic (ellipsoid phantom, Gaussian coil maps, radial B0 map)",
            "config": {"workload": "3D phyllotaxis radial 4096 x 512 (2,097,152 samples), "
                                   f"192^3, 16 coils, B0 in [-300, 300] Hz, 10 ms readout, "
                                   f"SVD interpolators L={L}, eps 1e-6"},
            "l_loop": {"value": round(1000.0 / ms_l, 3), "ms_per_step": round(ms_l, 3)},
            "speedup_vs_l_loop": round(ms_l / ms_b, 2),
            "rel_l2_batched_vs_l_loop": diff, "setup_s": round(setup, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cg-runs", type=int, default=20)
    ap.add_argument("--cg-iter", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="run one plain device copy kernel (float32 multiply by 1) over this "
                         "many MiB before the workload, as a known-size reference for HBM "
                         "byte counters")
    ap.add_argument("--serving", type=int, default=0, metavar="N",
                    help="instead of the operator benchmark, run the serving workload "
                         "(N mixed 2D/3D CG requests) served vs plain per-call")
    ap.add_argument("--serving-rate", type=float, default=20.0,
                    help="Poisson arrival rate (requests/s) of the serving workload")
    ap.add_argument("--serving-pass", type=int, default=12, metavar="N",
                    help="requests of the bounded serving pass reported under the operator "
                         "benchmark's 'serving' key (0: skip)")
    ap.add_argument("--stacked", action="store_true",
                    help="instead of the operator benchmark, time the stacked 2.5D operator on "
                         "a stack-of-stars 128^3, 8-coil workload against the full 3D operator "
                         "on the same samples")
    ap.add_argument("--orc", type=int, default=0, metavar="L",
                    help="instead of the operator benchmark, time the off-resonance corrected "
                         "operator with L interpolators on the 3D workload against L separate "
                         "operator calls")
    ap.add_argument("--autotune", action="store_true",
                    help="choose tile / group / split options with the autotuner "
                         "(tuning time is reported as part of plan setup)")
    ap.add_argument("--plan-opt", action="append", default=[], metavar="KEY=VALUE",
                    help="plan tuning option passed to the operator (spread_group, "
                         "interp_group, max_subproblem, spread_max, spread_method as "
                         "integers; tile as AxBxC)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build measures "
              "the single-GPU kernel path and has no multi-GPU mode.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import hipnufft

    if args.serving > 0:
        from hipnufft.serving import benchmark

        res = benchmark(args.serving, args.serving_rate, log=log)
        print(json.dumps({"metric": "served reconstructions/s (10-iteration CG, mixed 2D/3D)",
                          "value": round(res["saturation"]["served"]["requests_per_s"], 3),
                          "unit": "requests/s", "n_gpus": 1, "higher_is_better": True,
                          "data": "synthetic (seeded trajectories, images, coil maps)",
                          "serving": res}), flush=True)
        return
    if args.stacked:
        print(json.dumps(stacked_benchmark(args, torch, np, hipnufft)), flush=True)
        return
    if args.orc > 0:
        print(json.dumps(orc_benchmark(args, torch, np, hipnufft)), flush=True)
        return
    from hipnufft.trajectories import (ellipsoid_phantom, gaussian_smaps,
                                       initialize_3D_phyllotaxis_radial)

    if args.copy_calib_mib > 0:
        n = args.copy_calib_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_calib_mib} MiB read + {args.copy_calib_mib} MiB written")
        del src, dst

    shape = (192, 192, 192)
    C = 16
    log("building inputs")
    traj = initialize_3D_phyllotaxis_radial(4096, 512).reshape(-1, 3).astype(np.float32)
    img_np = ellipsoid_phantom(shape, seed=0)
    smaps_np = gaussian_smaps(C, shape, seed=1)
    dev = torch.device("cuda", 0)
    t = time.perf_counter()
    plan_opts = {}
    for k, v in (o.split("=", 1) for o in args.plan_opt):
        plan_opts[k] = tuple(int(x) for x in v.split("x")) if k == "tile" else int(v)
    if args.autotune:
        plan_opts["autotune"] = True
    op = hipnufft.get_operator("hipnufft")(traj, shape, n_coils=C, smaps=smaps_np, **plan_opts)
    if op.tuning is not None:
        log(f"autotune: {op.tuning}")
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t
    log(f"operator ready in {t_plan:.2f}s; n_trans={op.n_trans}; info={op.raw_op.plan.info()}")
    x = torch.from_numpy(img_np).to(dev)

    for _ in range(args.warmup):
        y = op.op(x)
        z = op.adj_op(y)
    torch.cuda.synchronize()
    log("warmup done")

    plan = op.raw_op.plan
    plan.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y = op.op(x)
        z = op.adj_op(y)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timings = plan.get_timing()
    plan.set_timing(False)
    elapsed = t1 - t0
    value = args.steps / elapsed
    log(f"timed {args.steps} op+adj in {elapsed:.3f}s -> {value:.2f}/s; stages={timings}")

    # roofline of the dominant stage
    dom = max(timings, key=lambda s: timings[s][0])
    ms, cnt = timings[dom]
    avg_s = ms / 1e3 / max(cnt, 1)
    T = op.n_trans
    alg = stage_bytes(dom, op, T)
    achieved = alg / avg_s / 1e9
    traffic, traffic_src = measured_traffic(dom)
    pbytes = path_bytes(op, C)
    stage_summary = {s: {"ms_total": round(v[0], 3), "launches": v[1],
                         "ms_per_launch": round(v[0] / max(v[1], 1), 4)}
                     for s, v in timings.items() if v[1]}

    acc = {}
    if not args.no_check:
        gen = torch.Generator().manual_seed(0)
        traj_rad = op.samples
        y2 = y.reshape(C, -1)
        acc["op_rel_l2_vs_fp64_ndft"] = ndft_op_check(torch, traj_rad, x, op.smaps, y2,
                                                      op.inv_norm_factor, 256, gen)
        acc["adj_rel_l2_vs_fp64_ndft"] = ndft_adj_check(torch, traj_rad, y2, op.smaps, z,
                                                        op.inv_norm_factor, 256, gen)
        acc["samples_checked"] = 256
        log(f"accuracy {acc}")

    # CG latency (20 iterations, graph-captured).  One untimed solve first:
    # the first solve of a process also pays one-time initialisation (kernel
    # loading, graph set-up); it is reported as first_ms
    lat = []
    y_obs = y.clone()
    first = None
    if args.cg_runs > 0:
        torch.cuda.synchronize()
        t = time.perf_counter()
        op.pinv_solver(y_obs, optim="cg", max_iter=args.cg_iter, graph=True)
        torch.cuda.synchronize()
        first = (time.perf_counter() - t) * 1e3
        log(f"first cg {first:.1f} ms")
    for r in range(args.cg_runs):
        torch.cuda.synchronize()
        t = time.perf_counter()
        rec = op.pinv_solver(y_obs, optim="cg", max_iter=args.cg_iter, graph=True)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t) * 1e3)
    lat_np = np.array(lat)
    cg = {"p50": float(np.percentile(lat_np, 50)) if lat else None,
          "p99": float(np.percentile(lat_np, 99)) if lat else None,
          "runs": len(lat), "iterations": args.cg_iter, "graph": True, "first_ms": first}
    if args.cg_runs > 0:
        # the step size every CG run estimates first: at most 10 power iterations
        # with the reference's stop rule (|norm_k - norm_k-1| < 1e-6, absolute);
        # stopped_early says whether a run of it met the rule before the tenth
        import warnings

        with warnings.catch_warnings(record=True) as caught:
            warnings.simplefilter("always")
            L = op.get_lipschitz_cst(seed=0)
        cg["lipschitz"] = {"value": L, "power_iterations": 10,
                           "stopped_early": not any("converge" in str(w.message)
                                                    for w in caught)}
    log(f"cg latency {cg}")

    # the same CG with Toeplitz gradients (FFT-only iterations); the untimed
    # first solve also builds the Toeplitz kernel
    lat_t = []
    rec_t = None
    first_t = None
    if args.cg_runs > 0:
        torch.cuda.synchronize()
        t = time.perf_counter()
        op.pinv_solver(y_obs, optim="cg", max_iter=args.cg_iter, graph=True, toeplitz=True)
        torch.cuda.synchronize()
        first_t = (time.perf_counter() - t) * 1e3
    for r in range(args.cg_runs):
        torch.cuda.synchronize()
        t = time.perf_counter()
        rec_t = op.pinv_solver(y_obs, optim="cg", max_iter=args.cg_iter, graph=True, toeplitz=True)
        torch.cuda.synchronize()
        lat_t.append((time.perf_counter() - t) * 1e3)
    if lat_t:
        lt = np.array(lat_t)
        cg["toeplitz"] = {"p50": float(np.percentile(lt, 50)), "p99": float(np.percentile(lt, 99)),
                          "runs": len(lat_t),
                          "first_ms": first_t,
                          "rel_l2_vs_nufft_cg": float(torch.linalg.vector_norm(rec_t - rec)
                                                      / torch.linalg.vector_norm(rec))}
        log(f"cg toeplitz {cg['toeplitz']}")
    if args.cg_runs > 0 and not args.no_check:
        t = time.perf_counter()
        cg["accuracy_reduced"] = reduced_cg_accuracy(torch, hipnufft)
        log(f"cg accuracy (reduced problem) {cg['accuracy_reduced']} in {time.perf_counter() - t:.1f}s")

    tuned = None if op.tuning is None else dict(op.tuning)
    serving = None
    if args.serving_pass > 0:
        from hipnufft.serving import benchmark as serving_benchmark

        # the headline operator's workspaces are not needed any more
        op = y = z = x = y_obs = rec = rec_t = plan = None
        torch.cuda.empty_cache()
        log(f"serving pass ({args.serving_pass} requests)")
        serving = serving_benchmark(args.serving_pass, args.serving_rate, log=log)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (bounded sample)")
        cpu = cpu_baseline(traj, img_np, smaps_np)
        log(f"cpu {cpu}")

    out = {
        "metric": "op+adj_op/s (3D 192³,16 coils) + p50/p99 CG-recon latency at rel err≤1e-4",
        "value": round(value, 3),
        "unit": "op+adj_op/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 accumulate), fp64 CG dots",
        "data": "synthetic (seeded phyllotaxis trajectory, ellipsoid phantom, Gaussian coil maps)",
        "config": {"workload": "3D phyllotaxis radial 4096x512 (2,097,152 pts), 192^3, "
                               "16 coils SENSE, eps 1e-6, sigma 2",
                   "n_trans": T, "plan_setup_s": round(t_plan, 3), "plan_opts": plan_opts,
                   "tuned": tuned},
        "cg_latency_ms": cg,
        "accuracy": acc,
        "stages": stage_summary,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": None if traffic is None else int(traffic),
                     "traffic_source": traffic_src,
                     "alg_bytes_per_launch": int(alg), "ms_per_launch": round(avg_s * 1e3, 4)},
        "path_roofline": {"alg_bytes_per_op_adj": int(pbytes),
                          "achieved": round(pbytes / (elapsed / args.steps) / 1e9, 1),
                          "unit": "GB/s",
                          "frac": round(pbytes / (elapsed / args.steps) / 1e9 / PEAK_HBM_GBS, 4)},
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
