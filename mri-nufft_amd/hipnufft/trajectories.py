"""Trajectory generators and synthetic inputs for benchmarks and examples.

Restated from the reference's generators (``src/mrinufft/trajectories``):

* ``initialize_2D_radial``   <- ``inits/radial.py:12-43``
* ``initialize_2D_spiral``   <- ``inits/spiral.py:26-130`` (algebraic spirals
  with power >= 1; the centre patch for powers < 1 is not restated)
* ``initialize_3D_phyllotaxis_radial`` <- ``inits/radial.py:46-94`` built, like
  the reference, from a 5-revolution Archimedes spiral conified to the
  cone of half-angle pi/2 (``inits/cones.py:58-143``, ``tools/transforms.py:192-268``)
  and precessed over a golden-angle Fibonacci sphere
  (``tools/transforms.py:113-189``, Rodrigues rotation ``maths/rotations.py:88-131``)
* ``stack``                  <- ``tools/transforms.py:17-64``

plus the seeded phantom and coil maps used by the benchmark workloads (the
reference has no phantom generator).  Every array is float64 in
[-0.5, 0.5] like the reference (``KMAX = 0.5``, ``trajectories/utils.py:18``).
"""

from __future__ import annotations

import numpy as np

KMAX = 0.5
GOLDEN = np.pi * (3 - np.sqrt(5))  # "golden" tilt (trajectories/utils.py:663-703)


def _tilt(tilt, n):
    if tilt is None or tilt == "none":
        return 0.0
    if isinstance(tilt, (int, float)):
        return float(tilt)
    return {
        "uniform": 2 * np.pi / n,
        "intergaps": np.pi / n / 2,
        "inverted": np.pi / n + np.pi,
        "golden": GOLDEN,
        "mri-golden": np.pi * (np.sqrt(5) - 1) / 2,
    }[tilt]


def _rot2d(theta):
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s], [s, c]])


def _rotz(theta):
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def initialize_2D_radial(Nc, Ns, tilt="uniform", in_out=False):
    """(Nc, Ns, 2) radial spokes, each the previous one rotated by the tilt."""
    radius = KMAX * np.linspace(-1 if in_out else 0, 1, Ns)
    traj = np.zeros((Nc, Ns, 2))
    traj[0, :, 0] = radius
    rot = _rot2d(_tilt(tilt, Nc) / (1 + in_out)).T
    for i in range(1, Nc):
        traj[i] = traj[i - 1] @ rot
    return traj


def initialize_2D_spiral(Nc, Ns, tilt="uniform", in_out=False, nb_revolutions=1.0,
                         spiral="archimedes"):
    """(Nc, Ns, 2) algebraic spiral ``r = theta^n`` (n = 1 Archimedes, 2 Galilean)."""
    power = {"archimedes": 1.0, "galilean": 2.0}.get(spiral, spiral)
    power = float(power)
    if power < 1:
        raise NotImplementedError("spiral powers below 1 need the centre patch")
    angles = 2 * np.pi * nb_revolutions * np.linspace(-1 if in_out else 0, 1, Ns)
    radius = np.abs(angles) ** power
    shot = np.sign(angles) * np.abs(radius) * np.exp(1j * np.abs(angles))
    traj = np.zeros((Nc, Ns, 2))
    traj[0] = np.stack([shot.real, shot.imag], axis=-1)
    rot = _rot2d(_tilt(tilt, Nc) / (1 + in_out)).T
    for i in range(1, Nc):
        traj[i] = traj[i - 1] @ rot
    return KMAX * traj / np.max(np.linalg.norm(traj, axis=-1))


def _rodrigues(v1, v2, eps=1e-8):
    """Rotation matrix taking v1 to v2 (v1, v2 unit vectors)."""
    cr = np.cross(v1, v2)
    if np.linalg.norm(cr) < eps:
        return np.sign(np.dot(v1, v2)) * np.identity(3)
    c = np.dot(v1, v2)
    k = np.array([[0.0, -cr[2], cr[1]], [cr[2], 0.0, -cr[0]], [-cr[1], cr[0], 0.0]])
    return np.identity(3) + k + k @ k / (1 + c)


def initialize_3D_phyllotaxis_radial(Nc, Ns, nb_interleaves=1):
    """(Nc, Ns, 3) centre-out radial spokes on a golden-angle Fibonacci sphere."""
    spiral = initialize_2D_spiral(1, Ns, nb_revolutions=5)[0]
    norms = np.linalg.norm(spiral, axis=-1)
    # cone of half-angle pi/2 (polar angle pi/2 + pi/2): the spiral collapses
    # onto the -z axis, keeping its radii
    sin_pi = np.sin(np.pi)
    cone = np.stack([spiral[:, 0] * sin_pi, spiral[:, 1] * sin_pi, norms * np.cos(np.pi)], -1)
    z = np.linspace(-1, 1, Nc)
    phi = GOLDEN * np.arange(Nc)
    r = np.sqrt(1 - z ** 2)
    vec = np.stack([np.cos(phi) * r, np.sin(phi) * r, z], -1)
    ez = np.array([0.0, 0.0, 1.0])
    traj = np.empty((Nc, Ns, 3))
    for i in range(Nc):
        traj[i] = cone @ _rodrigues(ez, vec[i]).T
    traj = traj.reshape((-1, nb_interleaves, Ns, 3)).swapaxes(0, 1).reshape((Nc, Ns, 3))
    return traj


def stack(trajectory, nb_stacks, hard_bounded=True):
    """Stack a 2D trajectory along kz (no z-rotation), ``tools/transforms.py:17-64``."""
    Nc, Ns = trajectory.shape[:2]
    if trajectory.shape[-1] == 2:
        trajectory = np.concatenate([trajectory, np.zeros((Nc, Ns, 1))], axis=-1)
    flat = trajectory.reshape((Nc * Ns, 3))
    ub, lb = KMAX / nb_stacks, -KMAX / nb_stacks
    if hard_bounded:
        ub = max(np.max(flat[..., 2]), ub)
        lb = min(np.min(flat[..., 2]), lb)
    z_axis = np.linspace(-KMAX - lb, KMAX - ub, nb_stacks)
    out = np.repeat(flat[None], nb_stacks, axis=0)
    out[:, :, 2] = z_axis[:, None] + flat[None, :, 2]
    return out.reshape(nb_stacks * Nc, Ns, 3)


# ----------------------------------------------------------------------------
# synthetic images and coil maps
# ----------------------------------------------------------------------------
def ellipsoid_phantom(shape, seed=0, n_ellipsoids=10, noise=0.01):
    """Seeded Shepp-Logan-like phantom: nested ellipsoids plus N(0, noise) complex noise."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.linspace(-1, 1, s, endpoint=False) for s in shape], indexing="ij")
    img = np.zeros(shape, dtype=np.float64)
    # outer skull and brain
    outer = sum((g / a) ** 2 for g, a in zip(grids, [0.69, 0.92, 0.81][-len(shape):]))
    inner = sum((g / a) ** 2 for g, a in zip(grids, [0.6624, 0.874, 0.78][-len(shape):]))
    img += 1.0 * (outer <= 1) - 0.8 * (inner <= 1)
    for _ in range(n_ellipsoids - 2):
        ctr = rng.uniform(-0.4, 0.4, len(shape))
        ax = rng.uniform(0.05, 0.25, len(shape))
        amp = rng.uniform(-0.2, 0.3)
        img += amp * (sum(((g - c) / a) ** 2 for g, c, a in zip(grids, ctr, ax)) <= 1)
    img = img.astype(np.complex128)
    img += noise * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    return img.astype(np.complex64)


def gaussian_smaps(n_coils, shape, seed=0, sigma=0.3, n_bumps=4):
    """Coil maps: sums of complex Gaussian bumps (sigma in FOV units), unit RSS."""
    rng = np.random.default_rng(seed)
    axes = [np.linspace(-0.5, 0.5, s, endpoint=False) for s in shape]
    maps = np.zeros((n_coils, *shape), dtype=np.complex64)
    for c in range(n_coils):
        acc = np.zeros(shape, dtype=np.complex128)
        for _ in range(n_bumps):
            ctr = rng.uniform(-0.5, 0.5, len(shape))
            amp = rng.standard_normal() + 1j * rng.standard_normal()
            # separable Gaussian: outer product of 1D factors
            g = np.ones((1,) * len(shape))
            for d, (a, x) in enumerate(zip(axes, ctr)):
                sh = [1] * len(shape)
                sh[d] = -1
                g = g * np.exp(-((a - x) ** 2) / (2 * sigma ** 2)).reshape(sh)
            acc += amp * g
        maps[c] = acc
    rss = np.sqrt(np.sum(np.abs(maps.astype(np.complex128)) ** 2, axis=0))
    return (maps / rss).astype(np.complex64)
