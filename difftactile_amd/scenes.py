"""Synthetic scenes for the benchmark / parity workloads.

The reference samples particles from meshes (``obj_loader.py:25-65``:
voxelised lattice at ``particle_density`` samples per unit); these generators
produce the same kind of lattice directly, seeded, so no mesh library is
needed.  All outputs are numpy float32 / int32 in the reference's layouts.
"""
import numpy as np

from .params import DRUCKER_PRAGER, ELASTIC, Material, MPMParams


def lattice_cube(n_per_axis, spacing, centre, jitter=0.0, seed=0):
    """``n_per_axis**3`` particles on a cubic lattice centred at ``centre``.

    ``jitter`` is the half-width of a seeded uniform perturbation per axis
    (numpy PCG64).
    """
    rng = np.random.default_rng(seed)
    idx = np.arange(n_per_axis, dtype=np.float64)
    g = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), axis=-1).reshape(-1, 3)
    p = (g - (n_per_axis - 1) / 2.0) * spacing + np.asarray(centre, dtype=np.float64)
    if jitter:
        p = p + rng.uniform(-jitter, jitter, size=p.shape)
    return p.astype(np.float32)


def sphere_lattice(n_target, spacing, centre, jitter=0.0, seed=0):
    """Lattice points inside a sphere sized to hold about ``n_target`` particles.

    Returns (positions, radius).  Points are taken in lattice order, truncated
    to exactly ``n_target`` (nearest-to-centre first).
    """
    r = (3.0 * n_target / (4.0 * np.pi)) ** (1.0 / 3.0) * spacing
    m = int(np.ceil(r / spacing)) + 1
    idx = np.arange(-m, m + 1, dtype=np.float64)
    g = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), axis=-1).reshape(-1, 3) * spacing
    d2 = (g * g).sum(1)
    order = np.argsort(d2, kind="stable")[:n_target]
    order.sort()
    p = g[order]
    rng = np.random.default_rng(seed)
    if jitter:
        p = p + rng.uniform(-jitter, jitter, size=p.shape)
    p = p + np.asarray(centre, dtype=np.float64)
    radius = float(np.sqrt(d2[order].max()))
    return p.astype(np.float32), radius


def config1_elastic_cube(seed=0):
    """Elastic cube: 64^3 grid, 8000 particles (20^3 lattice at dx/2 spacing,
    +-0.25 dx seeded jitter), fixed-corotated E=5e4 nu=0.3 rho=1, gravity -9.8,
    dt=1e-4, resting just above the floor wall band, v0 = 0."""
    prm = MPMParams(n_grid=64, rho=1.0, dt=1e-4, gravity=(0.0, -9.8, 0.0), sticky_floor=True)
    dx = prm.dx
    x = lattice_cube(20, dx / 2, (0.5, 9.5 * dx, 0.5), jitter=0.25 * dx, seed=seed)
    n = len(x)
    mats = [Material(5e4, 0.3, ELASTIC)]
    return prm, mats, x, np.zeros((n, 3), np.float32), np.zeros(n, np.int32)


def config3_multi_material(n_particles=1_000_000, n_grid=128, dt=1e-5, seed=0):
    """Multi-material sphere: n_particles at 2 per dx per axis on an n_grid^3
    grid; a rigid core (r < 0.5 R, fixed corotated E=1e7) inside an
    elastoplastic Drucker-Prager shell (E=1e5, nu=0.3, friction 30 deg).
    Material ids: 0 core, 1 shell.  rho = 1000 so that node masses clear the
    1e-5 occupancy threshold (at rho = 1 a node holds ~4e-6); dt = 1e-5 keeps
    dt * sqrt(E/rho) / dx < 0.13 for the core.  Gravity -9.8."""
    prm = MPMParams(n_grid=n_grid, rho=1000.0, dt=dt, gravity=(0.0, -9.8, 0.0))
    dx = prm.dx
    centre = (0.5, 0.5, 0.5)
    x, R = sphere_lattice(n_particles, dx / 2, centre, jitter=0.25 * dx, seed=seed)
    d = np.linalg.norm(x.astype(np.float64) - np.asarray(centre), axis=1)
    mat_id = np.where(d < 0.5 * R, 0, 1).astype(np.int32)
    mats = [Material(1e7, 0.3, ELASTIC),
            Material(1e5, 0.3, DRUCKER_PRAGER, friction_deg=30.0)]
    rng = np.random.default_rng(seed + 1)
    v0 = np.tile(np.array([0.0, -0.5, 0.0], np.float32), (len(x), 1))
    v0 += rng.uniform(-0.05, 0.05, size=v0.shape).astype(np.float32)
    return prm, mats, x, v0.astype(np.float32), mat_id


def identity_F(n):
    return np.tile(np.eye(3, dtype=np.float32), (n, 1, 1))
