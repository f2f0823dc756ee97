"""Synthetic scenes for the benchmark / parity workloads.

The reference samples particles from meshes (``obj_loader.py:25-65``:
voxelised lattice at ``particle_density`` samples per unit); these generators
produce the same kind of lattice directly, seeded, so no mesh library is
needed.  All outputs are numpy float32 / int32 in the reference's layouts.
"""
import numpy as np

from .params import DRUCKER_PRAGER, ELASTIC, VON_MISES, Material, MPMParams


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


def config4_population(n_env=64, seed=0, dt=1e-5, pads=False):
    """Batched population: n_env independent objects with log-uniform particle
    counts in [5e3, 2e5] (seeded), each a ball at 2 particles per dx per axis
    on its own grid: 64^3 when it holds fewer than 3e4 particles, else 128^3.
    Materials alternate between fixed corotated (E=5e4) and von Mises (E=5e4,
    yield 50); rho = 1000; each ball starts with a seeded velocity of about
    0.5 downwards.  Returns one dict per environment (prm, mats, x, v, mat_id).

    pads: each environment also gets a slab pad of 16 x 16 x 9 vertices
    (10,800 tets, spacing 1/128, stable Neo-Hookean E = 1e5, nu = 0.45) resting
    a quarter cell above its ball, off the grid's cell centres by (0.21, 0.43)
    of its spacing, and driven down at 0.4 m/s (the C2 press
    rate); the dict then carries pad_mesh (nodes, f2v, tris, layer), pad_out,
    pad (mu, lam, mass) and pad_cv (the per-vertex control velocity)."""
    rng = np.random.default_rng(seed)
    counts = np.round(np.exp(rng.uniform(np.log(5e3), np.log(2e5), n_env))).astype(int)
    envs = []
    for e, n in enumerate(counts):
        n_grid = 64 if n < 30000 else 128
        prm = MPMParams(n_grid=n_grid, rho=1000.0, dt=dt, gravity=(0.0, -9.8, 0.0))
        dx = prm.dx
        x, _ = sphere_lattice(int(n), dx / 2, (0.5, 0.5, 0.5), jitter=0.25 * dx, seed=seed + e)
        mats = ([Material(5e4, 0.3, ELASTIC)] if e % 2 == 0
                else [Material(5e4, 0.3, VON_MISES, yield_stress=50.0)])
        v = np.tile(np.array([0.0, -0.5, 0.0], np.float32), (len(x), 1))
        v += rng.uniform(-0.1, 0.1, size=(1, 3)).astype(np.float32)
        env = dict(prm=prm, mats=mats, x=x, v=v.astype(np.float32),
                   mat_id=np.zeros(len(x), np.int32))
        if pads:
            from .sensor_model.meshes import slab_mesh
            h = 1.0 / 128
            top = float(np.max(x[:, 1]))
            # moved off the 128^3 cell centres by (0.21, 0.43) h (see config2_pad_press)
            nodes, f2v, tris, layer, od = slab_mesh(16, 16, 9, h=h,
                                                    origin=(0.5 - 7.29 * h, top + 0.25 * dx,
                                                            0.5 - 7.07 * h))
            E_p, nu_p = 1e5, 0.45
            vol = (15 * h) * (15 * h) * (8 * h)
            cv = np.zeros((len(nodes), 3), np.float32)
            cv[:, 1] = -0.4
            env.update(pad_mesh=(nodes.astype(np.float32), f2v, tris, layer), pad_out=od,
                       pad=dict(mu=E_p / 2 / (1 + nu_p),
                                lam=E_p * nu_p / (1 + nu_p) / (1 - 2 * nu_p),
                                mass=1000.0 * vol / len(nodes)),
                       pad_cv=cv)
        envs.append(env)
    return envs


def identity_F(n):
    return np.tile(np.eye(3, dtype=np.float32), (n, 1, 1))


def config2_pad_press(n_particles=200_000, n_grid=128, dt=1e-5, n_substeps=500, seed=0,
                      pad_verts=(40, 40, 6), press_substeps=None):
    """Slab pad pressing an elastoplastic block.

    The object constants of the grasp task (``tasks/grasp_plastic.py:35-46``:
    space_scale 10, obj_scale 4, rho 4), whose units its contact constants are
    given in; a full node then holds a mass of ~4, far above the 1e-5
    occupancy threshold.  Domain 10 per side (dx = 10/128).  Block: n_particles
    on a lattice at dx/2 spacing (+-0.25 dx seeded jitter) resting on the floor
    band, von Mises yield 50 (E = 5e4, nu = 0.3).  Pad: 40 x 40 x 6
    vertex slab (45,630 tets, 3,042 contact triangles), spacing dx,
    Neo-Hookean E = 1e5, nu = 0.45, density 2000 as the reference sensor's
    (``sensor_model/fem_sensor.py:37``; with it the grasp task's shear
    damping kt is stable on the pad's vertices); contact face 0.25 dx above the
    block top; the fixed top layer is driven down 2 mm at constant velocity
    over n_substeps.  Penalty contact kn = 1e4 (kd, kt, friction from the
    grasp task).
    pad_verts = (nx, nz, ny) vertices of the slab; press_substeps: the press
    rate is 2 mm over that many substeps (default n_substeps), so a shorter
    run of the same scene keeps the rate.
    Returns a dict with everything the coupled runner needs.
    """
    from .params import VON_MISES
    from .sensor_model.meshes import slab_mesh
    space = 10.0
    prm = MPMParams(n_grid=n_grid, space_scale=space, obj_scale=4.0, rho=4.0, dt=dt,
                    gravity=(0.0, -9.8, 0.0))
    dx = prm.dx
    spacing = dx / 2
    # block: nx * ny * nz lattice with ny = nx / 2.5 (a slab-like block)
    nx = int(round((n_particles * 2.5) ** (1.0 / 3.0)))
    ny = max(1, -(-n_particles // (nx * nx)))
    rng = np.random.default_rng(seed)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nx), indexing="ij")
    p = np.stack([ii, jj, kk], -1).reshape(-1, 3)[:n_particles].astype(np.float64)
    base_y = (prm.bound + 1) * dx
    x = np.empty_like(p)
    x[:, 0] = 0.5 * space + (p[:, 0] - (nx - 1) / 2) * spacing
    x[:, 1] = base_y + p[:, 1] * spacing
    x[:, 2] = 0.5 * space + (p[:, 2] - (nx - 1) / 2) * spacing
    x += rng.uniform(-0.25 * dx, 0.25 * dx, size=x.shape)
    top = x[:, 1].max()
    mats = [Material(5e4, 0.3, VON_MISES, yield_stress=50.0)]
    # pad
    h = dx
    pnx, pnz, pny = pad_verts
    # centred, then moved by (0.21, 0.43) h.  Centred, a pad of an odd vertex
    # count has its vertices exactly at the cell centres where the contact
    # queries sit, and the closest triangle and the inside test are decided by
    # rounding alone; other shifts leave queries equidistant from two triangle
    # centroids (a shift of (0.3, 0.1) h does).  With this one the nearest
    # centroid of a flat pad is 0.27 h away from a query and the next 0.47 h,
    # and every query is 0.15 h or more from a triangle edge.
    ox = 0.5 * space - 0.5 * (pnx - 1) * h + 0.21 * h
    oz = 0.5 * space - 0.5 * (pnz - 1) * h + 0.43 * h
    nodes, f2v, tris, layer, out_dir = slab_mesh(pnx, pnz, pny, h=h,
                                                 origin=(ox, top + 0.25 * dx, oz))
    E, nu, rho_pad = 1e5, 0.45, 2000.0
    mu, lam = E / 2 / (1 + nu), E * nu / (1 + nu) / (1 - 2 * nu)
    vol = ((pnx - 1) * h) * ((pnz - 1) * h) * ((pny - 1) * h)
    mass = rho_pad * vol / len(nodes)
    press = 0.2 / ((press_substeps or n_substeps) * dt)
    cv = np.zeros((n_substeps + 1, len(nodes), 3), np.float32)
    cv[:, :, 1] = -press
    # camera: 1 cm behind the contact face, looking at it (pad-local +y toward the face)
    cam_origin = np.array([0.5 * space, top + 0.25 * dx + 1.0, 0.5 * space])
    Rwl = np.diag([1.0, -1.0, -1.0])     # world -> local rotation (180 deg about x)
    inv_T = np.eye(4)
    inv_T[:3, :3] = Rwl
    inv_T[:3, 3] = -Rwl @ cam_origin
    return dict(prm=prm, mats=mats, x=x.astype(np.float32),
                v=np.zeros((len(x), 3), np.float32), mesh=(nodes, f2v, tris, layer, out_dir),
                pad_mu=mu, pad_lam=lam, pad_mass=mass, control_vel=cv, inv_T=inv_T,
                contact=dict(kn=1e4, kd=269.44, kt=154.78, mu_f=53.85), n_substeps=n_substeps)
