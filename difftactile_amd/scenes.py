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


def _rot_z(deg):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def config5_gripper(n_particles=2_000_000, n_grid=256, dt=1e-5, n_substeps=1000,
                    pad_verts=(40, 40, 6), n_rope=2001, rope_length=30.0, close_speed=10.0,
                    seed=0):
    """Full gripper scene: two FEM pads on a parallel jaw, a position-based
    cable and an elastoplastic MPM object, coupled by penalty contact.

    Units of the grasp task (cm: space_scale 10, obj_scale 4, rho 4,
    ``tasks/grasp_plastic.py:35-46``), domain 10 per side, dx = 10 / n_grid.
    Object: n_particles at dx/2 spacing (+-0.25 dx seeded jitter) in a tall
    block (x : y : z = 2 : 5 : 1.5), an elastic core (E 5e4) inside a von Mises
    shell (E 5e4, yield 50), its bottom at mid-height of the pads.  Pads: two
    slab meshes of pad_verts vertices (spacing dx, 40 x 40 x 6 = 45,630 tets),
    stable Neo-Hookean E 1e5, nu 0.45, density 2000, their contact faces a
    quarter cell from the block's sides and facing each other along x.  The
    two pads are the same mesh turned 180 degrees apart, so their face
    diagonals run opposite ways; the contact query points (cell centres) sit
    at (0.45, 0.3) of a face cell, where the nearest triangle centroid is the
    triangle under the point for either diagonal, with no distance ties.
    They sit on the fingers of
    the reference's parallel-jaw gripper (``sensor_model/gripper_kinematics.py``:
    base 14 x 2, fingers 7.5 x 1.4; the grip width puts each finger frame on
    its pad's face), whose controls close the jaw at close_speed per finger.
    Cable: n_rope particles (n_rope - 1 segments) of rope_length along z under
    the block, through the jaw, touching the left pad; stretch and bending
    compliance 1e-6, the cable task's radius 0.25 and density 10
    (``tasks/cable_straightening.py:34-44``), particle 0 held.  Contact: the
    grasp task's kd / kt / friction with kn = 1e4 (grid), and the cable task's
    force and constants (kn 55.33, kd 239.97, kt 94.35, friction 4.9,
    cable_straightening.py:65-68) for the cable.  The reference has the object-pad and cable-pad halves in separate
    tasks (grasp_plastic.py:151-164, cable_straightening.py:166-173); the
    coupled substep composes them.
    Returns a dict with everything the coupled runner needs."""
    from .sensor_model.meshes import slab_mesh
    space = 10.0
    prm = MPMParams(n_grid=n_grid, space_scale=space, obj_scale=4.0, rho=4.0, dt=dt,
                    gravity=(0.0, -9.8, 0.0))
    dx = prm.dx
    sp = dx / 2
    # block lattice x : y : z = 2 : 5 : 1.5
    unit = (n_particles / (2 * 5 * 1.5)) ** (1.0 / 3.0)
    nx, nz = max(2, int(round(2 * unit))), max(2, int(round(1.5 * unit)))
    ny = -(-n_particles // (nx * nz))
    rng = np.random.default_rng(seed)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    p = np.stack([ii, jj, kk], -1).reshape(-1, 3)[:n_particles].astype(np.float64)
    pnx, pnz, pny = pad_verts
    h = dx
    pad_w = (pnx - 1) * h                     # pad extent along the block's y
    pad_d = (pnz - 1) * h                     # and z
    y_pad0 = (prm.bound + 2 + 0.05) * dx     # pads' lower edge: cell centres 0.45 h above
    y_block0 = y_pad0 + 0.5 * pad_w
    x = np.empty_like(p)
    x[:, 0] = 0.5 * space + (p[:, 0] - (nx - 1) / 2) * sp
    x[:, 1] = y_block0 + p[:, 1] * sp
    x[:, 2] = 0.5 * space + (p[:, 2] - (nz - 1) / 2) * sp
    x += rng.uniform(-0.25 * dx, 0.25 * dx, size=x.shape)
    # the block's left side on a grid node: the node below it, whose cell
    # centre (the contact query point) lies half a cell outside the block,
    # then sits a quarter cell behind the left pad's face, so both pads touch
    # the block from the first substep
    x[:, 0] += (np.ceil(x[:, 0].min() / dx) - x[:, 0].min() / dx) * dx
    half = 0.5 * (x[:, 0].max() - x[:, 0].min())
    centre = x.mean(0)
    r = np.linalg.norm((x - centre) / np.array([half, 0.5 * (x[:, 1].max() - x[:, 1].min()),
                                               0.5 * (x[:, 2].max() - x[:, 2].min())]), axis=1)
    mat_id = np.where(r < 0.5, 0, 1).astype(np.int32)
    from .params import VON_MISES
    mats = [Material(5e4, 0.3, ELASTIC), Material(5e4, 0.3, VON_MISES, yield_stress=50.0)]
    # pads: the slab's contact face (local y = 0, normal -y) turned to face +x
    # (left pad, rotation +90 about z) or -x (right pad, -90)
    gap = 0.25 * dx
    face_l = x[:, 0].min() - gap
    face_r = x[:, 0].max() + gap
    nodes0, f2v, tris, layer, _ = slab_mesh(pnx, pnz, pny, h=h, origin=(0.0, 0.0, 0.0))
    z0 = (np.floor((0.5 * space - 0.5 * pad_d) / h) + 0.2) * h   # cell centres 0.3 h in
    pads = []
    for side, face, deg, od in ((1, face_l, 90.0, (1.0, 0.0, 0.0)),
                                (2, face_r, -90.0, (-1.0, 0.0, 0.0))):
        R = _rot_z(deg)
        loc = nodes0.copy()
        loc[:, 0] -= 0.5 * pad_w              # centre the face on the block's x-y centre line
        w = loc @ R.T
        w[:, 0] += face
        w[:, 1] += y_pad0 + 0.5 * pad_w
        w[:, 2] += z0
        pads.append(dict(nodes=w, R=R, out_dir=np.asarray(od), face=face))
    E, nu, rho_pad = 1e5, 0.45, 2000.0
    mu, lam = E / 2 / (1 + nu), E * nu / (1 + nu) / (1 - 2 * nu)
    vol = pad_w * pad_d * ((pny - 1) * h)
    mass = rho_pad * vol / len(nodes0)
    # the jaw: ee on the pads' centre line, grip width placing finger 1 on the
    # left pad's face (finger 2 then sits on the right one by symmetry)
    from .sensor_model.gripper_kinematics import BASE_HEIGHT, BASE_WIDTH, FINGER_HEIGHT
    ee_t = (0.5 * (face_l + face_r), y_pad0 + 0.5 * pad_w + BASE_HEIGHT + FINGER_HEIGHT, z0 +
            0.5 * pad_d)
    grip_width = face_l - (ee_t[0] - BASE_WIDTH / 2)
    controls = np.zeros((n_substeps + 1, 7), np.float32)
    controls[:, 6] = close_speed
    # cable along z under the block's lower end, between the jaws, touching the
    # left pad (its surface a tenth of a cell into the pad at the start)
    r_rad, r_rho = 0.25, 10.0
    rest = rope_length / (n_rope - 1)
    rope_start = (face_l + r_rad - 0.1 * dx, y_pad0 + 0.25 * pad_w, 0.5 * space - 0.5 * rope_length)
    # markers: camera 1 unit behind each pad's face, looking at it (pad-local
    # +y toward the face, as config2_pad_press)
    inv_T = []
    for pd in pads:
        # slab-local (nodes0) -> world
        T = np.eye(4)
        T[:3, :3] = pd["R"]
        T[:3, 3] = pd["nodes"][0] - pd["R"] @ nodes0[0]
        Rwl = np.diag([1.0, -1.0, -1.0])
        cam = np.eye(4)
        cam[:3, :3] = Rwl
        cam[:3, 3] = -Rwl @ np.array([0.5 * pad_w, 1.0, 0.5 * pad_d])
        inv_T.append(cam @ np.linalg.inv(T))
    return dict(prm=prm, mats=mats, x=x.astype(np.float32), v=np.zeros((len(x), 3), np.float32),
                mat_id=mat_id, pads=[(pd["nodes"].astype(np.float32), f2v, tris, layer,
                                      pd["out_dir"]) for pd in pads],
                pad_mu=mu, pad_lam=lam, pad_mass=mass, inv_T=inv_T,
                pad_inv_rot=[pd["R"].T.copy() for pd in pads],
                gripper=dict(rot=(0.0, 0.0, 0.0), trans=ee_t, grip_width=grip_width, dt=dt),
                controls=controls,
                rope=dict(n=n_rope, p_rad=r_rad, p_rho=r_rho, rest_length=rest, start=rope_start,
                          ori=(0.0, 0.0, 1.0), compliance=1e-6),
                contact=dict(kn=1e4, kd=269.44, kt=154.78, mu_f=53.85),
                rope_contact=dict(kn=55.33, kd=239.97, kt=94.35, mu_f=4.9),
                n_substeps=n_substeps, target_com=(centre + np.array([0.0, 0.05, 0.0])))
