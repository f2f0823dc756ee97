"""Oracle: grid penalty contact between an MPM object and FEM pads.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Restates
``difftactile/sensor_model/fem_sensor.py:343-406`` (find_closest, find_sdf,
update_contact_force) and ``difftactile/tasks/grasp_plastic.py:191-299``
(clamp_grid, calculate_contact_force, check_collision, collision).

Pinning (tests/test_oracle_contact.py):
  find_closest_np -- float32 brute force with the declared rounding rule,
                     checked against a pure-Python scalar loop on small cases
                     and, away from ties, against a float64 argmin.
  find_sdf        -- closed forms: signed plane distance, barycentric
                     coordinates reproduce the projected point.
  contact_force   -- closed forms of the normal / tangential split; Coulomb
                     cap; the grasp and cable forms told apart by a separating
                     normal velocity and by a shear speed just below 1e-4.
  collision       -- Newton's third law: sum(grid_f)/dt + sum(ext_force) = 0;
                     the query velocity mv / (m + 1e-5) through closed-form
                     forces under a flat pad (rest, pure approach, pure shear).
"""
import numpy as np
import torch

THIRD_F32 = np.float32(1.0 / 3.0)


def find_closest_np(q, pos, tris, eps=1e-10):
    """Closest surface triangle by centroid distance (``fem_sensor.py:343-359``).

    Declared float32 rule (no fused multiply-add):
    c = f32(1/3) * ((p1 + p2) + p3); d = c - q;
    dist = sqrt(((d0*d0 + d1*d1) + d2*d2) + 1e-10f);
    strict ``<`` scan from (100.0, -1): the first minimum wins; -1 when no
    distance is below 100.
    """
    q = np.asarray(q, dtype=np.float32)
    pos = np.asarray(pos, dtype=np.float32)
    tris = np.asarray(tris, dtype=np.int64)
    p1, p2, p3 = pos[tris[:, 0]], pos[tris[:, 1]], pos[tris[:, 2]]
    c = THIRD_F32 * ((p1 + p2) + p3)                       # (T,3) float32
    out = np.empty(len(q), dtype=np.int32)
    chunk = 2048
    e = np.float32(eps)
    for s in range(0, len(q), chunk):
        d = c[None, :, :] - q[s:s + chunk, None, :]          # (Q,T,3)
        dd = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        dist = np.sqrt(dd + e)
        idx = np.argmin(dist, axis=1)
        best = dist[np.arange(len(idx)), idx]
        out[s:s + chunk] = np.where(best < np.float32(100.0), idx, -1)
    return out


def find_closest_scalar(qp, pos, tris, eps=1e-10):
    """Pure-Python restatement of the scalar loop (small cases only)."""
    f = np.float32
    best, best_k = f(100.0), -1
    for k in range(len(tris)):
        a, b, c = tris[k]
        cen = [f(THIRD_F32 * f(f(pos[a][i] + pos[b][i]) + pos[c][i])) for i in range(3)]
        d = [f(cen[i] - f(qp[i])) for i in range(3)]
        dd = f(f(f(d[0] * d[0]) + f(d[1] * d[1])) + f(d[2] * d[2]))
        dist = f(np.sqrt(f(dd + f(eps))))
        if dist < best:
            best, best_k = dist, k
    return best_k


def _normalized(n, eps):
    """Taichi ``normalized(eps)``: n / sqrt(n.n + eps)."""
    return n / torch.sqrt((n * n).sum(-1, keepdim=True) + eps)


def find_sdf(q, qv, idx, pos, vel, tris, out_dir, eps=1e-10, offset=0.0):
    """Signed distance, normal, relative velocity and contact flag (``fem_sensor.py:361-394``).

    n = normalized(cross(p2 - p1, p3 - p1)) * sign(n . out_direction);
    sdf = (q - p1) . n - offset; barycentric (u, v) of the projected point on
    (p3 - p1, p2 - p1); rel_v = qv - mean(vel of the 3 vertices);
    flag = sdf < 0 and u >= 0 and v >= 0 and u + v <= 1.  Returns
    (sdf, -n, rel_v, flag).
    """
    t = tris[idx]
    p1, p2, p3 = pos[t[:, 0]], pos[t[:, 1]], pos[t[:, 2]]
    n = _normalized(torch.cross(p2 - p1, p3 - p1, dim=-1), eps)
    sgn = torch.sign((n * out_dir).sum(-1, keepdim=True))
    n = sgn * n
    ds = q - p1
    sdf = (ds * n).sum(-1)
    proj = q - sdf[:, None] * n
    v0 = p3 - p1
    v1 = p2 - p1
    v2 = proj - p1
    dot00 = (v0 * v0).sum(-1)
    dot01 = (v0 * v1).sum(-1)
    dot02 = (v0 * v2).sum(-1)
    dot11 = (v1 * v1).sum(-1)
    dot12 = (v1 * v2).sum(-1)
    inv_d = 1 / (dot00 * dot11 - dot01 * dot01)
    u = (dot11 * dot02 - dot01 * dot12) * inv_d
    v = (dot00 * dot12 - dot01 * dot02) * inv_d
    sdf = sdf - offset
    rel_v = qv - (1 / 3) * (vel[t[:, 0]] + vel[t[:, 1]] + vel[t[:, 2]])
    flag = (sdf < 0) & (u >= 0) & (v >= 0) & (u + v <= 1)
    return sdf, -n, rel_v, flag


def contact_force(sdf, norm_v, rel_v, kn, kd, kt, mu_f, norm_eps=1e-11, clamp_vn=True):
    """Penalty contact force.

    Grasp form (``grasp_plastic.py:204-216``, the defaults):
    v_n = max(n . v_rel, 0); f_n = -(kn + kd v_n) sdf n;
    v_t = v_rel - (n . v_rel) n; |v_t| = sqrt(v_t.v_t + 1e-11);
    if |v_t| > 1e-4: f_t = v_t / |v_t| * min(kt |v_t|, mu_f |f_n|) (|f_n| without eps);
    f = f_n + f_t.

    Cable form (``cable_straightening.py:211-223``: ``clamp_vn=False,
    norm_eps=0``): v_n = n . v_rel (negative when separating) and
    |v_t| = sqrt(v_t.v_t) (Taichi's ``.norm()`` without eps).
    """
    vn_raw = (norm_v * rel_v).sum(-1)
    vn = torch.clamp(vn_raw, min=0) if clamp_vn else vn_raw
    fn = -(kn + kd * vn)[:, None] * sdf[:, None] * norm_v
    vt = rel_v - vn_raw[:, None] * norm_v
    s2 = (vt * vt).sum(-1) + norm_eps
    shear = torch.sqrt(s2.detach()) > 1e-4
    # the unselected branch is evaluated on 1 so that sqrt and the division
    # stay finite in reverse mode (no eps: |v_t| may be 0 there)
    vtn = torch.sqrt(torch.where(shear, s2, torch.ones_like(s2)))
    fnn = torch.sqrt((fn * fn).sum(-1))
    mag = torch.minimum(kt * vtn, mu_f * fnn)
    ft = torch.where(shear[:, None], vt / vtn[:, None] * mag[:, None], torch.zeros_like(vt))
    return fn + ft


def node_centres(prm, nodes):
    """Query point of a grid node: cell centre ((i+0.5) dx, ...) (``grasp_plastic.py:273``).

    Float32 rule: fl32(fl32(i) + 0.5f) * fl32(dx).
    """
    n = prm.n_grid
    i = nodes // (n * n)
    j = (nodes // n) % n
    k = nodes % n
    ijk = np.stack([i, j, k], axis=1).astype(np.float32)
    return (ijk + np.float32(0.5)) * np.float32(prm.dx)


def check_collision(prm, grid_m, sensors_pos, tris_list):
    """Closest triangle per occupied node and pad (``grasp_plastic.py:269-276``).

    Returns occupied node indices (ascending) and an int32 array
    ``(n_occ, n_sensors)``; unoccupied nodes keep -1 (``contact_idx.fill(-1)``).
    """
    m = grid_m.detach().cpu().numpy()
    nodes = np.nonzero(m > np.float32(prm.eps) if m.dtype == np.float32 else m > prm.eps)[0]
    q = node_centres(prm, nodes)
    idx = np.stack([find_closest_np(q, p.detach().cpu().numpy(), t)
                    for p, t in zip(sensors_pos, tris_list)], axis=1) if len(sensors_pos) else \
        np.zeros((len(nodes), 0), np.int32)
    return nodes, idx


def collision(prm, grid_mv, grid_m, nodes, contact_idx, sensors, kn, kd, kt, mu_f):
    """Penalty contact on occupied nodes (``grasp_plastic.py:279-299``).

    ``sensors`` is a list of (pos, vel, tris, out_dir).  Returns
    ``grid_f (n^3, 3)`` holding f*dt (``mpm_elastic.py:130-132``) and, per pad,
    the external force field with -f/3 added to each vertex of the chosen
    triangle (``fem_sensor.py:401-406``).
    """
    n = prm.n_grid
    dtype = grid_mv.dtype
    nodes_t = torch.as_tensor(nodes, dtype=torch.int64)
    q = torch.as_tensor(node_centres(prm, nodes)).to(dtype)
    qv = grid_mv[nodes_t] / (grid_m[nodes_t] + prm.eps)[:, None]
    grid_f = torch.zeros((n ** 3, 3), dtype=dtype)
    ext = []
    for s, (pos, vel, tris, out_dir) in enumerate(sensors):
        idx = torch.as_tensor(contact_idx[:, s], dtype=torch.int64)
        tris_t = torch.as_tensor(np.asarray(tris), dtype=torch.int64)
        ok = idx >= 0
        idx_c = torch.where(ok, idx, torch.zeros_like(idx))
        sdf, nv, rv, flag = find_sdf(q, qv, idx_c, pos, vel, tris_t, out_dir)
        flag = flag & ok
        f = contact_force(sdf, -nv, -rv, kn, kd, kt, mu_f)
        f = torch.where(flag[:, None], f, torch.zeros_like(f))
        grid_f = grid_f.index_add(0, nodes_t, f * prm.dt)
        t = tris_t[idx_c]
        e = torch.zeros_like(pos)
        for c in range(3):
            e = e.index_add(0, t[:, c], -f / 3)
        ext.append(e)
    return grid_f, ext


def clamp_grad(g, lim=1000.0):
    """``clamp_grid`` (``grasp_plastic.py:191-202``): clip an adjoint to [-1000, 1000]."""
    return torch.clamp(g, -lim, lim)
