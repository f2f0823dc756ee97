"""Tactile marker set-up on the host (camera model of the pad).

``project_points_np`` restates ``project_points_to_pix``
(``difftactile/sensor_model/fisheye_model.py:32-51``); ``interp_weights``
restates the 5-nearest inverse-distance interpolation of ``init_cam_model``
(``fem_sensor.py:130-159``).  The reference detects its marker positions in
``init.png`` with OpenCV blob detection; here the markers are a seeded
synthetic set (a square lattice over the projected contact surface, the
``n_markers`` points nearest its centre), since OpenCV is not available.
"""
import numpy as np

FISHEYE = (8.627e-04, 173913.04, 3.320e02, 2.400e02, 2.0 * 0.01)   # f, m, cx, cy, z offset


def project_points_np(a, f=FISHEYE[0], m=FISHEYE[1], cx=FISHEYE[2], cy=FISHEYE[3],
                      z_offset=FISHEYE[4]):
    a = np.array(a, dtype=np.float64, copy=True)
    a[:, 2] += z_offset
    cos = a[:, 2] / np.linalg.norm(a, axis=1)
    theta = np.arccos(np.clip(cos, -1.0, 1.0))
    omega = np.arctan2(a[:, 1], a[:, 0]) + np.pi
    r = m * f * theta
    return np.stack([r * np.cos(omega) + cx, r * np.sin(omega) + cy], axis=1)


def camera_points(surface_nodes_local):
    """(x, z, y) reordering of pad-local surface points (``fem_sensor.py:137``)."""
    s = np.asarray(surface_nodes_local, dtype=np.float64)
    return np.stack([s[:, 0], s[:, 2], s[:, 1]], axis=1)


def synthetic_markers(cam_pts_2d, n_markers=136):
    """Marker pixels: a square lattice over the projected surface, the n_markers
    lattice points nearest to the centre of the projection."""
    lo, hi = cam_pts_2d.min(0), cam_pts_2d.max(0)
    side = int(np.ceil(np.sqrt(n_markers * 2)))
    gx = np.linspace(lo[0], hi[0], side + 2)[1:-1]
    gy = np.linspace(lo[1], hi[1], side + 2)[1:-1]
    g = np.stack(np.meshgrid(gx, gy, indexing="ij"), -1).reshape(-1, 2)
    c = cam_pts_2d.mean(0)
    order = np.argsort(np.linalg.norm(g - c, axis=1), kind="stable")[:n_markers]
    return g[np.sort(order)]


def interp_weights(markers_2d, cam_pts_2d, k=5):
    """k nearest projected surface points per marker and inverse-distance weights
    (``fem_sensor.py:143-153``)."""
    idx, w = [], []
    for mk in markers_2d:
        off = np.linalg.norm(mk[0:2] - cam_pts_2d, axis=1)
        sm = np.argpartition(off, k)[:k]
        inv = 1 / np.maximum(off[sm], 1e-12)
        idx.append(sm)
        w.append(inv / inv.sum())
    return np.asarray(idx, dtype=np.int32), np.asarray(w, dtype=np.float32)
