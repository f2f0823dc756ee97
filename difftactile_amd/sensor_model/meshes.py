"""Tetrahedral pad meshes.

``dome_mesh`` restates the dome elastomer construction of
``difftactile/sensor_model/fem_sensor.py:280-322`` (Fibonacci-sphere layers,
scipy Delaunay tetrahedralisation of (x, layer height, z), outer-layer surface
triangles from the 2D Delaunay of (x, z)).  ``slab_mesh`` builds the box pad of
the slab workload (no reference counterpart): a vertex lattice with six
tetrahedra per hexahedral cell and the contact face triangulated two triangles
per cell.
"""
import numpy as np


def fibonacci_hemisphere(samples, scale=1.0):
    """Points on the upper hemisphere (``fem_sensor.py:280-291``)."""
    phi = np.pi * (np.sqrt(5.0) - 1.0)
    idx = np.arange(samples).astype(float)
    y = 1 - (idx / (samples - 1))
    radius = np.sqrt(1 - y * y)
    theta = phi * idx
    return scale * np.vstack((np.cos(theta) * radius, y, np.sin(theta) * radius)).T


def dome_mesh(n_node=200, n_layers=4, t_res=0.25, inner_radius=0.75):
    """(nodes, f2v, surface_f2v, layer_idxs) of the dome pad (``fem_sensor.py:293-322``)."""
    from scipy.spatial import Delaunay
    all_nodes, layer_h, layer_idx = [], [], []
    surface = None
    cur = 0
    for i in range(n_layers):
        rad = inner_radius + i * t_res
        n = int(n_node * (rad ** 2) / (inner_radius ** 2))
        pts = fibonacci_hemisphere(n, rad)
        all_nodes.append(pts)
        layer_h.append([rad] * n)
        layer_idx.append([i] * n)
        if i == n_layers - 1:
            surface = cur + Delaunay(np.vstack((pts[:, 0], pts[:, 2])).T).simplices
        cur += n
    layer_h = np.concatenate(layer_h)
    nodes = np.concatenate(all_nodes)
    f2v = Delaunay(np.array([nodes[:, 0], layer_h, nodes[:, 2]]).T).simplices
    return nodes, f2v.astype(np.int32), surface.astype(np.int32), np.concatenate(layer_idx).astype(np.int32)


_KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def slab_mesh(nx=40, nz=40, ny=6, h=1.0 / 128, origin=(0.0, 0.0, 0.0)):
    """Box pad of nx * ny * nz vertices, spacing h.

    Layer l = 0 (the top plane, y = origin_y + (ny-1) h) is the fixed base; layer
    ny-1 (the bottom plane, y = origin_y) is the contact face, whose outward
    direction is (0, -1, 0).  Each hex is split into the 6 Kuhn tetrahedra
    around its main diagonal; the contact face has 2 triangles per cell.
    Returns (nodes, f2v, surface_f2v, layer_idxs, out_direction).
    """
    ox, oy, oz = origin
    ii, ll, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    vid = lambda i, l, k: (i * ny + l) * nz + k
    nodes = np.stack([ox + ii * h, oy + (ny - 1 - ll) * h, oz + kk * h], -1).reshape(-1, 3)
    layer = ll.reshape(-1).astype(np.int32)
    tets = []
    for i in range(nx - 1):
        for l in range(ny - 1):
            for k in range(nz - 1):
                for perm in _KUHN:
                    c = [0, 0, 0]
                    corners = [vid(i, l, k)]
                    for ax in perm:
                        c[ax] += 1
                        corners.append(vid(i + c[0], l + c[1], k + c[2]))
                    tets.append(corners)
    f2v = np.asarray(tets, dtype=np.int32)
    tris = []
    l = ny - 1
    for i in range(nx - 1):
        for k in range(nz - 1):
            a, b = vid(i, l, k), vid(i + 1, l, k)
            c, d = vid(i, l, k + 1), vid(i + 1, l, k + 1)
            tris.append([a, b, d])
            tris.append([a, d, c])
    return (nodes.astype(np.float64), f2v, np.asarray(tris, dtype=np.int32), layer,
            np.array([0.0, -1.0, 0.0]))
