"""FEM tactile pad (``difftactile/sensor_model/fem_sensor.py``, class ``FEMDomeSensor``).

Same field layout (``pos``, ``vel``, ``virtual_pos``, ``external_force_field``
as ``[sub_steps, n_verts, 3]`` with ``.grad`` twins; ``B``, ``f2v``,
``contact_seg``, ``layer_id``, ``control_vel``, ``out_direction``, the pose
fields ``d_pos``, ``d_ori``, ``trans_*``, ``rot_*``, ``inv_*``, ``dtrans_h``) and
kernel methods (``init_pos``, ``update``/``update.grad``,
``update2``/``update2.grad``, ``set_pose_control``, ``set_pose_control_bp``,
``set_trans_h``, ``set_control_vel`` with their ``.grad``, ``set_vel``,
``reset_contact``, window and cache helpers).  Kernels run in
libdifftactile_hip.so.  The camera/marker model (``init_cam_model``, which
needs OpenCV blob detection on ``init.png``) is not part of this package.
"""
import numpy as np
import torch

from .. import _lib
from ..field import Field, Kernel
from ..params import lame
from .meshes import dome_mesh


def _rotvec_deg(rx, ry, rz):
    from scipy.spatial.transform import Rotation as R
    return R.from_rotvec(np.deg2rad([rx, ry, rz])).as_matrix()


class FEMDomeSensor:
    GRIPPER = False   # gripper pads: per-frame control velocity (gripper_fem.py:403-418)

    def __init__(self, dt=5e-5, sub_steps=50, init_img_path=None, *, mesh=None, E=0.8e4, nu=0.43,
                 rho=2000.0, mass=None, device="cuda"):
        self.sub_steps = int(sub_steps)
        self.dt = float(dt)
        self.dim = 3
        self.device = torch.device(device)
        self.N_node, self.N_t, self.t_res, self.inner_radius = 200, 4, 0.25, 0.75
        if mesh is None:
            nodes, f2v, surf, layers = dome_mesh(self.N_node, self.N_t, self.t_res,
                                                 self.inner_radius)
            out_local = np.array([0.0, 1.0, 0.0])
        else:
            nodes, f2v, surf, layers = mesh[:4]
            out_local = np.asarray(mesh[4]) if len(mesh) > 4 else np.array([0.0, 1.0, 0.0])
        self.all_nodes, self.all_f2v, self.surface_f2v, self.layer_idxs = nodes, f2v, surf, layers
        self._out_local = out_local
        self.n_verts = len(nodes)
        self.n_cells = len(f2v)
        self.num_triangles = len(surf)
        # mass per vertex (fem_sensor.py:35-38): area per inner node * thickness * rho
        self.dx = 2 * np.pi * self.inner_radius ** 2 / self.N_node
        self.vol = self.dx * self.t_res
        self.rho = rho
        self.mass = float(mass) if mass is not None else self.vol * self.rho
        self.eps = 1e-10
        d = self.device
        self.E_init = Field((), device=d, needs_grad=True)
        self.nu_init = Field((), device=d, needs_grad=True)
        self.mu = Field((), device=d, needs_grad=True)
        self.lam = Field((), device=d, needs_grad=True)
        self.set_material_params(E, nu)

        S, V = self.sub_steps, self.n_verts
        self.init_x = Field(V, (3,), device=d)
        self.init_x.from_numpy(nodes.astype(np.float32))
        self.layer_id = Field(V, dtype=torch.int32, device=d)
        self.layer_id.from_numpy(layers.astype(np.int32))
        self.surface_id_np = np.where(layers == (layers.max()))[0]
        self.f2v = Field(self.n_cells, (4,), dtype=torch.int32, device=d)
        self.f2v.from_numpy(f2v.astype(np.int32))
        self.contact_seg = Field(self.num_triangles, (3,), dtype=torch.int32, device=d)
        self.contact_seg.from_numpy(surf.astype(np.int32))
        self.virtual_pos = Field((S, V), (3,), device=d, needs_grad=True)
        self.pos = Field((S, V), (3,), device=d, needs_grad=True)
        self.vel = Field((S, V), (3,), device=d, needs_grad=True)
        self.B = Field(self.n_cells, (3, 3), device=d)
        self.external_force_field = Field((S, V), (3,), device=d, needs_grad=True)
        self.surf_f = Field(S, (3,), device=d, needs_grad=True)
        self.out_direction = Field((), (3,), device=d)
        # pose control (fem_sensor.py:93-107): d_pos / d_ori and the pose fields
        self.d_pos = Field((), (3,), device=d, needs_grad=True)
        self.d_ori = Field((), (3,), device=d, needs_grad=True)
        self.trans_world = Field((), (4, 4), device=d, needs_grad=True)
        self.trans_local = Field((), (4, 4), device=d, needs_grad=True)
        self.trans_h = Field((), (4, 4), device=d, needs_grad=True)
        self.inv_trans_h = Field((), (4, 4), device=d, needs_grad=True)
        self.dtrans_h = Field((), (4, 4), device=d, needs_grad=True)
        self.rot_world = Field((), (3, 3), device=d, needs_grad=True)
        self.rot_local = Field((), (3, 3), device=d, needs_grad=True)
        self.rot_h = Field((), (3, 3), device=d, needs_grad=True)
        self.inv_rot = Field((), (3, 3), device=d, needs_grad=True)
        self._pose_before = {}
        cv_shape = (S, V) if self.GRIPPER else V
        self.control_vel = Field(cv_shape, (3,), device=d, needs_grad=True)
        self.cache = dict()
        self._centroids = torch.zeros(self.num_triangles, 4, device=d)
        self.update = Kernel(self._update, self._update_grad)
        self.update2 = Kernel(self._update2, self._update2_grad)
        self.set_pose_control = Kernel(lambda: self._pose_step(0),
                                       lambda: self._pose_step_grad(0))
        self.set_pose_control_bp = Kernel(lambda: self._pose_step(2),
                                          lambda: self._pose_step_grad(2))
        self.set_trans_h = Kernel(self._set_trans_h, self._set_trans_h_grad)
        self.set_control_vel = Kernel(self._set_control_vel, self._set_control_vel_grad)

    # -- parameters --------------------------------------------------------
    def set_material_params(self, E, nu):
        """``set_material_params`` (``fem_sensor.py:189-193``)."""
        mu, lam = lame(E, nu)
        self.E_init[None], self.nu_init[None] = E, nu
        self.mu[None], self.lam[None] = mu, lam

    def _fem_params(self):
        p = _lib.DtFemParams()
        p.mu = float(self.mu[None])
        p.lam = float(self.lam[None])
        p.dt = self.dt
        p.mass = self.mass
        return p

    # -- set-up ------------------------------------------------------------
    def init(self, rot_x=0.0, rot_y=0.0, rot_z=0.0, t_dx=0.0, t_dy=0.0, t_dz=0.0):
        """Pose the pad (``fem_sensor.py:112-128``) and compute rest shapes."""
        rot = _rotvec_deg(rot_x, rot_y, rot_z)
        T = np.eye(4)
        T[:3, :3] = rot
        T[:3, 3] = (t_dx, t_dy, t_dz)
        self.init_matrix(T)

    def init_matrix(self, T):
        """Pose from a 4x4 matrix (``gripper_fem.py`` ``init(matrix)``; the
        world pose of ``init``, fem_sensor.py:112-128, with identity local
        parts)."""
        T = np.asarray(T, dtype=np.float64)
        self.trans_world[None] = T
        self.trans_local[None] = np.eye(4)
        self.trans_h[None] = T
        self.inv_trans_h[None] = np.linalg.inv(T)
        self.dtrans_h[None] = np.eye(4)
        self.rot_world[None] = T[:3, :3]
        self.rot_local[None] = np.eye(3)
        self.rot_h[None] = T[:3, :3]
        self.inv_rot[None] = np.linalg.inv(T[:3, :3])
        self.init_pos()

    # -- pose control (fem_sensor.py:196-266) ------------------------------
    def _pose_pack(self, grad=False):
        parts = []
        for name in _lib.DOME_POSE:
            f = getattr(self, name)
            parts.append((f.grad if grad else f).data.reshape(-1))
        return torch.cat(parts).contiguous()

    def _pose_unpack(self, buf, grad=False):
        for name, (off, shape) in _lib.DOME_POSE.items():
            f = getattr(self, name)
            (f.grad if grad else f).data.copy_(buf[off:off + shape[0] * shape[1]].view(shape))

    def _step_time(self):
        return self.dt * (self.sub_steps - 1)

    def _d6(self, grad=False):
        if grad:
            return torch.cat([self.d_pos.grad.data, self.d_ori.grad.data]).contiguous()
        return torch.cat([self.d_pos.data, self.d_ori.data]).contiguous()

    def _pose_call(self, mode, d6):
        pin = self._pose_pack()
        self._pose_before[mode] = (pin, d6)
        pout = torch.empty_like(pin)
        _lib.call("dt_dome_pose_control", mode, d6, self._step_time(), pin, pout, _lib.stream())
        self._pose_unpack(pout)

    def _pose_call_grad(self, mode):
        pin, d6 = self._pose_before[mode]
        g = self._pose_pack(grad=True)
        g6 = torch.zeros(6, device=self.device)
        _lib.call("dt_dome_pose_control_adj", mode, d6, self._step_time(), pin, g, g, g6,
                  _lib.stream())
        self._pose_unpack(g, grad=True)
        return g6

    def _pose_step(self, mode):
        """``set_pose_control`` (mode 0, local frame, 229-245) /
        ``set_pose_control_bp`` (mode 2, 248-256) with the current d_pos, d_ori."""
        self._pose_call(mode, self._d6())

    def _pose_step_grad(self, mode):
        g6 = self._pose_call_grad(mode)
        self.d_pos.grad.data += g6[:3]
        self.d_ori.grad.data += g6[3:]

    def _set_trans_h(self, d_pos, d_ori):
        """``set_trans_h`` (217-227): the world-frame pose step."""
        d6 = torch.tensor(np.concatenate([np.asarray(d_pos, np.float32),
                                          np.asarray(d_ori, np.float32)]), device=self.device)
        self._pose_call(1, d6)

    def _set_trans_h_grad(self, d_pos=None, d_ori=None):
        """Adjoint of set_trans_h: the pose fields' gradients; returns
        d/d(d_pos, d_ori) as a device (6,) (they are kernel arguments)."""
        return self._pose_call_grad(1)

    def _set_control_vel(self, f):
        """``set_control_vel`` (258-266): control_vel from dtrans_h and
        virtual_pos[f]."""
        pose = self._pose_pack()
        cv = self.control_vel.data[f] if self.GRIPPER else self.control_vel.data
        _lib.call("dt_dome_control_vel", self.n_verts, self.virtual_pos.data[f], pose,
                  self._step_time(), cv, _lib.stream())

    def _set_control_vel_grad(self, f):
        pose = self._pose_pack()
        gpose = self._pose_pack(grad=True)
        gcv = self.control_vel.grad.data[f] if self.GRIPPER else self.control_vel.grad.data
        scratch = torch.empty(12 * _lib.REDUCE_BLOCKS, device=self.device)
        _lib.call("dt_dome_control_vel_adj", self.n_verts, self.virtual_pos.data[f], pose,
                  gcv.contiguous(), self._step_time(), scratch, gpose,
                  self.virtual_pos.grad.data[f], _lib.stream())
        self._pose_unpack(gpose, grad=True)

    def init_pos(self):
        """``init_pos`` (``fem_sensor.py:324-340``): world rest positions, B, out_direction."""
        T = self.trans_h.data
        p0 = self.init_x.data @ T[:3, :3].T + T[:3, 3]
        self.pos.data[0] = p0
        self.virtual_pos.data[0] = p0
        _lib.call("dt_fem_init_B", self.n_cells, _lib.ptr(self.f2v.data, torch.int32),
                  _lib.ptr(self.pos.data[0]), _lib.ptr(self.B.data), _lib.stream())
        self.out_direction.data.copy_(self.rot_h.data @ torch.tensor(
            self._out_local, dtype=torch.float32, device=self.device))

    def out_dir_host(self):
        """Outward direction as a host float32[3] (the contact entry points read it on the host)."""
        return np.ascontiguousarray(self.out_direction.to_numpy(), dtype=np.float32)

    def set_vel(self, f):
        """``set_vel`` (``fem_sensor.py:212-215``): vel[f] = control_vel."""
        cv = self.control_vel.data[f] if self.GRIPPER else self.control_vel.data
        self.vel.data[f] = cv

    def reset_contact(self):
        self.external_force_field.fill(0.0)
        self.surf_f.fill(0.0)

    # -- kernels -----------------------------------------------------------
    def _update(self, f):
        _lib.call("dt_fem_update", self._fem_params(), self.n_cells,
                  _lib.ptr(self.f2v.data, torch.int32), _lib.ptr(self.B.data),
                  _lib.ptr(self.pos.data[f]), _lib.ptr(self.vel.data[f]), _lib.stream())

    def _update_grad(self, f):
        _lib.call("dt_fem_update_adj", self._fem_params(), self.n_cells,
                  _lib.ptr(self.f2v.data, torch.int32), _lib.ptr(self.B.data),
                  _lib.ptr(self.pos.data[f]), _lib.ptr(self.vel.grad.data[f]),
                  _lib.ptr(self.pos.grad.data[f]), _lib.stream())

    def _cv(self, f):
        if self.GRIPPER:
            return self.control_vel.data[f], self.control_vel.data[f + 1]
        return self.control_vel.data, None

    def _gcv(self, f):
        if self.GRIPPER:
            return self.control_vel.grad.data[f], self.control_vel.grad.data[f + 1]
        return self.control_vel.grad.data, None

    def _update2(self, f):
        cv, cv1 = self._cv(f)
        _lib.call("dt_fem_update2", self._fem_params(), self.n_verts,
                  _lib.ptr(self.layer_id.data, torch.int32), _lib.ptr(self.pos.data[f]),
                  _lib.ptr(self.vel.data[f]), _lib.ptr(self.virtual_pos.data[f]),
                  _lib.ptr(self.external_force_field.data[f]), _lib.ptr(cv), _lib.ptr(cv1),
                  _lib.ptr(self.pos.data[f + 1]), _lib.ptr(self.vel.data[f + 1]),
                  _lib.ptr(self.virtual_pos.data[f + 1]), _lib.stream())

    def _update2_grad(self, f):
        gcv, gcv1 = self._gcv(f)
        _lib.call("dt_fem_update2_adj", self._fem_params(), self.n_verts,
                  _lib.ptr(self.layer_id.data, torch.int32), _lib.ptr(self.pos.grad.data[f + 1]),
                  _lib.ptr(self.vel.grad.data[f + 1]), _lib.ptr(self.virtual_pos.grad.data[f + 1]),
                  _lib.ptr(self.pos.grad.data[f]), _lib.ptr(self.vel.grad.data[f]),
                  _lib.ptr(self.virtual_pos.grad.data[f]),
                  _lib.ptr(self.external_force_field.grad.data[f]), _lib.ptr(gcv), _lib.ptr(gcv1),
                  _lib.stream())

    def centroids(self, f):
        _lib.call("dt_contact_centroids", self.num_triangles,
                  _lib.ptr(self.contact_seg.data, torch.int32), _lib.ptr(self.pos.data[f]),
                  _lib.ptr(self._centroids), _lib.stream())
        return self._centroids

    def get_external_force(self, f):
        """``get_external_force`` (``fem_sensor.py:268-274``)."""
        seg = self.contact_seg.data.long()
        e = self.external_force_field.data[f]
        self.surf_f.data[f] += (1 / 3) * self.dx * e[seg].sum((0, 1))

    # -- window / cache helpers (fem_sensor.py:460-551) ---------------------
    def copy_frame(self, source, target):
        for fld in (self.pos, self.vel, self.virtual_pos):
            fld.data[target] = fld.data[source]

    def copy_grad(self, source, target):
        for fld in (self.pos, self.vel, self.virtual_pos):
            fld.grad.data[target] = fld.grad.data[source]

    def memory_to_cache(self, t):
        name = f"{t:06d}"
        self.cache[name] = {k: getattr(self, k).data[0].detach().to("cpu", copy=True)
                            for k in ("pos", "vel", "virtual_pos")}
        self.cache[name]["trans_h"] = self.trans_h.data.detach().cpu().clone()
        self.cache[name]["rot_h"] = self.rot_h.data.detach().cpu().clone()
        self.copy_frame(self.sub_steps - 1, 0)

    def memory_from_cache(self, t):
        name = f"{t:06d}"
        self.copy_frame(0, self.sub_steps - 1)
        self.copy_grad(0, self.sub_steps - 1)
        self.clear_step_grad(self.sub_steps - 1)
        c = self.cache[name]
        self.trans_h.data.copy_(c["trans_h"])
        self.rot_h.data.copy_(c["rot_h"])
        for k in ("pos", "vel", "virtual_pos"):
            getattr(self, k).data[0].copy_(c[k])

    def clear_loss_grad(self):
        """``clear_loss_grad`` (fem_sensor.py:530-546)."""
        for fld in (self.E_init, self.nu_init, self.mu, self.lam, self.control_vel, self.d_pos,
                    self.d_ori, self.rot_h, self.rot_local, self.inv_rot, self.trans_h,
                    self.trans_local, self.inv_trans_h, self.dtrans_h):
            fld.grad.fill(0.0)

    def clear_step_grad(self, f):
        self.surf_f.grad.fill(0.0)
        self.external_force_field.grad.fill(0.0)
        for fld in (self.pos, self.vel, self.virtual_pos):
            fld.grad.data[:f] = 0.0
