"""Grid-mass deformation losses on the GPU (``difftactile/tasks/deform_loss.py``).

``DeformLoss`` keeps the reference ``Deform_Loss`` surface: a target density
grid and its nearest-point distance field (``update_target``, 61-86), a
density L1 term and an SDF-weighted mass term with their weights (99-111),
and the soft IoU (139-153).  The object's grid mass comes from the mass-only
scatter of ``compute_grid_m_kernel`` (``mpm_plastic.py:168-178``).  ``loss(x)``
is differentiable with respect to the particle positions, replacing
``compute_loss_kernel`` / ``compute_loss_kernel_grad`` (113-136).
"""
import torch

from . import _lib


class DeformLoss:
    def __init__(self, prm, device="cuda"):
        if prm.n_env > 1:
            raise ValueError("DeformLoss works on one environment")
        self.prm = prm
        self.c_prm = prm.c_struct()
        self.device = torch.device(device)
        n = prm.n_grid
        self.n_grid, self.dx_0 = n, prm.dx
        nn = n ** 3
        z = lambda *s: torch.zeros(*s, device=self.device)
        self.target_density = z(nn)
        self.target_sdf = z(nn)
        self.nearest_point = z(nn, 3)
        self.grid_mass = z(nn)
        self.sdf_weight = 1.0
        self.density_weight = 1.0
        self.target_iou = None
        self._scratch = torch.empty(5 * _lib.REDUCE_BLOCKS, device=self.device)
        self._tmp = (torch.empty(nn, device=self.device), torch.empty(nn, 3, device=self.device))

    # -- target -------------------------------------------------------------
    def initialize(self, grids):
        self.sdf_weight, self.density_weight = 1.0, 1.0
        self.load_target_density(grids)

    def load_target_density(self, grids):
        """Target density [n^3] (or (n, n, n)); computes its distance field
        and the IoU of the target with itself."""
        t = torch.as_tensor(grids, dtype=torch.float32, device=self.device).reshape(-1)
        if t.numel() != self.n_grid ** 3:
            raise ValueError("target density must have n_grid^3 entries")
        self.target_density.copy_(t)
        self.update_target()
        self.grid_mass.copy_(t)
        self.target_iou = self.iou()

    def update_target(self, n_sweeps=None):
        """Nearest-point distance field of the target (2 n_grid sweeps)."""
        sweeps = 2 * self.n_grid if n_sweeps is None else int(n_sweeps)
        _lib.call("dt_deform_target_sdf", self.n_grid, self.dx_0, self.target_density, sweeps,
                  self.target_sdf, self.nearest_point, *self._tmp, _lib.stream())

    def set_weights(self, sdf, density):
        self.sdf_weight, self.density_weight = float(sdf), float(density)

    # -- object mass and losses -----------------------------------------------
    def compute_grid_mass(self, x):
        """grid_mass = mass-only scatter of positions x [n][3]."""
        xx = x.detach().to(self.device, torch.float32).contiguous()
        self.grid_mass.zero_()
        _lib.call("dt_mpm_grid_mass", self.c_prm, xx.shape[0], xx, self.grid_mass, _lib.stream())
        return self.grid_mass

    def loss_terms(self, grid_mass=None):
        """(density loss, sdf loss) of grid_mass (default: the last scatter)."""
        m = self.grid_mass if grid_mass is None else grid_mass
        out = torch.empty(2, device=self.device)
        _lib.call("dt_deform_loss", m.numel(), m, self.target_density, self.target_sdf,
                  self._scratch, out, _lib.stream())
        return out

    def loss(self, x):
        """density_weight * sum |m - t| + sdf_weight * sum sdf m for the grid
        mass of positions x [n][3]; differentiable w.r.t. x."""
        return _DeformLossFn.apply(self, x)

    def iou(self, grid_mass=None):
        m = self.grid_mass if grid_mass is None else grid_mass
        out = torch.empty(5, device=self.device)
        _lib.call("dt_grid_iou_terms", m.numel(), m, self.target_density, self._scratch, out,
                  _lib.stream())
        ma, mb, s_ab, s_a, s_b = (float(v) for v in out.double().cpu())
        inter = s_ab / ma / mb
        union = s_a / ma + s_b / mb
        return inter / (union - inter)


class _DeformLossFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dl, x):
        xx = x.detach().to(dl.device, torch.float32).contiguous()
        m = dl.compute_grid_mass(xx).clone()
        terms = dl.loss_terms(m)
        ctx.dl = dl
        ctx.save_for_backward(xx, m)
        return terms[0] * dl.density_weight + terms[1] * dl.sdf_weight

    @staticmethod
    def backward(ctx, g):
        dl = ctx.dl
        xx, m = ctx.saved_tensors
        gm = torch.zeros_like(m)
        gl = g.detach().to(dl.device, torch.float32).reshape(1).contiguous()
        _lib.call("dt_deform_loss_adj", m.numel(), m, dl.target_density, dl.target_sdf,
                  dl.density_weight, dl.sdf_weight, gl, gm, _lib.stream())
        gx = torch.zeros_like(xx)
        _lib.call("dt_mpm_grid_mass_adj", dl.c_prm, xx.shape[0], xx, gm, gx, _lib.stream())
        return None, gx
