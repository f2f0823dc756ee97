"""Oracle: FEM tactile pad (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

Restates ``difftactile/sensor_model/fem_sensor.py`` (dome) and
``gripper_fem.py`` (gripper pad) in torch on the CPU.

Pinning (tests/test_oracle_fem.py):
  init_B   -- B Dm = I (closed form).
  update   -- stable Neo-Hookean P(I) = 0; P = dPsi/dF by FP64 central
              differences of Psi = mu/2 (I_C - 3) + lam/2 (J - alpha)^2
              - mu/2 log(I_C + 1) (Smith et al. 2018); per-tet force sum = 0
              and torque sum = 0 (translation / rotation invariance).
  update2  -- closed form per vertex (Dirichlet base layer, explicit Euler).
"""
import torch


def init_B(pos0, f2v):
    """Rest-shape inverse ``B = [a-d, b-d, c-d]^-1`` (``fem_sensor.py:334-338``).

    ``pos0`` are world positions after the pad transform, as in the reference.
    """
    p = pos0[f2v]                                  # (T,4,3)
    Dm = torch.stack([p[:, 0] - p[:, 3], p[:, 1] - p[:, 3], p[:, 2] - p[:, 3]], dim=-1)
    return torch.linalg.inv(Dm)


def tet_forces(pos, f2v, B, mu, lam):
    """Per-tet nodal forces of the stable Neo-Hookean model (``fem_sensor.py:411-440``).

    D = [a-d, b-d, c-d]; V = |det D| / 6 (current volume); F = D B;
    J = det F; I_C = tr(F^T F); dJ/dF = [f1 x f2, f2 x f0, f0 x f1];
    alpha = 1 + 0.75 mu / lam; P = mu (1 - 1/(I_C+1)) F + lam (J - alpha) dJ/dF;
    H = -V P B^T.  Forces: column k of H on vertex k (k<3), minus their sum on
    vertex 3.  Returns (T, 4, 3).
    """
    p = pos[f2v]
    D = torch.stack([p[:, 0] - p[:, 3], p[:, 1] - p[:, 3], p[:, 2] - p[:, 3]], dim=-1)
    Vol = torch.abs(torch.linalg.det(D)) / 6
    F = D @ B
    J = torch.linalg.det(F)
    IC = (F * F).sum((-1, -2))
    f0, f1, f2 = F[..., 0], F[..., 1], F[..., 2]
    dJdF = torch.stack([torch.cross(f1, f2, dim=-1), torch.cross(f2, f0, dim=-1),
                        torch.cross(f0, f1, dim=-1)], dim=-1)
    alpha = 1 + 0.75 * mu / lam
    P = mu * (1 - 1 / (IC + 1))[:, None, None] * F + lam * (J - alpha)[:, None, None] * dJdF
    H = -Vol[:, None, None] * (P @ B.transpose(-1, -2))
    f012 = H.transpose(-1, -2)                     # rows = force on vertex k
    f3 = -f012.sum(1, keepdim=True)
    return torch.cat([f012, f3], dim=1)


def update(pos, vel, f2v, B, mu, lam, dt, mass):
    """``update(f)``: vel[f] += dt * force / mass, scattered per tet (``fem_sensor.py:408-440``)."""
    forces = tet_forces(pos, f2v, B, mu, lam)
    dv = (dt / mass) * forces
    return vel.index_add(0, f2v.reshape(-1), dv.reshape(-1, 3))


def update2_dome(pos, vel, vpos, ext_f, layer_id, control_vel, dt, mass):
    """Dome ``update2`` (``fem_sensor.py:443-457``).

    v = vel + dt ext_f / mass; base layer (layer_id == 0) takes control_vel;
    vel' = v; pos' = pos + dt v; virtual_pos' = virtual_pos + dt control_vel.
    """
    v = vel + dt * ext_f / mass
    v = torch.where((layer_id == 0)[:, None], control_vel, v)
    return pos + dt * v, v, vpos + dt * control_vel


def update2_gripper(pos, vel, vpos, ext_f, layer_id, cv_f, cv_f1, dt, mass):
    """Gripper-pad ``update2`` (``gripper_fem.py:403-418``).

    As the dome, but the control velocity is per frame and
    vel' = v + (control_vel[f+1] - control_vel[f]).
    """
    v = vel + dt * ext_f / mass
    v = torch.where((layer_id == 0)[:, None], cv_f, v)
    return pos + dt * v, v + (cv_f1 - cv_f), vpos + dt * cv_f


def neo_hookean_energy(F, mu, lam):
    """Stable Neo-Hookean energy density whose gradient is P (used to pin ``tet_forces``)."""
    J = torch.linalg.det(F)
    IC = (F * F).sum((-1, -2))
    alpha = 1 + 0.75 * mu / lam
    return 0.5 * mu * (IC - 3) + 0.5 * lam * (J - alpha) ** 2 - 0.5 * mu * torch.log(IC + 1)
