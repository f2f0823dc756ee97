"""Gripper-finger FEM pad (``difftactile/sensor_model/gripper_fem.py``, class ``FEMDomeSensor``).

Same kernels as the dome pad; the control velocity is per frame
(``control_vel[f, i]``) and ``update2`` adds ``control_vel[f+1] - control_vel[f]``
to the next velocity (``gripper_fem.py:403-418``).  Default material
E = 2.2e3, nu = 0.46 (``gripper_fem.py:43``).
"""
from .fem_sensor import FEMDomeSensor as _Dome


class FEMDomeSensor(_Dome):
    GRIPPER = True

    def __init__(self, dt=5e-5, sub_steps=50, init_img_path=None, *, E=0.22e4, nu=0.46, **kw):
        super().__init__(dt, sub_steps, init_img_path, E=E, nu=nu, **kw)

    def init(self, matrix):
        """``init(matrix)`` (``gripper_fem.py:105-111``)."""
        self.init_matrix(matrix)
