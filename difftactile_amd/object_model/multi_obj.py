"""Multi-material MLS-MPM object (``difftactile/object_model/multi_obj.py``, class ``MPMObj``).

Per-particle material id ``titles`` selects (mu, lamda) from 3-entry tables
(``multi_obj.py:65-74,227-252``); ids come from height bands of the rest shape
(``preprocess_obj``, ``multi_obj.py:104-112``).  Gravity defaults to -2
(``multi_obj.py:37``).
"""
import numpy as np
import torch

from ..field import Field
from ..params import lame
from ._mpm_base import MPMBase


class MPMObj(MPMBase):
    KIND = 0

    def __init__(self, *args, gravity=(0.0, -2.0, 0.0), **kwargs):
        super().__init__(*args, gravity=gravity, **kwargs)

    def _init_materials(self):
        d = self.device
        s = self.space_scale
        self.E_0 = Field(3, device=d, needs_grad=True)
        self.nu_0 = Field(3, device=d, needs_grad=True)
        self.mu_0 = Field(3, device=d, needs_grad=True)
        self.lamda_0 = Field(3, device=d, needs_grad=True)
        self.lambda_0 = self.lamda_0
        E = [4e5 * s, 4e1 * s, 4e3 * s]   # 0 lower, 1 middle, 2 upper
        for i in range(3):
            mu, lam = lame(E[i], 0.4)
            self.E_0[i], self.nu_0[i] = E[i], 0.4
            self.mu_0[i], self.lamda_0[i] = mu, lam
        self.titles = Field(self.n_particles, dtype=torch.int32, device=d)
        self.mat_id = self.titles
        self.preprocess_obj()

    def preprocess_obj(self):
        """Height-band material ids (``multi_obj.py:104-112``; bands at 30/43 and 33/43 of
        the object height, ``multi_obj.py:58-59``)."""
        p = self.particles.to_numpy() / np.float32(self.obj_scale)
        h = np.max(p[:, 1])
        border = np.array([2 * h * (30 / 43) - h, 2 * h * (33 / 43) - h]) * self.obj_scale
        y = p[:, 1] * self.obj_scale
        ids = np.where(y < border[0], 0, np.where(y < border[1], 1, 2)).astype(np.int32)
        self.titles.from_numpy(ids)

    def _mat_table(self):
        t = torch.zeros((3, 8), dtype=torch.float32, device=self.device)
        t[:, 0] = self.mu_0.data
        t[:, 1] = self.lamda_0.data
        return t
