"""Elastoplastic MLS-MPM object (``difftactile/object_model/mpm_plastic.py``, class ``MPMObj``).

Von Mises return map in Hencky strain before the fixed-corotated stress
(``mpm_plastic.py:150-166,221-256``); ``yield_stress`` defaults to 1000
(``mpm_plastic.py:69,73``) and gravity to -9.80 (``mpm_plastic.py:39``).
"""
import torch

from ..field import Field
from ._mpm_base import MPMBase


class MPMObj(MPMBase):
    KIND = 1

    def __init__(self, *args, gravity=(0.0, -9.80, 0.0), **kwargs):
        super().__init__(*args, gravity=gravity, **kwargs)

    def _init_materials(self):
        super()._init_materials()
        self.yield_stress = Field((), device=self.device)
        self.yield_stress[None] = 1000.0

    def _mat_table(self):
        t = super()._mat_table()
        t[0, 2] = self.yield_stress.data
        return t
