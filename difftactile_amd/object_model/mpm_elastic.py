"""Elastic MLS-MPM object (``difftactile/object_model/mpm_elastic.py``, class ``MPMObj``).

Fixed-corotated Kirchhoff stress with J = det(Sigma); all kernels run in
libdifftactile_hip.so.  See ``_mpm_base.MPMBase`` for the shared surface.
"""
from ._mpm_base import MPMBase


class MPMObj(MPMBase):
    KIND = 0
