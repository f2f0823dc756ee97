"""Taichi-field-like views over GPU tensors.

The reference's models expose ``ti.field`` / ``ti.Vector.field`` /
``ti.Matrix.field`` members with a ``.grad`` twin, ``[f, p]`` indexing,
``to_numpy`` / ``from_numpy`` and ``fill``.  ``Field`` gives the same surface
over a torch tensor on the GPU, with the same ``to_numpy()`` shape (scalar
fields ``shape``, vector fields ``shape + (n,)``, matrix fields
``shape + (n, m)``).
"""
import numpy as np
import torch


class Field:
    def __init__(self, shape, elem=(), dtype=torch.float32, device="cuda", needs_grad=False):
        shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.shape = shape
        self.elem = tuple(elem)
        self.data = torch.zeros(shape + self.elem, dtype=dtype, device=device)
        self.grad = Field(shape, elem, dtype, device, False) if needs_grad else None

    # -- Taichi-like API --------------------------------------------------
    def to_numpy(self):
        return self.data.detach().cpu().numpy()

    def from_numpy(self, a):
        a = np.asarray(a)
        if a.shape != tuple(self.data.shape):
            raise ValueError(f"shape {a.shape} != field shape {tuple(self.data.shape)}")
        self.data.copy_(torch.from_numpy(np.ascontiguousarray(a)).to(self.data.dtype))

    def to_torch(self, device=None):
        t = self.data.clone()
        return t if device is None else t.to(device)

    def from_torch(self, t):
        self.data.copy_(t)

    def fill(self, v):
        self.data.fill_(v)

    def _key(self, idx):
        if idx is None:
            return ()
        return idx

    def __getitem__(self, idx):
        v = self.data[self._key(idx)]
        if v.dim() == 0:
            return v.item()
        return v.detach().cpu().numpy()

    def __setitem__(self, idx, val):
        k = self._key(idx)
        if isinstance(val, torch.Tensor):
            self.data[k] = val.to(self.data.dtype)
        else:
            self.data[k] = torch.as_tensor(np.asarray(val), dtype=self.data.dtype)

    # -- frame views for the kernels ---------------------------------------
    def frame(self, f):
        """Contiguous view of frame f (first axis)."""
        return self.data[f]


class Kernel:
    """A model method with a ``.grad`` twin, like a Taichi kernel."""

    def __init__(self, fwd, grad):
        self._fwd = fwd
        self.grad = grad

    def __call__(self, *args, **kwargs):
        return self._fwd(*args, **kwargs)
