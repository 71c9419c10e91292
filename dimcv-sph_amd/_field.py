"""Device-backed stand-ins for the Taichi fields the reference classes expose.

The reference allocates ``ti.field`` / ``ti.Vector.field`` objects once at
capacity (particle_system.py:136-228, dimcv_sph.py:30-53) and its callers use
``[None]``, ``[i]``, ``.to_numpy()``, ``.from_numpy()``, ``.fill()`` and
``.shape`` on them (run_simulation.py:63,72-73; karman_vortex.py:39-47,136-145).
``Field`` offers the same protocol on a PyTorch tensor in HBM; the tensor's
``data_ptr()`` is what the C ABI receives.
"""
from __future__ import annotations

import numpy as np
import torch


class Field:
    """A per-particle (or per-cell) field: shape (n,) or (n, comps)."""

    def __init__(self, n: int, comps: int = 0, dtype=torch.float32, device="cuda",
                 as_bool: bool = False, view_shape=None):
        shape = (n, comps) if comps else (n,)
        self.t = torch.zeros(shape, dtype=dtype, device=device)
        self.as_bool = as_bool
        self.view_shape = view_shape  # e.g. (3, 3) for a matrix field stored as 9 floats

    # -- Taichi field protocol --------------------------------------------
    @property
    def shape(self):
        if self.view_shape:
            return (self.t.shape[0],) + tuple(self.view_shape)
        return tuple(self.t.shape)

    def _view(self):
        return self.t.view(self.shape) if self.view_shape else self.t

    def __getitem__(self, idx):
        v = self._view()[idx]
        if v.dim() == 0:
            return bool(v.item()) if self.as_bool else v.item()
        return v.cpu().numpy()

    def __setitem__(self, idx, value):
        self._view()[idx] = torch.as_tensor(np.asarray(value), dtype=self.t.dtype).to(self.t.device)

    def to_numpy(self) -> np.ndarray:
        a = self._view().cpu().numpy()
        return a.astype(bool) if self.as_bool else a

    def from_numpy(self, a) -> None:
        src = torch.from_numpy(np.ascontiguousarray(a)).to(self.t.dtype)
        self.t.copy_(src.reshape(self.t.shape))

    def fill(self, value) -> None:
        self.t.fill_(value)

    # -- ABI -----------------------------------------------------------------
    def ptr(self) -> int:
        return self.t.data_ptr()

    def swap(self, other: "Field") -> None:
        self.t, other.t = other.t, self.t


class CounterField:
    """0-d int field backed by one entry of the device counter array."""

    def __init__(self, counters: torch.Tensor, index: int):
        self._c = counters
        self._i = index

    shape = ()

    def __getitem__(self, idx):
        assert idx is None
        return int(self._c[self._i].item())

    def __setitem__(self, idx, value):
        assert idx is None
        self._c[self._i] = int(value)

    def to_numpy(self):
        return np.array(self[None], dtype=np.int32)

    def fill(self, value):
        self[None] = value


class DeviceScalarField:
    """0-d float field backed by one entry of the FP64 device scalar array."""

    def __init__(self, dscal: torch.Tensor, index: int):
        self._d = dscal
        self._i = index

    shape = ()

    def __getitem__(self, idx):
        assert idx is None
        return float(self._d[self._i].item())

    def __setitem__(self, idx, value):
        assert idx is None
        self._d[self._i] = float(value)

    def to_numpy(self):
        return np.array(self[None])

    def fill(self, value):
        self[None] = value


class HostScalarField:
    """0-d f32 field held on the host (solver parameters such as ``dt``).

    Reads return the Python float of the f32 value, as a Taichi f32 field does.
    """

    def __init__(self, value=0.0, on_change=None):
        self._v = np.float32(value)
        self._cb = on_change

    shape = ()

    def __getitem__(self, idx):
        assert idx is None
        return float(self._v)

    def __setitem__(self, idx, value):
        assert idx is None
        self._v = np.float32(value)
        if self._cb:
            self._cb(self._v)

    def to_numpy(self):
        return np.array(self._v)

    def fill(self, value):
        self[None] = value
