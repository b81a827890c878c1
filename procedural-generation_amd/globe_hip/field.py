"""FaceField: six face arrays with halos -- the erosion stage's input and
output type (bake/globe/field.py:24-124).

``data`` is (6, NE, NE) or (6, NE, NE, C) with NE = N + 2H; interior cell
(i, j) of face f is ``data[f, i + H, j + H]``.  Vector fields hold
contravariant cell components.  Host (numpy) fields are what the stage
reads and writes; the halo exchange of a field that lives on the GPU (a
torch tensor) runs the device gather kernel.  There is no CPU halo exchange
here: a host field's halos come with it (e.g. evaluated at the extended
cell centres) or are refreshed by moving it to the GPU.
"""
from __future__ import annotations

import numpy as np


class FaceField:
    __slots__ = ("grid", "data", "is_vector", "name")

    def __init__(self, grid, data, is_vector: bool = False, name: str = ""):
        if tuple(data.shape[:3]) != (6, grid.NE, grid.NE):
            raise ValueError(f"data shape {tuple(data.shape)} does not match grid NE={grid.NE}")
        if is_vector and (data.ndim != 4 or data.shape[3] != 2):
            raise ValueError("vector fields must have shape (6, NE, NE, 2)")
        self.grid = grid
        self.data = data
        self.is_vector = bool(is_vector)
        self.name = name

    @classmethod
    def zeros(cls, grid, ncomp: int = 1, dtype=np.float32, is_vector: bool = False, name: str = "") -> "FaceField":
        shape = (6, grid.NE, grid.NE) if ncomp == 1 else (6, grid.NE, grid.NE, ncomp)
        return cls(grid, np.zeros(shape, dtype=dtype), is_vector=is_vector, name=name)

    @classmethod
    def from_function(cls, grid, fn, dtype=np.float32, name: str = "") -> "FaceField":
        """``fn(points3)`` at every extended cell centre: halos are analytic."""
        return cls(grid, np.ascontiguousarray(np.asarray(fn(grid.centers), dtype=dtype)), name=name)

    @property
    def N(self) -> int:
        return self.grid.N

    @property
    def H(self) -> int:
        return self.grid.H

    @property
    def interior(self):
        H, N = self.grid.H, self.grid.N
        return self.data[:, H:H + N, H:H + N]

    def face(self, f: int):
        return self.interior[f]

    def exchange_halos(self, linear: bool = False) -> "FaceField":
        """Fill the halo cells from the neighbouring faces on the GPU
        (float64 torch tensor data only)."""
        import torch

        if not isinstance(self.data, torch.Tensor) or self.data.device.type != "cuda":
            raise TypeError("FaceField.exchange_halos runs on the GPU: move data to a cuda float64 tensor first")
        from .erosion.maps import HaloTables

        HaloTables(self.grid, self.data.device).exchange(self.data, vector=self.is_vector, linear=linear)
        return self
