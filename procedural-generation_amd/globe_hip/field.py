"""FaceField: six face arrays with halos -- the erosion stage's input and
output type (bake/globe/field.py:24-124).

``data`` is (6, NE, NE) or (6, NE, NE, C) with NE = N + 2H; interior cell
(i, j) of face f is ``data[f, i + H, j + H]``.  Vector fields hold
contravariant cell components.  Host (numpy) fields are what the stage
reads and writes; the halo exchange of a field that lives on the GPU (a
torch tensor) runs the device gather kernel.  There is no CPU halo exchange
here: a host field's halos come with it (e.g. evaluated at the extended
cell centres) or are refreshed on the GPU (``load``, ``from_interior``).

Files: ``save`` / ``load`` use the reference's layout (bake/globe/field.py:
312-342): the interior of face k as ``<dir>/<name>.f{k}.npy``.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

#: stored fields that are tangent vectors (contravariant cell components);
#: the reference's list (bake/globe/field.py:349), so files load the same way
VECTOR_FIELD_NAMES = frozenset({"wind", "momentum", "plate_vel", "momentum_track"})


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
    def from_interior(cls, grid, arr, is_vector: bool = False, name: str = "", exchange: bool = True,
                      device="cuda") -> "FaceField":
        """Build from a (6, N, N[, C]) interior array; the halos are filled by
        the device exchange (in float64, stored back in the array's dtype)
        unless exchange is False (halos zero)."""
        arr = np.asarray(arr)
        if tuple(arr.shape[:3]) != (6, grid.N, grid.N):
            raise ValueError(f"interior shape {arr.shape} does not match grid N={grid.N}")
        f = cls.zeros(grid, 1 if arr.ndim == 3 else arr.shape[3], arr.dtype, is_vector=is_vector, name=name)
        f.interior[...] = arr
        if exchange:
            import torch

            dev = FaceField(grid, torch.from_numpy(f.data.astype(np.float64)).to(device), is_vector=is_vector)
            dev.exchange_halos()
            f.data[...] = dev.data.cpu().numpy().astype(arr.dtype)
        return f

    def save(self, directory, name: str | None = None) -> list:
        """Write the interior of each face as ``<dir>/<name>.f{k}.npy``."""
        name = name or self.name
        if not name:
            raise ValueError("field needs a name to be saved")
        directory = Path(directory)
        directory.mkdir(parents=True, exist_ok=True)
        data = self.data.cpu().numpy() if hasattr(self.data, "cpu") else self.data
        H, N = self.grid.H, self.grid.N
        paths = []
        for k in range(6):
            p = directory / f"{name}.f{k}.npy"
            np.save(p, np.ascontiguousarray(data[k, H:H + N, H:H + N]))
            paths.append(p)
        return paths

    @classmethod
    def load(cls, directory, name: str, grid, is_vector: bool | None = None, exchange: bool = True,
             device="cuda") -> "FaceField":
        """Load ``<dir>/<name>.f{0..5}.npy`` and fill the halos."""
        directory = Path(directory)
        arr = np.stack([np.load(directory / f"{name}.f{k}.npy") for k in range(6)], axis=0)
        if is_vector is None:
            is_vector = arr.ndim == 4 and arr.shape[3] == 2 and name in VECTOR_FIELD_NAMES
        return cls.from_interior(grid, arr, is_vector=is_vector, name=name, exchange=exchange, device=device)

    @staticmethod
    def exists(directory, name: str) -> bool:
        directory = Path(directory)
        return all((directory / f"{name}.f{k}.npy").exists() for k in range(6))

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
