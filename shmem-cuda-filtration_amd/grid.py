"""Host-side field bundle: one numpy array per member of filt_fields.

Arrays have shape (Nz, Ny, localNx) in C order, so the flat index is the
reference's i + j*localNx + k*localNx*Ny (defines.h:41-47 and cpu.cpp).
"""
import ctypes

import numpy as np

from .abi import FIELD_NAMES, FiltFields


class HostGrid:
    def __init__(self, params, localNx=None, with_perm=False, with_poro=False):
        self.params = params
        self.nx = params.NX if localNx is None else localNx
        self.ny, self.nz = params.Ny, params.Nz
        shape = (self.nz, self.ny, self.nx)
        self.a = {n: np.zeros(shape, np.float64) for n in FIELD_NAMES}
        self.a["media"] = np.zeros(shape, np.int32)
        self.a["perm"] = np.zeros(shape, np.float64) if with_perm else None
        self.a["poro"] = np.zeros(shape, np.float64) if with_poro else None

    @property
    def n_cells(self):
        return self.nx * self.ny * self.nz

    def __getitem__(self, name):
        return self.a[name]

    def __setitem__(self, name, value):
        if self.a.get(name) is None:
            self.a[name] = np.ascontiguousarray(value, dtype=np.float64).reshape(
                self.nz, self.ny, self.nx)
        else:
            self.a[name][...] = value

    def fields(self):
        """filt_fields pointing at these arrays (valid while self lives)."""
        f = FiltFields()
        for n in FIELD_NAMES + ("media", "perm", "poro"):
            arr = self.a.get(n)
            setattr(f, n, None if arr is None else arr.ctypes.data_as(ctypes.c_void_p).value)
        return f

    def copy(self):
        g = HostGrid.__new__(HostGrid)
        g.params, g.nx, g.ny, g.nz = self.params, self.nx, self.ny, self.nz
        g.a = {n: (None if v is None else v.copy()) for n, v in self.a.items()}
        return g

    def interior(self, name):
        """View of the cells the stencil and Newton update (cpu.cpp:220)."""
        v = self.a[name]
        if self.nz < 2:
            return v[:, 1:-1, 1:-1]
        return v[1:-1, 1:-1, 1:-1]
