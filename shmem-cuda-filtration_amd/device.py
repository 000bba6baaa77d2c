"""Device grids driven through the C ABI of libfiltration_hip.so."""
import ctypes

import numpy as np

from . import abi
from .abi import FIELD_NAMES, FiltFields, check, load_library

_H2D, _D2H, _D2D = 0, 1, 2


class DeviceGrid:
    """One subdomain resident on the GPU: the 21 arrays of ptr_Arrays
    (filt_device_memory_alloc) plus optional per-cell perm / poro fields."""

    def __init__(self, backend, params, localNx=None, perm=False, poro=False):
        self.be = backend
        self.lib = backend.lib
        self.params = params.copy()
        self.nx = params.NX if localNx is None else localNx
        self.n_cells = self.nx * params.Ny * params.Nz
        self.f = FiltFields()
        self.buf = ctypes.c_void_p()
        check(self.lib, self.lib.filt_device_memory_alloc(
            ctypes.byref(self.f), ctypes.byref(self.buf), self.nx, ctypes.byref(self.params)),
            "filt_device_memory_alloc")
        self._extra = []
        for name, want in (("perm", perm), ("poro", poro)):
            if want:
                ptr = ctypes.c_void_p()
                check(self.lib, self.lib.filt_device_malloc(ctypes.byref(ptr), 8 * self.n_cells),
                      "filt_device_malloc")
                setattr(self.f, name, ptr.value)
                self._extra.append(ptr.value)

    # -- transfers ---------------------------------------------------------
    def _ptr(self, name):
        v = getattr(self.f, name)
        if v is None:
            raise KeyError(f"{name} not allocated on the device")
        return v

    def upload(self, g, names=None):
        names = names or [n for n in FIELD_NAMES + ("media", "perm", "poro")
                          if g.a.get(n) is not None and getattr(self.f, n)]
        for n in names:
            arr = np.ascontiguousarray(g.a[n])
            check(self.lib, self.lib.filt_memcpy(self._ptr(n), arr.ctypes.data, arr.nbytes, _H2D),
                  f"upload {n}")

    def download(self, g, names=None):
        names = names or [n for n in FIELD_NAMES + ("media",)]
        for n in names:
            arr = g.a[n]
            check(self.lib, self.lib.filt_memcpy(arr.ctypes.data, self._ptr(n), arr.nbytes, _D2H),
                  f"download {n}")

    def get(self, name):
        dt = np.int32 if name == "media" else np.float64
        out = np.empty((self.params.Nz, self.params.Ny, self.nx), dt)
        check(self.lib, self.lib.filt_memcpy(out.ctypes.data, self._ptr(name), out.nbytes, _D2H),
              f"get {name}")
        return out

    # -- stages ------------------------------------------------------------
    def stage(self, n, t=0.0, rank=0, size=1):
        L, f, p = self.lib, ctypes.byref(self.f), ctypes.byref(self.params)
        if n == 1:
            rc = L.filt_ro_P2_Xi_calculation(f, p, self.nx, rank, size)
        elif n == 2:
            rc = L.filt_u_calculation(f, p, self.nx, rank, size)
        elif n == 3:
            rc = L.filt_roS_calculation(f, p, t, self.nx, rank, size)
        elif n == 4:
            rc = L.filt_P1_S2_calculation(f, p, self.nx, rank, size)
        elif n == 5:
            rc = L.filt_boundary_conditions(f, p, self.nx, rank, size)
        else:
            raise ValueError(n)
        check(L, rc, f"stage {n}")

    def time_step(self, t, with_u=1, rank=0, size=1):
        check(self.lib, self.lib.filt_time_step(ctypes.byref(self.f), ctypes.byref(self.params),
                                                t, self.nx, rank, size, with_u), "time_step")

    def run(self, nsteps, step0=0, with_u=1):
        for s in range(step0 + 1, step0 + nsteps + 1):
            self.time_step(s * self.params.dt, with_u)

    def sync(self):
        check(self.lib, self.lib.filt_synchronize(), "filt_synchronize")

    def free(self):
        if self.f.x:
            self.lib.filt_device_memory_free(ctypes.byref(self.f), self.buf)
            self.f.x = None
        for ptr in self._extra:
            self.lib.filt_device_free_ptr(ptr)
        self._extra = []

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Backend:
    """Entry to the HIP backend.  Raises when the library is missing or no
    GPU is present: there is no CPU fallback."""

    def __init__(self, device=0):
        self.lib = load_library()
        bx, by, bz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        p = abi.default_params_py()
        p.NX = p.Ny = p.Nz = 4
        check(self.lib, self.lib.filt_device_initialize(device, 4, ctypes.byref(p),
                                                        ctypes.byref(bx), ctypes.byref(by),
                                                        ctypes.byref(bz)),
              "filt_device_initialize")

    def grid(self, params, localNx=None, perm=False, poro=False):
        return DeviceGrid(self, params, localNx, perm, poro)

    def sync(self):
        check(self.lib, self.lib.filt_synchronize(), "filt_synchronize")
