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

    def mass(self):
        """(water, NAPL) mass over the interior cells, summed on the device."""
        out = (ctypes.c_double * 2)()
        check(self.lib, self.lib.filt_mass(ctypes.byref(self.f), ctypes.byref(self.params),
                                           self.nx, out), "filt_mass")
        return out[0], out[1]

    # -- checkpoint / resume, plots (save, restore, save_data_plots) --------
    def save(self, path, j, truncate=True, rank=0, size=1):
        check(self.lib, self.lib.filt_save(str(path).encode(), ctypes.byref(self.f),
                                           ctypes.byref(self.params), self.nx, rank, size, j,
                                           1 if truncate else 0), "filt_save")

    def restore(self, path, rank=0, size=1):
        """Load a save file into this grid; returns its step j."""
        j = ctypes.c_int(-1)
        check(self.lib, self.lib.filt_restore(str(path).encode(), ctypes.byref(self.f),
                                              ctypes.byref(self.params), self.nx, rank, size,
                                              ctypes.byref(j)), "filt_restore")
        return j.value

    def save_plots(self, folder, t, rank=0, size=1):
        check(self.lib, self.lib.filt_save_plots(str(folder).encode(), ctypes.byref(self.f),
                                                 ctypes.byref(self.params), self.nx, rank, size,
                                                 t), "filt_save_plots")

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


PREC_FP64 = abi.FILT_PREC_FP64    # the stage path's expressions, same bits
PREC_FAST = abi.FILT_PREC_FAST    # all FP64, algebraic shortcuts
PREC_MIXED = abi.FILT_PREC_MIXED  # FP32 fluxes and ro*S increment, FP64 Newton
PRECISIONS = {"fp64": PREC_FP64, "fast": PREC_FAST, "mixed": PREC_MIXED}


class Ensemble:
    """Fused one-launch-per-step batch of grids (filt_ensemble_* in the C ABI)."""

    def __init__(self, backend, precision=PREC_FP64):
        self.lib = backend.lib
        self.h = ctypes.c_void_p()
        check(self.lib, self.lib.filt_ensemble_create(ctypes.byref(self.h), precision),
              "filt_ensemble_create")
        self.shapes = {}

    @staticmethod
    def _p(a):
        if a is None:
            return None
        assert a.flags.c_contiguous
        return a.ctypes.data

    def add(self, params, P1, S2, perm=None, poro=None, media=None):
        mid = ctypes.c_int()
        P1 = np.ascontiguousarray(P1, np.float64)
        S2 = np.ascontiguousarray(S2, np.float64)
        perm = None if perm is None else np.ascontiguousarray(perm, np.float64)
        poro = None if poro is None else np.ascontiguousarray(poro, np.float64)
        media = None if media is None else np.ascontiguousarray(media, np.int32)
        check(self.lib, self.lib.filt_ensemble_add(self.h, ctypes.byref(params), self._p(P1),
                                                   self._p(S2), self._p(perm), self._p(poro),
                                                   self._p(media), ctypes.byref(mid)),
              "filt_ensemble_add")
        self.shapes[mid.value] = (params.Nz, params.Ny, params.NX)
        return mid.value

    def remove(self, mid):
        check(self.lib, self.lib.filt_ensemble_remove(self.h, mid), "filt_ensemble_remove")
        self.shapes.pop(mid, None)

    def step(self, n=1):
        check(self.lib, self.lib.filt_ensemble_step(self.h, n), "filt_ensemble_step")

    def capture(self, n):
        check(self.lib, self.lib.filt_ensemble_capture(self.h, n), "filt_ensemble_capture")

    def replay(self):
        check(self.lib, self.lib.filt_ensemble_replay(self.h), "filt_ensemble_replay")

    def get(self, mid):
        shape = self.shapes[mid]
        P = np.empty(shape, np.float64)
        S = np.empty(shape, np.float64)
        check(self.lib, self.lib.filt_ensemble_get(self.h, mid, P.ctypes.data, S.ctypes.data),
              "filt_ensemble_get")
        return P, S

    def mass(self, mid):
        out = (ctypes.c_double * 2)()
        check(self.lib, self.lib.filt_ensemble_mass(self.h, mid, out), "filt_ensemble_mass")
        return out[0], out[1]

    def steps(self, mid):
        return self.lib.filt_ensemble_steps(self.h, mid)

    def set_tiling(self, kz, order=0, tile=(0, 0, 1)):
        """tile = (tx, ty[, cy]): columns per tile and columns per thread
        ((0, 0): 16x16, one column per thread); kz planes per z-chunk; order
        0 XCD-grouped / 1 hardware order."""
        tx, ty = tile[0], tile[1]
        cy = tile[2] if len(tile) > 2 else 1
        check(self.lib, self.lib.filt_ensemble_set_tiling(self.h, tx, ty, cy, kz, order),
              "set_tiling")

    def tiling(self):
        v = [ctypes.c_int() for _ in range(5)]
        check(self.lib, self.lib.filt_ensemble_get_tiling(self.h, *[ctypes.byref(x) for x in v]),
              "get_tiling")
        return {"tile": [v[0].value, v[1].value, v[2].value], "kz": v[3].value,
                "order": v[4].value}

    def autotune(self, reps=3):
        """Search tile shape x kz x order; returns the chosen tiling with its
        ms per step."""
        tx, ty, cy, kz, order = (ctypes.c_int() for _ in range(5))
        ms = ctypes.c_double()
        check(self.lib, self.lib.filt_ensemble_autotune(
            self.h, reps, ctypes.byref(tx), ctypes.byref(ty), ctypes.byref(cy), ctypes.byref(kz),
            ctypes.byref(order), ctypes.byref(ms)), "filt_ensemble_autotune")
        return {"tile": [tx.value, ty.value, cy.value], "kz": kz.value, "order": order.value,
                "ms_per_step": ms.value}

    def apply_tiling(self, t):
        self.set_tiling(t["kz"], t["order"], tuple(t["tile"]))

    def cells(self):
        return self.lib.filt_ensemble_cells(self.h)

    def device_bytes(self):
        return self.lib.filt_ensemble_device_bytes(self.h)

    def close(self):
        if self.h:
            self.lib.filt_ensemble_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
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

    def ensemble(self, precision=PREC_FP64):
        return Ensemble(self, precision)

    def sync(self):
        check(self.lib, self.lib.filt_synchronize(), "filt_synchronize")


class Snapshot:
    """Device-resident snapshot of a grid with an asynchronous writer
    (filt_snapshot_* in the C ABI)."""

    def __init__(self, grid, what=3, rank=0, size=1):
        self.lib = grid.lib
        self.h = ctypes.c_void_p()
        check(self.lib, self.lib.filt_snapshot_create(ctypes.byref(self.h),
                                                      ctypes.byref(grid.params), grid.nx, rank,
                                                      size, what), "filt_snapshot_create")

    def take(self, grid, j, t=0.0):
        check(self.lib, self.lib.filt_snapshot_take(self.h, ctypes.byref(grid.f), j, t),
              "filt_snapshot_take")

    def write_async(self, save_path=None, plot_dir=None, truncate=True):
        check(self.lib, self.lib.filt_snapshot_write_async(
            self.h, str(save_path).encode() if save_path else None, 1 if truncate else 0,
            str(plot_dir).encode() if plot_dir else None), "filt_snapshot_write_async")

    def wait(self):
        check(self.lib, self.lib.filt_snapshot_wait(self.h), "filt_snapshot_wait")

    def restore(self, grid):
        j = ctypes.c_int(-1)
        check(self.lib, self.lib.filt_snapshot_restore(self.h, ctypes.byref(grid.f),
                                                       ctypes.byref(j)), "filt_snapshot_restore")
        return j.value

    def close(self):
        if self.h:
            rc = self.lib.filt_snapshot_destroy(self.h)
            self.h = ctypes.c_void_p()
            check(self.lib, rc, "filt_snapshot_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
