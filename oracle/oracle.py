"""Python binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
this module; it is the checker, never the thing measured or shipped.

  Oracle   -- liboracle.so, the C restatement of the reference's CPU path
  RefCPU   -- _ref/libref.so, the reference's own cpu.cpp + main.cpp +
              no_communication.cpp compiled from its sources (present only
              where the reference tree was available at build time)
"""
import ctypes
import importlib
import os
import sys
from ctypes import POINTER, c_double, c_int, c_size_t, c_void_p, c_char_p

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")
# the reference driver linked against the HIP backend (drop-in check)
REF_HIP_LIB = os.path.join(HERE, "_ref", "libref_hip.so")


def _pkg():
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    return importlib.import_module("shmem-cuda-filtration_amd")


pkg = _pkg()
FiltParams = pkg.abi.FiltParams
FiltFields = pkg.abi.FiltFields
HostGrid = pkg.grid.HostGrid


class Oracle:
    """The C restatement (filt_oracle.c)."""

    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run make -C oracle")
        self.lib = lib = ctypes.CDLL(path)
        P, F = POINTER(FiltParams), POINTER(FiltFields)
        for name in ("orc_ro_P2_Xi_calculation", "orc_u_calculation",
                     "orc_P1_S2_calculation", "orc_boundary_conditions"):
            getattr(lib, name).argtypes = [F, P, c_int, c_int, c_int]
        lib.orc_roS_calculation.argtypes = [F, P, c_double, c_int, c_int, c_int]
        lib.orc_time_step.argtypes = [F, P, c_double, c_int, c_int, c_int, c_int]
        lib.orc_run.argtypes = [F, P, c_int, c_int, c_int, c_int, c_int, c_int]
        lib.orc_initial_data.argtypes = [F, P, c_int, c_int, c_int, c_int]
        lib.orc_i_to_I.argtypes = [c_int, c_int, c_int, P]
        lib.orc_i_to_I.restype = c_int
        lib.orc_N_to_local.argtypes = [POINTER(c_int), POINTER(c_int), c_int, c_int, P]
        lib.orc_is_active_point.argtypes = [c_int, c_int, c_int, c_int]
        lib.orc_is_active_point.restype = c_int
        lib.orc_mass.argtypes = [F, P, c_int, POINTER(c_double)]
        lib.orc_set_loop_order.argtypes = [c_int]
        lib.orc_set_threads.argtypes = [c_int]
        lib.orc_get_threads.restype = c_int

    def stage(self, n, g, p, t=0.0, rank=0, size=1):
        f = g.fields()
        if n == 1:
            self.lib.orc_ro_P2_Xi_calculation(f, p, g.nx, rank, size)
        elif n == 2:
            self.lib.orc_u_calculation(f, p, g.nx, rank, size)
        elif n == 3:
            self.lib.orc_roS_calculation(f, p, t, g.nx, rank, size)
        elif n == 4:
            self.lib.orc_P1_S2_calculation(f, p, g.nx, rank, size)
        elif n == 5:
            self.lib.orc_boundary_conditions(f, p, g.nx, rank, size)
        else:
            raise ValueError(n)

    def time_step(self, g, p, t, with_u=1, rank=0, size=1):
        self.lib.orc_time_step(g.fields(), p, t, g.nx, rank, size, with_u)

    def run(self, g, p, nsteps, step0=0, with_u=1, rank=0, size=1):
        self.lib.orc_run(g.fields(), p, g.nx, rank, size, step0, nsteps, with_u)

    def initial_data(self, g, p, rank=0, size=1):
        self.lib.orc_initial_data(g.fields(), p, g.nx, g.ny, rank, size)

    def mass(self, g, p):
        out = (c_double * 2)()
        self.lib.orc_mass(g.fields(), p, g.nx, out)
        return out[0], out[1]

    def i_to_I(self, i, rank, size, p):
        return self.lib.orc_i_to_I(i, rank, size, p)

    def N_to_local(self, size, rank, p):
        a, b = c_int(), c_int()
        self.lib.orc_N_to_local(ctypes.byref(a), ctypes.byref(b), size, rank, p)
        return a.value, b.value

    def is_active_point(self, i, localNx, rank, size):
        return self.lib.orc_is_active_point(i, localNx, rank, size)

    def set_loop_order(self, order):
        self.lib.orc_set_loop_order(order)

    def set_threads(self, n):
        self.lib.orc_set_threads(n)


# ---------------------------------------------------------------- reference --

class RefConsts(ctypes.Structure):
    """consts_tag of the reference (defines.h:50-61)."""

    _fields_ = [
        ("K", c_double * 2), ("lambda_", c_double * 2), ("S_wr", c_double * 2),
        ("m", c_double * 2), ("P_d", c_double * 2),
        ("h1", c_double), ("h2", c_double), ("h3", c_double), ("dt", c_double),
        ("tau", c_double), ("l_w", c_double), ("l_n", c_double), ("c", c_double),
        ("beta1", c_double), ("beta2", c_double),
        ("NX", c_int), ("Ny", c_int), ("Nz", c_int),
        ("source", c_int), ("save_plots", c_int), ("print_screen", c_int),
    ]


class RefArrays(ctypes.Structure):
    """ptr_Arrays_tag of the reference (defines.h:41-47)."""

    _fields_ = [(n, c_void_p) for n in pkg.abi.FIELD_NAMES] + [("media", c_void_p)]


REF_CONSTANTS = dict(mu1=0.001, mu2=0.0009, ro01=1000.0, ro02=1460.0, S2_gr=0.4,
                     P_atm=100000.0, g_const=9.8, newton_iterations=7)

_SCALARS = ("h1", "h2", "h3", "dt", "tau", "l_w", "l_n", "c", "beta1", "beta2",
            "NX", "Ny", "Nz", "source", "save_plots", "print_screen")


def to_ref_consts(p):
    """filt_params -> consts_tag.  Raises if p holds something the compiled
    reference cannot express (other constants, explicit patch bounds, more
    than two media)."""
    for k, v in REF_CONSTANTS.items():
        if getattr(p, k) != v:
            raise ValueError(f"reference fixes {k}={v}, params have {getattr(p, k)}")
    nx2, nz2 = p.NX // 2, p.Nz // 2
    if (p.src_x0, p.src_x1, p.src_z0, p.src_z1) != (nx2 - p.source, nx2 + p.source,
                                                   nz2 - p.source, nz2 + p.source):
        raise ValueError("patch bounds are not the reference's (2*source+1)^2 patch")
    c = RefConsts()
    for q in range(2):
        c.K[q], c.lambda_[q], c.S_wr[q] = p.K[q], p.lambda_[q], p.S_wr[q]
        c.m[q], c.P_d[q] = p.m[q], p.P_d[q]
    for k in _SCALARS:
        setattr(c, k, getattr(p, k))
    return c


def from_ref_consts(c, base=None):
    p = base.copy() if base is not None else pkg.abi.default_params_py()
    for q in range(2):
        p.K[q], p.lambda_[q], p.S_wr[q] = c.K[q], c.lambda_[q], c.S_wr[q]
        p.m[q], p.P_d[q] = c.m[q], c.P_d[q]
    for k in _SCALARS:
        setattr(p, k, getattr(c, k))
    return p


def ref_arrays(g):
    if g.a.get("perm") is not None or g.a.get("poro") is not None:
        raise ValueError("the reference has no per-cell K / porosity fields")
    r = RefArrays()
    for n in pkg.abi.FIELD_NAMES + ("media",):
        setattr(r, n, g.a[n].ctypes.data_as(c_void_p).value)
    return r


class RefCPU:
    """The reference's own CPU path (oracle/_ref/libref.so)."""

    def __init__(self, path=REF_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing (reference tree absent at build time)")
        self.lib = lib = ctypes.CDLL(path)
        C, A = POINTER(RefConsts), POINTER(RefArrays)
        lib.ref_sizeof_consts.restype = c_size_t
        lib.ref_sizeof_ptr_arrays.restype = c_size_t
        lib.ref_constants.argtypes = [POINTER(c_double)]
        lib.ref_stage.argtypes = [c_int, A, C, c_int, c_int, c_int, c_double]
        lib.ref_stage.restype = c_int
        lib.ref_run.argtypes = [A, C, c_int, c_int, c_int, c_int, c_int, c_int]
        lib.ref_initial_data.argtypes = [A, C, c_int, c_int, c_int, c_int]
        lib.ref_i_to_I.argtypes = [c_int, c_int, c_int, C]
        lib.ref_i_to_I.restype = c_int
        lib.ref_N_to_local.argtypes = [POINTER(c_int), POINTER(c_int), c_int, c_int, C]
        lib.ref_is_active_point.argtypes = [c_int, c_int, c_int, c_int]
        lib.ref_is_active_point.restype = c_int
        lib.ref_read_defines.argtypes = [c_char_p, C]
        lib.ref_read_defines.restype = c_int
        lib.ref_defines_file.restype = c_char_p
        lib.ref_drive.argtypes = [C, c_int, c_void_p, c_void_p, c_size_t]
        lib.ref_drive.restype = c_int
        lib.ref_exchange.argtypes = [c_int, c_void_p, c_void_p, c_int, c_int, c_int, C]
        if hasattr(lib, "ref_roS_divgrad"):  # absent where another backend replaces cpu.cpp
            lib.ref_roS_divgrad.argtypes = [A, C, c_double, c_int, c_int, c_int]
            lib.ref_run_divgrad.argtypes = [A, C, c_int, c_int, c_int, c_int, c_int]
        lib.ref_save.argtypes = [c_char_p, A, C, c_int, c_int, c_int, c_int]
        lib.ref_save.restype = c_int
        lib.ref_restore.argtypes = [c_char_p, A, C, POINTER(c_int), c_int, c_int, c_int]
        lib.ref_restore.restype = c_int
        lib.ref_save_data_plots.argtypes = [c_char_p, A, C, c_double, c_int, c_int, c_int]
        lib.ref_save_data_plots.restype = c_int
        assert lib.ref_sizeof_consts() == ctypes.sizeof(RefConsts)
        assert lib.ref_sizeof_ptr_arrays() == ctypes.sizeof(RefArrays)

    def exchange(self, which, array, buffer, localNx, rank, size, p):
        """The reference backend's load_exchange_data (which = 0) or
        save_exchange_data (which = 1) on host arrays."""
        self.lib.ref_exchange(which, array.ctypes.data, buffer.ctypes.data, localNx, rank, size,
                              ctypes.byref(to_ref_consts(p)))

    def roS_divgrad(self, g, p, t, rank=0, size=1):
        """roS_calculation's loop over assign_rS (cpu.cpp:162-216)."""
        self.lib.ref_roS_divgrad(ref_arrays(g), to_ref_consts(p), t, g.nx, rank, size)

    def run_divgrad(self, g, p, nsteps, step0=0, rank=0, size=1):
        """time_step_function with the divgrad roS stage."""
        self.lib.ref_run_divgrad(ref_arrays(g), to_ref_consts(p), g.nx, rank, size, step0,
                                 nsteps)

    def save(self, workdir, g, p, j, rank=0, size=1):
        """The reference's save (main.cpp:485-537): writes
        workdir/save/save.dat from the host grid g."""
        rc = self.lib.ref_save(str(workdir).encode(), ref_arrays(g), to_ref_consts(p), j, g.nx,
                               rank, size)
        assert rc == 0, rc

    def restore(self, workdir, g, p, rank=0, size=1):
        """The reference's restore (main.cpp:540-569) into g; returns j."""
        j = c_int(-1)
        rc = self.lib.ref_restore(str(workdir).encode(), ref_arrays(g), to_ref_consts(p),
                                  ctypes.byref(j), g.nx, rank, size)
        assert rc == 0, rc
        return j.value

    def save_data_plots(self, workdir, g, p, t, rank=0, size=1):
        """The reference's save_data_plots (main.cpp:319-347) in workdir."""
        rc = self.lib.ref_save_data_plots(str(workdir).encode(), ref_arrays(g), to_ref_consts(p),
                                          t, g.nx, rank, size)
        assert rc == 0, rc

    def drive(self, p, nsteps):
        """The reference driver end to end (Initialize, time_step_function x
        nsteps, load_data_to_host, Finalize) on this library's backend."""
        import numpy as np
        shape = (p.Nz, p.Ny, p.NX)
        P = np.empty(shape)
        S = np.empty(shape)
        rc = self.lib.ref_drive(ctypes.byref(to_ref_consts(p)), nsteps, P.ctypes.data,
                                S.ctypes.data, P.size)
        assert rc == 0
        return P, S

    def constants(self):
        out = (c_double * 8)()
        self.lib.ref_constants(out)
        return list(out)

    def stage(self, n, g, p, t=0.0, rank=0, size=1):
        rc = self.lib.ref_stage(n, ref_arrays(g), to_ref_consts(p), g.nx, rank, size, t)
        assert rc == 0

    def run(self, g, p, nsteps, step0=0, rank=0, size=1):
        self.lib.ref_run(ref_arrays(g), to_ref_consts(p), g.nx, g.ny, rank, size, step0, nsteps)

    def initial_data(self, g, p, rank=0, size=1):
        self.lib.ref_initial_data(ref_arrays(g), to_ref_consts(p), g.nx, g.ny, rank, size)

    def i_to_I(self, i, rank, size, p):
        return self.lib.ref_i_to_I(i, rank, size, to_ref_consts(p))

    def N_to_local(self, size, rank, p):
        a, b = c_int(), c_int()
        self.lib.ref_N_to_local(ctypes.byref(a), ctypes.byref(b), size, rank, to_ref_consts(p))
        return a.value, b.value

    def is_active_point(self, i, localNx, rank, size):
        return self.lib.ref_is_active_point(i, localNx, rank, size)

    def read_defines(self, ini_text, workdir):
        """Run the reference's read_defines on ini_text, laid out at the
        relative path DEFINES_FILE (defines.h:5) under workdir."""
        rel = self.lib.ref_defines_file().decode()
        path = os.path.join(workdir, rel)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "wb") as fh:
            fh.write(ini_text if isinstance(ini_text, bytes) else ini_text.encode())
        c = RefConsts()
        rc = self.lib.ref_read_defines(workdir.encode(), ctypes.byref(c))
        assert rc == 0
        return c


def ref_available():
    return os.path.exists(REF_LIB)
