"""ctypes mirror of include/filtration_hip.h and loader of the HIP backend.

The C ABI is the product boundary; this module only declares its structs and
signatures for Python callers (tests, bench, tools).  Loading fails loudly when
the library has not been built: there is no CPU fallback.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_size_t, c_void_p, c_char_p

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libfiltration_hip.so")

FILT_MAX_MEDIA = 8

FILT_OK = 0
FILT_ERR_INVALID = 1
FILT_ERR_HIP = 2
FILT_ERR_NOMEM = 3
FILT_ERR_UNSTABLE = 4
FILT_ERR_IO = 5
FILT_ERR_UNSUPPORTED = 6
FILT_ERR_NOT_FOUND = 7
FILT_SNAP_SAVE = 1
FILT_SNAP_PLOTS = 2
FILT_SCHEME_UPWIND = 0
FILT_SCHEME_DIVGRAD = 1
FILT_PREC_FP64 = 0
FILT_PREC_FAST = 1
FILT_PREC_MIXED = 2

_D = c_double * FILT_MAX_MEDIA


class FiltParams(ctypes.Structure):
    """filt_params (superset of consts_tag, defines.h:50-61)."""

    _fields_ = [
        ("K", _D), ("lambda_", _D), ("S_wr", _D), ("m", _D), ("P_d", _D),
        ("h1", c_double), ("h2", c_double), ("h3", c_double), ("dt", c_double),
        ("tau", c_double), ("l_w", c_double), ("l_n", c_double), ("c", c_double),
        ("beta1", c_double), ("beta2", c_double),
        ("NX", c_int), ("Ny", c_int), ("Nz", c_int),
        ("source", c_int), ("save_plots", c_int), ("print_screen", c_int),
        ("mu1", c_double), ("mu2", c_double), ("ro01", c_double), ("ro02", c_double),
        ("S2_gr", c_double), ("P_atm", c_double), ("g_const", c_double),
        ("newton_iterations", c_int),
        ("src_x0", c_int), ("src_x1", c_int), ("src_z0", c_int), ("src_z1", c_int),
        ("n_media", c_int),
        ("roS_scheme", c_int),
    ]

    def copy(self):
        q = FiltParams()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(self), ctypes.sizeof(FiltParams))
        return q


FIELD_NAMES = (
    "x", "y", "z", "P1", "P2", "S2", "ro1", "ro2",
    "u1x", "u1y", "u1z", "u2x", "u2y", "u2z",
    "Xi1", "Xi2", "roS1", "roS1_old", "roS2", "roS2_old",
)


class FiltFields(ctypes.Structure):
    """filt_fields: ptr_Arrays_tag's members (defines.h:41-47) + perm, poro."""

    _fields_ = [(n, c_void_p) for n in FIELD_NAMES] + [
        ("media", c_void_p), ("perm", c_void_p), ("poro", c_void_p)]


def default_params_py():
    """The reference's compile-time defaults (defines.h:8-15, 22-26) without
    loading any library; ini-controlled members are zero."""
    p = FiltParams()
    K = (6.64e-11, 7.15e-12)
    lam = (2.7, 2.0)
    swr = (0.09, 0.12)
    m = (0.4, 0.39)
    pd = (755.0, 2060.0)
    for q in range(2):
        p.K[q], p.lambda_[q], p.S_wr[q], p.m[q], p.P_d[q] = K[q], lam[q], swr[q], m[q], pd[q]
    p.mu1, p.mu2, p.ro01, p.ro02 = 0.001, 0.0009, 1000.0, 1460.0
    p.S2_gr, p.P_atm, p.g_const = 0.4, 100000.0, 9.8
    p.newton_iterations = 7
    p.src_x0 = p.src_x1 = p.src_z0 = p.src_z1 = -1
    p.n_media = 2
    return p


def finalize_patch_py(p):
    """Patch bounds from `source` as main.cpp:109 tests them (C int division)."""
    if p.src_x0 < 0:
        p.src_x0 = p.NX // 2 - p.source
        p.src_x1 = p.NX // 2 + p.source
        p.src_z0 = p.Nz // 2 - p.source
        p.src_z1 = p.Nz // 2 + p.source
    return p


_lib = None


def _sig(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)


def load_library(path=HIP_LIB):
    """Load libfiltration_hip.so and declare its signatures.  Raises
    FileNotFoundError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with __graft_entry__.build() "
            "(the HIP backend has no CPU fallback)")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    P, F = POINTER(FiltParams), POINTER(FiltFields)
    pi = POINTER(c_int)
    _sig(lib, "filt_default_params", None, P)
    _sig(lib, "filt_params_finalize", None, P)
    _sig(lib, "filt_read_defines", c_int, c_char_p, P)
    _sig(lib, "filt_stability_number", c_double, P, c_void_p, c_void_p, c_size_t)
    _sig(lib, "filt_gen_uniform", c_int, c_void_p, c_size_t, ctypes.c_uint64, c_double, c_double)
    _sig(lib, "filt_gen_lognormal", c_int, c_void_p, c_int, c_int, c_int, ctypes.c_uint64,
         c_double, c_double, c_double, c_double, c_int)
    _sig(lib, "filt_i_to_I", c_int, c_int, c_int, c_int, P)
    _sig(lib, "filt_N_to_local", None, pi, pi, c_int, c_int, P)
    _sig(lib, "filt_is_active_point", c_int, c_int, c_int, c_int, c_int)
    _sig(lib, "filt_initial_data", c_int, F, P, c_int, c_int, c_int)
    _sig(lib, "filt_device_initialize", c_int, c_int, c_int, P, pi, pi, pi)
    _sig(lib, "filt_device_memory_alloc", c_int, F, POINTER(c_void_p), c_int, P)
    _sig(lib, "filt_device_memory_free", c_int, F, c_void_p)
    _sig(lib, "filt_load_data_to_host", c_int, c_void_p, c_void_p, c_int, P)
    _sig(lib, "filt_load_data_to_device", c_int, c_void_p, c_void_p, c_int, P)
    _sig(lib, "filt_load_data_to_device_int", c_int, c_void_p, c_void_p, c_int, P)
    _sig(lib, "filt_memcpy", c_int, c_void_p, c_void_p, c_size_t, c_int)
    _sig(lib, "filt_device_copy", c_int, c_void_p, c_void_p, c_size_t)
    _sig(lib, "filt_step_zm_launches", ctypes.c_longlong)
    _sig(lib, "filt_mass", c_int, F, P, c_int, POINTER(c_double))
    _sig(lib, "filt_mass_noalloc", c_int, F, P, c_int, c_void_p)
    _sig(lib, "filt_device_malloc", c_int, POINTER(c_void_p), c_size_t)
    _sig(lib, "filt_device_free_ptr", c_int, c_void_p)
    for name in ("filt_ro_P2_Xi_calculation", "filt_u_calculation",
                 "filt_P1_S2_calculation", "filt_boundary_conditions"):
        _sig(lib, name, c_int, F, P, c_int, c_int, c_int)
    _sig(lib, "filt_roS_calculation", c_int, F, P, c_double, c_int, c_int, c_int)
    _sig(lib, "filt_time_step", c_int, F, P, c_double, c_int, c_int, c_int, c_int)
    _sig(lib, "filt_time_step_timed", c_int, F, P, c_double, c_int, c_int, c_int, c_int,
         POINTER(c_void_p), POINTER(c_int))
    _sig(lib, "filt_u_roS_P1_S2_calculation", c_int, F, P, c_double, c_int, c_int, c_int, c_int)
    _sig(lib, "filt_load_exchange_data", c_int, c_void_p, c_void_p, c_void_p, c_int, P)
    _sig(lib, "filt_save_exchange_data", c_int, c_void_p, c_void_p, c_void_p, c_int,
         c_int, c_int, P)
    _sig(lib, "filt_set_tile_config", c_int, c_int, c_int, c_int, c_int, c_int)
    _sig(lib, "filt_get_tile_config", c_int, c_int, c_int, c_int, c_int, pi, pi, pi, pi)
    _sig(lib, "filt_autotune_stencil", c_int, F, P, c_int, c_int, POINTER(c_double))
    _sig(lib, "filt_autotune_step", c_int, F, P, c_int, c_int, POINTER(c_double))
    E = c_void_p
    _sig(lib, "filt_ensemble_create", c_int, POINTER(c_void_p), c_int)
    _sig(lib, "filt_ensemble_destroy", c_int, E)
    _sig(lib, "filt_ensemble_check", c_int, P, c_int, c_void_p, c_void_p)
    _sig(lib, "filt_ensemble_add", c_int, E, P, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p, pi)
    _sig(lib, "filt_ensemble_remove", c_int, E, c_int)
    _sig(lib, "filt_ensemble_count", c_int, E)
    _sig(lib, "filt_ensemble_cells", ctypes.c_longlong, E)
    _sig(lib, "filt_ensemble_steps", ctypes.c_longlong, E, c_int)
    _sig(lib, "filt_ensemble_member_bytes", c_size_t, P, c_int, c_int, c_int, c_int)
    _sig(lib, "filt_ensemble_device_bytes", c_size_t, E)
    _sig(lib, "filt_ensemble_step", c_int, E, c_int)
    _sig(lib, "filt_ensemble_capture", c_int, E, c_int)
    _sig(lib, "filt_ensemble_replay", c_int, E)
    _sig(lib, "filt_ensemble_get", c_int, E, c_int, c_void_p, c_void_p)
    _sig(lib, "filt_ensemble_mass", c_int, E, c_int, POINTER(c_double))
    _sig(lib, "filt_ensemble_set_tiling", c_int, E, c_int, c_int, c_int, c_int, c_int)
    _sig(lib, "filt_ensemble_get_tiling", c_int, E, pi, pi, pi, pi, pi)
    _sig(lib, "filt_ensemble_autotune", c_int, E, c_int, pi, pi, pi, pi, pi, POINTER(c_double))
    _sig(lib, "filt_device_free_bytes", c_int, POINTER(c_size_t), POINTER(c_size_t))
    _sig(lib, "filt_serve_create", c_int, POINTER(c_void_p), c_int, c_size_t, c_int, c_int)
    _sig(lib, "filt_serve_destroy", c_int, c_void_p)
    _sig(lib, "filt_serve_submit", c_int, c_void_p, P, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p, ctypes.c_longlong, pi)
    _sig(lib, "filt_serve_round", c_int, c_void_p, pi, pi)
    _sig(lib, "filt_serve_status", c_int, c_void_p, pi, pi, pi, POINTER(ctypes.c_longlong),
         POINTER(ctypes.c_longlong))
    _sig(lib, "filt_serve_is_done", c_int, c_void_p, c_int)
    _sig(lib, "filt_serve_take", c_int, c_void_p, c_int, c_void_p, c_void_p, POINTER(c_double))
    for name in ("filt_save", "filt_save_host"):
        _sig(lib, name, c_int, c_char_p, F, P, c_int, c_int, c_int, c_int, c_int)
    for name in ("filt_restore", "filt_restore_host"):
        _sig(lib, name, c_int, c_char_p, F, P, c_int, c_int, c_int, pi)
    for name in ("filt_save_plots", "filt_save_plots_host"):
        _sig(lib, name, c_int, c_char_p, F, P, c_int, c_int, c_int, c_double)
    _sig(lib, "filt_snapshot_create", c_int, POINTER(c_void_p), P, c_int, c_int, c_int, c_int)
    _sig(lib, "filt_snapshot_destroy", c_int, c_void_p)
    _sig(lib, "filt_snapshot_take", c_int, c_void_p, F, c_int, c_double)
    _sig(lib, "filt_snapshot_write_async", c_int, c_void_p, c_char_p, c_int, c_char_p)
    _sig(lib, "filt_snapshot_wait", c_int, c_void_p)
    _sig(lib, "filt_snapshot_restore", c_int, c_void_p, F, pi)
    _sig(lib, "filt_synchronize", c_int)
    _sig(lib, "filt_set_stream", c_int, c_void_p)
    _sig(lib, "filt_get_stream", c_void_p)
    _sig(lib, "filt_last_error", c_char_p)
    _sig(lib, "filt_build_info", c_char_p)
    _lib = lib
    return lib


class FiltError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


def check(lib, code, where):
    if code != FILT_OK:
        raise FiltError(code, where, lib.filt_last_error().decode(errors="replace"))
    return code
