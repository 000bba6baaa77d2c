"""ctypes access to lib/libfiltration_defines.so, the defines.h backend shims.

The shims have C++ linkage and take the reference's structs by value
(defines.h:41-61, 68-82, 106-116; the layout is include/filtration_defines.h).
This module binds them under their mangled names so Python callers (tests,
bench) can make the calls the reference driver makes, in its order, on this
backend.  Loading fails loudly when the library has not been built.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_void_p

from . import abi

DEFINES_LIB = os.path.join(abi.LIB_DIR, "libfiltration_defines.so")


class PtrArrays(ctypes.Structure):
    """ptr_Arrays_tag (defines.h:41-47): twenty FP64 arrays and media."""

    _fields_ = [(n, c_void_p) for n in abi.FIELD_NAMES] + [("media", c_void_p)]


class Consts(ctypes.Structure):
    """consts_tag (defines.h:50-61)."""

    _fields_ = [
        ("K", c_double * 2), ("lambda_", c_double * 2), ("S_wr", c_double * 2),
        ("m", c_double * 2), ("P_d", c_double * 2),
        ("h1", c_double), ("h2", c_double), ("h3", c_double), ("dt", c_double),
        ("tau", c_double), ("l_w", c_double), ("l_n", c_double), ("c", c_double),
        ("beta1", c_double), ("beta2", c_double),
        ("NX", c_int), ("Ny", c_int), ("Nz", c_int),
        ("source", c_int), ("save_plots", c_int), ("print_screen", c_int),
    ]


_SCALARS = ("h1", "h2", "h3", "dt", "tau", "l_w", "l_n", "c", "beta1", "beta2",
            "NX", "Ny", "Nz", "source", "save_plots", "print_screen")


def consts_from_params(p):
    """The consts a caller of defines.h passes for filt_params p (the first
    two media; the header's compile-time constants are the shim's own)."""
    c = Consts()
    for q in range(2):
        c.K[q], c.lambda_[q], c.S_wr[q] = p.K[q], p.lambda_[q], p.S_wr[q]
        c.m[q], c.P_d[q] = p.m[q], p.P_d[q]
    for k in _SCALARS:
        setattr(c, k, getattr(p, k))
    return c


# mangled name, result, argument types (defines.h:68-82, 106-116)
_A, _C, _I, _PD, _PI = PtrArrays, Consts, c_int, POINTER(c_double), POINTER(c_int)
_SIGNATURES = {
    "device_memory_alloc": ("_Z19device_memory_allocP14ptr_Arrays_tagPPdi10consts_tag",
                            (POINTER(PtrArrays), POINTER(c_void_p), _I, _C)),
    "device_memory_free": ("_Z18device_memory_free14ptr_Arrays_tagPd", (_A, c_void_p)),
    "load_data_to_host": ("_Z17load_data_to_hostPdS_i10consts_tag",
                          (c_void_p, c_void_p, _I, _C)),
    "load_data_to_device": ("_Z19load_data_to_devicePdS_i10consts_tag",
                            (c_void_p, c_void_p, _I, _C)),
    "load_data_to_device_int": ("_Z23load_data_to_device_intPiS_i10consts_tag",
                                (c_void_p, c_void_p, _I, _C)),
    "Device_Initialize": ("_Z17Device_InitializeiPiS_S_i10consts_tag",
                          (_I, _PI, _PI, _PI, _I, _C)),
    "ro_P2_Xi_calculation": ("_Z20ro_P2_Xi_calculation14ptr_Arrays_tagS_10consts_tagiiiiii",
                             (_A, _A, _C, _I, _I, _I, _I, _I, _I)),
    "u_calculation": ("_Z13u_calculation14ptr_Arrays_tagS_iiiiii10consts_tag",
                      (_A, _A, _I, _I, _I, _I, _I, _I, _C)),
    "roS_calculation": ("_Z15roS_calculation14ptr_Arrays_tagS_10consts_tagdiiiiii",
                        (_A, _A, _C, c_double, _I, _I, _I, _I, _I, _I)),
    "P1_S2_calculation": ("_Z17P1_S2_calculation14ptr_Arrays_tagS_10consts_tagiiiiii",
                          (_A, _A, _C, _I, _I, _I, _I, _I, _I)),
    "boundary_conditions": ("_Z19boundary_conditions14ptr_Arrays_tagS_iiiiii10consts_tag",
                            (_A, _A, _I, _I, _I, _I, _I, _I, _C)),
}


class Shim:
    """The defines.h backend functions of libfiltration_defines.so."""

    def __init__(self, path=DEFINES_LIB):
        abi.load_library()  # the shims call into libfiltration_hip.so
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is missing: build it with __graft_entry__.build()")
        self.lib = ctypes.CDLL(path)
        for name, (mangled, args) in _SIGNATURES.items():
            fn = getattr(self.lib, mangled)
            fn.restype = None
            fn.argtypes = list(args)
            setattr(self, name, fn)

    def time_step(self, dev, c, t, localNx, rank=0, size=1, blocks=(0, 0, 0)):
        """The backend calls of time_step_function (main.cpp:83-94) in its
        order; the exchanges between them do nothing on one rank."""
        host = PtrArrays()
        bx, by, bz = blocks
        self.ro_P2_Xi_calculation(host, dev, c, localNx, rank, size, bx, by, bz)
        self.u_calculation(host, dev, localNx, rank, size, bx, by, bz, c)
        self.roS_calculation(host, dev, c, t, localNx, rank, size, bx, by, bz)
        self.P1_S2_calculation(host, dev, c, localNx, rank, size, bx, by, bz)
        self.boundary_conditions(host, dev, localNx, rank, size, bx, by, bz, c)
