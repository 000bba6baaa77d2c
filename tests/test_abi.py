"""The C-ABI libraries load and export every entry point the headers declare
(no GPU needed: nothing here launches device work)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import REPO

INC = os.path.join(REPO, "include")
LIB = os.path.join(REPO, "shmem-cuda-filtration_amd", "lib")


def declared_c_functions(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(filt_\w+)\s*\(", text)))


def exported(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_hip_library_exports_every_declared_function():
    path = os.path.join(LIB, "libfiltration_hip.so")
    lib = ctypes.CDLL(path)
    names = declared_c_functions("filtration_hip.h")
    assert len(names) > 30
    syms = exported(path)
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    for n in names:
        getattr(lib, n)


# The 13 functions gpu.cu provides, with the mangled names the reference's
# defines.h gives them (C++ linkage, struct tags ptr_Arrays_tag / consts_tag).
BACKEND_MANGLED = [
    "_Z20ro_P2_Xi_calculation14ptr_Arrays_tagS_10consts_tagiiiiii",
    "_Z13u_calculation14ptr_Arrays_tagS_iiiiii10consts_tag",
    "_Z15roS_calculation14ptr_Arrays_tagS_10consts_tagdiiiiii",
    "_Z17P1_S2_calculation14ptr_Arrays_tagS_10consts_tagiiiiii",
    "_Z19boundary_conditions14ptr_Arrays_tagS_iiiiii10consts_tag",
    "_Z17load_data_to_hostPdS_i10consts_tag",
    "_Z19load_data_to_devicePdS_i10consts_tag",
    "_Z23load_data_to_device_intPiS_i10consts_tag",
    "_Z19device_memory_allocP14ptr_Arrays_tagPPdi10consts_tag",
    "_Z18device_memory_free14ptr_Arrays_tagPd",
    "_Z17Device_InitializeiPiS_S_i10consts_tag",
    "_Z18load_exchange_dataPdS_S_S_iiiii10consts_tag",
    "_Z18save_exchange_dataPdS_S_S_iiiii10consts_tag",
]


def test_defines_backend_exports_reference_mangled_names():
    path = os.path.join(LIB, "libfiltration_defines.so")
    ctypes.CDLL(os.path.join(LIB, "libfiltration_hip.so"), mode=ctypes.RTLD_GLOBAL)
    ctypes.CDLL(path)
    syms = exported(path)
    missing = [n for n in BACKEND_MANGLED if n not in syms]
    assert not missing, missing


def test_backend_names_match_reference_backend(ref):
    """Every mangled name the reference's CPU backend exports for the
    device-facing half of defines.h is exported by the HIP backend."""
    syms = exported(os.path.join(LIB, "libfiltration_defines.so"))
    refsyms = exported(os.path.join(REPO, "oracle", "_ref", "libref.so"))
    for n in BACKEND_MANGLED:
        assert n in refsyms, n
        assert n in syms, n


def test_struct_layouts_match_header(pkg):
    """ctypes mirrors have the C sizes (filt_params 512 bytes, 23 pointers)."""
    assert ctypes.sizeof(pkg.abi.FiltParams) == 512
    assert ctypes.sizeof(pkg.abi.FiltFields) == 23 * 8


def test_library_refuses_bad_arguments_without_gpu(pkg):
    lib = pkg.abi.load_library()
    p = pkg.abi.default_params_py()
    p.NX = p.Ny = p.Nz = 0
    f = pkg.abi.FiltFields()
    rc = lib.filt_ro_P2_Xi_calculation(ctypes.byref(f), ctypes.byref(p), 0, 0, 1)
    assert rc == pkg.abi.FILT_ERR_INVALID
    assert b"bad grid" in lib.filt_last_error()


def test_no_cpu_fallback_in_product_sources():
    """The product sources never import, load or link the oracle."""
    pkgdir = os.path.join(REPO, "shmem-cuda-filtration_amd")
    banned = re.compile(r"import\s+oracle|liboracle|libref|filt_oracle|orc_\w+\(|oracle/")
    for root, _, files in os.walk(pkgdir):
        for fn in files:
            if fn.endswith((".py", ".hip", ".cpp", ".hpp", ".h")) or fn == "Makefile":
                text = open(os.path.join(root, fn)).read()
                assert not banned.search(text), fn


def test_defines_shim_module_binds_every_entry_point():
    """defines_shim binds the defines.h functions of libfiltration_defines.so
    under the reference's mangled names, with the reference's struct sizes
    (ptr_Arrays: 21 pointers; consts: 20 doubles and 6 ints, no padding)."""
    import ctypes
    import importlib
    m = importlib.import_module("shmem-cuda-filtration_amd.defines_shim")
    assert ctypes.sizeof(m.PtrArrays) == 21 * 8
    assert ctypes.sizeof(m.Consts) == 20 * 8 + 6 * 4
    s = m.Shim()
    for name in m._SIGNATURES:
        assert callable(getattr(s, name))
