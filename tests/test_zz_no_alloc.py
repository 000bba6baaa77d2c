"""filt_mass_noalloc (the no-allocation entry point) through the C ABI on the
GPU: bit for bit what filt_mass returns, on two calls in a row, the second
on a smaller grid (its fixed scratch keeps the first call's partials, which
the second must not read); grids beyond its fixed scratch are refused."""
import ctypes
import importlib

import numpy as np
import pytest

from cases import random_state, ref_params

pytestmark = pytest.mark.gpu

pkg = importlib.import_module("shmem-cuda-filtration_amd")
abi = pkg.abi


def mass_noalloc(hip, d):
    """filt_mass_noalloc into device memory this test owns, copied back."""
    lib = hip.lib
    out = ctypes.c_void_p()
    abi.check(lib, lib.filt_device_malloc(ctypes.byref(out), 16), "malloc")
    try:
        abi.check(lib, lib.filt_mass_noalloc(ctypes.byref(d.f), ctypes.byref(d.params), d.nx,
                                             out), "filt_mass_noalloc")
        host = np.zeros(2)
        abi.check(lib, lib.filt_memcpy(host.ctypes.data, out, 16, 1), "copy back")
        return float(host[0]), float(host[1])
    finally:
        lib.filt_device_free_ptr(out)


def test_noalloc_mass_equals_filt_mass_on_two_calls(hip):
    grids = []
    try:
        for dims, seed in (((96, 80, 64), 1), ((20, 17, 9), 2)):
            p = ref_params(*dims)
            g = random_state(p, np.random.default_rng(seed), poro=True)
            d = hip.grid(p, poro=True)
            d.upload(g)
            grids.append(d)
        for d in grids:  # the larger grid first, then the smaller one
            want = d.mass()
            got = mass_noalloc(hip, d)
            assert got == want, (d.nx, got, want)
        # and again in the other order, after other work on the stream
        grids[0].run(2, with_u=0)
        for d in reversed(grids):
            assert mass_noalloc(hip, d) == d.mass()
    finally:
        for d in grids:
            d.free()


def test_noalloc_mass_refuses_grids_beyond_its_scratch(hip):
    p = ref_params(9, 8, 7)
    d = hip.grid(p)
    try:
        big = p.copy()
        big.NX, big.Ny, big.Nz = 1024, 1024, 257  # 269,484,032 cells > 268,435,456
        out = ctypes.c_void_p()
        abi.check(hip.lib, hip.lib.filt_device_malloc(ctypes.byref(out), 16), "malloc")
        try:
            rc = hip.lib.filt_mass_noalloc(ctypes.byref(d.f), ctypes.byref(big), big.NX, out)
            assert rc == abi.FILT_ERR_INVALID
            assert b"fixed scratch" in hip.lib.filt_last_error()
        finally:
            hip.lib.filt_device_free_ptr(out)
    finally:
        d.free()
