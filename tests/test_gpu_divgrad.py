"""The divgrad roS scheme (assign_rS, cpu.cpp:162-216) on the GPU against
the oracle (itself bit-exact against the reference's assign_rS): bit for
bit per stage and over trajectories; the fused forms refuse it."""
import ctypes

import numpy as np
import pytest

from cases import FIELD_NAMES_ALL, HostGrid, random_state, ref_params

pytestmark = pytest.mark.gpu


def divgrad_params(*dims, **kw):
    p = ref_params(*dims, **kw)
    p.roS_scheme = 1
    return p


@pytest.mark.parametrize("dims", [(7, 6, 5), (70, 9, 6), (9, 8, 1), (33, 17, 11), (3, 3, 3),
                                  (130, 7, 70)])
@pytest.mark.parametrize("tstep", [1, 4])
def test_divgrad_stage_bitwise(hip, orc, dims, tstep):
    p = divgrad_params(*dims)
    g = random_state(p, np.random.default_rng(dims[0] + 100 * tstep))
    d = hip.grid(p)
    try:
        d.upload(g)
        d.stage(3, t=tstep * p.dt)
        out = g.copy()
        d.download(out)
    finally:
        d.free()
    orc.stage(3, g, p, t=tstep * p.dt)
    for n in FIELD_NAMES_ALL:
        assert np.array_equal(out[n], g[n]), (n, dims, tstep)


TILES = [(64, 4), (32, 8), (16, 16), (64, 8), (128, 4), (128, 8)]


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("kz,order", [(1, 0), (5, 1), (64, 0)])
def test_divgrad_tiled_every_config_bitwise(hip, orc, tile, kz, order):
    """The z-marching divgrad kernel at every tile shape, z-chunk and block
    order, with porosity per cell and by the media map, on both sides of the
    time test (plain Euler while t < 2 dt): the oracle's bits (partial tiles
    in x and y, chunk ends not aligned with the grid)."""
    lib = hip.lib
    p = divgrad_params(137, 19, 23)
    try:
        assert lib.filt_set_tile_config(0, tile[0], tile[1], kz, order) == 0
        for poro in (False, True):
            for tstep in (1, 4):
                rng = np.random.default_rng(kz * 1000 + tile[0] * 10 + order + 2 * tstep + poro)
                g = random_state(p, rng, poro=poro)
                d = hip.grid(p, poro=poro)
                try:
                    d.upload(g)
                    d.stage(3, t=tstep * p.dt)
                    out = g.copy()
                    d.download(out)
                finally:
                    d.free()
                orc.stage(3, g, p, t=tstep * p.dt)
                for n in FIELD_NAMES_ALL:
                    assert np.array_equal(out[n], g[n]), (n, tile, kz, order, poro, tstep)
    finally:
        lib.filt_set_tile_config(0, 0, 0, 0, 0)


def test_divgrad_untiled_switch_and_divisor_range(hip, orc):
    """The one-cell-per-thread kernel (tiled kernels switched off, or a
    divisor h^2 below the range of the tiled kernel's exact division) gives
    the same bits as the tiled one."""
    lib = hip.lib
    outs = []
    for mode in ("tiled", "untiled", "tiny_h"):
        kw = {"h1": 1e-20} if mode == "tiny_h" else {}
        p = divgrad_params(40, 12, 9, **kw)
        g = random_state(p, np.random.default_rng(5), poro=True)
        try:
            if mode == "untiled":
                assert lib.filt_set_tile_config(0, -1, 0, 0, 0) == 0
            d = hip.grid(p, poro=True)
            try:
                d.upload(g)
                d.stage(3, t=4 * p.dt)
                out = g.copy()
                d.download(out)
            finally:
                d.free()
        finally:
            lib.filt_set_tile_config(0, 0, 0, 0, 0)
        orc.stage(3, g, p, t=4 * p.dt)
        for n in FIELD_NAMES_ALL:
            assert np.array_equal(out[n], g[n]), (n, mode)
        outs.append(out)
    for n in ("roS1", "roS2", "roS1_old", "roS2_old"):
        assert np.array_equal(outs[0][n], outs[1][n]), n


def test_stencil_autotuner_with_divgrad_keeps_roS_arrays(hip, orc):
    """filt_autotune_stencil on a grid with the divgrad scheme times the
    divgrad kernel as the roS stage; it leaves the four roS arrays as they
    were (the kernel rewrites its own roS_old input) and u as one call gives;
    a stage call afterwards, with the tuned tile, gives the oracle's bits."""
    lib = hip.lib
    p = divgrad_params(70, 33, 29)
    g = random_state(p, np.random.default_rng(12), poro=True)
    d = hip.grid(p, poro=True)
    try:
        d.upload(g)
        best = (ctypes.c_double * 2)()
        assert lib.filt_autotune_stencil(ctypes.byref(d.f), ctypes.byref(d.params), d.nx, 2,
                                         best) == 0
        assert best[0] > 0 and best[1] > 0
        v = [ctypes.c_int() for _ in range(4)]
        assert lib.filt_get_tile_config(0, 70, 33, 29, *[ctypes.byref(x) for x in v]) == 0
        assert (v[0].value, v[1].value) in TILES
        out = g.copy()
        d.download(out)
        want = g.copy()
        orc.stage(2, want, p)
        for n in FIELD_NAMES_ALL:
            assert np.array_equal(out[n], want[n]), n
        d.stage(3, t=4 * p.dt)
        d.download(out)
    finally:
        d.free()
        lib.filt_set_tile_config(0, 0, 0, 0, 0)
        lib.filt_set_tile_config(1, 0, 0, 0, 0)
    orc.stage(3, want, p, t=4 * p.dt)
    for n in FIELD_NAMES_ALL:
        assert np.array_equal(out[n], want[n]), n


@pytest.mark.parametrize("dims,steps", [((20, 16, 12), 60), ((10, 9, 1), 30)])
def test_divgrad_trajectory_matches_oracle(hip, orc, dims, steps):
    """Stages 1 (pow: ulp-level), u, divgrad roS, Newton, boundary: the
    trajectory bar of the plain path, relative L-inf <= 1e-10."""
    from cases import max_rel
    p = divgrad_params(*dims, source=1)
    a = HostGrid(p)
    orc.initial_data(a, p)
    d = hip.grid(p)
    try:
        d.upload(a)
        d.run(steps, with_u=0)  # the scheme needs u: the library computes it anyway
        out = a.copy()
        d.download(out)
    finally:
        d.free()
    orc.run(a, p, steps)
    for n in ("P1", "S2", "roS1", "roS2", "u2x", "u2y"):
        assert max_rel(out[n], a[n]) <= 1e-10, n
    assert np.isfinite(out["P1"]).all()


def test_fused_forms_refuse_divgrad(hip):
    p = divgrad_params(8, 8, 8)
    e = hip.ensemble(0)
    try:
        P = np.full((8, 8, 8), 1e5)
        S = np.zeros((8, 8, 8))
        mid = ctypes.c_int()
        rc = hip.lib.filt_ensemble_add(e.h, ctypes.byref(p), P.ctypes.data, S.ctypes.data,
                                       None, None, None, ctypes.byref(mid))
        assert rc == 6  # FILT_ERR_UNSUPPORTED
    finally:
        e.close()


def test_unknown_scheme_is_rejected(hip):
    p = ref_params(6, 6, 6)
    p.roS_scheme = 7
    d = hip.grid(p)
    try:
        rc = hip.lib.filt_roS_calculation(ctypes.byref(d.f), ctypes.byref(d.params), 0.0, 6, 0, 1)
        assert rc == 1  # FILT_ERR_INVALID
    finally:
        d.free()
