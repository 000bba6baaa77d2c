"""The tau-relaxed "divgrad" roS scheme (assign_rS, cpu.cpp:162-216),
selected by filt_params.roS_scheme: the oracle against the reference's own
assign_rS compiled into oracle/_ref, bit for bit, per stage (both time
branches, degenerate and multi-rank grids) and over trajectories."""
import numpy as np
import pytest

from cases import FIELD_NAMES_ALL, HostGrid, random_state, ref_params


def divgrad_params(*dims, **kw):
    p = ref_params(*dims, **kw)
    p.roS_scheme = 1  # FILT_SCHEME_DIVGRAD
    return p


@pytest.mark.parametrize("dims", [(7, 6, 5), (9, 8, 1), (5, 5, 3), (12, 7, 6), (4, 4, 2)])
@pytest.mark.parametrize("tstep", [1, 2, 7])
def test_divgrad_stage_bitwise(orc, ref, dims, tstep):
    """t = dt takes the Euler branch (t < 2 dt); 2 dt and later the relaxed one."""
    p = divgrad_params(*dims)
    g = random_state(p, np.random.default_rng(dims[0] * 31 + tstep))
    want = g.copy()
    ref.roS_divgrad(want, p, tstep * p.dt)
    orc.stage(3, g, p, t=tstep * p.dt)
    for n in FIELD_NAMES_ALL:
        assert np.array_equal(g[n], want[n]), n
    inner = (slice(1, -1) if dims[2] > 1 else slice(None), slice(1, -1), slice(1, -1))
    if dims[0] > 2 and dims[1] > 2 and (dims[2] > 2 or dims[2] == 1):
        changed = g["roS1"][inner] != random_state(
            p, np.random.default_rng(dims[0] * 31 + tstep))["roS1"][inner]
        assert changed.all()  # every interior cell was updated


@pytest.mark.parametrize("rank,size", [(1, 3), (0, 2), (2, 3)])
def test_divgrad_stage_multirank(orc, ref, rank, size):
    p = divgrad_params(13, 6, 5)
    nx, _ = ref.N_to_local(size, rank, p)
    g = random_state(p, np.random.default_rng(rank + 10 * size), localNx=nx)
    want = g.copy()
    ref.roS_divgrad(want, p, 5 * p.dt, rank=rank, size=size)
    orc.stage(3, g, p, t=5 * p.dt, rank=rank, size=size)
    for n in FIELD_NAMES_ALL:
        assert np.array_equal(g[n], want[n]), n


def test_divgrad_stage_differs_from_upwind(orc):
    """The selector switches the arithmetic (guards against a scheme flag
    that is silently ignored)."""
    p = ref_params(8, 7, 6)
    q = divgrad_params(8, 7, 6)
    a = random_state(p, np.random.default_rng(3))
    b = a.copy()
    orc.stage(3, a, p, t=3 * p.dt)
    orc.stage(3, b, q, t=3 * p.dt)
    assert not np.array_equal(a["roS1"], b["roS1"])


@pytest.mark.parametrize("dims,steps", [((12, 10, 8), 40), ((10, 9, 1), 30), ((6, 5, 4), 25)])
def test_divgrad_trajectory_bitwise(orc, ref, dims, steps):
    """time_step_function with the divgrad stage: stage 1, u, divgrad roS,
    Newton, boundary, from the initial state."""
    p = divgrad_params(*dims, source=1)
    a = HostGrid(p)
    orc.initial_data(a, p)
    b = a.copy()
    orc.run(a, p, steps, with_u=0)  # the scheme forces u on
    ref.run_divgrad(b, p, steps)
    for n in ("P1", "S2", "roS1", "roS2", "roS1_old", "roS2_old", "u1x", "u2y"):
        assert np.array_equal(a[n], b[n], equal_nan=True), n
    assert np.isfinite(a["P1"]).all()
