"""The oracle (oracle/filt_oracle.c) against the reference's own CPU path
compiled from its sources (oracle/_ref): bit-for-bit on every stage, on
multi-step trajectories, on the driver helpers and on the ini parser.
Skipped where the reference tree was absent at build time; the committed
golden vectors (test_golden.py) pin the same facts there."""
import numpy as np
import pytest

from cases import FIELD_NAMES_ALL, HostGrid, random_state, ref_params

DIMS = [(7, 6, 5), (9, 8, 7), (5, 5, 1), (4, 6, 2), (2, 5, 5), (3, 3, 3), (12, 4, 9)]


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def test_constants_mirror(ref):
    import oracle
    c = ref.constants()
    want = oracle.REF_CONSTANTS
    assert c == [want["mu1"], want["mu2"], want["ro01"], want["ro02"], want["S2_gr"],
                 want["P_atm"], want["g_const"], want["newton_iterations"]]


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("stage", [1, 2, 3, 4, 5])
def test_stage_bitwise(orc, ref, dims, stage):
    p = ref_params(*dims, source=1)
    g = random_state(p, np.random.default_rng(hash((dims, stage)) & 0xFFFF))
    a, b = g.copy(), g.copy()
    orc.stage(stage, a, p, t=0.37)
    ref.stage(stage, b, p, t=0.37)
    for n in FIELD_NAMES_ALL:
        assert same(a[n], b[n]), (n, dims, stage)


@pytest.mark.parametrize("dims", [(9, 8, 7), (5, 5, 1), (3, 3, 3)])
def test_loop_order_and_threads_do_not_change_results(orc, dims):
    p = ref_params(*dims)
    g = random_state(p, np.random.default_rng(5))
    a, b = g.copy(), g.copy()
    orc.set_loop_order(1)
    orc.run(a, p, 5)
    orc.set_loop_order(0)
    orc.set_threads(4)
    orc.run(b, p, 5)
    orc.set_threads(1)
    for n in FIELD_NAMES_ALL:
        assert same(a[n], b[n]), n


@pytest.mark.parametrize("dims,source", [((15, 15, 15), 3), ((16, 12, 10), 2), ((9, 8, 1), 1)])
def test_trajectory_bitwise(orc, ref, dims, source):
    p = ref_params(*dims, source=source)
    a = HostGrid(p)
    orc.initial_data(a, p)
    b = HostGrid(p)
    ref.initial_data(b, p)
    for n in FIELD_NAMES_ALL:
        assert same(a[n], b[n]), ("initial", n)
    orc.run(a, p, 60)
    ref.run(b, p, 60)
    for n in FIELD_NAMES_ALL:
        assert same(a[n], b[n]), ("run", n)


def test_two_media_trajectory_bitwise(orc, ref):
    p = ref_params(16, 16, 16, source=2)
    rng = np.random.default_rng(11)
    a = HostGrid(p)
    orc.initial_data(a, p)
    a.a["media"][...] = rng.integers(0, 2, a["P1"].shape)
    b = a.copy()
    orc.run(a, p, 100)
    ref.run(b, p, 100)
    for n in FIELD_NAMES_ALL:
        assert same(a[n], b[n]), n


@pytest.mark.parametrize("NX,size", [(15, 1), (15, 2), (16, 3), (31, 4), (7, 7)])
def test_index_helpers(orc, ref, NX, size):
    p = ref_params(NX, 5, 5)
    for rank in range(size):
        assert orc.N_to_local(size, rank, p) == ref.N_to_local(size, rank, p)
        lnx, _ = orc.N_to_local(size, rank, p)
        for i in range(lnx):
            assert orc.i_to_I(i, rank, size, p) == ref.i_to_I(i, rank, size, p)
            assert orc.is_active_point(i, lnx, rank, size) == ref.is_active_point(i, lnx, rank, size)


def test_multi_rank_stage_masks(orc, ref):
    """Stages honour is_active_point on a middle rank's slab (halo cells
    untouched) exactly like the reference."""
    p = ref_params(20, 6, 5, source=1)
    rank, size = 1, 3
    lnx, _ = orc.N_to_local(size, rank, p)
    g = random_state(p, np.random.default_rng(3), localNx=lnx)
    for stage in (1, 2, 3, 4, 5):
        a, b = g.copy(), g.copy()
        orc.stage(stage, a, p, rank=rank, size=size)
        ref.stage(stage, b, p, rank=rank, size=size)
        for n in FIELD_NAMES_ALL:
            assert same(a[n], b[n]), (stage, n)
