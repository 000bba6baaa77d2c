"""The HIP path against the golden vectors written by the reference's own CPU
path (tests/golden/make_golden.py), with no oracle in between.

Bars as for the oracle comparisons: stages without pow (u, roS, boundary) and
the boundary known-answer test bit for bit; stages that call pow (ro_P2_Xi,
the Newton with P_d != 0, which every fixture has) to 1e-13 relative per
element; trajectories to relative L-inf 1e-10 on the arrays the fixtures hold
after step 1 and at the end.  The fused forms run the 3D trajectories: FP64
to 1e-10, the fast and mixed forms to their bars (S2 within 1e-4, P1 within 1e-6).  The
initial state of a trajectory comes from the oracle's initial_data, which
test_golden pins bit for bit."""
import numpy as np
import pytest

from cases import FIELD_NAMES_ALL, HostGrid, max_rel
from test_golden import load, params_of

pytestmark = pytest.mark.gpu

STAGE_OUT = {1: ("P2", "Xi1", "Xi2", "ro1", "ro2"),
             2: ("u1x", "u1y", "u1z", "u2x", "u2y", "u2z"),
             3: ("roS1", "roS2", "roS1_old", "roS2_old"),
             4: ("P1", "S2"),
             5: ("P1", "S2")}


def elem_rel(a, b):
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


@pytest.mark.parametrize("name", ["stage_kat_7x6x5", "stage_kat_9x8x7", "stage_kat_5x5x1",
                                  "stage_kat_4x6x2"])
@pytest.mark.parametrize("stage", [1, 2, 3, 4, 5])
def test_stage_kat_on_gpu(hip, name, stage):
    d, meta = load(name)
    p = params_of(meta)
    g = HostGrid(p)
    for n in FIELD_NAMES_ALL:
        g.a[n][...] = d[f"in_{n}"]
    dev = hip.grid(p)
    try:
        dev.upload(g)
        dev.stage(stage, t=meta["t"])
        dev.sync()
        out = g.copy()
        dev.download(out)
    finally:
        dev.free()
    exact = stage in (2, 3, 5) or (stage == 4 and all(x == 0 for x in meta["P_d"]))
    for n in FIELD_NAMES_ALL:
        key = f"stage{stage}_{n}"
        want = d[key] if key in d.files else d[f"in_{n}"]
        if n not in STAGE_OUT[stage] or exact:
            assert np.array_equal(out[n], want, equal_nan=True), (name, stage, n)
        else:
            assert elem_rel(out[n], want) <= 1e-13, (name, stage, n, elem_rel(out[n], want))


def test_boundary_kat_on_gpu(hip):
    d, meta = load("boundary_kat_6x5x7")
    p = params_of(meta)
    g = HostGrid(p)
    g["S2"], g["P1"], g["ro1"] = d["in_S2"], d["in_P1"], d["in_ro1"]
    dev = hip.grid(p)
    try:
        dev.upload(g)
        dev.stage(5)
        dev.sync()
        out = g.copy()
        dev.download(out)
    finally:
        dev.free()
    assert np.array_equal(out["S2"], d["out_S2"])
    assert np.array_equal(out["P1"], d["out_P1"])


@pytest.mark.parametrize("name", ["traj_15cube_s3", "traj_16cube_two_media", "traj_9x8x1",
                                  "traj_6x7x2", "traj_2x8x8"])
def test_trajectory_on_gpu(hip, orc, name):
    d, meta = load(name)
    p = params_of(meta)
    g = HostGrid(p)
    orc.initial_data(g, p)
    if "media" in d.files:
        g.a["media"][...] = d["media"]
    dev = hip.grid(p)
    try:
        dev.upload(g)
        dev.run(1)
        dev.sync()
        first = g.copy()
        dev.download(first)
        dev.run(meta["steps"] - 1, step0=1)
        dev.sync()
        last = g.copy()
        dev.download(last)
    finally:
        dev.free()
    checked = 0
    for key in d.files:
        for prefix, got in (("step1_", first), ("final_", last)):
            if key.startswith(prefix):
                assert max_rel(got[key[len(prefix):]], d[key]) <= 1e-10, (name, key)
                checked += 1
    assert checked >= 4


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("name", ["traj_15cube_s3", "traj_16cube_two_media"])
def test_fused_forms_trajectory_against_golden(hip, orc, name, precision):
    """The fused one-launch step on the 3D golden trajectories: FP64 to the
    trajectory bar (1e-10 relative L-inf), the fast and mixed forms to their own bars
    (S2 within 1e-4, P1 within 1e-6 relative)."""
    from test_gpu_fused import fused_run
    d, meta = load(name)
    p = params_of(meta)
    g = HostGrid(p)
    orc.initial_data(g, p)
    if "media" in d.files:
        g.a["media"][...] = d["media"]
    P, S, _ = fused_run(hip, p, g, meta["steps"], precision=precision)
    if precision == 0:
        assert max_rel(P, d["final_P1"]) <= 1e-10
        assert max_rel(S, d["final_S2"]) <= 1e-10
    else:
        assert float(np.max(np.abs(S - d["final_S2"]))) <= 1e-4
        assert max_rel(P, d["final_P1"]) <= 1e-6
