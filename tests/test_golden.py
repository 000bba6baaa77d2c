"""The oracle against golden vectors written by the reference's own CPU path
(tests/golden/make_golden.py).  Bit-exact.  Runs without the reference tree,
so the oracle stays pinned wherever the tests run."""
import glob
import json
import os

import numpy as np
import pytest

from cases import FIELD_NAMES_ALL, HostGrid, abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    return d, meta


def params_of(meta):
    p = abi.default_params_py()
    for k, v in meta.items():
        if not hasattr(p, k):
            continue
        if isinstance(v, list):
            arr = getattr(p, k)
            for q, x in enumerate(v):
                arr[q] = x
        else:
            setattr(p, k, v)
    return abi.finalize_patch_py(p)


def test_fixtures_present():
    names = {os.path.basename(f)[:-4] for f in glob.glob(os.path.join(GOLDEN, "*.npz"))}
    assert {"stage_kat_7x6x5", "stage_kat_9x8x7", "boundary_kat_6x5x7", "traj_15cube_s3",
            "traj_16cube_two_media", "stated_cube32_divergence"} <= names


@pytest.mark.parametrize("name", ["stage_kat_7x6x5", "stage_kat_9x8x7", "stage_kat_5x5x1",
                                  "stage_kat_4x6x2"])
@pytest.mark.parametrize("stage", [1, 2, 3, 4, 5])
def test_stage_kat(orc, name, stage):
    d, meta = load(name)
    p = params_of(meta)
    g = HostGrid(p)
    for n in FIELD_NAMES_ALL:
        g.a[n][...] = d[f"in_{n}"]
    orc.stage(stage, g, p, t=meta["t"])
    for n in FIELD_NAMES_ALL:
        key = f"stage{stage}_{n}"
        want = d[key] if key in d.files else d[f"in_{n}"]
        assert np.array_equal(g[n], want, equal_nan=True), (name, stage, n)


def test_boundary_kat(orc):
    d, meta = load("boundary_kat_6x5x7")
    p = params_of(meta)
    g = HostGrid(p)
    g["S2"], g["P1"], g["ro1"] = d["in_S2"], d["in_P1"], d["in_ro1"]
    orc.stage(5, g, p)
    assert np.array_equal(g["S2"], d["out_S2"])
    assert np.array_equal(g["P1"], d["out_P1"])


@pytest.mark.parametrize("name", ["traj_15cube_s3", "traj_16cube_two_media", "traj_9x8x1",
                                  "traj_6x7x2", "traj_2x8x8"])
def test_trajectory(orc, name):
    d, meta = load(name)
    p = params_of(meta)
    g = HostGrid(p)
    orc.initial_data(g, p)
    if "media" in d.files:
        g.a["media"][...] = d["media"]
    sums = []
    for s in range(meta["steps"]):
        orc.run(g, p, 1, step0=s)
        sums.append([g["P1"].sum(), g["S2"].sum()])
        if s == 0:
            for key in d.files:
                if key.startswith("step1_"):
                    assert np.array_equal(g[key[6:]], d[key]), key
    for key in d.files:
        if key.startswith("final_"):
            assert np.array_equal(g[key[6:]], d[key]), key
    assert np.array_equal(np.array(sums), d["checksums"])


def test_stated_compressibility_diverges(orc):
    """With beta = 4.4e-10 (the configurations' figure) the reference scheme
    blows up within a dozen steps; the oracle reproduces that run exactly."""
    d, meta = load("stated_cube32_divergence")
    p = params_of(meta)
    g = HostGrid(p)
    orc.initial_data(g, p)
    pmax = []
    for s in range(len(d["pmax"])):
        orc.run(g, p, 1, step0=s)
        with np.errstate(invalid="ignore"):
            pmax.append(np.nanmax(np.abs(g["P1"])) if np.isfinite(g["P1"]).any() else np.nan)
    assert np.array_equal(np.array(pmax), d["pmax"], equal_nan=True)
    assert d["pmax"][-1] > 1e100
