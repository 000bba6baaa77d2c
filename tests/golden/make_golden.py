"""Writes the golden vectors under tests/golden/ from the reference's own CPU
path (oracle/_ref/libref.so, compiled from the reference sources by
oracle/Makefile).  Run from the repository root where the reference tree is
present:  python tests/golden/make_golden.py

Each .npz holds inputs and expected outputs (FP64 / int32 arrays shaped
(Nz, Ny, NX), x fastest) plus a JSON 'meta' string with the parameters.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for p in (REPO, os.path.join(REPO, "oracle"), os.path.join(REPO, "tests")):
    sys.path.insert(0, p)

import oracle  # noqa: E402
from cases import FIELD_NAMES_ALL, HostGrid, random_state, ref_params  # noqa: E402

PARAM_KEYS = ("K", "lambda_", "S_wr", "m", "P_d", "h1", "h2", "h3", "dt", "tau", "l_w", "l_n",
              "c", "beta1", "beta2", "NX", "Ny", "Nz", "source", "save_plots", "print_screen")


def meta_of(p, **extra):
    d = {}
    for k in PARAM_KEYS:
        v = getattr(p, k)
        d[k] = list(v)[:2] if hasattr(v, "__len__") else v
    d.update(extra)
    return json.dumps(d)


def save(name, meta, arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, meta=np.array(meta), **arrays)
    print(f"{name}: {os.path.getsize(path)} bytes")


def stage_kat(ref, dims, seed):
    p = ref_params(*dims, source=1)
    g = random_state(p, np.random.default_rng(seed))
    arrays = {f"in_{n}": g[n] for n in FIELD_NAMES_ALL}
    for st in (1, 2, 3, 4, 5):
        out = g.copy()
        ref.stage(st, out, p, t=0.37)
        for n in FIELD_NAMES_ALL:
            if not np.array_equal(out[n], g[n], equal_nan=True):
                arrays[f"stage{st}_{n}"] = out[n]
    save(f"stage_kat_{dims[0]}x{dims[1]}x{dims[2]}", meta_of(p, seed=seed, t=0.37), arrays)


def boundary_kat(ref):
    p = ref_params(6, 5, 7, source=1)
    g = random_state(p, np.random.default_rng(99))
    n = g["S2"].size
    g["S2"] = (np.arange(n, dtype=np.float64) + 1).reshape(g["S2"].shape) * 1e-3
    g["P1"] = 1e5 + (np.arange(n, dtype=np.float64)[::-1] + 1).reshape(g["P1"].shape)
    out = g.copy()
    ref.stage(5, out, p)
    save("boundary_kat_6x5x7", meta_of(p), {"in_S2": g["S2"], "in_P1": g["P1"],
                                            "in_ro1": g["ro1"], "out_S2": out["S2"],
                                            "out_P1": out["P1"]})


def trajectory(ref, dims, source, steps, name, media_seed=None, keep=("P1", "S2")):
    p = ref_params(*dims, source=source)
    g = HostGrid(p)
    ref.initial_data(g, p)
    arrays = {}
    if media_seed is not None:
        g.a["media"][...] = np.random.default_rng(media_seed).integers(0, 2, g["P1"].shape)
        arrays["media"] = g.a["media"].copy()
    sums = []
    for s in range(steps):
        ref.run(g, p, 1, step0=s)
        sums.append([g["P1"].sum(), g["S2"].sum()])
        if s == 0:
            for n in keep:
                arrays[f"step1_{n}"] = g[n].copy()
    for n in keep:
        arrays[f"final_{n}"] = g[n].copy()
    arrays["checksums"] = np.array(sums)
    save(name, meta_of(p, steps=steps, media_seed=media_seed), arrays)


def stated_cube32_divergence(ref):
    """The configuration-1 compressibility 4.4e-10 on the reference itself
    (h = 1/31, dt = 1e-4, 32^3, its own 5x5 patch and S2_gr 0.4)."""
    p = ref_params(32, 32, 32, source=2, h1=1 / 31, h2=1 / 31, h3=1 / 31, dt=1e-4,
                   beta1=4.4e-10, beta2=4.4e-10)
    for q in range(2):
        p.K[q], p.lambda_[q], p.S_wr[q], p.m[q], p.P_d[q] = 6.64e-11, 2.0, 0.1, 0.4, 0.0
    g = HostGrid(p)
    ref.initial_data(g, p)
    pmax = []
    for s in range(12):
        ref.run(g, p, 1, step0=s)
        with np.errstate(invalid="ignore"):
            pmax.append(np.nanmax(np.abs(g["P1"])) if np.isfinite(g["P1"]).any() else np.nan)
    finite = [bool(np.isfinite(g["P1"]).all())]
    save("stated_cube32_divergence", meta_of(p), {"pmax": np.array(pmax), "finite_end": np.array(finite)})


def main():
    ref = oracle.RefCPU()
    stage_kat(ref, (7, 6, 5), 11)
    stage_kat(ref, (9, 8, 7), 12)
    stage_kat(ref, (5, 5, 1), 13)
    stage_kat(ref, (4, 6, 2), 14)
    boundary_kat(ref)
    trajectory(ref, (15, 15, 15), 3, 60, "traj_15cube_s3")
    trajectory(ref, (16, 16, 16), 2, 100, "traj_16cube_two_media", media_seed=11)
    trajectory(ref, (9, 8, 1), 1, 30, "traj_9x8x1", keep=("P1", "S2", "u1x", "u2y"))
    trajectory(ref, (6, 7, 2), 1, 30, "traj_6x7x2", keep=("P1", "S2", "u2y", "u2z"))
    trajectory(ref, (2, 8, 8), 1, 30, "traj_2x8x8", keep=("P1", "S2", "u2y"))
    stated_cube32_divergence(ref)


if __name__ == "__main__":
    main()
