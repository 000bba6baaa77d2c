"""The fused one-launch-per-step forms (filt_ensemble_*) on the GPU.

FP64 fused: same bits as the plain stage path on the GPU (same expressions),
relative L-inf <= 1e-10 on P1 / S2 and 1e-12 on mass against the oracle.
Fast (closed-form Newton, lambda = 2 powers as products, all FP64) and mixed
(FP32 fluxes and ro*S increment, FP64 Newton): errors against the oracle
stated per field and held to S2 L-inf <= 1e-4, P1 relative L-inf <= 1e-6,
and water, NAPL and total mass <= 1e-9 relative.  Every tile shape gives the same bits.
Batching: a member's result is the same alone and in a mix, bit for bit."""
import numpy as np
import pytest

from cases import HostGrid, cube32_params, max_rel, random_state, ref_params

pytestmark = pytest.mark.gpu


def plain_run(hip, p, g, steps):
    d = hip.grid(p, perm=g.a.get("perm") is not None, poro=g.a.get("poro") is not None)
    try:
        d.upload(g)
        d.run(steps, with_u=0)
        d.sync()
        return d.get("P1"), d.get("S2")
    finally:
        d.free()


FP64, FAST, MIXED = 0, 1, 2


def fused_run(hip, p, g, steps, precision=0, graph=False, kz=None, tile=(0, 0, 1)):
    e = hip.ensemble(precision)
    try:
        mid = e.add(p, g["P1"], g["S2"], g.a.get("perm"), g.a.get("poro"),
                    g.a["media"] if p.n_media > 1 else None)
        if kz or tile != (0, 0, 1):
            kz_, order = kz or (32, 0)
            e.set_tiling(kz_, order, tile)
        if graph:
            e.capture(steps)
            e.replay()
        else:
            e.step(steps)
        P, S = e.get(mid)
        return P, S, e.mass(mid)
    finally:
        e.close()


def lognormal_like(n, seed=5, sigma=1.0):
    import importlib
    W = importlib.import_module("shmem-cuda-filtration_amd").workloads
    p = W.lognormal_params(n)
    perm, poro = W.gen_fields(p, seed, sigma=sigma, corr=4.0, clip=2.5)
    return p, W.initial_grid(p, perm, poro)


@pytest.mark.parametrize("dims", [(32, 32, 32), (37, 19, 23), (3, 3, 3), (18, 5, 40)])
def test_fused_fp64_equals_plain_bitwise(hip, dims):
    p = cube32_params()
    p.NX, p.Ny, p.Nz = dims
    p.src_x0, p.src_x1 = max(0, dims[0] // 2 - 2), dims[0] // 2 + 1
    p.src_z0, p.src_z1 = max(0, dims[2] // 2 - 2), dims[2] // 2 + 1
    g = HostGrid(p)
    import oracle
    oracle.Oracle().initial_data(g, p)
    Pp, Sp = plain_run(hip, p, g, 25)
    Pf, Sf, _ = fused_run(hip, p, g, 25)
    assert np.array_equal(Pf, Pp)
    assert np.array_equal(Sf, Sp)


@pytest.mark.parametrize("dims", [(65, 9, 7), (66, 6, 5), (129, 5, 6), (130, 4, 4), (4, 7, 9)])
def test_fused_fp64_equals_plain_at_tile_row_ends(hip, dims):
    """x-tiles start at x = 0, so the last interior column x = nx - 2 falls on
    the first or second column of a tile (nx = 65, 66 with 64-wide tiles;
    129, 130 with 64- and 128-wide), and a 4-wide grid has one interior
    column: the fused FP64 form equals the plain path bit for bit for every
    tile width there."""
    p = cube32_params()
    p.NX, p.Ny, p.Nz = dims
    p.src_x0, p.src_x1 = max(0, dims[0] // 2 - 2), dims[0] // 2 + 1
    p.src_z0, p.src_z1 = max(0, dims[2] // 2 - 2), dims[2] // 2 + 1
    g = random_state(p, np.random.default_rng(sum(dims)), media=False)
    Pp, Sp = plain_run(hip, p, g, 6)
    for tile in ((64, 4, 1), (64, 8, 1), (128, 4, 1), (32, 8, 1), (16, 16, 1), (64, 8, 2)):
        Pf, Sf, _ = fused_run(hip, p, g, 6, kz=(5, 1), tile=tile)
        assert np.array_equal(Pf, Pp, equal_nan=True), tile
        assert np.array_equal(Sf, Sp, equal_nan=True), tile


def test_fused_fp64_equals_plain_per_cell_fields(hip):
    p, g = lognormal_like(40)
    Pp, Sp = plain_run(hip, p, g, 30)
    for kz in ((1, 0), (7, 1), (64, 0)):
        Pf, Sf, _ = fused_run(hip, p, g, 30, kz=kz)
        assert np.array_equal(Pf, Pp), kz
        assert np.array_equal(Sf, Sp), kz


def test_fused_fp64_one_step_random_state_reference_media(hip, orc):
    """Random state, two media with P_d != 0 (generic pow), one step, against
    the oracle's full step (u skipped: not on the P/S path)."""
    p = ref_params(21, 17, 13, source=2)
    g = random_state(p, np.random.default_rng(12))
    Pf, Sf, _ = fused_run(hip, p, g, 1)
    want = g.copy()
    orc.run(want, p, 1, with_u=0)
    assert max_rel(Pf, want["P1"]) <= 1e-12
    assert max_rel(Sf, want["S2"]) <= 1e-12


def test_fused_fp64_cube32_trajectory_vs_oracle(hip, orc):
    p = cube32_params()
    g = HostGrid(p)
    orc.initial_data(g, p)
    Pf, Sf, mf = fused_run(hip, p, g, 500, graph=True)
    orc.run(g, p, 500, with_u=0)
    assert max_rel(Pf, g["P1"]) <= 1e-10
    assert max_rel(Sf, g["S2"]) <= 1e-10
    mo = orc.mass(g, p)
    assert abs(mf[0] - mo[0]) <= 1e-12 * mo[0]
    assert abs(mf[1] - mo[1]) <= 1e-12 * mo[1]


def errors_vs_oracle(P, S, m, g, mo):
    """Per-field errors of a form's result against the oracle's state g
    (mass mo): S2 L-inf, P1 relative L-inf, water / NAPL / total mass."""
    return {"S2_linf": float(np.max(np.abs(S - g["S2"]))), "P1_rel_linf": max_rel(P, g["P1"]),
            "water_mass_rel": abs(m[0] - mo[0]) / mo[0],
            "napl_mass_rel": abs(m[1] - mo[1]) / mo[1] if mo[1] else abs(m[1]),
            "total_mass_rel": abs(m[0] + m[1] - mo[0] - mo[1]) / (mo[0] + mo[1])}


TILES = [(16, 16, 1), (32, 8, 1), (64, 4, 1), (32, 16, 1), (64, 8, 1), (64, 8, 2), (32, 16, 2),
         (64, 16, 2), (128, 4, 1)]


@pytest.mark.parametrize("precision", [FP64, FAST, MIXED])
def test_every_tile_shape_gives_the_same_bits(hip, precision):
    """The autotuner may pick any tile shape, z-chunk and block order: all
    give the same bits (each cell's arithmetic does not depend on them)."""
    p, g = lognormal_like(41)
    P0, S0, m0 = fused_run(hip, p, g, 20, precision=precision)
    for tile in TILES:
        for kz in ((5, 0), (64, 1)):
            P, S, m = fused_run(hip, p, g, 20, precision=precision, kz=kz, tile=tile)
            assert np.array_equal(P, P0) and np.array_equal(S, S0), (tile, kz)
            assert m == m0


def test_mixed_cube32_trajectory_error(hip, orc):
    """Mixed form against the FP64 oracle on the homogeneous 32^3 cube, 500 steps."""
    p = cube32_params()
    g = HostGrid(p)
    orc.initial_data(g, p)
    Pm, Sm, mm = fused_run(hip, p, g, 500, precision=MIXED)
    orc.run(g, p, 500, with_u=0)
    e = errors_vs_oracle(Pm, Sm, mm, g, orc.mass(g, p))
    print("mixed 32^3 cube 500 steps vs oracle:", e)
    assert e["S2_linf"] <= 1e-4 and e["P1_rel_linf"] <= 1e-6
    assert e["water_mass_rel"] <= 1e-9 and e["total_mass_rel"] <= 1e-9
    assert e["napl_mass_rel"] <= 1e-9


def test_fast_cube32_trajectory_error(hip, orc):
    """Fast form against the FP64 oracle on the homogeneous 32^3 cube, 500 steps."""
    p = cube32_params()
    g = HostGrid(p)
    orc.initial_data(g, p)
    Pm, Sm, mm = fused_run(hip, p, g, 500, precision=FAST)
    orc.run(g, p, 500, with_u=0)
    eS = float(np.max(np.abs(Sm - g["S2"])))
    eP = max_rel(Pm, g["P1"])
    mo = orc.mass(g, p)
    e1, e2 = abs(mm[0] - mo[0]) / mo[0], abs(mm[1] - mo[1]) / mo[1]
    et = abs(mm[0] + mm[1] - mo[0] - mo[1]) / (mo[0] + mo[1])
    print(f"fast 32^3 cube 500 steps: S2 Linf {eS:.3e}, P1 rel Linf {eP:.3e}, "
          f"water mass {e1:.3e}, NAPL mass {e2:.3e}, total mass {et:.3e}")
    assert eS <= 1e-4
    assert eP <= 1e-6
    assert e1 <= 1e-9 and et <= 1e-9
    assert e2 <= 1e-9


@pytest.mark.parametrize("precision", [FAST, MIXED])
def test_heterogeneous_error(hip, orc, precision):
    p, g = lognormal_like(48, seed=9, sigma=1.5)
    Pm, Sm, mm = fused_run(hip, p, g, 100, precision=precision)
    orc.set_threads(8)
    orc.run(g, p, 100, with_u=0)
    orc.set_threads(1)
    eS = float(np.max(np.abs(Sm - g["S2"])))
    mo = orc.mass(g, p)
    e1, e2 = abs(mm[0] - mo[0]) / mo[0], abs(mm[1] - mo[1]) / mo[1]
    et = abs(mm[0] + mm[1] - mo[0] - mo[1]) / (mo[0] + mo[1])
    print(f"precision {precision} 48^3 sigma 1.5, 100 steps: S2 Linf {eS:.3e}, "
          f"P1 rel {max_rel(Pm, g['P1']):.3e}, water mass {e1:.3e}, NAPL mass {e2:.3e}, "
          f"total mass {et:.3e}")
    assert eS <= 1e-4 and e1 <= 1e-9 and et <= 1e-9 and e2 <= 1e-9


@pytest.mark.parametrize("precision", [FAST, MIXED])
def test_reference_media_trajectory_error(hip, orc, precision):
    """The fast and mixed forms on the reference's own media table (lambda 2.7 / 2.0,
    P_d 755 / 2060: the general Brooks-Corey powers and the Newton with
    capillary terms) on a random 0/1 media map, 100 steps, against the FP64
    oracle, to the same bars (FP32 powers missed the NAPL one: 4.6e-8)."""
    p = ref_params(24, 20, 16, source=2)
    g = HostGrid(p)
    orc.initial_data(g, p)
    g.a["media"][...] = np.random.default_rng(3).integers(0, 2, g["P1"].shape)
    Pm, Sm, mm = fused_run(hip, p, g, 100, precision=precision)
    orc.run(g, p, 100, with_u=0)
    eS = float(np.max(np.abs(Sm - g["S2"])))
    eP = max_rel(Pm, g["P1"])
    mo = orc.mass(g, p)
    e1, e2 = abs(mm[0] - mo[0]) / mo[0], abs(mm[1] - mo[1]) / mo[1]
    et = abs(mm[0] + mm[1] - mo[0] - mo[1]) / (mo[0] + mo[1])
    print(f"precision {precision} reference media 100 steps: S2 Linf {eS:.3e}, "
          f"P1 rel Linf {eP:.3e}, water mass {e1:.3e}, NAPL mass {e2:.3e}, total mass {et:.3e}")
    assert mo[1] > 0
    assert eS <= 1e-4 and eP <= 1e-6
    assert e1 <= 1e-9 and e2 <= 1e-9 and et <= 1e-9


@pytest.mark.parametrize("precision", [FP64, FAST, MIXED])
def test_member_same_alone_and_in_mix(hip, precision):
    import importlib
    W = importlib.import_module("shmem-cuda-filtration_amd").workloads
    rng = np.random.default_rng(4)
    members = []
    for q, (nx, ny, nz) in enumerate([(24, 16, 20), (17, 33, 9), (40, 8, 12), (9, 9, 9)]):
        p = W.base_params(nx, ny, nz, h=W.INI_H)
        W.set_patch(p, int(rng.integers(0, nx - 4)), int(rng.integers(0, nz - 4)))
        perm, poro = W.gen_fields(p, 100 + q, sigma=float(rng.uniform(0.5, 2)), corr=4.0,
                                  clip=2.5)
        members.append((p, W.initial_grid(p, perm, poro)))
    alone = [fused_run(hip, p, g, 40, precision=precision) for p, g in members]
    e = hip.ensemble(precision)
    try:
        ids = [e.add(p, g["P1"], g["S2"], g["perm"], g["poro"]) for p, g in members]
        e.capture(20)
        e.replay()
        e.replay()
        for mid, (Pa, Sa, ma) in zip(ids, alone):
            P, S = e.get(mid)
            assert np.array_equal(P, Pa) and np.array_equal(S, Sa)
            assert e.mass(mid) == ma
            assert e.steps(mid) == 40
    finally:
        e.close()


@pytest.mark.parametrize("precision", [FP64, FAST, MIXED])
def test_member_with_fields_same_beside_member_without(hip, precision):
    """A batch whose members all carry per-cell K and porosity and one medium
    runs a kernel that reads both fields without tests; a member without them
    in the batch switches it to the kernel that tests for them.  The member
    with fields gives the same bits either way."""
    import importlib
    W = importlib.import_module("shmem-cuda-filtration_amd").workloads
    p = W.base_params(26, 18, 22, h=W.INI_H)
    W.set_patch(p, 7, 9)
    perm, poro = W.gen_fields(p, 71, sigma=1.5, corr=4.0, clip=2.5)
    g = W.initial_grid(p, perm, poro)
    q = W.base_params(20, 14, 16, h=W.INI_H)
    W.set_patch(q, 5, 6)
    h = W.initial_grid(q, None, None)
    alone = fused_run(hip, p, g, 30, precision=precision)
    e = hip.ensemble(precision)
    try:
        a = e.add(p, g["P1"], g["S2"], perm, poro)
        e.add(q, h["P1"], h["S2"], None, None)
        e.step(30)
        P, S = e.get(a)
        assert np.array_equal(P, alone[0]) and np.array_equal(S, alone[1])
        assert e.mass(a) == alone[2]
    finally:
        e.close()


@pytest.mark.parametrize("precision", [FP64, FAST, MIXED])
def test_lambda2_member_same_beside_reference_media_member(hip, orc, precision):
    """A batch that holds a member with lambda != 2 or P_d != 0 runs the
    kernel with the general Brooks-Corey powers; a lambda = 2, P_d = 0 member
    alone runs the kernel without them.  Its results are the same bits either
    way, so a member's result still does not depend on the other members."""
    import importlib
    W = importlib.import_module("shmem-cuda-filtration_amd").workloads
    p = W.base_params(20, 18, 14, h=W.INI_H)
    perm, poro = W.gen_fields(p, 7, sigma=1.2, corr=4.0, clip=2.5)
    g = W.initial_grid(p, perm, poro)
    Pa, Sa, ma = fused_run(hip, p, g, 30, precision=precision)
    q = ref_params(16, 12, 10, source=2)
    h = HostGrid(q)
    orc.initial_data(h, q)
    h.a["media"][...] = np.random.default_rng(5).integers(0, 2, h["P1"].shape)
    e = hip.ensemble(precision)
    try:
        a = e.add(p, g["P1"], g["S2"], g["perm"], g["poro"])
        e.add(q, h["P1"], h["S2"], None, None, h.a["media"])
        e.step(30)
        P, S = e.get(a)
        assert np.array_equal(P, Pa) and np.array_equal(S, Sa)
        assert e.mass(a) == ma
    finally:
        e.close()


def test_continuous_batching_add_remove_between_steps(hip):
    """Members admitted and retired while others run keep their own results."""
    import importlib
    W = importlib.import_module("shmem-cuda-filtration_amd").workloads
    grids = []
    for q, n in enumerate((20, 28, 12)):
        p = W.base_params(n, n + 3, n - 2, h=W.INI_H)
        perm, poro = W.gen_fields(p, 300 + q, sigma=1.0, corr=4.0, clip=2.5)
        grids.append((p, W.initial_grid(p, perm, poro)))
    alone = [fused_run(hip, p, g, 15, precision=MIXED) for p, g in grids]
    e = hip.ensemble(MIXED)
    try:
        a = e.add(grids[0][0], grids[0][1]["P1"], grids[0][1]["S2"], grids[0][1]["perm"],
                  grids[0][1]["poro"])
        e.step(5)
        b = e.add(grids[1][0], grids[1][1]["P1"], grids[1][1]["S2"], grids[1][1]["perm"],
                  grids[1][1]["poro"])
        e.step(10)
        Pa, Sa = e.get(a)
        e.remove(a)
        c = e.add(grids[2][0], grids[2][1]["P1"], grids[2][1]["S2"], grids[2][1]["perm"],
                  grids[2][1]["poro"])
        e.step(5)
        Pb, Sb = e.get(b)
        e.step(10)
        Pc, Sc = e.get(c)
        assert np.array_equal(Pa, alone[0][0]) and np.array_equal(Sa, alone[0][1])
        assert np.array_equal(Pb, alone[1][0]) and np.array_equal(Sb, alone[1][1])
        assert np.array_equal(Pc, alone[2][0]) and np.array_equal(Sc, alone[2][1])
    finally:
        e.close()


def test_graph_replays_match_eager_steps(hip):
    p, g = lognormal_like(26)
    Pe, Se, _ = fused_run(hip, p, g, 12)
    e = hip.ensemble(0)
    try:
        mid = e.add(p, g["P1"], g["S2"], g["perm"], g["poro"])
        e.capture(3)
        for r in range(4):
            e.replay()
            e.mass(mid)  # eager work and a sync between replays
        P, S = e.get(mid)
        assert np.array_equal(P, Pe) and np.array_equal(S, Se)
    finally:
        e.close()


def test_ensemble_rejects_degenerate_grid(hip):
    p = ref_params(9, 8, 2)
    g = HostGrid(p)
    e = hip.ensemble(0)
    try:
        with pytest.raises(Exception):
            e.add(p, g["P1"], g["S2"])
    finally:
        e.close()
