"""Checkpoint / resume and plot output against the reference's own driver
functions (compiled into oracle/_ref): save (main.cpp:485-537) and restore
(main.cpp:540-569) of save/save.dat, and save_data_plots (main.cpp:319-482),
byte for byte, on host arrays.  The device forms are tested on the GPU."""
import ctypes
import os

import numpy as np
import pytest

from cases import FIELD_NAMES_ALL, HostGrid, ref_params


def random_grid(p, localNx, seed, specials=False):
    """Every array distinct and random; media 0/1.  specials plants -0.0,
    NaN and infinities into the velocities the plots print."""
    rng = np.random.default_rng(seed)
    g = HostGrid(p, localNx)
    for n in FIELD_NAMES_ALL:
        if n == "media":
            g.a[n][...] = rng.integers(0, 2, g.a[n].shape)
        else:
            g.a[n][...] = rng.normal(0, 1, g.a[n].shape) * 10.0 ** rng.integers(-6, 6, g.a[n].shape)
    # positive P1, S2: the reference's TEST build prints a warning per cell
    # otherwise (test_correct_P1_S2, main.cpp:582-593)
    g.a["P1"][...] = np.abs(g.a["P1"]) + 1.0
    g.a["S2"][...] = np.abs(g.a["S2"])
    if specials:
        u = g.a["u2y"].reshape(-1)
        u[:4] = [-0.0, np.nan, np.inf, -np.inf]
        g.a["u2x"].reshape(-1)[5] = -np.nan
    return g


def files_under(d):
    out = {}
    for root, _, names in os.walk(d):
        for n in names:
            path = os.path.join(root, n)
            with open(path, "rb") as fh:
                out[os.path.relpath(path, d)] = fh.read()
    return out


def check(lib, rc):
    assert rc == 0, lib.filt_last_error().decode()


@pytest.mark.parametrize("dims,size", [((7, 6, 5), 1), ((11, 5, 4), 3), ((9, 4, 1), 2)])
def test_save_file_matches_reference(pkg, ref, tmp_path, dims, size):
    lib = pkg.abi.load_library()
    p = ref_params(*dims)
    (tmp_path / "ref").mkdir()
    ours = tmp_path / "ours" / "save" / "save.dat"
    for rank in range(size):
        nx, _ = ref.N_to_local(size, rank, p)
        g = random_grid(p, nx, 10 * rank + size)
        ref.save(tmp_path / "ref", g, p, j=1234 + rank, rank=rank, size=size)
        check(lib, lib.filt_save_host(str(ours).encode(), ctypes.byref(g.fields()), p, nx, rank,
                                      size, 1234 + rank, 1 if rank == 0 else 0))
    want = (tmp_path / "ref" / "save" / "save.dat").read_bytes()
    assert ours.read_bytes() == want
    n = sum(ref.N_to_local(size, r, p)[0] for r in range(size)) * dims[1] * dims[2]
    assert len(want) == size * 4 + n * (7 * 8 + 4)


def test_save_coordinates_default_to_initial_data(pkg, orc, tmp_path):
    """x, y, z left null: the block holds the initial_data coordinates."""
    lib = pkg.abi.load_library()
    p = ref_params(8, 5, 6)
    g = random_grid(p, 8, 3)
    orc.initial_data(g, p)  # x, y, z of main.cpp:113-116
    a, b = tmp_path / "a.dat", tmp_path / "b.dat"
    check(lib, lib.filt_save_host(str(a).encode(), ctypes.byref(g.fields()), p, 8, 0, 1, 7, 1))
    f = g.fields()
    f.x = f.y = f.z = None
    check(lib, lib.filt_save_host(str(b).encode(), ctypes.byref(f), p, 8, 0, 1, 7, 1))
    assert a.read_bytes() == b.read_bytes()


@pytest.mark.parametrize("size,rank", [(1, 0), (3, 1), (3, 2)])
def test_restore_matches_reference(pkg, ref, tmp_path, size, rank):
    lib = pkg.abi.load_library()
    p = ref_params(10, 5, 4)
    for r in range(size):  # a file written by the reference
        nx, _ = ref.N_to_local(size, r, p)
        ref.save(tmp_path, random_grid(p, nx, 50 + r), p, j=900 + r, rank=r, size=size)
    nx, _ = ref.N_to_local(size, rank, p)
    # the reference reads every earlier rank's block into the caller's
    # arrays, which its driver allocates NX/size+3 columns wide (main.cpp:214)
    wide = HostGrid(p, p.NX // size + 3)
    j_ref = ref.restore(tmp_path, wide, p, rank=rank, size=size)
    n_own = nx * p.Ny * p.Nz
    got = HostGrid(p, nx)
    j = ctypes.c_int(-1)
    path = str(tmp_path / "save" / "save.dat").encode()
    check(lib, lib.filt_restore_host(path, ctypes.byref(got.fields()), p, nx, rank, size,
                                     ctypes.byref(j)))
    assert j.value == j_ref == 900 + rank
    for n in ("P1", "S2", "x", "y", "z", "roS1_old", "roS2_old", "media"):
        assert np.array_equal(got[n].reshape(-1), wide[n].reshape(-1)[:n_own]), n


def test_restore_errors(pkg, tmp_path):
    lib = pkg.abi.load_library()
    p = ref_params(6, 4, 3)
    g = HostGrid(p)
    j = ctypes.c_int()
    missing = str(tmp_path / "none.dat").encode()
    assert lib.filt_restore_host(missing, ctypes.byref(g.fields()), p, 6, 0, 1,
                                 ctypes.byref(j)) == pkg.abi.FILT_ERR_IO
    short = tmp_path / "short.dat"
    short.write_bytes(b"\x01\x00\x00\x00" + b"\x00" * 100)
    assert lib.filt_restore_host(str(short).encode(), ctypes.byref(g.fields()), p, 6, 0, 1,
                                 ctypes.byref(j)) == pkg.abi.FILT_ERR_IO
    # a width that is not this rank's
    assert lib.filt_restore_host(missing, ctypes.byref(g.fields()), p, 5, 0, 1,
                                 ctypes.byref(j)) == pkg.abi.FILT_ERR_INVALID


@pytest.mark.parametrize("dims,size,t", [((7, 6, 5), 1, 0.25), ((12, 5, 4), 3, 1.5),
                                         ((6, 7, 1), 2, 12.0), ((5, 5, 5), 1, 3.125)])
def test_plots_match_reference(pkg, ref, tmp_path, dims, size, t):
    """Every plot file and media.dat equal the reference's, byte for byte
    (the reference's own index choices for the profile files included)."""
    lib = pkg.abi.load_library()
    p = ref_params(*dims)
    (tmp_path / "ref").mkdir()
    (tmp_path / "ours").mkdir()
    for rank in range(size):
        nx, _ = ref.N_to_local(size, rank, p)
        g = random_grid(p, nx, 7 * rank + 1, specials=(rank == 0))
        ref.save_data_plots(tmp_path / "ref", g, p, t, rank=rank, size=size)
        check(lib, lib.filt_save_plots_host(str(tmp_path / "ours").encode(),
                                            ctypes.byref(g.fields()), p, nx, rank, size, t))
    want, got = files_under(tmp_path / "ref"), files_under(tmp_path / "ours")
    assert sorted(got) == sorted(want)
    assert len(want) == 6
    for name in want:
        assert got[name] == want[name], name


def test_plots_many_threads_keep_line_order(pkg, ref, tmp_path):
    """A grid wide enough that the formatting splits over several rounds of
    threads: the files still equal the reference's."""
    lib = pkg.abi.load_library()
    p = ref_params(150, 6, 5)
    g = random_grid(p, 150, 99)
    (tmp_path / "ref").mkdir()
    ref.save_data_plots(tmp_path / "ref", g, p, 2.0)
    check(lib, lib.filt_save_plots_host(str(tmp_path / "o").encode(), ctypes.byref(g.fields()),
                                        p, 150, 0, 1, 2.0))
    assert files_under(tmp_path / "o") == files_under(tmp_path / "ref")
