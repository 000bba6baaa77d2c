"""Checkpoint / resume, plots and asynchronous snapshots through the device
entry points: the device forms write what the host forms (checked byte for
byte against the reference on the CPU) write for the same arrays; a run
resumed from a save file continues bit for bit; a snapshot taken while the
GPU keeps stepping writes the state of the moment it was taken."""
import ctypes
import os

import numpy as np
import pytest

from cases import HostGrid, ref_params, random_state

pytestmark = pytest.mark.gpu


def files_under(d):
    out = {}
    for root, _, names in os.walk(d):
        for n in names:
            with open(os.path.join(root, n), "rb") as fh:
                out[os.path.relpath(os.path.join(root, n), d)] = fh.read()
    return out


def host_copy(d, p):
    g = HostGrid(p, d.nx)
    d.download(g)
    return g


def check(lib, rc):
    assert rc == 0, lib.filt_last_error().decode()


def started(hip, p, steps, with_u=1):
    """A device grid at the initial state of main.cpp:97-138 after steps."""
    g = HostGrid(p)
    hip.lib.filt_initial_data(ctypes.byref(g.fields()), ctypes.byref(p), p.NX, 0, 1)
    d = hip.grid(p)
    d.upload(g)
    d.run(steps, with_u=with_u)
    return d


def test_device_save_and_plots_equal_host_forms(hip, tmp_path):
    lib = hip.lib
    p = ref_params(13, 9, 7, source=2)
    d = started(hip, p, 25)
    try:
        d.save(tmp_path / "dev" / "save.dat", j=25)
        d.save_plots(tmp_path / "devplots", 25 * p.dt)
        g = host_copy(d, p)
        f = g.fields()
        f.x = f.y = f.z = None  # the device form takes initial_data's coordinates
        check(lib, lib.filt_save_host(str(tmp_path / "host" / "save.dat").encode(),
                                      ctypes.byref(f), p, p.NX, 0, 1, 25, 1))
        check(lib, lib.filt_save_plots_host(str(tmp_path / "hostplots").encode(),
                                            ctypes.byref(f), p, p.NX, 0, 1, 25 * p.dt))
    finally:
        d.free()
    assert (tmp_path / "dev" / "save.dat").read_bytes() == \
        (tmp_path / "host" / "save.dat").read_bytes()
    a, b = files_under(tmp_path / "devplots"), files_under(tmp_path / "hostplots")
    assert len(a) == 6 and a == b


def test_resume_continues_bit_for_bit(hip, tmp_path):
    """20 steps straight == 12 steps, save, restore into a fresh grid, 8 more."""
    p = ref_params(16, 12, 10, source=2)
    rng = np.random.default_rng(4)
    media = rng.integers(0, 2, (p.Nz, p.Ny, p.NX)).astype(np.int32)
    g = HostGrid(p)
    hip.lib.filt_initial_data(ctypes.byref(g.fields()), ctypes.byref(p), p.NX, 0, 1)
    g.a["media"][...] = media
    a = hip.grid(p)
    b = hip.grid(p)
    try:
        a.upload(g)
        a.run(12)
        a.save(tmp_path / "save" / "save.dat", j=12)
        a.run(8, step0=12)
        j = b.restore(tmp_path / "save" / "save.dat")
        assert j == 12
        assert np.array_equal(b.get("media"), media)
        b.run(8, step0=j)
        for n in ("P1", "S2", "roS1", "roS2", "u2x", "u2y"):
            assert np.array_equal(a.get(n), b.get(n)), n
    finally:
        a.free()
        b.free()


def test_restore_missing_file_is_an_error(hip, tmp_path):
    p = ref_params(6, 5, 4)
    d = hip.grid(p)
    try:
        j = ctypes.c_int()
        rc = hip.lib.filt_restore(str(tmp_path / "none.dat").encode(), ctypes.byref(d.f),
                                  ctypes.byref(p), p.NX, 0, 1, ctypes.byref(j))
        assert rc == 5  # FILT_ERR_IO
    finally:
        d.free()


def test_u_on_output_steps_only_is_exact(hip):
    """A driver that skips u_calculation except on output steps gets the u
    the reference computes every step."""
    p = ref_params(14, 10, 9, source=2)
    a = started(hip, p, 0)
    b = started(hip, p, 0)
    try:
        a.run(15, with_u=1)
        b.run(14, with_u=0)
        b.time_step(15 * p.dt, with_u=1)
        for n in ("u1x", "u1y", "u1z", "u2x", "u2y", "u2z", "P1", "S2"):
            assert np.array_equal(a.get(n), b.get(n)), n
    finally:
        a.free()
        b.free()


def test_snapshot_writes_the_state_of_its_take(hip, pkg, tmp_path):
    """Take at step 10, keep stepping to 20, write asynchronously: the files
    equal the synchronous writers' output at step 10; restoring the snapshot
    rolls the grid back to step 10."""
    p = ref_params(20, 11, 9, source=2)
    d = started(hip, p, 10)
    s = pkg.device.Snapshot(d, what=pkg.abi.FILT_SNAP_SAVE | pkg.abi.FILT_SNAP_PLOTS)
    try:
        d.save(tmp_path / "sync" / "save.dat", j=10)
        d.save_plots(tmp_path / "sync_plots", 10 * p.dt)
        P10, S10 = d.get("P1"), d.get("S2")
        s.take(d, 10, 10 * p.dt)
        d.run(10, step0=10)  # the GPU keeps going while the writer works
        s.write_async(tmp_path / "async" / "save.dat", tmp_path / "async_plots")
        d.run(5, step0=20)
        s.wait()
        assert not np.array_equal(d.get("P1"), P10)
        assert s.restore(d) == 10
        assert np.array_equal(d.get("P1"), P10) and np.array_equal(d.get("S2"), S10)
    finally:
        s.close()
        d.free()
    assert (tmp_path / "async" / "save.dat").read_bytes() == \
        (tmp_path / "sync" / "save.dat").read_bytes()
    assert files_under(tmp_path / "async_plots") == files_under(tmp_path / "sync_plots")


def test_snapshot_reuse_and_kind_checks(hip, pkg, tmp_path):
    p = ref_params(9, 8, 7)
    d = started(hip, p, 3)
    s = pkg.device.Snapshot(d, what=pkg.abi.FILT_SNAP_SAVE)
    try:
        with pytest.raises(Exception):
            s.write_async(plot_dir=tmp_path / "x")  # no plot arrays in this snapshot
        for j in (3, 4):
            s.take(d, j)
            s.write_async(tmp_path / f"s{j}.dat")
            d.run(1, step0=j)
        s.wait()
        d.save(tmp_path / "now.dat", j=5)
        s.take(d, 5)
        s.write_async(tmp_path / "s5.dat")
        s.wait()
        assert (tmp_path / "s5.dat").read_bytes() == (tmp_path / "now.dat").read_bytes()
        assert (tmp_path / "s3.dat").read_bytes() != (tmp_path / "s4.dat").read_bytes()
    finally:
        s.close()
        d.free()


def test_failed_snapshot_write_is_reported_by_the_next_take(hip, pkg, tmp_path):
    """A write_async that fails (its folder cannot be created: a file stands
    in the way) is not lost when the caller takes the next snapshot before
    waiting: the take returns the write's FILT_ERR_IO once, and later calls
    succeed again."""
    p = ref_params(9, 8, 7)
    d = started(hip, p, 2)
    s = pkg.device.Snapshot(d, what=pkg.abi.FILT_SNAP_SAVE)
    blocker = tmp_path / "not_a_folder"
    blocker.write_bytes(b"x")
    try:
        s.take(d, 2)
        s.write_async(blocker / "save" / "save.dat")
        with pytest.raises(pkg.abi.FiltError) as err:
            s.take(d, 3)
        assert err.value.code == pkg.abi.FILT_ERR_IO
        assert "previous write failed" in str(err.value)
        s.take(d, 3)  # reported once; the snapshot works again
        s.write_async(tmp_path / "ok.dat")
        s.wait()
        d.save(tmp_path / "sync.dat", j=3)
        assert (tmp_path / "ok.dat").read_bytes() == (tmp_path / "sync.dat").read_bytes()
        # the same failure reported by wait, then nothing left to report
        s.write_async(blocker / "again.dat")
        with pytest.raises(pkg.abi.FiltError):
            s.wait()
        s.wait()
    finally:
        s.close()
        d.free()
