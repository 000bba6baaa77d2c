"""Host-side pieces of the product: the defines.ini reader, the driver
helpers and initial state, the seeded field generator, the stability guard.
Compared against the reference's own code (oracle/_ref) where it has a
counterpart, against closed forms otherwise."""
import ctypes
import os

import numpy as np
import pytest

from cases import FIELD_NAMES_ALL, HostGrid, ref_params

INI_TEXT = (
    "#######\r\n# comment\r\nH1=0.02045\r\n#\r\nH2=0.01969\r\nH3=0.01969\r\nDT=0.001\r\n"
    "TAU=0.002\r\nL_N=8.528784e-10\r\nL_W=2e-7\r\nC=1407\r\nBETA1=1.0e-4\r\nBETA2=1.0e-8\r\n"
    "#BETA2=1.0e-4\r\nNX=15\r\nNY=14\r\nNZ=13\r\nSOURCE=3" + " " * 120 + "\r\n#\r\n"
    "SAVE_PLOTS=100000\r\nPRINT_SCREEN=1000\r\nK[0]=6.64e-11\r\nUNKNOWN_KEY=5\r\nNX=16\r\n"
)


def read_ini(pkg, text, tmp_path, base=None):
    lib = pkg.abi.load_library()
    tmp_path.mkdir(parents=True, exist_ok=True)
    path = tmp_path / "defines.ini"
    path.write_bytes(text.encode())
    p = base.copy() if base is not None else pkg.abi.default_params_py()
    rc = lib.filt_read_defines(str(path).encode(), ctypes.byref(p))
    assert rc == 0
    return p


def test_read_defines_matches_reference_parser(pkg, ref, tmp_path):
    import oracle
    p = read_ini(pkg, INI_TEXT, tmp_path / "a")
    c = ref.read_defines(INI_TEXT, str(tmp_path / "b"))
    want = oracle.from_ref_consts(c)
    for name, _ in oracle.RefConsts._fields_:
        a, b = getattr(p, name), getattr(want, name)
        if hasattr(a, "__len__"):
            assert list(a)[:2] == list(b)[:2], name
        else:
            assert a == b, name
    assert p.NX == 16  # last occurrence wins
    assert p.source == 3


def test_read_defines_on_reference_ini_file(pkg, ref, tmp_path):
    """The reference's shipped defines.ini, read by both parsers."""
    src = "/root/reference/defines.ini"
    if not os.path.exists(src):
        pytest.skip("reference tree absent")
    import oracle
    text = open(src, "rb").read().decode("cp1251")
    p = read_ini(pkg, text, tmp_path / "a")
    want = oracle.from_ref_consts(ref.read_defines(text.encode("cp1251"), str(tmp_path / "b")))
    for name, _ in oracle.RefConsts._fields_:
        a, b = getattr(p, name), getattr(want, name)
        if hasattr(a, "__len__"):
            assert list(a)[:2] == list(b)[:2], name
        else:
            assert a == b, name
    # the documented-but-ignored keys carry the compiled defaults anyway
    assert p.mu1 == 0.001 and p.ro02 == 1460 and p.S2_gr == 0.4


def test_read_defines_extensions(pkg, tmp_path):
    text = ("NX=20\nNZ=10\nSOURCE=1\nS2_GR=0.5\nLAMBDA[3]=2.0\nP_D[1]=0\nNEWTON_ITERATIONS=9\n"
            "SRC_X0=8\nSRC_X1=11\nSRC_Z0=3\nSRC_Z1=6\nN_MEDIA=4\nMU2=0.0009\n")
    p = read_ini(pkg, text, tmp_path)
    assert p.S2_gr == 0.5 and p.lambda_[3] == 2.0 and p.P_d[1] == 0
    assert p.newton_iterations == 9 and p.n_media == 4
    assert (p.src_x0, p.src_x1, p.src_z0, p.src_z1) == (8, 11, 3, 6)
    q = read_ini(pkg, "NX=20\nNZ=10\nSOURCE=1\n", tmp_path)
    assert (q.src_x0, q.src_x1, q.src_z0, q.src_z1) == (9, 11, 4, 6)


def test_read_defines_missing_file(pkg):
    lib = pkg.abi.load_library()
    p = pkg.abi.default_params_py()
    assert lib.filt_read_defines(b"/nonexistent/defines.ini", ctypes.byref(p)) == \
        pkg.abi.FILT_ERR_IO


def test_default_params_match_reference_constants(pkg, ref):
    lib = pkg.abi.load_library()
    p = pkg.abi.FiltParams()
    lib.filt_default_params(ctypes.byref(p))
    c = ref.constants()
    assert [p.mu1, p.mu2, p.ro01, p.ro02, p.S2_gr, p.P_atm, p.g_const,
            p.newton_iterations] == c
    q = pkg.abi.default_params_py()
    assert bytes(p) == bytes(q)


@pytest.mark.parametrize("NX,size", [(15, 1), (16, 3), (31, 4)])
def test_driver_helpers_match_reference(pkg, ref, NX, size):
    lib = pkg.abi.load_library()
    p = ref_params(NX, 5, 5)
    for rank in range(size):
        a, b = ctypes.c_int(), ctypes.c_int()
        lib.filt_N_to_local(ctypes.byref(a), ctypes.byref(b), size, rank, ctypes.byref(p))
        assert (a.value, b.value) == ref.N_to_local(size, rank, p)
        for i in range(a.value):
            assert lib.filt_i_to_I(i, rank, size, ctypes.byref(p)) == ref.i_to_I(i, rank, size, p)
            assert lib.filt_is_active_point(i, a.value, rank, size) == \
                ref.is_active_point(i, a.value, rank, size)


@pytest.mark.parametrize("rank,size", [(0, 1), (1, 3), (2, 3)])
def test_initial_data_matches_reference(pkg, orc, ref, rank, size):
    lib = pkg.abi.load_library()
    p = ref_params(17, 9, 11, source=2)
    lnx, _ = ref.N_to_local(size, rank, p)
    a, b, c = HostGrid(p, lnx), HostGrid(p, lnx), HostGrid(p, lnx)
    f = a.fields()
    assert lib.filt_initial_data(ctypes.byref(f), ctypes.byref(p), lnx, rank, size) == 0
    ref.initial_data(b, p, rank, size)
    orc.initial_data(c, p, rank, size)
    for n in ("P1", "S2", "x", "y", "z", "media"):
        assert np.array_equal(a[n], b[n]), n
        assert np.array_equal(c[n], b[n]), n


def test_field_generator_statistics_and_determinism(pkg):
    W = pkg.workloads
    p = W.base_params(48, 40, 32)
    perm, poro = W.gen_fields(p, 7, sigma=1.3, corr=4.0, threads=1)
    perm2, poro2 = W.gen_fields(p, 7, sigma=1.3, corr=4.0, threads=5)
    assert np.array_equal(perm, perm2) and np.array_equal(poro, poro2)
    z = (np.log(perm) - np.log(W.K_MEAN)) / 1.3
    assert abs(z.mean()) < 1e-12 and abs(z.std() - 1) < 1e-12
    assert 0.35 <= poro.min() and poro.max() <= 0.45
    assert abs(poro.mean() - 0.40) < 2e-3
    # Gaussian covariance exp(-r^2/(2 l^2)) at lag l along x: e^-1/2 ~ 0.61
    c4 = np.mean(z * np.roll(z, -4, axis=2))
    assert 0.45 < c4 < 0.75, c4
    c16 = np.mean(z * np.roll(z, -16, axis=2))
    assert abs(c16) < 0.15, c16
    p3, _ = W.gen_fields(p, 8, sigma=1.3, corr=4.0)
    assert not np.array_equal(p3, perm)
    clipped, _ = W.gen_fields(p, 7, sigma=1.3, corr=4.0, clip=1.0)
    zc = (np.log(clipped) - np.log(W.K_MEAN)) / 1.3
    assert zc.max() <= 1 + 1e-12 and zc.min() >= -1 - 1e-12


def test_stability_guard(pkg):
    from cases import cube32_params
    W = pkg.workloads
    stated = cube32_params(beta1=4.4e-10, beta2=4.4e-10)
    r = W.stability(stated)
    assert 300 < r < 600  # ~435: 4*dt*K/(mu phi beta)*3/h^2
    assert W.stability(cube32_params()) < 2e-2
    p = W.base_params(16, 16, 16, h=W.INI_H)
    perm = np.full((16, 16, 16), 6.64e-11)
    perm[3, 4, 5] = 6.64e-11 * 1000
    poro = np.full((16, 16, 16), 0.4)
    assert W.stability(p, perm, poro) == pytest.approx(1000 * W.stability(p), rel=1e-12)


def test_workload_initial_state(pkg):
    W = pkg.workloads
    p = W.cube32_params()
    g = W.initial_grid(p)
    assert g["S2"].sum() == pytest.approx(16 * 0.5)
    assert np.all(g["S2"][14:18, 0, 14:18] == 0.5)
    assert g["P1"][0, 3, 0] == 100000 + 3 * 1000 * 9.8 * p.h2


@pytest.mark.parametrize("shape", [(512, 512, 256), (256, 512, 512)])
def test_slab_workload_passes_the_guard(pkg, shape):
    """The 512x512x256 slab (sigma 1.5, correlation 16) and its transposed
    orientation: with ln K clipped at +-2 sigma the guard stays below 2 (at
    +-2.5 sigma it would read ~2.6)."""
    W = pkg.workloads
    p, perm, poro = W.slab_workload(*shape)
    assert perm.shape == shape[::-1]
    r = W.stability(p, perm, poro)
    assert 0.5 < r < 1.5


def test_host_allocation_failure_is_a_status_not_an_abort(pkg):
    """C++ exceptions never cross the C ABI: a field of 10^15 cells cannot be
    allocated, and the call returns FILT_ERR_NOMEM with a message."""
    lib = pkg.abi.load_library()
    out = np.zeros(8)
    rc = lib.filt_gen_lognormal(out.ctypes.data, 100000, 100000, 100000, 1, 6.64e-11, 1.0, 4.0,
                                0.0, 1)
    assert rc == pkg.abi.FILT_ERR_NOMEM
    assert b"memory" in lib.filt_last_error()


def test_ensemble_check_refusals_without_a_device(pkg):
    """filt_ensemble_check (what filt_ensemble_add and filt_serve_submit
    refuse) runs on the host alone."""
    import ctypes
    W = pkg.workloads
    lib = pkg.abi.load_library()
    p = W.base_params(20, 16, 12, h=W.INI_H)
    chk = lambda q, prec=1, perm=None: lib.filt_ensemble_check(
        ctypes.byref(q), prec, None if perm is None else perm.ctypes.data, None)
    for prec in (0, 1, 2):
        assert chk(p, prec) == pkg.abi.FILT_OK
    assert chk(p, 3) == pkg.abi.FILT_ERR_INVALID
    q = p.copy()
    q.Nz = 2
    assert chk(q) == pkg.abi.FILT_ERR_UNSUPPORTED
    q = p.copy()
    q.roS_scheme = pkg.abi.FILT_SCHEME_DIVGRAD
    assert chk(q) == pkg.abi.FILT_ERR_UNSUPPORTED
    q = p.copy()
    q.NX, q.Ny, q.Nz = 1024, 1024, 512  # 4 GiB per FP64 array
    assert chk(q) == pkg.abi.FILT_ERR_UNSUPPORTED
    q = p.copy()
    q.mu2 = 1e31  # outside the FP64 form's exact-division range (an immobile NAPL)
    assert chk(q, 0) == pkg.abi.FILT_ERR_UNSUPPORTED
    assert chk(q, 1) == pkg.abi.FILT_OK
    q = p.copy()
    q.dt *= 1e5
    assert chk(q) == pkg.abi.FILT_ERR_UNSTABLE
    import numpy as np
    perm = np.full((12, 16, 20), W.K_MEAN)
    assert chk(p, 1, perm) == pkg.abi.FILT_OK
    perm[3, 4, 5] = W.K_MEAN * 1e6  # one cell far too permeable for dt
    assert chk(p, 1, perm) == pkg.abi.FILT_ERR_UNSTABLE
