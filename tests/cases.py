"""Parameter sets and seeded states shared by the tests."""
import importlib

import numpy as np

pkg = importlib.import_module("shmem-cuda-filtration_amd")
abi = pkg.abi
HostGrid = pkg.grid.HostGrid
FIELD_NAMES = abi.FIELD_NAMES
FIELD_NAMES_ALL = abi.FIELD_NAMES + ("media",)

# defines.ini of the reference (defines.ini:5-49): the ini-controlled members.
INI = dict(h1=0.02045, h2=0.01969, h3=0.01969, dt=0.001, tau=0.002,
           l_n=8.528784e-10, l_w=2e-7, c=1407.0, beta1=1.0e-4, beta2=1.0e-8,
           save_plots=100000, print_screen=1000)


def ref_params(nx, ny, nz, source=1, **kw):
    """Parameters the compiled reference can express: its compile-time
    constants and media table, ini values, a (2*source+1)^2 patch."""
    p = abi.default_params_py()
    for k, v in INI.items():
        setattr(p, k, v)
    p.NX, p.Ny, p.Nz, p.source = nx, ny, nz, source
    for k, v in kw.items():
        setattr(p, k, v)
    return abi.finalize_patch_py(p)


def cube32_params(n=32, beta1=1.0e-4, beta2=1.0e-8):
    """Configuration 1 (32^3 over a 1 m cube, h = 1/31): K = 6.64e-11,
    phi = 0.4, lambda = 2, S_wr = 0.1, P_d = 0, Sn = 0.5 on the 4x4 patch
    I, k in [14, 17], dt = 1e-4.  The stated compressibility 4.4e-10 is
    unstable for this explicit scheme (NaN by step 6); beta defaults to the
    reference defines.ini values."""
    p = abi.default_params_py()
    h = 1.0 / (n - 1)
    p.h1 = p.h2 = p.h3 = h
    p.dt = 1e-4
    p.beta1, p.beta2 = beta1, beta2
    p.NX = p.Ny = p.Nz = n
    for q in range(2):
        p.K[q], p.lambda_[q], p.S_wr[q], p.m[q], p.P_d[q] = 6.64e-11, 2.0, 0.1, 0.4, 0.0
    p.S2_gr = 0.5
    c = n // 2
    p.src_x0, p.src_x1, p.src_z0, p.src_z1 = c - 2, c + 1, c - 2, c + 1
    p.save_plots, p.print_screen = 100000, 1000
    return p


def random_state(p, rng, localNx=None, media=True, perm=False, poro=False):
    """Every field filled with plausible random values (P in [1e5, 1.3e5],
    S2 in [0, 0.6], a random 0/1 media map), so that one stage call exercises
    all of its branches."""
    g = HostGrid(p, localNx, with_perm=perm, with_poro=poro)
    sh = g["P1"].shape
    for n in abi.FIELD_NAMES:
        g[n] = rng.uniform(0.5, 2.0, sh)
    g["P1"] = rng.uniform(1e5, 1.3e5, sh)
    g["P2"] = g["P1"] + rng.uniform(0, 100, sh)
    g["S2"] = rng.uniform(0, 0.6, sh)
    g["ro1"] = rng.uniform(990, 1010, sh)
    g["ro2"] = rng.uniform(1450, 1470, sh)
    g["Xi1"] = -rng.uniform(1e-9, 1e-7, sh)
    g["Xi2"] = -rng.uniform(1e-9, 1e-7, sh)
    g["roS1"] = g["ro1"] * (1 - g["S2"]) * rng.uniform(0.999, 1.001, sh)
    g["roS2"] = g["ro2"] * g["S2"] * rng.uniform(0.999, 1.001, sh)
    if media:
        g.a["media"][...] = rng.integers(0, 2, sh)
    if perm:
        g["perm"] = 6.64e-11 * np.exp(rng.normal(0, 1, sh))
    if poro:
        g["poro"] = rng.uniform(0.35, 0.45, sh)
    return g


def distinct_media_params(nx, ny, nz, source=1):
    """Reference-expressible parameters whose two media differ in every
    entry (the reference's own defines.h table already does)."""
    return ref_params(nx, ny, nz, source)


def max_rel(a, b, mask=None):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    d = np.max(np.abs(a - b)) if a.size else 0.0
    s = np.max(np.abs(b)) if b.size else 1.0
    return d / s if s else d
