"""The benchmark configurations as parameter sets and seeded initial states.

Shared deviations from the configuration text, applied identically to every
path and to the oracle:
  * compressibility beta1 = 1e-4, beta2 = 1e-8 (the reference's defines.ini
    values) instead of 4.4e-10: with 4.4e-10 the explicit scheme's
    amplification number is ~435 at 32^3 (stable needs <= 2) and the state
    reaches NaN by step 6;
  * 256^3 and larger grids use the defines.ini node spacing
    (h1 = 0.02045 m, h2 = h3 = 0.01969 m) instead of a 1 m cube;
  * log-normal ln K is clipped at +-2.5 sigma: on the 5 m deep 256-cell
    column the artificial compressibility makes water ~6x denser at the
    bottom than at the top, which multiplies the diffusivity, and the
    unclipped tail (K up to ~e^4.7 K_mean) then drives the local
    amplification number to ~3.7 (NaN within 100 steps, oracle and GPU
    alike); clipped, the guard reads ~0.4.
filt_stability_number reports the guard value for each set.
"""
import ctypes

import numpy as np

from . import abi
from .grid import HostGrid

K_MEAN = 6.64e-11
INI_H = (0.02045, 0.01969, 0.01969)
STABLE_BETA = (1.0e-4, 1.0e-8)


def base_params(nx, ny, nz, h=None, dt=1e-4, beta=STABLE_BETA, sn=0.5, patch=4):
    """Fluids of the configurations: K = 6.64e-11, phi = 0.4, water
    ro0 = 1000 / mu = 1e-3, NAPL ro0 = 1460 / mu = 0.9e-3, Brooks-Corey
    lambda = 2, S_wr = 0.1, P_d = 0 (one medium); Sn on a patch x patch
    source centred on the top face."""
    p = abi.default_params_py()
    if h is None:
        h = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    p.h1, p.h2, p.h3 = h
    p.dt = dt
    p.beta1, p.beta2 = beta
    p.NX, p.Ny, p.Nz = nx, ny, nz
    for q in range(abi.FILT_MAX_MEDIA):
        p.K[q], p.lambda_[q], p.S_wr[q], p.m[q], p.P_d[q] = K_MEAN, 2.0, 0.1, 0.4, 0.0
    p.n_media = 1
    p.S2_gr = sn
    set_patch(p, nx // 2 - patch // 2, nz // 2 - patch // 2, patch)
    p.save_plots, p.print_screen = 100000, 1000
    return p


def set_patch(p, x0, z0, size=4):
    p.src_x0, p.src_x1 = x0, x0 + size - 1
    p.src_z0, p.src_z1 = z0, z0 + size - 1
    return p


def cube32_params():
    """32^3 over a 1 m cube, homogeneous, 500 steps."""
    return base_params(32, 32, 32)


def lognormal_params(n=256):
    """256^3, per-cell log-normal K (sigma 1, correlation 8 cells) and
    porosity U[0.35, 0.45], 200 steps."""
    return base_params(n, n, n, h=INI_H)


def slab_params(nx=512, ny=512, nz=256):
    """One rank's share: 512 x 512 x 256, sigma 1.5, correlation 16 cells."""
    return base_params(nx, ny, nz, h=INI_H)


def gen_fields(p, seed, sigma, corr, clip=0.0, poro=(0.35, 0.45), threads=0):
    """Per-cell permeability and porosity (x-fastest, shape (Nz, Ny, NX))."""
    lib = abi.load_library()
    shape = (p.Nz, p.Ny, p.NX)
    perm = np.empty(shape, np.float64)
    phi = np.empty(shape, np.float64)
    abi.check(lib, lib.filt_gen_lognormal(perm.ctypes.data, p.NX, p.Ny, p.Nz, seed, K_MEAN,
                                          sigma, corr, clip, threads), "filt_gen_lognormal")
    abi.check(lib, lib.filt_gen_uniform(phi.ctypes.data, phi.size, seed ^ 0x5F3759DF,
                                        poro[0], poro[1]), "filt_gen_uniform")
    return perm, phi


def initial_grid(p, perm=None, poro=None, coords=False):
    """HostGrid with the initial state of main.cpp:97-138 (filt_initial_data)
    and the property fields attached."""
    lib = abi.load_library()
    g = HostGrid(p, with_perm=perm is not None, with_poro=poro is not None)
    if perm is not None:
        g["perm"] = perm
    if poro is not None:
        g["poro"] = poro
    f = g.fields()
    if not coords:
        f.x = f.y = f.z = None
    abi.check(lib, lib.filt_initial_data(ctypes.byref(f), ctypes.byref(p), g.nx, 0, 1),
              "filt_initial_data")
    return g


def stability(p, perm=None, poro=None):
    lib = abi.load_library()
    n = 0 if perm is None else perm.size
    return lib.filt_stability_number(ctypes.byref(p),
                                     None if perm is None else perm.ctypes.data,
                                     None if poro is None else poro.ctypes.data, n)


def lognormal_workload(n=256, seed=2024):
    p = lognormal_params(n)
    perm, poro = gen_fields(p, seed, sigma=1.0, corr=8.0, clip=2.5)
    return p, perm, poro


def slab_workload(nx=512, ny=512, nz=256, seed=3033):
    """The slab with ln K clipped at +-2 sigma: its 10 m water column makes
    water ~11x denser at the bottom, and at +-2.5 sigma the guard reads 2.65
    (unstable); at +-2 sigma it reads 1.25."""
    p = slab_params(nx, ny, nz)
    perm, poro = gen_fields(p, seed, sigma=1.5, corr=16.0, clip=2.0)
    return p, perm, poro


ENSEMBLE_EDGES = (48, 64, 96, 128, 160)


def ensemble_member(rng, edges=ENSEMBLE_EDGES, corr=8.0, seed=0, shape=None, threads=0,
                    clip=2.0, r_max=1.5):
    """One realization of the batched-ensemble / serving configurations: each
    axis drawn from `edges` (or the given shape), sigma ~ U[0.5, 2],
    correlation `corr` cells, porosity U[0.35, 0.45], a 4x4 source patch at a
    random position on the top face, ln K clipped at +-`clip` sigma (with
    sigma = 2 on a 160-cell column the unclipped tail exceeds the stability
    limit).  The field keeps its stated sigma and clip; where the stability
    guard still reads above r_max the member's time step is lowered to
    dt * r_max / r (the guard is linear in dt), so every member is stable
    and carries the variance its sigma states.  Returns the member with its
    sigma, clip, guard value r (after any dt change) and dt."""
    if shape is None:
        shape = tuple(int(rng.choice(edges)) for _ in range(3))
    nx, ny, nz = shape
    p = base_params(nx, ny, nz, h=INI_H)
    set_patch(p, int(rng.integers(0, nx - 3)), int(rng.integers(0, nz - 3)))
    sigma = float(rng.uniform(0.5, 2.0))
    perm, poro = gen_fields(p, seed, sigma=sigma, corr=corr, clip=clip, threads=threads)
    r = stability(p, perm, poro)
    if r > r_max:
        p.dt = p.dt * r_max / r
        r = stability(p, perm, poro)
    return {"params": p, "perm": perm, "poro": poro, "sigma": sigma, "clip": clip,
            "stability_r": r, "dt": p.dt, "shape": shape}


def ensemble_members(n=96, seed=4096):
    """Configuration 4: n realizations with edges from {48,64,96,128,160}."""
    rng = np.random.default_rng(seed)
    return [ensemble_member(rng, seed=seed * 1000 + q) for q in range(n)]
