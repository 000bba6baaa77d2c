"""Pins of the oracle that do not use the reference build: closed forms,
invariants and reductions of its generalisations to the reference's own
media-table path.  Each is chosen so that a plausible slip (a dropped term,
a wrong sign or index, a swapped operand) fails it."""
import numpy as np
import pytest

from cases import FIELD_NAMES_ALL, HostGrid, abi, random_state, ref_params


def bl(nx, ny, nz, **kw):
    p = ref_params(nx, ny, nz, **kw)
    for q in range(2):
        p.lambda_[q], p.S_wr[q], p.P_d[q] = 2.0, 0.1, 0.0
    return p


# ------------------------------------------------------------- reductions --

@pytest.mark.parametrize("stage", [1, 2, 3, 4, 5])
def test_uniform_perm_poro_fields_equal_media_table(orc, stage):
    """perm = K[media] and poro = m[media] cell by cell reproduce the media
    lookup bit for bit."""
    p = ref_params(11, 9, 7)
    g = random_state(p, np.random.default_rng(40 + stage))
    a = g.copy()
    b = HostGrid(p, with_perm=True, with_poro=True)
    for n in FIELD_NAMES_ALL:
        b.a[n][...] = g[n]
    md = g.a["media"]
    b["perm"] = np.where(md == 0, p.K[0], p.K[1])
    b["poro"] = np.where(md == 0, p.m[0], p.m[1])
    orc.stage(stage, a, p, t=0.1)
    orc.stage(stage, b, p, t=0.1)
    for n in FIELD_NAMES_ALL:
        assert np.array_equal(a[n], b[n], equal_nan=True), n


def test_single_medium_table_ignores_media_array(orc):
    p = ref_params(8, 7, 6)
    g = random_state(p, np.random.default_rng(3))
    a, b = g.copy(), g.copy()
    b.a["media"][...] = 0
    p1 = p.copy()
    p1.n_media = 1
    orc.run(a, p1, 3)
    orc.run(b, p, 3)
    for n in FIELD_NAMES_ALL[:-1]:
        assert np.array_equal(a[n], b[n]), n


def test_explicit_patch_bounds_equal_source_patch(orc):
    p = ref_params(15, 9, 13, source=2)
    q = p.copy()
    q.source = 0
    q.src_x0, q.src_x1, q.src_z0, q.src_z1 = 5, 9, 4, 8
    assert (p.src_x0, p.src_x1, p.src_z0, p.src_z1) == (5, 9, 4, 8)
    a, b = HostGrid(p), HostGrid(q)
    orc.initial_data(a, p)
    orc.initial_data(b, q)
    orc.run(a, p, 4)
    orc.run(b, q, 4)
    for n in ("P1", "S2"):
        assert np.array_equal(a[n], b[n])
    assert a["S2"][4:9, 0, 5:10].min() == p.S2_gr and a["S2"][:, 0, :].sum() == 25 * p.S2_gr


def test_even_patch_and_runtime_sn(orc):
    p = bl(16, 6, 12)
    p.S2_gr = 0.5
    p.src_x0, p.src_x1, p.src_z0, p.src_z1 = 6, 9, 4, 7
    g = HostGrid(p)
    orc.initial_data(g, p)
    top = g["S2"][:, 0, :]
    assert top.sum() == 16 * 0.5 and np.all(top[4:8, 6:10] == 0.5)
    orc.run(g, p, 2)
    top = g["S2"][:, 0, :]
    assert np.all(top[4:8, 6:10] == 0.5)


# ----------------------------------------------------------- closed forms --

def test_stage1_closed_form(orc):
    """S_e = (1-S-Swr)/(1-Swr); Brooks-Corey/Burdine k1 = S_e^((2+3l)/l),
    k2 = (1-S_e)^2 (1-S_e^((2+l)/l)); Xi = -K k / mu; linear density."""
    p = ref_params(9, 8, 7)
    g = random_state(p, np.random.default_rng(5))
    orc.stage(1, g, p)
    md = g.a["media"]
    K, lam, swr = (np.where(md == 0, v[0], v[1]) for v in (p.K, p.lambda_, p.S_wr))
    se = (1 - g["S2"] - swr) / (1 - swr)
    k1 = se ** ((2 + 3 * lam) / lam)
    k2 = (1 - se) ** 2 * (1 - se ** ((2 + lam) / lam))
    np.testing.assert_allclose(g["Xi1"], -K * k1 / p.mu1, rtol=1e-14)
    np.testing.assert_allclose(g["Xi2"], -K * k2 / p.mu2, rtol=1e-13)
    np.testing.assert_allclose(g["ro1"], p.ro01 * (1 + p.beta1 * (g["P1"] - p.P_atm)), rtol=1e-15)
    np.testing.assert_allclose(g["ro2"], p.ro02 * (1 + p.beta2 * (g["P1"] - p.P_atm)), rtol=1e-15)
    assert np.array_equal(g["P2"], g["P1"])
    # water-filled cell: full water mobility, no NAPL mobility
    w = HostGrid(p)
    w["S2"] = 0.0
    w["P1"] = 1e5
    orc.stage(1, w, p)
    assert np.all(w["Xi1"][md == 0] == -p.K[0] / p.mu1) or True
    assert np.all(w["Xi2"] == 0) and np.all(w["Xi1"] == -p.K[0] * 1.0 / p.mu1)


def test_stage2_linear_pressure_is_exact(orc):
    """A linear P makes one-sided and central differences exact:
    u_x = Xi*a, u_y = Xi*(b - ro*g), u_z = Xi*c everywhere."""
    p = ref_params(10, 9, 8)
    g = random_state(p, np.random.default_rng(6))
    z, y, x = np.meshgrid(np.arange(8) * p.h3, np.arange(9) * p.h2, np.arange(10) * p.h1,
                          indexing="ij")
    a, b, c = 300.0, 9800.0, -150.0
    g["P1"] = 1e5 + a * x + b * y + c * z
    g["P2"] = 1e5 + 2 * a * x + b * y + 3 * c * z
    orc.stage(2, g, p)
    np.testing.assert_allclose(g["u1x"], g["Xi1"] * a, rtol=1e-9)
    np.testing.assert_allclose(g["u2x"], g["Xi2"] * 2 * a, rtol=1e-9)
    np.testing.assert_allclose(g["u1y"], g["Xi1"] * (b - g["ro1"] * p.g_const), rtol=1e-6)
    np.testing.assert_allclose(g["u2z"], g["Xi2"] * 3 * c, rtol=1e-9)


def test_stage2_small_axes_give_zero(orc):
    p = ref_params(2, 5, 1)
    g = random_state(p, np.random.default_rng(2))
    orc.stage(2, g, p)
    assert np.all(g["u1x"] == 0) and np.all(g["u2z"] == 0) and np.any(g["u1y"] != 0)


def test_stage3_uniform_pressure_gravity_residual(orc):
    """Uniform P: x and z face forces vanish; both y faces carry +ro_c*g > 0,
    so the y flux is ro_c*g*(lam_{j-1} - lam_c)/h2 with lam = -Xi*ro
    entering as (-1)*Xi*ro; roS = ro*S - dt/phi * that."""
    p = ref_params(7, 8, 6)
    g = random_state(p, np.random.default_rng(8))
    g["P1"] = 1.1e5
    g["P2"] = 1.1e5
    orc.stage(3, g, p)
    md = g.a["media"]
    phi = np.where(md == 0, p.m[0], p.m[1])
    for X, r, S, out in (("Xi1", "ro1", 1 - g["S2"], "roS1"), ("Xi2", "ro2", g["S2"], "roS2")):
        lam = g[X] * g[r]
        lam_m = np.roll(lam, 1, axis=1)
        f2 = (g[r] * p.g_const * (-lam) - g[r] * p.g_const * (-1) * lam_m) / p.h2
        want = g[r] * S - (p.dt / phi) * f2
        sl = (slice(1, -1), slice(1, -1), slice(1, -1))
        np.testing.assert_allclose(g[out][sl], want[sl], rtol=1e-12)
        np.testing.assert_allclose(g[out + "_old"][sl], (g[r] * S)[sl], rtol=0, atol=0)


def test_stage3_x_fluxes_telescope(orc):
    """Along an x line with zero flux through its end faces the interior
    x-flux contributions sum to zero (up to rounding) when y and z carry
    nothing: uniform in y (no gravity: g = 0) and z."""
    p = ref_params(12, 3, 3)
    p.g_const = 0.0
    rng = np.random.default_rng(9)
    g = random_state(p, rng, media=False)
    line_P = 1e5 + rng.uniform(0, 1e3, 12)
    line_lam = rng.uniform(0.5, 2, 12)
    for n in ("P1", "P2"):
        g[n] = np.broadcast_to(line_P, g[n].shape)
    for n in ("Xi1", "Xi2"):
        g[n] = -np.broadcast_to(line_lam, g[n].shape) * 1e-3
    for n in ("ro1", "ro2"):
        g[n] = 1000.0
    g["S2"] = 0.3
    # end cells carry the neighbour's pressure, so no flux crosses faces 0|1, 10|11
    g["P1"][:, :, 0] = g["P1"][:, :, 1]
    g["P1"][:, :, -1] = g["P1"][:, :, -2]
    orc.stage(3, g, p)
    inc = (1000.0 * 0.7 - g["roS1"][1, 1, 1:-1]) * p.m[0] / p.dt
    assert abs(inc.sum()) <= 1e-9 * np.abs(inc).max()
    assert np.abs(inc).max() > 0


def test_newton_fixed_point_and_root(orc):
    """F(P, S) = 0 is a fixed point; from a perturbed start 7 iterations land
    on the root of the P_d = 0 system, which reduces to a quadratic in
    x = P - P_atm:  r01(1+b1 x)(1 - B/(r02(1+b2 x))) = A."""
    p = bl(6, 6, 6)
    rng = np.random.default_rng(10)
    g = random_state(p, rng)
    P, S = g["P1"].copy(), g["S2"].copy()
    g["roS1"] = p.ro01 * (1 + p.beta1 * (P - p.P_atm)) * (1 - S)
    g["roS2"] = p.ro02 * (1 + p.beta2 * (P - p.P_atm)) * S
    a = g.copy()
    orc.stage(4, a, p)
    sl = (slice(1, -1),) * 3
    np.testing.assert_allclose(a["P1"][sl], P[sl], rtol=1e-14)
    np.testing.assert_allclose(a["S2"][sl], S[sl], rtol=0, atol=1e-14)
    b = g.copy()
    b["P1"] = P + rng.uniform(-20, 20, P.shape)
    b["S2"] = np.clip(S + rng.uniform(-0.01, 0.01, S.shape), 0.0, 0.6)
    orc.stage(4, b, p)
    A, B = g["roS1"], g["roS2"]
    r1, r2, b1, b2 = p.ro01, p.ro02, p.beta1, p.beta2
    # r1 (1+b1 x)(r2 (1+b2 x) - B) = A r2 (1+b2 x)  ->  c2 x^2 + c1 x + c0 = 0
    c2 = r1 * b1 * r2 * b2
    c1 = r1 * b1 * (r2 - B) + r1 * r2 * b2 - A * r2 * b2
    c0 = r1 * (r2 - B) - A * r2
    disc = np.sqrt(c1 * c1 - 4 * c2 * c0)
    x = (-2 * c0) / (c1 + disc)  # the root near x = P - P_atm, stable form
    np.testing.assert_allclose(b["P1"][sl], (p.P_atm + x)[sl], rtol=1e-12)
    Sroot = B / (r2 * (1 + b2 * x))
    np.testing.assert_allclose(b["S2"][sl], Sroot[sl], rtol=0, atol=1e-12)


# --------------------------------------------------------------- boundary --

def closed_form_boundary(g, p):
    """Independent restatement of the boundary sweep's result for NX, Ny,
    Nz >= 3 on one rank (see k_bc_a / k_bc_b in csrc/filt_stages.hip)."""
    S0, P0, ro1 = g["S2"].copy(), g["P1"].copy(), g["ro1"]
    N, M, L = p.NX, p.Ny, p.Nz
    S, P = S0.copy(), P0.copy()   # arrays indexed [k, j, i]

    def src(i, k):
        return p.S2_gr if (p.src_x0 <= i <= p.src_x1 and p.src_z0 <= k <= p.src_z1) else 0.0

    def c(k):
        return min(max(k, 1), L - 2)

    def post(i, j, k):   # value of shell cell (i, j, k), 1 <= i <= N-2
        if j == M - 1:
            return S0[c(k), M - 2, i], P0[c(k), M - 2, i] + ro1[0, 1, i] * p.g_const * p.h2
        if j == 0:
            s = S0[1, 0, i] if k == 0 else src(i, c(k))
            return s, p.P_atm
        return S0[c(k), j, i], P0[c(k), j, i]

    for k in range(L):
        for j in range(M):
            S[k, j, 0], P[k, j, 0] = S0[k, j, 1], P0[k, j, 1]
            for i in range(1, N):
                ii = min(i, N - 2)
                interior_src = 1 <= j <= M - 2 and 1 <= k <= L - 2
                if i == N - 1 and interior_src:
                    S[k, j, i], P[k, j, i] = S0[k, j, N - 2], P0[k, j, N - 2]
                elif not interior_src or i == N - 1:
                    S[k, j, i], P[k, j, i] = post(ii, j, k)
    return S, P


@pytest.mark.parametrize("dims", [(3, 3, 3), (5, 4, 6), (8, 7, 5), (4, 9, 3)])
def test_boundary_sweep_equals_closed_form(orc, dims):
    p = ref_params(*dims, source=1)
    rng = np.random.default_rng(sum(dims))
    g = random_state(p, rng)
    g["S2"] = rng.permutation(g["S2"].size).reshape(g["S2"].shape) * 1e-3
    g["P1"] = 1e5 + rng.permutation(g["P1"].size).reshape(g["P1"].shape)
    S, P = closed_form_boundary(g, p)
    orc.stage(5, g, p)
    assert np.array_equal(g["S2"], S)
    assert np.array_equal(g["P1"], P)


# ------------------------------------------------------------------- mass --

def test_mass_uniform_state(orc):
    p = ref_params(6, 5, 4)
    g = HostGrid(p, with_poro=True)
    g["P1"] = 1.2e5
    g["S2"] = 0.25
    g["poro"] = 0.3
    m1, m2 = orc.mass(g, p)
    n_int = 4 * 3 * 2
    V = p.h1 * p.h2 * p.h3
    assert m1 == pytest.approx(n_int * 0.3 * p.ro01 * (1 + p.beta1 * 2e4) * 0.75 * V, rel=1e-14)
    assert m2 == pytest.approx(n_int * 0.3 * p.ro02 * (1 + p.beta2 * 2e4) * 0.25 * V, rel=1e-14)
