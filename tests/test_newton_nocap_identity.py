"""The capillary-free Newton iteration of the device code drops the zero
capillary terms: with P_k = PkS = 0 the reference's iteration (cpu.cpp:300-317)
forms (P + P_k) - P_atm and ((P + P_k) - P_atm) + S*PkS, and the device code
uses P - P_atm for both, so F2S is the product F2 starts from.  That holds bit
for bit when P_atm != 0 and S > -inf.  This checks the identity in IEEE
double arithmetic (numpy, round to nearest, no fused operations, as the
device code is compiled) over signed zeros, infinities, NaN, subnormals and
random states, and that it fails where its conditions do not hold."""
import numpy as np

RO01, RO02, B1, B2 = 1000.0, 1460.0, 1.0e-4, 1.0e-8


def step_reference(P, S, roS1, roS2, Pa, P_k=0.0, PkS=0.0):
    with np.errstate(all="ignore"):
        F1 = RO01 * (1. + B1 * (P - Pa)) * (1. - S) - roS1
        F2 = RO02 * (1. + B2 * (P + P_k - Pa)) * S - roS2
        F1P = RO01 * B1 * (1 - S)
        F2P = RO02 * B2 * S
        F1S = (-1) * RO01 * (1 + B1 * (P - Pa))
        F2S = RO02 * (1 + B2 * (P + P_k - Pa + (S * PkS)))
        det = F1P * F2S - F1S * F2P
        return P - (1 / det) * (F2S * F1 - F1S * F2), S - (1 / det) * (F1P * F2 - F2P * F1)


def step_nocap(P, S, roS1, roS2, Pa):
    with np.errstate(all="ignore"):
        F1 = RO01 * (1. + B1 * (P - Pa)) * (1. - S) - roS1
        F2S = RO02 * (1 + B2 * (P - Pa))
        F2 = F2S * S - roS2
        F1P = RO01 * B1 * (1 - S)
        F2P = RO02 * B2 * S
        F1S = (-1) * RO01 * (1 + B1 * (P - Pa))
        det = F1P * F2S - F1S * F2P
        return P - (1 / det) * (F2S * F1 - F1S * F2), S - (1 / det) * (F1P * F2 - F2P * F1)


def same(a, b):
    """Bit-equal, or both NaN (NaN payloads are the hardware's business)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))


def grid_inputs():
    specials_P = np.array([0.0, -0.0, 1e5, -1e5, 1e5 + 1e-11, 5e-324, -5e-324, 1e308, -1e308,
                           np.inf, -np.inf, np.nan, 101325.0, 1.0, -1.0])
    specials_S = np.array([0.0, -0.0, 0.3, 1.0, 0.9, -0.1, 5e-324, -5e-324, 1e300, -1e300,
                           np.inf, np.nan])
    rng = np.random.default_rng(11)
    P = np.concatenate([specials_P, rng.uniform(9e4, 1.4e5, 2000)])
    S = np.concatenate([specials_S, rng.uniform(0.0, 0.9, 2000)])
    PP, SS = np.meshgrid(P, S, indexing="ij")
    roS1 = RO01 * (1 - SS) * 1.0003
    roS2 = RO02 * SS * 0.9997
    return PP.ravel(), SS.ravel(), roS1.ravel(), roS2.ravel()


def test_identity_holds_where_its_conditions_do():
    P, S, r1, r2 = grid_inputs()
    for Pa in (1e5, -3.0, 5e-324, 101325.0, P[17]):
        a = step_reference(P, S, r1, r2, Pa)
        b = step_nocap(P, S, r1, r2, Pa)
        assert same(a[0], b[0]).all() and same(a[1], b[1]).all(), Pa


def test_p_equal_to_p_atm_keeps_the_zero():
    """P - P_atm = +0 exactly: + S*0 leaves it +0 for S of either sign."""
    Pa = 123456.0
    P = np.full(4, Pa)
    S = np.array([0.25, -0.25, 0.0, -0.0])
    a = step_reference(P, S, RO01 * (1 - S), RO02 * S, Pa)
    b = step_nocap(P, S, RO01 * (1 - S), RO02 * S, Pa)
    assert same(a[0], b[0]).all() and same(a[1], b[1]).all()


def test_identity_fails_outside_its_conditions():
    """Planted counterexamples: P_atm = 0 with P = -0 (a signed-zero
    difference in P - P_atm), and S = -inf (S*0 is NaN, so the reference's
    F2S is NaN where the folded one is finite; the device code leaves such
    states to the general iteration)."""
    d_ref = (np.array([-0.0]) + 0.0) - 0.0
    d_new = np.array([-0.0]) - 0.0
    assert d_ref.view(np.uint64)[0] != d_new.view(np.uint64)[0]
    with np.errstate(all="ignore"):
        P, S, Pa = 1e5, -np.inf, 1e5
        f2s_ref = RO02 * (1 + B2 * (P + 0.0 - Pa + (S * 0.0)))
        f2s_new = RO02 * (1 + B2 * (P - Pa))
    assert np.isnan(f2s_ref) and np.isfinite(f2s_new)
    # a dropped term in F2S would fail the identity on ordinary states
    P, S, r1, r2 = grid_inputs()
    with np.errstate(all="ignore"):
        bad = step_reference(P, S, r1, r2, 1e5, P_k=1e-3)
    good = step_nocap(P, S, r1, r2, 1e5)
    assert not same(bad[0], good[0]).all()
    assert S is not None
