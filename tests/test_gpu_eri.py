"""The non-default Coulomb-integral semantics on the device against the oracle.

GapII screening, the unchecked BaseERI (Safe ERI F), the per-term range
flags of CasedERI (source/CCD.f90: each ERI call's DummyFlag) and CoreN
masking (run on the device as the frozen core; the oracle masks every
integral as the reference does).  Bars as the default-integral parity tests:
residual terms within 1e-12 relative, energies within 1e-12 Ha, the
reference-mode trajectory with identical NIter and energy within 1e-11 Ha.
"""
import numpy as np
import pytest

from tests.helpers import info_for, oracle_for

pytestmark = pytest.mark.gpu

TERM_TOL = 1e-12


def _ctx(info):
    from uegccd_amd.driver import CCDContext, init_heg
    heg = init_heg(info)
    return heg, CCDContext(heg, info)


def _random(orc, seed, core=0):
    e2, taa, tab, tba = orc.mp2()
    rng = np.random.default_rng(seed)
    o = orc.nocc
    mask = np.zeros(orc.shape, dtype=bool)
    for i in range(1, o + 1):
        for j in range(1, o + 1):
            for a in range(o + 1, orc.nao + 1):
                if orc.findindex(i, j, a) > o and min(i, j) > core:
                    mask[i - 1, j - 1, a - o - 1] = True
    a = np.where(mask, rng.normal(0, 1e-2, mask.shape), 0.0)
    b = np.where(mask, rng.normal(0, 1e-2, mask.shape), 0.0)
    return np.asfortranarray(a), np.asfortranarray(b)


def _check_terms(ctx, orc, tab, tba, what):
    g_ab, g_ba, r_ab, r_ba = ctx.residual(tab, tba)
    o_gab, o_gba, o_rab, o_rba = orc.terms(tab, tba)
    scale = max(np.abs(o_gab).max(), np.abs(o_gba).max())
    assert np.abs(g_ab - o_gab).max() / scale < TERM_TOL, (what, "G abab")
    assert np.abs(g_ba - o_gba).max() / scale < TERM_TOL, (what, "G abba")
    rs = max(np.abs(o_rab).max(), np.abs(o_rba).max(), scale)
    assert np.abs(r_ab - o_rab).max() / rs < 1e-11, (what, "R abab")
    assert np.abs(r_ba - o_rba).max() / rs < 1e-11, (what, "R abba")


def _check_all(info, solve=True, core=0):
    orc = oracle_for(info)
    heg, ctx = _ctx(info)
    e2, taa, tab, tba = ctx.mp2()
    o_e2, o_taa, o_tab, o_tba = orc.mp2()
    assert abs(e2 - o_e2) < 1e-12
    assert np.abs(tab - o_tab).max() < 1e-13 and np.abs(tba - o_tba).max() < 1e-13
    _check_terms(ctx, orc, o_tab, o_tba, "mp2")
    ra, rb = _random(orc, 7, core)
    _check_terms(ctx, orc, ra, rb, "random")
    assert abs(ctx.energy(ra, rb) - orc.energy(ra + rb, ra, rb)) < 1e-12
    if solve:
        r = ctx.solve(o_tab, o_tba)
        o = orc.ccd()
        assert r.niter == o["niter"], (r.niter, o["niter"])
        assert abs(r.ecorr - o["ecorr"]) < 1e-11, (r.ecorr, o["ecorr"])
    return ctx, orc


@pytest.mark.parametrize("safe", [True, False])
def test_gapii_screening(safe):
    """GapII below v(0) = -Madelung and the smallest transfers: those integrals
    vanish (GapERI with Safe ERI F, CasedERI with T), the hole-hole ladder's
    v(0) among them."""
    print(f"[gapii] safe={safe}", flush=True)
    info = info_for("n14_cut5", **{"GapII": 0.06, "Safe ERI": safe})
    _check_all(info)


def test_base_eri_without_checks():
    print("[base eri]", flush=True)
    _check_all(info_for("n14_cut5_tw", **{"Safe ERI": False}))


RANGES = [
    {"DRings Range": 1, "ExRings Range": 2},
    {"Mosaics Range": 1, "QuadRings Range": 2, "LinRings Range": 3},
    {"Ladders Range": 2},
    {"DLadders Range": 1, "QuadLadd Range": 2},
    {"Energy Range": 1, "DriverDir Range": 2, "DriverEx Range": 1},
    {"XRings Range": 1, "Rings Range": 2, "ExLadders Range": 3},
]


@pytest.mark.parametrize("k", range(len(RANGES)))
def test_range_flags(k):
    """Range flags of CasedERI per term: split and shared tables in the slice,
    ladder and residual kernels."""
    print(f"[range] {RANGES[k]}", flush=True)
    _check_all(info_for("n14_cut5_tw", **RANGES[k]), solve=(k in (0, 3, 4)))


def test_random_range_flags():
    from uegccd_amd.inputfile import RANGE_NAMES
    rng = np.random.default_rng(2024)
    for rep in range(3):
        over = {name: int(rng.integers(0, 3)) for name in RANGE_NAMES}
        print(f"[range random] {over}", flush=True)
        _check_all(info_for("n24_rs5", **over), solve=False)


def test_negative_v0_refused_only_where_met():
    """At small r_s the long-range kernel's v(0) is negative (the checked
    variants' -100 sentinel).  The ladders meet v(0) (c = a, i = k): refused.
    The drivers never do (v(n_i - n_a)): run, and equal to the oracle."""
    from uegccd_amd import lib as L
    info = info_for("n14_cut4", rSMin=0.01, rSMax=0.01, **{"Ladders Range": 2})
    with pytest.raises(L.UegccdError) as e:
        _ctx(info)
    assert e.value.status == L.ERR_UNSUPPORTED
    info = info_for("n14_cut4", rSMin=0.01, rSMax=0.01, **{"DriverDir Range": 2})
    orc = oracle_for(info)
    assert orc.eri(1, 2, 1, 2, 2) < 0.0  # the sentinel's cause is there
    _check_all(info, solve=False)


@pytest.mark.parametrize("kind", ["core", "cased", "nocase"])
def test_coren_frozen_core(kind):
    """CoreN = 2 (CoreFactor 0.3 of 7 pairs): CoreERI (Safe ERI F), CasedERI,
    NoCaseERI (with GapII): MP2, residual, energy and the whole reference
    trajectory against the oracle's masked integrals."""
    over = {"CoreFactor": 0.3, "Safe ERI": kind == "cased"}
    if kind == "nocase":
        over["GapII"] = 0.06
    print(f"[coren] {kind}", flush=True)
    info = info_for("n14_cut5", **over)
    assert info.CoreN == 2
    ctx, orc = _check_all(info, core=2)
    # connectivity leaves the core excitations out (source/MP2.f90:75-85)
    full = _ctx(info_for("n14_cut5"))[1].connectivity()
    assert ctx.connectivity().sum() < full.sum()


def test_coren_structure_factor_and_refusal():
    from uegccd_amd import lib as L
    info = info_for("n14_cut5_tw", CoreFactor=0.3)
    orc = oracle_for(info)
    heg, ctx = _ctx(info)
    _, _, tab, tba = orc.mp2()
    q, vals, e = ctx.structure_factor(tab, tba)
    oq, ovals, oe = orc.structure_factor(tab, tba)
    np.testing.assert_array_equal(q, oq)
    assert np.abs(vals - ovals).max() <= 1e-13 * np.abs(ovals).max()
    assert abs(e - oe) < 1e-12
    bad = tab.copy()
    o = orc.nocc
    a = next(a for a in range(o + 1, orc.nao + 1) if orc.findindex(1, 4, a) > o)
    bad[0, 3, a - o - 1] = 1e-3  # a valid amplitude with the core index i = 1
    with pytest.raises(L.UegccdError) as err:
        ctx.residual(bad, tba)
    assert err.value.status == L.ERR_INVALID


def test_coren_and_gap_drpa():
    print("[drpa coren gap]", flush=True)
    info = info_for("n14_cut5", CoreFactor=0.3, GapII=0.06)
    orc = oracle_for(info)
    heg, ctx = _ctx(info)
    r, es = ctx.drpa()
    o = orc.drpa()
    assert r.niter == o["niter"]
    assert abs(r.ecorr - o["ecorr"]) < 1e-12 and abs(es - o["esosex"]) < 1e-12
