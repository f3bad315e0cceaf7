"""Bases past a thousand orbitals through the C ABI.

- dRPA/SOSEX (uegccd_drpa) on the reference's basis scan
  (benchmarks/ZRun-dRPA/old-dRPA: N=14, r_s=1, Gamma, mosaics off, cutoffs
  5..665, up to 71,999 orbitals): both energies within the twelve printed
  decimals of every scanned row tested (half a unit of the last, 5e-13,
  plus 1e-13 for a different order of summation over up to 10^6 terms),
  and the basis size.
- CCD on N=14 |n|^2 <= 50 (1,503 orbitals): the residual terms at the MP2
  amplitudes against the oracle (1e-12 relative to the largest G entry), and
  the DIIS solve's amplitudes a fixed point of the oracle's equations (its
  max|R| there below the solve's tolerance, its energy there within 1e-12 Ha
  of the solve's).
- CCD on N=2 |n|^2 <= 120 (5,497 orbitals): its one pair-momentum block's
  virtual list (5,496 columns) goes through LDS in two parts; the residual
  terms at the MP2 amplitudes and the converged DIIS solve against the
  oracle as above.
- A basis with a virtual-occupied |dn|^2 past the slice blocks' LDS
  Coulomb table (N=2 |n|^2 <= 1050, 142,945 orbitals) is refused for CCD
  with UEGCCD_ERR_UNSUPPORTED, while MP2 and the energy run on it.
"""
import numpy as np
import pytest

from tests.helpers import oracle_for
from uegccd_amd.inputfile import parse_input, standard_input

pytestmark = pytest.mark.gpu

SCAN_TOL = 6e-13
SCAN_CUTOFFS = (5, 25, 50, 100, 150, 200, 300, 400, 500, 600, 665)


def _scan_info(cut, fixtures, **over):
    p = fixtures["drpa_scan"]["params"]
    kw = {"Do Mosaics": p["do_mosaic"], "DodRPASOSEX": True, "fcut": p["fcut"]}
    kw.update(over)
    return parse_input(standard_input(p["nel"], cut, p["rs"], tuple(p["twist"]), **kw))


@pytest.mark.parametrize("cut", SCAN_CUTOFFS)
def test_drpa_basis_scan(fixtures, cut):
    from uegccd_amd.driver import CCDContext, init_heg, options
    print(f"[scan] dRPA cutoff {cut}", flush=True)
    row = {r[0]: r for r in fixtures["drpa_scan"]["rows"]}[cut]
    info = _scan_info(cut, fixtures)
    heg = init_heg(info)
    assert 2 * heg.NAO == row[3]
    ctx = CCDContext(heg, info, options(mode="reference"))
    r, esosex = ctx.drpa()
    assert r.status == 0 and not r.failed
    assert abs(r.ecorr - row[1]) < SCAN_TOL
    assert abs(esosex - row[2]) < SCAN_TOL


def _ccd_against_oracle(info):
    from uegccd_amd.driver import CCDContext, init_heg, options
    heg = init_heg(info)
    orc = oracle_for(info)
    e2o, _, tab0, tba0 = orc.mp2()
    ctx = CCDContext(heg, info, options(mode="diis", ndiis=8))
    g_ab, g_ba, _, _ = ctx.residual(tab0, tba0)
    o_gab, o_gba, _, _ = orc.terms(tab0, tba0, mask=31)
    scale = max(np.abs(o_gab).max(), np.abs(o_gba).max())
    assert np.abs(g_ab - o_gab).max() / scale < 1e-12
    assert np.abs(g_ba - o_gba).max() / scale < 1e-12
    e2, _, tab, tba = ctx.mp2()
    assert abs(e2 - e2o) < 1e-12
    r = ctx.solve(tab, tba)
    assert r.status == 0 and not r.failed
    _, _, o_rab, o_rba = orc.terms(r.tab, r.tba, mask=31)
    assert max(np.abs(o_rab).max(), np.abs(o_rba).max()) < 1e-7
    assert abs(orc.energy(r.tab + r.tba, r.tab, r.tba) - r.ecorr) < 1e-12
    return heg


def test_ccd_past_a_thousand_orbitals(fixtures):
    print("[large] CCD N=14 cutoff 50", flush=True)
    heg = _ccd_against_oracle(_scan_info(50, fixtures, **{"Do Mosaics": True, "DodRPASOSEX": False}))
    assert heg.NAO == 1503


def test_ccd_virtual_list_in_parts():
    print("[large] CCD N=2 cutoff 120", flush=True)
    heg = _ccd_against_oracle(parse_input(standard_input(2, 120, 1.0, fcut=1000.0)))
    assert (heg.NAO, heg.NOcc) == (5497, 1)


def test_ccd_refused_past_the_slice_table():
    from uegccd_amd import lib as L
    from uegccd_amd.driver import CCDContext, init_heg, options
    print("[large] CCD refusal N=2 cutoff 1050", flush=True)
    info = parse_input(standard_input(2, 1050, 1.0, fcut=10000.0))
    heg = init_heg(info)
    assert heg.NAO > 140000
    ctx = CCDContext(heg, info, options(mode="reference"))
    e2, _, tab, tba = ctx.mp2()
    assert e2 < 0.0
    assert abs(ctx.energy(tab, tba) - e2) < 1e-12
    with pytest.raises(L.UegccdError) as e:
        ctx.solve(tab, tba)
    assert e.value.status == L.ERR_UNSUPPORTED
    with pytest.raises(L.UegccdError) as e:
        ctx.residual(tab, tba)
    assert e.value.status == L.ERR_UNSUPPORTED
