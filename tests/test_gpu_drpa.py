"""dRPA / SOSEX (DrvdRPA, source/dRPA.f90:23-143) through the C ABI
(uegccd_drpa) against the reference's dRPA benchmark output and the CPU
oracle.

Bars: the printed trajectory (E, NIter, dT) of benchmarks/ZRun-dRPA/Output
within its print precision, Ec(dRPA) and Ec(SOSEX) within the F18.12 print
precision, the oracle's trajectory within 1e-12 and its energies within
1e-12 Ha, equal NIter.
"""
import numpy as np
import pytest

from tests.helpers import info_for, oracle_for

pytestmark = pytest.mark.gpu

F18_12 = 5.1e-13
F15_10 = 5.1e-11


def _ctx(info, mode="reference"):
    from uegccd_amd.driver import CCDContext, init_heg, options
    return CCDContext(init_heg(info), info, options(mode=mode))


def test_drpa_reference_benchmark(fixtures):
    """N=14, r_s=1, cutoff 5 with DodRPASOSEX and mosaics off, as the
    reference's benchmarks/ZRun-dRPA Input."""
    print("[drpa] n14_cut5 benchmark", flush=True)
    c = fixtures["n14_cut5"]
    info = info_for("n14_cut5", **{"Do Mosaics": False, "DodRPASOSEX": True})
    r, esosex = _ctx(info).drpa()
    tr = np.array(c["drpa_trace"])
    assert r.status == 0 and not r.failed
    assert r.niter == c["drpa_niter"] == len(tr)
    np.testing.assert_allclose(r.trace_e, tr[:, 0], rtol=0, atol=F15_10)
    np.testing.assert_allclose(r.trace_dt, tr[:, 2], rtol=0, atol=F15_10)
    assert abs(r.ecorr - c["drpa_ecorr"]) < F18_12
    assert abs(esosex - c["drpa_esosex"]) < F18_12
    assert abs(r.maxres - c["drpa_maxres"]) < 1e-13


@pytest.mark.parametrize("name,mosaic", [("n14_cut4", True), ("n14_cut5_tw", True),
                                         ("n54_cut9", False)])
def test_drpa_vs_oracle(name, mosaic):
    """Mosaic terms on, a twisted basis, and 27 occupied orbitals (666
    transfers) against the oracle's literal DrvdRPA."""
    print(f"[drpa] {name} mosaic={mosaic}", flush=True)
    info = info_for(name, **{"Do Mosaics": mosaic, "DodRPASOSEX": True})
    o = oracle_for(info).drpa()
    ctx = _ctx(info)
    r, esosex = ctx.drpa()
    assert r.niter == o["niter"]
    np.testing.assert_allclose(r.trace_e, o["trace_e"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(r.trace_dt, o["trace_dt"], rtol=0, atol=1e-12)
    assert abs(r.ecorr - o["ecorr"]) < 1e-12
    assert abs(esosex - o["esosex"]) < 1e-12
    assert abs(r.maxres - o["maxres"]) < 1e-12 + 1e-6 * o["maxres"]
    np.testing.assert_allclose(r.tab, o["t2"], rtol=0, atol=1e-12)


def test_drpa_from_given_amplitudes_and_diis_context_refused():
    """A given start (the oracle's amplitudes after a few iterations) takes the
    same trajectory as the oracle from there; a DIIS-mode context is refused."""
    from uegccd_amd import lib as L
    print("[drpa] given start, DIIS refused", flush=True)
    info = info_for("n14_cut4", **{"DodRPASOSEX": True})
    orc = oracle_for(info)
    t3 = orc.drpa(maxiter=3)["t2"]
    o = orc.drpa(t2=t3)
    r, esosex = _ctx(info).drpa(t2=t3)
    assert r.niter == o["niter"]
    assert abs(r.ecorr - o["ecorr"]) < 1e-12 and abs(esosex - o["esosex"]) < 1e-12
    with pytest.raises(L.UegccdError) as e:
        _ctx(info, mode="diis").drpa()
    assert e.value.status == L.ERR_UNSUPPORTED
