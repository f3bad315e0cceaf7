"""The twist-averaging driver (source/Main.f90:99-255) against the reference's
own twist-averaging run (source/Test: N=24, r_s=5, cutoff 40, fcut 2, ten
twists, MP2 per twist, CCD at the special twist).  Every number of its fort.59
is reproduced: per-twist E_HF / E(2), averages, squared averages, standard
errors, the twist-averaged eigenvalues and connectivity, each twist's
connectivity distance, the special twist; then the final MP2 and the 84-step
CCD trajectory with the averaged eigenvalues."""
import re

import numpy as np
import pytest

from tests.helpers import info_for

pytestmark = pytest.mark.gpu
FLOAT = re.compile(r"^[-+]?\d*\.\d+(E[-+]\d+)?$")


def _nums(line):
    return [float(t) for t in line.split() if FLOAT.match(t)]


@pytest.fixture(scope="module")
def ta(fixtures):
    from uegccd_amd.twist_average import twist_average
    print("[ta] reference twist-averaging run", flush=True)
    info = info_for("n24_rs5")
    info.DoCalcTACCD = False
    info.TACalcN = 10
    return twist_average(info, twists=fixtures["n24"]["ta_twists"])


def test_per_twist_hf_mp2(ta, fixtures):
    ref = fixtures["n24"]["ta_fort59_lines"]
    for k, p in enumerate(ta.points):
        r = _nums(ref[k])
        assert tuple(r[:3]) == p.twist
        assert abs(p.hf - r[3]) < 1e-13 and abs(p.mp2 - r[4]) < 1e-13 and abs(p.ccd - r[5]) < 1e-13


def test_statistics(ta, fixtures):
    ref = fixtures["n24"]["ta_fort59_lines"]
    avg, sq, err = _nums(ref[10]), _nums(ref[11]), _nums(ref[12])
    np.testing.assert_allclose(list(ta.ntwist) + [ta.hf, ta.mp2, ta.ccd], avg, rtol=0, atol=1e-13)
    np.testing.assert_allclose(list(ta.ntwist2) + [ta.hf2, ta.mp22, ta.ccd2], sq, rtol=0, atol=1e-13)
    np.testing.assert_allclose(list(ta.ntwist_err) + [ta.hf_err, ta.mp2_err, ta.ccd_err], err,
                               rtol=1e-10, atol=0)
    np.testing.assert_allclose(ta.eigen_avg, _nums(ref[13]), rtol=0, atol=1e-14)
    conn = _nums(ref[14])
    np.testing.assert_array_equal(ta.conn_avg[:len(conn)], conn)
    assert not ta.conn_avg[len(conn):].any()


def test_special_twist(ta, fixtures):
    ref = fixtures["n24"]["ta_fort59_lines"]
    d2 = [_nums(line)[0] for line in ref[15:25]]
    np.testing.assert_allclose(ta.conn_diff2, d2, rtol=1e-12)
    assert tuple(_nums(ref[25])) == ta.special_twist == tuple(fixtures["n24"]["ta_special_twist"])
    assert ta.special_index == 5     # twists 6 and 7 tie; the first is kept
    assert abs(ta.min_diff2 - _nums(ref[26])[0]) < 1e-12 * ta.min_diff2
    for ours, theirs in zip(ta.fort59[25:], ref[25:]):
        assert ours.split()[:3] == theirs.split()[:3]


def test_final_ccd_with_averaged_eigenvalues(ta, fixtures):
    n24 = fixtures["n24"]
    assert abs(ta.final_e2 - n24["ta_final_e2"]) < 1e-12
    r = ta.final
    assert r.niter == n24["ta_final_niter"] == 84
    ref = np.array(n24["ta_final_trace"])
    np.testing.assert_allclose(r.trace_e, ref[:, 0], rtol=0, atol=1e-10)
    np.testing.assert_allclose(r.trace_dt, ref[:, 2], rtol=0, atol=1e-10)
    assert abs(r.ecorr - n24["ta_final_ecorr"]) < 1e-12


def test_served_ccd_per_twist(fixtures):
    """With DoCalcTACCD the per-twist CCDs run as one served batch; each
    equals the same twist's CCD solved alone."""
    from uegccd_amd.driver import CCDContext, init_heg, options
    from uegccd_amd.twist_average import twist_average
    info = info_for("n24_rs5")
    info.DoCalcTACCD = True
    twists = fixtures["n24"]["ta_twists"][:4]
    res = twist_average(info, twists=twists, ccd_opts=options(mode="diis", ndiis=8))
    for tw, p in zip(twists, res.points):
        ctx = CCDContext(init_heg(info, twist=tw, device=0), info, options(mode="diis", ndiis=8))
        _, _, tab, tba = ctx.mp2()
        alone = ctx.solve(tab, tba)
        assert abs(p.ccd - alone.ecorr) < 1e-12
        ctx.close()
