"""The CPU oracle against every golden vector the reference's stored outputs
hold for this path (tests/golden/reference_fixtures.json), plus closed forms
and invariants that pin what the fixtures leave open."""
import numpy as np
import pytest

from oracle import OracleParams, OracleSystem

# Printed precision of the reference's formats: F18.12 (energies), F15.10 /
# F14.10 (per-iteration lines), ES24.16 (eigenvalues), list-directed (17 digits).
F18_12 = 5.1e-13
F15_10 = 5.1e-11


def _sys(params, **kw):
    p = dict(params)
    p.update(kw)
    return OracleSystem(OracleParams(**p))


@pytest.fixture(scope="module")
def n24(fixtures):
    return fixtures["n24"]


def test_n24_basis_eigenvalues_and_hf(n24):
    s = _sys(n24["params"])
    assert (s.nocc, s.nao) == (12, 34)
    np.testing.assert_allclose(s.eigen, n24["eigen"], rtol=0, atol=2e-16 * 10)
    assert abs(s.ehf - n24["ehf"]) < F18_12
    assert abs(s.madelung - n24["madelung_rs5"]) <= 1e-17
    assert abs(s.L - n24["L_rs5"]) <= 1e-14
    tab = np.array(n24["basis_rs5"])
    np.testing.assert_array_equal(s.nvec, tab[:, :3].astype(int))
    np.testing.assert_allclose(s.eigen, tab[:, 3], rtol=0, atol=1e-16 * 5)


def test_n24_basis_at_rs1(n24):
    p = dict(n24["params"], rs=1.0)
    s = _sys(p)
    tab = np.array(n24["basis_rs1"])
    np.testing.assert_array_equal(s.nvec, tab[:, :3].astype(int))
    np.testing.assert_allclose(s.eigen, tab[:, 3], rtol=0, atol=1e-15)
    assert abs(s.madelung - n24["madelung_rs1"]) <= 1e-16


def test_n24_mp2_and_ccd_trajectory(n24):
    s = _sys(n24["params"])
    e2, taa, tab, tba = s.mp2()
    assert abs(e2 - n24["e2"]) < F18_12
    r = s.ccd()
    assert r["niter"] == n24["ccd_niter"] == 70
    ref = np.array(n24["ccd_trace"])
    np.testing.assert_array_equal(ref[:, 1], np.arange(1, 71))
    np.testing.assert_allclose(r["trace_e"], ref[:, 0], rtol=0, atol=F15_10)
    np.testing.assert_allclose(r["trace_dt"], ref[:, 2], rtol=0, atol=F15_10)
    assert abs(r["ecorr"] - n24["ccd_ecorr"]) < F18_12
    assert abs(r["maxres"] - n24["ccd_maxres"]) < 5e-21
    assert not r["failed"]


def test_twist_averaging_mp2_per_twist(n24):
    """The 10 twists of the twist-averaging run: E(SCF), E(2) (fort.59, 17
    digits) and the eigenvalue listings of each MP2 block."""
    for tw, (hf, mp2), blk in zip(n24["ta_twists"], n24["ta_fort59_hf_mp2"], n24["ta_mp2"]):
        s = _sys(n24["params"], twist=tuple(tw))
        e2, *_ = s.mp2()
        assert abs(s.ehf - hf) < 1e-13 * max(1, abs(hf))
        assert abs(e2 - mp2) < 1e-14
        np.testing.assert_allclose(s.eigen, blk["eig"], rtol=0, atol=1e-15)


def test_twist_averaging_final_ccd(n24):
    """The CCD at the special twist with twist-averaged eigenvalues
    (source/Main.f90:233-251): 84 iterations, Ec = -0.187005573201."""
    s = _sys(n24["params"], twist=tuple(n24["ta_special_twist"]))
    assert abs(s.ehf - n24["ta_final_ehf"]) < F18_12
    s.set_eigen(np.array(n24["ta_avg_eigen"]))
    e2, *_ = s.mp2()
    assert abs(e2 - n24["ta_final_e2"]) < F18_12
    r = s.ccd()
    assert r["niter"] == n24["ta_final_niter"] == 84
    ref = np.array(n24["ta_final_trace"])
    np.testing.assert_allclose(r["trace_e"], ref[:, 0], rtol=0, atol=F15_10)
    np.testing.assert_allclose(r["trace_dt"], ref[:, 2], rtol=0, atol=F15_10)
    assert abs(r["ecorr"] - n24["ta_final_ecorr"]) < F18_12
    assert abs(r["maxres"] - n24["ta_final_maxres"]) < 5e-21


def test_c0_system_hf_mp2(fixtures):
    """N=14, r_s=1, Gamma, cutoff 5 (57 orbitals): E(SCF), eigenvalues, E(2)."""
    c = fixtures["n14_cut5"]
    s = _sys(c["params"])
    assert s.nao == 57
    assert abs(s.ehf - c["ehf"]) < F18_12
    assert abs(s.madelung - c["madelung"]) < 1e-16
    np.testing.assert_allclose(s.eigen, c["eigen"], rtol=0, atol=5e-16 * 4)
    bas = np.array(c["basis"])
    np.testing.assert_array_equal(s.nvec[:len(bas)], bas[:, :3].astype(int))
    e2, *_ = s.mp2()
    assert abs(e2 - c["e2"]) < F18_12


def test_v6_first_twist_ccd(fixtures):
    """N=14 cutoff 2 at the first v6 twist: E_HF, E(2), Ec(CCD) to 17 digits."""
    v = fixtures["v6_twist1"]
    s = _sys(v["params"])
    assert s.nao == 19
    assert abs(s.ehf - v["ehf"]) < 1e-14
    np.testing.assert_allclose(s.eigen, v["eigen"], rtol=0, atol=1e-15)
    e2, *_ = s.mp2()
    assert abs(e2 - v["e2"]) < 1e-16
    r = s.ccd()
    assert r["niter"] == v["ccd_niter"] == 33
    assert abs(r["ecorr"] - v["ccd_ecorr"]) < 1e-16
    ref = np.array(v["ccd_trace"])
    np.testing.assert_allclose(r["trace_e"], ref[:, 0], rtol=0, atol=F15_10)


# ---------------------------------------------------------------------------
# What the fixtures leave open: closed forms, invariants, term sensitivity.

@pytest.mark.parametrize("nel,rs", [(14, 1.0), (54, 2.0), (24, 5.0)])
def test_madelung_closed_form(nel, rs):
    """Simple-cubic Madelung energy: v_M = -2.837297479 / L."""
    s = _sys(dict(nel=nel, maxkpoint=1, rs=rs))
    assert abs(s.madelung - (-2.837297479 / s.L)) < 1e-9 / s.L


def test_eri_closed_form():
    s = _sys(dict(nel=14, maxkpoint=5, rs=1.3))
    rng = np.random.default_rng(7)
    n = s.nvec
    for _ in range(200):
        i, j, a = (int(x) for x in rng.integers(1, s.nao + 1, size=3))
        b = s.findindex(i, j, a)
        if b == 0:
            continue
        dq = n[a - 1] - n[i - 1]
        q2 = int(dq @ dq)
        want = -s.madelung if q2 == 0 else 1.0 / (np.pi * s.L * q2)
        assert s.eri(a, b, i, j, 0) == pytest.approx(want, rel=1e-15)
        assert np.all(n[b - 1] == n[i - 1] + n[j - 1] - n[a - 1])


def test_mp2_closed_form_numpy():
    """E(2) from a vectorised numpy restatement of the MP2 formula."""
    s = _sys(dict(nel=14, maxkpoint=9, rs=2.0, twist=(0.1, -0.2, 0.3)))
    o, n = s.nocc, s.nao
    nv, eps = s.nvec, s.eigen
    key = {tuple(x): p for p, x in enumerate(nv)}

    def v(d):
        q2 = int(d @ d)
        return -s.madelung if q2 == 0 else 1.0 / (np.pi * s.L * q2)

    e2 = 0.0
    for i in range(o):
        for j in range(o):
            for a in range(o, n):
                b = key.get(tuple(nv[i] + nv[j] - nv[a]), -1)
                if b < o:
                    continue
                vab, vba = v(nv[a] - nv[i]), v(nv[b] - nv[i])
                d = eps[i] + eps[j] - eps[a] - eps[b]
                tab, tba = vab / d, -vba / d
                e2 += 0.5 * (tab + tba) * (vab - vba) + 0.5 * tab * vab - 0.5 * tba * vba
    got, *_ = s.mp2()
    assert got == pytest.approx(e2, rel=1e-13)


def test_amplitude_symmetries_preserved():
    """T^ab(i,j,a) = T^ab(j,i,b) and T^ba(i,j,a) = -T^ab(i,j,b) at MP2 and after
    the iteration; invalid entries stay exactly zero."""
    s = _sys(dict(nel=14, maxkpoint=5, rs=1.0, twist=(0.05, 0.1, -0.2)))
    o = s.nocc
    r = s.ccd(maxiter=15)
    for tab, tba in ((s.mp2()[2], s.mp2()[3]), (r["tab"], r["tba"])):
        for i in range(1, o + 1):
            for j in range(1, o + 1):
                for a in range(o + 1, s.nao + 1):
                    b = s.findindex(i, j, a)
                    x = tab[i - 1, j - 1, a - o - 1]
                    if b <= o:
                        assert x == 0.0 and tba[i - 1, j - 1, a - o - 1] == 0.0
                        continue
                    assert abs(x - tab[j - 1, i - 1, b - o - 1]) < 1e-14
                    assert abs(tba[i - 1, j - 1, a - o - 1] + tab[i - 1, j - 1, b - o - 1]) < 1e-14


@pytest.mark.parametrize("drop", [2, 4, 8, 16])
def test_each_term_moves_the_fixture(n24, drop):
    """Dropping any one term from the residual changes the first CCD step's
    energy far beyond the fixture's printed precision, so the trajectory
    fixture pins every term (mosaics, ladders, rings, crossed rings)."""
    s = _sys(n24["params"])
    _, taa, tab, tba = s.mp2()
    for mask, sign in ((31, 0), (31 & ~drop, 1)):
        gab, gba, _, _ = s.terms(tab, tba, mask=mask)
        t1ab, t1ba = tab.copy(order="F"), tba.copy(order="F")
        from oracle import _lib, _dp
        _lib().orc_solvecc(s._h, _dp(t1ab), _dp(gab))
        _lib().orc_solvecc(s._h, _dp(t1ba), _dp(gba))
        e = s.energy(t1ab + t1ba, t1ab, t1ba)
        ref = n24["ccd_trace"][1][0]
        if sign == 0:
            assert abs(e - ref) < F15_10
        else:
            assert abs(e - ref) > 1e-6


def test_c0_drpa_sosex(fixtures):
    """DrvdRPA on the reference's dRPA benchmark Input (N=14, r_s=1, cutoff 5,
    DodRPASOSEX, mosaics off): the printed trajectory (E, NIter, dT),
    Ec(dRPA), Ec(SOSEX) and the max residual of benchmarks/ZRun-dRPA/Output."""
    c = fixtures["n14_cut5"]
    s = _sys(c["params"], do_mosaic=False)
    r = s.drpa()
    tr = np.array(c["drpa_trace"])
    assert r["niter"] == c["drpa_niter"] == len(tr)
    np.testing.assert_array_equal(tr[:, 1], np.arange(1, len(tr) + 1))
    np.testing.assert_allclose(r["trace_e"], tr[:, 0], rtol=0, atol=F15_10)
    np.testing.assert_allclose(r["trace_dt"], tr[:, 2], rtol=0, atol=F15_10)
    assert abs(r["ecorr"] - c["drpa_ecorr"]) < F18_12
    assert abs(r["esosex"] - c["drpa_esosex"]) < F18_12
    assert abs(r["maxres"] - c["drpa_maxres"]) < 5e-21
    assert not r["failed"]


def _drpa_reduced(s, t2):
    """The dRPA G in the form the device computes it: every ring term carries
    v(n_i - n_a), so with R(i,a) = sum_k T(i,k,a) and Q(j, q) = sum over
    (l, d) with n_d - n_l = q of T(l,j,d),
      G = v_ia (1 + 2 R(i,a) + 2 R(j,b) + 4 R(i,a) Q(j, n_a - n_i))
          - 1/2 (Jvv(a) + Jvv(b) + Joo(i) + Joo(j)) T,
    Jvv(c) = sum_k 2 v(n_c - n_k) R(k,c), Joo(k) = sum_c 2 v(n_c - n_k) R(k,c)."""
    o, nv = s.nocc, s.nvir
    nvec = s.nvec.astype(int)
    idx = {tuple(nvec[p]): p for p in range(s.nao)}
    def v(p, q):  # v(n_p - n_q), 0-based orbitals
        return s.eri(p + 1, q + 1, q + 1, p + 1)
    R = t2.sum(axis=1)  # R[i, a-o]
    qsum = {}
    for l in range(o):
        for d in range(o, s.nao):
            q = tuple(nvec[d] - nvec[l])
            qsum.setdefault(q, np.zeros(o))
            qsum[q] += t2[l, :, d - o]
    jvv = np.array([sum(2 * v(o + c, k) * R[k, c] for k in range(o)) for c in range(nv)])
    joo = np.array([sum(2 * v(o + c, k) * R[k, c] for c in range(nv)) for k in range(o)])
    g = np.zeros_like(t2)
    for i in range(o):
        for j in range(o):
            for av in range(nv):
                a = o + av
                b = idx.get(tuple(nvec[i] + nvec[j] - nvec[a]), -1)
                if b < o:
                    continue
                vq = v(i, a)
                qj = qsum.get(tuple(nvec[a] - nvec[i]), np.zeros(o))[j]
                ring = vq * (2 * R[i, av] + 2 * R[j, b - o] + 4 * R[i, av] * qj)
                mos = 0.5 * (jvv[av] + jvv[b - o] + joo[i] + joo[j]) * t2[i, j, av]
                g[i, j, av] = vq + (ring if s.params.do_ring else 0.0) - (mos if s.params.do_mosaic else 0.0)
    return g


@pytest.mark.parametrize("mosaic", [False, True])
def test_drpa_reduced_form(fixtures, mosaic):
    """The row-sum / transfer-sum form of the dRPA terms (what the device
    computes) against the oracle's loop-literal GetG2Drivers/Mosaics/Rings on
    random amplitudes of the valid elements."""
    c = fixtures["n14_cut5"]
    s = _sys(c["params"], do_mosaic=mosaic)
    _, taa, _, _ = s.mp2()
    rng = np.random.default_rng(7)
    t2 = np.where(taa != 0.0, rng.standard_normal(taa.shape) * 0.05, 0.0)
    g, _ = s.drpa_terms(np.asfortranarray(t2))
    gr = _drpa_reduced(s, t2)
    assert np.abs(g).max() > 0.1
    np.testing.assert_allclose(gr, g, rtol=0, atol=1e-14 * np.abs(g).max())


@pytest.mark.parametrize("amps", ["mp2", "random"])
def test_structure_factor_identities(fixtures, amps):
    """CCEnergyAndSF (source/CCD.f90:240-351) restated in the oracle, pinned by
    what the binning fixes: the cells are exactly the transfers n_a - n_i and
    n_a - n_j of the valid elements (brute force over the n-vectors), in the
    reference's print order; SfG = 2 pi |q| / L and SfV = v(q) = 1/(pi L |q|^2)
    in closed form; sum_q SfV SfSum = 2 E (the energy pinned to the reference's
    Output); sum_q SfV SfDSum and sum_q SfV SfXSum equal the element sums
    2 (v_ia T^ab - v_ja T^ba) and v_ia T^ba - v_ja T^ab; SfDSum + SfXSum =
    SfSum cell by cell."""
    c = fixtures["n14_cut5"]
    s = _sys(c["params"])
    _, taa, tab, tba = s.mp2()
    if amps == "random":
        rng = np.random.default_rng(11)
        mask = taa != 0.0
        tab = np.where(mask, rng.standard_normal(taa.shape) * 0.05, 0.0)
        tba = np.where(mask, rng.standard_normal(taa.shape) * 0.05, 0.0)
        taa = tab + tba
    q, vals, e = s.structure_factor(tab, tba)
    o = s.nocc
    nvec = s.nvec.astype(int)
    idx = {tuple(nvec[p]): p for p in range(s.nao)}
    cells = set()
    sd = sx = 0.0
    for i in range(o):
        for j in range(o):
            for a in range(o, s.nao):
                if idx.get(tuple(nvec[i] + nvec[j] - nvec[a]), -1) < o:
                    continue
                qi, qj = tuple(nvec[a] - nvec[i]), tuple(nvec[a] - nvec[j])
                cells.add(qi)
                cells.add(qj)
                vi, vj = 1 / (np.pi * s.L * np.dot(qi, qi)), 1 / (np.pi * s.L * np.dot(qj, qj))
                x = (i, j, a - o)
                sd += 2 * (vi * tab[x] - vj * tba[x])
                sx += vi * tba[x] - vj * tab[x]
    assert [tuple(r) for r in q] == sorted(cells)
    q2 = (q.astype(float) ** 2).sum(axis=1)
    np.testing.assert_allclose(vals[:, 0], 2 * np.pi / s.L * np.sqrt(q2), rtol=1e-15)
    np.testing.assert_allclose(vals[:, 2], 1 / (np.pi * s.L * q2), rtol=1e-14)
    e_ref = s.energy(taa, tab, tba)
    assert abs(e - e_ref) < 1e-15
    assert abs(vals[:, 2] @ vals[:, 1] - 2 * e_ref) < 1e-13
    assert abs(vals[:, 2] @ vals[:, 3] - sd) < 1e-13
    assert abs(vals[:, 2] @ vals[:, 4] - sx) < 1e-13
    np.testing.assert_allclose(vals[:, 3] + vals[:, 4], vals[:, 1], rtol=0, atol=1e-15)
    if amps == "mp2":
        assert abs(e_ref - c["e2"]) < 1e-10


def test_openmp_build_gives_the_same_bits():
    """liboracle_omp.so (the all-cores CPU baseline) splits the term loops
    over threads by the outer occupied index; each element keeps its serial
    summation order, so residuals and the whole trajectory are identical."""
    from tests.helpers import info_for, oracle_for
    s = oracle_for(info_for("n14_cut5_tw"))
    omp = OracleSystem(s.params, variant="omp")
    _, _, tab, tba = s.mp2()
    for a, b in zip(s.terms(tab, tba), omp.terms(tab, tba)):
        np.testing.assert_array_equal(a, b)
    r, ro = s.ccd(), omp.ccd()
    assert r["niter"] == ro["niter"] and r["ecorr"] == ro["ecorr"]
    np.testing.assert_array_equal(r["trace_e"], ro["trace_e"])


def test_mp2_structure_factor_pins(fixtures):
    """DrvMBPT's binning (unit 81): the cells and (SfG, SfSum, SfV) equal the
    first columns of CCEnergyAndSF's on the MP2 amplitudes (pinned above),
    sum_q SfV SfSum = 2 E(2), and E(2) is the reference's printed value."""
    n24 = fixtures["n24"]
    s = _sys(n24["params"])
    q, vals, e2 = s.mp2_structure_factor()
    assert abs(e2 - n24["e2"]) < F18_12
    _, _, tab, tba = s.mp2()
    qs, vs, _ = s.structure_factor(tab, tba)
    np.testing.assert_array_equal(q, qs)
    np.testing.assert_array_equal(vals, vs[:, :3])
    assert abs((vals[:, 2] * vals[:, 1]).sum() - 2.0 * e2) < 1e-13


def test_exchange_structure_factor_pins(fixtures):
    """print_ex_SF (unit 82): the counts are the ordered occupied pairs by
    transfer (brute force), ExV is the Coulomb kernel 1/(pi L |q|^2), and
    sum_q count ExV is minus the HF exchange energy, which with the kinetic
    energy and the Madelung term gives the reference's printed E(SCF)."""
    n24 = fixtures["n24"]
    s = _sys(n24["params"])
    q, cnt, exv = s.exchange_structure_factor()
    no = s.nocc
    n = s.nvec[:no]
    pairs = {}
    for a in range(no):
        for b in range(no):
            if a != b:
                k = tuple(n[a] - n[b])
                pairs[k] = pairs.get(k, 0) + 1
    assert sorted(pairs) == [tuple(x) for x in q]
    assert [pairs[tuple(x)] for x in q] == list(cnt)
    q2 = (q.astype(float) ** 2).sum(axis=1)
    np.testing.assert_allclose(exv, 1.0 / (np.pi * s.L * q2), rtol=1e-13)
    tw = np.array(n24["params"]["twist"])
    kin = ((2.0 * np.pi / s.L) ** 2 * ((n + tw) ** 2).sum(axis=1)).sum()  # 2 x sum 1/2 |k|^2
    xhf = -(cnt * exv).sum()
    assert abs(kin + xhf + s.madelung * 0.5 * 2 * no - n24["ehf"]) < F18_12


def test_drpa_scan_energies(fixtures):
    """The reference's dRPA/SOSEX basis scan (benchmarks/ZRun-dRPA/old-dRPA:
    N=14, r_s=1, Gamma, mosaics off): Ec(dRPA) and Ec(SOSEX) within their
    twelve printed decimals, and the basis size, on the rows up to cutoff 50
    (1,503 orbitals)."""
    sc = fixtures["drpa_scan"]
    rows = [r for r in sc["rows"] if r[0] <= 50]
    assert len(rows) == 10
    for cut, ed, es, two_m in rows:
        s = _sys(sc["params"], maxkpoint=cut)
        assert 2 * s.nao == two_m
        r = s.drpa()
        assert not r["failed"]
        assert abs(r["ecorr"] - ed) < F18_12, cut
        assert abs(r["esosex"] - es) < F18_12, cut
