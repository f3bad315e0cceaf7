"""The ZCode Fortran host (ISO_C_BINDING over libuegccd) on reference Inputs:
its Output file carries the reference's blocks and numbers."""
import os
import re
import subprocess

import numpy as np
import pytest

from uegccd_amd.inputfile import standard_input

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ZCODE = os.path.join(ROOT, "uegccd_amd", "host", "ZCode")
FLOAT = r"[-+]?\d*\.\d+(?:[EeDd][-+]?\d+)?"


def run_zcode(tmp_path, text, env=None):
    if not os.path.exists(ZCODE):
        pytest.fail("ZCode was not built (uegccd_amd/build.py zcode)")
    (tmp_path / "Input").write_text(text)
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([ZCODE], cwd=tmp_path, capture_output=True, text=True, timeout=120, env=e)
    assert "Finishing calculation after single point calculation" in (r.stdout + r.stderr), r.stdout + r.stderr
    return (tmp_path / "Output").read_text().splitlines()


def parse(lines):
    out = {"eig": [], "trace": []}
    for ln in lines:
        if "E(SCF) is" in ln:
            out["ehf"] = float(re.findall(FLOAT, ln)[0])
        elif "E(SCF) + E(2)" in ln:
            pass
        elif "E(2) is" in ln:
            out["e2"] = float(re.findall(FLOAT, ln)[0])
        elif re.match(r"^\s+\*\s+" + FLOAT + r"\s+\*\s*$", ln):
            out["eig"].append(float(re.findall(FLOAT, ln)[0]))
        elif re.match(r"^\s+\*\s+" + FLOAT + r"\s+\d+\s+" + FLOAT + r"\s+\*", ln):
            t = ln.split()
            out["trace"].append((float(t[1]), int(t[2]), float(t[3])))
        elif "converged in" in ln:
            out["niter"] = int(re.findall(r"in\s+(\d+)", ln)[0])
        elif "Ec(CCD) is" in ln or "Ec(dRPA) is" in ln:
            out["ecorr"] = float(re.findall(FLOAT, ln)[0])
        elif "Ec(SOSEX) is" in ln:
            out["esosex"] = float(re.findall(FLOAT, ln)[0])
        elif "Max CCD Residual" in ln:
            out["maxres"] = float(re.findall(FLOAT, ln)[0])
    return out


def test_zcode_n24_reference_output(tmp_path, fixtures):
    print("[zcode] n24", flush=True)
    fx = fixtures["n24"]
    lines = run_zcode(tmp_path, standard_input(24, 40, 5.0, (0.1, 0.2, 0.45), fcut=2.0))
    o = parse(lines)
    assert o["ehf"] == pytest.approx(fx["ehf"], abs=1e-12)
    assert o["e2"] == pytest.approx(fx["e2"], abs=1e-12)
    np.testing.assert_allclose(o["eig"], fx["eigen"], rtol=0, atol=1e-15)
    assert o["niter"] == 70
    ref = np.array(fx["ccd_trace"])
    got = np.array(o["trace"])
    np.testing.assert_array_equal(got[:, 1], ref[:, 1])
    np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=0, atol=1.1e-10)
    np.testing.assert_allclose(got[:, 2], ref[:, 2], rtol=0, atol=1.1e-10)
    assert o["ecorr"] == pytest.approx(fx["ccd_ecorr"], abs=1e-12)
    assert o["maxres"] == pytest.approx(fx["ccd_maxres"], rel=1e-6)


@pytest.mark.parametrize("drpa", [False, True])
def test_zcode_matches_python_host(tmp_path, drpa):
    """Same Input through ZCode (Fortran) and the Python host: same Output text
    (CCD, and dRPA/SOSEX with mosaics on)."""
    print(f"[zcode] n14 cut5 vs python host drpa={drpa}", flush=True)
    from uegccd_amd.driver import single_point
    text = standard_input(14, 5, 1.0, DodRPASOSEX=drpa)
    zlines = run_zcode(tmp_path, text)
    ptext, heg, e2, r = single_point(text)
    assert zlines == ptext.splitlines()


def test_zcode_diis_mode(tmp_path):
    """DIIS mode through the environment knob converges to the same energy."""
    import json
    cfg = json.load(open(os.path.join(ROOT, "tests", "golden", "config_energies.json")))
    lines = run_zcode(tmp_path, standard_input(14, 5, 1.0), env={"UEGCCD_MODE": "diis"})
    o = parse(lines)
    assert o["ecorr"] == pytest.approx(cfg["n14_cut5"]["ecorr_tight"], abs=1e-9)
    assert o["niter"] < cfg["n14_cut5"]["niter_reference_mode"]


def test_zcode_drpa_reference_output(tmp_path, fixtures):
    """The reference's dRPA benchmark Input (DodRPASOSEX T, mosaics off):
    ZCode's Output carries the benchmark's dRPA trajectory, Ec(dRPA),
    Ec(SOSEX) and max residual (benchmarks/ZRun-dRPA/Output)."""
    print("[zcode] dRPA benchmark", flush=True)
    fx = fixtures["n14_cut5"]
    lines = run_zcode(tmp_path, standard_input(14, 5, 1.0, **{"Do Mosaics": False, "DodRPASOSEX": True}))
    assert any("dRPA summary follows" in ln for ln in lines)
    o = parse(lines)
    assert o["e2"] == pytest.approx(fx["e2"], abs=1e-12)
    ref = np.array(fx["drpa_trace"])
    got = np.array(o["trace"])
    assert o["niter"] == fx["drpa_niter"] == len(got)
    np.testing.assert_array_equal(got[:, 1], ref[:, 1])
    np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=0, atol=1.1e-10)
    np.testing.assert_allclose(got[:, 2], ref[:, 2], rtol=0, atol=1.1e-10)
    assert o["ecorr"] == pytest.approx(fx["drpa_ecorr"], abs=1e-12)
    assert o["esosex"] == pytest.approx(fx["drpa_esosex"], abs=1e-12)
    assert o["maxres"] == pytest.approx(fx["drpa_maxres"], rel=1e-6)


def test_zcode_structure_factor_unit80(tmp_path):
    """DoSFCalcMP2 and DoSFCalcCCD: ZCode writes the two unit-80 blocks of
    CCEnergyAndSF (source/CCD.f90:65,109), equal to the library's structure
    factor of the MP2 guess and of the converged amplitudes."""
    print("[zcode] n14 cut5 structure factor", flush=True)
    from uegccd_amd.driver import CCDContext, init_heg, options
    from uegccd_amd.inputfile import parse_input
    text = standard_input(14, 5, 1.0, DoSFCalcMP2=True, DoSFCalcCCD=True)
    run_zcode(tmp_path, text)
    # list-directed records may wrap over several lines: a block is every
    # number after its header, eight to a cell
    blocks, cur = [], None
    for line in (tmp_path / "fort.80").read_text().splitlines():
        if "Starting" in line:
            cur = []
            blocks.append(cur)
        elif "--------" not in line and cur is not None:
            cur.extend(float(x) for x in line.split())
    assert len(blocks) == 2
    blocks = [np.array(b).reshape(-1, 8) for b in blocks]
    info = parse_input(text)
    ctx = CCDContext(init_heg(info), info, options(mode="reference"))
    _, _, tab, tba = ctx.mp2()
    expect = [ctx.structure_factor(tab, tba)]
    r = ctx.solve(tab, tba)
    expect.append(ctx.structure_factor(r.tab, r.tba))
    for got, (q, vals, _) in zip(blocks, expect):
        got = np.array(got)
        assert got.shape == (len(q), 8)
        assert (got[:, :3].astype(int) == q).all()
        np.testing.assert_allclose(got[:, 3:], vals, rtol=0, atol=1e-13 * np.abs(vals).max())


def _sf_blocks(path, ncol):
    """The blocks of a structure-factor unit: every number after each
    header, ncol to a cell (list-directed records may wrap)."""
    blocks, cur = [], None
    for line in path.read_text().splitlines():
        if "Starting" in line:
            cur = []
            blocks.append(cur)
        elif "--------" not in line and cur is not None:
            cur.extend(float(x) for x in line.split())
    return [np.array(b).reshape(-1, ncol) for b in blocks]


def test_zcode_structure_factor_units_81_82(tmp_path):
    """Every DrvMBPT call writes DrvMBPT's structure factor to unit 81
    (source/MP2.f90:105-117) and every basis set-up print_ex_SF's two blocks
    to unit 82, at r_s = 1 and at the Input's r_s (source/HEG.f90:319,531);
    ZCode's blocks equal the library's and the oracle's."""
    print("[zcode] units 81 and 82", flush=True)
    from tests.helpers import oracle_for
    from uegccd_amd.driver import CCDContext, exchange_structure_factor, init_heg, options
    from uegccd_amd.inputfile import parse_input
    text = standard_input(14, 5, 2.0, (0.1, -0.2, 0.3))
    run_zcode(tmp_path, text)
    info = parse_input(text)
    (b81,) = _sf_blocks(tmp_path / "fort.81", 6)
    q, vals, _ = oracle_for(info).mp2_structure_factor()
    assert (b81[:, :3].astype(int) == q).all()
    np.testing.assert_allclose(b81[:, 3:], vals, rtol=0, atol=1e-13 * np.abs(vals).max())
    ctx = CCDContext(init_heg(info), info, options(mode="reference"))
    lq, lvals, _ = ctx.mp2_structure_factor()
    np.testing.assert_allclose(b81[:, 3:], lvals, rtol=0, atol=1e-13 * np.abs(vals).max())
    blocks = _sf_blocks(tmp_path / "fort.82", 5)
    assert len(blocks) == 2
    for blk, rs in zip(blocks, (1.0, info.rSMin)):
        q, cnt, v = exchange_structure_factor(info, rs=rs)
        assert (blk[:, :3].astype(int) == q).all() and (blk[:, 3].astype(int) == cnt).all()
        np.testing.assert_allclose(blk[:, 4], v, rtol=1e-13)


def test_zcode_twist_average_reference_run(tmp_path, fixtures):
    """DoSkipTA F: ZCode runs the reference's twist averaging
    (source/Main.f90:99-255) on the reference's own TA Input (N=24, r_s=5,
    cutoff 40, fcut 2) over its ten recorded twists (source/Test/fort.59):
    every fort.59 number (per-twist E_HF and E(2), averages, squared
    averages, errors, averaged eigenvalues and connectivity, connectivity
    distances, the special twist with its tie between twists 6 and 7), the
    per-twist connectivity in fort.58 and the final MP2 + 84-iteration CCD
    with the averaged eigenvalues in Output."""
    print("[zcode] n24 twist averaging", flush=True)
    fx = fixtures["n24"]
    (tmp_path / "twists").write_text("\n".join(" ".join(repr(x) for x in t) for t in fx["ta_twists"]) + "\n")
    text = standard_input(24, 40, 5.0, (0.1, 0.2, 0.45), fcut=2.0, DoSkipTA=False, TACalcN=10)
    (tmp_path / "Input").write_text(text)
    e = dict(os.environ, UEGCCD_TWISTS=str(tmp_path / "twists"))
    r = subprocess.run([ZCODE], cwd=tmp_path, capture_output=True, text=True, timeout=300, env=e)
    assert r.returncode == 0, r.stdout + r.stderr
    ours = (tmp_path / "fort.59").read_text().splitlines()
    ref = fx["ta_fort59_lines"]
    assert len(ours) == len(ref) == 27
    for k in range(27):   # gfortran's list-directed layout, field by field
        assert re.sub(r"\d", "0", ours[k]) == re.sub(r"\d", "0", ref[k]), (k, ours[k], ref[k])
    num = re.compile(r"^[-+]?\d*\.\d+(E[-+]\d+)?$")
    nums = lambda ln: [float(t) for t in ln.split() if num.match(t)]  # noqa: E731
    for k in range(10):   # per twist: the twist (as printed), E_HF, E(2), E(2) again (no TA CCD)
        assert ours[k].split()[:3] == ref[k].split()[:3]
        np.testing.assert_allclose(nums(ours[k])[3:], nums(ref[k])[3:], rtol=0, atol=1e-13)
    for k in (10, 11):    # averages, squared averages
        np.testing.assert_allclose(nums(ours[k]), nums(ref[k]), rtol=0, atol=1e-13)
    np.testing.assert_allclose(nums(ours[12]), nums(ref[12]), rtol=1e-10)      # errors
    np.testing.assert_allclose(nums(ours[13]), nums(ref[13]), rtol=0, atol=1e-14)  # eigenvalues
    assert nums(ours[14]) == nums(ref[14])                                      # connectivity
    for k in range(15, 25):
        assert int(ours[k].split()[0]) == k - 14
        np.testing.assert_allclose(nums(ours[k])[-1:], nums(ref[k])[-1:], rtol=1e-12)
    assert ours[25].split() == ref[25].split()      # special twist angle found (twist 6)
    assert abs(nums(ours[26])[0] - nums(ref[26])[0]) < 1e-12 * nums(ref[26])[0]
    assert len((tmp_path / "fort.58").read_text().splitlines()) == 10
    # the final MP2 and CCD (the last blocks of Output)
    lines = (tmp_path / "Output").read_text().splitlines()
    last = max(k for k, ln in enumerate(lines) if "E(2) is" in ln)
    o = parse(lines[last - 2:])
    assert o["e2"] == pytest.approx(fx["ta_final_e2"], abs=1e-12)
    assert o["niter"] == fx["ta_final_niter"] == 84
    tr = np.array(o["trace"])
    refp = np.array(fx["ta_final_trace"])
    np.testing.assert_allclose(tr[:, 0], refp[:, 0], rtol=0, atol=1.1e-10)
    np.testing.assert_allclose(tr[:, 2], refp[:, 2], rtol=0, atol=1.1e-10)
    assert o["ecorr"] == pytest.approx(fx["ta_final_ecorr"], abs=1e-12)


def test_zcode_twist_average_served_ccd(tmp_path):
    """DoCalcTACCD T: every twist's CCD runs through the library's request
    server; each per-twist record's CCD energy equals that twist's solve
    alone."""
    print("[zcode] twist averaging with served CCD", flush=True)
    from uegccd_amd.driver import CCDContext, init_heg, options
    from uegccd_amd.inputfile import parse_input
    twists = [(0.1, -0.2, 0.3), (-0.25, 0.05, 0.4), (0.33, 0.12, -0.41)]
    (tmp_path / "twists").write_text("\n".join(" ".join(repr(x) for x in t) for t in twists) + "\n")
    text = standard_input(14, 5, 1.0, DoSkipTA=False, DoCalcTACCD=True, TACalcN=3)
    (tmp_path / "Input").write_text(text)
    e = dict(os.environ, UEGCCD_TWISTS=str(tmp_path / "twists"))
    r = subprocess.run([ZCODE], cwd=tmp_path, capture_output=True, text=True, timeout=300, env=e)
    assert r.returncode == 0, r.stdout + r.stderr
    info = parse_input(text)
    rec = (tmp_path / "fort.59").read_text().splitlines()
    for k, tw in enumerate(twists):
        ctx = CCDContext(init_heg(info, twist=tw, device=0), info, options(mode="reference"))
        _, _, tab, tba = ctx.mp2()
        alone = ctx.solve(tab, tba)
        assert abs(float(rec[k].split()[5]) - alone.ecorr) < 1e-12
        ctx.close()
