"""CPU-side checks of the product: the C-ABI library loads and exports what
include/uegccd.h declares, the host basis builder agrees with the oracle, the
Input parser follows source/IO.f90, and calls fail loudly without a device."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests.helpers import SYSTEMS, info_for, oracle_for
from uegccd_amd import lib as L
from uegccd_amd.inputfile import InputError, parse_input, standard_input

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "uegccd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(uegccd_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = header_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    assert set(L.EXPORTS) <= set(names)


def test_strerror_and_defaults():
    lib = L.load()
    for st in range(0, -9, -1):
        assert lib.uegccd_strerror(st)
    o = L.Options()
    lib.uegccd_default_options(ctypes.byref(o))
    assert o.mode == L.MODE_REFERENCE and o.max_iter == 1000 and o.precision == L.PREC_FP64


@pytest.mark.parametrize("name", ["n14_cut4", "n14_cut5_tw", "n24_rs5", "n54_cut16", "n14_cut25"])
def test_heg_build_matches_oracle(name):
    from uegccd_amd.driver import init_heg
    info = info_for(name)
    heg = init_heg(info)
    orc = oracle_for(info)
    assert heg.NAO == orc.nao and heg.NOcc == orc.nocc
    np.testing.assert_array_equal(heg.nVec, orc.nvec)
    np.testing.assert_allclose(heg.Eigen, orc.eigen, rtol=0, atol=1e-14)
    assert heg.L == pytest.approx(orc.L, rel=1e-15)
    assert heg.Madelung == pytest.approx(orc.madelung, rel=1e-13)
    assert heg.EHF == pytest.approx(orc.ehf, rel=1e-13, abs=1e-13)
    assert heg.DenomFactor == orc.denomfactor and heg.TolMax == orc.tolmax


def test_basis_sizes_of_the_configs():
    from uegccd_amd.driver import init_heg
    sizes = {n: init_heg(info_for(n)).NAO for n in ("n14_cut4", "n14_cut5", "n14_cut25", "n54_cut16")}
    assert sizes == {"n14_cut4": 33, "n14_cut5": 57, "n14_cut25": 515, "n54_cut16": 257}


def test_no_device_fails_loudly():
    """Without a HIP device the solver refuses to run (no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from uegccd_amd.driver import CCDContext, init_heg
    info = info_for("n14_cut4")
    with pytest.raises(L.UegccdError) as e:
        CCDContext(init_heg(info), info)
    assert e.value.status in (L.ERR_HIP, L.ERR_INVALID)


def test_footprint_is_host_only_and_grows_with_history():
    """uegccd_footprint plans a context without a device: it covers at least
    the amplitudes, the DIIS history (T and error vectors of the valid
    elements only), the residual outputs and the ring intermediates; the
    history grows by two vectors of the valid amplitudes per entry (each
    history array rounded to 256 B).  The valid elements (those with a
    virtual momentum partner) are counted with the oracle's FindIndex."""
    from uegccd_amd.driver import footprint, init_heg, options
    info = info_for("n54_cut9")
    heg = init_heg(info)
    o, v = heg.NOcc, heg.NAO - heg.NOcc
    n3 = o * o * v
    orc = oracle_for(info)
    assert (orc.nocc, orc.nvir) == (o, v)
    nvalid = sum(1 for a in range(o + 1, o + v + 1) for j in range(1, o + 1)
                 for i in range(1, o + 1) if orc.findindex(i, j, a) > o)
    assert 0 < nvalid < n3
    b4 = footprint(heg, info, options(mode="diis", ndiis=4))
    b8 = footprint(heg, info, options(mode="diis", ndiis=8))
    vec = (16 * n3 + 255) // 256 * 256

    def hist(nd):
        return 2 * ((16 * nvalid * nd + 255) // 256 * 256)

    assert b8 - b4 == hist(8) - hist(4)
    assert b8 >= vec * 3 + hist(8) + 2 * 32 * o * v * o
    assert b8 < 4 * b4
    with pytest.raises(L.UegccdError):
        footprint(heg, info, options(mode="diis", ndiis=17))


def test_server_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from uegccd_amd.driver import Server
    with pytest.raises(L.UegccdError) as e:
        Server()
    assert e.value.status in (L.ERR_HIP, L.ERR_INVALID)


# --- Input parser (source/IO.f90) ------------------------------------------

def test_parse_standard_input():
    info = parse_input(standard_input(54, 16, 2.0))
    assert (info.NElectron, info.NOcc, info.MaxKPoint, info.rSMin) == (54, 27, 16, 2.0)
    assert info.SafeERI and info.DoRing and info.DoXRing and info.DoLadder and info.DoMosaic
    assert info.irange == [0] * 15 and info.exchangefactor == 1.0 and info.DoSkipTA


def test_parse_reference_test_input_layout():
    # the layout of source/Test/Input: comments after values, NTwist, odd spacing
    text = standard_input(14, 2, 1.0, (0.1, 0.1, 0.1), DoSingleCalc=False)
    text = text.replace("corefcut:          1", "corefcut:         1     ! e.g. 1.414213562")
    info = parse_input(text)
    assert info.NTwist == (0.1, 0.1, 0.1) and info.corefcut == 1.0 and not info.DoSingleCalc


@pytest.mark.parametrize("mutate,msg", [
    (lambda t: t + "Bogus Key: 1\n", "not recognized"),
    (lambda t: t + "Safe ERI: T\n", "multiply set"),
    (lambda t: t.replace("DoKS:", "!DoKS:"), "DoKS not set"),
    (lambda t: t.replace("EndInput:          T", "EndInput:          F"), "EndInput must be T"),
    (lambda t: t.replace("# Electrons:       14", "# Electrons:       13"), "closed shell"),
    (lambda t: t.replace("# rS Points:       1", "# rS Points:       2"), "more than one rs"),
    (lambda t: t.replace("DoTruncCoulombHF:  F", "DoTruncCoulombHF:  T"), "Not Implemented"),
])
def test_parse_errors(mutate, msg):
    text = mutate(standard_input(14, 2, 1.0))
    with pytest.raises(InputError) as e:
        parse_input(text)
    assert msg in str(e.value)


def test_corefactor_and_fcut2ecut_derivations():
    t = standard_input(14, 2, 1.0, CoreFactor=0.5, fcut2ecut=True, fcut=3.0)
    info = parse_input(t)
    assert info.CoreN == 3  # int(14/2 * 0.5)
    # Nfermi = 14^(2/3)/4 < 10 -> MaxKPoint = int(10 * fcut * 2)
    assert info.MaxKPoint == 60


# --- fort.59 records (Fortran list-directed output) --------------------------

def _f59_items(line):
    import re
    toks = line.split()
    items, text = [], []
    for t in toks:
        if re.match(r"^[-+]?\d+$", t):
            items.append(int(t))
        elif re.match(r"^[-+]?\d*\.\d+(E[-+]\d+)?$", t):
            items.append(float(t))
        else:
            text.append(t)
    return items, " ".join(text)


def test_fort59_records_reproduce_reference_bytes(fixtures):
    """Every line of the reference's twist-averaging fort.59, re-printed from
    its own numbers by uegccd_amd.fortran_io, is byte-identical."""
    from uegccd_amd.fortran_io import record
    lines = fixtures["n24"]["ta_fort59_lines"]
    assert len(lines) == 27
    for line in lines:
        items, text = _f59_items(line)
        if line.strip().startswith(("special", "lowest")):
            out = record(text, items)
        elif text:
            out = record(items, text)
        else:
            out = record(*items)
        assert out == line, (out, line)


@pytest.mark.parametrize("name", ["n14_cut5", "n24_rs5", "n54_cut16", "n14_cut5_tw"])
def test_exchange_structure_factor_vs_oracle(name):
    """uegccd_exchange_structure_factor (print_ex_SF, unit 82; host code)
    against the oracle at the Input's r_s and at r_s = 1, the two blocks
    print_HF writes per basis (source/HEG.f90:319,531)."""
    from uegccd_amd.driver import exchange_structure_factor
    info = info_for(name)
    for rs in (info.rSMin, 1.0):
        q, cnt, v = exchange_structure_factor(info, rs=rs)
        o = oracle_for(info_for(name, rSMin=rs, rSMax=rs))
        oq, ocnt, ov = o.exchange_structure_factor()
        np.testing.assert_array_equal(q, oq)
        np.testing.assert_array_equal(cnt, ocnt)
        np.testing.assert_allclose(v, ov, rtol=1e-14)


def _twisted_sphere(ec, twist):
    """Every integer n with |n + twist|^2 <= ec, in the stable order of
    |n + twist|^2 over the enumeration (a, then b, then c ascending): the
    basis of the twisted cutoff before any fcut truncation, by brute force."""
    cell = int(np.sqrt(ec)) + 3
    pts = []
    for a in range(-cell, cell + 1):
        for b in range(-cell, cell + 1):
            for c in range(-cell, cell + 1):
                x, y, z = a + twist[0], b + twist[1], c + twist[2]
                k2 = x * x + y * y + z * z
                if k2 <= ec:
                    pts.append((k2, len(pts), (a, b, c)))
    pts.sort(key=lambda p: (p[0], p[1]))
    return np.array([p[2] for p in pts], dtype=np.int32)


TWISTED = [(14, 5, (0.11, -0.23, 0.05)), (38, 9, (0.5, 0.5, 0.5)), (54, 16, (-0.31, 0.42, 0.07)),
           (14, 25, (0.5, 0.0, -0.25))]


@pytest.mark.parametrize("nel,ec,twist", TWISTED)
def test_twisted_cutoff_basis(nel, ec, twist):
    """The opt-in twisted sphere |n + twist|^2 <= E_c: the oracle's basis is
    the brute-force lattice sphere around -twist in SetupBasis's stable order,
    the host builder gives the oracle's n-vectors and eigenvalues, and every
    orbital lies inside the sphere, none outside."""
    from uegccd_amd.driver import init_heg
    from uegccd_amd.twist_stream import request_info
    info = request_info((nel, ec, 1.5, twist), cutoff_twisted=True)
    orc = oracle_for(info)
    want = _twisted_sphere(ec, twist)
    assert orc.nao == len(want)
    np.testing.assert_array_equal(orc.nvec, want)
    heg = init_heg(info)
    assert heg.NAO == orc.nao
    np.testing.assert_array_equal(heg.nVec, orc.nvec)
    np.testing.assert_allclose(heg.Eigen, orc.eigen, rtol=0, atol=1e-14)
    assert heg.EHF == pytest.approx(orc.ehf, rel=1e-13, abs=1e-13)
    nt2 = ((heg.nVec + np.array(twist)) ** 2).sum(axis=1)
    assert nt2.max() <= ec
    # the untwisted cutoff of the same request is a different set of plane waves
    plain = init_heg(request_info((nel, ec, 1.5, twist)))
    assert {tuple(v) for v in plain.nVec} != {tuple(v) for v in heg.nVec}


def test_twisted_cutoff_at_gamma_is_the_reference_cutoff():
    from uegccd_amd.driver import init_heg
    from uegccd_amd.twist_stream import request_info
    for nel, ec in ((14, 5), (54, 16)):
        a = init_heg(request_info((nel, ec, 2.0, (0.0, 0.0, 0.0))))
        b = init_heg(request_info((nel, ec, 2.0, (0.0, 0.0, 0.0)), cutoff_twisted=True))
        np.testing.assert_array_equal(a.nVec, b.nVec)
        np.testing.assert_array_equal(a.Eigen, b.Eigen)


def test_cutoff_flag_is_checked():
    lib = L.load()
    p = L.HegParams()
    p.nel, p.maxkpoint, p.rs, p.fcut, p.corefcut = 14, 5, 1.0, 100.0, 1.0
    p.madelungfactor, p.exchangefactor = 1.0, 1.0
    p.cutoff_twisted = 2
    n = ctypes.c_int32()
    assert lib.uegccd_heg_build(ctypes.byref(p), ctypes.byref(n), None, None, None) == L.ERR_INVALID


@pytest.mark.parametrize("twisted", [False, True])
def test_oracle_findindex_on_both_cutoffs(twisted):
    """The oracle's FindIndex returns the (1-based) orbital of n_m + n_n - n_l
    whenever that momentum is in the basis and 0 otherwise, by brute force
    over every triple of a twisted system under either cutoff (the twisted
    sphere holds plane waves outside |n|^2 <= E_c, which the reference's early
    exit would miss)."""
    from uegccd_amd.twist_stream import request_info
    info = request_info((14, 5, 1.0, (0.41, -0.23, 0.35)), cutoff_twisted=twisted)
    orc = oracle_for(info)
    where = {tuple(v): k + 1 for k, v in enumerate(orc.nvec)}
    if twisted:
        assert max(int((v ** 2).sum()) for v in orc.nvec) > 5  # outside the untwisted sphere
    n = orc.nao
    for m in range(1, n + 1):
        for l in range(1, n + 1):
            for nn in range(1, 8):
                want = where.get(tuple(orc.nvec[m - 1] + orc.nvec[nn - 1] - orc.nvec[l - 1]), 0)
                assert orc.findindex(m, nn, l) == want


def _scan_info(cut, fixtures):
    p = fixtures["drpa_scan"]["params"]
    return parse_input(standard_input(p["nel"], cut, p["rs"], tuple(p["twist"]), **{
        "Do Mosaics": p["do_mosaic"], "DodRPASOSEX": True, "fcut": p["fcut"]}))


def test_basis_sizes_of_the_drpa_scan(fixtures):
    """Every row of the reference's dRPA basis scan (N=14, r_s=1, cutoffs
    5..665): the host builder's basis has the scan's 2M spin orbitals."""
    from uegccd_amd.driver import init_heg
    for cut, _, _, two_m in fixtures["drpa_scan"]["rows"]:
        assert 2 * init_heg(_scan_info(cut, fixtures)).NAO == two_m, cut


def test_planning_takes_the_largest_scan_basis(fixtures):
    """A context of the scan's largest basis (71,999 orbitals) plans on the
    host: its footprint is known without a device, and the amplitude arrays
    (o*o*v entries) dominate it."""
    from uegccd_amd.driver import footprint, init_heg, options
    info = _scan_info(665, fixtures)
    heg = init_heg(info)
    assert heg.NAO == 71999
    n3 = heg.NOcc ** 2 * (heg.NAO - heg.NOcc)
    b = footprint(heg, info, options(mode="reference"))
    assert 16 * n3 * 4 < b < 16 * n3 * 40
