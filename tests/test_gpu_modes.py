"""Mixed precision and batched solves against the plain FP64 solve.

Bars: mixed precision within 1e-6 Ha of the FP64 fixed point (with at least
one FP32 ladder iteration); a batched solve equals the same system solved
alone to 1e-12 Ha with the same iteration count (the kernels reduce each
system's blocks in a fixed order regardless of the batch it sits in).
"""
import json
import os

import numpy as np
import pytest

from tests.helpers import info_for

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = json.load(open(os.path.join(ROOT, "tests", "golden", "config_energies.json")))


def _ctx(info, **opt):
    from uegccd_amd.driver import CCDContext, init_heg, options
    heg = init_heg(info)
    return CCDContext(heg, info, options(**opt))


@pytest.mark.parametrize("name", ["n14_cut5", "n54_cut9", "n38_cut9", "n54_cut16"])
def test_mixed_precision_energy(name):
    """n54_cut16 is the mixed-precision workload itself: N=54, r_s=2, |n|^2<=16,
    FP32 ladders while max|R| >= 1e-5, then FP64."""
    print(f"[mixed] {name}", flush=True)
    info = info_for(name)
    c64 = _ctx(info, mode="diis", ndiis=8)
    _, _, tab, tba = c64.mp2()
    r64 = c64.solve(tab, tba)
    cmx = _ctx(info, mode="diis", ndiis=8, precision="mixed")
    rmx = cmx.solve(tab, tba)
    assert rmx.status == 0 and not rmx.failed
    assert rmx.iters_fp32 >= 1
    assert abs(rmx.ecorr - r64.ecorr) < 1e-6
    assert abs(rmx.ecorr - CFG[name]["ecorr_tight"]) < 1e-6


def test_mixed_precision_reference_mode_switches():
    """In the reference iteration the FP32 phase ends once max|R| < 1e-5 and the
    FP64 iterations that follow bring the energy back to the FP64 trajectory's
    fixed point."""
    info = info_for("n14_cut5")
    c = _ctx(info, mode="reference", precision="mixed")
    _, _, tab, tba = c.mp2()
    r = c.solve(tab, tba)
    assert 1 <= r.iters_fp32 < r.niter
    assert abs(r.ecorr - CFG["n14_cut5"]["ecorr_reference_mode"]) < 1e-6


def test_batch_equals_alone():
    print("[batch] mixed systems", flush=True)
    from uegccd_amd.driver import solve_batch
    names = [("n14_cut5", {}), ("n14_cut5_tw", {}), ("n38_cut9", {}), ("n24_rs5", {})]
    ctxs, guesses, alone = [], [], []
    for name, over in names:
        c = _ctx(info_for(name, **over), mode="diis", ndiis=8)
        _, _, tab, tba = c.mp2()
        alone.append(c.solve(tab, tba))
        ctxs.append(c)
        guesses.append((tab, tba))
    res = solve_batch(ctxs, [g[0] for g in guesses], [g[1] for g in guesses])
    for a, b in zip(alone, res):
        assert b.status == 0
        assert b.niter == a.niter
        assert abs(b.ecorr - a.ecorr) < 1e-12
        np.testing.assert_allclose(b.tab, a.tab, rtol=0, atol=1e-14)


def test_nonfinite_is_isolated_in_a_batch():
    """A request whose eigenvalues hold a NaN fails alone; the others in the
    batch are unchanged."""
    from uegccd_amd.driver import CCDContext, init_heg, options, solve_batch
    good_info = info_for("n14_cut5")
    good = _ctx(good_info, mode="diis", ndiis=8)
    _, _, tab, tba = good.mp2()
    ref = good.solve(tab, tba)
    bad_info = info_for("n14_cut4")
    heg = init_heg(bad_info)
    heg.Eigen = heg.Eigen.copy()
    heg.Eigen[3] = np.nan
    bad = CCDContext(heg, bad_info, options(mode="diis", ndiis=8))
    btab = np.zeros(bad.shape, order="F")
    res = solve_batch([good, bad], [tab, btab], [tba, btab])
    assert res[0].status == 0 and abs(res[0].ecorr - ref.ecorr) < 1e-12
    assert res[1].status != 0


@pytest.mark.parametrize("name,opt", [("n54_cut16", dict(mode="diis", ndiis=8)),
                                      ("n24_rs5", dict(mode="reference")),
                                      ("n38_cut9", dict(mode="diis", ndiis=6, precision="mixed"))])
def test_graph_replay_equals_launches(name, opt):
    """Chunks of passes replayed from captured HIP graphs give the same bits
    as the same passes launched kernel by kernel.  Each solve replays its
    chunk graphs several times with the host's state polls (copies and
    synchronisations) between them, and three solves in a row reuse the
    captured graphs."""
    print(f"[graph] {name} {opt}", flush=True)
    info = info_for(name)
    ref = _ctx(info, graph=0, **opt)
    _, _, tab, tba = ref.mp2()
    r0 = ref.solve(tab, tba)
    g = _ctx(info, graph=1, **opt)
    for _ in range(3):
        r = g.solve(tab, tba)
        assert r.niter == r0.niter and r.ecorr == r0.ecorr
        np.testing.assert_array_equal(r.tab, r0.tab)
        np.testing.assert_array_equal(r.tba, r0.tba)
        np.testing.assert_array_equal(r.trace_e, r0.trace_e)


@pytest.mark.parametrize("name,opt", [("n54_cut16", dict(mode="diis", ndiis=8)),
                                      ("n14_cut5", dict(mode="reference")),
                                      ("n38_cut9", dict(mode="diis", ndiis=6, precision="mixed"))])
def test_persistent_kernel_equals_launches(name, opt):
    """The persistent pass kernel (a chunk of passes in one cooperative launch,
    grid barriers between the phases) gives the same bits as the five
    launches per pass: the same work items, each computed in the same order."""
    print(f"[persistent] {name} {opt}", flush=True)
    info = info_for(name)
    ref = _ctx(info, graph=0, **opt)
    _, _, tab, tba = ref.mp2()
    r0 = ref.solve(tab, tba)
    p = _ctx(info, persistent=1, **opt)
    for _ in range(2):
        r = p.solve(tab, tba)
        assert r.niter == r0.niter and r.ecorr == r0.ecorr
        np.testing.assert_array_equal(r.tab, r0.tab)
        np.testing.assert_array_equal(r.tba, r0.tba)


def test_tile_choices_give_the_same_bits():
    """The tuned kernels' tile choices change how loads are grouped, not the
    order of the accumulation: every choice gives the solve's bits; invalid
    choices are refused."""
    import ctypes
    from uegccd_amd import lib as L
    info = info_for("n38_cut9")
    ref = _ctx(info, mode="diis", ndiis=8)
    _, _, tab, tba = ref.mp2()
    r0 = ref.solve(tab, tba)
    lib = L.load()
    for g in (1, 2, 4):
        for qs in (4, 8):
            for hs in (4, 8):
                L.check(lib.uegccd_set_tiles(ref.handle, ctypes.byref(L.Tiles(g, qs, hs))), "set_tiles")
                r = ref.solve(tab, tba)
                assert r.niter == r0.niter and r.ecorr == r0.ecorr, (g, qs, hs)
                np.testing.assert_array_equal(r.tab, r0.tab)
    assert lib.uegccd_set_tiles(ref.handle, ctypes.byref(L.Tiles(3, 8, 8))) == L.ERR_INVALID
    L.check(lib.uegccd_set_tiles(ref.handle, None), "set_tiles reset")
