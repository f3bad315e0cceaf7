"""The allocation-free solve (uegccd_solve_noalloc, through the C ABI) against
uegccd_ctx_create + uegccd_solve on the same systems: two calls in a row on
one device, the second with a smaller system (the scratch keeps the first
call's contents), and a system larger than the scratch refused with
UEGCCD_ERR_SCRATCH.  Bar: the same kernels in the same order, so the same
NIter and energy within 1e-12 Ha, amplitudes within 1e-13."""
import ctypes

import numpy as np
import pytest

from tests.helpers import info_for

pytestmark = pytest.mark.gpu


def _noalloc(info, opts):
    from uegccd_amd import lib as L
    from uegccd_amd.driver import _system, init_heg
    heg = init_heg(info)
    s, keep = _system(heg, info)
    shape = (heg.NOcc, heg.NOcc, heg.NAO - heg.NOcc)
    return heg, s, keep, shape


def _compare(name, opts_kw):
    from uegccd_amd import lib as L
    from uegccd_amd.driver import CCDContext, options
    info = info_for(name)
    opts = options(**opts_kw)
    heg, s, keep, shape = _noalloc(info, opts)
    ctx = CCDContext(heg, info, opts)
    _, _, tab, tba = ctx.mp2()
    ref = ctx.solve(tab, tba)
    ctx.close()
    a, b, aa = (np.array(x, order="F", copy=True) for x in (tab, tba, np.zeros(shape)))
    res = L.Result()
    lib = L.load()
    print(f"[noalloc] {name} {opts_kw}", flush=True)
    st = lib.uegccd_solve_noalloc(ctypes.byref(s), ctypes.byref(opts), L.dptr(aa), L.dptr(a),
                                  L.dptr(b), ctypes.byref(res))
    assert st == L.OK, L.strerror(st)
    assert res.niter == ref.niter, (res.niter, ref.niter)
    assert abs(res.ecorr - ref.ecorr) < 1e-12, (res.ecorr, ref.ecorr)
    assert np.abs(a - ref.tab).max() < 1e-13 and np.abs(b - ref.tba).max() < 1e-13
    assert np.abs(aa - (ref.tab + ref.tba)).max() < 1e-13
    assert res.iters_fp32 == ref.iters_fp32 and res.failed == ref.failed


def test_noalloc_equals_solve_twice():
    # the N=54 |n|^2<=16 DIIS(8) solve (56 MB of the 64 MiB scratch), then a
    # much smaller system in the same scratch, in both iteration modes
    _compare("n54_cut16", dict(mode="diis", ndiis=8))
    _compare("n14_cut5", dict(mode="reference"))
    _compare("n38_cut9", dict(mode="diis", ndiis=6, precision="mixed"))
    _compare("n14_cut4", dict(mode="diis", ndiis=8))


def test_noalloc_refuses_oversize():
    from uegccd_amd import lib as L
    from uegccd_amd.driver import footprint, init_heg, options
    from uegccd_amd.inputfile import parse_input, standard_input
    info = parse_input(standard_input(54, 25, 2.0, (0.0, 0.0, 0.0)))
    opts = options(mode="diis", ndiis=8)
    assert footprint(init_heg(info), info, opts) > (L.NOALLOC_SCRATCH_MB << 20)
    heg, s, keep, shape = _noalloc(info, opts)
    a, b = np.zeros(shape, order="F"), np.zeros(shape, order="F")
    res = L.Result()
    st = L.load().uegccd_solve_noalloc(ctypes.byref(s), ctypes.byref(opts), None, L.dptr(a),
                                       L.dptr(b), ctypes.byref(res))
    assert st == L.ERR_SCRATCH
