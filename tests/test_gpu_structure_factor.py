"""Structure factor on the device (uegccd_structure_factor, the reference's
CCEnergyAndSF, source/CCD.f90:240-351) against the oracle's restatement, on
the MP2 guess and on converged CCD amplitudes."""
import numpy as np
import pytest

from tests.helpers import info_for, oracle_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,converged", [("n14_cut4", False), ("n14_cut5_tw", True), ("n54_cut9", False)])
def test_structure_factor_vs_oracle(name, converged):
    from uegccd_amd.driver import CCDContext, init_heg, options
    print(f"[sf] {name} converged={converged}", flush=True)
    info = info_for(name)
    ctx = CCDContext(init_heg(info), info, options(mode="reference"))
    _, _, tab, tba = ctx.mp2()
    if converged:
        r = ctx.solve(tab, tba)
        tab, tba = r.tab, r.tba
    q, vals, e = ctx.structure_factor(tab, tba)
    qo, vo, eo = oracle_for(info).structure_factor(tab, tba)
    assert q.shape == qo.shape and (q == qo).all()
    scale = np.abs(vo).max(axis=0)
    np.testing.assert_allclose(vals, vo, rtol=0, atol=1e-13 * scale.max())
    np.testing.assert_allclose(vals[:, 0], vo[:, 0], rtol=1e-15)
    np.testing.assert_allclose(vals[:, 2], vo[:, 2], rtol=1e-15)
    assert abs(e - eo) < 1e-12


@pytest.mark.parametrize("name,over", [("n14_cut4", {}), ("n14_cut5_tw", {}), ("n54_cut9", {}),
                                       ("n14_cut5_tw", {"CoreFactor": 0.3}),
                                       ("n14_cut5", {"Energy Range": 1})])
def test_mp2_structure_factor_vs_oracle(name, over):
    """uegccd_mp2_structure_factor (DrvMBPT's binning, unit 81) against the
    oracle's restatement: the same cells, SfG and SfV to the last bits, SfSum
    within 1e-13 of the largest value, E(2) within 1e-12 Ha.  With an energy
    range flag SfV stays DrvMBPT's flag-0 integral; with CoreN the masked core
    elements still bin their cells."""
    from uegccd_amd.driver import CCDContext, init_heg, options
    print(f"[mp2 sf] {name} {over}", flush=True)
    info = info_for(name, **over)
    ctx = CCDContext(init_heg(info), info, options(mode="reference"))
    q, vals, e2 = ctx.mp2_structure_factor()
    qo, vo, eo = oracle_for(info).mp2_structure_factor()
    assert q.shape == qo.shape and (q == qo).all()
    np.testing.assert_allclose(vals[:, 0], vo[:, 0], rtol=1e-15)
    np.testing.assert_allclose(vals[:, 2], vo[:, 2], rtol=1e-15)
    np.testing.assert_allclose(vals[:, 1], vo[:, 1], rtol=0, atol=1e-13 * np.abs(vo[:, 1]).max())
    assert abs(e2 - eo) < 1e-12
