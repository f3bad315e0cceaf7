"""Shared builders for the tests: the same system through the oracle (checker)
and through the product host path (uegccd_amd)."""
from __future__ import annotations

import numpy as np

from oracle import OracleParams, OracleSystem
from uegccd_amd.inputfile import parse_input, standard_input

SYSTEMS = {
    # name: (nel, cutoff, rs, twist)
    "n14_cut4": (14, 4, 1.0, (0.0, 0.0, 0.0)),
    "n14_cut5": (14, 5, 1.0, (0.0, 0.0, 0.0)),
    "n14_cut5_tw": (14, 5, 2.0, (0.17, -0.31, 0.07)),
    "n24_rs5": (24, 40, 5.0, (0.1, 0.2, 0.45)),
    "n38_cut9": (38, 9, 1.0, (0.0, 0.0, 0.0)),
    "n54_cut9": (54, 9, 2.0, (0.0, 0.0, 0.0)),
    "n14_cut25": (14, 25, 1.0, (0.0, 0.0, 0.0)),
    "n54_cut16": (54, 16, 2.0, (0.0, 0.0, 0.0)),
}


def info_for(name, **over):
    nel, cut, rs, tw = SYSTEMS[name]
    extra = {}
    if name == "n24_rs5":
        extra["fcut"] = 2.0
    extra.update(over)
    return parse_input(standard_input(nel, cut, rs, tw, **extra))


def oracle_for(info, eigen=None) -> OracleSystem:
    p = OracleParams(nel=info.NElectron, maxkpoint=info.MaxKPoint, rs=info.rSMin,
                     twist=tuple(info.NTwist), fcut=info.fcut, corefcut=info.corefcut,
                     madelungfactor=info.madelungfactor, gap=info.Gap, gapii=info.GapII,
                     exchangefactor=info.exchangefactor, coren=info.CoreN,
                     safe_eri=info.SafeERI, do_ring=info.DoRing, do_xring=info.DoXRing,
                     do_ladder=info.DoLadder, do_mosaic=info.DoMosaic,
                     irange=tuple(info.irange), cutoff_twisted=bool(info.CutoffTwisted))
    s = OracleSystem(p)
    if eigen is not None:
        s.set_eigen(eigen)
    return s


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)
