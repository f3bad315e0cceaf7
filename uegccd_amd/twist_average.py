"""Twist averaging, the reference's Main path after the single point
(source/Main.f90:99-255), on libuegccd.

For each twist: the plane-wave basis, E_HF, the MP2 guess and E(2) and the MP2
connectivity histogram, on the GPU; with DoCalcTACCD, the CCD of every twist,
iterated together as one served batch.  Then the reference's statistics
(averages, squared averages, standard errors, the twist-averaged eigenvalues
and connectivity), the special twist whose connectivity is closest to the
average (first minimum of sum_k ((C_t(k) - <C>(k)) / k^2)^2), and the MP2 and
CCD at that twist with the twist-averaged eigenvalues.  fort.59 and fort.58
are produced record for record in the reference's list-directed layout.

The reference draws twists with gfortran's rand() (source/HEG.f90:535-549);
here they are given, or drawn from numpy's generator with the same form
(a single-precision uniform number minus 0.5 per component).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import lib as L
from .driver import (CCDContext, DrvCCD, Server, SolveResult, init_heg, options,
                     format_ccd_block, format_mp2_block)
from .fortran_io import record
from .inputfile import UEGInfo

CONNECTIVITY_LIMIT = 400     # the reference's histogram length (source/MP2.f90:26)


def generate_twists(n: int, seed: int = 86456):
    """n twists in [-0.5, 0.5)^3, each component a single-precision uniform
    number minus 0.5 (the form of GenerateTwist)."""
    rng = np.random.default_rng(seed)
    return [tuple(float(x) - 0.5 for x in rng.random(3, dtype=np.float32)) for _ in range(n)]


@dataclass
class TwistPoint:
    twist: tuple
    hf: float
    mp2: float
    ccd: float
    eigen: np.ndarray = field(repr=False)
    connectivity: np.ndarray = field(repr=False)


@dataclass
class TwistAverage:
    points: list
    ntwist: np.ndarray
    hf: float
    mp2: float
    ccd: float
    ntwist2: np.ndarray
    hf2: float
    mp22: float
    ccd2: float
    ntwist_err: np.ndarray
    hf_err: float
    mp2_err: float
    ccd_err: float
    eigen_avg: np.ndarray = field(repr=False)
    conn_avg: np.ndarray = field(repr=False)
    conn_diff2: list = field(repr=False)
    special_index: int = 0
    special_twist: tuple = ()
    min_diff2: float = 0.0
    final_e2: float = 0.0
    final: SolveResult | None = field(default=None, repr=False)
    fort59: list = field(default_factory=list, repr=False)
    fort58: list = field(default_factory=list, repr=False)
    output: str = field(default="", repr=False)


def _last_nonzero(c):
    nz = np.nonzero(c)[0]
    return int(nz[-1]) + 1 if nz.size else 0


def twist_average(info: UEGInfo, twists=None, seed: int = 86456, device: int = 0,
                  opts: L.Options | None = None, ccd_opts: L.Options | None = None) -> TwistAverage:
    """The twist-averaging run of source/Main.f90:99-255 for `info` over
    `twists` (default: info.TACalcN generated twists).  opts are the solver
    options of the final CCD (default: the reference's iteration); ccd_opts
    those of the per-twist CCDs when info.DoCalcTACCD (default: DIIS(8))."""
    if twists is None:
        twists = generate_twists(info.TACalcN, seed)
    twists = [tuple(float(x) for x in t) for t in twists]
    n = len(twists)
    if n < 2:
        raise ValueError("twist averaging needs at least two twists")
    fort59, fort58 = [], []
    points = []
    served = {}
    if info.DoCalcTACCD:
        srv = Server(device=device, max_active=min(64, n), iters_per_step=4)
        copt = ccd_opts if ccd_opts is not None else options(mode="diis", ndiis=8)
        served = {k: srv.submit(info, copt, twist=tw) for k, tw in enumerate(twists)}
        srv.drain()
        served = {k: srv.poll(rid) for k, rid in served.items()}
        srv.close()
    conn_max = None
    for k, tw in enumerate(twists):
        heg = init_heg(info, twist=tw, device=device)
        ctx = CCDContext(heg, info)
        e2, *_ = ctx.mp2()
        conn = ctx.connectivity(CONNECTIVITY_LIMIT)
        ctx.close()
        ccd = e2
        if info.DoCalcTACCD:
            r = served[k]
            if r.status not in (L.OK, L.ERR_NOT_CONVERGED):
                raise L.UegccdError(r.status, f"CCD at twist {k + 1}")
            ccd = r.ecorr
        p = TwistPoint(tw, heg.EHF, e2, ccd, heg.Eigen.copy(), conn)
        points.append(p)
        fort59.append(record(list(tw), p.hf, p.mp2, p.ccd))
        if conn_max is None:
            conn_max = _last_nonzero(conn)
        fort58.append(record(conn[:conn_max].tolist()))
    nt = np.array([p.twist for p in points])
    hf = np.array([p.hf for p in points])
    mp2 = np.array([p.mp2 for p in points])
    ccd = np.array([p.ccd for p in points])
    # sums in twist order, as the reference accumulates them
    def acc(vals):
        s = 0.0 * vals[0]
        for v in vals:
            s = s + v
        return s
    ntwist, hfa, mp2a, ccda = acc(nt) / n, acc(hf) / n, acc(mp2) / n, acc(ccd) / n
    ntwist2, hf2, mp22, ccd22 = acc(nt ** 2) / n, acc(hf ** 2) / n, acc(mp2 ** 2) / n, acc(ccd ** 2) / n
    err = lambda x2, x: ((x2 - x ** 2) / (n - 1)) ** 0.5  # noqa: E731
    conn_sum = acc(np.array([p.connectivity for p in points], dtype=np.int64))
    conn_avg = conn_sum.astype(np.float64) / n
    eigen_avg = acc(np.array([p.eigen for p in points])) / n
    res = TwistAverage(points, ntwist, hfa, mp2a, ccda, ntwist2, hf2, mp22, ccd22,
                       err(ntwist2, ntwist), err(hf2, hfa), err(mp22, mp2a), err(ccd22, ccda),
                       eigen_avg, conn_avg, [])
    fort59.append(record(list(ntwist), hfa, mp2a, ccda, "averages"))
    fort59.append(record(list(ntwist2), hf2, mp22, ccd22, "squared averages"))
    fort59.append(record(list(res.ntwist_err), res.hf_err, res.mp2_err, res.ccd_err, "errors"))
    fort59.append(record(list(eigen_avg), "eigenvalue averages"))
    fort59.append(record(list(conn_avg[:_last_nonzero(conn_sum)]), "conn"))
    # the twist whose connectivity is closest to the average
    weight = 1.0 / (np.arange(1, CONNECTIVITY_LIMIT + 1, dtype=np.float64) ** 2.0)
    best = None
    for k, p in enumerate(points):
        diff = (p.connectivity - conn_avg) * weight
        d2 = 0.0
        for v in diff:
            d2 += v * v
        res.conn_diff2.append(d2)
        if best is None or d2 < best[1]:
            best = (k, d2)
        fort59.append(record(k + 1, d2))
    res.special_index, res.min_diff2 = best
    res.special_twist = points[best[0]].twist
    fort59.append(record("special twist angle found", list(res.special_twist)))
    fort59.append(record("lowest connectivity diff squared", res.min_diff2))
    # MP2 and CCD at the special twist with the twist-averaged eigenvalues
    heg = init_heg(info, twist=res.special_twist, device=device)
    heg.Eigen = eigen_avg.copy()
    ctx = CCDContext(heg, info, opts)
    e2, taa, tab, tba = ctx.mp2()
    res.final_e2 = e2
    heg.ECorr = e2
    mp2_block = format_mp2_block(heg, e2)
    res.final = DrvCCD(heg, info, taa, tab, tba, ctx=ctx)
    res.output = mp2_block + format_ccd_block(heg, info, res.final)
    ctx.close()
    res.fort59, res.fort58 = fort59, fort58
    return res
