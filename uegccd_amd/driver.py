"""Host mirror of the reference's solver interface, running on libuegccd.

Names follow the reference so a caller can switch over routine by routine:
  init_heg      init_HEG + change_rs      (source/HEG.f90:16-43, 494-533)
  DrvMBPT       MP2 guess and E(2)        (source/MP2.f90:11-157)
  DrvCCD        the CCD solve             (source/CCD.f90:18-183)
  CCDContext    a solver bound to one basis (residual / energy / solve)
  single_point  Main's DoSingleCalc path  (source/Main.f90:44-97)
Amplitude arrays are numpy float64 arrays of shape (NOcc, NOcc, NAO-NOcc) in
Fortran order, the reference's T(NOcc,NOcc,NOcc+1:NAO).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import lib as L
from .inputfile import UEGInfo, parse_input


@dataclass
class HEGData:
    """HEGDataType (source/Types.f90:7-27) at the calculation's r_s."""
    Nel: int
    NOcc: int
    NAO: int
    L: float
    Omega: float
    Madelung: float
    EHF: float
    ScreeningDistance: float
    rS: float
    nTwist: tuple
    nVec: np.ndarray          # (NAO, 3) int32
    Eigen: np.ndarray         # (NAO,) float64
    DenomFactor: float = 2.0
    TolMax: float = 1e-8
    FailRatio: float = 100.0
    ECorr: float = 0.0


def init_heg(info: UEGInfo, twist=None, device: int | None = None) -> HEGData:
    """Basis, eigenvalues, Madelung and E_HF: built on HIP device `device`
    (uegccd_heg_build_device), or by the host builder (uegccd_heg_build) when
    device is None; both give the same bits."""
    lib = L.load()
    p = L.HegParams()
    p.nel, p.maxkpoint, p.rs = info.NElectron, info.MaxKPoint, info.rSMin
    tw = info.NTwist if twist is None else twist
    for d in range(3):
        p.twist[d] = float(tw[d])
    p.fcut, p.corefcut = info.fcut, info.corefcut
    p.madelungfactor, p.gap, p.exchangefactor = info.madelungfactor, info.Gap, info.exchangefactor
    p.cutoff_twisted = int(info.CutoffTwisted)
    if device is None:
        build = lambda *a: lib.uegccd_heg_build(ctypes.byref(p), *a)  # noqa: E731
        name = "uegccd_heg_build"
    else:
        build = lambda *a: lib.uegccd_heg_build_device(ctypes.byref(p), int(device), *a)  # noqa: E731
        name = "uegccd_heg_build_device"
    nao = ctypes.c_int32()
    L.check(build(ctypes.byref(nao), None, None, None), name)
    n = nao.value
    nvec_cm = np.zeros(3 * n, dtype=np.int32)
    eig = np.zeros(n)
    sc = np.zeros(8)
    L.check(build(ctypes.byref(nao), L.iptr(nvec_cm), L.dptr(eig), L.dptr(sc)), name)
    info.NAO = n
    return HEGData(Nel=info.NElectron, NOcc=info.NElectron // 2, NAO=n, L=sc[0], Omega=sc[1],
                   Madelung=sc[2], EHF=sc[3], ScreeningDistance=sc[4], rS=info.rSMin,
                   nTwist=tuple(tw), nVec=nvec_cm.reshape(3, n).T.copy(), Eigen=eig,
                   DenomFactor=sc[5], TolMax=sc[6], FailRatio=sc[7])


def exchange_structure_factor(info: UEGInfo, rs: float | None = None, twist=None):
    """print_ex_SF (source/HEG.f90:617-674; unit 82) of the basis at r_s (the
    Input's rSMin by default): (q [n,3], count [n], ExV [n])."""
    lib = L.load()
    p = L.HegParams()
    p.nel, p.maxkpoint = info.NElectron, info.MaxKPoint
    p.rs = info.rSMin if rs is None else rs
    tw = info.NTwist if twist is None else twist
    for d in range(3):
        p.twist[d] = float(tw[d])
    p.fcut, p.corefcut = info.fcut, info.corefcut
    p.madelungfactor, p.gap, p.exchangefactor = info.madelungfactor, info.Gap, info.exchangefactor
    p.cutoff_twisted = int(info.CutoffTwisted)
    n = ctypes.c_int32(0)
    L.check(lib.uegccd_exchange_structure_factor(ctypes.byref(p), ctypes.byref(n), None, None, None),
            "uegccd_exchange_structure_factor")
    q = np.zeros((max(n.value, 1), 3), dtype=np.int32)
    cnt = np.zeros(max(n.value, 1), dtype=np.int32)
    v = np.zeros(max(n.value, 1))
    L.check(lib.uegccd_exchange_structure_factor(ctypes.byref(p), ctypes.byref(n), L.iptr(q), L.iptr(cnt),
                                                 L.dptr(v)), "uegccd_exchange_structure_factor")
    return q[:n.value], cnt[:n.value], v[:n.value]


def options(mode="reference", ndiis=8, precision="fp64", **kw) -> L.Options:
    lib = L.load()
    o = L.Options()
    lib.uegccd_default_options(ctypes.byref(o))
    o.mode = {"reference": L.MODE_REFERENCE, "diis": L.MODE_DIIS}[mode]
    o.ndiis = ndiis
    o.precision = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED}[precision]
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class CCDContext:
    """A device context for one basis: uegccd_ctx_create .. uegccd_ctx_destroy."""

    def __init__(self, heg: HEGData, info: UEGInfo, opts: L.Options | None = None):
        self._lib = L.load()
        self.heg, self.info = heg, info
        self.nocc, self.nao = heg.NOcc, heg.NAO
        self.shape = (self.nocc, self.nocc, self.nao - self.nocc)
        s = L.System()
        self._nvec = np.ascontiguousarray(heg.nVec.T, dtype=np.int32).reshape(-1)  # nVec(NAO,3)
        self._eig = np.ascontiguousarray(heg.Eigen, dtype=np.float64)
        s.nocc, s.nao = self.nocc, self.nao
        s.nvec, s.eigen = L.iptr(self._nvec), L.dptr(self._eig)
        s.L, s.madelung, s.rs, s.gapii = heg.L, heg.Madelung, heg.rS, info.GapII
        s.coren, s.safe_eri = info.CoreN, int(info.SafeERI)
        for k in range(15):
            s.irange[k] = int(info.irange[k])
        s.do_ring, s.do_xring = int(info.DoRing), int(info.DoXRing)
        s.do_ladder, s.do_mosaic = int(info.DoLadder), int(info.DoMosaic)
        s.screening = heg.ScreeningDistance
        self.opts = opts if opts is not None else options()
        h = ctypes.c_void_p()
        L.check(self._lib.uegccd_ctx_create(ctypes.byref(s), ctypes.byref(self.opts),
                                            ctypes.byref(h)), "uegccd_ctx_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.uegccd_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def zeros(self):
        return np.zeros(self.shape, order="F")

    def info_sizes(self):
        no, na, nb = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        nv = ctypes.c_int64()
        L.check(self._lib.uegccd_ctx_info(self._h, ctypes.byref(no), ctypes.byref(na),
                                          ctypes.byref(nv), ctypes.byref(nb)))
        return dict(nocc=no.value, nao=na.value, nvalid=nv.value, nblocks=nb.value)

    def _f(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"amplitude shape {a.shape} != {self.shape}")
        return a

    def mp2(self):
        taa, tab, tba = self.zeros(), self.zeros(), self.zeros()
        e2 = ctypes.c_double()
        L.check(self._lib.uegccd_mp2(self._h, L.dptr(taa), L.dptr(tab), L.dptr(tba),
                                     ctypes.byref(e2)), "uegccd_mp2")
        return e2.value, taa, tab, tba

    def residual(self, tab, tba):
        tab, tba = self._f(tab), self._f(tba)
        out = [self.zeros() for _ in range(4)]
        L.check(self._lib.uegccd_residual(self._h, L.dptr(tab), L.dptr(tba),
                                          *[L.dptr(x) for x in out]), "uegccd_residual")
        return tuple(out)

    def energy(self, tab, tba):
        tab, tba = self._f(tab), self._f(tba)
        e = ctypes.c_double()
        L.check(self._lib.uegccd_energy(self._h, L.dptr(tab), L.dptr(tba), ctypes.byref(e)),
                "uegccd_energy")
        return e.value

    def solve(self, tab, tba, allow_unconverged=False):
        tab = np.array(self._f(tab), order="F", copy=True)
        tba = np.array(self._f(tba), order="F", copy=True)
        taa = self.zeros()
        r = L.Result()
        st = self._lib.uegccd_solve(self._h, L.dptr(taa), L.dptr(tab), L.dptr(tba),
                                    ctypes.byref(r))
        if st not in (L.OK, L.ERR_NOT_CONVERGED) or (st != L.OK and not allow_unconverged):
            if not (allow_unconverged and st == L.ERR_NOT_CONVERGED):
                raise L.UegccdError(st, "uegccd_solve")
        return SolveResult.from_c(r, taa, tab, tba, self.trace(r.niter))

    def structure_factor(self, tab, tba):
        """uegccd_structure_factor: (q [n,3] int, vals [n,5] = SfG, SfSum, SfV,
        SfDSum, SfXSum, energy) of the given amplitudes (CCEnergyAndSF,
        source/CCD.f90:240-351)."""
        tab = np.asfortranarray(tab, dtype=np.float64)
        tba = np.asfortranarray(tba, dtype=np.float64)
        n = ctypes.c_int32(0)
        L.check(self._lib.uegccd_structure_factor(self._h, None, None, ctypes.byref(n), None, None, None),
                "uegccd_structure_factor")
        q = np.zeros((max(n.value, 1), 3), dtype=np.int32)
        vals = np.zeros((max(n.value, 1), 5))
        e = ctypes.c_double(0.0)
        L.check(self._lib.uegccd_structure_factor(
            self._h, L.dptr(tab), L.dptr(tba), ctypes.byref(n),
            q.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), L.dptr(vals), ctypes.byref(e)),
            "uegccd_structure_factor")
        return q[:n.value], vals[:n.value], e.value

    def mp2_structure_factor(self):
        """uegccd_mp2_structure_factor: DrvMBPT's binning of the MP2 amplitudes
        (source/MP2.f90:86-117; unit 81): (q [n,3], vals [n,3] = SfG, SfSum,
        SfV, E(2))."""
        n = ctypes.c_int32(0)
        L.check(self._lib.uegccd_mp2_structure_factor(self._h, ctypes.byref(n), None, None, None),
                "uegccd_mp2_structure_factor")
        q = np.zeros((max(n.value, 1), 3), dtype=np.int32)
        vals = np.zeros((max(n.value, 1), 3))
        e = ctypes.c_double(0.0)
        L.check(self._lib.uegccd_mp2_structure_factor(
            self._h, ctypes.byref(n), q.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), L.dptr(vals),
            ctypes.byref(e)), "uegccd_mp2_structure_factor")
        return q[:n.value], vals[:n.value], e.value

    def drpa(self, t2=None, allow_unconverged=False):
        """DrvdRPA (source/dRPA.f90:23-143) on this basis: the ring-only CCD on
        the spin-summed T2 (default start: the MP2 T2aaaa), iterated as the
        reference does.  Returns (SolveResult with ecorr = Ec(dRPA) and t2 in
        its tab field, Ec(SOSEX))."""
        t2 = self.zeros() if t2 is None else np.array(self._f(t2), order="F", copy=True)
        start = L.dptr(t2) if np.any(t2) else None
        r = L.Result()
        es = ctypes.c_double()
        st = self._lib.uegccd_drpa(self._h, start, ctypes.byref(r), ctypes.byref(es))
        if st != L.OK and not (allow_unconverged and st == L.ERR_NOT_CONVERGED):
            raise L.UegccdError(st, "uegccd_drpa")
        if start is None:
            L.check(self._lib.uegccd_get_amplitudes(self._h, None, L.dptr(t2), None),
                    "uegccd_get_amplitudes")
        zero = self.zeros()
        return SolveResult.from_c(r, t2, t2, zero, self.trace(r.niter)), es.value

    def connectivity(self, nbins=400):
        """DrvMBPT's connectivity histogram (uegccd_connectivity): bin k-1
        counts the excitations with |n_i - n_a|^2 = k and |n_i - n_b|^2 = k."""
        h = np.zeros(nbins, dtype=np.int32)
        L.check(self._lib.uegccd_connectivity(self._h, L.iptr(h), nbins), "uegccd_connectivity")
        return h

    def trace(self, n):
        e, dt = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        L.check(self._lib.uegccd_trace(self._h, L.dptr(e), L.dptr(dt), n), "uegccd_trace")
        return e[:n], dt[:n]


@dataclass
class SolveResult:
    ecorr: float
    niter: int
    dt: float
    maxres: float
    failed: bool
    iters_fp32: int
    status: int
    solve_ms: float
    taa: np.ndarray = field(repr=False, default=None)
    tab: np.ndarray = field(repr=False, default=None)
    tba: np.ndarray = field(repr=False, default=None)
    trace_e: np.ndarray = field(repr=False, default=None)
    trace_dt: np.ndarray = field(repr=False, default=None)

    @classmethod
    def from_c(cls, r, taa, tab, tba, trace):
        return cls(ecorr=r.ecorr, niter=r.niter, dt=r.dt, maxres=r.maxres, failed=bool(r.failed),
                   iters_fp32=r.iters_fp32, status=r.status, solve_ms=r.solve_ms,
                   taa=taa, tab=tab, tba=tba, trace_e=trace[0], trace_dt=trace[1])


def solve_batch(ctxs, tabs, tbas):
    """uegccd_solve_batch over several contexts (one launch per kernel per pass)."""
    lib = L.load()
    n = len(ctxs)
    P = ctypes.c_void_p * n
    D = ctypes.POINTER(ctypes.c_double) * n
    tabs = [np.array(np.asfortranarray(t, dtype=np.float64), order="F", copy=True) for t in tabs]
    tbas = [np.array(np.asfortranarray(t, dtype=np.float64), order="F", copy=True) for t in tbas]
    res = (L.Result * n)()
    st = lib.uegccd_solve_batch(P(*[c.handle.value for c in ctxs]), n,
                                D(*[L.dptr(t) for t in tabs]), D(*[L.dptr(t) for t in tbas]), res)
    # per-system failures (not converged, non-finite) are reported in each result
    if st not in (L.OK, L.ERR_NOT_CONVERGED, L.ERR_NONFINITE):
        raise L.UegccdError(st, "uegccd_solve_batch")
    out = []
    for k in range(n):
        r = res[k]
        out.append(SolveResult.from_c(r, tabs[k] + tbas[k], tabs[k], tbas[k],
                                      ctxs[k].trace(r.niter)))
    return out


def _system(heg: HEGData, info: UEGInfo):
    """uegccd_system for a basis; the arrays it points into are returned too."""
    nvec = np.ascontiguousarray(heg.nVec.T, dtype=np.int32).reshape(-1)
    eig = np.ascontiguousarray(heg.Eigen, dtype=np.float64)
    s = L.System()
    s.nocc, s.nao = heg.NOcc, heg.NAO
    s.nvec, s.eigen = L.iptr(nvec), L.dptr(eig)
    s.L, s.madelung, s.rs, s.gapii = heg.L, heg.Madelung, heg.rS, info.GapII
    s.coren, s.safe_eri = info.CoreN, int(info.SafeERI)
    for k in range(15):
        s.irange[k] = int(info.irange[k])
    s.do_ring, s.do_xring = int(info.DoRing), int(info.DoXRing)
    s.do_ladder, s.do_mosaic = int(info.DoLadder), int(info.DoMosaic)
    s.screening = heg.ScreeningDistance
    return s, (nvec, eig)


def footprint(heg: HEGData, info: UEGInfo, opts: L.Options | None = None) -> int:
    """Device bytes a solver context for this basis holds (uegccd_footprint)."""
    lib = L.load()
    s, keep = _system(heg, info)
    o = opts if opts is not None else options()
    b = ctypes.c_int64()
    L.check(lib.uegccd_footprint(ctypes.byref(s), ctypes.byref(o), ctypes.byref(b)),
            "uegccd_footprint")
    del keep
    return b.value


@dataclass
class RequestResult:
    state: int
    status: int
    nao: int
    nocc: int
    niter: int
    iters_fp32: int
    ehf: float
    e2: float
    ecorr: float
    maxres: float
    t_submit_ms: float
    t_start_ms: float
    t_done_ms: float
    footprint_bytes: int
    plan_ms: float

    @property
    def latency_ms(self):
        return self.t_done_ms - self.t_submit_ms


class Server:
    """Continuous-batching request loop (uegccd_server_*): twist averaging and
    r_s scans submitted as independent solves, iterated together on one GPU."""

    def __init__(self, device=0, hbm_budget_bytes=0, max_active=64, iters_per_step=4,
                 planner_threads=0):
        self._lib = L.load()
        o = L.ServerOptions()
        o.device, o.hbm_budget_bytes = device, int(hbm_budget_bytes)
        o.max_active, o.iters_per_step = max_active, iters_per_step
        o.planner_threads = planner_threads
        h = ctypes.c_void_p()
        L.check(self._lib.uegccd_server_create(ctypes.byref(o), ctypes.byref(h)),
                "uegccd_server_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.uegccd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, info: UEGInfo, opts: L.Options | None = None, twist=None) -> int:
        q = L.Request()
        p = q.heg
        p.nel, p.maxkpoint, p.rs = info.NElectron, info.MaxKPoint, info.rSMin
        tw = info.NTwist if twist is None else twist
        for d in range(3):
            p.twist[d] = float(tw[d])
        p.fcut, p.corefcut = info.fcut, info.corefcut
        p.madelungfactor, p.gap, p.exchangefactor = info.madelungfactor, info.Gap, info.exchangefactor
        p.cutoff_twisted = int(info.CutoffTwisted)
        q.gapii, q.coren, q.safe_eri = info.GapII, info.CoreN, int(info.SafeERI)
        for k in range(15):
            q.irange[k] = int(info.irange[k])
        q.do_ring, q.do_xring = int(info.DoRing), int(info.DoXRing)
        q.do_ladder, q.do_mosaic = int(info.DoLadder), int(info.DoMosaic)
        q.opt = opts if opts is not None else options()
        rid = ctypes.c_int64()
        L.check(self._lib.uegccd_server_submit(self._h, ctypes.byref(q), ctypes.byref(rid)),
                "uegccd_server_submit")
        return rid.value

    def step(self):
        na, nq = ctypes.c_int32(), ctypes.c_int32()
        L.check(self._lib.uegccd_server_step(self._h, ctypes.byref(na), ctypes.byref(nq)),
                "uegccd_server_step")
        return na.value, nq.value

    def drain(self):
        L.check(self._lib.uegccd_server_drain(self._h), "uegccd_server_drain")

    def poll(self, rid: int) -> RequestResult:
        r = L.RequestResult()
        L.check(self._lib.uegccd_server_poll(self._h, int(rid), ctypes.byref(r)),
                "uegccd_server_poll")
        return RequestResult(**{k: getattr(r, k) for k, _ in L.RequestResult._fields_})

    def clock_ms(self) -> float:
        return self._lib.uegccd_server_clock_ms(self._h)


def DrvMBPT(heg: HEGData, info: UEGInfo, ctx: CCDContext | None = None):
    """MP2 amplitudes (T2aaaa, T2abab, T2abba); sets heg.ECorr = E(2)."""
    c = ctx or CCDContext(heg, info)
    e2, taa, tab, tba = c.mp2()
    heg.ECorr = e2
    return taa, tab, tba


def DrvCCD(heg: HEGData, info: UEGInfo, T2aaaa, T2abab, T2abba, opts: L.Options | None = None,
           ctx: CCDContext | None = None, sf_out: list | None = None):
    """Solve in place from the given amplitudes; sets heg.ECorr; returns SolveResult.
    With sf_out, the unit-80 structure-factor blocks the reference writes with
    DoSFCalcMP2 (of the starting amplitudes, source/CCD.f90:65) and DoSFCalcCCD
    (of the converged ones, source/CCD.f90:109) are appended to it."""
    c = ctx or CCDContext(heg, info, opts)
    if sf_out is not None and info.DoSFCalcMP2:
        sf_out.append(format_sf_block(*c.structure_factor(T2abab, T2abba)[:2]))
    r = c.solve(T2abab, T2abba, allow_unconverged=True)
    if sf_out is not None and info.DoSFCalcCCD:
        sf_out.append(format_sf_block(*c.structure_factor(r.tab, r.tba)[:2]))
    T2abab[...] = r.tab
    T2abba[...] = r.tba
    T2aaaa[...] = r.tab + r.tba
    heg.ECorr = r.ecorr
    return r


def format_sf_block(q, vals) -> str:
    """Unit 80 of CCEnergyAndSF (source/CCD.f90:265-267, 334-342): three header
    lines, then one list-directed row I, J, A, SfG, SfSum, SfV, SfDSum, SfXSum
    per cell (integers in 12 columns, reals to 17 significant digits; the
    spacing of list-directed output is the compiler's)."""
    out = [" --------", " Starting a new structure factor", " --------"]
    for (x, y, z), r in zip(q, vals):
        out.append(f"{x:12d}{y:12d}{z:12d}" + "".join(f"{v:26.17E}" for v in r))
    return "\n".join(out) + "\n"


def DrvdRPA(heg: HEGData, info: UEGInfo, T2, opts: L.Options | None = None,
            ctx: CCDContext | None = None):
    """DrvdRPA (source/dRPA.f90:23-143) in place from the spin-summed T2 (the
    reference passes T2aaaa); sets heg.ECorr as the reference leaves it
    (Ec(SOSEX) after a converged run, else Ec(dRPA)); returns (SolveResult,
    Ec(SOSEX))."""
    c = ctx or CCDContext(heg, info, opts)
    r, esosex = c.drpa(T2, allow_unconverged=True)
    T2[...] = r.tab
    heg.ECorr = r.ecorr if r.failed else esosex
    return r, esosex


# --- Output blocks, formats of source/MP2.f90:143-150, source/CCD.f90:127-140
# and source/dRPA.f90:122-139 ---
_STAR = " " * 14 + "*" * 50
_DASH = " " * 14 + "*" + "-" * 48 + "*"


def format_mp2_block(heg: HEGData, e2: float) -> str:
    out = [_STAR, " " * 14 + f"* E(SCF) is {'':12}{heg.EHF:18.12f} a.u.  *", _DASH,
           " " * 14 + "*                RHF Eigenvalues:                *", _DASH]
    for e in heg.Eigen[:heg.NOcc]:
        out.append(" " * 14 + "*" + " " * 12 + f"{e:24.16E}" + " " * 12 + "*")
    out.append(" " * 14 + "*" + " " * 12 + "-" * 24 + " " * 12 + "*")
    for e in heg.Eigen[heg.NOcc:]:
        out.append(" " * 14 + "*" + " " * 12 + f"{e:24.16E}" + " " * 12 + "*")
    out += [_STAR, " " * 14 + f"* E(2) is {'':14}{e2:18.12f} a.u.  *",
            " " * 14 + f"* E(SCF) + E(2) is {'':5}{e2 + heg.EHF:18.12f} a.u.  *", _STAR]
    return "\n".join(out) + "\n"


def format_ccd_block(heg: HEGData, info: UEGInfo, r: SolveResult) -> str:
    out = [" " * 14 + "*               CCD summary follows              *", _DASH]
    if info.DoRing:
        out.append(" " * 14 + "* Including ring diagrams...                     *")
    if info.DoXRing:
        out.append(" " * 14 + "* Including crossed ring diagrams...             *")
    if info.DoLadder:
        out.append(" " * 14 + "* Including ladder diagrams...                   *")
    if info.DoMosaic:
        out.append(" " * 14 + "* Including mosaic diagrams...                   *")
    out += [_DASH, " " * 14 + "*  CCD Energy      Iteration    Biggest change   *"]
    for k in range(r.niter):
        out.append(" " * 14 + f"* {r.trace_e[k]:15.10f}  {k + 1:5d}{'':7}{r.trace_dt[k]:14.10f}    *")
    out.append(_DASH)
    if r.failed:
        out.append(" " * 14 + "* Final CCD Energy is   DID NOT CONVERGE         *")
    else:
        out.append(" " * 14 + f"* CCD has converged in {r.niter:3d} iterations{'':12}*")
        out.append(" " * 14 + f"* Ec(CCD) is {'':10}{r.ecorr:18.12f} a.u.   *")
        out.append(" " * 14 + f"* Final CCD Energy is  {r.ecorr + heg.EHF:18.12f} a.u.   *")
    out.append(" " * 14 + f"* Max CCD Residual is  {'':4}{r.maxres:18.12E}    *")
    out.append(_STAR)
    return "\n".join(out) + "\n"


def format_drpa_block(heg: HEGData, info: UEGInfo, r: SolveResult, esosex: float) -> str:
    out = [" " * 14 + "*              dRPA summary follows              *", _DASH]
    if info.DoRing:
        out.append(" " * 14 + "* Including ring diagrams...                     *")
    if info.DoMosaic:
        out.append(" " * 14 + "* Including mosaic diagrams...                   *")
    out += [_DASH, " " * 14 + "* dRPA Energy      Iteration    Biggest change   *"]
    for k in range(r.niter):
        out.append(" " * 14 + f"* {r.trace_e[k]:15.10f}  {k + 1:5d}{'':7}{r.trace_dt[k]:14.10f}    *")
    out.append(_DASH)
    if r.failed:
        out.append(" " * 14 + "* Final dRPA Energy is   DID NOT CONVERGE        *")
    else:
        out.append(" " * 14 + f"* dRPA has converged in {r.niter:3d} iterations{'':11}*")
        out.append(" " * 14 + f"* Ec(dRPA) is {'':9}{r.ecorr:18.12f} a.u.   *")
        out.append(" " * 14 + f"* Final dRPA Energy is {r.ecorr + heg.EHF:18.12f} a.u.   *")
        out.append(" " * 14 + f"* Ec(SOSEX) is {'':8}{esosex:18.12f} a.u.   *")
        out.append(" " * 14 + f"* Final SOSEX Energy is{esosex + heg.EHF:18.12f} a.u.   *")
    out.append(" " * 14 + f"* Max CCD Residual is  {'':4}{r.maxres:18.12E}    *")
    out.append(_STAR)
    return "\n".join(out) + "\n"


def single_point(input_text: str, opts: L.Options | None = None):
    """Main's DoSingleCalc path (source/Main.f90:44-97): basis, MP2, then CCD
    or, with DodRPASOSEX, dRPA/SOSEX on T2aaaa (source/Main.f90:84-92); returns
    (Output text, HEGData, E(2), SolveResult)."""
    info = parse_input(input_text)
    heg = init_heg(info, device=0)
    ctx = CCDContext(heg, info, opts)
    taa, tab, tba = DrvMBPT(heg, info, ctx)
    e2 = heg.ECorr
    text = format_mp2_block(heg, e2)
    if info.DodRPASOSEX:
        r, esosex = DrvdRPA(heg, info, taa, ctx=ctx)
        text += format_drpa_block(heg, info, r, esosex)
    else:
        r = DrvCCD(heg, info, taa, tab, tba, ctx=ctx)
        text += format_ccd_block(heg, info, r)
    return text, heg, e2, r
