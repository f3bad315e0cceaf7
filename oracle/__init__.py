"""CPU oracle for the CCD path -- TEST INFRASTRUCTURE ONLY.

A ctypes wrapper over ``liboracle.so`` (built from ``uegccd_oracle.c`` by the
Makefile beside it), the loop-literal C restatement of the reference's
source/HEG.f90, source/MP2.f90, source/CCD.f90 and source/DIISCC.f90.

Only the tests, the smoke entry point and the CPU-baseline leg of bench.py
import this package.  The product (``uegccd_amd``) never does.

Parity status of each function (what pins it; see tests/test_oracle_*.py):
  * basis/eigenvalues/E_HF/Madelung  -- source/Test/fort.60 and Output
    eigenvalue listings, benchmarks/ZRun-dRPA/Output, closed-form Madelung.
  * MP2 amplitudes and E(2)          -- source/Test/Output E(2) lines for 11
    systems, benchmarks/ZRun-dRPA/Output:65, closed form.
  * CCD terms + damped Jacobi + DIIS -- source/Test/Output 70- and 84-iteration
    trajectories (every printed E and dT), benchmarks/ZRun-v6-N14 twist 1;
    brute-force invariants (T^ab(ij,a) = T^ab(ji,b), T^ba(ij,a) = -T^ab(ij,b)).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
# "serial": liboracle.so, the -O3 single-thread build (the checker and the
# one-core CPU baseline); "omp": liboracle_omp.so, the same code with the term
# loops over OpenMP threads (the all-cores baseline; the same results).
_VARIANTS = {"serial": "liboracle.so", "omp": "liboracle_omp.so"}

IRANGE_KEYS = ("Rings Range", "XRings Range", "Ladders Range", "Mosaics Range",
               "DriverDir Range", "DriverEx Range", "Energy Range", "LinRings Range",
               "QuadRings Range", "DRings Range", "ExRings Range", "LinLadd Range",
               "QuadLadd Range", "DLadders Range", "ExLadders Range")


class _Params(ctypes.Structure):
    _fields_ = [("nel", ctypes.c_int), ("maxkpoint", ctypes.c_int), ("coren", ctypes.c_int),
                ("safe_eri", ctypes.c_int), ("do_ring", ctypes.c_int),
                ("do_xring", ctypes.c_int), ("do_ladder", ctypes.c_int),
                ("do_mosaic", ctypes.c_int), ("irange", ctypes.c_int * 15),
                ("rs", ctypes.c_double), ("twist", ctypes.c_double * 3),
                ("fcut", ctypes.c_double), ("corefcut", ctypes.c_double),
                ("madelungfactor", ctypes.c_double), ("gap", ctypes.c_double),
                ("gapii", ctypes.c_double), ("exchangefactor", ctypes.c_double),
                ("cutoff_twisted", ctypes.c_int)]


@dataclass
class OracleParams:
    """The UEGInfo fields the path reads (source/Types.f90:29-82)."""
    nel: int
    maxkpoint: int
    rs: float
    twist: tuple = (0.0, 0.0, 0.0)
    fcut: float = 100.0
    corefcut: float = 1.0
    madelungfactor: float = 1.0
    gap: float = 0.0
    gapii: float = -0.01
    exchangefactor: float = 1.0
    coren: int = 0
    safe_eri: bool = True
    do_ring: bool = True
    do_xring: bool = True
    do_ladder: bool = True
    do_mosaic: bool = True
    irange: tuple = field(default_factory=lambda: (0,) * 15)
    cutoff_twisted: bool = False


def _lib(variant="serial"):
    if variant not in _LIBS:
        path = os.path.join(_HERE, _VARIANTS[variant])
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        lib = ctypes.CDLL(path)
        P, D, I = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        lib.orc_build.restype = P
        lib.orc_build.argtypes = [ctypes.POINTER(_Params)]
        lib.orc_free.argtypes = [P]
        for f in ("orc_nocc", "orc_nao", "orc_cell"):
            getattr(lib, f).restype = ctypes.c_int
            getattr(lib, f).argtypes = [P]
        lib.orc_scalars.argtypes = [P, D]
        lib.orc_basis.argtypes = [P, I, D]
        lib.orc_set_eigen.argtypes = [P, D]
        lib.orc_findindex.restype = ctypes.c_int
        lib.orc_findindex.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.orc_eri.restype = ctypes.c_double
        lib.orc_eri.argtypes = [P] + [ctypes.c_int] * 5
        lib.orc_mp2.restype = ctypes.c_double
        lib.orc_mp2.argtypes = [P, D, D, D]
        lib.orc_energy.restype = ctypes.c_double
        lib.orc_energy.argtypes = [P, D, D, D]
        lib.orc_terms.argtypes = [P, ctypes.c_int, D, D, D, D, D, D, D]
        lib.orc_ccd.restype = ctypes.c_int
        lib.orc_ccd.argtypes = [P, ctypes.c_double, ctypes.c_int, D, D, D, D, D, I, D, D]
        lib.orc_solvecc.argtypes = [P, D, D]
        lib.orc_drpa_energy.restype = ctypes.c_double
        lib.orc_sf.restype = ctypes.c_int
        lib.orc_sf.argtypes = [P, D, D, D, ctypes.c_int, ctypes.POINTER(ctypes.c_int), D, D]
        lib.orc_drpa_energy.argtypes = [P, D, ctypes.c_double]
        lib.orc_mp2_sf.restype = ctypes.c_int
        lib.orc_mp2_sf.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), D, D]
        lib.orc_exsf.restype = ctypes.c_int
        lib.orc_exsf.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int), D]
        lib.orc_drpa_terms.argtypes = [P, ctypes.c_int, D, D, D]
        lib.orc_drpa.restype = ctypes.c_int
        lib.orc_drpa.argtypes = [P, ctypes.c_double, ctypes.c_int, D, D, D, D, I, D, D]
        _LIBS[variant] = lib
    return _LIBS[variant]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleSystem:
    """One HEG system (HEGDataType) built by the oracle's restated init_HEG+change_rs."""

    def __init__(self, p: OracleParams, variant: str = "serial"):
        self.params = p
        cp = _Params()
        cp.nel, cp.maxkpoint, cp.coren = p.nel, p.maxkpoint, p.coren
        cp.safe_eri = int(p.safe_eri)
        cp.do_ring, cp.do_xring = int(p.do_ring), int(p.do_xring)
        cp.do_ladder, cp.do_mosaic = int(p.do_ladder), int(p.do_mosaic)
        for k in range(15):
            cp.irange[k] = int(p.irange[k])
        cp.rs = p.rs
        for d in range(3):
            cp.twist[d] = p.twist[d]
        cp.fcut, cp.corefcut = p.fcut, p.corefcut
        cp.madelungfactor, cp.gap, cp.gapii = p.madelungfactor, p.gap, p.gapii
        cp.exchangefactor = p.exchangefactor
        cp.cutoff_twisted = int(p.cutoff_twisted)
        self._lib = _lib(variant)
        self._h = self._lib.orc_build(ctypes.byref(cp))
        if not self._h:
            raise ValueError("oracle: basis construction failed")
        self.nocc = self._lib.orc_nocc(self._h)
        self.nao = self._lib.orc_nao(self._h)
        self.nvir = self.nao - self.nocc
        sc = np.zeros(9)
        self._lib.orc_scalars(self._h, _dp(sc))
        (self.L, self.omega, self.madelung, self.ehf, self.screening, self.rs,
         self.denomfactor, self.tolmax, self.failratio) = sc.tolist()
        self.nvec = np.zeros((self.nao, 3), dtype=np.int32)
        self.eigen = np.zeros(self.nao)
        self._lib.orc_basis(self._h, self.nvec.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            _dp(self.eigen))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.orc_free(h)
            self._h = None

    @property
    def shape(self):
        """Dense amplitude shape, Fortran order T(NOcc,NOcc,NOcc+1:NAO)."""
        return (self.nocc, self.nocc, self.nvir)

    def zeros(self):
        return np.zeros(self.shape, order="F")

    def set_eigen(self, eigen):
        e = np.ascontiguousarray(eigen, dtype=np.float64)
        self._lib.orc_set_eigen(self._h, _dp(e))
        self.eigen = e.copy()

    def findindex(self, m, n, l):
        return self._lib.orc_findindex(self._h, m, n, l)

    def eri(self, a, b, i, j, flag=0):
        return self._lib.orc_eri(self._h, a, b, i, j, flag)

    def mp2(self):
        taa, tab, tba = self.zeros(), self.zeros(), self.zeros()
        e2 = self._lib.orc_mp2(self._h, _dp(taa), _dp(tab), _dp(tba))
        return e2, taa, tab, tba

    def energy(self, taa, tab, tba):
        taa, tab, tba = (np.asfortranarray(x, dtype=np.float64) for x in (taa, tab, tba))
        return self._lib.orc_energy(self._h, _dp(taa), _dp(tab), _dp(tba))

    def structure_factor(self, tab, tba):
        """CCEnergyAndSF: (q [n,3] int, vals [n,5] = SfG, SfSum, SfV, SfDSum,
        SfXSum, energy), rows in the reference's print order."""
        tab = np.asfortranarray(tab, dtype=np.float64)
        tba = np.asfortranarray(tba, dtype=np.float64)
        taa = np.asfortranarray(tab + tba)
        e = ctypes.c_double(0.0)
        n = self._lib.orc_sf(self._h, _dp(taa), _dp(tab), _dp(tba), 0, None, None, ctypes.byref(e))
        q = np.zeros((max(n, 1), 3), dtype=np.int32)
        vals = np.zeros((max(n, 1), 5))
        self._lib.orc_sf(self._h, _dp(taa), _dp(tab), _dp(tba), n,
                         q.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(vals), ctypes.byref(e))
        return q[:n], vals[:n], e.value

    def mp2_structure_factor(self):
        """DrvMBPT's structure factor: (q [n,3], vals [n,3] = SfG, SfSum, SfV, E(2))."""
        e = ctypes.c_double(0.0)
        n = self._lib.orc_mp2_sf(self._h, 0, None, None, ctypes.byref(e))
        q = np.zeros((max(n, 1), 3), dtype=np.int32)
        vals = np.zeros((max(n, 1), 3))
        self._lib.orc_mp2_sf(self._h, n, q.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(vals),
                             ctypes.byref(e))
        return q[:n], vals[:n], e.value

    def exchange_structure_factor(self):
        """print_ex_SF at this system's r_s: (q [n,3], count [n], ExV [n])."""
        I = ctypes.POINTER(ctypes.c_int)
        n = self._lib.orc_exsf(self._h, 0, None, None, None)
        q = np.zeros((max(n, 1), 3), dtype=np.int32)
        cnt = np.zeros(max(n, 1), dtype=np.int32)
        v = np.zeros(max(n, 1))
        self._lib.orc_exsf(self._h, n, q.ctypes.data_as(I), cnt.ctypes.data_as(I), _dp(v))
        return q[:n], cnt[:n], v[:n]

    def terms(self, tab, tba, mask=31):
        """G and R for the given amplitudes (TermsWrapper)."""
        tab = np.asfortranarray(tab, dtype=np.float64)
        tba = np.asfortranarray(tba, dtype=np.float64)
        taa = np.asfortranarray(tab + tba)
        gab, gba, rab, rba = self.zeros(), self.zeros(), self.zeros(), self.zeros()
        self._lib.orc_terms(self._h, mask, _dp(taa), _dp(tab), _dp(tba),
                            _dp(gab), _dp(gba), _dp(rab), _dp(rba))
        return gab, gba, rab, rba

    def ccd(self, tab=None, tba=None, tolmax=0.0, maxiter=1000):
        """Reference-trajectory CCD from the MP2 guess (or given amplitudes)."""
        if tab is None:
            _, taa, tab, tba = self.mp2()
        else:
            tab = np.array(tab, dtype=np.float64, order="F")
            tba = np.array(tba, dtype=np.float64, order="F")
            taa = np.asfortranarray(tab + tba)
        te = np.zeros(max(maxiter, 1) + 1)
        tdt = np.zeros(max(maxiter, 1) + 1)
        ec = ctypes.c_double()
        mr = ctypes.c_double()
        fl = ctypes.c_int()
        niter = self._lib.orc_ccd(self._h, tolmax, maxiter, _dp(taa), _dp(tab), _dp(tba),
                                  ctypes.byref(ec), ctypes.byref(mr), ctypes.byref(fl),
                                  _dp(te), _dp(tdt))
        return dict(ecorr=ec.value, niter=niter, maxres=mr.value, failed=bool(fl.value),
                    trace_e=te[:niter].copy(), trace_dt=tdt[:niter].copy(),
                    tab=tab, tba=tba, taa=taa)

    def drpa_energy(self, t2, csosex=0.0):
        """dRPA (csosex 0) or SOSEX (csosex 1) energy of the spin-summed T2."""
        t2 = np.asfortranarray(t2, dtype=np.float64)
        return self._lib.orc_drpa_energy(self._h, _dp(t2), float(csosex))

    def drpa_terms(self, t2, mask=11):
        """G and R of the dRPA TermsWrapper for the spin-summed T2."""
        t2 = np.asfortranarray(t2, dtype=np.float64)
        g, r = self.zeros(), self.zeros()
        self._lib.orc_drpa_terms(self._h, mask, _dp(t2), _dp(g), _dp(r))
        return g, r

    def drpa(self, t2=None, tolmax=0.0, maxiter=1000):
        """DrvdRPA from the MP2 T2aaaa (or the given spin-summed T2)."""
        if t2 is None:
            _, t2, _, _ = self.mp2()
        t2 = np.array(t2, dtype=np.float64, order="F")
        te = np.zeros(max(maxiter, 1) + 1)
        tdt = np.zeros(max(maxiter, 1) + 1)
        ec, es, mr = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        fl = ctypes.c_int()
        niter = self._lib.orc_drpa(self._h, tolmax, maxiter, _dp(t2), ctypes.byref(ec),
                                   ctypes.byref(es), ctypes.byref(mr), ctypes.byref(fl),
                                   _dp(te), _dp(tdt))
        return dict(ecorr=ec.value, esosex=es.value, niter=niter, maxres=mr.value,
                    failed=bool(fl.value), trace_e=te[:niter].copy(),
                    trace_dt=tdt[:niter].copy(), t2=t2)
