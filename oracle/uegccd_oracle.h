/*
 * uegccd_oracle.h -- CPU restatement of the reference's CCD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP product is compared
 * against; it is never linked into, loaded by or called from the product
 * library (uegccd_amd/csrc) or the ZCode host.  Only tests/, the smoke() entry
 * and bench.py's CPU-baseline leg may load it.
 *
 * Everything here follows the reference loop for loop, with 1-based orbital
 * indices exactly as in the Fortran (occupied 1..NOcc, virtual NOcc+1..NAO).
 * Amplitude arrays are dense, column-major T(NOcc,NOcc,NOcc+1:NAO):
 *     T(i,j,a) -> (i-1) + o*(j-1) + o*o*(a-o-1)
 */
#ifndef UEGCCD_ORACLE_H
#define UEGCCD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the fields of UEGInfoType that the path reads (source/Types.f90:29-82). */
typedef struct {
    int    nel;            /* # Electrons                                    */
    int    maxkpoint;      /* Momentum Cutoff (integer |n|^2 limit)           */
    int    coren;          /* int(N/2*CoreFactor) (source/IO.f90:232)         */
    int    safe_eri;       /* Safe ERI                                        */
    int    do_ring, do_xring, do_ladder, do_mosaic;
    int    irange[15];     /* Rings, XRings, Ladders, Mosaics, DriverDir,
                              DriverEx, Energy, LinRings, QuadRings, DRings,
                              ExRings, LinLadd, QuadLadd, DLadders, ExLadders */
    double rs;             /* rSMin                                           */
    double twist[3];       /* NTwist                                          */
    double fcut, corefcut;
    double madelungfactor, gap, gapii, exchangefactor;
    int    cutoff_twisted; /* 0: |n|^2 <= maxkpoint (the reference's cutoff);
                              1: |n+twist|^2 <= maxkpoint (opt-in)            */
} orc_params;

/* Opaque system (HEGDataType + the ERI pointer choice). */
typedef struct orc_system orc_system;

orc_system* orc_build(const orc_params* p);        /* init_HEG + change_rs */
void   orc_free(orc_system* s);
int    orc_nocc(const orc_system* s);
int    orc_nao(const orc_system* s);
int    orc_cell(const orc_system* s);
/* scalars: L, Omega, Madelung, EHF, ScreeningDistance, rS, DenomFactor, TolMax, FailRatio */
void   orc_scalars(const orc_system* s, double* out9);
void   orc_basis(const orc_system* s, int* nvec /* [nao*3], row p = (n1,n2,n3) */,
                 double* eigen /* [nao] */);
/* Replace eigenvalues (the twist-averaging driver does this, source/Main.f90:243). */
void   orc_set_eigen(orc_system* s, const double* eigen);
/* Enumerate the basis after an arbitrary twist change (source/HEG.f90:551-578). */
int    orc_findindex(const orc_system* s, int m, int n, int l);
double orc_eri(const orc_system* s, int a, int b, int i, int j, int flag);

/* DrvMBPT amplitudes and E(2) (source/MP2.f90:56-98). */
double orc_mp2(const orc_system* s, double* taaaa, double* tabab, double* tabba);

/* CCEnergy (source/CCD.f90:186-237). */
int    orc_sf(const orc_system* s, const double* taaaa, const double* tabab, const double* tabba,
              int cap, int* q, double* vals, double* ecorr);
/* DrvMBPT's structure factor (source/MP2.f90:56-117): rows (SfG, SfSum, SfV). */
int    orc_mp2_sf(const orc_system* s, int cap, int* q, double* vals, double* e2);
/* print_ex_SF (source/HEG.f90:617-674) at the current r_s: count and ExV per cell. */
int    orc_exsf(const orc_system* s, int cap, int* q, int* count, double* exv);
double orc_energy(const orc_system* s, const double* taaaa, const double* tabab,
                  const double* tabba);

/* TermsWrapper (source/CCD.f90:146-181): G from drivers+mosaics+ladders+rings+
 * xrings, then R = D*T - G.  term_mask bit0 drivers, bit1 mosaics, bit2
 * ladders, bit3 rings, bit4 xrings; each further gated by the Do* flags. */
void   orc_terms(const orc_system* s, int term_mask,
                 const double* taaaa, const double* tabab, const double* tabba,
                 double* gabab, double* gabba, double* rabab, double* rabba);

/* DrvCCD (source/CCD.f90:18-183) with the reference's DIIS (source/DIISCC.f90).
 * tolmax <= 0 uses FindTol's value.  trace_e/trace_dt (may be NULL) receive
 * the printed (E, dT) per iteration, index NIter-1, up to maxiter entries.
 * Returns NIter; *ecorr, *maxres, *failed are the summary values. */
int    orc_ccd(const orc_system* s, double tolmax, int maxiter,
               double* taaaa, double* tabab, double* tabba,
               double* ecorr, double* maxres, int* failed,
               double* trace_e, double* trace_dt);

/* One damped-Jacobi step as SolveCC (source/CCD.f90:354-386), in place. */
void   orc_solvecc(const orc_system* s, double* t2, const double* g2);

/* dRPA / rCCD on the spin-summed T2 (source/dRPA.f90).  Energy with the
 * SOSEX coefficient csosex (0: dRPA, 1: SOSEX; source/dRPA.f90:181-216). */
double orc_drpa_energy(const orc_system* s, const double* t2, double csosex);
/* The dRPA TermsWrapper (source/dRPA.f90:146-175): G and R = D*T - G (r2 may
 * be NULL); term_mask bit0 drivers, bit1 mosaics, bit3 rings. */
void   orc_drpa_terms(const orc_system* s, int term_mask, const double* t2, double* g2,
                      double* r2);
/* DrvdRPA (source/dRPA.f90:23-143): as orc_ccd, from t2 in place; *ecorr is
 * Ec(dRPA), *esosex Ec(SOSEX) of the final amplitudes. */
int    orc_drpa(const orc_system* s, double tolmax, int maxiter, double* t2, double* ecorr,
                double* esosex, double* maxres, int* failed, double* trace_e, double* trace_dt);

#ifdef __cplusplus
}
#endif
#endif
