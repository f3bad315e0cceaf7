/*
 * uegccd_oracle.c -- loop-literal CPU restatement of the reference's CCD path
 * (TEST INFRASTRUCTURE ONLY: the checker and the CPU baseline; the product
 * never links or loads this file).
 *
 * Each routine names the reference lines it restates.  Orbital indices are
 * 1-based exactly as in the Fortran; arrays are laid out column-major.
 */
#include "uegccd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323846264338327950288

/* The all-cores CPU baseline (liboracle_omp.so, built with -DORC_OMP
 * -fopenmp) splits the outer occupied loop of the term routines over
 * threads.  Every element (i,j,a) is written only by the iteration of its own
 * i, in the same order as the serial loop, so the results are the same bits. */
#ifdef ORC_OMP
#define ORC_PAR_I _Pragma("omp parallel for schedule(dynamic)")
#else
#define ORC_PAR_I
#endif
enum { ERI_CASED, ERI_NOCASE, ERI_CORE, ERI_GAP, ERI_BASE };
/* irange[] slots, in Input-key order (source/IO.f90:88-102) */
enum { R_RING, R_XRING, R_LADDER, R_MOSAIC, R_DRVDIR, R_DRVEX, R_ENERGY, R_LINRING,
       R_QUADRING, R_DRING, R_EXRING, R_LINLADD, R_QUADLADD, R_DLADD, R_EXLADD };

struct orc_system {
    orc_params p;
    int nel, nocc, nao, cell, maxk;
    double L, omega, madelung, density, kf, screening, rs, ehf, xhf;
    double denomfactor, tolmax, failratio;
    int *nvec;      /* nvec(p,d) at [(p-1)*3+d] */
    double *kvec;   /* kvec(p,d) at [(p-1)*3+d] */
    double *eigen;  /* eigen(p) at [p-1] */
    int *kp2bf;     /* kPointToBasisFn(-cell:cell,...) */
    int eri_kind;
};

static inline int kidx(const orc_system* s, int x, int y, int z) {
    int w = 2 * s->cell + 1;
    return (x + s->cell) + w * ((y + s->cell) + w * (z + s->cell));
}

/* CalcMadelung, source/HEG.f90:89-148 (Ewald sum, Fraser et al.). */
static double calc_madelung(double omega, double L) {
    double kappa = 2.8 / pow(omega, 1.0 / 3.0);
    double term2 = -PI / (pow(kappa, 2.0) * pow(L, 3.0));
    double term4 = -2.0 * kappa / sqrt(PI);
    double recipsum2 = 0.0, realsum2 = 0.0;
    for (int l4 = 1; l4 <= 10; ++l4) {
        recipsum2 = 0.0;
        for (int l1 = -l4; l1 <= l4; ++l1)
            for (int l2 = -l4; l2 <= l4; ++l2)
                for (int l3 = -l4; l3 <= l4; ++l3) {
                    int n2 = l1 * l1 + l2 * l2 + l3 * l3;
                    double k2 = (1.0 / pow(L, 2.0)) * (double)(l1 * l1 + l2 * l2 + l3 * l3);
                    double ek2 = (1.0 / pow(L, 3.0)) * (1.0 / (PI * k2)) *
                                 exp(-pow(PI, 2.0) * k2 / pow(kappa, 2.0));
                    if (n2 != 0) recipsum2 += ek2;
                }
    }
    for (int l4 = 1; l4 <= 10; ++l4) {
        realsum2 = 0.0;
        for (int l1 = -l4; l1 <= l4; ++l1)
            for (int l2 = -l4; l2 <= l4; ++l2)
                for (int l3 = -l4; l3 <= l4; ++l3) {
                    double modr = L * sqrt((double)(l1 * l1 + l2 * l2 + l3 * l3));
                    if (modr != 0.0) realsum2 += erfc(kappa * modr) / modr;
                }
    }
    return realsum2 + recipsum2 + term2 + term4;
}

/* FcutToBasisFn, source/HEG.f90:326-342. */
static int fcut_to_basis(double f, int nel) {
    int m = (int)(pow(f, 1.5) * nel / 2.0);
    double mt = pow(f, 1.5) * nel / 2.0;
    if (fabs(m - mt) > fabs(m + 1 - mt)) m = m + 1;
    return m;
}

/* Eigenvalues, source/HEG.f90:291-304 and 514-529. */
static void form_eigen(orc_system* s) {
    for (int l1 = 1; l1 <= s->nao; ++l1) {
        const double* k1 = &s->kvec[(l1 - 1) * 3];
        double e = 0.5 * (k1[0] * k1[0] + k1[1] * k1[1] + k1[2] * k1[2]);
        for (int l2 = 1; l2 <= s->nel / 2; ++l2) {
            if (l2 != l1) {
                const double* k2 = &s->kvec[(l2 - 1) * 3];
                double q0 = k1[0] - k2[0], q1 = k1[1] - k2[1], q2 = k1[2] - k2[2];
                e = e - (1.0 / s->omega) * (4.0 * PI / (q0 * q0 + q1 * q1 + q2 * q2)) *
                            s->p.exchangefactor;
            } else {
                e = e + s->madelung - s->p.gap;
            }
        }
        s->eigen[l1 - 1] = e;
    }
}

/* print_HF, source/HEG.f90:581-615 (E_HF including +Madelung*N/2). */
static void form_ehf(orc_system* s) {
    double ehf = 0.0, xhf = 0.0;
    for (int l1 = 1; l1 <= s->nel / 2; ++l1) {
        const double* k1 = &s->kvec[(l1 - 1) * 3];
        ehf = ehf + 0.5 * (k1[0] * k1[0] + k1[1] * k1[1] + k1[2] * k1[2]);
        for (int l2 = 1; l2 <= s->nel / 2; ++l2) {
            if (l2 != l1) {
                const double* k2 = &s->kvec[(l2 - 1) * 3];
                double q0 = k1[0] - k2[0], q1 = k1[1] - k2[1], q2 = k1[2] - k2[2];
                double t = 0.5 * (1.0 / s->omega) * (4.0 * PI / (q0 * q0 + q1 * q1 + q2 * q2));
                ehf = ehf - t;
                xhf = xhf - t;
            }
        }
    }
    ehf = ehf * 2.0;
    xhf = xhf * 2.0;
    s->ehf = ehf + s->madelung * 0.5 * s->nel;
    s->xhf = xhf;
}

typedef struct { int n[3]; double nt[3]; double nt2; } g1_t;

/* SetupBasis, source/HEG.f90:150-324, at the current (L, twist). */
static int setup_basis(orc_system* s) {
    int cell = s->cell, w = 2 * cell + 1;
    g1_t* g1 = (g1_t*)malloc(sizeof(g1_t) * (size_t)w * w * w);
    int M = 0;
    for (int l1 = -cell; l1 <= cell; ++l1)
        for (int l2 = -cell; l2 <= cell; ++l2)
            for (int l3 = -cell; l3 <= cell; ++l3) {
                g1_t g;
                g.n[0] = l1; g.n[1] = l2; g.n[2] = l3;
                g.nt[0] = l1 + s->p.twist[0];
                g.nt[1] = l2 + s->p.twist[1];
                g.nt[2] = l3 + s->p.twist[2];
                g.nt2 = g.nt[0] * g.nt[0] + g.nt[1] * g.nt[1] + g.nt[2] * g.nt[2];
                /* source/HEG.f90:190-194 cuts on the untwisted |n|^2; the
                   opt-in twisted sphere cuts on |n+twist|^2 instead */
                const int in = s->p.cutoff_twisted ? (g.nt2 <= (double)s->maxk)
                                                   : (l1 * l1 + l2 * l2 + l3 * l3 <= s->maxk);
                if (in) g1[M++] = g;
            }
    int nbasis = M;
    /* stable ascending sort by twisted |n+t|^2: the bubble sort at
       source/HEG.f90:206-217 swaps only on strict '<', so ties keep the
       enumeration order.  Insertion sort gives the identical permutation. */
    for (int i = 1; i < nbasis; ++i) {
        g1_t x = g1[i];
        int j = i - 1;
        while (j >= 0 && x.nt2 < g1[j].nt2) { g1[j + 1] = g1[j]; --j; }
        g1[j + 1] = x;
    }
    int mfcut = fcut_to_basis(s->p.fcut, s->nel);
    if (mfcut < nbasis) nbasis = mfcut;
    int mcore = fcut_to_basis(s->p.corefcut, s->nel);
    if ((s->p.corefcut - 1.0) < 1e-6 && mcore != s->nel / 2) {
        fprintf(stderr, "oracle: corefcut is less than 1 rather than equal to 1\n");
        free(g1);
        return -1;
    }
    nbasis = nbasis - mcore + s->nel / 2;
    s->nao = nbasis;
    s->kvec = (double*)calloc((size_t)nbasis * 3, sizeof(double));
    s->eigen = (double*)calloc((size_t)nbasis, sizeof(double));
    s->nvec = (int*)calloc((size_t)nbasis * 3, sizeof(int));
    s->kp2bf = (int*)calloc((size_t)w * w * w, sizeof(int));
    for (int l1 = 1; l1 <= nbasis; ++l1) {
        int off = (l1 > s->nel / 2) ? mcore - s->nel / 2 : 0;
        const g1_t* g = &g1[l1 + off - 1];
        for (int d = 0; d < 3; ++d) {
            s->kvec[(l1 - 1) * 3 + d] = g->nt[d] * (2.0 * PI / s->L);
            s->nvec[(l1 - 1) * 3 + d] = g->n[d];
        }
    }
    form_eigen(s);
    for (int l1 = 1; l1 <= nbasis; ++l1) {
        int off = (l1 > s->nel / 2) ? mcore - s->nel / 2 : 0;
        const g1_t* g = &g1[l1 + off - 1];
        s->kp2bf[kidx(s, g->n[0], g->n[1], g->n[2])] = l1;
    }
    form_ehf(s);
    free(g1);
    return 0;
}

/* change_rs, source/HEG.f90:494-533. */
static void change_rs(orc_system* s, double rs_new) {
    double rescale = rs_new / s->rs;
    s->L = s->L * rescale;
    s->omega = s->omega * pow(rescale, 3.0);
    for (int i = 0; i < s->nao * 3; ++i) s->kvec[i] = s->kvec[i] * (1.0 / rescale);
    s->madelung = calc_madelung(s->omega, s->L) * s->p.madelungfactor;
    s->density = s->nel / pow(s->L, 3.0);
    s->kf = pow(3.0 * pow(PI, 2.0) * s->density, 1.0 / 3.0);
    s->screening = pow(4.0 / PI * s->kf, 0.5);
    form_eigen(s);
    form_ehf(s);
    s->rs = rs_new;
}

/* FindTol, source/HEG.f90:1073-1086 (DenomFactor starts at 2, source/CCD.f90:11). */
static void find_tol(orc_system* s) {
    double rs = s->rs;
    s->tolmax = 1.0e-8;
    s->failratio = 100;
    s->denomfactor = 2.0;
    if (rs > 5.1) s->tolmax = 1.0e-6;
    if (rs > 5.1) s->denomfactor = 5.0;
    if (rs > 9.9) s->tolmax = 1.0e-5;
    if (rs > 80.0) s->tolmax = 1.0e-3;
    if (rs < 0.95) s->failratio = 1e6;
}

/* SetPointers, source/Main.f90:274-318. */
static int select_eri(const orc_params* p) {
    int sum = 0;
    for (int i = 0; i < 15; ++i) sum += abs(p->irange[i]);
    if ((sum > 0 && p->gapii > 0.0 && p->coren > 0) || p->safe_eri) return ERI_CASED;
    if (p->gapii > 0.0 && p->coren > 0) return ERI_NOCASE;
    if (p->coren > 0) return ERI_CORE;
    if (p->gapii > 0.0) return ERI_GAP;
    return ERI_BASE;
}

orc_system* orc_build(const orc_params* p) {
    orc_system* s = (orc_system*)calloc(1, sizeof(orc_system));
    s->p = *p;
    s->nel = p->nel;
    s->nocc = p->nel / 2;
    s->maxk = p->maxkpoint;
    /* init_HEG, source/HEG.f90:16-43 (at r_s = 1) */
    s->rs = 1.0;
    s->L = s->rs * pow(4.0 / 3.0 * PI * s->nel, 1.0 / 3.0);
    s->omega = pow(s->L, 3);
    s->madelung = calc_madelung(s->omega, s->L) * p->madelungfactor;
    s->cell = (int)sqrt((double)p->maxkpoint) + 3;
    s->density = s->nel / pow(s->L, 3.0);
    s->kf = pow(3.0 * pow(PI, 2.0) * s->density, 1.0 / 3.0);
    s->screening = pow(4.0 / PI * s->kf, 0.5);
    if (setup_basis(s) != 0) { orc_free(s); return NULL; }
    s->eri_kind = select_eri(p);
    /* source/Main.f90:78-79 */
    change_rs(s, p->rs);
    find_tol(s);
    return s;
}

void orc_free(orc_system* s) {
    if (!s) return;
    free(s->nvec); free(s->kvec); free(s->eigen); free(s->kp2bf);
    free(s);
}

int orc_nocc(const orc_system* s) { return s->nocc; }
int orc_nao(const orc_system* s) { return s->nao; }
int orc_cell(const orc_system* s) { return s->cell; }

void orc_scalars(const orc_system* s, double* o) {
    o[0] = s->L; o[1] = s->omega; o[2] = s->madelung; o[3] = s->ehf; o[4] = s->screening;
    o[5] = s->rs; o[6] = s->denomfactor; o[7] = s->tolmax; o[8] = s->failratio;
}

void orc_basis(const orc_system* s, int* nvec, double* eigen) {
    if (nvec) memcpy(nvec, s->nvec, sizeof(int) * (size_t)s->nao * 3);
    if (eigen) memcpy(eigen, s->eigen, sizeof(double) * (size_t)s->nao);
}

void orc_set_eigen(orc_system* s, const double* eigen) {
    memcpy(s->eigen, eigen, sizeof(double) * (size_t)s->nao);
}

/* FindIndex, source/HEG.f90:357-404. */
static inline int find_index(const orc_system* s, int m, int n, int l) {
    const int* nm = &s->nvec[(m - 1) * 3];
    const int* nn = &s->nvec[(n - 1) * 3];
    const int* nl = &s->nvec[(l - 1) * 3];
    int k0 = nm[0] + nn[0] - nl[0], k1 = nm[1] + nn[1] - nl[1], k2 = nm[2] + nn[2] - nl[2];
    if (abs(k0) > s->cell || abs(k1) > s->cell || abs(k2) > s->cell) return 0;
    /* the reference's early exit outside its cutoff sphere |n|^2 <= MaxKPoint;
       with the opt-in twisted sphere the same test on |n + twist|^2 */
    if (s->p.cutoff_twisted) {
        const double x = k0 + s->p.twist[0], y = k1 + s->p.twist[1], z = k2 + s->p.twist[2];
        if (x * x + y * y + z * z > (double)s->maxk) return 0;
    } else if (k0 * k0 + k1 * k1 + k2 * k2 > s->maxk) {
        return 0;
    }
    int k = s->kp2bf[kidx(s, k0, k1, k2)];
    if (k > s->nao) return 0;
    return k;
}
int orc_findindex(const orc_system* s, int m, int n, int l) { return find_index(s, m, n, l); }

/* CheckMomentumSymmetry, source/HEG.f90:438-450. */
static inline int momentum_ok(const orc_system* s, int i, int j, int a, int b) {
    const int* ni = &s->nvec[(i - 1) * 3];
    const int* nj = &s->nvec[(j - 1) * 3];
    const int* na = &s->nvec[(a - 1) * 3];
    const int* nb = &s->nvec[(b - 1) * 3];
    for (int d = 0; d < 3; ++d)
        if (-ni[d] + na[d] != -(-nj[d] + nb[d])) return 0;
    return 1;
}

/* The five ERI variants, source/HEG.f90:689-978, argument order (A,B,I,J).
 * Only CasedERI reads the range flag. */
static inline double eri(const orc_system* s, int a, int b, int i, int j, int flag) {
    const int kind = s->eri_kind;
    double v;
    if (kind != ERI_BASE) {
        if (a <= 0 || b <= 0 || i <= 0 || j <= 0) return -100.0;
        if (!momentum_ok(s, i, j, a, b)) return -100.0;
    }
    const int* ni = &s->nvec[(i - 1) * 3];
    const int* na = &s->nvec[(a - 1) * 3];
    int q0 = -ni[0] + na[0], q1 = -ni[1] + na[1], q2 = -ni[2] + na[2];
    if (kind != ERI_CASED) flag = 0;
    if (q0 == 0 && q1 == 0 && q2 == 0) {
        switch (flag) {
            case 0: v = -s->madelung; break;
            case 1: v = (1.0 / pow(s->L, 3.0)) * 4.0 * PI / pow(s->screening, 2.0); break;
            case 2: v = -s->madelung - (1.0 / pow(s->L, 3.0)) * 4.0 * PI / pow(s->screening, 2.0); break;
            case 3: v = 0.0; break;
            default: return -100.0;
        }
    } else {
        double qq = (double)(q0 * q0 + q1 * q1 + q2 * q2);
        double qs = qq * pow(2.0 * PI / s->L, 2.0);
        switch (flag) {
            case 0: v = 1.0 / (s->L * PI * qq); break;
            case 1: v = (1.0 / pow(s->L, 3.0)) * 4.0 * PI / (pow(s->screening, 2.0) + qs); break;
            case 2: v = (1.0 / pow(s->L, 3.0)) * 4.0 * PI / (pow(s->screening, 2.0) + qs) *
                        pow(s->screening, 2.0) / qs; break;
            case 3: v = 0.0; break;
            default: v = 0.0; break;  /* unset in the reference */
        }
    }
    if (kind == ERI_BASE) return v;
    const int cn = s->p.coren;
    const int incore = (a <= cn) || (b <= cn) || (i <= cn) || (j <= cn);
    if (kind == ERI_CORE) {
        if (v < 0.0) return -100.0;
        if (incore) v = 0.0;
        return v;
    }
    if (kind == ERI_GAP) {
        if (v < 0.0) return -100.0;
        if (v > s->p.gapii && s->p.gapii > 0.0) v = 0.0;
        return v;
    }
    /* CASED and NOCASE */
    if (incore) v = 0.0;
    if (v < 0.0) return -100.0;
    if (v > s->p.gapii && s->p.gapii > 0.0) v = 0.0;
    return v;
}
double orc_eri(const orc_system* s, int a, int b, int i, int j, int flag) {
    return eri(s, a, b, i, j, flag);
}

static inline int imax3(int a, int b, int c) { int m = a > b ? a : b; return m > c ? m : c; }

#define IDX(i, j, a) ((size_t)((i) - 1) + (size_t)o * ((size_t)((j) - 1) + (size_t)o * (size_t)((a) - o - 1)))

/* DrvMBPT amplitudes + E(2), source/MP2.f90:51-98. */
double orc_mp2(const orc_system* s, double* taaaa, double* tabab, double* tabba) {
    const int o = s->nocc, n = s->nao;
    size_t sz = (size_t)o * o * (n - o);
    memset(tabab, 0, sz * sizeof(double));
    memset(tabba, 0, sz * sizeof(double));
    memset(taaaa, 0, sz * sizeof(double));
    double ecorr = 0.0;
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                double vab = eri(s, i, j, a, b, 0);
                double vba = eri(s, i, j, b, a, 0);
                if (fmin(vab, vba) < -80.0) continue;
                double den = s->eigen[i - 1] + s->eigen[j - 1] - s->eigen[a - 1] - s->eigen[b - 1];
                size_t x = IDX(i, j, a);
                tabab[x] = vab / den;
                tabba[x] = -vba / den;
                taaaa[x] = tabab[x] + tabba[x];
                ecorr = ecorr + 0.5 * taaaa[x] * (vab - vba) + 0.5 * tabab[x] * (vab) -
                        0.5 * tabba[x] * (vba);
            }
    return ecorr;
}

/* CCEnergy, source/CCD.f90:186-237. */
double orc_energy(const orc_system* s, const double* taaaa, const double* tabab,
                  const double* tabba) {
    const int o = s->nocc, n = s->nao;
    const int fl = s->p.irange[R_ENERGY];
    double ecorr = 0.0;
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                double vab = eri(s, a, b, i, j, fl);
                double vba = eri(s, a, b, j, i, fl);
                if (fmin(vab, vba) < -80.0) continue;
                size_t x = IDX(i, j, a);
                double de = vab * (taaaa[x] + tabab[x]) - vba * (taaaa[x] + tabba[x]);
                ecorr = ecorr + 0.5 * de;
            }
    return ecorr;
}

/* CCEnergyAndSF, source/CCD.f90:240-351 (with dK and dK2scaled,
   source/HEG.f90:1012-1054, and SendForAllocation, 1056-1070): the energy
   and the five grids over q = n_a - n_i (direct) and n_a - n_j (exchange) of
   the valid elements, written in the reference's print order (q_x slowest)
   for the cells with SfG > 1e-10.  Rows k < cap go to q[3k..] and
   vals[5k..] = (SfG, SfSum, SfV, SfDSum, SfXSum); returns the row count. */
int orc_sf(const orc_system* s, const double* taaaa, const double* tabab, const double* tabba,
           int cap, int* q, double* vals, double* ecorr) {
    const int o = s->nocc, n = s->nao;
    const int fl = s->p.irange[R_ENERGY];
    int mx = s->nvec[0];
    for (int k = 0; k < 3 * n; ++k) mx = s->nvec[k] > mx ? s->nvec[k] : mx;
    const int len = mx * 2 + 1, w = 2 * len + 1;
    const size_t cells = (size_t)w * w * w;
    double* g[5];
    for (int k = 0; k < 5; ++k) g[k] = (double*)calloc(cells, sizeof(double));
    enum { SUM, G, V, DSUM, XSUM };
    const double sc = pow(2.0 * 3.14159265358979323846 / s->L, 2.0);
    double e = 0.0;
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                double vab = eri(s, a, b, i, j, fl);
                double vba = eri(s, a, b, j, i, fl);
                if (fmin(vab, vba) < -80.0) continue;
                size_t x = IDX(i, j, a);
                double de = vab * (taaaa[x] + tabab[x]) - vba * (taaaa[x] + tabba[x]);
                e = e + 0.5 * de;
                for (int ex = 0; ex < 2; ++ex) {
                    const int r = ex ? j : i;  /* dK(a,b,i,j) or dK(a,b,j,i) */
                    int qv[3], q2 = 0;
                    for (int d = 0; d < 3; ++d) {
                        qv[d] = s->nvec[(a - 1) * 3 + d] - s->nvec[(r - 1) * 3 + d];
                        q2 += qv[d] * qv[d];
                    }
                    const size_t c = ((size_t)(qv[0] + len) * w + (qv[1] + len)) * w + (qv[2] + len);
                    if (!ex) {
                        g[SUM][c] += taaaa[x] + tabab[x];
                        g[DSUM][c] += 2.0 * tabab[x];
                        g[XSUM][c] += taaaa[x] - tabab[x];
                        g[V][c] = vab;
                    } else {
                        g[SUM][c] -= taaaa[x] + tabba[x];
                        g[DSUM][c] -= 2.0 * tabba[x];
                        g[XSUM][c] -= taaaa[x] - tabba[x];
                        g[V][c] = vba;
                    }
                    g[G][c] = sqrt(q2 * sc);
                }
            }
    int rows = 0;
    for (size_t c = 0; c < cells; ++c) {
        if (!(g[G][c] > 1.0e-10)) continue;
        if (rows < cap) {
            q[3 * rows] = (int)(c / ((size_t)w * w)) - len;
            q[3 * rows + 1] = (int)((c / w) % w) - len;
            q[3 * rows + 2] = (int)(c % w) - len;
            double* r = vals + 5 * (size_t)rows;
            r[0] = g[G][c]; r[1] = g[SUM][c]; r[2] = g[V][c]; r[3] = g[DSUM][c]; r[4] = g[XSUM][c];
        }
        ++rows;
    }
    for (int k = 0; k < 5; ++k) free(g[k]);
    if (ecorr) *ecorr = e;
    return rows;
}

/* DrvMBPT's structure factor, source/MP2.f90:56-117: the MP2 amplitudes of
   each valid (i,j,a) binned at q = n_a - n_i (+(T^aa + T^ab)) and at
   q = n_a - n_j (-(T^aa + T^ba)), SfG = |q| 2 pi / L and SfV the integral of
   the cell's last writer; the cells with SfG > 0 in print order, vals[3k..] =
   (SfG, SfSum, SfV).  Returns the row count; *e2 the energy. */
int orc_mp2_sf(const orc_system* s, int cap, int* q, double* vals, double* e2) {
    const int o = s->nocc, n = s->nao;
    int mx = s->nvec[0];
    for (int k = 0; k < 3 * n; ++k) mx = s->nvec[k] > mx ? s->nvec[k] : mx;
    const int len = mx * 2 + 1, w = 2 * len + 1;
    const size_t cells = (size_t)w * w * w;
    double* sum = (double*)calloc(cells, sizeof(double));
    double* sg = (double*)calloc(cells, sizeof(double));
    double* sv = (double*)calloc(cells, sizeof(double));
    const double sc = pow(2.0 * PI / s->L, 2.0);
    double e = 0.0;
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                double vab = eri(s, i, j, a, b, 0);
                double vba = eri(s, i, j, b, a, 0);
                if (fmin(vab, vba) < -80.0) continue;
                double den = s->eigen[i - 1] + s->eigen[j - 1] - s->eigen[a - 1] - s->eigen[b - 1];
                double tab = vab / den, tba = -vba / den, taa = tab + tba;
                e = e + 0.5 * taa * (vab - vba) + 0.5 * tab * vab - 0.5 * tba * vba;
                for (int ex = 0; ex < 2; ++ex) {
                    const int r = ex ? j : i;
                    int qv[3], q2 = 0;
                    for (int d = 0; d < 3; ++d) {
                        qv[d] = s->nvec[(a - 1) * 3 + d] - s->nvec[(r - 1) * 3 + d];
                        q2 += qv[d] * qv[d];
                    }
                    const size_t c = ((size_t)(qv[0] + len) * w + (qv[1] + len)) * w + (qv[2] + len);
                    if (!ex) { sum[c] += taa + tab; sv[c] = vab; }
                    else { sum[c] -= taa + tba; sv[c] = vba; }
                    sg[c] = sqrt(q2 * sc);
                }
            }
    int rows = 0;
    for (size_t c = 0; c < cells; ++c) {
        if (!(sg[c] > 0.0)) continue;
        if (rows < cap) {
            q[3 * rows] = (int)(c / ((size_t)w * w)) - len;
            q[3 * rows + 1] = (int)((c / w) % w) - len;
            q[3 * rows + 2] = (int)(c % w) - len;
            vals[3 * rows] = sg[c]; vals[3 * rows + 1] = sum[c]; vals[3 * rows + 2] = sv[c];
        }
        ++rows;
    }
    free(sum); free(sg); free(sv);
    if (e2) *e2 = e;
    return rows;
}

/* print_ex_SF, source/HEG.f90:617-674, at the system's current r_s: over
   ordered pairs of distinct occupied orbitals, the count at q = n_l1 - n_l2
   and ExV = (1/Omega) 4 pi / |k_l1 - k_l2|^2 (the last pair's), cells with a
   count in print order.  Returns the row count. */
int orc_exsf(const orc_system* s, int cap, int* q, int* count, double* exv) {
    const int no = s->nel / 2;
    int qmax[3] = {0, 0, 0};
    for (int l1 = 1; l1 <= no; ++l1)
        for (int l2 = 1; l2 <= no; ++l2) {
            if (l2 == l1) continue;
            for (int d = 0; d < 3; ++d) {
                int t = abs(s->nvec[(l1 - 1) * 3 + d] - s->nvec[(l2 - 1) * 3 + d]);
                if (t > qmax[d]) qmax[d] = t;
            }
        }
    int maxq = qmax[0];
    if (qmax[1] > maxq) maxq = qmax[1];
    if (qmax[2] > maxq) maxq = qmax[2];
    maxq += 1;
    const int w = 2 * maxq + 1;
    int* cnt = (int*)calloc((size_t)w * w * w, sizeof(int));
    double* v = (double*)calloc((size_t)w * w * w, sizeof(double));
    for (int l1 = 1; l1 <= no; ++l1)
        for (int l2 = 1; l2 <= no; ++l2) {
            if (l2 == l1) continue;
            int qv[3];
            double dk[3], dd = 0.0;
            for (int d = 0; d < 3; ++d) {
                qv[d] = s->nvec[(l1 - 1) * 3 + d] - s->nvec[(l2 - 1) * 3 + d];
                dk[d] = s->kvec[(l1 - 1) * 3 + d] - s->kvec[(l2 - 1) * 3 + d];
            }
            dd = dk[0] * dk[0] + dk[1] * dk[1] + dk[2] * dk[2];
            const size_t c = ((size_t)(qv[0] + maxq) * w + (qv[1] + maxq)) * w + (qv[2] + maxq);
            cnt[c] += 1;
            v[c] = (1.0 / s->omega) * (4.0 * PI / dd);
        }
    int rows = 0;
    for (size_t c = 0; c < (size_t)w * w * w; ++c) {
        if (!(cnt[c] > 1.0e-10)) continue;
        if (rows < cap) {
            q[3 * rows] = (int)(c / ((size_t)w * w)) - maxq;
            q[3 * rows + 1] = (int)((c / w) % w) - maxq;
            q[3 * rows + 2] = (int)(c % w) - maxq;
            count[rows] = cnt[c];
            exv[rows] = v[c];
        }
        ++rows;
    }
    free(cnt); free(v);
    return rows;
}

/* GetG2Drivers, source/CCD.f90:465-504. */
static void g2_drivers(const orc_system* s, double* gab, double* gba) {
    const int o = s->nocc, n = s->nao;
    size_t sz = (size_t)o * o * (n - o);
    memset(gab, 0, sz * sizeof(double));
    memset(gba, 0, sz * sizeof(double));
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                double vab = eri(s, i, j, a, b, s->p.irange[R_DRVDIR]);
                double vba = eri(s, i, j, b, a, s->p.irange[R_DRVEX]);
                if (fmin(vab, vba) < -80.0) continue;
                gab[IDX(i, j, a)] = vab;
                gba[IDX(i, j, a)] = -vba;
            }
}

/* GetG2Mosaics, source/CCD.f90:507-585. */
static void g2_mosaics(const orc_system* s, double* gab, double* gba, const double* taa,
                       const double* tab, const double* tba) {
    const int o = s->nocc, n = s->nao;
    const int fl = s->p.irange[R_MOSAIC];
    double* joo = (double*)calloc((size_t)o + 1, sizeof(double));
    double* jvv = (double*)calloc((size_t)n + 1, sizeof(double));
    for (int k = 1; k <= o; ++k)
        for (int l = 1; l <= o; ++l)
            for (int c = o + 1; c <= n; ++c) {
                int d = find_index(s, k, l, c);
                if (d <= o) continue;
                double vkl = eri(s, c, d, k, l, fl);
                double vlk = eri(s, c, d, l, k, fl);
                if (fmin(vkl, vlk) < -80.0) continue;
                size_t x = IDX(k, l, c);
                jvv[c] = jvv[c] + vkl * (taa[x] + tab[x]) - vlk * (taa[x] + tba[x]);
                joo[k] = joo[k] + vkl * (taa[x] + tab[x]) - vlk * (taa[x] + tba[x]);
            }
    ORC_PAR_I
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                size_t x = IDX(i, j, a);
                gab[x] = gab[x] - 0.5 * (jvv[a] * tab[x] + jvv[b] * tab[x]);
                gba[x] = gba[x] - 0.5 * (jvv[a] * tba[x] + jvv[b] * tba[x]);
                gab[x] = gab[x] - 0.5 * (joo[i] * tab[x] + joo[j] * tab[x]);
                gba[x] = gba[x] - 0.5 * (joo[i] * tba[x] + joo[j] * tba[x]);
            }
    free(joo);
    free(jvv);
}

/* GetG2Ladders, source/CCD.f90:588-686. */
static void g2_ladders(const orc_system* s, double* gab, double* gba, const double* tab,
                       const double* tba) {
    const int o = s->nocc, n = s->nao;
    const int* r = s->p.irange;
    const int f_j2 = imax3(r[R_LADDER], r[R_DLADD], r[R_LINLADD]);
    const int f_j3 = imax3(r[R_LADDER], r[R_EXLADD], r[R_LINLADD]);
    const int f_qd = imax3(r[R_LADDER], r[R_DLADD], r[R_QUADLADD]);
    const int f_qx = imax3(r[R_LADDER], r[R_EXLADD], r[R_QUADLADD]);
    ORC_PAR_I
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j) {
            for (int k = 1; k <= o; ++k) {
                int l = find_index(s, i, j, k);
                if (l > o || l <= 0) continue;
                double j2 = eri(s, i, j, k, l, f_j2);
                double j3 = -eri(s, i, j, l, k, f_j3);
                if (fmin(j2, -j3) < -80.0) continue;
                for (int c = o + 1; c <= n; ++c) {
                    int d = find_index(s, i, j, c);
                    if (d <= o) continue;
                    double vkl = eri(s, c, d, k, l, f_qd);
                    double vlk = eri(s, c, d, l, k, f_qx);
                    if (fmin(vkl, vlk) < -80.0) continue;
                    size_t x = IDX(i, j, c);
                    j2 = j2 + 0.5 * (vkl * tab[x] - vlk * tba[x]);
                    j3 = j3 + 0.5 * (vkl * tba[x] - vlk * tab[x]);
                }
                for (int a = o + 1; a <= n; ++a) {
                    int b = find_index(s, i, j, a);
                    if (b <= o) continue;
                    size_t x = IDX(i, j, a), y = IDX(k, l, a);
                    gab[x] = gab[x] + 0.5 * (j2 * tab[y] + j3 * tba[y]);
                    gba[x] = gba[x] + 0.5 * (j2 * tba[y] + j3 * tab[y]);
                }
            }
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                size_t x = IDX(i, j, a);
                for (int c = o + 1; c <= n; ++c) {
                    int d = find_index(s, i, j, c);
                    if (d <= o) continue;
                    double vab = eri(s, c, d, a, b, f_j2);
                    double vba = eri(s, c, d, b, a, f_j3);
                    if (fmin(vab, vba) < -80.0) continue;
                    size_t y = IDX(i, j, c);
                    gab[x] = gab[x] + 0.5 * (vab * tab[y] - vba * tba[y]);
                    gba[x] = gba[x] + 0.5 * (vab * tba[y] - vba * tab[y]);
                }
            }
        }
}

/* GetG2Rings, source/CCD.f90:689-814. */
static void g2_rings(const orc_system* s, double* gab, double* gba, const double* taa,
                     const double* tab, const double* tba) {
    const int o = s->nocc, n = s->nao;
    const int* r = s->p.irange;
    const int f_ld = imax3(r[R_RING], r[R_DRING], r[R_LINRING]);
    const int f_lx = imax3(r[R_RING], r[R_EXRING], r[R_LINRING]);
    const int f_qd = imax3(r[R_RING], r[R_DRING], r[R_QUADRING]);
    const int f_qx = imax3(r[R_RING], r[R_EXRING], r[R_QUADRING]);
    ORC_PAR_I
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                size_t x = IDX(i, j, a);
                for (int k = 1; k <= o; ++k) {
                    int c = find_index(s, i, k, a);
                    if (c <= o) continue;
                    double vkb = eri(s, c, j, k, b, f_ld);
                    double vbk = eri(s, c, j, b, k, f_lx);
                    if (fmin(vkb, vbk) < -80.0) continue;
                    size_t y = IDX(i, k, a);
                    gab[x] = gab[x] + (vkb - vbk) * tab[y] + vkb * taa[y];
                    gba[x] = gba[x] - vbk * tba[y];
                }
                for (int k = 1; k <= o; ++k) {
                    int c = find_index(s, j, k, b);
                    if (c <= o) continue;
                    double vka = eri(s, c, i, k, a, f_ld);
                    double vak = eri(s, c, i, a, k, f_lx);
                    if (fmin(vka, vak) < -80.0) continue;
                    size_t y = IDX(j, k, b);
                    gab[x] = gab[x] + (vka - vak) * tab[y] + vka * taa[y];
                    gba[x] = gba[x] - vak * tba[y];
                }
            }
    ORC_PAR_I
    for (int i = 1; i <= o; ++i)
        for (int a = o + 1; a <= n; ++a)
            for (int l = 1; l <= o; ++l) {
                double j1 = 0.0, j2 = 0.0, j3 = 0.0;
                for (int k = 1; k <= o; ++k) {
                    int c = find_index(s, i, k, a);
                    if (c <= o) continue;
                    int d = find_index(s, k, l, c);
                    if (d <= o) continue;
                    double vkl = eri(s, c, d, k, l, f_qd);
                    double vlk = eri(s, c, d, l, k, f_qx);
                    if (fmin(vkl, vlk) < -80.0) continue;
                    size_t y = IDX(i, k, a);
                    j1 = j1 + taa[y] * (vkl - vlk) + tab[y] * vkl;
                    j2 = j2 + tab[y] * (vkl - vlk) + taa[y] * vkl;
                    j3 = j3 - tba[y] * vlk;
                }
                for (int j = 1; j <= o; ++j) {
                    int b = find_index(s, i, j, a);
                    if (b <= o) continue;
                    int d = find_index(s, j, l, b);
                    if (d <= o) continue;
                    size_t x = IDX(i, j, a), y = IDX(l, j, d);
                    gab[x] = gab[x] + j2 * taa[y] + j1 * tab[y];
                    gba[x] = gba[x] + j3 * tba[y];
                }
            }
}

/* GetG2XRings, source/CCD.f90:817-940. */
static void g2_xrings(const orc_system* s, double* gab, double* gba, const double* taa,
                      const double* tab, const double* tba) {
    const int o = s->nocc, n = s->nao;
    const int* r = s->p.irange;
    const int f_lx = imax3(r[R_XRING], r[R_EXRING], r[R_LINRING]);
    const int f_ld = imax3(r[R_XRING], r[R_DRING], r[R_LINRING]);
    const int f_qd = imax3(r[R_XRING], r[R_DRING], r[R_QUADRING]);
    const int f_qx = imax3(r[R_XRING], r[R_EXRING], r[R_QUADRING]);
    ORC_PAR_I
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                size_t x = IDX(i, j, a);
                for (int k = 1; k <= o; ++k) {
                    int c = find_index(s, k, i, b);
                    if (c <= o) continue;
                    double vka = eri(s, j, c, k, a, f_lx);
                    double vak = eri(s, j, c, a, k, f_ld);
                    if (fmin(vka, vak) < -80.0) continue;
                    size_t y = IDX(k, i, b);
                    gab[x] = gab[x] - vka * tab[y];
                    gba[x] = gba[x] + (vak - vka) * tba[y] + vak * taa[y];
                }
                for (int k = 1; k <= o; ++k) {
                    int c = find_index(s, k, j, a);
                    if (c <= o) continue;
                    double vkb = eri(s, i, c, k, b, f_lx);
                    double vbk = eri(s, i, c, b, k, f_ld);
                    if (fmin(vkb, vbk) < -80.0) continue;
                    size_t y = IDX(k, j, a);
                    gab[x] = gab[x] - vkb * tab[y];
                    gba[x] = gba[x] + (vbk - vkb) * tba[y] + vbk * taa[y];
                }
            }
    ORC_PAR_I
    for (int i = 1; i <= o; ++i)
        for (int b = o + 1; b <= n; ++b)
            for (int l = 1; l <= o; ++l) {
                double j1 = 0.0, j2 = 0.0, j3 = 0.0;
                for (int k = 1; k <= o; ++k) {
                    int c = find_index(s, i, k, b);
                    if (c <= o) continue;
                    int d = find_index(s, k, l, c);
                    if (d <= o) continue;
                    double vkl = eri(s, c, d, k, l, f_qd);
                    double vlk = eri(s, c, d, l, k, f_qx);
                    if (fmin(vkl, vlk) < -80.0) continue;
                    size_t y = IDX(k, i, b);
                    j1 = j1 + taa[y] * (vkl - vlk) + tba[y] * vkl;
                    j2 = j2 + tba[y] * (vkl - vlk) + taa[y] * vkl;
                    j3 = j3 + tab[y] * vlk;
                }
                for (int j = 1; j <= o; ++j) {
                    int a = find_index(s, j, i, b);
                    if (a <= o) continue;
                    int d = find_index(s, l, j, a);
                    if (d <= o) continue;
                    size_t x = IDX(i, j, a), y = IDX(l, j, a);
                    gba[x] = gba[x] - j1 * tba[y] - j2 * taa[y];
                    gab[x] = gab[x] + j3 * tab[y];
                }
            }
}

/* GetRes, source/CCD.f90:389-424: R = -G + D*T on valid entries. */
static void get_res(const orc_system* s, const double* t, const double* g, double* rr) {
    const int o = s->nocc, n = s->nao;
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                double den = s->eigen[i - 1] + s->eigen[j - 1] - s->eigen[a - 1] - s->eigen[b - 1];
                size_t x = IDX(i, j, a);
                rr[x] = -g[x] + den * t[x];
            }
}

void orc_terms(const orc_system* s, int mask, const double* taa, const double* tab,
               const double* tba, double* gab, double* gba, double* rab, double* rba) {
    if (mask & 1) {
        g2_drivers(s, gab, gba);
    } else {
        size_t sz = (size_t)s->nocc * s->nocc * (s->nao - s->nocc);
        memset(gab, 0, sz * sizeof(double));
        memset(gba, 0, sz * sizeof(double));
    }
    if ((mask & 2) && s->p.do_mosaic) g2_mosaics(s, gab, gba, taa, tab, tba);
    if ((mask & 4) && s->p.do_ladder) g2_ladders(s, gab, gba, tab, tba);
    if ((mask & 8) && s->p.do_ring) g2_rings(s, gab, gba, taa, tab, tba);
    if ((mask & 16) && s->p.do_xring) g2_xrings(s, gab, gba, taa, tab, tba);
    if (rab) get_res(s, tab, gab, rab);
    if (rba) get_res(s, tba, gba, rba);
}

/* SolveCC, source/CCD.f90:354-386. */
void orc_solvecc(const orc_system* s, double* t2, const double* g2) {
    const int o = s->nocc, n = s->nao;
    const double df = s->denomfactor;
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                double den = s->eigen[i - 1] + s->eigen[j - 1] - s->eigen[a - 1] - s->eigen[b - 1];
                size_t x = IDX(i, j, a);
                t2[x] = (g2[x] / den - (1.0 - df) * t2[x]) / df;
            }
}

/* LU with partial pivoting, returning LAPACK-style info (0, or the first
 * column whose pivot is exactly zero); solves only when info == 0, as the
 * reference's call to DGeSV (source/DIISCC.f90:114). */
static int lu_solve(int nn, double* m /* col-major nn x nn */, double* rhs) {
    int info = 0;
    int* piv = (int*)malloc(sizeof(int) * (size_t)nn);
    for (int c = 0; c < nn; ++c) {
        int p = c;
        double best = fabs(m[c + c * nn]);
        for (int r2 = c + 1; r2 < nn; ++r2)
            if (fabs(m[r2 + c * nn]) > best) { best = fabs(m[r2 + c * nn]); p = r2; }
        piv[c] = p;
        if (m[p + c * nn] == 0.0) { if (info == 0) info = c + 1; continue; }
        if (p != c)
            for (int k = 0; k < nn; ++k) { double t = m[c + k * nn]; m[c + k * nn] = m[p + k * nn]; m[p + k * nn] = t; }
        for (int r2 = c + 1; r2 < nn; ++r2) {
            double f = m[r2 + c * nn] / m[c + c * nn];
            m[r2 + c * nn] = f;
            for (int k = c + 1; k < nn; ++k) m[r2 + k * nn] -= f * m[c + k * nn];
        }
    }
    if (info == 0) {
        for (int c = 0; c < nn; ++c) if (piv[c] != c) { double t = rhs[c]; rhs[c] = rhs[piv[c]]; rhs[piv[c]] = t; }
        for (int r2 = 0; r2 < nn; ++r2) for (int k = 0; k < r2; ++k) rhs[r2] -= m[r2 + k * nn] * rhs[k];
        for (int r2 = nn - 1; r2 >= 0; --r2) {
            for (int k = r2 + 1; k < nn; ++k) rhs[r2] -= m[r2 + k * nn] * rhs[k];
            rhs[r2] /= m[r2 + r2 * nn];
        }
    }
    free(piv);
    return info;
}

/* GetTFromDIIS, source/DIISCC.f90:52-131, including its B-matrix overwrite. */
static void get_t_from_diis(size_t sz, int ndiis, int niter, double* t2, const double* oldt2,
                            double* r2) {
    const int iwait = 5;
    double c[16] = {0};
    c[ndiis - 1] = 1.0;
    if (niter >= ndiis + iwait) {
        int nn = ndiis + 1;
        double m[17 * 17] = {0}, v[17] = {0};
        for (int p = 0; p < ndiis; ++p)
            for (int q = 0; q < ndiis; ++q) {
                double acc = m[p + q * nn];
                for (size_t x = 0; x < sz; ++x) acc += r2[x + sz * p] * r2[x + sz * q];
                m[p + q * nn] = acc;
            }
        for (int p = 0; p < ndiis; ++p)
            for (int q = 0; q < nn; ++q) m[p + q * nn] = 1.0;      /* M(1:NDIIS,1:) = One */
        for (int q = 0; q < nn; q += ndiis) m[0 + q * nn] = 1.0;   /* M(1,1::NDIIS) = One */
        v[ndiis] = 1.0;
        if (lu_solve(nn, m, v) == 0)
            for (int p = 0; p < ndiis; ++p) c[p] = v[p];
        for (size_t x = 0; x < sz; ++x) t2[x] = 0.0;
        for (int p = 0; p < ndiis; ++p)
            for (size_t x = 0; x < sz; ++x) t2[x] = t2[x] + c[p] * oldt2[x + sz * p];
    } else {
        memcpy(t2, oldt2 + sz * (ndiis - 1), sz * sizeof(double));
    }
    memmove(r2, r2 + sz, sz * (ndiis - 1) * sizeof(double));
}

/* DrvCCD, source/CCD.f90:18-183 (Max_N_DIIS = 3, source/DIISCC.f90:19). */
int orc_ccd(const orc_system* s, double tolmax, int maxiter, double* taa, double* tab,
            double* tba, double* ecorr_out, double* maxres_out, int* failed,
            double* trace_e, double* trace_dt) {
    const int o = s->nocc, n = s->nao, ndiis = 3;
    const size_t sz = (size_t)o * o * (n - o);
    if (tolmax <= 0.0) tolmax = s->tolmax;
    if (maxiter <= 0) maxiter = 1000;
    double* gab = (double*)calloc(sz, sizeof(double));
    double* gba = (double*)calloc(sz, sizeof(double));
    double* oldab = (double*)calloc(sz * ndiis, sizeof(double));
    double* oldba = (double*)calloc(sz * ndiis, sizeof(double));
    double* rab = (double*)calloc(sz * ndiis, sizeof(double));
    double* rba = (double*)calloc(sz * ndiis, sizeof(double));
    memcpy(oldab + sz * (ndiis - 1), tab, sz * sizeof(double));
    memcpy(oldba + sz * (ndiis - 1), tba, sz * sizeof(double));

    int niter = 1;
    double dt = 0.0;
    for (size_t x = 0; x < sz; ++x) {
        dt = fmax(dt, fabs(tab[x]));
        dt = fmax(dt, fabs(tba[x]));
    }
    double ecorr = orc_energy(s, taa, tab, tba);
    if (trace_e) trace_e[0] = ecorr;
    if (trace_dt) trace_dt[0] = dt;
    while (dt >= tolmax && niter < maxiter) {
        get_t_from_diis(sz, ndiis, niter, tab, oldab, rab);
        get_t_from_diis(sz, ndiis, niter, tba, oldba, rba);
        for (size_t x = 0; x < sz; ++x) taa[x] = tab[x] + tba[x];
        orc_terms(s, 31, taa, tab, tba, gab, gba, rab + sz * (ndiis - 1), rba + sz * (ndiis - 1));
        orc_solvecc(s, tab, gab);
        orc_solvecc(s, tba, gba);
        for (size_t x = 0; x < sz; ++x) taa[x] = tab[x] + tba[x];
        /* CnvrgCC, source/CCD.f90:427-462 */
        niter = niter + 1;
        memmove(oldab, oldab + sz, sz * (ndiis - 1) * sizeof(double));
        memmove(oldba, oldba + sz, sz * (ndiis - 1) * sizeof(double));
        double dab = 0.0, dba = 0.0;
        for (size_t x = 0; x < sz; ++x) {
            dab = fmax(dab, fabs(tab[x] - oldab[x + sz * (ndiis - 1)]));
            dba = fmax(dba, fabs(tba[x] - oldba[x + sz * (ndiis - 1)]));
        }
        dt = fmax(dab, dba);
        memcpy(oldab + sz * (ndiis - 1), tab, sz * sizeof(double));
        memcpy(oldba + sz * (ndiis - 1), tba, sz * sizeof(double));
        ecorr = orc_energy(s, taa, tab, tba);
        if (trace_e && niter <= maxiter) trace_e[niter - 1] = ecorr;
        if (trace_dt && niter <= maxiter) trace_dt[niter - 1] = dt;
    }
    orc_terms(s, 31, taa, tab, tba, gab, gba, rab + sz * (ndiis - 1), rba + sz * (ndiis - 1));
    double maxres = 0.0;
    for (size_t x = 0; x < sz; ++x) {
        maxres = fmax(maxres, fabs(rab[x + sz * (ndiis - 1)]));
        maxres = fmax(maxres, fabs(rba[x + sz * (ndiis - 1)]));
    }
    if (ecorr_out) *ecorr_out = ecorr;
    if (maxres_out) *maxres_out = maxres;
    if (failed) *failed = maxres > s->failratio * tolmax;
    free(gab); free(gba); free(oldab); free(oldba); free(rab); free(rba);
    return niter;
}

/* ------------------------------------------------------------------------
 * dRPA / rCCD (source/dRPA.f90): the ring-only CCD on the single spin-summed
 * T2 that DrvdRPA iterates (source/Main.f90:85-91 passes T2aaaa).
 */

/* CCEnergy of the dRPA module, source/dRPA.f90:181-216. */
double orc_drpa_energy(const orc_system* s, const double* t2, double csosex) {
    const int o = s->nocc, n = s->nao;
    double ecorr = 0.0;
    for (int i = 1; i <= o; ++i)
        for (int j = 1; j <= o; ++j)
            for (int a = o + 1; a <= n; ++a) {
                int b = find_index(s, i, j, a);
                if (b <= o) continue;
                double vijab = eri(s, a, b, i, j, 0);
                double vijba = eri(s, b, a, i, j, 0);
                size_t x = IDX(i, j, a);
                double de = vijab * t2[x] * 2 - vijba * t2[x] * csosex;
                ecorr = ecorr + de;
            }
    return ecorr;
}

/* TermsWrapper of the dRPA module, source/dRPA.f90:146-175: GetG2Drivers
 * (source/dRPA.f90:364-396), GetG2Mosaics (399-450), GetG2Rings (453-551),
 * then GetRes (282-312).  term_mask bit0 drivers, bit1 mosaics, bit3 rings,
 * each further gated by the Do* flags as in TermsWrapper. */
void orc_drpa_terms(const orc_system* s, int term_mask, const double* t2, double* g2,
                    double* r2) {
    const int o = s->nocc, n = s->nao;
    const size_t sz = (size_t)o * o * (n - o);
    memset(g2, 0, sz * sizeof(double));
    if (term_mask & 1)
        for (int i = 1; i <= o; ++i)
            for (int j = 1; j <= o; ++j)
                for (int a = o + 1; a <= n; ++a) {
                    int b = find_index(s, i, j, a);
                    if (b <= o) continue;
                    double v = eri(s, i, j, a, b, s->p.irange[R_DRVDIR]);
                    if (v < -80.0) continue;
                    g2[IDX(i, j, a)] = v;
                }
    if ((term_mask & 2) && s->p.do_mosaic) {
        double* joo = (double*)calloc((size_t)o + 1, sizeof(double));
        double* jvv = (double*)calloc((size_t)n + 1, sizeof(double));
        for (int k = 1; k <= o; ++k)
            for (int l = 1; l <= o; ++l)
                for (int c = o + 1; c <= n; ++c) {
                    int d = find_index(s, k, l, c);
                    if (d <= o) continue;
                    double v = eri(s, c, d, k, l, s->p.irange[R_MOSAIC]);
                    if (v < -80.0) continue;
                    jvv[c] = jvv[c] + 2 * v * t2[IDX(k, l, c)];
                    joo[k] = joo[k] + 2 * v * t2[IDX(k, l, c)];
                }
        for (int i = 1; i <= o; ++i)
            for (int j = 1; j <= o; ++j)
                for (int a = o + 1; a <= n; ++a) {
                    int b = find_index(s, i, j, a);
                    if (b <= o) continue;
                    size_t x = IDX(i, j, a);
                    g2[x] = g2[x] - 0.5 * (jvv[a] * t2[x] + jvv[b] * t2[x]);
                    g2[x] = g2[x] - 0.5 * (joo[i] * t2[x] + joo[j] * t2[x]);
                }
        free(joo);
        free(jvv);
    }
    if ((term_mask & 8) && s->p.do_ring) {
        const int flin = imax3(s->p.irange[R_RING], s->p.irange[R_DRING], s->p.irange[R_LINRING]);
        const int fquad = imax3(s->p.irange[R_RING], s->p.irange[R_DRING], s->p.irange[R_QUADRING]);
        for (int i = 1; i <= o; ++i)
            for (int j = 1; j <= o; ++j)
                for (int a = o + 1; a <= n; ++a) {
                    int b = find_index(s, i, j, a);
                    if (b <= o) continue;
                    size_t x = IDX(i, j, a);
                    for (int k = 1; k <= o; ++k) {
                        int c = find_index(s, i, k, a);
                        if (c <= o) continue;
                        double v = eri(s, c, j, k, b, flin);
                        if (v < -80.0) continue;
                        g2[x] = g2[x] + 2 * v * t2[IDX(i, k, a)];
                    }
                    for (int k = 1; k <= o; ++k) {
                        int c = find_index(s, j, k, b);
                        if (c <= o) continue;
                        double v = eri(s, c, i, k, a, flin);
                        if (v < -80.0) continue;
                        g2[x] = g2[x] + 2 * v * t2[IDX(j, k, b)];
                    }
                }
        for (int i = 1; i <= o; ++i)
            for (int a = o + 1; a <= n; ++a)
                for (int l = 1; l <= o; ++l) {
                    double jidal = 0.0;
                    for (int k = 1; k <= o; ++k) {
                        int c = find_index(s, i, k, a);
                        if (c <= o) continue;
                        int d = find_index(s, k, l, c);
                        if (d <= o) continue;
                        double v = eri(s, c, d, k, l, fquad);
                        if (v < -80.0) continue;
                        jidal = jidal + 2 * t2[IDX(i, k, a)] * v;
                    }
                    for (int j = 1; j <= o; ++j) {
                        int b = find_index(s, i, j, a);
                        if (b <= o) continue;
                        int d = find_index(s, j, l, b);
                        if (d <= o) continue;
                        g2[IDX(i, j, a)] = g2[IDX(i, j, a)] + 2 * jidal * t2[IDX(l, j, d)];
                    }
                }
    }
    if (r2) get_res(s, t2, g2, r2);
}

/* DrvdRPA, source/dRPA.f90:23-143, with the DIIS module's GetTFromDIIS
 * (Max_N_DIIS = 3, source/DIISCC.f90:19, 52-131) and the dRPA module's own
 * SolveCC (219-250) and CnvrgCC (315-344).  Returns NIter; *ecorr is
 * Ec(dRPA), *esosex Ec(SOSEX) of the final amplitudes. */
int orc_drpa(const orc_system* s, double tolmax, int maxiter, double* t2, double* ecorr_out,
             double* esosex_out, double* maxres_out, int* failed, double* trace_e,
             double* trace_dt) {
    const int o = s->nocc, n = s->nao, ndiis = 3;
    const size_t sz = (size_t)o * o * (n - o);
    if (tolmax <= 0.0) tolmax = s->tolmax;
    if (maxiter <= 0) maxiter = 1000;
    double* g2 = (double*)calloc(sz, sizeof(double));
    double* old = (double*)calloc(sz * ndiis, sizeof(double));
    double* r2 = (double*)calloc(sz * ndiis, sizeof(double));
    memcpy(old + sz * (ndiis - 1), t2, sz * sizeof(double));
    int niter = 1;
    double dt = 0.0;
    for (size_t x = 0; x < sz; ++x) dt = fmax(dt, fabs(t2[x]));
    double ecorr = orc_drpa_energy(s, t2, 0.0);
    if (trace_e) trace_e[0] = ecorr;
    if (trace_dt) trace_dt[0] = dt;
    while (dt >= tolmax && niter < maxiter) {
        get_t_from_diis(sz, ndiis, niter, t2, old, r2);
        orc_drpa_terms(s, 11, t2, g2, r2 + sz * (ndiis - 1));
        orc_solvecc(s, t2, g2);
        niter = niter + 1;
        memmove(old, old + sz, sz * (ndiis - 1) * sizeof(double));
        dt = 0.0;
        for (size_t x = 0; x < sz; ++x) dt = fmax(dt, fabs(t2[x] - old[x + sz * (ndiis - 1)]));
        memcpy(old + sz * (ndiis - 1), t2, sz * sizeof(double));
        ecorr = orc_drpa_energy(s, t2, 0.0);
        if (trace_e && niter <= maxiter) trace_e[niter - 1] = ecorr;
        if (trace_dt && niter <= maxiter) trace_dt[niter - 1] = dt;
    }
    orc_drpa_terms(s, 11, t2, g2, r2 + sz * (ndiis - 1));
    double maxres = 0.0;
    for (size_t x = 0; x < sz; ++x) maxres = fmax(maxres, fabs(r2[x + sz * (ndiis - 1)]));
    if (ecorr_out) *ecorr_out = ecorr;
    if (esosex_out) *esosex_out = orc_drpa_energy(s, t2, 1.0);
    if (maxres_out) *maxres_out = maxres;
    if (failed) *failed = maxres > 10000.0 * tolmax;
    free(g2); free(old); free(r2);
    return niter;
}
