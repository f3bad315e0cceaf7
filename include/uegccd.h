/*
 * uegccd.h -- C ABI of the MI355X CCD amplitude solver (libuegccd.so).
 *
 * The library replaces the body of the reference's CCD driver,
 *     Subroutine DrvCCD(HEGData, UEGInfo, T2aaaa, T2abab, T2abba)
 *         (source/CCD.f90:18-32, public at source/CCD.f90:13-14),
 * together with what that body calls per iteration: TermsWrapper
 * (source/CCD.f90:146-181), SolveCC/GetRes/CnvrgCC (source/CCD.f90:354-462),
 * CCEnergy (source/CCD.f90:186-237) and the DIIS module (source/DIISCC.f90).
 * The MP2 initial guess (DrvMBPT, source/MP2.f90:11-157) and the plane-wave
 * basis (SetupBasis/change_rs, source/HEG.f90:150-324,494-533) are exported as
 * well so a host can drive a whole single-point calculation through it.
 *
 * Conventions
 *  - Amplitude arrays are the reference's explicit-shape FP64 arrays
 *    T(NOcc,NOcc,NOcc+1:NAO) in Fortran (column-major) order:
 *        element (i,j,a) at (i-1) + NOcc*(j-1) + NOcc*NOcc*(a-NOcc-1).
 *    Host pointers; the library copies in and out.  Entries whose momentum
 *    partner b = k_i + k_j - k_a is not a virtual orbital are zero on output.
 *  - Every function returns UEGCCD_OK (0) or a negative status; nothing calls
 *    exit/stop.  uegccd_strerror() names each status.
 *  - A context owns its device buffers and one HIP stream; contexts are
 *    independent (no globals), calls on one context must be serialised by the
 *    caller.  Nothing falls back to the CPU: without a usable HIP device the
 *    calls fail with UEGCCD_ERR_HIP.
 */
#ifndef UEGCCD_H
#define UEGCCD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    UEGCCD_OK = 0,
    UEGCCD_ERR_INVALID = -1,      /* bad argument or inconsistent system      */
    UEGCCD_ERR_ALLOC = -2,        /* device or host allocation failed         */
    UEGCCD_ERR_HIP = -3,          /* HIP runtime error / no device            */
    UEGCCD_ERR_UNSUPPORTED = -4,  /* an Input option this build does not run  */
    UEGCCD_ERR_NOT_CONVERGED = -5,/* solve stopped at max_iter (soft failure)  */
    UEGCCD_ERR_NONFINITE = -6,    /* NaN/Inf in amplitudes or residual         */
    UEGCCD_ERR_BUSY = -7,         /* server: no free slot / HBM budget now     */
    UEGCCD_ERR_NOT_FOUND = -8,    /* server: unknown request id               */
    UEGCCD_ERR_SCRATCH = -9       /* larger than the fixed device scratch of
                                     uegccd_solve_noalloc                      */
};

/* Iteration schemes. */
enum {
    /* The reference's own iteration: damped Jacobi with the DIIS of
       source/DIISCC.f90 (whose B-matrix is overwritten, so it never
       extrapolates), stopped at the first max|dT| < tol_dt exactly as the
       loop at source/CCD.f90:80-99.  Same NIter and trajectory. */
    UEGCCD_MODE_REFERENCE = 0,
    /* Pulay DIIS of length ndiis on the device; stops when |dE| < tol_e and
       max|R| < tol_r (R = D*T - G of the input amplitudes). */
    UEGCCD_MODE_DIIS = 1
};

enum { UEGCCD_PREC_FP64 = 0, UEGCCD_PREC_MIXED = 1 };

/*
 * The data of HEGDataType/UEGInfoType the solver reads (source/Types.f90:7-82).
 * nvec is the reference's nVec(NAO,3) array, column-major: component d of
 * orbital p (1-based) at nvec[(d-1)*nao + (p-1)].
 */
typedef struct {
    int32_t nocc;          /* UEGInfo%NOcc                                  */
    int32_t nao;           /* UEGInfo%NAO                                   */
    const int32_t* nvec;   /* HEGData%nVec(NAO,3)                            */
    const double* eigen;   /* HEGData%Eigen(NAO)                             */
    double L;              /* HEGData%L (box length at r_s)                  */
    double madelung;       /* HEGData%Madelung (already x MadelungFactor)    */
    double rs;             /* HEGData%rS                                     */
    double gapii;          /* UEGInfo%GapII                                  */
    int32_t coren;         /* UEGInfo%CoreN                                  */
    int32_t safe_eri;      /* UEGInfo%SafeERI                                */
    int32_t irange[15];    /* IRange* in Input-key order (source/IO.f90:88-102) */
    int32_t do_ring, do_xring, do_ladder, do_mosaic;
    double screening;      /* HEGData%ScreeningDistance (source/HEG.f90:512);
                              read only when a range flag selects the short- or
                              long-range kernel (flags 1, 2 under CasedERI)  */
} uegccd_system;

typedef struct {
    int32_t mode;          /* UEGCCD_MODE_*                                  */
    int32_t ndiis;         /* DIIS length (DIIS mode), 2..16                 */
    int32_t diis_start;    /* first NIter at which DIIS extrapolates         */
    int32_t max_iter;      /* <= 0: 1000 (source/CCD.f90:36)                 */
    double tol_dt;         /* <= 0: FindTol's TolMax (source/HEG.f90:1073)    */
    double tol_e;          /* DIIS mode energy tolerance (Ha)                */
    double tol_r;          /* DIIS mode max|R| tolerance                     */
    double damping;        /* <= 0: FindTol's DenomFactor (2, or 5 if rs>5.1)
                              in reference mode, 1 (undamped) in DIIS mode   */
    int32_t precision;     /* UEGCCD_PREC_*                                  */
    double mixed_switch;   /* MIXED: FP32 ladders while max|R| >= this       */
    int32_t device;        /* HIP device ordinal                             */
    int32_t iters_per_sync;/* iterations launched between host polls (>=1)  */
    int32_t graph;         /* 1: each chunk of passes between host polls is
                              captured once as a HIP graph and replayed
                              (single-context solves); 0: launched kernel by
                              kernel                                        */
    int32_t persistent;    /* 1: each chunk of passes of a single-context solve
                              runs as one cooperative launch whose phases are
                              separated by grid-wide barriers (the slice,
                              contraction, residual and state-update phases
                              of a pass); 0: off                            */
} uegccd_options;

typedef struct {
    double ecorr;          /* Ec(CCD)                                        */
    double ehf_plus_ec;    /* unused by the library (host fills it)          */
    double dt;             /* last max|dT|                                   */
    double maxres;         /* max|D*T - G| at the returned amplitudes        */
    int32_t niter;         /* NIter as the reference counts it               */
    int32_t failed;        /* MaxRes > FailRatio*TolMax (source/CCD.f90:107) */
    int32_t iters_fp32;    /* iterations whose ladders ran in FP32           */
    int32_t status;        /* UEGCCD_OK / NOT_CONVERGED / NONFINITE          */
    double solve_ms;       /* device wall time of the iteration loop         */
} uegccd_result;

typedef struct uegccd_ctx uegccd_ctx;

/* Fills *opt with the reference behaviour: MODE_REFERENCE, FindTol values,
 * FP64, max_iter 1000. */
void uegccd_default_options(uegccd_options* opt);

/* Builds the per-basis index maps (momentum partners, pair-momentum blocks,
 * Coulomb table) and device buffers.  Replaces SetUpCCScr/SetUpDIISCC
 * (source/CCScr.f90:9-30, source/DIISCC.f90:10-45) and the FindIndex/ERI
 * machinery (source/HEG.f90:357-404, 689-978).  Any basis with NOcc <= 64
 * whose element map fits 64 bits (about 10^6 orbitals at N=14) is taken;
 * the CCD passes (uegccd_solve*, uegccd_residual, uegccd_kernel_time,
 * served requests) return UEGCCD_ERR_UNSUPPORTED for a basis with a
 * virtual-occupied |n_a - n_k|^2 above 1022, while uegccd_drpa, uegccd_mp2,
 * uegccd_energy and the structure factors run on it. */
int uegccd_ctx_create(const uegccd_system* sys, const uegccd_options* opt, uegccd_ctx** out);
void uegccd_ctx_destroy(uegccd_ctx* ctx);

/* Sizes: NOcc, NAO, number of valid (i,j,a), number of pair-momentum blocks. */
int uegccd_ctx_info(const uegccd_ctx* ctx, int32_t* nocc, int32_t* nao, int64_t* nvalid,
                    int32_t* nblocks);

/* One residual evaluation, TermsWrapper (source/CCD.f90:146-181):
 * G = drivers + mosaics + ladders + rings + xrings, R = D*T - G.
 * Any of g2abab/g2abba/r2abab/r2abba may be NULL.  Like uegccd_energy and
 * uegccd_structure_factor, it copies its amplitudes into the context: they
 * become the amplitudes uegccd_solve_resident(from_mp2 = 0) starts from, and
 * the iteration state is reset. */
int uegccd_residual(uegccd_ctx* ctx, const double* t2abab, const double* t2abba,
                    double* g2abab, double* g2abba, double* r2abab, double* r2abba);

/* CCEnergy (source/CCD.f90:186-237).  The amplitudes become the context's
 * resident amplitudes (see uegccd_residual). */
int uegccd_energy(uegccd_ctx* ctx, const double* t2abab, const double* t2abba, double* ecorr);

/* Structure factor of the amplitudes (host arrays, the layout above): replaces
 * CCEnergyAndSF (source/CCD.f90:240-351), which the reference calls on the MP2
 * guess (source/CCD.f90:65, DoSFCalcMP2) and on the converged amplitudes
 * (source/CCD.f90:109, DoSFCalcCCD).  One row per momentum transfer q = n_a - n_i
 * or n_a - n_j of a valid element, in the reference's print order (q_x, then q_y,
 * then q_z ascending): qvec[3k..3k+2] = q and vals[5k..5k+4] = (SfG = 2 pi |q| / L,
 * SfSum, SfV = v(q), SfDSum, SfXSum).  With qvec or vals NULL, *ncell receives the
 * row count and nothing is computed; otherwise *ncell is the capacity in rows on
 * entry (UEGCCD_ERR_INVALID if too small) and the row count on return.  *ecorr,
 * when given, receives the energy the reference routine also sets.  The
 * amplitudes become the context's resident amplitudes (see uegccd_residual). */
int uegccd_structure_factor(uegccd_ctx* ctx, const double* t2abab, const double* t2abba, int32_t* ncell,
                            int32_t* qvec, double* vals, double* ecorr);

/* DrvMBPT's structure factor (source/MP2.f90:86-117, written to unit 81 on
 * every MP2 call): the MP2 amplitudes are written on the device (as
 * uegccd_mp2, which they replace in the context) and binned as
 * uegccd_structure_factor bins, with three values per cell, vals[3k..3k+2] =
 * (SfG, SfSum, SfV); SfV is DrvMBPT's integral, which carries no range flag.
 * Size query and capacity as uegccd_structure_factor; *e2 (may be NULL)
 * receives E(2). */
int uegccd_mp2_structure_factor(uegccd_ctx* ctx, int32_t* ncell, int32_t* qvec, double* vals,
                                double* e2);

/* DrvCCD (source/CCD.f90:18-183): iterate from the given amplitudes (the MP2
 * guess in the reference) to convergence.  t2aaaa is written as
 * t2abab + t2abba on output (may be NULL). */
int uegccd_solve(uegccd_ctx* ctx, double* t2aaaa, double* t2abab, double* t2abba,
                 uegccd_result* res);

/* As uegccd_solve, but the amplitudes stay in device memory: from_mp2 != 0
 * first writes the MP2 guess on the device (the caller-side copy of
 * source/Main.f90:85-87 without a host round trip); otherwise the solve starts
 * from the amplitudes the context holds: those of the last call that left
 * amplitudes there (a solve's result, uegccd_mp2's guess, or the input of
 * uegccd_residual, uegccd_energy or uegccd_structure_factor).
 * Read them with uegccd_get_amplitudes.  No counterpart in the reference; for
 * callers that keep T2 resident across solves (serving, benchmarks). */
int uegccd_solve_resident(uegccd_ctx* ctx, int32_t from_mp2, uegccd_result* res);
int uegccd_get_amplitudes(uegccd_ctx* ctx, double* t2aaaa, double* t2abab, double* t2abba);

/* Average device time (microseconds, HIP events on the context's stream) of
 * one kernel of the residual pass over the context's current amplitudes:
 * 0 per-slice amplitudes and ring/mosaic products, 1 contractions
 * (particle-particle ladder tiles, hole-hole W rows and ladder, quadratic
 * rings per momentum transfer, Joo), 3 residual assembly + update (2, the
 * former separate ladder kernel, is rejected); the ladders in the context's
 * starting precision.  No counterpart in the reference; measurement
 * support. */
int uegccd_kernel_time(uegccd_ctx* ctx, int32_t kernel, int32_t reps, double* avg_us);

/* Tile choices of the tuned kernels: k-steps per group of loads in the
 * ladder tiles (1, 2 or 4) and per round of loads in the ring (4 or 8) and
 * hole-hole (4 or 8) contraction tiles.  A context starts from the tuned
 * table compiled into the library (uegccd_amd/tuned_tiles.json, written by
 * the offline sweep python -m uegccd_amd.tune), by the class of its largest
 * pair-momentum block and its precision; uegccd_set_tiles overrides that
 * (NULL restores it), uegccd_get_tiles reports it with the block sizes that
 * select the class.  A batch runs the tiles of its solve with the most ladder
 * tiles.  No counterpart in the reference; tuning support. */
typedef struct {
    int32_t ladder_group, quad_steps, hh_steps;
} uegccd_tiles;
int uegccd_set_tiles(uegccd_ctx* ctx, const uegccd_tiles* tiles);
int uegccd_get_tiles(const uegccd_ctx* ctx, uegccd_tiles* tiles, int32_t* noo_max, int32_t* nvv_max);

/* In-solve timing of one kernel (numbering as uegccd_kernel_time; -1 turns
 * it off; the per-solve reductions, convergence test and DIIS bookkeeping of
 * a pass run at the start of the next pass's kernel 0): in every pass of the context's later solves, HIP stop events on the
 * context's stream are written by that kernel's dispatch and by the dispatch
 * before it (kernels of a stream run back to back, so the difference is the
 * kernel's duration as a kernel trace shows it; when it is the first kernel
 * of a pass the first pass of each host poll is not timed).
 * uegccd_kernel_stats returns the summed device time and the number of such
 * launches since the last call (passes in which some solve was still
 * iterating; the passes launched after the last solve finished return at
 * once and are not counted), and resets them.  No
 * counterpart in the reference; measurement support. */
int uegccd_time_kernel(uegccd_ctx* ctx, int32_t kernel);
int uegccd_kernel_stats(uegccd_ctx* ctx, double* total_us, int64_t* launches);

/* Per-iteration (E, dT) of the last solve, index NIter-1 (the lines the
 * reference writes with format 1040, source/CCD.f90:76,97). */
int uegccd_trace(const uegccd_ctx* ctx, double* e, double* dt, int32_t n);

/* DrvMBPT amplitudes and E(2) (source/MP2.f90:56-98).  Any array may be NULL. */
int uegccd_mp2(uegccd_ctx* ctx, double* t2aaaa, double* t2abab, double* t2abba, double* e2);

/* Replaces DrvdRPA (source/dRPA.f90:23-143, called from source/Main.f90:91,
 * 139, 253 when DodRPASOSEX): the ring-only CCD (dRPA + SOSEX) on the
 * spin-summed amplitudes T2 (the reference passes its T2aaaa array), with the
 * ring and mosaic terms the system enables, iterated as the reference does
 * (damped Jacobi; its DIIS is inert) until max|dT| < TolMax, then the residual
 * check (failed when max|R| > 10000 TolMax).  t2 (layout as the other
 * amplitude arrays): in, the starting amplitudes (NULL: the MP2 T2aaaa); out,
 * the final ones.  res->ecorr is Ec(dRPA) of the last iteration, *esosex (may
 * be NULL) Ec(SOSEX) of the final amplitudes; uegccd_trace holds the
 * (E, dT) lines.  The context must be FP64 in UEGCCD_MODE_REFERENCE, else
 * UEGCCD_ERR_UNSUPPORTED.  Returns the status as uegccd_solve. */
int uegccd_drpa(uegccd_ctx* ctx, double* t2, uegccd_result* res, double* esosex);

/* DrvMBPT's connectivity (source/MP2.f90:75-85): counts[k] is the number of
 * (i,j,a,b) excitations with |n_i - n_a|^2 = k+1 plus those with
 * |n_i - n_b|^2 = k+1, over the valid (i,j,a); counted on the device.
 * nbins <= 4096 (the reference uses 400); a transfer beyond nbins returns
 * UEGCCD_ERR_INVALID, where the reference stops ("Connectivity greater than
 * current ConnectivityLimit"). */
int uegccd_connectivity(uegccd_ctx* ctx, int32_t* counts, int32_t nbins);

/* Plane-wave basis of init_HEG + change_rs (source/HEG.f90:16-43,150-324,
 * 494-533): enumeration |n|^2 <= maxkpoint, stable ordering by |n+twist|^2,
 * fcut/corefcut truncation, HF eigenvalues with the Madelung term, E_HF.
 * cutoff_twisted = 1 (no counterpart in the reference, which cuts on the
 * untwisted |n|^2 at source/HEG.f90:190-194) enumerates |n+twist|^2 <=
 * maxkpoint instead: a sphere of the same radius around -twist, so that the
 * basis of every twist holds the same kinetic-energy window.
 * Call with nvec == NULL to query nao; then nvec holds nao*3 (column-major
 * like nVec(NAO,3)) and eigen nao entries.  scalars (may be NULL) receives
 * L, Omega, Madelung, EHF, ScreeningDistance, DenomFactor, TolMax, FailRatio. */
typedef struct {
    int32_t nel;
    int32_t maxkpoint;
    double rs;
    double twist[3];
    double fcut, corefcut;
    double madelungfactor, gap, exchangefactor;
    int32_t cutoff_twisted; /* 0: |n|^2 <= maxkpoint; 1: |n+twist|^2 <= maxkpoint */
} uegccd_heg_params;

int uegccd_heg_build(const uegccd_heg_params* p, int32_t* nao, int32_t* nvec, double* eigen,
                     double* scalars);

/* The same basis built on HIP device `device` (the same results to the last
 * bit): the |n|^2 <= maxkpoint cube is keyed by |n+twist|^2 in a coalesced
 * kernel, sorted with its enumeration index in LDS by one block (the stable
 * order of SetupBasis, source/HEG.f90:180-217), and the orbitals, HF
 * eigenvalues and E_HF (source/HEG.f90:150-324, 581-615) come from one
 * kernel; the Madelung sum and the FindTol scalars stay on the host.  Cutoffs
 * with (2*floor(sqrt(maxkpoint)) + 7)^3 > 8192 return UEGCCD_ERR_UNSUPPORTED. */
int uegccd_heg_build_device(const uegccd_heg_params* p, int32_t device, int32_t* nao,
                            int32_t* nvec, double* eigen, double* scalars);

/* print_ex_SF (source/HEG.f90:617-674), the exchange structure factor
 * print_HF writes to unit 82 after each basis set-up and change of r_s: for
 * the basis of p at r_s = p->rs, over ordered pairs (l1, l2) of distinct
 * occupied orbitals, count[k] pairs with n_l1 - n_l2 = qvec[3k..3k+2] and
 * exv[k] = (1/Omega) 4 pi / |k_l1 - k_l2|^2 of the cell's last pair in the
 * reference's loop order, cells in its print order (q_x, then q_y, then q_z
 * ascending).  With qvec, count or exv NULL, *ncell receives the cell count;
 * else *ncell is the capacity on entry and the count on return.  A host
 * computation (no device). */
int uegccd_exchange_structure_factor(const uegccd_heg_params* p, int32_t* ncell, int32_t* qvec,
                                     int32_t* count, double* exv);

const char* uegccd_strerror(int status);

/* ------------------------------------------------------------------------
 * Allocation-free solve.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same solve as
 * uegccd_ctx_create + uegccd_solve + uegccd_ctx_destroy (same kernels, same
 * results), with the context's device memory (maps, amplitudes, DIIS
 * history, intermediates, state) taken from a fixed per-device scratch array
 * of UEGCCD_NOALLOC_SCRATCH_MB MiB that the library holds from the first call
 * on a device until the process ends.  It calls no hipMalloc/hipFree, makes
 * no stream, event or pinned buffer (it runs on the null stream and polls
 * into pageable memory); host-side planning uses the C++ heap.  A system
 * whose context needs more than the scratch returns UEGCCD_ERR_SCRATCH
 * (uegccd_footprint gives the size; the N=54 |n|^2<=16 solve with DIIS 8
 * needs 56 MB).  Every call on a device uses the same array and leaves its
 * contents behind (each call clears what it needs): calls that overlap on one
 * device race and must be serialised by the caller.  Arguments as
 * uegccd_ctx_create (sys, opt) and uegccd_solve (amplitudes in/out, res). */
#define UEGCCD_NOALLOC_SCRATCH_MB 64
int uegccd_solve_noalloc(const uegccd_system* sys, const uegccd_options* opt, double* t2aaaa,
                         double* t2abab, double* t2abba, uegccd_result* res);

/* ------------------------------------------------------------------------
 * Batched solves (twist averaging / r_s scans): several independent systems
 * iterated together, each kernel launch covering every system's tiles.  No
 * counterpart in the reference (which runs twists one after another,
 * source/Main.f90:112-168); each system's result equals its uegccd_solve.
 * ---------------------------------------------------------------------- */
int uegccd_solve_batch(uegccd_ctx** ctxs, int32_t n, double** t2abab, double** t2abba,
                       uegccd_result* res);

/* Device bytes a context for this system and these options would hold (the
 * admission measure of the server).  No counterpart in the reference. */
int uegccd_footprint(const uegccd_system* sys, const uegccd_options* opt, int64_t* bytes);

/* ------------------------------------------------------------------------
 * Serving: a request loop that admits solves (twist averaging, r_s scans) as
 * device memory frees up, keeps each one's amplitudes and DIIS history
 * resident, and iterates every admitted solve together (continuous batching:
 * solves join and leave between passes).  Each request's result equals its
 * solve alone.  No counterpart in the reference, whose twist loop is serial
 * (source/Main.f90:112-168).  One caller thread drives it with
 * uegccd_server_step / uegccd_server_drain; the server's own planner threads
 * prepare queued requests (basis, maps) meanwhile.
 * ---------------------------------------------------------------------- */
typedef struct uegccd_server uegccd_server;

typedef struct {
    int32_t device;
    int64_t hbm_budget_bytes;  /* admission limit on resident solves (<= 0: 64 GiB) */
    int32_t max_active;        /* most solves iterated together (<= 0: 64)        */
    int32_t iters_per_step;    /* passes launched per scheduling step (<= 0: 4)   */
    int32_t planner_threads;   /* host threads building bases and maps of queued
                                  requests (0: min(8, cores/2); < 0: none)       */
} uegccd_server_options;

typedef struct {
    uegccd_heg_params heg;     /* the system; the server builds its basis       */
    double gapii;              /* the UEGInfo fields of uegccd_system            */
    int32_t coren, safe_eri;
    int32_t irange[15];
    int32_t do_ring, do_xring, do_ladder, do_mosaic;
    uegccd_options opt;        /* this request's solver options                  */
} uegccd_request;

enum { UEGCCD_REQ_QUEUED = 0, UEGCCD_REQ_RUNNING = 1, UEGCCD_REQ_DONE = 2 };

typedef struct {
    int32_t state;             /* UEGCCD_REQ_*                                   */
    int32_t status;            /* final status of a done request                 */
    int32_t nao, nocc;
    int32_t niter, iters_fp32;
    double ehf, e2, ecorr, maxres;
    double t_submit_ms, t_start_ms, t_done_ms;   /* server clock                 */
    int64_t footprint_bytes;
    double plan_ms;            /* host time spent planning the request (basis,
                                  maps, constant image) on a planner thread    */
} uegccd_request_result;

int uegccd_server_create(const uegccd_server_options* opt, uegccd_server** out);
void uegccd_server_destroy(uegccd_server* srv);
int uegccd_server_submit(uegccd_server* srv, const uegccd_request* req, int64_t* id);
/* One scheduling step: admit what fits, run iters_per_step passes of every
 * admitted solve, retire the finished ones. */
int uegccd_server_step(uegccd_server* srv, int32_t* n_active, int32_t* n_queued);
int uegccd_server_drain(uegccd_server* srv);
int uegccd_server_poll(const uegccd_server* srv, int64_t id, uegccd_request_result* res);
double uegccd_server_clock_ms(const uegccd_server* srv);

#ifdef __cplusplus
}
#endif
#endif
