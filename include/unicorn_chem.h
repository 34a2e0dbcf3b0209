/* unicorn_chem.h -- C ABI of the MI355X (gfx950) chemistry step for UNICORN.
 *
 * Two layers:
 *
 * 1. The Fortran drop-in (ISO_C_BINDING / plain symbol override).  These keep
 *    the reference routines' names, argument lists and COMMON-block array
 *    layout; a UNICORN build links them instead of INCLUDE-ing
 *    src/lib/ckrates.f, solvesp.f and solveh.f (src/lib/spgsolv.f:2-4):
 *
 *      ckrates_  replaces SUBROUTINE CKRATES                      (src/lib/ckrates.f:1)
 *      solvesp_  replaces SUBROUTINE SOLVESP(ISOR1,RELXSP,TOLRSP,SPNORM)
 *                                                                 (src/lib/solvesp.f:1)
 *      solveh_   replaces SUBROUTINE SOLVEH(ISOR2,RELXH,TOLRH,SIGSP,HNORM)
 *                                                                 (src/lib/solveh.f:1)
 *
 *    or, for a build that drops spgsolv.f as a whole,
 *
 *      spgsolv_  replaces SUBROUTINE SPGSOLV(ISOR1,RELXSP,TOLRSP,RESDC,
 *                ISOR2,RELXH,TOLRH,RESDH,SIGH,SIGSP)             (src/lib/spgsolv.f:6)
 *                with the coefficient assembly on the device too
 *
 *    They live in libunicorn_fshim.so (unicorn_amd/csrc/fortran_shim.hip),
 *    built for the grid PARAMETER(LI,LJ) of the Fortran build, and read and
 *    write the COMMON blocks /CB01/ /CB02/ /CB03/ /CB04/ /CB06/ /CB08/ /REAC/
 *    /HFORM/ /HRLS/ /DUMD/ /SOOT/ /DUMY/ /DUMS/ /DUMR/ /DUMMY/ /DUM0/ /DUMX/ (and
 *    /CB07/ for spgsolv_)
 *    exactly as the reference routines do, with or without the transport
 *    coefficients SPGSOLV assembles (E/W/P/Q, ZE..ZS, STE..STS, the RDHSP inputs)
 *    and with the velocities of T(h); everything runs on the GPU (there is no CPU
 *    fallback).
 *
 * 2. The device API (libunicorn_chem.so), used by the shim and directly by
 *    host programs: a context holds the scales and scratch, a step advances a
 *    device-resident field by one chemistry step.
 *
 * Field layout (both layers): the reference's, i.e. for an LI x LJ grid the
 * flat index is c = (I-1)*LJ + (J-1) (J fastest); species arrays are
 * [52][LI*LJ].  Temperature TK and enthalpy HT are the reference's
 * nondimensional variables (T/TSTR, h/VSTR^2).
 *
 * Errors: every function returns 0 on success or a negative UC_E_* code; the
 * message of the last error on the calling thread is uc_last_error().
 */
#ifndef UNICORN_CHEM_H
#define UNICORN_CHEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UC_ABI_VERSION 4

enum {
    UC_OK = 0,
    UC_E_ARG = -1,       /* bad argument (sizes, null pointers, unknown mode) */
    UC_E_HIP = -2,       /* a HIP runtime call failed */
    UC_E_NOGPU = -3,     /* no gfx950 device */
    UC_E_UNSUPPORTED = -4 /* input outside the implemented contract */
};

/* sweep control: UC_MODE_GLOBAL reproduces the reference's global stops
 * (solvesp.f:2834, solveh.f:1162); UC_MODE_PERCELL stops each cell on its
 * own mean |dY| <= TOLRSP and |dh| <= TOLRH. */
enum { UC_MODE_GLOBAL = 0, UC_MODE_PERCELL = 2 };

/* /CB06/ scales, /CB02/ WM and /HFORM/ HFO as the reference's main program
 * sets them (src/unicornd-mild-sd-5.f:421-464). */
typedef struct {
    double alstr, vstr, rstr, pstr, tstr, tmstr, amstr, acstr, dstr, hstr,
        cpstr, akstr, epstr, wmstr, gasc, beta1, beta2, beta3, beta4, rei;
    double pback;     /* PREF/PSTR (main:429) */
    double wm[52];    /* W_k/1000/WMSTR */
    double hfo[52];   /* HCONV*a6_low(k)/WM(k) */
} uc_scales;

/* Solver controls of one step (the SPGSOLV arguments, spgsolv.f:6-7). */
typedef struct {
    double dt;        /* /CB03/ DT, nondimensional */
    double relxsp, tolrsp, relxh, tolrh;
    int32_t isor1, isor2; /* maximum species / enthalpy sweeps */
    int32_t mode;         /* UC_MODE_* */
    int32_t predict;      /* global mode: predicted sweep count (0 = isor1) */
} uc_controls;

/* Per-step results. */
typedef struct {
    int32_t isor1, isor2;   /* sweeps used, reference convention (max+1 = not converged) */
    double resdc, resdh;    /* SPGSOLV residuals (spgsolv.f:1299-1416) */
    int64_t cell_sweeps;    /* species sweeps executed, summed over cells and launches */
    int64_t redo_cells;     /* cells re-run because the decided count fell before the
                               snapshot window of the last segment */
    double ms_rates, ms_sweeps, ms_sources, ms_total; /* device time: rate coefficients,
                               sweep segments, sources; whole step */
    int32_t sweep_launches; /* sweep-segment launches this step (coupled step: wavefront steps,
                               one k_sweep_cpl launch each) */
    int32_t first_segment;  /* sweeps in the first segment (the previous count minus one) */
    int32_t t_out_of_limit; /* cells whose enthalpy lies above the high-range fit at 5000 K
                               (the condition of solveh.f:1200-1204, message 712, which the
                               reference never prints: IOVER starts at 1, solveh.f:1170) */
    int32_t newton_failed;  /* cells whose T(h) Newton iteration did not converge in 100 steps
                               (solveh.f:1206-1219; the reference prints message 714 for the
                               first of them in a call and DEBUG(9427) for each) */
    int32_t first_failed;   /* flat index (I-1)*LJ+(J-1) of the first of those in I-then-J
                               order, or -1 */
} uc_stats;

/* Device-resident field (all pointers are device pointers). */
typedef struct {
    int32_t li, lj;         /* grid, including the boundary frame */
    double *rho;            /* [n] */
    double *y;              /* [52][n] */
    double *tk, *ht;        /* [n] */
    double *stmf, *stnd;    /* [n] soot mass fraction, number density */
    const int32_t *iskip;   /* [n] >= 3: skipped by the chemistry (ckrates.f:93) */
    double *qdot;           /* [n] out: heat release (solveh.f:1142) */
    double *qdsp;           /* [52][n] out: species production (solveh.f:1137) */
} uc_field;

/* Tuning of a step (uc_autotune chooses it; uc_set_tuning / uc_get_tuning). */
typedef struct {
    int32_t bins;       /* 1: cells in natural order; 2..32: each grid's cells ordered
                           stiffest first by the previous step's sweeps in this many bins */
    int32_t lpt;        /* 1: tiles of 64 cells run stiffest first (longest-processing-time) */
    int32_t persistent; /* 1: the sweep kernel runs persistent waves over the tile queue */
    int32_t block;      /* block size of the one-thread-per-cell kernels: 64, 128, 256 or 512 */
    int32_t graph;      /* 1: each sweep segment is replayed from a captured hipGraph (default 0:
                           launched directly, which measured faster) */
    int32_t lanes;      /* lanes per cell of the chemistry-only sweep: 1 (thread per cell), or
                           2 or 4, one lane in each of that many waves of a workgroup, which
                           split each species block's reaction terms (not with persistent) */
    int32_t rate_parts; /* lanes per cell of the rate-coefficient kernel: 1, 2, 4 or 8, one
                           lane in each of that many workgroups, each forming its share of
                           the reactions' slots; with 8 the T(h) kernel also takes 8 lanes
                           per cell (the same bits for every value); 0 (the default): from
                           the step's cell count, 2 for a full field, up to 8 for a few tiles */
} uc_tuning;

/* What SPGSOLV hands the chemistry beyond the state (spgsolv.f:98-1290); device pointers
 * in the field layout, each may be NULL (zero).  coef[4m+f][n] is face f (0 E: I+1,
 * 1 W: I-1, 2 P: J+1, 3 Q: J-1) of species m+1, in the order of /DUMMY/ then /DUM0/
 * (solvesp.f:24-58: EA0, WA0, PA0, QA0, EA1, ...).  The flux inputs (emu, ekt, tmu,
 * vdf) need the geometry; emu NULL is REI*BETA3*RHO as SPGSOLV sets it (spgsolv.f:175). */
typedef struct {
    const double *coef;          /* [204][n] species coefficients (solvesp.f:424-426) */
    const double *stc;           /* [4][n] STE, STW, STN, STS (solvesp.f:2812-2825) */
    const double *zc;            /* [4][n] ZE, ZW, ZN, ZS (solveh.f:1154-1158) */
    const double *emu, *ekt, *tmu;  /* [n] RDHSP inputs (solveh.f:52-53) */
    const double *vdf;           /* [52][n] */
    const double *ygeo;          /* [n] /CB03/ Y (radius) */
    const double *xxc, *xxs;     /* [li] /CB03/ XXC, XXS */
    const double *yyc, *yys;     /* [lj] /CB03/ YYC, YYS */
    const double *u, *v, *w;     /* [n] velocities: ENTH = (FPZ-(U*U+V*V+W*W)/2)*HSTR+... */
    double sigsp;                /* SOLVEH's SIGSP (main:346: 0.9) */
    int32_t isym;                /* /CB01/ ISYM */
} uc_coupling;

/* SPGSOLV's inputs to its coefficient assembly beyond the state (spgsolv.f:98-349);
 * device pointers in the field layout.  U(J,I) lies on the face between I-1 and I,
 * V(J,I) on the face between J-1 and J (/CB04/).  sigsp and sigh are SPGSOLV's SIGSP
 * and SIGH arguments (main:346-347, both 0.9). */
typedef struct {
    const double *vtc, *tmu;     /* [n] /CB07/ VTC, TMU */
    const double *vdf;           /* [52][n] /CB07/ VDF (species 1..51 used) */
    const double *stdf;          /* [n] /SOOT/ STDF */
    const double *u, *v;         /* [n] /CB04/ U, V */
    const double *ygeo;          /* [n] /CB03/ Y */
    const double *xxc, *xxs;     /* [li] */
    const double *yyc, *yys;     /* [lj] */
    double sigsp, sigh;
    int32_t isym;                /* /CB01/ ISYM */
} uc_transport;

typedef struct uc_context uc_context;

int uc_abi_version(void);
const char *uc_last_error(void);

/* Scales from the reference-values card (main:156): VREF, TREF, PREF, the
 * geometry card's RTIN/RTOT, FO2IN and IGRAV. */
int uc_scales_from_card(double vref, double tref, double pref, double rtin, double rtot,
                        double fo2in, int32_t igrav, uc_scales *out);

/* A context for fields of up to max_cells cells (at most 4,000,000) on the
 * current HIP device.  It owns the per-cell device memory of a step (RC slabs,
 * saved iterates and snapshots, sources; about 9.7 KB per cell) and the
 * sweep-residual partial sums; it allocates with hipMalloc. */
int uc_create(const uc_scales *scales, int64_t max_cells, uc_context **out);

/* A context for batches of up to max_grids grids of max_cells cells in total
 * (max_cells + 64*max_grids <= 4,000,000).  uc_create is the one-grid case. */
int uc_create_batch(const uc_scales *scales, int64_t max_cells, int32_t max_grids, uc_context **out);
int uc_destroy(uc_context *ctx);

/* One chemistry step on a device-resident field, on `stream` (a hipStream_t,
 * NULL = default stream).  Sequence (the reference around SPGSOLV):
 * main:821-828 clip/closure, CKRATES, SOLVESP, SOLVEH, the SPGSOLV epilogue,
 * and the ideal-gas density update main:1056-1073.  Synchronises the stream
 * (the global stops are decided on the host).  The list of active cells is
 * cached per (ISKIP pointer, LI, LJ): a caller that changes ISKIP in place
 * calls uc_cells_changed. */
int uc_step(uc_context *ctx, const uc_field *field, const uc_controls *ctl, void *stream,
            uc_stats *stats);

/* The same step with the coupling inputs (uc_coupling; NULL: uc_step).  Non-zero
 * coefficients run the reference's lexicographic Gauss-Seidel (J outer, I inner; every
 * cell reads its neighbours' latest iterates, solvesp.f:424-426, solveh.f:1154-1160) as a
 * pipelined anti-diagonal wavefront, exactly; the flux inputs add the species-diffusion
 * enthalpy flux to RHSZ (solveh.f:59-115).  Global stops only (UC_MODE_GLOBAL); one grid. */
int uc_step_coupled(uc_context *ctx, const uc_field *field, const uc_controls *ctl,
                    const uc_coupling *coupling, void *stream, uc_stats *stats);

/* SPGSOLV's coefficient assembly (spgsolv.f:98-1290; replaces that part of SUBROUTINE
 * SPGSOLV, spgsolv.f:6-1290, whose outputs are the /DUMMY/, /DUM0/, /DUMX/, /DUMS/ and
 * /DUMY/ EKT arrays): EKT of every cell into ekt[n], and for the interior cells with
 * ISKIP < 10 the E/W/P/Q coefficients of species 1..51 into coef[204][n] (uc_coupling's
 * order), STE..STS into stc[4][n] and ZE..ZS into zc[4][n]; other cells are left as they
 * were.  Reads field->rho, y, tk, iskip and ctl->dt.  Equal bit for bit to the reference
 * routine.  Asynchronous on `stream`; allocates nothing. */
int uc_assemble(uc_context *ctx, const uc_field *field, const uc_controls *ctl, const uc_transport *tr,
                double *coef, double *stc, double *zc, double *ekt, void *stream);

/* Same, on host arrays (copies in and out; the PCIe cost is in ms_total). */
int uc_step_host(uc_context *ctx, const uc_field *host_field, const uc_controls *ctl,
                 uc_stats *stats);

/* One chemistry step of ngrids independent grids (each its own SPGSOLV call: its own
 * DT, controls and global stops) in one set of launches; stats[ngrids].  The grids'
 * cells are packed into the context, binned by stiffness and run as one work queue.
 * No counterpart in the reference (its DT is a single COMMON scalar, /CB03/). */
int uc_step_batch(uc_context *ctx, const uc_field *fields, const uc_controls *ctls,
                  int32_t ngrids, void *stream, uc_stats *stats);

/* Choose the tuning for these grids by timing steps on a copy of their state (the
 * fields are left as they were); the choice stays in the context. */
int uc_autotune(uc_context *ctx, const uc_field *fields, const uc_controls *ctls,
                int32_t ngrids, void *stream, uc_tuning *chosen);
int uc_set_tuning(uc_context *ctx, const uc_tuning *tuning);

/* Mixture-averaged transport properties of every cell (replaces the inline block
 * src/unicornd-mild-sd-5.f:843-1005 of PROGRAM UNICORN with IPROP = 1, whose outputs are
 * /CB07/ VMU, VTC, VDFSP and /SOOT/ STDF): vmu[n], vtc[n], vdf[52][n] and stdf[n] from
 * field->y (after main:821-828) and field->tk, in the reference's term order and
 * float32 literals, equal bit for bit to the reference.  TMU (k-epsilon runs,
 * main:1006-1010) is not part of it.  The outputs feed uc_transport.  Asynchronous on
 * `stream`; the context builds its pair tables (main:469-482) at the first call. */
int uc_transport_props(uc_context *ctx, const uc_field *field, double *vmu, double *vtc, double *vdf,
                       double *stdf, void *stream);

/* uc_transport_props without a context and without allocating device memory.  No
 * counterpart in the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The kernel's
 * table block (about 152 KB: pair tables, fits, constants) lives in one fixed device
 * array per GPU, kept until the process ends; its address is looked up at the first
 * call on each device, and the tables for `scales` are uploaded there when they differ
 * from the last call's (that call synchronises `stream`; later ones only launch).  Limit:
 * LI*LJ <= 4,000,000 cells.  Calls that overlap on one GPU with different scales race
 * on the tables.  Same outputs as uc_transport_props, bit for bit. */
int uc_transport_props_noalloc(const uc_scales *scales, const uc_field *field, double *vmu, double *vtc,
                               double *vdf, double *stdf, void *stream);

/* The ISKIP contents of the grids changed in place: rebuild the cell list at the
 * next step. */
int uc_cells_changed(uc_context *ctx);

/* main:821-828 on one device field (Y<0 -> 0, Y_N2 = 1 - sum Y_1..51, every cell):
 * the first phase of uc_step, on its own.  uc_set_clip(ctx, 0) makes the context's
 * steps start without it, for a caller that has applied it already (the surrogate
 * step clips before its prediction and runs the FP64 step on the rejected cells
 * after).  No separate counterpart in the reference: main:821-828 runs in the time
 * loop before CKRATES. */
int uc_clip(const uc_field *field, void *stream);
int uc_set_clip(uc_context *ctx, int32_t on);

/* on = 1: the context keeps every cell's rate coefficients of the last step that advanced it
 * (as /DUMR/ RC keeps CKRATES' output, ckrates.f:94-1333; zero before the first such step),
 * about 4.4 KB of device memory per cell, and each step also forms QDOT and QDSP of the interior
 * cells with 3 <= ISKIP < 10, which SOLVEH evaluates from those coefficients and SPGSOLV's
 * prologue copies of the species although CKRATES and SOLVESP skip them (solveh.f:119-1142).
 * on = 0 (the default): those cells' QDOT and QDSP are left as they are.  The Fortran
 * drop-in runs with it on. */
int uc_keep_rates(uc_context *ctx, int32_t on);
int uc_get_tuning(const uc_context *ctx, uc_tuning *tuning);

/* Per-cell status of the last step's T(h) (the conditions behind SOLVEH's messages,
 * solveh.f:1200-1227; the reference prints them and continues): flags[li*lj] (device
 * memory, field layout) of grid `grid` (0 for uc_step) receives 1 where the enthalpy lies
 * above the high-range fit at 5000 K, 2 where the Newton iteration did not converge in
 * 100 steps (3: both), 0 elsewhere and at every cell the step did not advance.  The flags
 * are the context's until its next step; asynchronous on `stream`.  uc_stats carries
 * their counts. */
int uc_th_flags(const uc_context *ctx, int32_t grid, uint8_t *flags, void *stream);

/* The species sweeps' residual history of the last step of grid `grid`: RSD/SPNORM after
 * each sweep up to the decided count, the quantity SOLVESP tests against TOLRSP
 * (solvesp.f:2834) and keeps to itself.  Writes at most `cap` values to `out` (host
 * memory) and returns how many the step has (a count >= 0), or a negative error code.
 * Empty for the per-cell stop mode. */
int uc_sweep_residuals(const uc_context *ctx, int32_t grid, double *out, int32_t cap);

/* ---- Fortran drop-in (libunicorn_fshim.so) ------------------------------ */
void ckrates_(void);
/* ISOR1 at most 127 (the rows of the per-sweep residual partials) */
void solvesp_(int32_t *isor1, double *relxsp, double *tolrsp, double *spnorm);
/* ISOR2 at most 1000.  Called without a preceding ckrates_ (ICHEM <= 0 with ITHRM > 0,
 * spgsolv.f:1291-1298) it reads the species SPGSOLV's prologue left in /DUMY/ SSP and the
 * rate coefficients /DUMR/ holds, as SOLVEH does. */
void solveh_(int32_t *isor2, double *relxh, double *tolrh, double *sigsp, double *hnorm);
/* Replaces SUBROUTINE SPGSOLV(ISOR1,RELXSP,TOLRSP,RESDC,ISOR2,RELXH,TOLRH,RESDH,SIGH,SIGSP)
 * (src/lib/spgsolv.f:6) as a whole, for a Fortran build that drops spgsolv.f instead of its
 * three INCLUDE lines: the coefficient assembly on the device (uc_assemble) from the
 * transport COMMONs, the three routines above with those coefficients, and the epilogue
 * (RESDC, RESDH, FA = SA, HT = FPZ, STMF/STND = SST).  The scratch COMMONs SPGSOLV fills on
 * the way (/DUMMY/, /DUM0/, /DUMX/, /DUMS/ STE..STS, /DUMY/ EMU and EKT) are not written.
 * Needs ICHEM > 0 and ITHRM > 0 (stops with a message otherwise). */
void spgsolv_(int32_t *isor1, double *relxsp, double *tolrsp, double *resdc, int32_t *isor2, double *relxh,
              double *tolrh, double *resdh, double *sigh, double *sigsp);

#ifdef __cplusplus
}
#endif
#endif
