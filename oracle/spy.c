/* Test infrastructure (oracle): NOT part of the shipped product.
 *
 * CKRATES / SOLVESP / SOLVEH for the reference program itself (PROGRAM UNICORN,
 * src/unicornd-mild-sd-5.f) built with spgsolv.f's three INCLUDEs
 * (src/lib/spgsolv.f:2-4) resolved to nothing.  Each entry point writes the
 * COMMON state the real program holds at that moment to raw float64/int32
 * files, then runs the flang-built reference routine (renamed ref_ckrates_,
 * ref_solvesp_, ref_solveh_ by objcopy), so the run is the reference's own.
 *
 * What it pins (tests/test_program.py):
 *   - /CB06/ scales, /CB02/ WM, /HFORM/ HFO, /REAC/ and /CB08/ as main:327-468
 *     and the BLOCK DATA leave them, before the first chemistry call;
 *   - the transport properties /CB07/ VMU, VTC, VDF and /SOOT/ STDF that
 *     main:829-1005 computed from the state at the same moment;
 *   - the post-step density (main:1056-1073, 1188-1190): RHO at the next step's
 *     CKRATES entry against the state SOLVEH left;
 *   - the state at every entry and exit, for drop-in comparisons.
 *
 * Environment: UC_SPY_DIR (output directory, required); UC_SPY_STOP=n exits
 * the program at the entry of the n-th CKRATES call, after its dump; UC_SPY_HIN=1
 * also dumps the state at SOLVEH's entry. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ref_commons.h"

extern void ref_ckrates_(void);
extern void ref_solvesp_(int32_t *isor1, double *relxsp, double *tolrsp, double *spnorm);
extern void ref_solveh_(int32_t *isor2, double *relxh, double *tolrh, double *sigsp, double *hnorm);

static int n_rates, n_sp, n_h;

static void put(const char *tag, int n, const char *name, const void *p, size_t bytes) {
  const char *dir = getenv("UC_SPY_DIR");
  if (!dir) {
    fprintf(stderr, "spy: UC_SPY_DIR is not set\n");
    exit(2);
  }
  char path[4096];
  snprintf(path, sizeof path, "%s/%s%d_%s.bin", dir, tag, n, name);
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(p, 1, bytes, f) != bytes || fclose(f) != 0) {
    fprintf(stderr, "spy: cannot write %s\n", path);
    exit(2);
  }
}

#define PUT(tag, n, name, arr) put(tag, n, name, arr, sizeof(arr))
#define PUTV(tag, n, name, v) put(tag, n, name, &(v), sizeof(v))

static void dump_constants(void) {
  PUTV("const", 0, "cb06", cb06_);        /* 20 doubles, main:421-446 */
  PUTV("const", 0, "cb02", cb02_);        /* RREF VREF TREF PREF WM(52), main:461 */
  PUT("const", 0, "hfo", hform_.hfo);     /* main:462 */
  PUTV("const", 0, "cb08", cb08_);        /* TPOL1 TPOL2 POLSP CISP */
  PUTV("const", 0, "reac", reac_);        /* after main:364-413 */
  PUTV("const", 0, "rsoot", soot_.rsoot);
}

static void dump_state(const char *tag, int n) {
  PUT(tag, n, "iskip", cb01_.iskip);
  PUT(tag, n, "rho", cb04_.rho);
  PUT(tag, n, "fsp", cb04_.fsp);
  PUT(tag, n, "uvw", cb04_.u);            /* U; V and W follow it in the block */
  put(tag, n, "v", cb04_.v, sizeof(cb04_.v));
  put(tag, n, "w", cb04_.w, sizeof(cb04_.w));
  put(tag, n, "p", cb04_.p, sizeof(cb04_.p));
  PUT(tag, n, "ht", cb04_.ht);
  PUT(tag, n, "tk", cb04_.tk);
  PUT(tag, n, "vmu", cb07_.vmu);
  PUT(tag, n, "vtc", cb07_.vtc);
  PUT(tag, n, "vdf", cb07_.vdf);
  PUT(tag, n, "tmu", cb07_.tmu);
  PUT(tag, n, "stmf", soot_.stmf);
  PUT(tag, n, "stnd", soot_.stnd);
  PUT(tag, n, "stdf", soot_.stdf);
  PUTV(tag, n, "dt", cb03_.dt);
  PUT(tag, n, "x", cb03_.x);
  PUT(tag, n, "y", cb03_.y);
  PUT(tag, n, "ssp", dumy_.ssp);
  PUT(tag, n, "fpz", dumy_.fpz);
  PUT(tag, n, "qdot", hrls_.qdot);
  PUT(tag, n, "qdsp", dumd_.qdsp);
  PUT(tag, n, "sst1", dums_.sst1);
  PUT(tag, n, "sst2", dums_.sst2);
  int32_t flags[3] = {cb01_.ichem, cb01_.ithrm, cb01_.itr};
  PUT(tag, n, "flags", flags);
}

void ckrates_(void) {
  ++n_rates;
  if (n_rates == 1) dump_constants();
  dump_state("rates", n_rates);
  const char *stop = getenv("UC_SPY_STOP");
  if (stop && n_rates >= atoi(stop)) exit(0);
  ref_ckrates_();
}

void solvesp_(int32_t *isor1, double *relxsp, double *tolrsp, double *spnorm) {
  ++n_sp;
  double ctl[4] = {(double)*isor1, *relxsp, *tolrsp, *spnorm};
  ref_solvesp_(isor1, relxsp, tolrsp, spnorm);
  ctl[0] = ctl[0] * 1e6 + *isor1;   /* sweeps allowed * 1e6 + sweeps used */
  PUT("sp", n_sp, "ctl", ctl);
}

void solveh_(int32_t *isor2, double *relxh, double *tolrh, double *sigsp, double *hnorm) {
  ++n_h;
  double ctl[5] = {(double)*isor2, *relxh, *tolrh, *sigsp, *hnorm};
  if (getenv("UC_SPY_HIN")) dump_state("hin", n_h);
  ref_solveh_(isor2, relxh, tolrh, sigsp, hnorm);
  ctl[0] = ctl[0] * 1e6 + *isor2;
  PUT("h", n_h, "ctl", ctl);
  dump_state("hout", n_h);
}
