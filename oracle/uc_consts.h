/* Test infrastructure (oracle): NOT part of the shipped product.
 *
 * CPU restatement of the reference scales and mechanism pre-processing that
 * UNICORN's main program applies before the chemistry step runs.  Shared by
 * the reference driver (ref_driver.c, linked with the flang-compiled reference
 * routines) and by the C restatement of the chemistry step (chem_oracle.c).
 *
 * Follows src/unicornd-mild-sd-5.f (cited below as main:LINE).  Every default
 * REAL literal of the reference (no D exponent) is a float32 constant widened
 * to REAL*8; F32(x) reproduces that rounding.
 */
#ifndef UC_CONSTS_H
#define UC_CONSTS_H

#include <math.h>

#define UC_LSP 52
#define UC_LRX 544
#define F32(x) ((double)(float)(x))

/* molar masses WSP, g/mol (main:76-86) */
static const double UC_WSP[UC_LSP] = {
    2.016, 32.00, 1.008, 16.00, 17.01, 18.02, 33.01, 34.01, 28.01, 44.01,
    29.02, 30.03, 16.04, 15.04, 14.03, 14.03, 28.05, 31.03, 29.06, 30.07,
    13.02, 26.04, 27.05, 43.05, 44.05, 42.04, 41.03, 25.03, 31.03, 32.04,
    46.07, 44.05, 45.06, 45.06, 43.05, 45.06, 40.07, 39.06, 41.07, 42.08,
    44.10, 43.09, 43.09, 54.09, 29.02, 68.12, 100.2, 56.11, 70.14, 39.95,
    4.003, 28.01};

/* Reference scales of /CB06/ and the derived groups (main:331-468). */
typedef struct {
    /* inputs of the reference-values card (main:156), docs/input.uni:5 */
    double vref, tref, pref, rtin, rtot, fo2in;
    int igrav;
    /* /CB06/ */
    double alstr, vstr, rstr, pstr, tstr, tmstr, amstr, acstr, dstr, hstr,
        cpstr, akstr, epstr, wmstr, gasc, beta1, beta2, beta3, beta4, rei;
    /* /CB02/ RREF, and the derived */
    double rref, pback, hconv;
    double wm[UC_LSP];  /* /CB02/ WM  (main:462) */
    double hfo[UC_LSP]; /* /HFORM/ HFO (main:463); needs POLSP */
} uc_scales;

/* main:331-462; hfo is filled by uc_scales_hfo once POLSP is known. */
static inline void uc_scales_init(uc_scales *s, double vref, double tref,
                                  double pref, double rtin, double rtot,
                                  double fo2in, int igrav) {
    const double GASC = F32(8.3144), GAMA = F32(1.4), GFORCE = F32(9.80665);
    s->vref = vref; s->tref = tref; s->pref = pref; s->rtin = rtin;
    s->rtot = rtot; s->fo2in = fo2in; s->igrav = igrav;
    double wmref = 1.0 / (fo2in / UC_WSP[1] + (1.0 - fo2in) / UC_WSP[UC_LSP - 1]) / 1000.0;
    double amref = 5.69428e-07 * pow(tref, F32(0.62245));
    double acref = 2.78763e-04 * pow(tref, F32(0.796783));
    double dref = 1.2e-04;
    double rref = pref / (GASC / wmref) / tref;
    s->gasc = GASC;
    s->rref = rref;
    s->alstr = rtot - rtin;
    s->vstr = vref;
    s->rstr = rref;
    s->pstr = rref * vref * vref;
    s->pback = pref / s->pstr;
    s->tstr = tref;
    s->hstr = vref * vref;
    s->akstr = vref * vref;
    s->epstr = vref * vref * vref / s->alstr;
    s->cpstr = vref * vref / tref;
    s->acstr = acref;
    s->tmstr = s->alstr / vref;
    s->amstr = amref;
    s->dstr = dref;
    s->wmstr = wmref;
    s->rei = amref / rref / vref / s->alstr;
    s->beta1 = GASC * s->rstr * s->tstr / s->pstr / s->wmstr;
    s->beta2 = s->tstr * s->acstr / s->amstr / s->vstr / s->vstr;
    s->beta3 = s->rstr * s->dstr / s->amstr;
    s->beta4 = (double)igrav * GFORCE * s->alstr / s->vstr / s->vstr;
    if (igrav <= -1) s->beta4 = GFORCE * s->alstr / s->vstr / s->vstr / (double)(-igrav);
    s->hconv = GASC / s->wmstr / s->hstr;
    (void)GAMA;
    for (int i = 0; i < UC_LSP; ++i) s->wm[i] = UC_WSP[i] / 1000.0 / s->wmstr;
}

/* HFO(I) = HCONV*POLSP(6,I)/WM(I)  (main:463); polsp is POLSP(14,52) column-major. */
static inline void uc_scales_hfo(uc_scales *s, const double *polsp) {
    for (int i = 0; i < UC_LSP; ++i) s->hfo[i] = s->hconv * polsp[14 * i + 5] / s->wm[i];
}

/* The docs/input.uni reference-values card: VREF, TREF, PREF, (RREF), ...,
 * FO2IN; RTIN/RTOT from the geometry card; IGRAV from the physics card. */
static inline void uc_scales_default(uc_scales *s) {
    uc_scales_init(s, 10.0, 294.0, 1.0133e+05, 0.00, 0.100, 0.233, 1);
}

#endif
