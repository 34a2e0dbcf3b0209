/* Test infrastructure (oracle): NOT part of the shipped product.
 *
 * COMMON blocks of the reference chemistry (src/lib/ckrates.f:6-29,
 * solvesp.f:24-82, solveh.f:20-33, spgsolv.f:82) for the grid PARAMETER
 * (LI=711, LJ=131, src/lib/ckrates.f:4).  Declared extern by default; a file
 * that defines REF_COMMON as empty before including this one defines them
 * (oracle/shim_commons.c, for a driver built without the Fortran routines).
 * /CB08/ and /REAC/ are always extern: the flang-built BLOCK DATA defines them. */
#ifndef REF_COMMONS_H
#define REF_COMMONS_H
#include <stdint.h>

#define LI 711
#define LJ 131
#define LE (LI * LJ)
#define LSP 52
#define LRX 544

#ifndef REF_COMMON
#define REF_COMMON extern
#endif

REF_COMMON struct { int32_t isym, iflow, iswrl, ithrm, ichem, igrav, itr; int32_t iskip[LE]; } cb01_;
REF_COMMON struct { double rref, vref, tref, pref, wm[LSP]; } cb02_;
REF_COMMON struct { double dt, x[LE], y[LE], xxc[LI], xxs[LI], yyc[LJ], yys[LJ]; } cb03_;
REF_COMMON struct { double rho[LE], fsp[LSP][LE], u[LE], v[LE], w[LE], p[LE], ht[LE], tk[LE], ak[LE], eps[LE]; } cb04_;
REF_COMMON struct { double alstr, vstr, rstr, pstr, tstr, tmstr, amstr, acstr, dstr, hstr, cpstr, akstr,
                epstr, wmstr, gasc, beta1, beta2, beta3, beta4, rei; } cb06_;
REF_COMMON struct { double vmu[LE], vtc[LE], vdf[LSP][LE], tmu[LE]; } cb07_;
extern struct { double tpol1, tpol2, polsp[LSP][14], cisp[LSP][12]; } cb08_;
extern struct { int32_t lrev, nrlind, itbef[LRX]; double alf[LRX], akf[LRX], exaf[LRX], allow[LRX],
                aklow[LRX], ealow[LRX], troe[4][LRX]; } reac_;
REF_COMMON struct { double hfo[LSP]; } hform_;
REF_COMMON struct { double qdot[LE]; } hrls_;
REF_COMMON struct { double qdsp[LSP][LE]; } dumd_;
REF_COMMON struct { double rsoot, stmf[LE], stnd[LE], stdf[LE]; } soot_;
REF_COMMON struct { double a[55][LE]; } dummy_;
REF_COMMON struct { double a[149][LE]; } dum0_;
REF_COMMON struct { double ze[LE], zw[LE], zn[LE], zs[LE]; } dumx_;
REF_COMMON struct { double ssp[LSP][LE], fpz[LE], emu[LE], ekt[LE]; } dumy_;
REF_COMMON struct { double rc[LRX][LE]; } dumr_;
REF_COMMON struct { double ste[LE], stw[LE], stn[LE], sts[LE], sst1[LE], sst2[LE], sour[6][LE]; } dums_;

#endif
