// fortran_shim.hip -- the Fortran drop-in: CKRATES, SOLVESP and SOLVEH with the
// reference's names, argument lists and COMMON-block layout, executed on gfx950.
//
// A UNICORN build links libunicorn_fshim.so in place of the three routines that
// src/lib/spgsolv.f:2-4 INCLUDEs (delete those three lines); SPGSOLV, its
// coefficient assembly and epilogue, and the main program stay as they are.  The
// program must export its COMMON blocks to shared libraries (link it with
// -Wl,--export-dynamic).  The grid is the Fortran build's PARAMETER (LI, LJ),
// given here at compile time (UC_LI, UC_LJ; src/lib/ckrates.f:4 has 711 x 131).
//
// Division of work (one call each per SPGSOLV call, in this order):
//   ckrates_  upload the old-level state; rate coefficients, rho*Y and soot sources
//             into the device slabs (ckrates.f:92-1352).  RC stays on the device:
//             only SOLVESP and SOLVEH read it (/DUMR/ is scratch elsewhere).
//   solvesp_  sweeps to the global stop, N2 closure (solvesp.f:83-2914); also the
//             SOLVEH sources, which depend only on RC, the final iterates, T and rho.
//             Writes /DUMY/ SSP, /DUMS/ SST1 SST2, /CB04/ FCB (N2) and ISOR1.
//   solveh_   enthalpy relaxation to the global stop and T(h) (solveh.f:1147-1234).
//             Writes /DUMY/ FPZ, /CB04/ TK, /HRLS/ QDOT, /DUMD/ QDSP and ISOR2.
//
// Contract: chemistry only.  The transport coefficients SPGSOLV assembles
// (/DUMMY/ and /DUM0/ E,W,P,Q of every species, /DUMX/ ZE..ZS, /DUMS/ STE..STS) must
// be zero at every cell the chemistry advances; otherwise the program stops with a
// message (the neighbour-coupled relaxation is not implemented; there is no CPU
// fallback).  Errors of the device layer also stop the program, after printing
// uc_last_error().  Cells with 3 <= ISKIP < 10 keep their QDOT/QDSP: the reference
// evaluates them there from RC values CKRATES did not compute in that step.
#include "chem_kernels.hip"

#include <cstdlib>

#ifndef UC_LI
#define UC_LI 711
#endif
#ifndef UC_LJ
#define UC_LJ 131
#endif

namespace shim {
constexpr int64_t LI = UC_LI, LJ = UC_LJ, LE = LI * LJ;
constexpr int LSP = 52;
}  // namespace shim

// COMMON blocks, as the reference declares them (src/lib/ckrates.f:6-29,
// solvesp.f:24-82, solveh.f:20-33, spgsolv.f:82); only the members used here.
extern "C" {
extern struct { int32_t isym, iflow, iswrl, ithrm, ichem, igrav, itr; int32_t iskip[shim::LE]; } cb01_;
extern struct { double rref, vref, tref, pref, wm[shim::LSP]; } cb02_;
extern struct { double dt; } cb03_;
extern struct { double rho[shim::LE], fsp[shim::LSP][shim::LE], u[shim::LE], v[shim::LE], w[shim::LE],
                p[shim::LE], ht[shim::LE], tk[shim::LE]; } cb04_;
extern struct { double alstr, vstr, rstr, pstr, tstr, tmstr, amstr, acstr, dstr, hstr, cpstr, akstr,
                epstr, wmstr, gasc, beta1, beta2, beta3, beta4, rei; } cb06_;
extern struct { double hfo[shim::LSP]; } hform_;
extern struct { double qdot[shim::LE]; } hrls_;
extern struct { double qdsp[shim::LSP][shim::LE]; } dumd_;
extern struct { double rsoot, stmf[shim::LE], stnd[shim::LE]; } soot_;
extern struct { double a[55][shim::LE]; } dummy_;
extern struct { double a[149][shim::LE]; } dum0_;
extern struct { double z[4][shim::LE]; } dumx_;
extern struct { double ssp[shim::LSP][shim::LE], fpz[shim::LE], emu[shim::LE], ekt[shim::LE]; } dumy_;
extern struct { double vmu[shim::LE], vtc[shim::LE], vdf[shim::LSP][shim::LE], tmu[shim::LE]; } cb07_;
extern struct { double st[4][shim::LE], sst1[shim::LE], sst2[shim::LE]; } dums_;
}

namespace {

// k_finish and k_sources write per-list-position rows; scatter them to field layout
__global__ void k_scatter_rows(const double* __restrict__ src, int64_t npad, const int32_t* __restrict__ cells,
                               int32_t rows, double* __restrict__ dst, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npad) return;
  const int32_t c = cells[q];
  if (c < 0) return;  // padding lane
  for (int r = 0; r < rows; ++r) dst[r * n + c] = src[r * npad + q];
}

struct Shim {
  uc_context* ctx = nullptr;
  uc_scales sc{};
  hipStream_t s = nullptr;
  double *rho = nullptr, *y = nullptr, *tk = nullptr, *ht = nullptr, *stmf = nullptr, *stnd = nullptr;
  double *qdot = nullptr, *qdsp = nullptr, *ssp = nullptr, *sst = nullptr, *fpz = nullptr;
  int32_t* iskip = nullptr;
  uc_field f{};
  int stage = 0;  // 1 after ckrates_, 2 after solvesp_
};
Shim g;

[[noreturn]] void stop(const std::string& where, const std::string& msg) {
  std::fprintf(stderr, "unicorn_fshim: %s: %s\n", where.c_str(), msg.c_str());
  std::fflush(stderr);
  std::exit(3);
}

void check(int rc, const char* where) {
  if (rc != UC_OK) stop(where, uc_last_error());
}

void hipcheck(hipError_t e, const char* where) {
  if (e != hipSuccess) stop(where, hipGetErrorString(e));
}

// /CB06/, /CB02/ WM and PREF, /HFORM/ as the main program set them (main:421-464)
uc_scales scales_from_commons() {
  uc_scales s{};
  s.alstr = cb06_.alstr; s.vstr = cb06_.vstr; s.rstr = cb06_.rstr; s.pstr = cb06_.pstr; s.tstr = cb06_.tstr;
  s.tmstr = cb06_.tmstr; s.amstr = cb06_.amstr; s.acstr = cb06_.acstr; s.dstr = cb06_.dstr;
  s.hstr = cb06_.hstr; s.cpstr = cb06_.cpstr; s.akstr = cb06_.akstr; s.epstr = cb06_.epstr;
  s.wmstr = cb06_.wmstr; s.gasc = cb06_.gasc; s.beta1 = cb06_.beta1; s.beta2 = cb06_.beta2;
  s.beta3 = cb06_.beta3; s.beta4 = cb06_.beta4; s.rei = cb06_.rei;
  s.pback = cb02_.pref / cb06_.pstr;  // main:429
  for (int k = 0; k < shim::LSP; ++k) { s.wm[k] = cb02_.wm[k]; s.hfo[k] = hform_.hfo[k]; }
  return s;
}

void ensure_context() {
  const uc_scales sc = scales_from_commons();
  if (g.ctx && std::memcmp(&sc, &g.sc, sizeof(sc)) == 0) return;
  if (g.ctx) check(uc_destroy(g.ctx), "ckrates_");
  g.sc = sc;
  check(uc_create(&g.sc, shim::LE, &g.ctx), "ckrates_");
  if (!g.s) {
    hipcheck(hipStreamCreate(&g.s), "ckrates_: hipStreamCreate");
    const size_t b1 = sizeof(double) * shim::LE, b52 = b1 * shim::LSP;
    for (double** p : {&g.rho, &g.tk, &g.ht, &g.stmf, &g.stnd, &g.qdot, &g.fpz})
      hipcheck(hipMalloc(p, b1), "ckrates_: hipMalloc");
    for (double** p : {&g.y, &g.qdsp, &g.ssp}) hipcheck(hipMalloc(p, b52), "ckrates_: hipMalloc");
    hipcheck(hipMalloc(&g.sst, 2 * b1), "ckrates_: hipMalloc");
    hipcheck(hipMalloc(&g.iskip, sizeof(int32_t) * shim::LE), "ckrates_: hipMalloc");
  }
  g.f = uc_field{(int32_t)shim::LI, (int32_t)shim::LJ, g.rho, g.y, g.tk, g.ht, g.stmf, g.stnd, g.iskip,
                 g.qdot, g.qdsp};
}

bool active(int64_t c) {
  const int64_t i = c / shim::LJ, j = c % shim::LJ;
  return i > 0 && i < shim::LI - 1 && j > 0 && j < shim::LJ - 1 && cb01_.iskip[c] < 3;
}

// the chemistry-only contract: no transport coefficient at any cell the chemistry advances
void check_contract() {
  auto nonzero = [](const double* a) {
    for (int64_t c = 0; c < shim::LE; ++c)
      if (a[c] != 0.0 && active(c)) return true;
    return false;
  };
  for (int k = 0; k < 55; ++k)
    if (nonzero(dummy_.a[k])) stop("ckrates_", "non-zero transport coefficient in /DUMMY/ (array " +
                                                std::to_string(k + 1) + "): only the chemistry-only contract is implemented");
  for (int k = 0; k < 149; ++k)
    if (nonzero(dum0_.a[k])) stop("ckrates_", "non-zero transport coefficient in /DUM0/ (array " +
                                               std::to_string(k + 1) + "): only the chemistry-only contract is implemented");
  for (int k = 0; k < 4; ++k) {
    if (nonzero(dumx_.z[k])) stop("ckrates_", "non-zero enthalpy coefficient in /DUMX/");
    if (nonzero(dums_.st[k])) stop("ckrates_", "non-zero soot coefficient in /DUMS/");
  }
  // RDHSP = (EMU VDF + REI TMU / SIGSP - EKT) HSSP (solveh.f:52-53) feeds the species-
  // diffusion enthalpy flux of RHSZ (solveh.f:59-115) through every cell's neighbours;
  // this drop-in computes RHSZ with that flux zero, so RDHSP must vanish everywhere:
  // EKT = TMU = 0, and VDF = 0 wherever EMU is not (SPGSOLV sets EMU even without
  // transport)
  for (int64_t c = 0; c < shim::LE; ++c) {
    bool flux = dumy_.ekt[c] != 0.0 || cb07_.tmu[c] != 0.0;
    for (int k = 0; k < shim::LSP && !flux && dumy_.emu[c] != 0.0; ++k) flux = cb07_.vdf[k][c] != 0.0;
    if (flux)
      stop("ckrates_", "non-zero species-diffusion enthalpy flux input (EKT, TMU, or VDF with EMU): the flux "
                       "term of SOLVEH is not implemented, only the chemistry-only contract");
  }
}

void h2d(void* dst, const void* src, size_t b, const char* where) {
  hipcheck(hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, g.s), where);
}
void d2h(void* dst, const void* src, size_t b, const char* where) {
  hipcheck(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToHost, g.s), where);
}

}  // namespace

extern "C" {

// SUBROUTINE CKRATES (src/lib/ckrates.f:1)
void ckrates_(void) {
  check_contract();  // host COMMONs only: refuse before touching the GPU
  ensure_context();
  const size_t b1 = sizeof(double) * shim::LE, b52 = b1 * shim::LSP;
  h2d(g.rho, cb04_.rho, b1, "ckrates_");
  h2d(g.y, cb04_.fsp, b52, "ckrates_");  // FA = SA on entry (spgsolv.f:169-220)
  h2d(g.tk, cb04_.tk, b1, "ckrates_");
  h2d(g.ht, dumy_.fpz, b1, "ckrates_");   // FPZ = HT on entry
  h2d(g.fpz, dumy_.fpz, b1, "ckrates_");
  h2d(g.stmf, soot_.stmf, b1, "ckrates_");
  h2d(g.stnd, soot_.stnd, b1, "ckrates_");
  h2d(g.iskip, cb01_.iskip, sizeof(int32_t) * shim::LE, "ckrates_");
  h2d(g.ssp, dumy_.ssp, b52, "ckrates_");
  h2d(g.sst, dums_.sst1, b1, "ckrates_");
  h2d(g.sst + shim::LE, dums_.sst2, b1, "ckrates_");
  h2d(g.qdot, hrls_.qdot, b1, "ckrates_");
  h2d(g.qdsp, dumd_.qdsp, b52, "ckrates_");
  hipcheck(hipStreamSynchronize(g.s), "ckrates_");
  g.ctx->lay.ngrids = 0;  // ISKIP may change between steps: rebuild the cell list
  // SOLVESP/SOLVEH supply the controls and norms; these are placeholders until then
  const uc_controls ctl{cb03_.dt, 0.99, 1e-8, 0.99, 1e-8, 100, 100, UC_MODE_GLOBAL, 0};
  check(ph_begin(g.ctx, &g.f, &ctl, 1, g.s, false, nullptr, nullptr), "ckrates_");
  check(ph_rates(g.ctx), "ckrates_");
  g.stage = 1;
}

// SUBROUTINE SOLVESP(ISOR1,RELXSP,TOLRSP,SPNORM) (src/lib/solvesp.f:1)
void solvesp_(int32_t* isor1, double* relxsp, double* tolrsp, double* spnorm) {
  if (g.stage != 1) stop("solvesp_", "called without a preceding CKRATES");
  StepRun& R = g.ctx->run;
  if (*isor1 < 1 || *isor1 > 127) stop("solvesp_", "ISOR1 must be 1..127");
  if (!(*relxsp > 0.0 && *relxsp <= 1.0) || !(*spnorm > 0.0)) stop("solvesp_", "RELXSP or SPNORM out of range");
  R.ctl[0].isor1 = *isor1;
  R.ctl[0].relxsp = *relxsp;
  R.ctl[0].tolrsp = *tolrsp;
  R.spnorm[0] = *spnorm;
  GridTab& G = g.ctx->gt[0];
  G.sor1 = *relxsp;
  G.sor2 = 1.0 - *relxsp;
  G.tolrsp = *tolrsp;
  G.isor1 = *isor1;
  check(ph_species(g.ctx), "solvesp_");
  *isor1 = R.out[0].isor1;
  const int64_t npad = g.ctx->lay.npad, nb = blocks_for(npad, 256);
  k_scatter_rows<<<nb, 256, 0, g.s>>>(g.ctx->sa, npad, g.ctx->cells, shim::LSP, g.ssp, shim::LE);
  k_scatter_rows<<<nb, 256, 0, g.s>>>(g.ctx->sst, npad, g.ctx->cells, 2, g.sst, shim::LE);
  // the closure also sets the old-level N2, FCB = SCB (solvesp.f:2912)
  k_scatter_rows<<<nb, 256, 0, g.s>>>(g.ctx->sa + 51 * npad, npad, g.ctx->cells, 1, g.y + 51 * shim::LE,
                                      shim::LE);
  hipcheck(hipGetLastError(), "solvesp_");
  const size_t b1 = sizeof(double) * shim::LE;
  d2h(dumy_.ssp, g.ssp, b1 * shim::LSP, "solvesp_");
  d2h(cb04_.fsp[51], g.y + 51 * shim::LE, b1, "solvesp_");
  d2h(dums_.sst1, g.sst, b1, "solvesp_");
  d2h(dums_.sst2, g.sst + shim::LE, b1, "solvesp_");
  hipcheck(hipStreamSynchronize(g.s), "solvesp_");
  g.stage = 2;
}

// SUBROUTINE SOLVEH(ISOR2,RELXH,TOLRH,SIGSP,HNORM) (src/lib/solveh.f:1); SIGSP enters
// only the species-diffusion enthalpy flux, zero under the chemistry-only contract
void solveh_(int32_t* isor2, double* relxh, double* tolrh, double* sigsp, double* hnorm) {
  (void)sigsp;
  if (g.stage != 2) stop("solveh_", "called without a preceding SOLVESP");
  StepRun& R = g.ctx->run;
  if (*isor2 < 1 || *isor2 > 1000) stop("solveh_", "ISOR2 must be 1..1000");
  if (!(*relxh > 0.0 && *relxh <= 1.0) || !(*hnorm > 0.0)) stop("solveh_", "RELXH or HNORM out of range");
  R.ctl[0].isor2 = *isor2;
  R.ctl[0].relxh = *relxh;
  R.ctl[0].tolrh = *tolrh;
  R.hnorm[0] = *hnorm;
  GridTab& G = g.ctx->gt[0];
  G.zsor1 = *relxh;
  G.zsor2 = 1.0 - *relxh;
  G.tolrh = *tolrh;
  G.isor2 = *isor2;
  check(ph_enthalpy(g.ctx), "solveh_");
  check(ph_finish(g.ctx, g.fpz), "solveh_");
  uc_stats st{};
  check(ph_end(g.ctx, &st), "solveh_");
  *isor2 = st.isor2;
  const size_t b1 = sizeof(double) * shim::LE;
  d2h(dumy_.fpz, g.fpz, b1, "solveh_");
  d2h(cb04_.tk, g.tk, b1, "solveh_");
  d2h(hrls_.qdot, g.qdot, b1, "solveh_");
  d2h(dumd_.qdsp, g.qdsp, b1 * shim::LSP, "solveh_");
  hipcheck(hipStreamSynchronize(g.s), "solveh_");
  g.stage = 0;
}

}  // extern "C"
