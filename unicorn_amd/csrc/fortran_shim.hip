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
//   ckrates_  upload the old-level state (SA, SST1/SST2 and FPZ are taken as SPGSOLV's
//             prologue leaves them, FA, STMF/STND and HT, spgsolv.f:169-231, and copied on
//             the device); rate coefficients, rho*Y and soot sources
//             into the device slabs (ckrates.f:92-1352).  RC stays on the device:
//             only SOLVESP and SOLVEH read it (/DUMR/ is scratch elsewhere).
//   solvesp_  sweeps to the global stop, N2 closure (solvesp.f:83-2914); also the
//             SOLVEH sources, which depend only on RC, the final iterates, T and rho.
//             Writes /DUMY/ SSP, /DUMS/ SST1 SST2, /CB04/ FCB (N2) and ISOR1.
//   solveh_   enthalpy relaxation to the global stop and T(h) (solveh.f:1147-1234).
//             Writes /DUMY/ FPZ, /CB04/ TK, /HRLS/ QDOT, /DUMD/ QDSP and ISOR2.
//             Called alone (ICHEM <= 0 with ITHRM > 0, spgsolv.f:1291-1298), it uploads
//             the state and reads the rate coefficients /DUMR/ holds, as SOLVEH does.
//
// Coupling.  SPGSOLV's coefficient assembly stays in the Fortran; CKRATES reads what it
// left in the COMMON blocks: /DUMMY/ and /DUM0/ E,W,P,Q of every species, /DUMS/
// STE..STS, /DUMX/ ZE..ZS, the RDHSP inputs (/DUMY/ EMU, EKT, /CB07/ VDF, TMU), the
// geometry (/CB03/, ISYM) and the velocities (/CB04/ U, V, W).  Where they are zero at
// every cell the chemistry advances (the flux inputs: everywhere), the step takes the
// chemistry-only path; otherwise the coupled one (uc_step_coupled's phases: the
// reference's Gauss-Seidel order as a wavefront), exactly as the reference.  There is no
// CPU fallback: errors of the device layer stop the program after printing
// uc_last_error().  Cells with 3 <= ISKIP < 10 keep their QDOT/QDSP: the reference
// evaluates them there from RC values CKRATES did not compute in that step.
#include "chem_kernels.hip"

#include <atomic>
#include <cstdlib>
#include <thread>

#ifndef UC_LI
#define UC_LI 711
#endif
#ifndef UC_LJ
#define UC_LJ 131
#endif

namespace shim {
constexpr int64_t LI = UC_LI, LJ = UC_LJ, LE = LI * LJ;
constexpr int LSP = 52;
constexpr int LRX = 544;
}  // namespace shim

// COMMON blocks, as the reference declares them (src/lib/ckrates.f:6-29,
// solvesp.f:24-82, solveh.f:20-33, spgsolv.f:82); only the members used here.
extern "C" {
extern struct { int32_t isym, iflow, iswrl, ithrm, ichem, igrav, itr; int32_t iskip[shim::LE]; } cb01_;
extern struct { double rref, vref, tref, pref, wm[shim::LSP]; } cb02_;
extern struct { double dt, x[shim::LE], y[shim::LE], xxc[shim::LI], xxs[shim::LI], yyc[shim::LJ], yys[shim::LJ]; } cb03_;
extern struct { double rho[shim::LE], fsp[shim::LSP][shim::LE], u[shim::LE], v[shim::LE], w[shim::LE],
                p[shim::LE], ht[shim::LE], tk[shim::LE]; } cb04_;
extern struct { double alstr, vstr, rstr, pstr, tstr, tmstr, amstr, acstr, dstr, hstr, cpstr, akstr,
                epstr, wmstr, gasc, beta1, beta2, beta3, beta4, rei; } cb06_;
extern struct { double hfo[shim::LSP]; } hform_;
extern struct { double qdot[shim::LE]; } hrls_;
extern struct { double qdsp[shim::LSP][shim::LE]; } dumd_;
extern struct { double rsoot, stmf[shim::LE], stnd[shim::LE], stdf[shim::LE]; } soot_;
extern struct { double a[55][shim::LE]; } dummy_;
extern struct { double a[149][shim::LE]; } dum0_;
extern struct { double z[4][shim::LE]; } dumx_;
extern struct { double ssp[shim::LSP][shim::LE], fpz[shim::LE], emu[shim::LE], ekt[shim::LE]; } dumy_;
extern struct { double vmu[shim::LE], vtc[shim::LE], vdf[shim::LSP][shim::LE], tmu[shim::LE]; } cb07_;
extern struct { double st[4][shim::LE], sst1[shim::LE], sst2[shim::LE]; } dums_;
extern struct { double rc[shim::LRX][shim::LE]; } dumr_;
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
  // coupling inputs on the device (allocated at first use) and what this step uses
  double *coef = nullptr, *stc = nullptr, *zc = nullptr, *emu = nullptr, *ekt = nullptr, *tmu = nullptr,
         *vdf = nullptr, *ygeo = nullptr, *xxc = nullptr, *xxs = nullptr, *yyc = nullptr, *yys = nullptr,
         *u = nullptr, *v = nullptr, *w = nullptr;
  double *vtc = nullptr, *stdf = nullptr;  // spgsolv_'s further assembly inputs
  double* rcf = nullptr;   // /DUMR/ RC for SOLVEH called alone (ICHEM <= 0)
  uc_coupling cp{};
  bool coupled = false;
  // QDOT/QDSP on the device mirror the COMMONs once uploaded: only SOLVEH writes them
  // (solveh.f:1138,1142; every other routine only reads them) and solveh_ downloads both
  // whole, so later calls need not upload them again (UNICORN_FSHIM_UPLOAD_QDOT=1 in the
  // environment uploads them on every call, for a program that writes them elsewhere)
  bool q_resident = false;
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

// x under the Fortran edit descriptor D12.5: [-]0.ddddd, then D and a signed two-digit
// exponent (a signed three-digit one without the D beyond 99), right-justified in 12
std::string fortran_d12_5(double x) {
  int e = 0;
  std::string digits = "00000";
  if (x != 0.0) {
    char b[32];
    std::snprintf(b, sizeof b, "%.4e", std::fabs(x));   // d.dddde+XX: the 5 significant digits
    digits = std::string(1, b[0]) + std::string(b + 2, 4);
    e = std::atoi(b + 7) + 1;
  }
  char ex[8];
  if (std::abs(e) <= 99) std::snprintf(ex, sizeof ex, "D%c%02d", e < 0 ? '-' : '+', std::abs(e));
  else std::snprintf(ex, sizeof ex, "%c%03d", e < 0 ? '-' : '+', std::abs(e));
  std::string s = std::string(std::signbit(x) && x != 0.0 ? "-" : "") + "0." + digits + ex;
  return s.size() < 12 ? std::string(12 - s.size(), ' ') + s : s;
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
  g.q_resident = false;
  check(uc_create(&g.sc, shim::LE, &g.ctx), "ckrates_");
  // /DUMR/ RC as the reference keeps it: SOLVEH forms QDOT/QDSP of 3 <= ISKIP < 10 cells
  // from the coefficients of the last step that advanced them (solveh.f:119, ckrates.f:94)
  check(uc_keep_rates(g.ctx, 1), "ckrates_");
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

// Which coupling inputs this step has (host COMMONs only): coefficients and velocities
// matter where the chemistry advances; the flux inputs (RDHSP = (EMU VDF + REI TMU/SIGSP -
// EKT) HSSP, solveh.f:52-53) everywhere, because the flux reads neighbours.
struct Present {
  bool coef = false, st = false, z = false, flux = false, vel = false;
};

// A chemistry-only run reads every coefficient array to find them all zero (about 150 MB at
// the reference grid), so the arrays are scanned by several threads.
Present scan_coupling() {
  auto nonzero = [](const double* a) {
    for (int64_t c = 0; c < shim::LE; ++c)
      if (a[c] != 0.0 && active(c)) return true;
    return false;
  };
  // tasks 0..203: species coefficient arrays; 204..207 STE..STS; 208..211 ZE..ZS;
  // 212..214 U, V, W; 215: the flux inputs
  constexpr int NTASK = 216;
  std::atomic<int> next{0};
  std::atomic<bool> coef{false}, st{false}, z{false}, vel{false}, flux{false};
  auto work = [&] {
    for (int t; (t = next.fetch_add(1)) < NTASK;) {
      if (t < 204) {
        if (!coef && nonzero(t < 55 ? dummy_.a[t] : dum0_.a[t - 55])) coef = true;
      } else if (t < 208) {
        if (!st && nonzero(dums_.st[t - 204])) st = true;
      } else if (t < 212) {
        if (!z && nonzero(dumx_.z[t - 208])) z = true;
      } else if (t < 215) {
        const double* a = t == 212 ? cb04_.u : t == 213 ? cb04_.v : cb04_.w;
        if (!vel && nonzero(a)) vel = true;
      } else {
        bool f = false;
        for (int64_t c = 0; c < shim::LE && !f; ++c) {
          f = dumy_.ekt[c] != 0.0 || cb07_.tmu[c] != 0.0;
          for (int k = 0; k < shim::LSP && !f && dumy_.emu[c] != 0.0; ++k) f = cb07_.vdf[k][c] != 0.0;
        }
        if (f) flux = true;
      }
    }
  };
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned i = 1; i < nt; ++i) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  Present p;
  p.coef = coef; p.st = st; p.z = z; p.vel = vel; p.flux = flux;
  return p;
}

void h2d(void* dst, const void* src, size_t b, const char* where) {
  hipcheck(hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, g.s), where);
}
void d2h(void* dst, const void* src, size_t b, const char* where) {
  hipcheck(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToHost, g.s), where);
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

double* dalloc(double*& p, size_t bytes, const char* where) {
  if (!p) hipcheck(hipMalloc(&p, bytes), where);
  return p;
}

// the state CKRATES reads, and SPGSOLV's prologue copies (spgsolv.f:169-231) made on the
// device: SA = FA, SST1 = STMF, SST2 = STND and FPZ = HT at every cell
void upload_state() {
  const size_t b1 = sizeof(double) * shim::LE, b52 = b1 * shim::LSP;
  h2d(g.rho, cb04_.rho, b1, "ckrates_");
  h2d(g.y, cb04_.fsp, b52, "ckrates_");
  h2d(g.tk, cb04_.tk, b1, "ckrates_");
  h2d(g.ht, dumy_.fpz, b1, "ckrates_");   // FPZ = HT on entry
  h2d(g.stmf, soot_.stmf, b1, "ckrates_");
  h2d(g.stnd, soot_.stnd, b1, "ckrates_");
  h2d(g.iskip, cb01_.iskip, sizeof(int32_t) * shim::LE, "ckrates_");
  // SPGSOLV's prologue (spgsolv.f:169-231) set SA = FA, SST1 = STMF, SST2 = STND and
  // FPZ = HT at every cell just before this call: those copies are made on the device
  hipcheck(hipMemcpyAsync(g.ssp, g.y, b52, hipMemcpyDeviceToDevice, g.s), "ckrates_");
  hipcheck(hipMemcpyAsync(g.sst, g.stmf, b1, hipMemcpyDeviceToDevice, g.s), "ckrates_");
  hipcheck(hipMemcpyAsync(g.sst + shim::LE, g.stnd, b1, hipMemcpyDeviceToDevice, g.s), "ckrates_");
  hipcheck(hipMemcpyAsync(g.fpz, g.ht, b1, hipMemcpyDeviceToDevice, g.s), "ckrates_");
  static const bool q_always = [] {
    const char* e = std::getenv("UNICORN_FSHIM_UPLOAD_QDOT");
    return e && *e && *e != '0';
  }();
  if (!g.q_resident || q_always) {
    h2d(g.qdot, hrls_.qdot, b1, "ckrates_");
    h2d(g.qdsp, dumd_.qdsp, b52, "ckrates_");
    g.q_resident = true;
  }
}

// CKRATES on the device after the coupling inputs are set: the step's cell list, the
// rate coefficients
void begin_rates(const char* where) {
  hipcheck(hipStreamSynchronize(g.s), where);
  g.ctx->lay.ngrids = 0;  // ISKIP may change between steps: rebuild the cell list
  // SOLVESP/SOLVEH supply the controls and norms; these are placeholders until then
  const uc_controls ctl{cb03_.dt, 0.99, 1e-8, 0.99, 1e-8, 100, 100, UC_MODE_GLOBAL, 0};
  check(ph_begin(g.ctx, &g.f, &ctl, 1, g.s, false, nullptr, nullptr, g.coupled ? &g.cp : nullptr), where);
  check(ph_rates(g.ctx), where);
  g.stage = 1;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// the coupling inputs this step has, uploaded from the COMMONs SPGSOLV filled
void upload_coupling(const Present& pr) {
  const size_t b1 = sizeof(double) * shim::LE, b52 = b1 * shim::LSP;
  // the coupling inputs this step has
  auto dev = [&](double*& p, size_t bytes) { return dalloc(p, bytes, "ckrates_: hipMalloc"); };
  g.cp = uc_coupling{};
  if (pr.coef) {
    dev(g.coef, 204 * b1);
    h2d(g.coef, dummy_.a, 55 * b1, "ckrates_");   // EA0..PAD
    h2d(g.coef + 55 * shim::LE, dum0_.a, 149 * b1, "ckrates_");  // QAD..QCA
    g.cp.coef = g.coef;
  }
  if (pr.st) { h2d(dev(g.stc, 4 * b1), dums_.st, 4 * b1, "ckrates_"); g.cp.stc = g.stc; }
  if (pr.z) { h2d(dev(g.zc, 4 * b1), dumx_.z, 4 * b1, "ckrates_"); g.cp.zc = g.zc; }
  if (pr.flux) {
    h2d(dev(g.emu, b1), dumy_.emu, b1, "ckrates_");
    h2d(dev(g.ekt, b1), dumy_.ekt, b1, "ckrates_");
    h2d(dev(g.tmu, b1), cb07_.tmu, b1, "ckrates_");
    h2d(dev(g.vdf, b52), cb07_.vdf, b52, "ckrates_");
    h2d(dev(g.ygeo, b1), cb03_.y, b1, "ckrates_");
    h2d(dev(g.xxc, sizeof(double) * shim::LI), cb03_.xxc, sizeof(double) * shim::LI, "ckrates_");
    h2d(dev(g.xxs, sizeof(double) * shim::LI), cb03_.xxs, sizeof(double) * shim::LI, "ckrates_");
    h2d(dev(g.yyc, sizeof(double) * shim::LJ), cb03_.yyc, sizeof(double) * shim::LJ, "ckrates_");
    h2d(dev(g.yys, sizeof(double) * shim::LJ), cb03_.yys, sizeof(double) * shim::LJ, "ckrates_");
    g.cp.emu = g.emu; g.cp.ekt = g.ekt; g.cp.tmu = g.tmu; g.cp.vdf = g.vdf;
    g.cp.ygeo = g.ygeo; g.cp.xxc = g.xxc; g.cp.xxs = g.xxs; g.cp.yyc = g.yyc; g.cp.yys = g.yys;
  }
  if (pr.vel) {
    h2d(dev(g.u, b1), cb04_.u, b1, "ckrates_");
    h2d(dev(g.v, b1), cb04_.v, b1, "ckrates_");
    h2d(dev(g.w, b1), cb04_.w, b1, "ckrates_");
    g.cp.u = g.u; g.cp.v = g.v; g.cp.w = g.w;
  }
  g.cp.sigsp = 0.9;  // SOLVEH's argument, set there
  g.cp.isym = cb01_.isym;
  g.coupled = pr.coef || pr.st || pr.z || pr.flux || pr.vel;
}

}  // namespace

extern "C" {

// SUBROUTINE CKRATES (src/lib/ckrates.f:1)
void ckrates_(void) {
  const Present pr = scan_coupling();  // host COMMONs only
  ensure_context();
  upload_state();
  upload_coupling(pr);
  begin_rates("ckrates_");
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
  G.spnorm = *spnorm;
  const size_t b1 = sizeof(double) * shim::LE;
  if (R.cpl_sp) {  // the closed iterates of every cell are in the field-layout iterate array
    check(ph_species_cpl(g.ctx), "solvesp_");
    *isor1 = R.out[0].isor1;
    const double* it = g.ctx->cpl_it;
    d2h(dumy_.ssp, it, b1 * shim::LSP, "solvesp_");
    d2h(cb04_.fsp[51], it + 51 * shim::LE, b1, "solvesp_");  // FCB = SCB (solvesp.f:2912)
    d2h(dums_.sst1, it + 52 * shim::LE, b1, "solvesp_");
    d2h(dums_.sst2, it + 53 * shim::LE, b1, "solvesp_");
    hipcheck(hipStreamSynchronize(g.s), "solvesp_");
    g.stage = 2;
    return;
  }
  check(ph_species(g.ctx), "solvesp_");
  *isor1 = R.out[0].isor1;
  const int64_t npad = g.ctx->lay.npad, nb = blocks_for(npad, 256);
  k_scatter_rows<<<nb, 256, 0, g.s>>>(g.ctx->sa, npad, g.ctx->cells, shim::LSP, g.ssp, shim::LE);
  k_scatter_rows<<<nb, 256, 0, g.s>>>(g.ctx->sst, npad, g.ctx->cells, 2, g.sst, shim::LE);
  // the closure also sets the old-level N2, FCB = SCB (solvesp.f:2912)
  k_scatter_rows<<<nb, 256, 0, g.s>>>(g.ctx->sa + 51 * npad, npad, g.ctx->cells, 1, g.y + 51 * shim::LE,
                                      shim::LE);
  hipcheck(hipGetLastError(), "solvesp_");
  d2h(dumy_.ssp, g.ssp, b1 * shim::LSP, "solvesp_");
  d2h(cb04_.fsp[51], g.y + 51 * shim::LE, b1, "solvesp_");
  d2h(dums_.sst1, g.sst, b1, "solvesp_");
  d2h(dums_.sst2, g.sst + shim::LE, b1, "solvesp_");
  hipcheck(hipStreamSynchronize(g.s), "solvesp_");
  g.stage = 2;
}

// SUBROUTINE SOLVEH(ISOR2,RELXH,TOLRH,SIGSP,HNORM) (src/lib/solveh.f:1); SIGSP enters
// the species-diffusion enthalpy flux (solveh.f:52-53)
void solveh_(int32_t* isor2, double* relxh, double* tolrh, double* sigsp, double* hnorm) {
  // SPGSOLV calls SOLVEH alone when ICHEM <= 0 and ITHRM > 0 (spgsolv.f:1291-1298,
  // main:1027): the species are the prologue copies SSP = FA, and the sources read whatever
  // rate coefficients /DUMR/ holds (zero if CKRATES never ran)
  const bool alone = g.stage == 0;
  if (alone) {
    if (cb01_.ichem > 0) stop("solveh_", "called without a preceding CKRATES and SOLVESP while ICHEM > 0");
    const Present pr = scan_coupling();
    ensure_context();
    upload_state();
    upload_coupling(pr);
    const size_t brc = sizeof(double) * shim::LE * shim::LRX;
    h2d(dalloc(g.rcf, brc, "solveh_: hipMalloc"), dumr_.rc, brc, "solveh_");
    hipcheck(hipStreamSynchronize(g.s), "solveh_");
    g.ctx->lay.ngrids = 0;  // ISKIP may change between steps: rebuild the cell list
    const uc_controls ctl{cb03_.dt, 0.99, 1e-8, *relxh, *tolrh, 100, *isor2 >= 1 ? *isor2 : 1, UC_MODE_GLOBAL, 0};
    check(ph_begin(g.ctx, &g.f, &ctl, 1, g.s, false, nullptr, nullptr, g.coupled ? &g.cp : nullptr), "solveh_");
    g.stage = 2;
  } else if (g.stage != 2) {
    stop("solveh_", "called without a preceding SOLVESP");
  }
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
  G.hnorm = *hnorm;
  if (alone) {
    if (R.cpl_flux && !(*sigsp > 0.0)) stop("solveh_", "SIGSP must be positive");
    R.cp.sigsp = *sigsp;
    check(ph_solveh_sources(g.ctx, g.rcf), "solveh_");
  } else if (R.cpl_sp) {  // the sources with the enthalpy flux, which needs SIGSP
    if (R.cpl_flux && !(*sigsp > 0.0)) stop("solveh_", "SIGSP must be positive");
    R.cp.sigsp = *sigsp;
    check(ph_sources_cpl(g.ctx), "solveh_");
  }
  check(R.cpl_h ? ph_enthalpy_cpl(g.ctx) : ph_enthalpy(g.ctx), "solveh_");
  check(ph_finish(g.ctx, g.fpz), "solveh_");
  uc_stats st{};
  check(ph_end(g.ctx, &st), "solveh_");
  *isor2 = st.isor2;
  if (st.newton_failed > 0) {
    // the reference's messages (solveh.f:1170-1227): over the cells in I-then-J order,
    // IOVER starts at 1, a cell above the high-range fit sets it to 1, and a cell whose
    // Newton iteration fails prints FORMAT 714 with its HT (which SOLVEH leaves as it was)
    // while IOVER <= 1, then sets it to 2.  The DEBUG(9427) line the reference prints for
    // every failing cell is not reproduced (uc_stats counts those cells).
    std::vector<uint8_t> fl(shim::LE);
    d2h(fl.data(), g.ctx->thflag, shim::LE, "solveh_");
    hipcheck(hipStreamSynchronize(g.s), "solveh_");
    int iover = 1;
    std::fflush(stdout);
    for (int64_t i = 1; i < shim::LI - 1; ++i)
      for (int64_t j = 1; j < shim::LJ - 1; ++j) {
        const int64_t c = i * shim::LJ + j;
        if (cb01_.iskip[c] >= 3) continue;
        if (fl[c] & 1) iover = 1;
        if (fl[c] & 2) {
          if (iover <= 1) std::printf("     UNABLE TO FIND THE ROOT%4d  %s\n", (int)(i + 1), fortran_d12_5(cb04_.ht[c]).c_str());
          iover = 2;
        }
      }
    std::fflush(stdout);
  }
  const size_t b1 = sizeof(double) * shim::LE;
  d2h(dumy_.fpz, g.fpz, b1, "solveh_");
  d2h(cb04_.tk, g.tk, b1, "solveh_");
  d2h(hrls_.qdot, g.qdot, b1, "solveh_");
  d2h(dumd_.qdsp, g.qdsp, b1 * shim::LSP, "solveh_");
  hipcheck(hipStreamSynchronize(g.s), "solveh_");
  g.stage = 0;
}

// SUBROUTINE SPGSOLV(ISOR1,RELXSP,TOLRSP,RESDC,ISOR2,RELXH,TOLRH,RESDH,SIGH,SIGSP)
// (src/lib/spgsolv.f:6), for a Fortran build that links this library in place of the whole
// of spgsolv.f: the coefficient assembly (spgsolv.f:98-1290) runs on the device from the
// transport COMMONs (/CB07/ VTC, TMU, VDF; /SOOT/ STDF; /CB04/ U, V; /CB03/ geometry), then
// CKRATES, SOLVESP and SOLVEH as above with those coefficients, then the epilogue
// (spgsolv.f:1299-1416) on the host COMMONs.  The scratch COMMONs SPGSOLV fills on the way
// (/DUMMY/, /DUM0/, /DUMX/, /DUMS/ STE..STS, /DUMY/ EMU and EKT) are not written: the main
// program reuses them as work arrays.  Needs ICHEM > 0 and ITHRM > 0.
void spgsolv_(int32_t* isor1, double* relxsp, double* tolrsp, double* resdc, int32_t* isor2, double* relxh,
              double* tolrh, double* resdh, double* sigh, double* sigsp) {
  if (cb01_.ichem <= 0 || cb01_.ithrm <= 0) stop("spgsolv_", "the device SPGSOLV needs ICHEM > 0 and ITHRM > 0");
  ensure_context();
  // SPGSOLV's FPZ = HT at every cell (spgsolv.f:176), which CKRATES's upload reads as HT
  std::memcpy(dumy_.fpz, cb04_.ht, sizeof(double) * shim::LE);
  upload_state();
  const size_t b1 = sizeof(double) * shim::LE, b52 = b1 * shim::LSP;
  const size_t bi = sizeof(double) * shim::LI, bj = sizeof(double) * shim::LJ;
  const char* W = "spgsolv_";
  h2d(dalloc(g.vtc, b1, W), cb07_.vtc, b1, W);
  h2d(dalloc(g.tmu, b1, W), cb07_.tmu, b1, W);
  h2d(dalloc(g.vdf, b52, W), cb07_.vdf, b52, W);
  h2d(dalloc(g.stdf, b1, W), soot_.stdf, b1, W);
  h2d(dalloc(g.u, b1, W), cb04_.u, b1, W);
  h2d(dalloc(g.v, b1, W), cb04_.v, b1, W);
  h2d(dalloc(g.w, b1, W), cb04_.w, b1, W);
  h2d(dalloc(g.ygeo, b1, W), cb03_.y, b1, W);
  h2d(dalloc(g.xxc, bi, W), cb03_.xxc, bi, W);
  h2d(dalloc(g.xxs, bi, W), cb03_.xxs, bi, W);
  h2d(dalloc(g.yyc, bj, W), cb03_.yyc, bj, W);
  h2d(dalloc(g.yys, bj, W), cb03_.yys, bj, W);
  // the assembly writes the interior cells with ISKIP < 10; elsewhere the device buffers
  // hold zeros (the reference leaves face conductances there, spgsolv.f:233-348, which only
  // cells with ISKIP < 3 read, so no result depends on them)
  hipcheck(hipMemsetAsync(dalloc(g.coef, 204 * b1, W), 0, 204 * b1, g.s), W);
  hipcheck(hipMemsetAsync(dalloc(g.stc, 4 * b1, W), 0, 4 * b1, g.s), W);
  hipcheck(hipMemsetAsync(dalloc(g.zc, 4 * b1, W), 0, 4 * b1, g.s), W);
  dalloc(g.ekt, b1, W);
  const uc_transport tr{g.vtc, g.tmu, g.vdf, g.stdf, g.u, g.v, g.ygeo, g.xxc, g.xxs, g.yyc, g.yys, *sigsp, *sigh,
                        cb01_.isym};
  const uc_controls ctl{cb03_.dt, 0.99, 1e-8, 0.99, 1e-8, 100, 100, UC_MODE_GLOBAL, 0};
  check(uc_assemble(g.ctx, &g.f, &ctl, &tr, g.coef, g.stc, g.zc, g.ekt, g.s), W);
  // EMU null: REI*BETA3*RHO on the device, as spgsolv.f:175 sets it
  g.cp = uc_coupling{g.coef, g.stc, g.zc, nullptr, g.ekt, g.tmu, g.vdf, g.ygeo, g.xxc, g.xxs, g.yyc, g.yys,
                     g.u, g.v, g.w, *sigsp, cb01_.isym};
  g.coupled = true;
  begin_rates(W);
  double spnorm = (double)shim::LE * (double)(shim::LSP - 1), hnorm = (double)shim::LE;  // spgsolv.f:100-101
  solvesp_(isor1, relxsp, tolrsp, &spnorm);
  solveh_(isor2, relxh, tolrh, sigsp, &hnorm);
  // spgsolv.f:1299-1416: residuals over the cells that advanced, then the new state
  double rh = 0.0, rc = 0.0;
  for (int64_t i = 1; i < shim::LI - 1; ++i)
    for (int64_t j = 1; j < shim::LJ - 1; ++j) {
      const int64_t c = i * shim::LJ + j;
      if (cb01_.iskip[c] >= 3) continue;
      rh = rh + std::fabs(dumy_.fpz[c] - cb04_.ht[c]);
      rc = rc + std::fabs(dums_.sst1[c] - soot_.stmf[c]);
      for (int k = 0; k < shim::LSP - 1; ++k) rc = rc + std::fabs(dumy_.ssp[k][c] - cb04_.fsp[k][c]);
      cb04_.ht[c] = dumy_.fpz[c];
      soot_.stmf[c] = dums_.sst1[c];
      soot_.stnd[c] = dums_.sst2[c];
      for (int k = 0; k < shim::LSP - 1; ++k) cb04_.fsp[k][c] = dumy_.ssp[k][c];
    }
  *resdh = rh / hnorm;
  *resdc = rc / spnorm;
}

}  // extern "C"
