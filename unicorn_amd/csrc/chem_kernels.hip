// MI355X (gfx950) kernels and host C ABI of UNICORN's chemistry step.
//
// Step = main:821-828 clip/closure -> CKRATES -> SOLVESP -> SOLVEH -> SPGSOLV
// epilogue -> main:1056-1073 density, for the chemistry-only contract.
//
// Kernels (one thread per cell, FP64 throughout):
//   k_clip      all cells: Y<0 -> 0, Y_N2 = 1 - sum(Y_1..51)            main:821-828
//   k_chem      fused, per active interior cell: rate coefficients into the
//               thread's RC slab, Gauss-Seidel sweeps with the slab re-read
//               each sweep, N2 closure, SOLVEH sources (radiation, RA/PRD,
//               QDOT, QDSP, RHSZ).  A cell stops sweeping once a sweep leaves
//               every iterate bit-for-bit unchanged (later sweeps would repeat
//               it exactly); per-sweep residual sums are accumulated per wave
//               so the host can find the reference's global stop N*.
//   k_hres      per cell: enthalpy relaxation residual trajectory            solveh.f:1147-1167
//   k_finish    per cell: enthalpy relaxation with the decided sweep count,
//               Newton T(h), soot number density completion, SPGSOLV
//               copy-back and residuals, density                             solveh.f:1170-1230
//   k_density / k_density_edge  density of the remaining cells + extrapolation main:1056-1068
//
// RC slabs: thread t of a k_chem launch owns rc[s*G + t], s = 0..543 (G
// threads in the launch), so every RC load/store of a wave is one contiguous
// 512-byte segment.  A slab is written by the rate stage and read back by
// every sweep of the same thread, so it stays in the L2 / Infinity Cache.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/unicorn_chem.h"

namespace {

constexpr int NSP = 52;
constexpr int NRX = 544;
constexpr int WAVE = 64;

struct UcParams {
  double tstr, alstr, vstr, rstr, wmstr, hstr, gasc, pback, beta1;
  double wm[NSP], rwm[NSP], hfo[NSP];
  double nasa[NSP][14];
  double rcfac[NRX];  // C_m / prod W of the slot's reactants (ckrates.f:83-85, :796-1333)
  double dtw[NSP];    // DT * WM(k) of the current step
};

typedef const __attribute__((address_space(4))) UcParams* UcP;
typedef __attribute__((ext_vector_type(2))) unsigned int u32x2;

// A wave's slab tile: slot s of lane l at tile[s*64 + l].  Accessed with buffer
// instructions: one lane-offset VGPR shared by all slots, the slot offset as a
// scalar constant, and the 128-bit resource in SGPRs.  Out-of-range offsets
// read as zero / are dropped by the hardware bounds check.
struct RcSlab {
  __amdgpu_buffer_rsrc_t r;
  uint32_t lane8;
  __device__ __forceinline__ double ld(int s) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, lane8, s * 512, 0));
  }
  __device__ __forceinline__ void st(int s, double v) const {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, lane8, s * 512, 0);
  }
  // opaque redefinition: later slab loads can be neither merged with earlier ones
  // nor hoisted out of the sweep loop
  __device__ __forceinline__ void fence() { asm volatile("" : "+s"(r)); }
};

__device__ __forceinline__ uint64_t wave_uniform(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

constexpr int SLAB_SLOTS = 664;  // = UC_SLAB_SLOTS of chem_gen.inc

__device__ __forceinline__ RcSlab wave_slab(double* slabs) {
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  double* tile = (double*)wave_uniform((uint64_t)(slabs + wave * (uint64_t)SLAB_SLOTS * 64));
  return RcSlab{__builtin_amdgcn_make_buffer_rsrc(tile, 0, SLAB_SLOTS * 512, 0x00020000),
                (threadIdx.x & 63u) * 8u};
}

// A field array ([k][n] species-major, or [n]) of one cell: the cell's byte
// offset is the only per-lane part; k*n*8 goes in the scalar offset.
struct FieldBuf {
  __amdgpu_buffer_rsrc_t r;
  uint32_t c8;
  __device__ __forceinline__ double ld(int64_t k_n) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, c8, (int)(k_n * 8), 0));
  }
  __device__ __forceinline__ void st(int64_t k_n, double v) const {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, c8, (int)(k_n * 8), 0);
  }
};

__device__ __forceinline__ FieldBuf field_buf(const double* base, int64_t bytes, int64_t c) {
  return FieldBuf{__builtin_amdgcn_make_buffer_rsrc((void*)base, 0, (int)bytes, 0x00020000), (uint32_t)(c * 8)};
}

#include "chem_gen.inc"
static_assert(UC_SLAB_SLOTS == SLAB_SLOTS, "slab layout");

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      return fail(UC_E_HIP, std::string(#x ": ") + hipGetErrorString(e_));               \
  } while (0)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

// ------------------------------------------------------------------ kernels
__global__ void k_clip(double* __restrict__ y, int64_t n) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double tmass = 0.0;
  for (int k = 0; k < NSP - 1; ++k) {
    double v = y[k * n + c];
    if (v < 0.0) {
      v = 0.0;
      y[k * n + c] = 0.0;
    }
    tmass += v;
  }
  y[(NSP - 1) * n + c] = 1.0 - tmass;
}

struct ChemArgs {
  const double* rho; const double* y; const double* tk; const double* ht;
  const double* stmf; const double* stnd;
  const int32_t* cells; int32_t ncells; int64_t n;
  double* sa;        // [52][n]   SA* after closure
  double* sst;       // [4][n]    SST1, SST2, SST2 before the last sweep, SST2 increment b
  int32_t* nsw;      // [n]       sweeps executed
  double* rhsz; double* qdot; double* qdsp;
  double* rcs;       // RC slabs, G*544
  double* rsd_part;  // [isor1][nwaves] per-wave sums of the per-sweep residual (may be null)
  int32_t nmax, mode;
  int32_t acc_from;  // residuals of sweeps >= acc_from are accumulated (when rsd_part is set)
  int32_t rsd_stride;  // row stride of rsd_part: waves of a full-size launch (>= this launch's)
  double dt, sor1, sor2, tolrsp;
};

// nsw word: sweeps executed, plus NSW_CAPPED when the cell stopped on the sweep
// cap instead of at a bitwise fixed point
constexpr int32_t NSW_CAPPED = 1 << 30;
constexpr int32_t NSW_MASK = NSW_CAPPED - 1;

__global__ void __launch_bounds__(256) k_chem(const UcParams* __restrict__ Pg, ChemArgs a) {
  UcP P = (UcP)Pg;
  const int64_t G = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = (int)(tid / WAVE);
  const int64_t rstride = a.rsd_stride;
  RcSlab rc = wave_slab(a.rcs);
  const int64_t n = a.n;
  const int64_t b1 = n * 8, b52 = n * 8 * NSP;
  // every lane of a wave runs the same number of outer iterations
  const int64_t iters = (a.ncells + G - 1) / G;
  for (int64_t it = 0; it < iters; ++it) {
    asm volatile("" : "+s"(P));  // parameter loads stay inside the cell loop
    const int64_t q = it * G + tid;
    const bool valid = q < a.ncells;
    const int64_t c = valid ? a.cells[q] : a.cells[0];
    const FieldBuf Y = field_buf(a.y, b52, c);
    const double rho = a.rho[c], tk = a.tk[c], ht = a.ht[c], stmf = a.stmf[c], stnd = a.stnd[c];
    double tbody;
    {
      double fa[NSP];
#pragma unroll
      for (int k = 0; k < NSP; ++k) fa[k] = Y.ld(k * n);
      uc_rates(P, tk, rho, fa, rc, tbody);
    }
    uc_soot(P, rc, tk, rho, stmf, stnd, tbody);
    double s[54];
#pragma unroll
    for (int k = 0; k < NSP; ++k) s[k] = Y.ld(k * n);  // SA = FA (spgsolv.f:169-220), re-read
    s[52] = stmf;
    s[53] = stnd;
    bool active = valid, fixed = false;
    int32_t nsw = 0;
    double sst2_prev = stnd;
    for (int sw = 0; sw < a.nmax; ++sw) {
      if (!__any(active)) break;
      double r = 0.0;
      if (active) {
        sst2_prev = s[53];
        double d2;
        r = uc_sweep(P, rc, s, rho, stmf, stnd, a.dt, a.sor1, a.sor2, &d2);
        ++nsw;
        if (a.mode == UC_MODE_PERCELL) {
          if (r / 51.0 <= a.tolrsp) active = false;
        } else if (r == 0.0) {
          active = false;  // bitwise fixed point: every later sweep repeats this one
          fixed = true;
        }
      }
      if (a.rsd_part && sw >= a.acc_from) {
        const double wsum = wave_sum(r);
        if ((threadIdx.x & (WAVE - 1)) == 0) a.rsd_part[sw * rstride + wave] += wsum;
      }
    }
    if (!valid) continue;
    // SST2 increment with the final species, before the closure rescales them (for
    // completing the decided sweep count later, solvesp.f:2820-2831)
    const double sst2_b = (rho * stnd + a.dt * rc.ld(UC_SLOT_SOUR + 1) * s[21]) * a.sor1
                          / (rho + a.dt * rc.ld(UC_SLOT_SOUR + 5));
    // N2 closure (solvesp.f:2839-2914)
    double ftot = 0.0;
#pragma unroll
    for (int k = 0; k < 51; ++k) ftot += s[k];
    ftot += s[52];
    s[51] = 1.0 - ftot;
    if (ftot > 1.0) {
#pragma unroll
      for (int k = 0; k < 51; ++k) s[k] = s[k] / ftot;
      s[52] = s[52] / ftot;
      s[51] = 0.0;
    }
    // SOLVEH sources with the final iterates (solveh.f:34-1142)
    const double T = tk * P->tstr;
    const double T2 = T * T, T3 = T * T2, T4 = T * T3, T5 = T * T4;
    const int ip = T > 1000.0 ? 7 : 0;
    const double hconvr = P->gasc / P->wmstr / P->hstr;
    double tmole = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) tmole += s[k] * P->rwm[k];
    const double tkdi = 1000.0 / T, tkdis = tkdi * tkdi;
    const double p1 = (double)6.6334f + T * (-3.5686e-03 + T * (1.6682e-08 + T * (2.5611e-10 - 2.6558e-14 * T)));
    const double p2 = 1.8741e+01 - tkdi * (1.2131e+02 - 2.7350e+02 * tkdi + 1.9405e+02 * tkdis
                                           - 5.6310e+01 * tkdis * tkdi + (double)5.81690f * tkdis * tkdis);
    const double p3 = T <= 750.0
        ? (double)4.7869f + T * (-6.953e-02 + T * (2.95775e-04 + T * (-4.25732e-07 + 2.02894e-10 * T)))
        : 1.009e+01 + T * (-1.183e-02 + T * (4.7753e-06 + T * (-5.87209e-10 - 2.5334e-14 * T)));
    const double p4 = -(double)0.23093f - tkdi * ((double)1.12390f - (double)9.41530f * tkdi
                      + (double)2.99880f * tkdis - (double)0.51382f * tkdis * tkdi + 1.86840e-05 * tkdis * tkdis);
    const double plankc = (p1 * s[12] * P->rwm[12] + p2 * s[9] * P->rwm[9] + p3 * s[8] * P->rwm[8]
                           + p4 * s[5] * P->rwm[5]) / tmole;
    const double crad = P->alstr / P->rstr / P->vstr / P->vstr / P->vstr;
    const double qrad = crad * (double)0.4f * (4.0 * 5.669e-08 * plankc * (pow(T, 4.0) - 8.1e+09)
                                              + 3.334e-04 * stmf * (rho / 1.9e+03) * T5);
    double rhsz = -qrad * a.dt;
    const double rconv = P->vstr / P->wmstr / P->alstr * 1.0e-06;
    double prod = 0.0, hrr = 0.0;
    const FieldBuf QD = field_buf(a.qdsp, b52, c);
    const bool hiT = ip != 0;
    uc_prd(P, rc, s, [&](UcP Q, int k, double pr) {
      // per-species enthalpy HENTH (solveh.f:39-57) with this cell's NASA range
      double pk[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) pk[i] = hiT ? Q->nasa[k][7 + i] : Q->nasa[k][i];
      const double hssp = (pk[0] * T + pk[1] * T2 / 2.0 + pk[2] * T3 / 3.0 + pk[3] * T4 / 4.0
                           + pk[4] * T5 / 5.0) * Q->rwm[k] * hconvr;
      const double henth = hssp * Q->wmstr * Q->hstr + pk[5] * Q->rwm[k] * Q->gasc;
      const double prd = Q->wm[k] * Q->hfo[k] * pr;
      prod += prd;
      hrr += prd * henth / Q->hfo[k];
      QD.st(k * n, -rconv * prd / Q->hfo[k]);
    });
    rhsz = rhsz + rho * ht - a.dt * prod;
    a.rhsz[c] = rhsz;
    a.qdot[c] = -hrr * rconv;
    const FieldBuf SA = field_buf(a.sa, b52, c);
#pragma unroll
    for (int k = 0; k < NSP; ++k) SA.st(k * n, s[k]);
    const FieldBuf SS = field_buf(a.sst, 4 * b1, c);
    SS.st(0, s[52]);
    SS.st(n, s[53]);
    SS.st(2 * n, sst2_prev);
    SS.st(3 * n, sst2_b);
    a.nsw[c] = nsw | (fixed ? 0 : NSW_CAPPED);
  }
}

struct HresArgs {
  const double* rho; const double* ht; const double* rhsz;
  const int32_t* cells; int32_t ncells;
  double* part;  // [isor2][nwaves]
  int32_t nmax;
  double zsor1, zsor2;
};

__global__ void __launch_bounds__(256) k_hres(HresArgs a) {
  const int64_t G = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = (int)(tid / WAVE), nwaves = (int)(G / WAVE);
  const int64_t iters = (a.ncells + G - 1) / G;
  for (int64_t it = 0; it < iters; ++it) {
    const int64_t q = it * G + tid;
    const bool valid = q < a.ncells;
    const int64_t c = valid ? a.cells[q] : a.cells[0];
    double fpz = a.ht[c];
    const double inc = a.rhsz[c] * a.zsor1 / a.rho[c];
    bool active = valid;
    for (int sw = 0; sw < a.nmax; ++sw) {
      if (!__any(active)) break;
      double d = 0.0;
      if (active) {
        const double fnew = a.zsor2 * fpz + inc;
        d = fabs(fpz - fnew);
        fpz = fnew;
        if (d == 0.0) active = false;
      }
      const double wsum = wave_sum(d);
      if ((threadIdx.x & (WAVE - 1)) == 0) a.part[(int64_t)sw * nwaves + wave] += wsum;
    }
  }
}

struct FinishArgs {
  double* rho; double* y; double* tk; double* ht; double* stmf; double* stnd;
  const double* sa; const double* sst; const int32_t* nsw; const double* rhsz;
  const int32_t* cells; int32_t ncells; int64_t n;
  int32_t nsweeps;   // decided species sweep count (global mode)
  int32_t nh;        // decided enthalpy sweep count (global mode) / cap (per-cell)
  int32_t mode;
  double zsor1, zsor2, tolrh, sor2;
  double* res_part;  // [2][nblocks] RESDC, RESDH partial sums
};

__global__ void __launch_bounds__(256) k_finish(const UcParams* __restrict__ Pg, FinishArgs a) {
  UcP P = (UcP)Pg;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rc_ = 0.0, rh_ = 0.0;
  if (q < a.ncells) {
    const int64_t c = a.cells[q], n = a.n;
    const FieldBuf SA = field_buf(a.sa, n * 8 * NSP, c), Y = field_buf(a.y, n * 8 * NSP, c);
    double s[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) s[k] = SA.ld(k * n);
    const double rho = a.rho[c], ht = a.ht[c];
    // enthalpy relaxation (solveh.f:1147-1167), zero transport: PPP = RHO
    double fpz = ht;
    const double inc = a.rhsz[c] * a.zsor1 / rho;
    for (int sw = 0; sw < a.nh; ++sw) {
      const double fnew = a.zsor2 * fpz + inc;
      const double d = fabs(fpz - fnew);
      fpz = fnew;
      if (a.mode == UC_MODE_PERCELL && d <= a.tolrh) break;
    }
    // T from h (solveh.f:1170-1230)
    double enth = fpz * P->hstr;
#pragma unroll
    for (int k = 0; k < NSP; ++k) enth += s[k] * P->hfo[k] * P->hstr;
    bool hiT = false;
    double tpoly = 1000.0;
    double tkd = a.tk[c] * P->tstr;
    double a1, a2, a3, a4, a5, a6;
    for (;;) {
      if (tkd > tpoly) tkd = tpoly / 2.0;
      a1 = a2 = a3 = a4 = a5 = a6 = 0.0;
      asm volatile("" : "+s"(P));
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        const double sw = s[k] * P->rwm[k];
        a1 += (hiT ? P->nasa[k][7] : P->nasa[k][0]) * sw;
        a2 += (hiT ? P->nasa[k][8] : P->nasa[k][1]) * sw / 2.0;
        a3 += (hiT ? P->nasa[k][9] : P->nasa[k][2]) * sw / 3.0;
        a4 += (hiT ? P->nasa[k][10] : P->nasa[k][3]) * sw / 4.0;
        a5 += (hiT ? P->nasa[k][11] : P->nasa[k][4]) * sw / 5.0;
        a6 += (hiT ? P->nasa[k][12] : P->nasa[k][5]) * sw;
      }
      const double funt = ((a1 + (a2 + (a3 + (a4 + a5 * tpoly) * tpoly) * tpoly) * tpoly) * tpoly + a6)
                          * P->gasc / P->wmstr - enth;
      if (funt >= 1.0e-07 || hiT) break;
      hiT = true;
      if (tkd <= tpoly) tkd = 5000.0 / 2.0;
      tpoly = 5000.0;
    }
    const double eh = enth * P->wmstr / P->gasc;
    for (int itn = 1; itn <= 100; ++itn) {
      const double dtkd = ((a1 + (a2 + (a3 + (a4 + a5 * tkd) * tkd) * tkd) * tkd) * tkd + a6 - eh)
                          / (a1 + (2.0 * a2 + (3.0 * a3 + (4.0 * a4 + 5.0 * a5 * tkd) * tkd) * tkd) * tkd);
      if (fabs(dtkd) <= (double)0.001f) break;
      tkd = tkd - dtkd;
    }
    if (tkd > 5000.0) tkd = 5000.0;
    if (tkd < P->tstr - 1.0) tkd = P->tstr - 1.0;
    const double tnew = tkd / P->tstr;
    // soot number density: complete the decided sweep count (solvesp.f:2820-2831)
    const FieldBuf SS = field_buf(a.sst, n * 8 * 4, c);
    double sst2 = SS.ld(n);
    if (a.mode != UC_MODE_PERCELL) {
      const int32_t done = a.nsw[c] & NSW_MASK;
      if (a.nsweeps == done - 1) {
        sst2 = SS.ld(2 * n);
      } else {
        const double b = SS.ld(3 * n);
        for (int sw = done; sw < a.nsweeps; ++sw) {
          const double fnew = a.sor2 * sst2 + b;
          sst2 = fnew < 0.0 ? 0.0 : fnew;
        }
      }
    }
    // SPGSOLV epilogue (spgsolv.f:1299-1416)
    const double sst1 = SS.ld(0);
    rh_ = fabs(fpz - ht);
    rc_ = fabs(sst1 - a.stmf[c]);
    double rbar = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      if (k < 51) rc_ += fabs(s[k] - Y.ld(k * n));
      Y.st(k * n, s[k]);
      rbar += s[k] * P->rwm[k];
    }
    a.ht[c] = fpz;
    a.stmf[c] = sst1;
    a.stnd[c] = sst2;
    a.tk[c] = tnew;
    a.rho[c] = P->pback / (P->beta1 * tnew * rbar);  // main:1056-1061
  }
  __shared__ double red[2][256 / WAVE];
  rc_ = wave_sum(rc_);
  rh_ = wave_sum(rh_);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    red[0][threadIdx.x / WAVE] = rc_;
    red[1][threadIdx.x / WAVE] = rh_;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, z = 0.0;
    for (int w = 0; w < (int)(blockDim.x / WAVE); ++w) { x += red[0][w]; z += red[1][w]; }
    a.res_part[blockIdx.x] = x;
    a.res_part[gridDim.x + blockIdx.x] = z;
  }
}

// density of every cell the chemistry did not update (main:1056-1061)
__global__ void k_density(const UcParams* __restrict__ P, double* __restrict__ rho, const double* __restrict__ y,
                          const double* __restrict__ tk, const uint8_t* __restrict__ done, int64_t n) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || done[c]) return;
  double rbar = 0.0;
  for (int k = 0; k < NSP; ++k) rbar += y[k * n + c] * P->rwm[k];
  rho[c] = P->pback / (P->beta1 * tk[c] * rbar);
}

// main:1062-1068: last I column extrapolated, then the J=1 row copied from J=2
__global__ void k_density_edge(double* __restrict__ rho, int32_t li, int32_t lj) {
  const int t = threadIdx.x;
  const int64_t n = (int64_t)li * lj;
  for (int j = t; j < lj; j += blockDim.x) {
    const int64_t c = n - lj + j;
    rho[c] = 2.0 * rho[c - lj] - rho[c - 2 * lj];
  }
  __syncthreads();
  for (int i = t; i < li; i += blockDim.x) rho[(int64_t)i * lj] = rho[(int64_t)i * lj + 1];
}

__global__ void k_zero(double* __restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

}  // namespace

// ---------------------------------------------------------------- host side
struct uc_context {
  uc_scales sc{};
  UcParams hp{};
  UcParams* dp = nullptr;
  int64_t max_cells = 0;
  int device = 0;
  int64_t G = 0;           // k_chem threads (RC slabs)
  double* rcs = nullptr;
  double* sa = nullptr;
  double* sst = nullptr;
  double* rhsz = nullptr;
  int32_t* nsw = nullptr;
  int32_t* cells = nullptr;
  int32_t* redo = nullptr;
  uint8_t* done = nullptr;
  double* rsd_part = nullptr;
  int64_t rsd_part_len = 0;
  double* res_part = nullptr;
  int64_t res_part_len = 0;
  double* rows = nullptr;  // per-sweep residual sums (<= 1024 sweeps)
  hipEvent_t ev[4]{};
  double last_dt = -1.0;
  int32_t last_nstar = 0;   // previous step's species sweep count (first-pass cap)
  int blocks_per_cu = 1;
  // cell list cache
  const int32_t* mask_ptr = nullptr;
  int32_t mask_li = 0, mask_lj = 0;
  std::vector<int32_t> cell_list;
  std::vector<int32_t> prev_nsw;
};

namespace {

// rcfac[s] = C_m / prod W of the slot's reactants (incl. the C3H8 collider of 459/460)
void fill_params(const uc_scales& sc, UcParams& p);

}  // namespace

#include "chem_host.inc"

extern "C" {

int uc_abi_version(void) { return UC_ABI_VERSION; }

const char* uc_last_error(void) { return g_err.c_str(); }

int uc_scales_from_card(double vref, double tref, double pref, double rtin, double rtot, double fo2in,
                        int32_t igrav, uc_scales* o) {
  if (!o) return fail(UC_E_ARG, "uc_scales_from_card: null output");
  if (!(vref > 0 && tref > 0 && pref > 0 && rtot > rtin && fo2in >= 0 && fo2in <= 1))
    return fail(UC_E_ARG, "uc_scales_from_card: reference values out of range");
  scales_from_card(vref, tref, pref, rtin, rtot, fo2in, igrav, *o);
  return UC_OK;
}

int uc_create(const uc_scales* scales, int64_t max_cells, uc_context** out) {
  if (!scales || !out || max_cells < 9) return fail(UC_E_ARG, "uc_create: bad arguments");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(UC_E_NOGPU, "uc_create: no HIP device");
  auto ctx = std::make_unique<uc_context>();
  HIPCHK(hipGetDevice(&ctx->device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
    return fail(UC_E_NOGPU, std::string("uc_create: device is ") + prop.gcnArchName + ", built for gfx950");
  ctx->sc = *scales;
  fill_params(ctx->sc, ctx->hp);
  ctx->max_cells = max_cells;
  // RC slabs: enough threads to fill the device a few waves deep, capped by the field
  // one RC slab per resident thread of k_chem: fill every CU to the kernel's occupancy
  const int64_t cus = prop.multiProcessorCount;
  int bpc = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_chem, 256, 0));
  ctx->blocks_per_cu = std::max(1, bpc);
  ctx->G = std::max<int64_t>(256, std::min<int64_t>(((max_cells + 255) / 256) * 256,
                                                     cus * ctx->blocks_per_cu * 256));
  HIPCHK(hipMalloc(&ctx->dp, sizeof(UcParams)));
  HIPCHK(hipMemcpy(ctx->dp, &ctx->hp, sizeof(UcParams), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&ctx->rcs, sizeof(double) * SLAB_SLOTS * ctx->G));
  HIPCHK(hipMalloc(&ctx->sa, sizeof(double) * NSP * max_cells));
  HIPCHK(hipMalloc(&ctx->sst, sizeof(double) * 4 * max_cells));
  HIPCHK(hipMalloc(&ctx->rhsz, sizeof(double) * max_cells));
  HIPCHK(hipMalloc(&ctx->nsw, sizeof(int32_t) * max_cells));
  HIPCHK(hipMalloc(&ctx->cells, sizeof(int32_t) * max_cells));
  HIPCHK(hipMalloc(&ctx->redo, sizeof(int32_t) * max_cells));
  HIPCHK(hipMalloc(&ctx->done, max_cells));
  ctx->rsd_part_len = (int64_t)1024 * (ctx->G / WAVE);  // isor1 <= 127, isor2 <= 1000 rows
  HIPCHK(hipMalloc(&ctx->rsd_part, sizeof(double) * ctx->rsd_part_len));
  ctx->res_part_len = 2 * ((max_cells + 255) / 256 + 1);
  HIPCHK(hipMalloc(&ctx->res_part, sizeof(double) * ctx->res_part_len));
  HIPCHK(hipMalloc(&ctx->rows, sizeof(double) * 1024));
  for (auto& e : ctx->ev) HIPCHK(hipEventCreate(&e));
  *out = ctx.release();
  return UC_OK;
}

int uc_destroy(uc_context* ctx) {
  if (!ctx) return UC_OK;
  for (void* p : {(void*)ctx->dp, (void*)ctx->rcs, (void*)ctx->sa, (void*)ctx->sst, (void*)ctx->rhsz,
                  (void*)ctx->nsw, (void*)ctx->cells, (void*)ctx->redo, (void*)ctx->done,
                  (void*)ctx->rsd_part, (void*)ctx->res_part, (void*)ctx->rows})
    (void)hipFree(p);
  for (auto& e : ctx->ev) (void)hipEventDestroy(e);
  delete ctx;
  return UC_OK;
}

int uc_step(uc_context* ctx, const uc_field* f, const uc_controls* ctl, void* stream, uc_stats* st) {
  return step_impl(ctx, f, ctl, (hipStream_t)stream, st);
}

int uc_step_host(uc_context* ctx, const uc_field* hf, const uc_controls* ctl, uc_stats* st) {
  return step_host_impl(ctx, hf, ctl, st);
}

}  // extern "C"
