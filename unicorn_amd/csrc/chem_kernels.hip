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
constexpr int64_t UC_MAX_CELLS = 4000000;  // 54 doubles x cells < 2^31 bytes

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

// nsw word of a cell: species sweeps executed, plus NSW_FIXED when the cell
// stopped at a bitwise fixed point (a sweep that changed nothing; in the
// per-cell mode: when it met its own tolerance)
constexpr int32_t NSW_FIXED = 1 << 30;
constexpr int32_t NSW_MASK = NSW_FIXED - 1;

// Per-cell buffers are indexed by the cell's position q in the cell list;
// species-major arrays have a row stride of npad (the list length rounded up
// to whole waves).  The slab of cell q is tile q/64, lane q%64 -- thread q of a
// one-thread-per-cell launch, so the tile is the launch's wave index.

struct RatesArgs {
  const double* rho; const double* y; const double* tk; const double* stmf; const double* stnd;
  const int32_t* cells; int32_t ncells; int64_t n;
  double* slabs;
};

// rate coefficients, rho*Y_k and soot sources of every cell into its slab
// (ckrates.f:92-1352)
__global__ void __launch_bounds__(256) k_rates(const UcParams* __restrict__ Pg, RatesArgs a) {
  UcP P = (UcP)Pg;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c = a.cells[q < a.ncells ? q : a.ncells - 1];  // tail lanes redo the last cell
  RcSlab rc = wave_slab(a.slabs);
  const FieldBuf Y = field_buf(a.y, a.n * 8 * NSP, c);
  const double rho = a.rho[c], tk = a.tk[c];
  double tbody;
  {
    double fa[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) fa[k] = Y.ld(k * a.n);
    uc_rates(P, tk, rho, fa, rc, tbody);
  }
  uc_soot(P, rc, tk, rho, a.stmf[c], a.stnd[c], tbody);
}

struct SweepArgs {
  const double* rho; const double* y; const double* stmf; const double* stnd;
  const int32_t* cells; int32_t ncells; int64_t n; int64_t npad;
  double* slabs;
  double* st;          // [54][npad] iterates where each cell stopped
  double* snap;        // [nsnap][54][npad] iterates after sweeps snap_from.. (may be null)
  double* sst2p;       // [npad] SST2 before a fixed cell's last sweep
  int32_t* nsw;        // [npad] NSW word
  double* rsd_part;    // [isor1][rsd_stride] per-wave residual sums of each sweep (may be null)
  int64_t rsd_stride;
  unsigned long long* swcount;  // sweeps executed (+=)
  int32_t* redo_count;          // cells re-run (+=, redo launches)
  int32_t start, count;  // sweeps start+1 .. start+count (counts are 1-based)
  int32_t snap_from;     // first sweep count that is snapshotted
  int32_t nsnap;
  int32_t redo;          // > 0: re-run from scratch, to this count, the cells whose iterates
                         // at that count are neither in st nor in a snapshot
  int32_t mode;
  double dt, sor1, sor2, tolrsp;
};

// A segment of Gauss-Seidel sweeps (solvesp.f:87-2831) for every cell that has
// not stopped: from Y (start = 0) or from the saved iterates, with the slab's
// RC.  Sweeps are deterministic functions of the iterates and the slab, so a
// segment continues the previous one bit for bit.
// One wave per block; the block's dynamic LDS (UC_SWEEP_LDS doubles per lane) is
// the sweep's second-level cache of slab values.
__global__ void __launch_bounds__(WAVE) k_sweep(const UcParams* __restrict__ Pg, SweepArgs a) {
  extern __shared__ double uc_lds[];
  double* const lds = uc_lds + threadIdx.x;
  UcP P = (UcP)Pg;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = q < a.ncells;
  const int64_t qq = valid ? q : a.ncells - 1;
  const int64_t wave = q / WAVE;
  const int32_t w0 = a.nsw[qq];
  bool active;
  if (a.redo > 0) {
    const int32_t m = w0 & NSW_MASK;
    const bool fx = (w0 & NSW_FIXED) != 0;
    active = valid && !(fx && m - 1 <= a.redo) && !(!fx && m == a.redo);
  } else {
    active = valid && (a.start == 0 || (w0 & NSW_FIXED) == 0);
  }
  if (!__any(active)) return;  // this wave's residual rows stay zero
  const bool was_active = active;
  if (a.redo > 0) {
    const unsigned long long b = __ballot(active);
    if ((threadIdx.x & (WAVE - 1)) == 0) atomicAdd(a.redo_count, (int32_t)__popcll(b));
  }
  RcSlab rc = wave_slab(a.slabs);
  const int64_t c = a.cells[qq];
  const double rho = a.rho[c], stmf = a.stmf[c], stnd = a.stnd[c];
  double s[54];
  int32_t nsw;
  if (a.start == 0) {  // SA = FA (spgsolv.f:169-220), soot from the field
    const FieldBuf Y = field_buf(a.y, a.n * 8 * NSP, c);
#pragma unroll
    for (int k = 0; k < NSP; ++k) s[k] = Y.ld(k * a.n);
    s[52] = stmf;
    s[53] = stnd;
    nsw = 0;
  } else {
    const FieldBuf ST = field_buf(a.st, a.npad * 8 * 54, qq);
#pragma unroll
    for (int k = 0; k < 54; ++k) s[k] = ST.ld(k * a.npad);
    nsw = w0 & NSW_MASK;
  }
  bool fixed = false;
  double sst2_prev = 0.0;
  uint32_t executed = 0;
  for (int32_t sw = a.start; sw < a.start + a.count; ++sw) {
    if (!__any(active)) break;
    double r = 0.0;
    if (active) {
      sst2_prev = s[53];
      double d2;
      r = uc_sweep(P, rc, lds, s, rho, stmf, stnd, a.dt, a.sor1, a.sor2, &d2);
      ++nsw;
      ++executed;
      if (a.mode == UC_MODE_PERCELL) {
        if (r / 51.0 <= a.tolrsp) { active = false; fixed = true; }
      } else if (r == 0.0) {
        active = false;  // every later sweep repeats this one exactly
        fixed = true;
      }
      if (a.snap && sw + 1 >= a.snap_from) {
        const FieldBuf SN = field_buf(a.snap + (int64_t)(sw + 1 - a.snap_from) * 54 * a.npad,
                                      a.npad * 8 * 54, qq);
#pragma unroll
        for (int k = 0; k < 54; ++k) SN.st(k * a.npad, s[k]);
      }
    }
    if (a.rsd_part) {
      const double ws = wave_sum(r);
      if ((threadIdx.x & (WAVE - 1)) == 0) a.rsd_part[(int64_t)sw * a.rsd_stride + wave] = ws;
    }
  }
  {
    unsigned long long e = executed;
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0) atomicAdd(a.swcount, e);
  }
  if (!was_active) return;
  const FieldBuf ST = field_buf(a.st, a.npad * 8 * 54, q);
#pragma unroll
  for (int k = 0; k < 54; ++k) ST.st(k * a.npad, s[k]);
  a.nsw[q] = nsw | (fixed ? NSW_FIXED : 0);
  if (fixed) a.sst2p[q] = sst2_prev;
}

struct SourcesArgs {
  const double* rho; const double* tk; const double* ht; const double* stmf; const double* stnd;
  const int32_t* cells; int32_t ncells; int64_t n; int64_t npad;
  const double* slabs; const double* st; const double* snap; const double* sst2p; const int32_t* nsw;
  int32_t nstar;        // decided sweep count (global mode)
  int32_t snap_from;    // sweep count of snapshot 0 of the last segment
  int32_t nsnap;
  int32_t mode;
  double dt, sor1, sor2;
  double* sa;           // [52][npad] species after the closure
  double* sst;          // [2][npad]  SST1, SST2
  double* rhsz;         // [npad]
  double* qdot; double* qdsp;  // field arrays (solveh.f:1137-1142)
};

// Iterates at the decided count, SST2 completion, N2 closure (solvesp.f:2839-2914)
// and the SOLVEH sources: radiation, RA/PRD, QDOT, QDSP, RHSZ (solveh.f:34-1142).
__global__ void __launch_bounds__(256) k_sources(const UcParams* __restrict__ Pg, SourcesArgs a) {
  UcP P = (UcP)Pg;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.ncells) return;
  RcSlab rc = wave_slab((double*)a.slabs);
  const int64_t c = a.cells[q], n = a.n;
  const double rho = a.rho[c], tk = a.tk[c], ht = a.ht[c], stmf = a.stmf[c], stnd = a.stnd[c];
  const int32_t w = a.nsw[q];
  const int32_t m = w & NSW_MASK;
  const bool fx = (w & NSW_FIXED) != 0;
  const bool global = a.mode != UC_MODE_PERCELL;
  const bool use_st = !global || (fx && m - 1 <= a.nstar) || (!fx && m == a.nstar);
  const FieldBuf SRC = use_st ? field_buf(a.st, a.npad * 8 * 54, q)
                              : field_buf(a.snap + (int64_t)(a.nstar - a.snap_from) * 54 * a.npad,
                                          a.npad * 8 * 54, q);
  double s[54];
#pragma unroll
  for (int k = 0; k < 54; ++k) s[k] = SRC.ld(k * a.npad);
  if (global && fx && m - 1 <= a.nstar) {
    // a fixed cell's species repeat; only SST2 keeps relaxing (solvesp.f:2820-2831)
    if (a.nstar == m - 1) {
      s[53] = a.sst2p[q];
    } else {
      const double b = (rho * stnd + a.dt * rc.ld(UC_SLOT_SOUR + 1) * s[21]) * a.sor1
                       / (rho + a.dt * rc.ld(UC_SLOT_SOUR + 5));
      for (int32_t i = m; i < a.nstar; ++i) {
        const double fnew = a.sor2 * s[53] + b;
        s[53] = fnew < 0.0 ? 0.0 : fnew;
      }
    }
  }
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
  const FieldBuf QD = field_buf(a.qdsp, n * 8 * NSP, c);
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
  a.rhsz[q] = rhsz;
  a.qdot[c] = -hrr * rconv;
  const FieldBuf SA = field_buf(a.sa, a.npad * 8 * NSP, q);
#pragma unroll
  for (int k = 0; k < NSP; ++k) SA.st(k * a.npad, s[k]);
  a.sst[q] = s[52];
  a.sst[a.npad + q] = s[53];
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
    const int64_t qq = valid ? q : 0;
    const int64_t c = a.cells[qq];
    double fpz = a.ht[c];
    const double inc = a.rhsz[qq] * a.zsor1 / a.rho[c];
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
  const double* sa; const double* sst; const double* rhsz;  // [52][npad], [2][npad], [npad]
  const int32_t* cells; int32_t ncells; int64_t n; int64_t npad;
  int32_t nh;        // decided enthalpy sweep count (global mode) / cap (per-cell)
  int32_t mode;
  double zsor1, zsor2, tolrh;
  double* res_part;  // [2][nblocks] RESDC, RESDH partial sums
  double* fpz_out;   // null: SPGSOLV epilogue + density; else FPZ here and TK only
};

__global__ void __launch_bounds__(256) k_finish(const UcParams* __restrict__ Pg, FinishArgs a) {
  UcP P = (UcP)Pg;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rc_ = 0.0, rh_ = 0.0;
  if (q < a.ncells) {
    const int64_t c = a.cells[q], n = a.n;
    const FieldBuf SA = field_buf(a.sa, a.npad * 8 * NSP, q), Y = field_buf(a.y, n * 8 * NSP, c);
    double s[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) s[k] = SA.ld(k * a.npad);
    const double rho = a.rho[c], ht = a.ht[c];
    // enthalpy relaxation (solveh.f:1147-1167), zero transport: PPP = RHO
    double fpz = ht;
    const double inc = a.rhsz[q] * a.zsor1 / rho;
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
    if (a.fpz_out) {  // SOLVEH's outputs only (solveh.f:1167, :1230)
      a.fpz_out[c] = fpz;
      a.tk[c] = tnew;
      goto reduce;
    }
    {
    // SPGSOLV epilogue (spgsolv.f:1299-1416)
    const double sst1 = a.sst[q], sst2 = a.sst[a.npad + q];
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
  }
reduce:
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
namespace {
// bookkeeping of the step in progress (chem_host.inc: ph_begin .. ph_end)
struct StepRun {
  const uc_field* f = nullptr;
  uc_controls ctl{};
  hipStream_t s = nullptr;
  int64_t n = 0, npad = 0, nwc = 0, nb = 0;
  int32_t ncells = 0, nstar = 0, nh = 0, snap_from = 1;
  bool global = true;
  double sor1 = 0.0, sor2 = 0.0, spnorm = 0.0, hnorm = 0.0;
  float ms_rates = 0.f, ms_sweeps = 0.f, ms_sources = 0.f;
  uc_stats out{};
  std::vector<double> r = std::vector<double>(1024, 0.0);  // per-sweep residual sums
};
}  // namespace

struct uc_context {
  uc_scales sc{};
  UcParams hp{};
  UcParams* dp = nullptr;
  int64_t max_cells = 0;
  int64_t max_pad = 0;     // max_cells rounded up to whole waves
  int device = 0;
  int32_t nsnap = 0;       // snapshot slots per cell
  double* slabs = nullptr; // [max_pad/64][SLAB_SLOTS][64]
  double* st = nullptr;    // [54][max_pad]
  double* snap = nullptr;  // [nsnap][54][max_pad]
  double* sst2p = nullptr; // [max_pad]
  double* sa = nullptr;    // [52][max_pad]
  double* sst = nullptr;   // [2][max_pad]
  double* rhsz = nullptr;  // [max_pad]
  int32_t* nsw = nullptr;  // [max_pad]
  int32_t* cells = nullptr;
  uint8_t* done = nullptr;
  double* rsd_part = nullptr;  // per-sweep, per-wave residual sums
  int64_t rsd_part_len = 0;
  int64_t hres_threads = 0;    // k_hres grid (persistent)
  double* res_part = nullptr;
  int64_t res_part_len = 0;
  double* rows = nullptr;      // per-sweep residual sums (<= 1024 sweeps)
  unsigned long long* counters = nullptr;  // [0] sweeps executed, [1] (int) cells re-run
  hipEvent_t ev[8]{};
  double last_dt = -1.0;
  int32_t last_nstar = 0;  // previous step's species sweep count (first segment length)
  // cell list cache
  const int32_t* mask_ptr = nullptr;
  int32_t mask_li = 0, mask_lj = 0;
  std::vector<int32_t> cell_list;
  StepRun run;
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
  // buffer instructions address each per-cell array with 32-bit byte offsets
  if (max_cells > UC_MAX_CELLS)
    return fail(UC_E_ARG, "uc_create: max_cells above " + std::to_string(UC_MAX_CELLS));
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
  ctx->max_pad = (max_cells + WAVE - 1) / WAVE * WAVE;
  ctx->nsnap = 8;
  const int64_t mp = ctx->max_pad;
  HIPCHK(hipMalloc(&ctx->dp, sizeof(UcParams)));
  HIPCHK(hipMemcpy(ctx->dp, &ctx->hp, sizeof(UcParams), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&ctx->slabs, sizeof(double) * SLAB_SLOTS * mp));
  HIPCHK(hipMalloc(&ctx->st, sizeof(double) * 54 * mp));
  HIPCHK(hipMalloc(&ctx->snap, sizeof(double) * 54 * mp * ctx->nsnap));
  HIPCHK(hipMalloc(&ctx->sst2p, sizeof(double) * mp));
  HIPCHK(hipMalloc(&ctx->sa, sizeof(double) * NSP * mp));
  HIPCHK(hipMalloc(&ctx->sst, sizeof(double) * 2 * mp));
  HIPCHK(hipMalloc(&ctx->rhsz, sizeof(double) * mp));
  HIPCHK(hipMalloc(&ctx->nsw, sizeof(int32_t) * mp));
  HIPCHK(hipMalloc(&ctx->cells, sizeof(int32_t) * max_cells));
  HIPCHK(hipMalloc(&ctx->done, max_cells));
  ctx->hres_threads = std::min<int64_t>(mp, (int64_t)prop.multiProcessorCount * 1024);
  ctx->rsd_part_len = std::max<int64_t>(128 * (mp / WAVE), 1000 * (ctx->hres_threads / WAVE));
  HIPCHK(hipMalloc(&ctx->rsd_part, sizeof(double) * ctx->rsd_part_len));
  ctx->res_part_len = 2 * ((max_cells + 255) / 256 + 1);
  HIPCHK(hipMalloc(&ctx->res_part, sizeof(double) * ctx->res_part_len));
  HIPCHK(hipMalloc(&ctx->rows, sizeof(double) * 1024));
  HIPCHK(hipMalloc(&ctx->counters, sizeof(unsigned long long) * 2));
  if (UC_SWEEP_LDS * 512 > 65536)
    HIPCHK(hipFuncSetAttribute((const void*)k_sweep, hipFuncAttributeMaxDynamicSharedMemorySize,
                               UC_SWEEP_LDS * 512));
  for (auto& e : ctx->ev) HIPCHK(hipEventCreate(&e));
  *out = ctx.release();
  return UC_OK;
}

int uc_destroy(uc_context* ctx) {
  if (!ctx) return UC_OK;
  for (void* p : {(void*)ctx->dp, (void*)ctx->slabs, (void*)ctx->st, (void*)ctx->snap, (void*)ctx->sst2p,
                  (void*)ctx->sa, (void*)ctx->sst, (void*)ctx->rhsz, (void*)ctx->nsw, (void*)ctx->cells,
                  (void*)ctx->done, (void*)ctx->rsd_part, (void*)ctx->res_part, (void*)ctx->rows,
                  (void*)ctx->counters})
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
