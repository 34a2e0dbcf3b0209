// MI355X (gfx950) kernels and host C ABI of UNICORN's chemistry step.
//
// Step = main:821-828 clip/closure -> CKRATES -> SOLVESP -> SOLVEH -> SPGSOLV
// epilogue -> main:1056-1073 density, on one grid or a batch of grids (each with its
// own DT, controls and global stops); or, with the transport coefficients SPGSOLV
// assembles, on one grid in the reference's Gauss-Seidel order (k_sweep_cpl and
// friends, below k_zero).
//
// Kernels (FP64 throughout; one thread per position of the cell list, whose
// grids are padded to whole waves):
//   k_clip      all cells: Y<0 -> 0, Y_N2 = 1 - sum(Y_1..51)            main:821-828
//   k_rates     rate coefficients, rho*Y_k and soot sources into the cell's
//               slab                                                      ckrates.f:92-1352
//   k_sweep     a segment of Gauss-Seidel sweeps per tile of 64 cells, with
//               per-tile per-sweep residual sums; a cell stops once a sweep
//               leaves every iterate bit-for-bit unchanged               solvesp.f:87-2838
//   k_sources   iterates at the decided count, N2 closure, SOLVEH sources  solvesp.f:2839-2914,
//                                                                         solveh.f:34-1145
//   k_hres      enthalpy relaxation residual trajectories                  solveh.f:1147-1167
//   k_finish    enthalpy relaxation, Newton T(h), SPGSOLV epilogue, density solveh.f:1170-1234
//   k_density / k_density_edge  density of the remaining cells, extrapolation main:1056-1068
//   k_pack, k_bin_*, k_tile_keys, k_save_nsw   batches and the stiffness-sorted work queue
//
// The step is orchestrated by chem_host.inc (phases ph_begin .. ph_end).
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

#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "../../include/unicorn_chem.h"

namespace {

constexpr int NSP = 52;
constexpr int NRX = 544;
constexpr int WAVE = 64;
constexpr int64_t UC_MAX_CELLS = 4000000;  // 54 doubles x cells < 2^31 bytes

struct UcParams {
  double tstr, alstr, vstr, rstr, wmstr, hstr, gasc, pback, beta1, rei, beta3;
  double wm[NSP], rwm[NSP], hfo[NSP];
  double nasa[NSP][14];
  double rcfac[NRX];  // C_m / prod W of the slot's reactants (ckrates.f:83-85, :796-1333)
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

__device__ __forceinline__ RcSlab tile_rsrc(const double* base, int64_t tile, int slots) {
  double* t = (double*)wave_uniform((uint64_t)(base + tile * (int64_t)slots * 64));
  return RcSlab{__builtin_amdgcn_make_buffer_rsrc(t, 0, slots * 512, 0x00020000), (threadIdx.x & 63u) * 8u};
}

__device__ __forceinline__ RcSlab tile_slab(double* slabs, int64_t tile) { return tile_rsrc(slabs, tile, SLAB_SLOTS); }

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

// The long straight-line kernels (70-380 KB of code each, larger than the instruction
// cache) start on 64 KB boundaries of the code object.  Their speed depended on where the
// linker happened to place them: k_sweep ran 3.5% slower at some offsets with the same
// instructions (1001x401 field, 20 steps: 110.8 against 107.2 ms of sweeps), and at a
// 64 KB boundary it runs at its best whatever else the code object holds.
#define UC_HOT __attribute__((aligned(65536)))
// Waves per SIMD the register allocation must allow, where more waves hide more latency
// than the spills they cost (1001x401 field, 20 steps): k_sources at 2 (84 B of spills;
// 12.6 -> 11.1 ms, at 3 it spills 652 B and takes 13.2), k_rates at 3 (324 B; 14.2 ->
// 13.8 ms), k_finish at 3 (20 B; about 0.5 ms less).
#define UC_WAVES(n) __attribute__((amdgpu_waves_per_eu(n, n)))

// Reads of the coupled sweep (uc_sweep_cpl): slab slots, the cell's coupling slots
// (UC_CPL + s, a tile of UC_CPL_SLOTS per wave like the slab), and the neighbours'
// current iterates (UC_NBR + 4*row + dir, dir 0: J+1, 1: J-1, 2: I+1, 3: I-1) from the
// field-layout iterate array gs [54][n].  Slot numbers are compile-time constants, so
// each ld() inlines to one buffer load.
struct RcCpl {
  RcSlab slab, cpl;
  __amdgpu_buffer_rsrc_t g;
  uint32_t nb8[4];  // byte offsets of the four neighbour cells
  int32_t n8;       // row stride of gs in bytes
  __device__ __forceinline__ double ld(int s) const {
    if (s < UC_CPL) return slab.ld(s);
    if (s < UC_NBR) return cpl.ld(s - UC_CPL);
    const int r = (s - UC_NBR) >> 2, d = (s - UC_NBR) & 3;
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(g, nb8[d], r * n8, 0));
  }
  __device__ __forceinline__ void st(int s, double v) const { slab.st(s, v); }
  __device__ __forceinline__ void fence() {
    slab.fence();
    cpl.fence();
    asm volatile("" : "+s"(g));
  }
};

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

// Per-grid parameters and stop state of a (batched) step.  Grids are packed:
// grid g's cells are field cells off .. off+li*lj-1 of the packed arrays.
struct GridTab {
  double dt, sor1, sor2, tolrsp, zsor1, zsor2, tolrh;
  double spnorm, hnorm;                   // global-stop norms (solvesp.f:2834, solveh.f:1162)
  int32_t mode, isor1, isor2;
  int32_t start, count, snap_from, redo;  // current sweep segment (counts are 1-based)
  int32_t nstar, nh;                      // decided species / enthalpy sweep counts
  int32_t hfrom;                          // first enthalpy sweep of the current k_hres chunk
  int32_t pred, grow, nsnap;              // segment plan: predicted count, next growth, snapshots
  int32_t isor1_out, isor2_out, hcount;   // reported counts (limit + 1: not converged); residuals kept
  int32_t tile0, tile1;                   // the grid's tiles of 64 list positions
  int32_t li, lj;
  int64_t off;
};

constexpr int32_t UC_HIST = 128;  // species residuals kept per grid (isor1 <= 127)

// After a sweep segment: each grid's global stop (solvesp.f:2834: the first sweep with
// RSD/SPNORM <= TOLRSP) from its per-sweep row sums, and its next segment: one sweep,
// then 2, 4, ... up to the snapshot depth after a predicted first segment, the snapshot
// depth without a prediction.  A grid that stops keeps the snapshot window of its last
// segment; when the decided count lies before that window, redo is set to it (the
// cells not covered re-run from scratch).  hist [grid][UC_HIST] receives RSD/SPNORM of
// every sweep read.  One thread per grid; the step needs no host round trip per segment.
__global__ void k_decide_sp(GridTab* __restrict__ gt, const double* __restrict__ rows, int32_t nrows,
                            int32_t ngrids, double* __restrict__ hist) {
  for (int32_t g = threadIdx.x; g < ngrids; g += blockDim.x) {
    GridTab& G = gt[g];
    if (G.count <= 0) continue;
    const int32_t end = G.start + G.count;
    const bool percell = G.mode == UC_MODE_PERCELL;
    int32_t conv = 0;
    if (!percell) {
      const double* r = rows + (int64_t)g * nrows;
      for (int32_t i = G.start; i < end; ++i) {
        const double v = r[i - G.start] / G.spnorm;
        hist[(int64_t)g * UC_HIST + i] = v;
        if (v <= G.tolrsp) { conv = i + 1; break; }
      }
    }
    if (percell || conv || end >= G.isor1) {
      G.nstar = conv ? conv : G.isor1;
      G.isor1_out = conv ? conv : G.isor1 + 1;
      G.hcount = percell ? 0 : (conv ? conv : end);
      G.count = 0;
      G.redo = !percell && G.nstar < G.snap_from ? G.nstar : 0;
      continue;
    }
    const int32_t len = min(G.isor1 - end, G.pred > 0 ? G.grow : G.nsnap);
    G.grow = min(2 * G.grow, G.nsnap);
    G.start = end;
    G.count = len;
    G.snap_from = end + 1;
  }
}

// After a k_hres chunk of H sweeps: each grid's enthalpy stop (solveh.f:1162: the first
// sweep with RSDH/HNORM <= TOLRH) from its row sums; hfrom advances to the next chunk
// or becomes -1 with the decided count in nh.
__global__ void k_decide_h(GridTab* __restrict__ gt, const double* __restrict__ rows, int32_t H, int32_t ngrids) {
  for (int32_t g = threadIdx.x; g < ngrids; g += blockDim.x) {
    GridTab& G = gt[g];
    if (G.hfrom < 0) continue;
    const int32_t end = min(G.hfrom + H, G.isor2);
    int32_t conv = 0;
    for (int32_t i = G.hfrom; i < end; ++i)
      if (rows[(int64_t)g * H + i - G.hfrom] / G.hnorm <= G.tolrh) { conv = i + 1; break; }
    if (conv || end >= G.isor2) {
      G.nh = conv ? conv : G.isor2;
      G.isor2_out = conv ? conv : G.isor2 + 1;
      G.hfrom = -1;
    } else {
      G.hfrom = end;
    }
  }
}

// The cell list: position q holds a packed field index, or -1-c for a padding lane
// (c: a real cell of the same grid, so the lane has valid inputs).  Every grid's
// positions are padded to whole tiles of 64, so a wave belongs to one grid.
// Per-position buffers (slabs, st, snap, sa, ...) are indexed by q; species-major
// ones have a row stride of npad (the list length).  The slab of position q is
// tile q/64, lane q%64.
__device__ __forceinline__ int64_t list_cell(const int32_t* cells, int64_t q, bool& valid) {
  const int32_t v = cells[q];
  valid = v >= 0;
  return valid ? v : -1 - (int64_t)v;
}

__device__ __forceinline__ int32_t wave_grid(const int32_t* gid, int64_t q) {
  return __builtin_amdgcn_readfirstlane(gid[q]);
}

struct RatesArgs {
  const double* rho; const double* y; const double* tk; const double* stmf; const double* stnd;
  const int32_t* cells; int64_t npad; int64_t n;
  double* slabs;
};

// rate coefficients, rho*Y_k and soot sources of every cell into its slab
// (ckrates.f:92-1352)
__global__ void __launch_bounds__(512) UC_WAVES(3) UC_HOT k_rates(const UcParams* __restrict__ Pg, RatesArgs a) {
  UcP P = (UcP)Pg;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.npad) return;  // whole waves only: npad is a multiple of 64
  bool valid;
  const int64_t c = list_cell(a.cells, q, valid);
  RcSlab rc = tile_slab(a.slabs, q / WAVE);
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

// The same slab in NP parts: blockIdx.y is the part, which forms the slots of its share
// of the reactions (part 0 also rho*Y_k and the soot sources), so a cell's rates take
// NP lanes in as many workgroups and a small cell set finishes in about 1/NP of the
// one-lane time.  Each slot's expression is uc_rates' own: the slab is the same bits.
template <int NP>
__global__ void __launch_bounds__(512) UC_HOT k_rates_split(const UcParams* __restrict__ Pg, RatesArgs a) {
  UcP P = (UcP)Pg;
  const int part = blockIdx.y;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.npad) return;  // whole waves only: npad is a multiple of 64
  bool valid;
  const int64_t c = list_cell(a.cells, q, valid);
  RcSlab rc = tile_slab(a.slabs, q / WAVE);
  const FieldBuf Y = field_buf(a.y, a.n * 8 * NSP, c);
  const double rho = a.rho[c], tk = a.tk[c];
  double tbody;
  if constexpr (NP == 2) uc_rates_split2(part, P, tk, rho, Y, a.n, rc, tbody);
  else if constexpr (NP == 4) uc_rates_split4(part, P, tk, rho, Y, a.n, rc, tbody);
  else uc_rates_split8(part, P, tk, rho, Y, a.n, rc, tbody);
  if (part == 0) uc_soot(P, rc, tk, rho, a.stmf[c], a.stnd[c], tbody);
}

struct SweepArgs {
  const double* rho; const double* y; const double* stmf; const double* stnd;
  const int32_t* cells; const int32_t* gid; const GridTab* gt;
  int64_t npad; int64_t n;
  double* slabs;
  double* st;          // [54][npad] iterates where each cell stopped
  double* snap;        // [nsnap][54][npad] iterates after sweeps snap_from.. (may be null)
  double* sst2p;       // [npad] SST2 before a fixed cell's last sweep
  int32_t* nsw;        // [npad] NSW word
  double* rsd_part;    // [isor1][ntiles] per-tile residual sums of each sweep (may be null)
  unsigned long long* swcount;  // [ngrids] sweeps executed (+=)
  int32_t* redo_count;          // [ngrids] cells re-run (+=, redo launches)
  const int32_t* order;  // tile order (null: natural)
  int32_t* queue;        // persistent launches: next position in `order` (null: one tile per block)
  int32_t ntiles;
  int32_t nsnap;
  int32_t redo;          // 1: re-run from scratch, to each grid's redo count, the cells whose
                         // iterates at that count are neither in st nor in a snapshot
};

// A segment of Gauss-Seidel sweeps (solvesp.f:87-2831) for one tile of 64 list
// positions: every cell that has not stopped continues from Y (segment start 0) or
// from its saved iterates, with the slab's RC.  Sweeps are deterministic functions
// of the iterates and the slab, so a segment continues the previous one bit for bit.
// W lanes per cell: the block's W waves each run one part of the split sweep on the
// same 64 cells (uc_sweep_mwW) and hold identical iterates; wave 0 writes the results.
// lds: this lane's first position of its wave's LDS cache; xch: this lane's column of
// the exchange area (W > 1).
template <int W>
__device__ __forceinline__ void sweep_tile(UcP P, const SweepArgs& a, double* lds, double* xch, int64_t tile) {
  const int part = W == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int64_t q = tile * WAVE + (threadIdx.x & (WAVE - 1));
  const int32_t g = wave_grid(a.gid, q);
  const GridTab& G = a.gt[g];
  const int32_t start = G.start, count = a.redo ? G.redo : G.count;
  if (count <= 0) return;  // grid done (its residual rows are not read)
  bool valid;
  const int64_t c = list_cell(a.cells, q, valid);
  const int32_t w0 = a.nsw[q];
  bool active;
  if (a.redo) {
    const int32_t m = w0 & NSW_MASK;
    const bool fx = (w0 & NSW_FIXED) != 0;
    active = valid && !(fx && m - 1 <= G.redo) && !(!fx && m == G.redo);
  } else {
    active = valid && (start == 0 || (w0 & NSW_FIXED) == 0);
  }
  const int32_t first = a.redo ? 0 : start;
  if (a.redo && part == 0) {
    const unsigned long long b = __ballot(active);
    if ((threadIdx.x & (WAVE - 1)) == 0 && b) atomicAdd(&a.redo_count[g], (int32_t)__popcll(b));
  }
  if (!__any(active)) {  // residual rows of this tile: zero
    if (a.rsd_part && threadIdx.x == 0)
      for (int32_t sw = first; sw < first + count; ++sw) a.rsd_part[(int64_t)sw * a.ntiles + tile] = 0.0;
    return;
  }
  const bool was_active = active;
  RcSlab rc = tile_slab(a.slabs, tile);
  const double rho = a.rho[c], stmf = a.stmf[c], stnd = a.stnd[c];
  const double dt = G.dt, sor1 = G.sor1, sor2 = G.sor2, tol = G.tolrsp;
  const bool percell = G.mode == UC_MODE_PERCELL;
  double s[54];
  int32_t nsw;
  if (first == 0) {  // SA = FA (spgsolv.f:169-220), soot from the field
    const FieldBuf Y = field_buf(a.y, a.n * 8 * NSP, c);
#pragma unroll
    for (int k = 0; k < NSP; ++k) s[k] = Y.ld(k * a.n);
    s[52] = stmf;
    s[53] = stnd;
    nsw = 0;
  } else {
    const FieldBuf ST = field_buf(a.st, a.npad * 8 * 54, q);
#pragma unroll
    for (int k = 0; k < 54; ++k) s[k] = ST.ld(k * a.npad);
    nsw = w0 & NSW_MASK;
  }
  bool fixed = false;
  double sst2_prev = 0.0;
  uint32_t executed = 0;
  const int32_t snap_from = G.snap_from;
  for (int32_t sw = first; sw < first + count; ++sw) {
    double r = 0.0;
    if (__any(active)) {
      if (active) {
        sst2_prev = s[53];
        double d2;
        if constexpr (W == 1)
          r = uc_sweep(P, rc, lds, s, rho, stmf, stnd, dt, sor1, sor2, &d2);
        else if constexpr (W == 2)
          r = uc_sweep_mw2(part, P, rc, lds, xch, s, rho, stmf, stnd, dt, sor1, sor2, &d2);
        else
          r = uc_sweep_mw4(part, P, rc, lds, xch, s, rho, stmf, stnd, dt, sor1, sor2, &d2);
        ++nsw;
        ++executed;
        if (percell) {
          if (r / 51.0 <= tol) { active = false; fixed = true; }
        } else if (r == 0.0) {
          active = false;  // every later sweep repeats this one exactly
          fixed = true;
        }
        if (part == 0 && a.snap && sw + 1 >= snap_from) {
          const FieldBuf SN = field_buf(a.snap + (int64_t)(sw + 1 - snap_from) * 54 * a.npad,
                                        a.npad * 8 * 54, q);
#pragma unroll
          for (int k = 0; k < 54; ++k) SN.st(k * a.npad, s[k]);
        }
      }
    }
    if (a.rsd_part && part == 0) {
      const double ws = wave_sum(r);
      if ((threadIdx.x & (WAVE - 1)) == 0) a.rsd_part[(int64_t)sw * a.ntiles + tile] = ws;
    }
  }
  if (part != 0) return;
  {
    unsigned long long e = executed;
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0 && e) atomicAdd(&a.swcount[g], e);
  }
  if (!was_active) return;
  const FieldBuf ST = field_buf(a.st, a.npad * 8 * 54, q);
#pragma unroll
  for (int k = 0; k < 54; ++k) ST.st(k * a.npad, s[k]);
  a.nsw[q] = nsw | (fixed ? NSW_FIXED : 0);
  if (fixed) a.sst2p[q] = sst2_prev;
}

// One wave per block; the block's dynamic LDS (UC_SWEEP_LDS doubles per lane) is
// the sweep's second-level cache of slab values.  Tiles come from `order` (the
// stiffness-sorted device work queue): one per block, or, in the persistent form,
// as many as each wave takes from the shared queue position.
template <bool PERSISTENT>
__global__ void __launch_bounds__(WAVE) UC_HOT k_sweep(const UcParams* __restrict__ Pg, SweepArgs a) {
  extern __shared__ double uc_lds[];
  double* const lds = uc_lds + threadIdx.x;
  UcP P = (UcP)Pg;
  if (!PERSISTENT) {
    const int64_t t = blockIdx.x;
    sweep_tile<1>(P, a, lds, nullptr, a.order ? a.order[t] : t);
    return;
  }
  for (;;) {
    int32_t t = 0;
    if (threadIdx.x == 0) t = atomicAdd(a.queue, 1);
    t = __builtin_amdgcn_readfirstlane(t);  // lane 0's ticket, uniform
    if (t >= a.ntiles) return;  // every wave reaches the end of the queue
    sweep_tile<1>(P, a, lds, nullptr, a.order ? __builtin_amdgcn_readfirstlane(a.order[t]) : t);
    asm volatile("" : "+s"(P));
  }
}

// W lanes per cell: W waves per block, one tile per block.  Dynamic LDS: each wave's
// UC_MW<W>_LDS positions per lane, then the exchange area [2][2W+1][64].
template <int W>
__global__ void __launch_bounds__(W * WAVE) UC_HOT k_sweep_mw(const UcParams* __restrict__ Pg, SweepArgs a) {
  extern __shared__ double uc_lds[];
  constexpr int KL = W == 2 ? UC_MW2_LDS : UC_MW4_LDS;
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  double* const lds = uc_lds + wv * KL * WAVE + lane;
  double* const xch = uc_lds + W * KL * WAVE + lane;
  const int64_t t = blockIdx.x;
  sweep_tile<W>((UcP)Pg, a, lds, xch, a.order ? a.order[t] : t);
}

constexpr size_t mw_lds_bytes(int W, int KL) { return (size_t)(W * KL + 2 * (2 * W + 1)) * WAVE * 8; }

struct SourcesArgs {
  const double* rho; const double* tk; const double* ht; const double* stmf; const double* stnd;
  const int32_t* cells; const int32_t* gid; const GridTab* gt;
  int64_t npad; int64_t n;
  const double* slabs; const double* st; const double* snap; const double* sst2p; const int32_t* nsw;
  int32_t nsnap;
  double* sa;           // [52][npad] species after the closure
  double* sst;          // [2][npad]  SST1, SST2
  double* rhsz;         // [npad]
  double* qdot; double* qdsp;  // packed field arrays (solveh.f:1137-1142)
  const double* flux;   // [npad] DT/2 x species-diffusion enthalpy flux divergence (null: zero)
  // SOLVEH without SOLVESP (ICHEM <= 0, spgsolv.f:1291-1298): the species and soot scalars
  // as given, [54][n] field layout (SSP 1..52, SST1, SST2); null: the sweeps' iterates
  const double* given;
  // cells the sweeps did not advance (3 <= ISKIP < 10): the species of the field, y [52][n]
  // with STMF and STND, which SPGSOLV's prologue copies into SSP/SST (spgsolv.f:103-231)
  const double* fy;
};

// Iterates at the decided count, SST2 completion, N2 closure (solvesp.f:2839-2914)
// and the SOLVEH sources: radiation, RA/PRD, QDOT, QDSP, RHSZ (solveh.f:34-1142).
// One wave per block: the block's dynamic LDS (UC_PRD_LDS doubles per lane) is the
// second-level cache of the RA values uc_prd forms.
__global__ void __launch_bounds__(WAVE) UC_WAVES(2) UC_HOT k_sources(const UcParams* __restrict__ Pg, SourcesArgs a) {
  extern __shared__ double uc_lds[];
  double* const lds = uc_lds + threadIdx.x;
  UcP P = (UcP)Pg;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.npad) return;
  bool valid;
  const int64_t c = list_cell(a.cells, q, valid);
  if (!valid) return;
  const GridTab& G = a.gt[a.gid[q]];
  RcSlab rc = tile_slab((double*)a.slabs, q / WAVE);
  const int64_t n = a.n;
  const double rho = a.rho[c], tk = a.tk[c], ht = a.ht[c], stmf = a.stmf[c], stnd = a.stnd[c];
  const double dt = G.dt;
  double s[54];
  if (a.given) {
#pragma unroll
    for (int k = 0; k < 54; ++k) s[k] = a.given[(int64_t)k * n + c];
  } else if (a.fy) {
#pragma unroll
    for (int k = 0; k < NSP; ++k) s[k] = a.fy[(int64_t)k * n + c];
    s[52] = stmf;
    s[53] = stnd;
  } else {
    const int32_t w = a.nsw[q];
    const int32_t m = w & NSW_MASK;
    const bool fx = (w & NSW_FIXED) != 0;
    const bool global = G.mode != UC_MODE_PERCELL;
    const int32_t nstar = G.nstar;
    const bool use_st = !global || (fx && m - 1 <= nstar) || (!fx && m == nstar);
    const FieldBuf SRC = use_st ? field_buf(a.st, a.npad * 8 * 54, q)
                                : field_buf(a.snap + (int64_t)(nstar - G.snap_from) * 54 * a.npad,
                                            a.npad * 8 * 54, q);
    const double sor1 = G.sor1, sor2 = G.sor2;
#pragma unroll
    for (int k = 0; k < 54; ++k) s[k] = SRC.ld(k * a.npad);
    if (global && fx && m - 1 <= nstar) {
      // a fixed cell's species repeat; only SST2 keeps relaxing (solvesp.f:2820-2831)
      if (nstar == m - 1) {
        s[53] = a.sst2p[q];
      } else {
        const double b = (rho * stnd + dt * rc.ld(UC_SLOT_SOUR + 1) * s[21]) * sor1
                         / (rho + dt * rc.ld(UC_SLOT_SOUR + 5));
        for (int32_t i = m; i < nstar; ++i) {
          const double fnew = sor2 * s[53] + b;
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
  }
  // SOLVEH sources with the final iterates (solveh.f:34-1142)
  const double T = tk * P->tstr;
  const double T2 = T * T, T3 = T * T2, T4 = T * T3, T5 = T * T4;
  const int ip = T > 1000.0 ? 7 : 0;
  const double hconvr = P->gasc / P->wmstr / P->hstr;
  double tmole = 0.0;
#pragma unroll
  for (int k = 0; k < NSP; ++k) tmole += s[k] / P->wm[k];
  const double tkdi = 1000.0 / T, tkdis = tkdi * tkdi;
  const double p1 = (double)6.6334f + T * (-3.5686e-03 + T * (1.6682e-08 + T * (2.5611e-10 - 2.6558e-14 * T)));
  const double p2 = 1.8741e+01 - tkdi * (1.2131e+02 - 2.7350e+02 * tkdi + 1.9405e+02 * tkdis
                                         - 5.6310e+01 * tkdis * tkdi + (double)5.81690f * tkdis * tkdis);
  const double p3 = T <= 750.0
      ? (double)4.7869f + T * (-6.953e-02 + T * (2.95775e-04 + T * (-4.25732e-07 + 2.02894e-10 * T)))
      : 1.009e+01 + T * (-1.183e-02 + T * (4.7753e-06 + T * (-5.87209e-10 - 2.5334e-14 * T)));
  const double p4 = -(double)0.23093f - tkdi * ((double)1.12390f - (double)9.41530f * tkdi
                    + (double)2.99880f * tkdis - (double)0.51382f * tkdis * tkdi + 1.86840e-05 * tkdis * tkdis);
  const double plankc = (p1 * s[12] / P->wm[12] + p2 * s[9] / P->wm[9] + p3 * s[8] / P->wm[8]
                         + p4 * s[5] / P->wm[5]) / tmole;
  const double crad = P->alstr / P->rstr / P->vstr / P->vstr / P->vstr;
  const double qrad = crad * (double)0.4f * (4.0 * 5.669e-08 * plankc * (pow(T, 4.0) - 8.1e+09)
                                            + 3.334e-04 * stmf * (rho / 1.9e+03) * T5);
  double rhsz = a.flux ? a.flux[q] - qrad * dt : -qrad * dt;  // solveh.f:113-115
  const double rconv = P->vstr / P->wmstr / P->alstr * 1.0e-06;
  double prod = 0.0, hrr = 0.0;
  const FieldBuf QD = field_buf(a.qdsp, n * 8 * NSP, c);
  const bool hiT = ip != 0;
  uc_prd(P, rc, s, lds, [&](UcP Q, int k, double pr) {
    // per-species enthalpy HENTH (solveh.f:39-57) with this cell's NASA range
    double pk[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) pk[i] = hiT ? Q->nasa[k][7 + i] : Q->nasa[k][i];
    const double hssp = (pk[0] * T + pk[1] * T2 / 2.0 + pk[2] * T3 / 3.0 + pk[3] * T4 / 4.0
                         + pk[4] * T5 / 5.0) / Q->wm[k] * hconvr;
    const double henth = hssp * Q->wmstr * Q->hstr + pk[5] / Q->wm[k] * Q->gasc;
    // PRD(50..52) = 0.0 as written (solveh.f:1129-1131), not WM*HFO*0.0: the sign of a zero
    // QDSP follows the reference's expression
    const double prd = k >= 49 ? 0.0 : Q->wm[k] * Q->hfo[k] * pr;
    prod += prd;
    hrr += prd * henth / Q->hfo[k];
    QD.st(k * n, -rconv * prd / Q->hfo[k]);
  });
  rhsz = rhsz + rho * ht - dt * prod;
  a.qdot[c] = -hrr * rconv;
  if (!a.rhsz) return;  // QDOT/QDSP only (cells the step does not advance)
  a.rhsz[q] = rhsz;
  const FieldBuf SA = field_buf(a.sa, a.npad * 8 * NSP, q);
#pragma unroll
  for (int k = 0; k < NSP; ++k) SA.st(k * a.npad, s[k]);
  a.sst[q] = s[52];
  a.sst[a.npad + q] = s[53];
}

// The rate coefficients RC(1..544) of the listed cells from their slabs into a field-layout
// array rcf [544][n], which keeps each cell's coefficients of the last step that advanced it,
// as /DUMR/ RC does for the reference (CKRATES writes RC only where ISKIP < 3, ckrates.f:94).
__global__ void k_slab_to_rc(const double* __restrict__ slabs, const int32_t* __restrict__ cells, int64_t npad,
                             int64_t n, double* __restrict__ rcf) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npad) return;
  bool valid;
  const int64_t c = list_cell(cells, q, valid);
  if (!valid) return;
  const double* tile = slabs + (q / WAVE) * (int64_t)SLAB_SLOTS * WAVE + (q % WAVE);
  for (int r = 0; r < NRX; ++r) rcf[(int64_t)r * n + c] = tile[(int64_t)r * WAVE];
}


struct HresArgs {
  const double* rho; const double* ht; const double* rhsz;
  const int32_t* cells; const int32_t* gid; const GridTab* gt;
  int64_t npad;
  double* part;  // [H][ntiles] per-tile sums of enthalpy sweeps hfrom .. hfrom+H-1
  int32_t H, ntiles;
};

// enthalpy relaxation residual trajectory of every cell (solveh.f:1147-1167):
// sweeps G.hfrom .. G.hfrom+H-1 of its grid, per-tile sums (zero-transport: PPP = RHO)
__global__ void __launch_bounds__(256) k_hres(HresArgs a) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.npad) return;
  const int64_t tile = q / WAVE;
  const GridTab& G = a.gt[wave_grid(a.gid, q)];
  if (G.hfrom < 0) return;  // grid decided
  bool valid;
  const int64_t c = list_cell(a.cells, q, valid);
  double fpz = a.ht[c];
  const double inc = a.rhsz[q] * G.zsor1 / a.rho[c];
  const double zsor2 = G.zsor2;
  bool active = valid;
  for (int32_t sw = 0; sw < G.hfrom + a.H; ++sw) {
    double d = 0.0;
    if (active) {
      const double fnew = zsor2 * fpz + inc;
      d = fabs(fpz - fnew);
      fpz = fnew;
      if (d == 0.0) active = false;
    }
    if (sw >= G.hfrom) {
      const double ws = wave_sum(d);
      if ((threadIdx.x & (WAVE - 1)) == 0) a.part[(int64_t)(sw - G.hfrom) * a.ntiles + tile] = ws;
    }
  }
}

struct FinishArgs {
  double* rho; double* y; double* tk; double* ht; double* stmf; double* stnd;
  const double* sa; const double* sst; const double* rhsz;  // [52][npad], [2][npad], [npad]
  const int32_t* cells; const int32_t* gid; const GridTab* gt;
  int64_t npad; int64_t n;
  double* res_part;  // [2][ntiles] per-tile RESDC, RESDH partial sums
  int32_t ntiles;
  double* fpz_out;   // null: SPGSOLV epilogue + density; else FPZ here and TK only
  const double* fpz_in;  // [npad] relaxed FPZ of the coupled relaxation (null: relax here)
  const double* u; const double* v; const double* w;  // [n] velocities of T(h) (null: zero)
  int32_t* fails;    // [3][max_grids]: cells out of limit, Newton failures, first failed cell
  uint8_t* thflag;   // [n] per cell: 1 out of limit, 2 no root (written at every listed cell)
  int32_t max_grids;
};

// One cell's enthalpy relaxation, T(h) and SPGSOLV epilogue as k_finish forms them, on
// the cell's 8 adjacent lanes (part = lane % 8): lane p < 6 sums coefficient A(p+1) of
// T(h) over the species in the reference's order (dividing by p+1 where the reference
// does; the other sums divide by 1.0, which is exact), the sums reach the cell's other
// lanes by shuffles, every lane runs the rest of k_finish's arithmetic and part 0 writes.
// The results are k_finish's bits; this copy of its arithmetic exists because the
// one-lane kernel sits at the register file's limit and any shared form of the two
// spilled it.  limit/noroot: the conditions of solveh.f:1200-1204 and :1217-1219.
__device__ __forceinline__ void finish_cell_p8(UcP P, const FinishArgs& a, int64_t q, int64_t c, int part,
                                               const GridTab& G, bool& limit, bool& noroot, double& rc_, double& rh_) {
#pragma clang fp contract(off)   // the reference's operations in its order
  const bool writer = part == 0;
  const int64_t n = a.n;
  const FieldBuf SA = field_buf(a.sa, a.npad * 8 * NSP, q), Y = field_buf(a.y, n * 8 * NSP, c);
  double s[NSP];
#pragma unroll
  for (int k = 0; k < NSP; ++k) s[k] = SA.ld(k * a.npad);
  const double rho = a.rho[c], ht = a.ht[c];
  // enthalpy relaxation (solveh.f:1147-1167), zero transport: PPP = RHO
  double fpz = ht;
  if (a.fpz_in) {
    fpz = a.fpz_in[q];
  } else {
    const double inc = a.rhsz[q] * G.zsor1 / rho;
    const bool percell = G.mode == UC_MODE_PERCELL;
    for (int32_t sw = 0; sw < G.nh; ++sw) {
      const double fnew = G.zsor2 * fpz + inc;
      const double d = fabs(fpz - fnew);
      fpz = fnew;
      if (percell && d <= G.tolrh) break;
    }
  }
  // T from h (solveh.f:1170-1230); ENTH = (FPZ - (U*U+V*V+W*W)/2)*HSTR + sum Y*HFO*HSTR
  const double uc = a.u ? a.u[c] : 0.0, vc = a.v ? a.v[c] : 0.0, wc = a.w ? a.w[c] : 0.0;
  double enth = (fpz - 0.5 * (uc * uc + vc * vc + wc * wc)) * P->hstr;
#pragma unroll
  for (int k = 0; k < NSP; ++k) enth += s[k] * P->hfo[k] * P->hstr;
  bool hiT = false;
  double tpoly = 1000.0;
  double tkd = a.tk[c] * P->tstr;
  double a1, a2, a3, a4, a5, a6;
  for (;;) {
    if (tkd > tpoly) tkd = tpoly / 2.0;
    asm volatile("" : "+s"(P));
    {
      // A(i) = sum over k of POLSP(i,k)*Y(k)/WM(k)/i (/1 for A1 and A6: exact), lane i-1
      const int i = part < 6 ? part : 5;
      const double div = i == 0 || i == 5 ? 1.0 : (double)(i + 1);
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        const double wk = P->wm[k];
        // this lane's coefficient from the row's scalar loads (i and hiT vary by lane)
        double cf = hiT ? P->nasa[k][7] : P->nasa[k][0];
#pragma unroll
        for (int m = 1; m < 6; ++m) cf = i == m ? (hiT ? P->nasa[k][7 + m] : P->nasa[k][m]) : cf;
        acc = acc + cf * s[k] / wk / div;
      }
      a1 = __shfl(acc, 0, 8);
      a2 = __shfl(acc, 1, 8);
      a3 = __shfl(acc, 2, 8);
      a4 = __shfl(acc, 3, 8);
      a5 = __shfl(acc, 4, 8);
      a6 = __shfl(acc, 5, 8);
    }
    const double funt = ((a1 + (a2 + (a3 + (a4 + a5 * tpoly) * tpoly) * tpoly) * tpoly) * tpoly + a6)
                        * P->gasc / P->wmstr - enth;
    if (funt >= 1.0e-07) break;
    if (hiT) { limit = true; break; }   // above the high-range fit at TPOL2 (message 712)
    hiT = true;
    if (tkd <= tpoly) tkd = 5000.0 / 2.0;
    tpoly = 5000.0;
  }
  const double eh = enth * P->wmstr / P->gasc;
  noroot = true;   // until a correction falls within the tolerance (message 714 otherwise)
  for (int itn = 1; itn <= 100; ++itn) {
    const double dtkd = ((a1 + (a2 + (a3 + (a4 + a5 * tkd) * tkd) * tkd) * tkd) * tkd + a6 - eh)
                        / (a1 + (2.0 * a2 + (3.0 * a3 + (4.0 * a4 + 5.0 * a5 * tkd) * tkd) * tkd) * tkd);
    if (fabs(dtkd) <= (double)0.001f) { noroot = false; break; }
    tkd = tkd - dtkd;
  }
  if (tkd > 5000.0) tkd = 5000.0;
  if (tkd < P->tstr - 1.0) tkd = P->tstr - 1.0;
  const double tnew = tkd / P->tstr;
  if (a.fpz_out) {  // SOLVEH's outputs only (solveh.f:1167, :1230)
    if (writer) {
      a.fpz_out[c] = fpz;
      a.tk[c] = tnew;
    }
    return;
  }
  if (!writer) return;
  // SPGSOLV epilogue (spgsolv.f:1299-1416)
  const double sst1 = a.sst[q], sst2 = a.sst[a.npad + q];
  rh_ = fabs(fpz - ht);
  rc_ = fabs(sst1 - a.stmf[c]);
  double rbar = 0.0;
#pragma unroll
  for (int k = 0; k < NSP; ++k) {
    if (k < 51) rc_ += fabs(s[k] - Y.ld(k * n));
    Y.st(k * n, s[k]);
    rbar = rbar + s[k] / P->wm[k];   // main:1058
  }
  a.ht[c] = fpz;
  a.stmf[c] = sst1;
  a.stnd[c] = sst2;
  a.tk[c] = tnew;
  a.rho[c] = P->pback / (P->beta1 * tnew * rbar);  // main:1056-1061
}

// T(h) failure counts of a grid and its first failing cell (field index within the grid)
__device__ __forceinline__ void count_fails(const FinishArgs& a, int32_t g, const GridTab& G, int64_t q, bool limit,
                                            bool noroot) {
  const unsigned long long bl = __ballot(limit), bn = __ballot(noroot);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    if (bl) atomicAdd(&a.fails[g], (int32_t)__popcll(bl));
    if (bn) atomicAdd(&a.fails[a.max_grids + g], (int32_t)__popcll(bn));
  }
  if (noroot) {
    bool valid;
    const int64_t c = list_cell(a.cells, q, valid);
    atomicMin(&a.fails[2 * a.max_grids + g], (int32_t)(c - G.off));
  }
}

__global__ void __launch_bounds__(512) UC_WAVES(3) UC_HOT k_finish(const UcParams* __restrict__ Pg, FinishArgs a) {
#pragma clang fp contract(off)   // the reference's operations in its order
  UcP P = (UcP)Pg;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.npad) return;
  bool valid;
  const int64_t c = list_cell(a.cells, q, valid);
  const GridTab& G = a.gt[wave_grid(a.gid, q)];
  double rc_ = 0.0, rh_ = 0.0;
  bool limit = false, noroot = false;   // the conditions of solveh.f:1200-1204 and :1217-1219
  if (valid) {
    const int64_t n = a.n;
    const FieldBuf SA = field_buf(a.sa, a.npad * 8 * NSP, q), Y = field_buf(a.y, n * 8 * NSP, c);
    double s[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) s[k] = SA.ld(k * a.npad);
    const double rho = a.rho[c], ht = a.ht[c];
    // enthalpy relaxation (solveh.f:1147-1167), zero transport: PPP = RHO
    double fpz = ht;
    if (a.fpz_in) {
      fpz = a.fpz_in[q];
    } else {
      const double inc = a.rhsz[q] * G.zsor1 / rho;
      const bool percell = G.mode == UC_MODE_PERCELL;
      for (int32_t sw = 0; sw < G.nh; ++sw) {
        const double fnew = G.zsor2 * fpz + inc;
        const double d = fabs(fpz - fnew);
        fpz = fnew;
        if (percell && d <= G.tolrh) break;
      }
    }
    // T from h (solveh.f:1170-1230); ENTH = (FPZ - (U*U+V*V+W*W)/2)*HSTR + sum Y*HFO*HSTR
    const double uc = a.u ? a.u[c] : 0.0, vc = a.v ? a.v[c] : 0.0, wc = a.w ? a.w[c] : 0.0;
    double enth = (fpz - 0.5 * (uc * uc + vc * vc + wc * wc)) * P->hstr;
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
      for (int k = 0; k < NSP; ++k) {   // A1..A6 in the reference's order (solveh.f:1190-1197)
        const double wk = P->wm[k];
        a1 = a1 + (hiT ? P->nasa[k][7] : P->nasa[k][0]) * s[k] / wk;
        a2 = a2 + (hiT ? P->nasa[k][8] : P->nasa[k][1]) * s[k] / wk / 2.0;
        a3 = a3 + (hiT ? P->nasa[k][9] : P->nasa[k][2]) * s[k] / wk / 3.0;
        a4 = a4 + (hiT ? P->nasa[k][10] : P->nasa[k][3]) * s[k] / wk / 4.0;
        a5 = a5 + (hiT ? P->nasa[k][11] : P->nasa[k][4]) * s[k] / wk / 5.0;
        a6 = a6 + (hiT ? P->nasa[k][12] : P->nasa[k][5]) * s[k] / wk;
      }
      const double funt = ((a1 + (a2 + (a3 + (a4 + a5 * tpoly) * tpoly) * tpoly) * tpoly) * tpoly + a6)
                          * P->gasc / P->wmstr - enth;
      if (funt >= 1.0e-07) break;
      if (hiT) { limit = true; break; }   // above the high-range fit at TPOL2 (message 712)
      hiT = true;
      if (tkd <= tpoly) tkd = 5000.0 / 2.0;
      tpoly = 5000.0;
    }
    const double eh = enth * P->wmstr / P->gasc;
    noroot = true;   // until a correction falls within the tolerance (message 714 otherwise)
    for (int itn = 1; itn <= 100; ++itn) {
      const double dtkd = ((a1 + (a2 + (a3 + (a4 + a5 * tkd) * tkd) * tkd) * tkd) * tkd + a6 - eh)
                          / (a1 + (2.0 * a2 + (3.0 * a3 + (4.0 * a4 + 5.0 * a5 * tkd) * tkd) * tkd) * tkd);
      if (fabs(dtkd) <= (double)0.001f) { noroot = false; break; }
      tkd = tkd - dtkd;
    }
    if (tkd > 5000.0) tkd = 5000.0;
    if (tkd < P->tstr - 1.0) tkd = P->tstr - 1.0;
    const double tnew = tkd / P->tstr;
    if (a.fpz_out) {  // SOLVEH's outputs only (solveh.f:1167, :1230)
      a.fpz_out[c] = fpz;
      a.tk[c] = tnew;
      goto done;
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
      rbar = rbar + s[k] / P->wm[k];   // main:1058
    }
    a.ht[c] = fpz;
    a.stmf[c] = sst1;
    a.stnd[c] = sst2;
    a.tk[c] = tnew;
    a.rho[c] = P->pback / (P->beta1 * tnew * rbar);  // main:1056-1061
    }
  }
done:
  if (valid) a.thflag[c] = (uint8_t)((limit ? 1 : 0) | (noroot ? 2 : 0));
  {
    const int32_t g = wave_grid(a.gid, q);
    const unsigned long long bl = __ballot(limit), bn = __ballot(noroot);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
      if (bl) atomicAdd(&a.fails[g], (int32_t)__popcll(bl));
      if (bn) atomicAdd(&a.fails[a.max_grids + g], (int32_t)__popcll(bn));
    }
    if (noroot) atomicMin(&a.fails[2 * a.max_grids + g], (int32_t)(c - G.off));
  }
  rc_ = wave_sum(rc_);
  rh_ = wave_sum(rh_);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    a.res_part[q / WAVE] = rc_;
    a.res_part[a.ntiles + q / WAVE] = rh_;
  }
}

// The same with 8 lanes per cell for a few tiles, where one lane's chain of divisions
// through the T(h) coefficients bounds the step: a block of 512 threads is one tile, cell
// slot threadIdx.x / 8.  The tile's residual partials are summed by one wave in the
// one-lane kernel's order.
__global__ void __launch_bounds__(512) UC_HOT k_finish_p8(const UcParams* __restrict__ Pg, FinishArgs a) {
  __shared__ double part_rc[WAVE], part_rh[WAVE];
  UcP P = (UcP)Pg;
  const int slot = threadIdx.x / 8, part = threadIdx.x % 8;
  const int64_t tile = blockIdx.x;
  const int64_t q = tile * WAVE + slot;   // grid: npad / 64 blocks, so q < npad
  const int32_t g = wave_grid(a.gid, q);   // a tile belongs to one grid
  const GridTab& G = a.gt[g];
  double rc_ = 0.0, rh_ = 0.0;
  bool limit = false, noroot = false;
  bool valid;
  const int64_t c = list_cell(a.cells, q, valid);
  if (valid) finish_cell_p8(P, a, q, c, part, G, limit, noroot, rc_, rh_);
  if (valid && part == 0) a.thflag[c] = (uint8_t)((limit ? 1 : 0) | (noroot ? 2 : 0));
  count_fails(a, g, G, q, limit && part == 0, noroot && part == 0);
  if (part == 0) {
    part_rc[slot] = rc_;
    part_rh[slot] = rh_;
  }
  __syncthreads();
  if (threadIdx.x < WAVE) {
    const double r1 = wave_sum(part_rc[threadIdx.x]), r2 = wave_sum(part_rh[threadIdx.x]);
    if (threadIdx.x == 0) {
      a.res_part[tile] = r1;
      a.res_part[a.ntiles + tile] = r2;
    }
  }
}

// density of every cell the chemistry did not update (main:1056-1061)
__global__ void k_density(const UcParams* __restrict__ P, double* __restrict__ rho, const double* __restrict__ y,
                          const double* __restrict__ tk, const uint8_t* __restrict__ done, int64_t n) {
#pragma clang fp contract(off)   // the reference's operations in its order
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || done[c]) return;
  double rbar = 0.0;
  for (int k = 0; k < NSP; ++k) rbar = rbar + y[k * n + c] / P->wm[k];   // main:1058
  rho[c] = P->pback / (P->beta1 * tk[c] * rbar);
}

// main:1062-1068, per grid (block g): last I column extrapolated, then the J=1 row
// copied from J=2
__global__ void k_density_edge(double* __restrict__ rho, const GridTab* __restrict__ gt) {
  const GridTab& G = gt[blockIdx.x];
  double* r = rho + G.off;
  const int64_t li = G.li, lj = G.lj, n = li * lj;
  for (int64_t j = threadIdx.x; j < lj; j += blockDim.x) {
    const int64_t c = n - lj + j;
    r[c] = 2.0 * r[c - lj] - r[c - 2 * lj];
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < li; i += blockDim.x) r[i * lj] = r[i * lj + 1];
}

__global__ void k_zero(double* __restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ---- the coupled step (transport coefficients non-zero): the reference's lexicographic
// Gauss-Seidel (J outer, I inner) as a pipelined anti-diagonal wavefront.  A cell
// (i, j) of sweep s reads (i-1, j) and (i, j-1) of sweep s and (i+1, j), (i, j+1) of
// sweep s-1, so wavefront step t updates diagonal d = t - 2s (d = i + j - 2) of every
// sweep s of the segment at once: the diagonals of one step are never adjacent, every
// value a cell reads was written in an earlier step, and the result is the
// reference's sweep order exactly.  The cell list is in diagonal order (j ascending);
// dfirst[d] .. dfirst[d+1]-1 are the list positions of diagonal d.

// SA = FA, SST1 = STMF, SST2 = STND and FPZ = HT at every cell (spgsolv.f:169-231)
// Rate coefficients held in the field layout RC(LE, 544) (ckrates.f:27) into the slabs of
// the listed cells, for SOLVEH called without CKRATES (spgsolv.f:1291-1298 with ICHEM <= 0),
// which reads whatever /DUMR/ holds.
__global__ void k_rc_to_slab(const double* __restrict__ rcf, const int32_t* __restrict__ cells, int64_t npad,
                             int64_t n, double* __restrict__ slabs) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npad) return;
  bool valid;
  const int64_t c = list_cell(cells, q, valid);
  if (!valid) return;
  double* tile = slabs + (q / WAVE) * (int64_t)SLAB_SLOTS * WAVE + (q % WAVE);
  for (int r = 0; r < NRX; ++r) tile[(int64_t)r * WAVE] = rcf[(int64_t)r * n + c];
}

__global__ void k_cpl_init(const double* __restrict__ y, const double* __restrict__ stmf,
                           const double* __restrict__ stnd, const double* __restrict__ ht, int64_t n,
                           double* __restrict__ gs, double* __restrict__ gz) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  for (int k = 0; k < NSP; ++k) gs[k * n + c] = y[k * n + c];
  gs[52 * n + c] = stmf[c];
  gs[53 * n + c] = stnd[c];
  gz[c] = ht[c];
}

// The cell's coupling tile: per group g (species 1..51, the soot scalars, the enthalpy)
// PPP = RHO-(E+W+P+Q) (solvesp.f:424, :2812; solveh.f:1154) and P, Q, E, W.  coef:
// [204][n] E, W, P, Q per species (/DUMMY/ + /DUM0/ order); stc, zc: [4][n] E, W, N(P),
// S(Q).  A null array is zero.
__global__ void k_cpl_coef(const double* __restrict__ rho, const int32_t* __restrict__ cells, int64_t npad,
                           int64_t n, const double* __restrict__ coef, const double* __restrict__ stc,
                           const double* __restrict__ zc, double* __restrict__ cplc) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npad) return;
  bool valid;
  const int64_t c = list_cell(cells, q, valid);
  double* t = cplc + (q / WAVE) * (int64_t)UC_CPL_SLOTS * WAVE + (q & (WAVE - 1));
  const double r = rho[c];
  for (int g = 0; g < 53; ++g) {
    const double* a = g < 51 ? (coef ? coef + (int64_t)4 * g * n : nullptr) : (g == 51 ? stc : zc);
    const double e = a ? a[c] : 0.0, w = a ? a[n + c] : 0.0, p = a ? a[2 * n + c] : 0.0,
                 qq = a ? a[3 * n + c] : 0.0;
    t[(5 * g + 0) * WAVE] = r - (e + w + p + qq);
    t[(5 * g + 1) * WAVE] = p;
    t[(5 * g + 2) * WAVE] = qq;
    t[(5 * g + 3) * WAVE] = e;
    t[(5 * g + 4) * WAVE] = w;
  }
}

struct WaveArgs {
  const double* rho; const double* stmf; const double* stnd;
  const int32_t* cells; const int32_t* dfirst; const GridTab* gt;
  int64_t npad, n;
  int32_t lj, ndiag, maxtiles, ntiles;
  int32_t t;          // wavefront step
  int32_t S;          // sweeps of the segment
  double* slabs; const double* cplc;
  double* gs;         // [54][n] species iterates, SST1, SST2 (field layout)
  double* snap;       // [S][54][npad] iterates after each sweep of the segment
  double* rsd_part;   // rows start+s, accumulated per tile (a tile may span diagonals)
  // enthalpy (k_hsweep_cpl)
  const double* rhsz; double* gz; double* zsnap;  // [npad], [n], [S][npad]
};

// Block (si, ti): tile ti of diagonal t - 2*si in sweep start+si; lanes outside the
// diagonal compute but write nothing.
__device__ __forceinline__ bool wave_cell(const WaveArgs& a, int32_t& si, int64_t& tile, int64_t& q,
                                          bool& mine, int64_t& c) {
  si = (int32_t)blockIdx.x / a.maxtiles;
  const int32_t ti = (int32_t)blockIdx.x % a.maxtiles;
  const int32_t d = a.t - 2 * si;
  if (si >= a.S || d < 0 || d >= a.ndiag) return false;
  const int32_t q0 = a.dfirst[d], q1 = a.dfirst[d + 1];
  tile = q0 / WAVE + ti;
  if (tile * WAVE >= q1) return false;
  q = tile * WAVE + (threadIdx.x & (WAVE - 1));
  bool valid;
  c = list_cell(a.cells, q, valid);
  mine = valid && q >= q0 && q < q1;
  return true;
}

__global__ void __launch_bounds__(WAVE) UC_HOT k_sweep_cpl(const UcParams* __restrict__ Pg, WaveArgs a) {
  extern __shared__ double uc_lds[];
  double* const lds = uc_lds + threadIdx.x;
  UcP P = (UcP)Pg;
  int32_t si;
  int64_t tile, q, c;
  bool mine;
  if (!wave_cell(a, si, tile, q, mine, c)) return;
  const GridTab& G = a.gt[0];
  const int64_t n = a.n;
  const FieldBuf GS = field_buf(a.gs, n * 8 * 54, c);
  double s[54];
#pragma unroll
  for (int k = 0; k < 54; ++k) s[k] = GS.ld(k * n);
  const int64_t lj = a.lj;
  RcCpl rc{tile_slab(a.slabs, tile), tile_rsrc(a.cplc, tile, UC_CPL_SLOTS),
           __builtin_amdgcn_make_buffer_rsrc((void*)a.gs, 0, (int)(n * 8 * 54), 0x00020000),
           {(uint32_t)((c + 1) * 8), (uint32_t)((c - 1) * 8), (uint32_t)((c + lj) * 8), (uint32_t)((c - lj) * 8)},
           (int32_t)(n * 8)};
  double d2;
  const double r = uc_sweep_cpl(P, rc, lds, s, a.rho[c], a.stmf[c], a.stnd[c], G.dt, G.sor1, G.sor2, &d2);
  if (mine) {
#pragma unroll
    for (int k = 0; k < 54; ++k) GS.st(k * n, s[k]);
    const FieldBuf SN = field_buf(a.snap + (int64_t)si * 54 * a.npad, a.npad * 8 * 54, q);
#pragma unroll
    for (int k = 0; k < 54; ++k) SN.st(k * a.npad, s[k]);
  }
  const double ws = wave_sum(mine ? r : 0.0);
  if ((threadIdx.x & (WAVE - 1)) == 0) a.rsd_part[(int64_t)(G.start + si) * a.ntiles + tile] += ws;
}

// enthalpy: FNEW = SOR2*FPZ + (RHSZ-ZN*FPZ(J+1)-ZS*FPZ(J-1)-ZE*FPZ(I+1)-ZW*FPZ(I-1))*SOR1/PPP
// (solveh.f:1154-1160); rows of rsd_part are the segment's sweeps (hfrom based)
__global__ void __launch_bounds__(WAVE) k_hsweep_cpl(WaveArgs a) {
  int32_t si;
  int64_t tile, q, c;
  bool mine;
  if (!wave_cell(a, si, tile, q, mine, c)) return;
  const GridTab& G = a.gt[0];
  double dz = 0.0;
  if (mine) {
    const double* t = a.cplc + tile * (int64_t)UC_CPL_SLOTS * WAVE + (q & (WAVE - 1));
    const int64_t lj = a.lj;
    const double ppp = t[260 * WAVE], zn = t[261 * WAVE], zs = t[262 * WAVE], ze = t[263 * WAVE],
                 zw = t[264 * WAVE];
    const double fpz = a.gz[c];
    double num = a.rhsz[q];
    num = num - zn * a.gz[c + 1];
    num = num - zs * a.gz[c - 1];
    num = num - ze * a.gz[c + lj];
    num = num - zw * a.gz[c - lj];
    const double fnew = G.zsor2 * fpz + num * G.zsor1 / ppp;
    dz = fabs(fpz - fnew);
    a.gz[c] = fnew;
    a.zsnap[(int64_t)si * a.npad + q] = fnew;
  }
  const double ws = wave_sum(dz);
  if ((threadIdx.x & (WAVE - 1)) == 0) a.rsd_part[(int64_t)si * a.ntiles + tile] += ws;
}

// N2 closure (solvesp.f:2839-2914) of the iterates at the decided count, into gs: SSP
// of every cell for the enthalpy flux (active cells closed, the others as entered)
__global__ void k_cpl_close(const double* __restrict__ src, const int32_t* __restrict__ cells, int64_t npad,
                            int64_t n, double* __restrict__ gs) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npad) return;
  bool valid;
  const int64_t c = list_cell(cells, q, valid);
  if (!valid) return;
  double s[54];
  for (int k = 0; k < 54; ++k) s[k] = src[k * npad + q];
  double ftot = 0.0;
  for (int k = 0; k < 51; ++k) ftot += s[k];
  ftot += s[52];
  s[51] = 1.0 - ftot;
  if (ftot > 1.0) {
    for (int k = 0; k < 51; ++k) s[k] = s[k] / ftot;
    s[52] = s[52] / ftot;
    s[51] = 0.0;
  }
  for (int k = 0; k < 54; ++k) gs[k * n + c] = s[k];
}

// RDHSP = (EMU*VDF + REI*TMU/SIGSP - EKT)*HSSP at every cell (solveh.f:34-57); null
// inputs are zero, a null emu is REI*BETA3*RHO (spgsolv.f:175)
__global__ void k_cpl_rdhsp(const UcParams* __restrict__ Pg, const double* __restrict__ tk,
                            const double* __restrict__ rho, const double* __restrict__ emu,
                            const double* __restrict__ vdf, const double* __restrict__ tmu,
                            const double* __restrict__ ekt, double sigsp, int64_t n, double* __restrict__ rd) {
#pragma clang fp contract(off)   // the reference's operations in its order
  UcP P = (UcP)Pg;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double T = tk[c] * P->tstr;
  const double T2 = T * T, T3 = T * T2, T4 = T * T3, T5 = T * T4;
  const bool hiT = T > 1000.0;
  const double hconvr = P->gasc / P->wmstr / P->hstr;
  const double e = emu ? emu[c] : P->rei * P->beta3 * rho[c];
  const double tm = tmu ? tmu[c] : 0.0, ek = ekt ? ekt[c] : 0.0;
  for (int k = 0; k < NSP; ++k) {
    double pk[5];
    for (int i = 0; i < 5; ++i) pk[i] = hiT ? P->nasa[k][7 + i] : P->nasa[k][i];
    const double hssp = (pk[0] * T + pk[1] * T2 / 2.0 + pk[2] * T3 / 3.0 + pk[3] * T4 / 4.0 + pk[4] * T5 / 5.0)
                        / P->wm[k] * hconvr;   // solveh.f:49-51
    const double v = vdf ? vdf[k * n + c] : 0.0;
    rd[k * n + c] = (e * v + P->rei * tm / sigsp - ek) * hssp;
  }
}

// DT*0.5 times the species-diffusion enthalpy flux divergence (solveh.f:63-115) of every
// active cell, from RDHSP and the closed species (gs) of the cell and its neighbours
__global__ void k_cpl_flux(const double* __restrict__ rd, const double* __restrict__ gs,
                           const int32_t* __restrict__ cells, int64_t npad, int64_t n, int32_t lj,
                           const double* __restrict__ xxc, const double* __restrict__ xxs,
                           const double* __restrict__ yyc, const double* __restrict__ yys,
                           const double* __restrict__ ygeo, double sym, double dt, double* __restrict__ flux) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npad) return;
  bool valid;
  const int64_t ia = list_cell(cells, q, valid);
  if (!valid) return;
  const int64_t i = ia / lj, j = ia % lj;
  double rdhe = 0.0, rdhw = 0.0, rdhn = 0.0, rdhs = 0.0, rdhp = 0.0;
  for (int k = 0; k < NSP; ++k) {
    const double* r = rd + k * n;
    const double* y = gs + k * n;
    rdhe = rdhe + (r[ia + lj] + r[ia]) * (y[ia + lj] - y[ia]);
    rdhw = rdhw + (r[ia] + r[ia - lj]) * (y[ia] - y[ia - lj]);
    rdhn = rdhn + (r[ia + 1] + r[ia]) * (y[ia + 1] - y[ia]);
    rdhs = rdhs + (r[ia] + r[ia - 1]) * (y[ia] - y[ia - 1]);
    rdhp = rdhp + r[ia] * (y[ia + 1] - y[ia - 1]);
  }
  const double x = (rdhe / xxs[i + 1] - rdhw / xxs[i]) / xxc[i] + (rdhn / yys[j + 1] - rdhs / yys[j]) / yyc[j]
                   + sym * rdhp / ygeo[ia] / yyc[j];
  flux[q] = dt * 0.5 * x;
}

// ---- batches: per-grid field arrays <-> the packed arrays (grid g at offset off_g)
struct PackPtrs {
  double *rho, *y, *tk, *ht, *stmf, *stnd, *qdot, *qdsp;
  const int32_t* iskip;
};

// block (x, g): cells x*256.. of grid g; dir 0 packs (grid -> packed), 1 unpacks.  QDOT and
// QDSP are formed only at interior cells with ISKIP < qlim (3, or 10 where the kept rate
// coefficients give the cells SOLVEH alone visits theirs, solveh.f:119); elsewhere the
// grid's arrays keep what they held, as the reference's do.
__global__ void k_pack(const PackPtrs* __restrict__ grids, const GridTab* __restrict__ gt, PackPtrs pk,
                       int32_t* __restrict__ pk_iskip, int64_t N, int dir, int32_t qlim) {
  const int32_t g = blockIdx.y;
  const GridTab& G = gt[g];
  const int64_t ng = (int64_t)G.li * G.lj;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const PackPtrs& f = grids[g];
  const int64_t d = G.off + c;
  if (dir == 0) {
    pk.rho[d] = f.rho[c]; pk.tk[d] = f.tk[c]; pk.ht[d] = f.ht[c];
    pk.stmf[d] = f.stmf[c]; pk.stnd[d] = f.stnd[c]; pk_iskip[d] = f.iskip[c];
    for (int k = 0; k < NSP; ++k) pk.y[k * N + d] = f.y[k * ng + c];
  } else {
    f.rho[c] = pk.rho[d]; f.tk[c] = pk.tk[d]; f.ht[c] = pk.ht[d];
    f.stmf[c] = pk.stmf[d]; f.stnd[c] = pk.stnd[d];
    for (int k = 0; k < NSP; ++k) f.y[k * ng + c] = pk.y[k * N + d];
    const int64_t i = c / G.lj, j = c % G.lj;
    if (i > 0 && i < G.li - 1 && j > 0 && j < G.lj - 1 && pk_iskip[d] < qlim) {
      f.qdot[c] = pk.qdot[d];
      for (int k = 0; k < NSP; ++k) f.qdsp[k * ng + c] = pk.qdsp[k * N + d];
    }
  }
}

// ---- stiffness binning: sort keys (grid, bin of the previous step's sweeps)
__global__ void k_bin_keys(const int32_t* __restrict__ base, const int32_t* __restrict__ base_gid,
                           const int32_t* __restrict__ prev_nsw, int32_t nbase, int32_t bins, int32_t maxsw,
                           uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbase) return;
  const int32_t sw = prev_nsw[base[i]];
  int32_t b = (int32_t)((int64_t)sw * bins / (maxsw + 1));
  b = b < 0 ? 0 : (b >= bins ? bins - 1 : b);
  keys[i] = ((uint32_t)base_gid[i] << 6) | (uint32_t)(bins - 1 - b);  // stiffest first within a grid
  vals[i] = base[i];
}

// sorted (grid-major) cells into the padded list: grid g's cells start at tile0*64,
// padding lanes repeat the grid's first cell as -1-c
__global__ void k_bin_place(const int32_t* __restrict__ sorted_cells, const int32_t* __restrict__ base_off,
                            const GridTab* __restrict__ gt, int32_t ngrids, int32_t* __restrict__ cells,
                            int32_t* __restrict__ gid, int64_t npad) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npad) return;
  // the grid of list position q: binary search on tile0
  int32_t lo = 0, hi = ngrids - 1;
  const int32_t t = (int32_t)(q / WAVE);
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) / 2;
    if (gt[mid].tile0 <= t) lo = mid; else hi = mid - 1;
  }
  const int32_t g = lo;
  const int64_t r = q - (int64_t)gt[g].tile0 * WAVE;       // rank within the grid
  const int64_t cnt = base_off[g + 1] - base_off[g];
  gid[q] = g;
  cells[q] = r < cnt ? sorted_cells[base_off[g] + r] : -1 - sorted_cells[base_off[g]];
}

// after a step: each cell's sweeps, by field cell, for the next step's bins
__global__ void k_save_nsw(const int32_t* __restrict__ cells, const int32_t* __restrict__ nsw, int64_t npad,
                           int32_t* __restrict__ prev_nsw) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npad) return;
  const int32_t v = cells[q];
  if (v >= 0) prev_nsw[v] = nsw[q] & NSW_MASK;
}

// tile costs (stiffest bin of the tile) for the work-queue order
__global__ void k_tile_keys(const int32_t* __restrict__ cells, const int32_t* __restrict__ prev_nsw,
                            int32_t ntiles, uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  int32_t m = 0;
  for (int l = 0; l < WAVE; ++l) {
    const int32_t v = cells[t * WAVE + l];
    if (v >= 0) m = max(m, prev_nsw[v]);
  }
  keys[t] = 0x7fffffffu - (uint32_t)m;  // descending cost
  vals[t] = (int32_t)t;
}

}  // namespace

// ---------------------------------------------------------------- host side
namespace {
// bookkeeping of the step in progress (chem_host.inc: ph_begin .. ph_end)
struct StepRun {
  hipStream_t s = nullptr;
  int32_t ngrids = 0;
  bool clip = false;
  // coupled step (one grid): coupling inputs, which parts run as wavefronts, and the
  // relaxed FPZ the coupled enthalpy relaxation leaves for k_finish
  uc_coupling cp{};
  bool cpl = false, cpl_sp = false, cpl_h = false, cpl_flux = false, vel = false;
  const double* fpz_in = nullptr;
  int64_t cpl_sweeps = 0;
  float ms_rates = 0.f, ms_sweeps = 0.f, ms_sources = 0.f;
  int32_t sweep_launches = 0;
  std::vector<uc_field> fields;       // the caller's grids
  std::vector<uc_controls> ctl;       // their controls
  std::vector<double> spnorm, hnorm;  // global-stop norms per grid
  std::vector<uc_stats> out;          // per-grid statistics
  std::vector<std::vector<double>> hist;  // per grid: RSD/SPNORM of each species sweep the stop read
  bool hist_dev = false;              // the history is the device's (k_decide_sp), not `hist`
  bool span[3] = {false, false, false};  // rates, sweeps, sources timed by ctx->tev pairs
};

// the cell list of the current grids (rebuilt when the grids change)
struct Layout {
  int32_t ngrids = 0;
  std::vector<const void*> key;  // per grid: ISKIP pointer, li, lj
  int64_t N = 0;                 // packed cells
  int64_t npad = 0;              // list positions (grids padded to whole tiles)
  int32_t ntiles = 0;
  int32_t nbase = 0;             // active cells
  bool packed = false;           // batch: grids copied to the context's packed arrays
  bool have_prev = false;        // prev_nsw holds the previous step's sweeps
  bool diag = false;             // coupled step: cells in anti-diagonal order (ctx->cpl_dfirst)
  int32_t ndiag = 0, maxtiles = 0;  // diagonals; most tiles any diagonal's positions touch
  bool cells_natural = false;    // ctx->cells/gid hold the list in natural order (base_list)
};

// tuning of a step: the autotuner's choices (uc_tuning)
struct Tuning {
  int32_t bins = 1;        // stiffness bins of the cell order (1: natural order)
  int32_t lpt = 0;         // tiles stiffest first
  int32_t persistent = 1;  // k_sweep as persistent waves over the tile queue
  int32_t block = 256;     // block size of the one-thread-per-cell kernels
  int32_t graph = 0;       // sweep segments replayed from a captured hipGraph (measured slower than
                           // direct launches: two cross-stream event waits per replay)
  int32_t lanes = 1;       // lanes per cell of the chemistry-only sweep (1, 2, 4)
  int32_t rate_parts = 0;  // lanes per cell of the rate-coefficient kernel (1, 2, 4, 8; 0: by cell count)
};

// the sweep segment (grid-table upload, queue reset, k_sweep, per-grid row sums and
// their copy to the host) captured once per layout and tuning, replayed per segment
struct SegGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> key;
};
}  // namespace

#include "transport.inc"

struct uc_context {
  uc_scales sc{};
  bool clip = true;        // uc_step* start with main:821-828 (uc_set_clip)
  UcParams hp{};
  UcParams* dp = nullptr;
  int64_t max_cells = 0;   // packed cells over all grids
  int32_t max_grids = 1;
  int64_t max_pad = 0;     // list positions: max_cells + 64 per grid
  int device = 0;
  int cus = 0;
  int32_t nsnap = 0;       // snapshot slots per cell
  double* slabs = nullptr; // [max_pad/64][SLAB_SLOTS][64]
  double* st = nullptr;    // [54][max_pad]
  double* snap = nullptr;  // [nsnap][54][max_pad]
  double* sst2p = nullptr; // [max_pad]
  double* sa = nullptr;    // [52][max_pad]
  double* sst = nullptr;   // [2][max_pad]
  double* rhsz = nullptr;  // [max_pad]
  int32_t* nsw = nullptr;  // [max_pad]
  int32_t* cells = nullptr;  // [max_pad] packed field index, or -1-c for padding
  int32_t* gid = nullptr;    // [max_pad] grid of each list position
  int32_t* base = nullptr;   // [max_cells] active cells, grid-major natural order
  int32_t* base_gid = nullptr;
  int32_t* base_off = nullptr;  // [max_grids+1]
  int32_t* base_list = nullptr; // [max_pad] the natural-order padded list (bins off)
  int32_t* base_list_gid = nullptr;
  int32_t* prev_nsw = nullptr;  // [max_cells] previous step's sweeps per packed cell
  uint32_t* keys[2]{};          // radix sort buffers [max_cells]
  int32_t* vals[2]{};
  int32_t* order = nullptr;     // [max_tiles] tile order (the work queue)
  int32_t* queue = nullptr;     // persistent k_sweep: next queue position
  void* sort_tmp = nullptr;     // scratch of the radix sorts and of the cell selection
  size_t sort_tmp_bytes = 0;
  int32_t* nsel = nullptr;      // selected-cell count (device) and its pinned host copy
  int32_t* nsel_h = nullptr;
  uint8_t* done = nullptr;      // [max_cells]
  uint8_t* thflag = nullptr;    // [max_cells] T(h) conditions of the last step (k_finish)
  GridTab* gt_dev = nullptr;    // [max_grids]
  GridTab* gt = nullptr;        // pinned host mirror
  PackPtrs* grid_ptrs = nullptr;  // [max_grids] device copy of the grids' pointers (batches)
  PackPtrs pk{};                // packed field arrays (batches; allocated on first use)
  int32_t* pk_iskip = nullptr;
  double* rsd_part = nullptr;   // [128][max_tiles] per-sweep per-tile residual sums
  double* res_part = nullptr;   // [2][max_tiles]
  double* rows = nullptr;       // [max_grids][128] row sums
  double* rows_h = nullptr;     // pinned host copy
  char* end_h = nullptr;        // pinned host copies of res_part, counters and fails (ph_end)
  double* hist_dev = nullptr;   // [max_grids][UC_HIST] species RSD/SPNORM per sweep (k_decide_sp)
  // uc_keep_rates: every cell's RC of the last step that advanced it ([544][max_cells], zero
  // until then), and the list of cells with 3 <= ISKIP < 10 whose QDOT/QDSP SOLVEH forms from it
  bool keep_rc = false;
  double* rc_keep = nullptr;
  int32_t *stale_cells = nullptr, *stale_gid = nullptr;
  double* stale_slabs = nullptr;
  int64_t stale_npad = 0, stale_cap = 0;
  int64_t lay_builds = 0, stale_builds = -1;  // layouts built; the layout the stale list follows
  unsigned long long* counters = nullptr;  // [2][max_grids] sweeps executed, cells re-run
  int32_t* fails = nullptr;     // [3][max_grids] T(h): cells out of limit, no root, first without root
  hipEvent_t ev[8]{};
  hipEvent_t tev[6]{};  // start/end of the rates, sweeps and sources phases, read at ph_end
  std::vector<int32_t> last_nstar;  // per grid: previous step's species sweep count
  Layout lay;
  Tuning tune;
  SegGraph seg;
  hipStream_t gs = nullptr;  // the context's stream for captured sweep segments
  StepRun run;
  // coupled step (allocated at its first use)
  double* cpl_it = nullptr;     // [54][max_cells] iterates SA*, SST1, SST2 in field layout
  double* cpl_fpz = nullptr;    // [max_cells] FPZ in field layout
  double* cpl_rd = nullptr;     // [52][max_cells] RDHSP
  double* cpl_coef = nullptr;   // [max_pad/64][UC_CPL_SLOTS][64] PPP, P, Q, E, W per group
  double* cpl_flux = nullptr;   // [max_pad]
  double* cpl_snap = nullptr;   // [NSNAP_CPL][54][max_pad]
  double* cpl_zsnap = nullptr;  // [NSNAP_CPL][max_pad]
  int32_t* cpl_dfirst = nullptr;  // [ndiag+1] first list position of each diagonal
  int64_t cpl_dfirst_cap = 0;
  int32_t last_nh = 0;          // previous coupled step's enthalpy sweep count
  TrTables* tr = nullptr;       // transport-property tables (built at first use)
};

namespace {

// rcfac[s] = C_m / prod W of the slot's reactants (incl. the C3H8 collider of 459/460)
void fill_params(const uc_scales& sc, UcParams& p);

}  // namespace

#include "chem_host.inc"


// ---- SPGSOLV's coefficient assembly (spgsolv.f:98-1290), the input side of the coupled
// step.  FP64 with contraction off, so every value is the reference's operations in its
// order and equals the flang-built routine bit for bit.
struct AsmArgs {
  const double *rho, *y, *tk;
  const int32_t* iskip;
  const double *vtc, *tmu, *vdf, *stdf, *u, *v, *ygeo, *xxc, *xxs, *yyc, *yys;
  double *coef, *stc, *zc, *ekt;
  int64_t n;
  int32_t li, lj;
  double dt, sigsp, sigh, sym, rei, beta2, beta3, tstr, gasc, wmstr, hstr;
};

// EKT of every cell (spgsolv.f:102-176): cp of the mixture from POLSP (P->nasa), then
// EKT = REI*(BETA2*VTC/CP + TMU/SIGH)
__global__ void __launch_bounds__(256) k_asm_ekt(const UcParams* __restrict__ P, AsmArgs a) {
#pragma clang fp contract(off)
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n) return;
  const double tkd = a.tk[c] * a.tstr;
  const double tkd2 = tkd * tkd, tkd3 = tkd * tkd2, tkd4 = tkd * tkd3;
  const int ip = tkd > 1000.0 ? 7 : 0;  // /CB08/ TPOL1
  double tcp = 0.0;
#pragma unroll
  for (int k = 0; k < NSP; ++k) {  // unrolled: the 52 Y loads are in flight together
    const double* p = &P->nasa[k][ip];
    const double cpk = (p[0] + p[1] * tkd + p[2] * tkd2 + p[3] * tkd3 + p[4] * tkd4) / P->wm[k];
    tcp = tcp + cpk * a.y[(int64_t)k * a.n + c];
  }
  tcp = tcp / a.wmstr * a.gasc * a.tstr / a.hstr;
  a.ekt[c] = a.rei * (a.beta2 * a.vtc[c] / tcp + a.tmu[c] / a.sigh);
}

// E, W, P (J+1), Q (J-1) of the 51 species, the soot scalars (STE..STS) and the enthalpy
// (ZE..ZS) of one interior cell with ISKIP < 10 (spgsolv.f:1113-1290): hybrid upwinding of
// the face mass flux against the face conductance (spgsolv.f:233-349, formed here from the
// two cells beside the face) and the axisymmetric source on the J faces.  The soot W
// coefficient keeps the reference's comparison with the enthalpy's upwinded value
// (spgsolv.f:1208).  One thread per cell, J fastest across lanes.
__global__ void __launch_bounds__(256) k_asm_coef(AsmArgs a) {
#pragma clang fp contract(off)
  // XCD-aware order: the hardware deals blocks to the 8 XCDs round-robin (block b to XCD
  // b % 8), each with its own L2; renumbered so that each XCD takes a contiguous band of
  // rows, a cell's I-1 and I+1 neighbours are read through the same L2 as the cell itself
  const int64_t nb = gridDim.x, b = blockIdx.x, x = b % 8, q = nb / 8, r = nb % 8;
  const int64_t lb = x * q + (x < r ? x : r) + b / 8;
  const int64_t t = lb * blockDim.x + threadIdx.x;
  const int32_t nj = a.lj - 2;
  if (t >= (int64_t)(a.li - 2) * nj) return;
  const int32_t i = 1 + (int32_t)(t / nj), j = 1 + (int32_t)(t % nj);
  const int64_t c = (int64_t)i * a.lj + j, n = a.n;
  if (a.iskip[c] >= 10) return;
  const int64_t ce = c + a.lj, cw = c - a.lj, cn = c + 1, cs = c - 1;
  const double dyn = a.yys[j + 1], dys = a.yys[j];
  const double dtbrx = a.dt / a.xxc[i], dtbry = a.dt / a.yyc[j];
  const double rhoue = 0.5 * (a.rho[ce] + a.rho[c]) * a.u[ce];
  const double rhouw = 0.5 * (a.rho[c] + a.rho[cw]) * a.u[c];
  const double rhovn = 0.5 * (a.rho[cn] + a.rho[c]) * a.v[cn];
  const double rhovs = 0.5 * (a.rho[c] + a.rho[cs]) * a.v[c];
  const double rb = a.rei * a.beta3;
  const double emc = rb * a.rho[c], eme = rb * a.rho[ce], emw = rb * a.rho[cw], emn = rb * a.rho[cn],
               ems = rb * a.rho[cs];
  const double yg = a.ygeo[c];
  const double symz = a.sym * a.dt * 0.5 * a.ekt[c] / yg;
  const double sym1 = a.sym * a.dt * 0.5 * emc / yg;
  const double sym2 = a.sym * a.dt * 0.5 * a.rei * a.tmu[c] / a.sigsp / yg;
  const double turw_e = a.rei * (a.tmu[ce] + a.tmu[c]) / a.sigsp;
  const double turw_w = a.rei * (a.tmu[c] + a.tmu[cw]) / a.sigsp;
  const double turs_n = a.rei * (a.tmu[cn] + a.tmu[c]) / a.sigsp;
  const double turs_s = a.rei * (a.tmu[c] + a.tmu[cs]) / a.sigsp;
  const double xe = a.xxs[i + 1], xw = a.xxs[i];
  const double are = fabs(rhoue), arw = fabs(rhouw), arn = fabs(rhovn), ars = fabs(rhovs);
  // one scalar's four coefficients from its face conductances and symmetry term
  auto put = [&](double* o, int64_t stride, double de, double dw, double dn, double ds, double symc,
                 double wconv_ref, bool quirk) {
    double conv = are;
    if (de >= conv) conv = de;
    o[0] = dtbrx * 0.5 * (rhoue - conv);
    conv = arw;
    if (dw >= (quirk ? wconv_ref : conv)) conv = dw;
    o[stride] = dtbrx * 0.5 * (-rhouw - conv);
    conv = arn;
    if (dn >= conv) conv = dn;
    o[2 * stride] = dtbry * 0.5 * (rhovn - conv) - symc / dyn;
    conv = ars;
    if (ds >= conv) conv = ds;
    o[3 * stride] = dtbry * 0.5 * (-rhovs - conv) + symc / dys;
  };
  // enthalpy (its W upwinded value is the soot W quirk's reference)
  const double ekc = a.ekt[c];
  const double zdw = (ekc + a.ekt[cw]) / xw;
  double zconv_w = arw;
  if (zdw >= zconv_w) zconv_w = zdw;
  put(a.zc + c, n, (a.ekt[ce] + ekc) / xe, zdw, (a.ekt[cn] + ekc) / dyn, (ekc + a.ekt[cs]) / dys, symz, 0.0, false);
  // soot scalars (STDF) and the species (VDF); unrolled so several scalars' loads overlap
#pragma unroll 4
  for (int s = 51; s >= 0; --s) {
    const double* df = s == 51 ? a.stdf : a.vdf + (int64_t)s * n;
    const double dc = df[c];
    const double de = (eme * df[ce] + emc * dc + turw_e) / xe;
    const double dw = (emc * dc + emw * df[cw] + turw_w) / xw;
    const double dn = (emn * df[cn] + emc * dc + turs_n) / dyn;
    const double ds = (emc * dc + ems * df[cs] + turs_s) / dys;
    double* o = s == 51 ? a.stc + c : a.coef + (int64_t)(4 * s) * n + c;
    put(o, n, de, dw, dn, ds, sym1 * dc + sym2, zconv_w, s == 51);
  }
}

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

int uc_create_batch(const uc_scales* scales, int64_t max_cells, int32_t max_grids, uc_context** out) {
  if (!scales || !out || max_cells < 9 || max_grids < 1 || max_grids > (1 << 20))
    return fail(UC_E_ARG, "uc_create: bad arguments");
  // buffer instructions address each per-cell array with 32-bit byte offsets
  if (max_cells + (int64_t)WAVE * max_grids > UC_MAX_CELLS)
    return fail(UC_E_ARG, "uc_create: max_cells + 64*max_grids above " + std::to_string(UC_MAX_CELLS));
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(UC_E_NOGPU, "uc_create: no HIP device");
  auto ctx = std::make_unique<uc_context>();
  HIPCHK(hipGetDevice(&ctx->device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
    return fail(UC_E_NOGPU, std::string("uc_create: device is ") + prop.gcnArchName + ", built for gfx950");
  ctx->cus = prop.multiProcessorCount;
  ctx->sc = *scales;
  fill_params(ctx->sc, ctx->hp);
  ctx->max_cells = max_cells;
  ctx->max_grids = max_grids;
  ctx->max_pad = (max_cells + (int64_t)WAVE * max_grids + WAVE - 1) / WAVE * WAVE;
  ctx->nsnap = 8;
  const int64_t mp = ctx->max_pad, mt = mp / WAVE;
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
  HIPCHK(hipMalloc(&ctx->cells, sizeof(int32_t) * mp));
  HIPCHK(hipMalloc(&ctx->gid, sizeof(int32_t) * mp));
  HIPCHK(hipMalloc(&ctx->base_list, sizeof(int32_t) * mp));
  HIPCHK(hipMalloc(&ctx->base_list_gid, sizeof(int32_t) * mp));
  HIPCHK(hipMalloc(&ctx->base, sizeof(int32_t) * max_cells));
  HIPCHK(hipMalloc(&ctx->base_gid, sizeof(int32_t) * max_cells));
  HIPCHK(hipMalloc(&ctx->base_off, sizeof(int32_t) * (max_grids + 1)));
  HIPCHK(hipMalloc(&ctx->prev_nsw, sizeof(int32_t) * max_cells));
  for (int i = 0; i < 2; ++i) {
    HIPCHK(hipMalloc(&ctx->keys[i], sizeof(uint32_t) * std::max(max_cells, mt)));
    HIPCHK(hipMalloc(&ctx->vals[i], sizeof(int32_t) * std::max(max_cells, mt)));
  }
  HIPCHK(hipMalloc(&ctx->order, sizeof(int32_t) * mt));
  HIPCHK(hipMalloc(&ctx->queue, sizeof(int32_t) * 64));
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, ctx->sort_tmp_bytes, ctx->keys[0], ctx->keys[1], ctx->vals[0],
                                            ctx->vals[1], (int)std::max(max_cells, mt), 0, 32));
  {
    size_t sel = 0;
    HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, sel, rocprim::counting_iterator<int32_t>(0), ctx->done,
                                         ctx->base, ctx->nsel, (int)max_cells));
    ctx->sort_tmp_bytes = std::max(ctx->sort_tmp_bytes, sel);
  }
  HIPCHK(hipMalloc(&ctx->sort_tmp, std::max<size_t>(ctx->sort_tmp_bytes, 256)));
  HIPCHK(hipMalloc(&ctx->nsel, sizeof(int32_t) * 64));
  HIPCHK(hipHostMalloc(&ctx->nsel_h, sizeof(int32_t) * 64, hipHostMallocDefault));
  HIPCHK(hipMalloc(&ctx->done, max_cells));
  HIPCHK(hipMalloc(&ctx->thflag, (max_cells + 3) / 4 * 4));   // cleared as int32 words
  HIPCHK(hipMalloc(&ctx->gt_dev, sizeof(GridTab) * max_grids));
  HIPCHK(hipHostMalloc(&ctx->gt, sizeof(GridTab) * max_grids, hipHostMallocDefault));
  HIPCHK(hipMalloc(&ctx->grid_ptrs, sizeof(PackPtrs) * max_grids));
  HIPCHK(hipMalloc(&ctx->rsd_part, sizeof(double) * 128 * mt));
  HIPCHK(hipMalloc(&ctx->res_part, sizeof(double) * 2 * mt));
  HIPCHK(hipMalloc(&ctx->rows, sizeof(double) * 128 * max_grids));
  HIPCHK(hipHostMalloc(&ctx->rows_h, sizeof(double) * 128 * max_grids, hipHostMallocDefault));
  HIPCHK(hipHostMalloc(&ctx->end_h, sizeof(double) * 2 * (mt + 2 * max_grids) + sizeof(int32_t) * 3 * max_grids,
                       hipHostMallocDefault));
  HIPCHK(hipMalloc(&ctx->hist_dev, sizeof(double) * UC_HIST * max_grids));
  HIPCHK(hipMalloc(&ctx->counters, sizeof(unsigned long long) * 2 * max_grids));
  HIPCHK(hipMalloc(&ctx->fails, sizeof(int32_t) * 3 * max_grids));
  if (UC_SWEEP_LDS * 512 > 65536) {
    HIPCHK(hipFuncSetAttribute((const void*)k_sweep<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               UC_SWEEP_LDS * 512));
    HIPCHK(hipFuncSetAttribute((const void*)k_sweep<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               UC_SWEEP_LDS * 512));
    HIPCHK(hipFuncSetAttribute((const void*)k_sweep_cpl, hipFuncAttributeMaxDynamicSharedMemorySize,
                               UC_SWEEP_LDS * 512));
  }
  if (UC_PRD_LDS * 512 > 65536)
    HIPCHK(hipFuncSetAttribute((const void*)k_sources, hipFuncAttributeMaxDynamicSharedMemorySize, UC_PRD_LDS * 512));
  HIPCHK(hipFuncSetAttribute((const void*)k_sweep_mw<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)mw_lds_bytes(2, UC_MW2_LDS)));
  HIPCHK(hipFuncSetAttribute((const void*)k_sweep_mw<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)mw_lds_bytes(4, UC_MW4_LDS)));

  for (auto& e : ctx->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDefault));
  for (auto& e : ctx->tev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDefault));
  HIPCHK(hipStreamCreateWithFlags(&ctx->gs, hipStreamNonBlocking));  // default priority
  *out = ctx.release();
  return UC_OK;
}

int uc_create(const uc_scales* scales, int64_t max_cells, uc_context** out) {
  return uc_create_batch(scales, max_cells, 1, out);
}

int uc_destroy(uc_context* ctx) {
  if (!ctx) return UC_OK;
  for (void* p : {(void*)ctx->dp, (void*)ctx->slabs, (void*)ctx->st, (void*)ctx->snap, (void*)ctx->sst2p,
                  (void*)ctx->sa, (void*)ctx->sst, (void*)ctx->rhsz, (void*)ctx->nsw, (void*)ctx->cells,
                  (void*)ctx->gid, (void*)ctx->base_list, (void*)ctx->base_list_gid, (void*)ctx->base,
                  (void*)ctx->base_gid, (void*)ctx->base_off, (void*)ctx->prev_nsw, (void*)ctx->keys[0],
                  (void*)ctx->keys[1], (void*)ctx->vals[0], (void*)ctx->vals[1], (void*)ctx->order,
                  (void*)ctx->queue, ctx->sort_tmp, (void*)ctx->done, (void*)ctx->gt_dev, (void*)ctx->grid_ptrs,
                  (void*)ctx->rsd_part, (void*)ctx->res_part, (void*)ctx->rows, (void*)ctx->counters,
                  (void*)ctx->fails, (void*)ctx->thflag, (void*)ctx->pk.rho, (void*)ctx->pk.y, (void*)ctx->pk.tk, (void*)ctx->pk.ht, (void*)ctx->pk.stmf,
                  (void*)ctx->pk.stnd, (void*)ctx->pk.qdot, (void*)ctx->pk.qdsp, (void*)ctx->pk_iskip,
                  (void*)ctx->nsel, (void*)ctx->cpl_it, (void*)ctx->cpl_fpz, (void*)ctx->cpl_rd,
                  (void*)ctx->cpl_coef, (void*)ctx->cpl_flux, (void*)ctx->cpl_snap, (void*)ctx->cpl_zsnap,
                  (void*)ctx->cpl_dfirst, (void*)ctx->hist_dev, (void*)ctx->rc_keep, (void*)ctx->stale_cells,
                  (void*)ctx->stale_gid, (void*)ctx->stale_slabs})
    if (p) (void)hipFree(p);
  if (ctx->nsel_h) (void)hipHostFree(ctx->nsel_h);
  if (ctx->gs) (void)hipStreamSynchronize(ctx->gs);
  if (ctx->seg.exec) (void)hipGraphExecDestroy(ctx->seg.exec);
  if (ctx->seg.graph) (void)hipGraphDestroy(ctx->seg.graph);
  if (ctx->gt) (void)hipHostFree(ctx->gt);
  if (ctx->rows_h) (void)hipHostFree(ctx->rows_h);
  if (ctx->end_h) (void)hipHostFree(ctx->end_h);
  for (auto& e : ctx->ev) (void)hipEventDestroy(e);
  for (auto& e : ctx->tev) (void)hipEventDestroy(e);
  if (ctx->gs) (void)hipStreamDestroy(ctx->gs);
  delete ctx;
  return UC_OK;
}

int uc_step(uc_context* ctx, const uc_field* f, const uc_controls* ctl, void* stream, uc_stats* st) {
  return step_impl(ctx, f, ctl, (hipStream_t)stream, st);
}

int uc_assemble(uc_context* ctx, const uc_field* f, const uc_controls* ctl, const uc_transport* tr, double* coef,
                double* stc, double* zc, double* ekt, void* stream) {
  if (!ctx || !f || !ctl || !tr) return fail(UC_E_ARG, "uc_assemble: null argument");
  if (f->li < 3 || f->lj < 3) return fail(UC_E_ARG, "uc_assemble: grid must be at least 3 x 3");
  const int64_t n = (int64_t)f->li * f->lj;
  if (n > ctx->max_cells) return fail(UC_E_ARG, "uc_assemble: the grid does not fit the context");
  if (!f->rho || !f->y || !f->tk || !f->iskip)
    return fail(UC_E_ARG, "uc_assemble: null field array (rho, y, tk and iskip are read)");
  if (!tr->vtc || !tr->tmu || !tr->vdf || !tr->stdf || !tr->u || !tr->v || !tr->ygeo || !tr->xxc || !tr->xxs ||
      !tr->yyc || !tr->yys)
    return fail(UC_E_ARG, "uc_assemble: null transport array");
  if (!coef || !stc || !zc || !ekt) return fail(UC_E_ARG, "uc_assemble: null output array");
  if (!(ctl->dt > 0.0) || !(tr->sigsp > 0.0) || !(tr->sigh > 0.0))
    return fail(UC_E_ARG, "uc_assemble: dt, sigsp and sigh must be positive");
  HIPCHK(hipSetDevice(ctx->device));
  AsmArgs a{};
  a.rho = f->rho; a.y = f->y; a.tk = f->tk; a.iskip = f->iskip;
  a.vtc = tr->vtc; a.tmu = tr->tmu; a.vdf = tr->vdf; a.stdf = tr->stdf; a.u = tr->u; a.v = tr->v;
  a.ygeo = tr->ygeo; a.xxc = tr->xxc; a.xxs = tr->xxs; a.yyc = tr->yyc; a.yys = tr->yys;
  a.coef = coef; a.stc = stc; a.zc = zc; a.ekt = ekt;
  a.n = n; a.li = f->li; a.lj = f->lj;
  a.dt = ctl->dt; a.sigsp = tr->sigsp; a.sigh = tr->sigh; a.sym = (double)tr->isym;
  const uc_scales& S = ctx->sc;
  a.rei = S.rei; a.beta2 = S.beta2; a.beta3 = S.beta3; a.tstr = S.tstr; a.gasc = S.gasc; a.wmstr = S.wmstr;
  a.hstr = S.hstr;
  hipStream_t s = (hipStream_t)stream;
  k_asm_ekt<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(ctx->dp, a);
  const int64_t ni = (int64_t)(f->li - 2) * (f->lj - 2);
  k_asm_coef<<<(unsigned)((ni + 255) / 256), 256, 0, s>>>(a);
  HIPCHK(hipGetLastError());
  return UC_OK;
}

int uc_transport_props(uc_context* ctx, const uc_field* f, double* vmu, double* vtc, double* vdf, double* stdf,
                       void* stream) {
  if (!ctx || !f) return fail(UC_E_ARG, "uc_transport_props: null argument");
  if (f->li < 3 || f->lj < 3) return fail(UC_E_ARG, "uc_transport_props: grid must be at least 3 x 3");
  const int64_t n = (int64_t)f->li * f->lj;
  if (n > ctx->max_cells) return fail(UC_E_ARG, "uc_transport_props: the grid does not fit the context");
  if (!f->y || !f->tk) return fail(UC_E_ARG, "uc_transport_props: null field array (y and tk are read)");
  if (!vmu || !vtc || !vdf || !stdf) return fail(UC_E_ARG, "uc_transport_props: null output array");
  HIPCHK(hipSetDevice(ctx->device));
  if (!ctx->tr) {
    std::unique_ptr<TrTables> h(new TrTables);
    fill_transport_tables(ctx->sc, *h);
    HIPCHK(hipMalloc(&ctx->tr, sizeof(TrTables)));
    HIPCHK(hipMemcpy(ctx->tr, h.get(), sizeof(TrTables), hipMemcpyHostToDevice));
  }
  k_transport<<<blocks_for(n, 64), 64, 0, (hipStream_t)stream>>>(ctx->tr, f->y, f->tk, n, vmu, vtc, vdf, stdf);
  HIPCHK(hipGetLastError());
  return UC_OK;
}

}  // extern "C"

// ---- uc_transport_props without allocation -------------------------------------------
// The transport kernel's only scratch is its table block (pair tables, fits, constants).
// For callers that cannot allocate device memory, it lives in this fixed device array,
// one per GPU for the life of the process; its address is looked up once per device.
namespace {
constexpr size_t TR_SCRATCH_DOUBLES = (sizeof(TrTables) + 7) / 8;   // about 152 KB
constexpr int TR_MAX_DEVICES = 64;
struct TrScratchState {
  double* ptr = nullptr;   // device address of this GPU's copy of `scratch`
  bool loaded = false;     // tables for `key` are in it
  uc_scales key{};
};
TrScratchState tr_state[TR_MAX_DEVICES];
std::mutex tr_mutex;
TrTables tr_host;          // host staging of the tables (static storage)
}  // namespace

static __device__ __align__(256) double scratch[TR_SCRATCH_DOUBLES];

extern "C" {

int uc_transport_props_noalloc(const uc_scales* scales, const uc_field* f, double* vmu, double* vtc, double* vdf,
                               double* stdf, void* stream) {
  if (!scales || !f) return fail(UC_E_ARG, "uc_transport_props_noalloc: null argument");
  if (f->li < 3 || f->lj < 3) return fail(UC_E_ARG, "uc_transport_props_noalloc: grid must be at least 3 x 3");
  const int64_t n = (int64_t)f->li * f->lj;
  if (n > UC_MAX_CELLS) return fail(UC_E_ARG, "uc_transport_props_noalloc: more than 4,000,000 cells");
  if (!f->y || !f->tk) return fail(UC_E_ARG, "uc_transport_props_noalloc: null field array (y and tk are read)");
  if (!vmu || !vtc || !vdf || !stdf) return fail(UC_E_ARG, "uc_transport_props_noalloc: null output array");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= TR_MAX_DEVICES) return fail(UC_E_UNSUPPORTED, "uc_transport_props_noalloc: device index");
  hipStream_t s = (hipStream_t)stream;
  TrScratchState& st = tr_state[dev];
  {
    std::lock_guard<std::mutex> lock(tr_mutex);
    if (!st.ptr) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(UC_E_HIP, std::string("uc_transport_props_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
      st.ptr = (double*)p;
    }
    if (!st.loaded || std::memcmp(&st.key, scales, sizeof(uc_scales)) != 0) {
      fill_transport_tables(*scales, tr_host);
      HIPCHK(hipMemcpyAsync(st.ptr, &tr_host, sizeof(TrTables), hipMemcpyHostToDevice, s));
      HIPCHK(hipStreamSynchronize(s));   // tr_host is reused by the next upload
      st.key = *scales;
      st.loaded = true;
    }
  }
  k_transport<<<blocks_for(n, 64), 64, 0, s>>>((const TrTables*)st.ptr, f->y, f->tk, n, vmu, vtc, vdf, stdf);
  HIPCHK(hipGetLastError());
  return UC_OK;
}

int uc_step_coupled(uc_context* ctx, const uc_field* f, const uc_controls* ctl, const uc_coupling* cp, void* stream,
                    uc_stats* st) {
  return step_cpl_impl(ctx, f, ctl, cp, (hipStream_t)stream, st);
}

int uc_step_batch(uc_context* ctx, const uc_field* fields, const uc_controls* ctls, int32_t ngrids, void* stream,
                  uc_stats* stats) {
  return step_batch_impl(ctx, fields, ctls, ngrids, (hipStream_t)stream, stats);
}

int uc_step_host(uc_context* ctx, const uc_field* hf, const uc_controls* ctl, uc_stats* st) {
  return step_host_impl(ctx, hf, ctl, st);
}

int uc_autotune(uc_context* ctx, const uc_field* fields, const uc_controls* ctls, int32_t ngrids, void* stream,
                uc_tuning* chosen) {
  return autotune_impl(ctx, fields, ctls, ngrids, (hipStream_t)stream, chosen);
}

int uc_set_tuning(uc_context* ctx, const uc_tuning* t) {
  if (!ctx || !t) return fail(UC_E_ARG, "uc_set_tuning: null argument");
  if (t->bins < 1 || t->bins > 32 || (t->block != 64 && t->block != 128 && t->block != 256 && t->block != 512))
    return fail(UC_E_ARG, "uc_set_tuning: bins must be 1..32 and block 64, 128, 256 or 512");
  if (t->lanes != 1 && t->lanes != 2 && t->lanes != 4)
    return fail(UC_E_ARG, "uc_set_tuning: lanes must be 1, 2 or 4");
  if (t->lanes > 1 && t->persistent)
    return fail(UC_E_ARG, "uc_set_tuning: the split sweep (lanes > 1) runs one tile per workgroup, not persistent");
  if (t->rate_parts != 0 && t->rate_parts != 1 && t->rate_parts != 2 && t->rate_parts != 4 && t->rate_parts != 8)
    return fail(UC_E_ARG, "uc_set_tuning: rate_parts must be 0 (by cell count), 1, 2, 4 or 8");
  ctx->tune = Tuning{t->bins, t->lpt ? 1 : 0, t->persistent ? 1 : 0, t->block, t->graph ? 1 : 0, t->lanes,
                     t->rate_parts};
  return UC_OK;
}

int uc_cells_changed(uc_context* ctx) {
  if (!ctx) return fail(UC_E_ARG, "uc_cells_changed: null context");
  ctx->lay.ngrids = 0;
  return UC_OK;
}

int uc_keep_rates(uc_context* ctx, int32_t on) {
  if (!ctx) return fail(UC_E_ARG, "uc_keep_rates: null context");
  if (on && !ctx->rc_keep) {
    HIPCHK(hipSetDevice(ctx->device));
    const size_t bytes = sizeof(double) * NRX * (size_t)ctx->max_cells;
    HIPCHK(hipMalloc(&ctx->rc_keep, bytes));
    HIPCHK(hipMemset(ctx->rc_keep, 0, bytes));   // /DUMR/ RC before any CKRATES: zero
    HIPCHK(hipDeviceSynchronize());
  }
  ctx->keep_rc = on != 0;
  ctx->stale_builds = -1;
  return UC_OK;
}

int uc_set_clip(uc_context* ctx, int32_t on) {
  if (!ctx) return fail(UC_E_ARG, "uc_set_clip: null context");
  ctx->clip = on != 0;
  return UC_OK;
}

int uc_clip(const uc_field* f, void* stream) {
  if (!f || !f->y || f->li < 3 || f->lj < 3) return fail(UC_E_ARG, "uc_clip: bad field");
  const int64_t n = (int64_t)f->li * f->lj;
  k_clip<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(f->y, n);
  HIPCHK(hipGetLastError());
  return UC_OK;
}

int uc_sweep_residuals(const uc_context* ctx, int32_t grid, double* out, int32_t cap) {
  if (!ctx || cap < 0 || (cap > 0 && !out)) return fail(UC_E_ARG, "uc_sweep_residuals: bad argument");
  if (grid < 0 || grid >= (int32_t)ctx->run.hist.size())
    return fail(UC_E_ARG, "uc_sweep_residuals: no step of grid " + std::to_string(grid));
  if (ctx->run.hist_dev) {
    const int32_t n = ctx->gt[grid].hcount;
    if (cap > 0 && n > 0)
      HIPCHK(hipMemcpy(out, ctx->hist_dev + (int64_t)grid * UC_HIST, sizeof(double) * std::min(cap, n),
                       hipMemcpyDeviceToHost));
    return n;
  }
  const std::vector<double>& h = ctx->run.hist[grid];
  for (int32_t i = 0; i < cap && i < (int32_t)h.size(); ++i) out[i] = h[i];
  return (int32_t)h.size();
}

int uc_th_flags(const uc_context* ctx, int32_t grid, uint8_t* flags, void* stream) {
  if (!ctx || !flags) return fail(UC_E_ARG, "uc_th_flags: null argument");
  if (grid < 0 || grid >= ctx->run.ngrids) return fail(UC_E_ARG, "uc_th_flags: no such grid in the last step");
  const GridTab& G = ctx->gt[grid];
  HIPCHK(hipMemcpyAsync(flags, ctx->thflag + G.off, (size_t)G.li * G.lj, hipMemcpyDeviceToDevice,
                        (hipStream_t)stream));
  return UC_OK;
}

int uc_get_tuning(const uc_context* ctx, uc_tuning* t) {
  if (!ctx || !t) return fail(UC_E_ARG, "uc_get_tuning: null argument");
  *t = uc_tuning{ctx->tune.bins, ctx->tune.lpt, ctx->tune.persistent, ctx->tune.block, ctx->tune.graph,
                 ctx->tune.lanes, ctx->tune.rate_parts};
  return UC_OK;
}

}  // extern "C"
