// Chemistry side of the surrogate (MI355X / gfx950): training pairs from FP64 steps,
// network features, and the per-cell gate that applies accepted predictions.  No
// counterpart in the reference.
//
// Features (54, then padded to 64): T in K, ln(dt / s), and the Box-Cox transform
// (lambda = 0.1) of the 52 mass fractions, each standardised with the training set's
// mean and standard deviation.  Targets (53): the step's dT (K) and, per species, the
// change of the Box-Cox transform, BC(Y_k after) - BC(Y_k before) (which resolves
// minor species relative to their size and keeps every predicted
// Y_k = BC^-1(BC(Y_k) + change) >= 0); each target d is compressed as asinh(d / s)
// with s its median |d| over the training set (linear below s, logarithmic above, so
// the small changes of a relaxed state and the large ones of an ignition are resolved
// alike) and then standardised.
#include "mlp_kernels.hip"

namespace {

constexpr int NSPS = 52;
constexpr int NF = 54, NT = 53, NPAD = 64;
// k_gate: threads per block (a multiple of 64; four lanes per cell)
#ifndef UC_GATE_BLOCK
#define UC_GATE_BLOCK 256
#endif
#ifndef UC_GATE_WAVES
#define UC_GATE_WAVES 2
#endif
#define GATE_WAVES UC_GATE_WAVES  // waves per SIMD the gate's registers must allow (3 spills)
constexpr int GATE_BLOCK = UC_GATE_BLOCK;
// k_network: waves per block (32 cells each).  The block holds half of a hidden layer's
// weights in LDS at a time (64 KB), so two blocks share a CU and one stages while the
// other computes: two waves per SIMD.
constexpr int NET_WAVES = 4;
constexpr int NET_LDS = 128 * 256 * 2;   // bytes: half a hidden layer, 128 x 256 bf16

__device__ __forceinline__ double boxcox(double y) { return (pow(y > 0.0 ? y : 0.0, 0.1) - 1.0) / 0.1; }
// inverse on the range of boxcox: (1 + 0.1 b)^10, 0 below b = -10
__device__ __forceinline__ double boxcox_inv(double b) {
  double u = 1.0 + 0.1 * b;
  u = u > 0.0 ? u : 0.0;
  const double u2 = u * u, u4 = u2 * u2, u8 = u4 * u4;
  return u8 * u2;
}

// raw features (T, ln dt, Y) and raw targets (dT, dY) of the listed cells of one field
// step: rows row0 .. row0+ncells-1
__global__ void k_collect(const double* __restrict__ y0, const double* __restrict__ tk0,
                          const double* __restrict__ y1, const double* __restrict__ tk1, int64_t n,
                          const int32_t* __restrict__ cells, int32_t ncells, double tstr, double log_dt,
                          float* __restrict__ xraw, float* __restrict__ traw, int64_t row0) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncells) return;
  const int64_t c = cells[q], r = row0 + q;
  float* x = xraw + r * NF;
  float* t = traw + r * NT;
  x[0] = (float)(tk0[c] * tstr);
  x[1] = (float)log_dt;
  t[0] = (float)((tk1[c] - tk0[c]) * tstr);
  for (int k = 0; k < NSPS; ++k) {
    x[2 + k] = (float)y0[k * n + c];
    t[1 + k] = (float)(boxcox(y1[k * n + c]) - boxcox(y0[k * n + c]));
  }
}

// transformed value of raw feature j (Box-Cox for the mass fractions)
__device__ __forceinline__ double feat(const float* x, int j) { return j < 2 ? (double)x[j] : boxcox((double)x[j]); }

// per-column sums, sums of squares, min and max of the transformed features (kind 0)
// or of the compressed targets asinh(d / tscale[j]) (kind 1), over rows; one block per
// (column, chunk) -> partials
__global__ void k_colstats(const float* __restrict__ raw, int64_t rows, int ncol, int kind,
                           const float* __restrict__ tscale, int64_t chunk, double* __restrict__ part) {
  __shared__ double s1[256], s2[256], mn[256], mx[256];
  const int j = blockIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * chunk, r1 = min(rows, r0 + chunk);
  double a = 0.0, b = 0.0, lo = 1e300, hi = -1e300;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    const double v = kind == 0 ? feat(raw + r * ncol, j) : asinh((double)raw[r * ncol + j] / tscale[j]);
    a += v;
    b += v * v;
    lo = fmin(lo, v);
    hi = fmax(hi, v);
  }
  s1[threadIdx.x] = a; s2[threadIdx.x] = b; mn[threadIdx.x] = lo; mx[threadIdx.x] = hi;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
      mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + o]);
      mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* p = part + ((int64_t)blockIdx.y * ncol + j) * 4;
    p[0] = s1[0]; p[1] = s2[0]; p[2] = mn[0]; p[3] = mx[0];
  }
}

struct Norm {  // standardisation: (v - mean) / sd; targets first compressed as asinh(d / tscale)
  float xmean[NF], xsd[NF], tmean[NT], tsd[NT], tscale[NT];
};

// a training batch: rows idx[0..B) of the raw set -> bf16 X [B][64], X^T [64][B] and
// fp32 standardised targets [B][64] (padding zero)
__global__ void k_batch(const float* __restrict__ xraw, const float* __restrict__ traw, const int64_t* __restrict__ idx,
                        int B, const Norm* __restrict__ nm, bf16_t* __restrict__ X, bf16_t* __restrict__ XT,
                        float* __restrict__ T) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * NPAD) return;
  const int b = (int)(i / NPAD), j = (int)(i % NPAD);
  const int64_t r = idx[b];
  float xv = 0.f, tv = 0.f;
  if (j < NF) xv = (float)((feat(xraw + r * NF, j) - nm->xmean[j]) / nm->xsd[j]);
  if (j < NT) tv = (float)((asinh((double)traw[r * NT + j] / nm->tscale[j]) - nm->tmean[j]) / nm->tsd[j]);
  X[(int64_t)b * NPAD + j] = to_bf16(xv);
  XT[(int64_t)j * B + b] = to_bf16(xv);
  T[(int64_t)b * NPAD + j] = tv;
}

// ---- inference on a field
struct Thermo {  // what the gate needs to turn (T, Y) into the field's variables
  double tstr, hstr, gasc, wmstr, pback, beta1;
  double rwm[NSPS], hfo[NSPS];
  double wm[NSPS];         // /CB02/ WM, for the density in the reference's order
  double nasah[NSPS][14];  // NASA-7 a_i / i (i = 1..5) and a_6, per range: sensible-enthalpy form
  double mel[3][NSPS];   // element (C, H, O) mass per unit mass of each species
  double xlo[NF], xhi[NF];  // training range of the transformed features (with a margin)
  double tlo[NT], thi[NT];  // training range of the targets (with a margin)
  double tmin, tmax;        // the range SOLVEH's Newton clamps T to (solveh.f:38,1229-1230), K
  double stmf_max, stnd_max;  // training range of the soot scalars
};

// features of the listed cells -> bf16 X [Bpad][64] (rows >= ncells zero)
__global__ void k_features(const double* __restrict__ y, const double* __restrict__ tk, int64_t n,
                           const int32_t* __restrict__ cells, int32_t ncells, int32_t bpad, double tstr,
                           double log_dt, const Norm* __restrict__ nm, bf16_t* __restrict__ X) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)bpad * NPAD) return;
  const int q = (int)(i / NPAD), j = (int)(i % NPAD);
  float v = 0.f;
  if (q < ncells && j < NF) {
    const int64_t c = cells[q];
    const double raw = j == 0 ? tk[c] * tstr : (j == 1 ? log_dt : boxcox(y[(j - 2) * n + c]));
    v = (float)((raw - nm->xmean[j]) / nm->xsd[j]);
  }
  X[i] = to_bf16(v);
}

// The gate of one cell, from its 64 standardised network outputs o (53 used).
// Un-standardise dT and the Box-Cox changes; Y1_k = BC^-1(BC(Y0_k) + change); over
// species 1..49 correct dY = Y1 - Y0 by the change of least weighted norm (weights
// w_k = max(Y0_k, Y1_k), so the abundant species carry it) that conserves C, H and O
// (AR, HE and N2 hold none); clip negative mass fractions; close N2 as SOLVESP does
// (solvesp.f:2839-2914: ftot = sum Y_1..51 + STMF, Y_N2 = 1 - ftot, and when
// ftot > 1 every Y and STMF divided by ftot and Y_N2 = 0).  Accept when the features
// lie in the training hull (`inside`, the caller's), STMF/STND in theirs, every output
// in the range of the training targets and the new T in [tmin, tmax] (a NaN fails
// these), the C/H/O change of the final Y is within tol_el, |sum Y + STMF - 1|
// within tol_sum, and the chemistry is slow: |dT| <= tol_dt, every |dY_k| <= tol_dy and
// every predicted |change of BC(Y_k)| <= tol_bc, a relative measure that also flags
// the radical build-up before an ignition (a cell whose step changes fast goes to the
// FP64 step, where the network's relative error would be a large absolute one); then
// write Y, TK and HT (the sensible enthalpy of the new T and Y, whose inverse is
// solveh.f:1170-1230).  Otherwise the cell is left as it was.  The soot scalars are
// not predicted: an accepted cell keeps its STMF and STND.  Returns 0 accepted, 1 out of range, 2 residual, 3 sum, 4 active;
// rr: C/H/O residual of the raw prediction.
// Output j of the network: u = v tsd + tmean (un-standardised, still compressed), in
// float, the network's own precision; the change it stands for is d = tscale sinh(u).
// ok: u lies in the training targets' range (a NaN fails); slow: for a species
// (j >= 1), |d| <= tol_bc.
__device__ __forceinline__ float out_change(float v, int j, const Norm* __restrict__ nm,
                                            const Thermo* __restrict__ th, double tol_bc, bool& ok, bool& slow) {
  const float u = v * nm->tsd[j] + nm->tmean[j];
  ok = (double)u >= th->tlo[j] && (double)u <= th->thi[j];
  const float d = nm->tscale[j] * sinhf(u);
  slow = j == 0 || fabs((double)d) <= tol_bc;
  return d;
}

// Y0^0.1, the base of the Box-Cox transform (Y < 0 and NaN taken as 0), as
// exp(0.1 ln Y): within a few ulp of pow, at a fraction of its cost (0 at Y = 0)
__device__ __forceinline__ double bc_base(double y) { return exp(0.1 * log(y > 0.0 ? y : 0.0)); }

// Y after a Box-Cox change d, from P = Y0^0.1: BC^-1(BC(Y0) + d) = (P + 0.1 d)^10 (0 when
// the base falls below 0)
__device__ __forceinline__ double y_after(double P, float d) {
  double u = P + 0.1 * (double)d;
  u = u > 0.0 ? u : 0.0;
  const double u2 = u * u, u4 = u2 * u2, u8 = u4 * u4;
  return u8 * u2;
}

// The gate runs four lanes per cell: lane l of a wave serves cell slot l / 4 and part
// l % 4 of its species, k = part + 4 j (j = 0..12), whose Y0, predicted Y and tables it
// holds in registers, so Y is read from HBM once and written once.  Sums over species
// are each part's own in index order, then combined across the cell's quad of lanes as
// (p0 + p1) + (p2 + p3) by DPP moves, which every lane of the cell evaluates to the same
// bits.
constexpr int GPART = 4, GSP = NSPS / GPART;  // parts per cell, species per part
constexpr int GATE_CELLS = GATE_BLOCK / 64 * 16;  // cells per block and round


// per-species constants of the gate, staged in LDS: species k, network output 1 + k
struct GateSp {
  double mel[3], rwm, hfo, nasah[14], xlo, xhi, tlo, thi;
  float tsd, tmean, tscale;
};

// The value of another lane of the same quad (DPP quad_perm CTRL: 0xB1 lane ^ 1, 0x4E
// lane ^ 2, 0xFF lane 3 of the quad); a VALU move, no LDS round trip.
template <int CTRL>
__device__ __forceinline__ int quad_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double quad_d(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)quad_i<CTRL>((int)(uint32_t)b), hi = (uint32_t)quad_i<CTRL>((int)(uint32_t)(b >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double part_sum(double v) {
  v += quad_d<0xB1>(v);
  return v + quad_d<0x4E>(v);
}
__device__ __forceinline__ double part_max(double v) {
  v = fmax(v, quad_d<0xB1>(v));
  return fmax(v, quad_d<0x4E>(v));
}
__device__ __forceinline__ bool part_all(bool b) {
  int v = b ? 1 : 0;
  v &= quad_i<0xB1>(v);
  v &= quad_i<0x4E>(v);
  return v != 0;
}

// Statistics of the gate: stats [0] accepted, [1] out of range, [2] residual, [3]
// sum, [4] active, [5..7] max |raw C/H/O residual| over the accepted cells (bits of a
// double; non-negative doubles order as their bits).  A block sums its lanes' counts in LDS and adds to
// stats once per counter, since same-address atomics from every wave serialise in L2.
struct GateCount {
  unsigned int n[5] = {0, 0, 0, 0, 0};
  double rmax[3] = {0.0, 0.0, 0.0};
  __device__ void add(int why, const double* rr) {
#pragma unroll
    for (int k = 0; k < 5; ++k) n[k] += why == k;
    if (why == 0)   // the raw residual of the predictions the gate accepts
#pragma unroll
      for (int e = 0; e < 3; ++e) rmax[e] = fmax(rmax[e], fabs(rr[e]));
  }
};
__device__ __forceinline__ void gate_flush(const GateCount& gc, unsigned long long* s_st,
                                           unsigned long long* __restrict__ stats) {
  const int l = threadIdx.x % 64;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    unsigned int v = gc.n[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (l == 0 && v) atomicAdd(&s_st[k], (unsigned long long)v);
  }
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    double v = gc.rmax[e];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if (l == 0 && v > 0.0) atomicMax(&s_st[5 + e], (unsigned long long)__double_as_longlong(v));
  }
  __syncthreads();
  if (threadIdx.x < 5 && s_st[threadIdx.x]) atomicAdd(&stats[threadIdx.x], s_st[threadIdx.x]);
  if (threadIdx.x >= 5 && threadIdx.x < 8 && s_st[threadIdx.x]) atomicMax(&stats[threadIdx.x], s_st[threadIdx.x]);
}

// The gate as a kernel, on the network's standardised outputs: output j of list
// position q at out[q * sq + j * sj] (row-major [q][64] from the separate-kernel
// network, species-major [64][ldo] from k_network).  pbuf: Y0^0.1 as [52][ldo]
// (k_network writes it; null: computed here, by the same function).  reject[c] = 1
// where rejected, 0 where accepted (may be null); mask[c] = 10 where accepted, iskip[c]
// where rejected (may be null): the ISKIP with which the FP64 step then advances only
// the rejected cells.  Every lane runs every shuffle, so the arithmetic of a cell is
// done by all four of its lanes whether or not the cell is live or accepted; only the
// stores are guarded.
template <bool PBUF>
__global__ void __launch_bounds__(GATE_BLOCK) __attribute__((amdgpu_waves_per_eu(GATE_WAVES))) k_gate(const float* __restrict__ out, int64_t sq, int64_t sj,
                                              const double* __restrict__ pbuf, int64_t ldo,
                                              const Norm* __restrict__ nm, const Thermo* __restrict__ th,
                                              double* __restrict__ y, double* __restrict__ tk,
                                              double* __restrict__ ht, const double* __restrict__ stmf,
                                              const double* __restrict__ stnd, int64_t n,
                                              const int32_t* __restrict__ cells, int32_t ncells, double log_dt,
                                              double tol_el, double tol_sum, double tol_dt, double tol_dy,
                                              double tol_bc, int32_t* __restrict__ reject,
                                              const int32_t* __restrict__ iskip, int32_t* __restrict__ mask,
                                              unsigned long long* __restrict__ stats) {
  __shared__ GateSp sp[NSPS];
  for (int k = threadIdx.x; k < NSPS; k += blockDim.x) {
    GateSp& e = sp[k];
    for (int i = 0; i < 3; ++i) e.mel[i] = th->mel[i][k];
    e.rwm = th->rwm[k];
    e.hfo = th->hfo[k];
    for (int i = 0; i < 14; ++i) e.nasah[i] = th->nasah[k][i];
    e.xlo = th->xlo[k + 2];
    e.xhi = th->xhi[k + 2];
    e.tlo = th->tlo[k + 1];
    e.thi = th->thi[k + 1];
    e.tsd = nm->tsd[k + 1];
    e.tmean = nm->tmean[k + 1];
    e.tscale = nm->tscale[k + 1];
  }
  __shared__ unsigned long long s_st[8];
  if (threadIdx.x < 8) s_st[threadIdx.x] = 0;
  __syncthreads();
  const int l = threadIdx.x % 64, part = l % 4;
  GateCount gc;
  const int64_t q0 = (int64_t)blockIdx.x * GATE_CELLS + (threadIdx.x / 64) * 16 + l / 4;
  const bool live = q0 < ncells;
  // a lane past the list repeats the last cell's arithmetic; its stores are skipped
  const int64_t q = live ? q0 : ncells - 1, c = cells[q];
  double y0[GSP], y1[GSP];
  // features in the hull, outputs in range, Box-Cox changes slow (this part's species)
  bool ok = true, slow = true;
#pragma unroll
  for (int j = 0; j < GSP; ++j) {
    const int k = part + GPART * j;
    const GateSp& e = sp[k];
    y0[j] = y[(int64_t)k * n + c];
    // Y0^0.1: from the network's features, or computed here by the same function
    const double P = PBUF ? pbuf[(int64_t)k * ldo + q] : bc_base(y0[j]);
    const double v = (P - 1.0) / 0.1;
    ok = ok && v >= e.xlo && v <= e.xhi;
    // output 1 + k: u = v tsd + tmean, in range; the change d = tscale sinh(u)
    const float u = out[q * sq + (1 + k) * sj] * e.tsd + e.tmean;
    ok = ok && (double)u >= e.tlo && (double)u <= e.thi;
    const float d = e.tscale * sinhf(u);
    slow = slow && fabs((double)d) <= tol_bc;
    y1[j] = k < 51 ? y_after(P, d) : 0.0;
  }
  ok = part_all(ok);
  slow = part_all(slow);
  // dT (output 0) and the cell's own hull tests
  const float u0 = out[q * sq] * nm->tsd[0] + nm->tmean[0];
  const double dT = (double)(nm->tscale[0] * sinhf(u0));
  const double T0 = tk[c] * th->tstr;
  ok = ok && (double)u0 >= th->tlo[0] && (double)u0 <= th->thi[0] && T0 >= th->xlo[0] && T0 <= th->xhi[0] &&
       log_dt >= th->xlo[1] && log_dt <= th->xhi[1];
  // raw residual r = A dY and the weighted normal matrix M = A W A^T (symmetric 3x3),
  // over species 1..49 (AR, HE and N2 hold no C, H or O)
  double rr[3] = {0.0, 0.0, 0.0}, m[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < GSP; ++j) {
    const int k = part + GPART * j;
    if (k < 49) {
      const GateSp& e = sp[k];
      const double w = fmax(y0[j], y1[j]), a0 = e.mel[0], a1 = e.mel[1], a2 = e.mel[2];
      rr[0] += a0 * (y1[j] - y0[j]);
      rr[1] += a1 * (y1[j] - y0[j]);
      rr[2] += a2 * (y1[j] - y0[j]);
      m[0] += w * a0 * a0; m[1] += w * a0 * a1; m[2] += w * a0 * a2;
      m[3] += w * a1 * a1; m[4] += w * a1 * a2; m[5] += w * a2 * a2;
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) rr[i] = part_sum(rr[i]);
#pragma unroll
  for (int i = 0; i < 6; ++i) m[i] = part_sum(m[i]);
  // lambda = M^-1 r (Cramer); the correction of species k is -w_k A_k^T lambda
  const double c00 = m[3] * m[5] - m[4] * m[4], c01 = m[2] * m[4] - m[1] * m[5], c02 = m[1] * m[4] - m[2] * m[3];
  const double c11 = m[0] * m[5] - m[2] * m[2], c12 = m[1] * m[2] - m[0] * m[4], c22 = m[0] * m[3] - m[1] * m[1];
  const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  double lam[3] = {0.0, 0.0, 0.0};
  if (det > 0.0) {
    lam[0] = (c00 * rr[0] + c01 * rr[1] + c02 * rr[2]) / det;
    lam[1] = (c01 * rr[0] + c11 * rr[1] + c12 * rr[2]) / det;
    lam[2] = (c02 * rr[0] + c12 * rr[1] + c22 * rr[2]) / det;
  }
  // Y1 of species k < 51: predicted, corrected, clipped; its sum, C/H/O change and
  // largest change
  double tot = 0.0, dm[3] = {0.0, 0.0, 0.0}, m1[3] = {0.0, 0.0, 0.0}, dmax = 0.0;
#pragma unroll
  for (int j = 0; j < GSP; ++j) {
    const int k = part + GPART * j;
    if (k < 51) {
      const GateSp& e = sp[k];
      double v = y1[j];
      if (k < 49) v -= fmax(y0[j], v) * (e.mel[0] * lam[0] + e.mel[1] * lam[1] + e.mel[2] * lam[2]);
      v = v < 0.0 ? 0.0 : v;
      y1[j] = v;
      tot += v;
      dmax = fmax(dmax, fabs(v - y0[j]));
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        dm[i] += e.mel[i] * (v - y0[j]);
        m1[i] += e.mel[i] * v;
      }
    }
  }
  tot = part_sum(tot);
  dmax = part_max(dmax);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    dm[i] = part_sum(dm[i]);
    m1[i] = part_sum(m1[i]);
  }
  // N2 closure as SOLVESP does (solvesp.f:2839-2914); N2 is species 51 = part 3, j = 12
  const double stm = stmf[c], stn = stnd[c];
  const double ftot = tot + stm;
  const bool over = ftot > 1.0;
  const double yn = over ? 0.0 : 1.0 - ftot;
  const double y0n = quad_d<0xFF>(y0[GSP - 1]);   // part 3 of this cell
  dmax = fmax(dmax, fabs(yn - y0n));
  double res = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    // y1/ftot - y0 = (y1 - y0) - y1 (1 - 1/ftot) over species 1..51; N2 holds no C/H/O
    const double dd = dm[i] - (over ? m1[i] * (1.0 - 1.0 / ftot) : 0.0) + sp[51].mel[i] * (yn - y0n);
    res = fmax(res, fabs(dd));
  }
  const double sum = over ? tot / ftot + stm / ftot : tot + stm + yn;
  const double T = T0 + dT;
  const bool in = ok && stm <= th->stmf_max && stn <= th->stnd_max && T >= th->tmin && T <= th->tmax;
  const bool fast = !(slow && fabs(dT) <= tol_dt && dmax <= tol_dy);
  const int why = !live ? -1 : (!in ? 1 : (!(res <= tol_el) ? 2 : (!(fabs(sum - 1.0) <= tol_sum) ? 3 : (fast ? 4 : 0))));
  // the applied state: Y (N2 closed, renormalised when over), and HT, the sensible
  // enthalpy of the new T and Y, whose inverse is solveh.f:1170-1230
  const int ip = T > 1000.0 ? 7 : 0;
  double pp[6] = {0, 0, 0, 0, 0, 0}, hf = 0.0;
#pragma unroll
  for (int j = 0; j < GSP; ++j) {
    const int k = part + GPART * j;
    const GateSp& e = sp[k];
    const double v = k < 51 ? (over ? y1[j] / ftot : y1[j]) : yn;
    const double sw = v * e.rwm;
#pragma unroll
    for (int i = 0; i < 6; ++i) pp[i] += e.nasah[ip + i] * sw;
    hf += v * e.hfo * th->hstr;
    if (why == 0) y[(int64_t)k * n + c] = v;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) pp[i] = part_sum(pp[i]);
  hf = part_sum(hf);
  const double H =
      ((pp[0] + (pp[1] + (pp[2] + (pp[3] + pp[4] * T) * T) * T) * T) * T + pp[5]) * th->gasc / th->wmstr;
  if (part == 0 && why == 0) {
    tk[c] = T / th->tstr;
    ht[c] = (H - hf) / th->hstr;
  }
  if (part == 0 && live) {
    if (reject) reject[c] = why != 0;
    if (mask) mask[c] = why == 0 ? 10 : iskip[c];
    gc.add(why, rr);
  }
  gate_flush(gc, s_st, stats);
}

// ---- the network on a field: features -> MLP -> outputs, 32 cells per wave in two tiles
// of 16, one launch (k_gate follows).
// The network runs transposed, Z^T = W H^T, on v_mfma_f32_16x16x32_bf16: W is the A
// operand (rows = output neurons) and H^T the B operand (columns = a tile's 16 cells).
// Lane l = 16g + cl holds cell cl of each tile; a layer's result tile t gives it neurons
// 16t + 4g + r (r = 0..3), which, after bias, GELU and rounding to bf16, are the next
// layer's B operand without leaving registers: k-step s takes tile 2s (j = 0..3) and
// tile 2s+1 (j = 4..7), and the next layer's weight columns are stored in that order
// (k_pack_infer, perm = 1).  The weights pass through LDS in groups of output tiles
// (infer_layer), the activations never leave registers; the 64 outputs per cell go to out
// and the Y^0.1 of its features to pbuf, both species-major for the gate's coalesced loads.
struct InferArgs {
  const bf16_t* w[5];  // packed weights: [256][64], 3 x [256][256] (perm), [64][256] (perm)
  const float* b[5];   // biases: 4 x [256], [64]
  const Norm* nm;
  const double *y, *tk;
  float* out;          // [64][ldo] standardised outputs, species-major
  double* pbuf;        // [52][ldo] Y^0.1, for the gate
  int64_t ldo;
  const int32_t* cells;
  int64_t n;
  int32_t ncells;
  double tstr, log_dt;
};

// Output tiles T0 .. T0+NT-1 of a layer (rows 16 T0 .. 16 (T0+NT) - 1 of its weights) into
// the block's LDS: [KSTEPS][16 NT][4] chunks of 16 bytes (8 bf16 of one row and k-step),
// chunk c of row r at position c ^ ((r >> 2) & 3), so that the 16 lanes one 16-byte LDS
// read serves at once (rows cl = 0..15 of a tile, one chunk) fall on 16 different bank
// groups.  The global layout is [ROWS][K] (k_pack_infer).  Every thread of the block
// calls this; the caller synchronises before and after.
template <int NT, int KSTEPS>
__device__ __forceinline__ void stage_tiles(const bf16_t* __restrict__ w, int t0, bf16_t* __restrict__ lds) {
  constexpr int ROWS = NT * 16, KC = KSTEPS * 4, NCH = ROWS * KC, NTH = 64 * NET_WAVES, R = 4;
  static_assert(NCH % (NTH * R) == 0, "whole rounds of chunks");
  const bf16_t* src = w + (int64_t)t0 * 16 * KSTEPS * 32;
#pragma unroll 1
  for (int rd = 0; rd < NCH / (NTH * R); ++rd) {
    bf16x8 v[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int ch = (rd * R + i) * NTH + threadIdx.x;   // consecutive threads: consecutive 16 B of a row
      v[i] = *(const bf16x8*)(src + (int64_t)ch * 8);
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int ch = (rd * R + i) * NTH + threadIdx.x;
      const int row = ch / KC, kc = ch % KC, st = kc >> 2, c = kc & 3;
      *(bf16x8*)(lds + ((st * ROWS + row) * 4 + (c ^ ((row >> 2) & 3))) * 8) = v[i];
    }
  }
}

// Output tiles T0 .. T0+NT-1 of a layer, Z^T = W H^T, for the wave's two tiles of 16
// cells, from the weights stage_tiles put in LDS: each A fragment (16 bytes a lane) serves
// both cell tiles' MFMAs.  Per output, the k-steps accumulate in order, so the results do
// not depend on how the tiles are grouped.
template <int NT, int KSTEPS>
__device__ __forceinline__ void tiles_mfma(const bf16_t* __restrict__ lds, int t0, const bf16x8 (*in)[8],
                                           f32x4 (*acc)[16], int cl, int g) {
  constexpr int ROWS = NT * 16;
  const int sw = g ^ ((cl >> 2) & 3);
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[u][t0 + t][r] = 0.f;
#pragma unroll
  for (int st = 0; st < KSTEPS; ++st)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const bf16x8 wa = *(const bf16x8*)(lds + ((st * ROWS + t * 16 + cl) * 4 + sw) * 8);
      acc[0][t0 + t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, in[0][st], acc[0][t0 + t], 0, 0, 0);
      acc[1][t0 + t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, in[1][st], acc[1][t0 + t], 0, 0, 0);
    }
}

// a layer of NTILES output tiles in LDS-sized groups of NT tiles; the block synchronises
// around each group's staging
template <int NTILES, int NT, int KSTEPS>
__device__ __forceinline__ void infer_layer(const bf16_t* __restrict__ w, bf16_t* __restrict__ lds,
                                            const bf16x8 (*in)[8], f32x4 (*acc)[16], int cl, int g) {
#pragma unroll
  for (int t0 = 0; t0 < NTILES; t0 += NT) {
    stage_tiles<NT, KSTEPS>(w, t0, lds);
    __syncthreads();
    tiles_mfma<NT, KSTEPS>(lds, t0, in, acc, cl, g);
    __syncthreads();
  }
}

// bias, GELU and bf16: a layer's result tiles become the next layer's B operand
__device__ __forceinline__ void infer_activate(const f32x4* acc, const float* __restrict__ b, int g, bf16x8* act) {
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int t = 2 * s + j / 4, r = j % 4;
      act[s][j] = (__bf16)gelu(acc[t][r] + b[16 * t + 4 * g + r]);
    }
}

__global__ void __launch_bounds__(64 * NET_WAVES, 2) k_network(InferArgs a) {
  extern __shared__ __attribute__((aligned(16))) bf16_t wlds[];  // a group of a layer's weight tiles (stage_tiles)
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64, g = l / 16, cl = l % 16;
  const Norm* nm = a.nm;
  // features: this lane supplies k = 32 s + 8 g + j of cell cl of each of its two tiles
  bf16x8 act[2][8];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int64_t q = ((int64_t)blockIdx.x * NET_WAVES + wv) * 32 + 16 * u + cl;
    const bool live = q < a.ncells;
    const int64_t c = live ? a.cells[q] : 0;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int f = 32 * s + 8 * g + j;
        float v = 0.f;
        if (live && f < NF) {
          double raw;
          if (f >= 2) {
            const double P = bc_base(a.y[(int64_t)(f - 2) * a.n + c]);
            a.pbuf[(int64_t)(f - 2) * a.ldo + q] = P;
            raw = (P - 1.0) / 0.1;
          } else {
            raw = f == 0 ? a.tk[c] * a.tstr : a.log_dt;
          }
          v = (float)((raw - nm->xmean[f]) / nm->xsd[f]);
        }
        act[u][s][j] = (__bf16)v;
      }
  }
  f32x4 acc[2][16];
  infer_layer<16, 16, 2>(a.w[0], wlds, act, acc, cl, g);
#pragma unroll 1
  for (int L = 1; L < 4; ++L) {
#pragma unroll
    for (int u = 0; u < 2; ++u) infer_activate(acc[u], a.b[L - 1], g, act[u]);
    infer_layer<16, 8, 8>(a.w[L], wlds, act, acc, cl, g);
  }
#pragma unroll
  for (int u = 0; u < 2; ++u) infer_activate(acc[u], a.b[3], g, act[u]);
  infer_layer<4, 4, 8>(a.w[4], wlds, act, acc, cl, g);
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int64_t q = ((int64_t)blockIdx.x * NET_WAVES + wv) * 32 + 16 * u + cl;
    if (q < a.ncells)
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int o = 16 * t + 4 * g + r;
          a.out[o * a.ldo + q] = acc[u][t][r] + a.b[4][o];
        }
  }
}

// density of every cell from (Y, T) (main:1056-1061), then the grid's edges: the last
// I row extrapolated, the J=1 column copied from J=2 (main:1062-1068).  Used after a
// step in which the gate accepted every cell, so that no FP64 finish ran.
__global__ void k_density_all(const Thermo* __restrict__ th, const double* __restrict__ y,
                              const double* __restrict__ tk, double* __restrict__ rho, int64_t n) {
#pragma clang fp contract(off)
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double rbar = 0.0;
  for (int k = 0; k < NSPS; ++k) rbar = rbar + y[k * n + c] / th->wm[k];   // main:1058
  rho[c] = th->pback / (th->beta1 * tk[c] * rbar);
}

__global__ void k_density_edges(double* __restrict__ rho, int64_t li, int64_t lj) {
  const int64_t n = li * lj;
  for (int64_t j = threadIdx.x; j < lj; j += blockDim.x) rho[n - lj + j] = 2.0 * rho[n - 2 * lj + j] - rho[n - 3 * lj + j];
  __syncthreads();
  for (int64_t i = threadIdx.x; i < li; i += blockDim.x) rho[i * lj] = rho[i * lj + 1];
}

}  // namespace

extern "C" {

int us_collect(const double* y0, const double* tk0, const double* y1, const double* tk1, int64_t n,
               const int32_t* cells, int32_t ncells, double tstr, double log_dt, float* xraw, float* traw,
               int64_t row0, void* stream) {
  k_collect<<<blocks(ncells, 256), 256, 0, (hipStream_t)stream>>>(y0, tk0, y1, tk1, n, cells, ncells, tstr, log_dt,
                                                                   xraw, traw, row0);
  HIPCHK(hipGetLastError());
  return 0;
}

// per-column partial statistics: part[(chunk * ncol + j) * 4 + {sum, sumsq, min, max}]
int us_colstats(const float* raw, int64_t rows, int ncol, int kind, const float* tscale, int64_t chunk, double* part,
                void* stream) {
  if (kind == 1 && !tscale) return fail(-1, "us_colstats: targets need their scales");
  const int64_t nchunk = (rows + chunk - 1) / chunk;
  k_colstats<<<dim3(ncol, nchunk), 256, 0, (hipStream_t)stream>>>(raw, rows, ncol, kind, tscale, chunk, part);
  HIPCHK(hipGetLastError());
  return (int)nchunk;
}

int us_batch(const float* xraw, const float* traw, const int64_t* idx, int B, const void* norm, uint16_t* X,
             uint16_t* XT, float* T, void* stream) {
  k_batch<<<blocks((int64_t)B * NPAD, 256), 256, 0, (hipStream_t)stream>>>(xraw, traw, idx, B, (const Norm*)norm,
                                                                          (bf16_t*)X, (bf16_t*)XT, T);
  HIPCHK(hipGetLastError());
  return 0;
}

int us_features(const double* y, const double* tk, int64_t n, const int32_t* cells, int32_t ncells, int32_t bpad,
                double tstr, double log_dt, const void* norm, uint16_t* X, void* stream) {
  k_features<<<blocks((int64_t)bpad * NPAD, 256), 256, 0, (hipStream_t)stream>>>(y, tk, n, cells, ncells, bpad, tstr,
                                                                                log_dt, (const Norm*)norm,
                                                                                (bf16_t*)X);
  HIPCHK(hipGetLastError());
  return 0;
}

int us_gate(const float* out, const void* norm, const void* thermo, double* y, double* tk, double* ht,
            const double* stmf, const double* stnd, int64_t n, const int32_t* cells, int32_t ncells, double log_dt,
            double tol_el, double tol_sum, double tol_dt, double tol_dy, double tol_bc, int32_t* reject,
            unsigned long long* stats, void* stream) {
  if (ncells <= 0) return 0;
  k_gate<false><<<blocks(ncells, GATE_CELLS), GATE_BLOCK, 0, (hipStream_t)stream>>>(out, NPAD, 1, nullptr, 0, (const Norm*)norm,
                                                                (const Thermo*)thermo, y, tk, ht, stmf, stnd, n, cells,
                                                                ncells, log_dt, tol_el, tol_sum, tol_dt, tol_dy, tol_bc,
                                                                reject, nullptr, nullptr, stats);
  HIPCHK(hipGetLastError());
  return 0;
}

// the surrogate's prediction on a field: k_network (features and MLP; outputs to out,
// [64][ldo] floats, Y^0.1 to pbuf, [52][ldo] doubles, ldo >= ncells) then k_gate (mask
// and stats); w and b: five device pointers each, as packed by um_pack_infer and the
// network's fp32 biases
int us_infer(const uint16_t* const* w, const float* const* b, const void* norm, const void* thermo, double* y,
             double* tk, double* ht, const double* stmf, const double* stnd, const int32_t* iskip, int32_t* mask,
             float* out, double* pbuf, int64_t ldo, int64_t n, const int32_t* cells, int32_t ncells, double tstr,
             double log_dt, double tol_el, double tol_sum, double tol_dt, double tol_dy, double tol_bc,
             unsigned long long* stats, void* stream) {
  if (ncells < 0 || n < 1 || !out || !pbuf || !mask || ldo < ncells) return fail(-1, "us_infer: bad sizes or buffers");
  if (ncells == 0) return 0;
  InferArgs a{};
  for (int i = 0; i < 5; ++i) {
    a.w[i] = (const bf16_t*)w[i];
    a.b[i] = b[i];
  }
  a.nm = (const Norm*)norm;
  a.y = y; a.tk = tk; a.out = out; a.pbuf = pbuf; a.ldo = ldo; a.cells = cells; a.n = n; a.ncells = ncells;
  a.tstr = tstr; a.log_dt = log_dt;
  hipStream_t s = (hipStream_t)stream;
  // the weights' LDS goes beyond the default 64 KB (set on the current device)
  HIPCHK(hipFuncSetAttribute((const void*)k_network, hipFuncAttributeMaxDynamicSharedMemorySize, NET_LDS));
  k_network<<<blocks(ncells, 32 * NET_WAVES), 64 * NET_WAVES, NET_LDS, s>>>(a);
  k_gate<true><<<blocks(ncells, GATE_CELLS), GATE_BLOCK, 0, s>>>(out, 1, ldo, pbuf, ldo, (const Norm*)norm, (const Thermo*)thermo, y, tk,
                                              ht, stmf, stnd, n, cells, ncells, log_dt, tol_el, tol_sum, tol_dt,
                                              tol_dy, tol_bc, nullptr, iskip, mask, stats);
  HIPCHK(hipGetLastError());
  return 0;
}

int us_density(const void* thermo, const double* y, const double* tk, double* rho, int64_t li, int64_t lj,
               void* stream) {
  if (li < 3 || lj < 2) return fail(-1, "us_density: grid too small");
  const int64_t n = li * lj;
  k_density_all<<<blocks(n, 256), 256, 0, (hipStream_t)stream>>>((const Thermo*)thermo, y, tk, rho, n);
  k_density_edges<<<1, 256, 0, (hipStream_t)stream>>>(rho, li, lj);
  HIPCHK(hipGetLastError());
  return 0;
}

int us_sizes(int64_t* out) {  // sizes of the parameter blocks, for the host layer to check
  out[0] = sizeof(Norm);
  out[1] = sizeof(Thermo);
  return 0;
}

}  // extern "C"
