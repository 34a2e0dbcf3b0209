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
// solveh.f:1170-1230).  Otherwise the cell is left as it was.  Two passes over the
// species (the second recomputes the first's values with the same arithmetic) keep
// the registers few.  The soot scalars are not predicted: an accepted cell keeps its
// STMF and STND.  Returns 0 accepted, 1 out of range, 2 residual, 3 sum, 4 active;
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

// Y0^0.1, the base of the Box-Cox transform (Y < 0 taken as 0)
__device__ __forceinline__ double bc_base(double y) { return pow(y > 0.0 ? y : 0.0, 0.1); }

// Y after a Box-Cox change d, from P = Y0^0.1: BC^-1(BC(Y0) + d) = (P + 0.1 d)^10 (0 when
// the base falls below 0)
__device__ __forceinline__ double y_after(double P, float d) {
  double u = P + 0.1 * (double)d;
  u = u > 0.0 ? u : 0.0;
  const double u2 = u * u, u4 = u2 * u2, u8 = u4 * u4;
  return u8 * u2;
}

// The gate's arithmetic for one cell.  D d: d(j), the change of output j (j = 0: dT in K;
// j = 1 + k: the Box-Cox change of Y_k); P p: p(k) = Y0_k^0.1; Y0 y0f: y0f(k) = Y0_k.
// flags_ok: the features lie in the hull and every output in range; slow_bc: every
// species' Box-Cox change within tol_bc.
template <class D, class P, class Y0>
__device__ int gate_cell(D d, P p, Y0 y0f, const Thermo* __restrict__ th, double* __restrict__ y, double* __restrict__ tk,
                         double* __restrict__ ht, double stm, double stn, int64_t n, int64_t c, bool flags_ok,
                         bool slow_bc, double tol_el, double tol_sum, double tol_dt, double tol_dy, double* rr) {
  // raw residual r = A dY and the weighted normal matrix M = A W A^T (symmetric 3x3)
  double m[6] = {0, 0, 0, 0, 0, 0};
  rr[0] = rr[1] = rr[2] = 0.0;
  for (int k = 0; k < 49; ++k) {
    const double y0 = y0f(k), y1 = y_after(p(k), d(1 + k));
    const double w = fmax(y0, y1), a0 = th->mel[0][k], a1 = th->mel[1][k], a2 = th->mel[2][k];
    rr[0] += a0 * (y1 - y0);
    rr[1] += a1 * (y1 - y0);
    rr[2] += a2 * (y1 - y0);
    m[0] += w * a0 * a0; m[1] += w * a0 * a1; m[2] += w * a0 * a2;
    m[3] += w * a1 * a1; m[4] += w * a1 * a2; m[5] += w * a2 * a2;
  }
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
  // Y1 of species k < 51: predicted, corrected, clipped (both passes use this)
  auto y1_of = [&](int k, double y0) {
    double y1 = y_after(p(k), d(1 + k));
    if (k < 49) y1 -= fmax(y0, y1) * (th->mel[0][k] * lam[0] + th->mel[1][k] * lam[1] + th->mel[2][k] * lam[2]);
    return y1 < 0.0 ? 0.0 : y1;
  };
  double tot = 0.0, dm[3] = {0.0, 0.0, 0.0}, m1[3] = {0.0, 0.0, 0.0}, dmax = 0.0;
  for (int k = 0; k < 51; ++k) {
    const double y0 = y0f(k), y1 = y1_of(k, y0);
    tot += y1;
    dmax = fmax(dmax, fabs(y1 - y0));
    for (int e = 0; e < 3; ++e) {
      dm[e] += th->mel[e][k] * (y1 - y0);
      m1[e] += th->mel[e][k] * y1;
    }
  }
  const double ftot = tot + stm;
  const bool over = ftot > 1.0;
  const double yn = over ? 0.0 : 1.0 - ftot;
  const double y0n = y0f(51);
  dmax = fmax(dmax, fabs(yn - y0n));
  double res = 0.0;
  for (int e = 0; e < 3; ++e) {
    // y1/ftot - y0 = (y1 - y0) - y1 (1 - 1/ftot) over species 1..51; N2 holds no C/H/O
    const double dd = dm[e] - (over ? m1[e] * (1.0 - 1.0 / ftot) : 0.0) + th->mel[e][51] * (yn - y0n);
    res = fmax(res, fabs(dd));
  }
  const double sum = over ? tot / ftot + stm / ftot : tot + stm + yn;
  const double dT = (double)d(0), T = tk[c] * th->tstr + dT;
  const bool in = flags_ok && stm <= th->stmf_max && stn <= th->stnd_max && T >= th->tmin && T <= th->tmax;
  const bool slow = slow_bc && fabs(dT) <= tol_dt && dmax <= tol_dy;
  const int why = !in ? 1 : (!(res <= tol_el) ? 2 : (!(fabs(sum - 1.0) <= tol_sum) ? 3 : (!slow ? 4 : 0)));
  if (why) return why;
  const int ip = T > 1000.0 ? 7 : 0;
  double pp[6] = {0, 0, 0, 0, 0, 0}, hf = 0.0;
  for (int k = 0; k < NSPS; ++k) {
    double y1 = yn;
    if (k < 51) {
      y1 = y1_of(k, y0f(k));
      if (over) y1 = y1 / ftot;
    }
    const double sw = y1 * th->rwm[k];
    for (int i = 0; i < 6; ++i) pp[i] += th->nasah[k][ip + i] * sw;
    hf += y1 * th->hfo[k] * th->hstr;
    y[(int64_t)k * n + c] = y1;
  }
  const double H =
      ((pp[0] + (pp[1] + (pp[2] + (pp[3] + pp[4] * T) * T) * T) * T) * T + pp[5]) * th->gasc / th->wmstr;
  tk[c] = T / th->tstr;
  ht[c] = (H - hf) / th->hstr;
  return 0;
}

// per-wave statistics of the gate: stats [0] accepted, [1] out of range, [2] residual,
// [3] sum, [4] active (why = -1: no cell), [5..7] max |raw C/H/O residual| (bits of a
// double); one atomic per wave and counter
__device__ __forceinline__ void gate_stats(int why, const double* rr, unsigned long long* stats) {
  const int l = threadIdx.x % 64;
  for (int k = 0; k < 5; ++k) {
    const unsigned long long m = __ballot(why == k);
    if (l == 0 && m) atomicAdd(&stats[k], (unsigned long long)__popcll(m));
  }
  for (int e = 0; e < 3; ++e) {
    double v = why >= 0 ? fabs(rr[e]) : 0.0;
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if (l == 0 && v > 0.0) atomicMax(&stats[5 + e], (unsigned long long)__double_as_longlong(v));
  }
}

// the gate as a kernel of its own, on the outputs of the separate-kernel network
// (out [ncells][64]); reject[c] = 1 where rejected, 0 where accepted
__global__ void k_gate(const float* __restrict__ out, const Norm* __restrict__ nm, const Thermo* __restrict__ th,
                       double* __restrict__ y, double* __restrict__ tk, double* __restrict__ ht,
                       const double* __restrict__ stmf, const double* __restrict__ stnd, int64_t n,
                       const int32_t* __restrict__ cells, int32_t ncells, double log_dt, double tol_el,
                       double tol_sum, double tol_dt, double tol_dy, double tol_bc, int32_t* __restrict__ reject,
                       unsigned long long* __restrict__ stats) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int why = -1;
  double rr[3] = {0.0, 0.0, 0.0};
  if (q < ncells) {
    const int64_t c = cells[q];
    const double T0 = tk[c] * th->tstr;
    bool ok = true, slow = true;
    for (int j = 0; j < NF && ok; ++j) {
      const double v = j == 0 ? T0 : (j == 1 ? log_dt : (bc_base(y[(int64_t)(j - 2) * n + c]) - 1.0) / 0.1);
      ok = v >= th->xlo[j] && v <= th->xhi[j];
    }
    const float* o = out + q * NPAD;
    float dv[NT];
    for (int j = 0; j < NT; ++j) {
      bool a, b;
      dv[j] = out_change(o[j], j, nm, th, tol_bc, a, b);
      ok = ok && a;
      slow = slow && b;
    }
    why = gate_cell([&](int j) { return dv[j]; }, [&](int k) { return bc_base(y[(int64_t)k * n + c]); },
                    [&](int k) { return y[(int64_t)k * n + c]; }, th, y, tk, ht, stmf[c], stnd[c], n, c, ok, slow,
                    tol_el, tol_sum, tol_dt, tol_dy, rr);
    reject[c] = why != 0;
  }
  gate_stats(why, rr, stats);
}

// ---- fused inference: features -> network -> gate, 32 cells per wave, one launch.
// The network runs transposed, Z^T = W H^T: W is the MFMA's A operand (rows = output
// neurons) and H^T its B operand (columns = the wave's 32 cells), so a layer's result
// fragment (lane l: cell l%32, neurons 8(r/4) + 4(l/32) + r%4) is, after bias, GELU
// and rounding to bf16, the next layer's B operand without leaving registers: k-step
// 2t+p takes registers 8p..8p+7 of result tile t, and the next layer's weight columns
// are stored in that order (k_pack_infer, perm = 1).  Only the 64 outputs per cell pass
// through LDS, to the lane that runs the cell's FP64 gate.
struct InferArgs {
  const bf16_t* w[5];  // packed weights: [256][64], 3 x [256][256] (perm), [64][256] (perm)
  const float* b[5];   // biases: 4 x [256], [64]
  const Norm* nm;
  const Thermo* th;
  double *y, *tk, *ht;  // rho is left to the step's finish (FP64 fallback or us_density)
  const double *stmf, *stnd;
  const int32_t* iskip;
  int32_t* mask;       // ISKIP for the FP64 fallback: 10 where accepted, ISKIP[c] where rejected
  float* out;          // optional [ncells][64] standardised outputs (tests), may be null
  const int32_t* cells;
  int64_t n;
  int32_t ncells;
  double log_dt, tol_el, tol_sum, tol_dt, tol_dy, tol_bc;
  unsigned long long* stats;  // as k_gate
};

template <int NTILES, int KSTEPS>
__device__ __forceinline__ void infer_layer(const bf16_t* __restrict__ w, const bf16x8* in, f32x16* acc, int cl,
                                            int h) {
  constexpr int K = KSTEPS * 16;
#pragma unroll
  for (int t = 0; t < NTILES; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[t][r] = 0.f;
#pragma unroll
  for (int s = 0; s < KSTEPS; ++s) {
    bf16x8 wa[NTILES];
#pragma unroll
    for (int t = 0; t < NTILES; ++t) wa[t] = *(const bf16x8*)(w + (t * 32 + cl) * K + 16 * s + 8 * h);
#pragma unroll
    for (int t = 0; t < NTILES; ++t) acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[t], in[s], acc[t], 0, 0, 0);
  }
}

__device__ __forceinline__ void infer_activate(const f32x16* acc, const float* __restrict__ b, int h, bf16x8* act) {
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) act[2 * t + r / 8][r % 8] = (__bf16)gelu(acc[t][r] + b[32 * t + 8 * (r / 4) + 4 * h + r % 4]);
}

__global__ void __launch_bounds__(256) k_infer(InferArgs a) {
  __shared__ float sh[4][32][65];    // outputs, then the changes they stand for
  __shared__ double shp[4][32][53];  // Y^0.1 of the 52 species
  __shared__ double shy[4][32][53];  // Y of the 52 species
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64, h = l / 32, cl = l % 32;
  const int64_t q = ((int64_t)blockIdx.x * 4 + wv) * 32 + cl;
  const bool live = q < a.ncells;
  const int64_t c = live ? a.cells[q] : 0;
  const Norm* nm = a.nm;
  const Thermo* th = a.th;
  const double T0 = live ? a.tk[c] * th->tstr : 0.0;
  // features: this lane supplies k = 16 s + 8 h + j of its cell; Y^0.1 of its species
  // kept in LDS for the gate
  bf16x8 x[4];
  int inside = 1;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int f = 16 * s + 8 * h + j;
      float v = 0.f;
      if (live && f < NF) {
        double raw;
        if (f >= 2) {
          const double yv = a.y[(int64_t)(f - 2) * a.n + c];
          const double P = bc_base(yv);
          shy[wv][cl][f - 2] = yv;
          shp[wv][cl][f - 2] = P;
          raw = (P - 1.0) / 0.1;
        } else {
          raw = f == 0 ? T0 : a.log_dt;
        }
        inside = inside && raw >= th->xlo[f] && raw <= th->xhi[f];
        v = (float)((raw - nm->xmean[f]) / nm->xsd[f]);
      }
      x[s][j] = (__bf16)v;
    }
  inside = inside && __shfl_xor(inside, 32);
  f32x16 acc[8];
  bf16x8 act[16];
  infer_layer<8, 4>(a.w[0], x, acc, cl, h);
  infer_activate(acc, a.b[0], h, act);
#pragma unroll 1
  for (int L = 1; L < 4; ++L) {
    infer_layer<8, 16>(a.w[L], act, acc, cl, h);
    infer_activate(acc, a.b[L], h, act);
  }
  infer_layer<2, 16>(a.w[4], act, acc, cl, h);
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = 32 * t + 8 * (r / 4) + 4 * h + r % 4;
      const float v = acc[t][r] + a.b[4][o];
      sh[wv][cl][o] = v;
      if (a.out && live) a.out[q * NPAD + o] = v;
    }
  __syncthreads();
  // outputs -> changes, range and slow-chemistry flags: both lanes of a cell, alternate
  // outputs
  bool ok = inside != 0, slow = true;
  if (live)
    for (int j = h; j < NT; j += 2) {
      bool r0, r1;
      sh[wv][cl][j] = out_change(sh[wv][cl][j], j, nm, th, a.tol_bc, r0, r1);
      ok = ok && r0;
      slow = slow && r1;
    }
  ok = ok && __shfl_xor((int)ok, 32);
  slow = slow && __shfl_xor((int)slow, 32);
  __syncthreads();
  // the gate, one lane per cell
  int why = -1;
  double rr[3] = {0.0, 0.0, 0.0};
  if (live && h == 0) {
    const float* dv = sh[wv][cl];
    const double* pv = shp[wv][cl];
    const double* yv = shy[wv][cl];
    why = gate_cell([&](int j) { return dv[j]; }, [&](int k) { return pv[k]; }, [&](int k) { return yv[k]; }, th,
                    a.y, a.tk, a.ht, a.stmf[c], a.stnd[c], a.n, c, ok, slow, a.tol_el, a.tol_sum, a.tol_dt, a.tol_dy,
                    rr);
    a.mask[c] = why == 0 ? 10 : a.iskip[c];
  }
  gate_stats(why, rr, a.stats);
}

// density of every cell from (Y, T) (main:1056-1061), then the grid's edges: the last
// I row extrapolated, the J=1 column copied from J=2 (main:1062-1068).  Used after a
// step in which the gate accepted every cell, so that no FP64 finish ran.
__global__ void k_density_all(const Thermo* __restrict__ th, const double* __restrict__ y,
                              const double* __restrict__ tk, double* __restrict__ rho, int64_t n) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double rbar = 0.0;
  for (int k = 0; k < NSPS; ++k) rbar += y[k * n + c] * th->rwm[k];
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
  k_gate<<<blocks(ncells, 128), 128, 0, (hipStream_t)stream>>>(out, (const Norm*)norm, (const Thermo*)thermo, y, tk, ht,
                                                               stmf, stnd, n, cells, ncells, log_dt, tol_el, tol_sum,
                                                               tol_dt, tol_dy, tol_bc, reject, stats);
  HIPCHK(hipGetLastError());
  return 0;
}

// the fused step (k_infer); w and b: five device pointers each, as packed by
// um_pack_infer and the network's fp32 biases
int us_infer(const uint16_t* const* w, const float* const* b, const void* norm, const void* thermo, double* y,
             double* tk, double* ht, const double* stmf, const double* stnd, const int32_t* iskip,
             int32_t* mask, float* out, int64_t n, const int32_t* cells, int32_t ncells, double log_dt, double tol_el,
             double tol_sum, double tol_dt, double tol_dy, double tol_bc, unsigned long long* stats, void* stream) {
  if (ncells < 0 || n < 1) return fail(-1, "us_infer: bad sizes");
  if (ncells == 0) return 0;
  InferArgs a{};
  for (int i = 0; i < 5; ++i) {
    a.w[i] = (const bf16_t*)w[i];
    a.b[i] = b[i];
  }
  a.nm = (const Norm*)norm;
  a.th = (const Thermo*)thermo;
  a.y = y; a.tk = tk; a.ht = ht; a.stmf = stmf; a.stnd = stnd;
  a.iskip = iskip; a.mask = mask; a.out = out; a.cells = cells; a.n = n; a.ncells = ncells;
  a.log_dt = log_dt; a.tol_el = tol_el; a.tol_sum = tol_sum; a.tol_dt = tol_dt; a.tol_dy = tol_dy; a.tol_bc = tol_bc;
  a.stats = stats;
  k_infer<<<blocks(ncells, 128), 256, 0, (hipStream_t)stream>>>(a);
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
