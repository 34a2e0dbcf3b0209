// Chemistry side of the surrogate (MI355X / gfx950): training pairs from FP64 steps,
// network features, and the per-cell gate that applies accepted predictions.  No
// counterpart in the reference.
//
// Features (54, then padded to 64): T in K, ln(dt / s), and the Box-Cox transform
// (lambda = 0.1) of the 52 mass fractions, each standardised with the training set's
// mean and standard deviation.  Targets (53): the step's dT (K) and dY_k, standardised
// the same way.
#include "mlp_kernels.hip"

namespace {

constexpr int NSPS = 52;
constexpr int NF = 54, NT = 53, NPAD = 64;

__device__ __forceinline__ double boxcox(double y) { return (pow(y > 0.0 ? y : 0.0, 0.1) - 1.0) / 0.1; }

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
    t[1 + k] = (float)(y1[k * n + c] - y0[k * n + c]);
  }
}

// transformed value of raw feature j (Box-Cox for the mass fractions)
__device__ __forceinline__ double feat(const float* x, int j) { return j < 2 ? (double)x[j] : boxcox((double)x[j]); }

// per-column sums, sums of squares, min and max of the transformed features (kind 0)
// or of the targets (kind 1), over rows; one block per (column, chunk) -> partials
__global__ void k_colstats(const float* __restrict__ raw, int64_t rows, int ncol, int kind, int64_t chunk,
                           double* __restrict__ part) {
  __shared__ double s1[256], s2[256], mn[256], mx[256];
  const int j = blockIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * chunk, r1 = min(rows, r0 + chunk);
  double a = 0.0, b = 0.0, lo = 1e300, hi = -1e300;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    const double v = kind == 0 ? feat(raw + r * ncol, j) : (double)raw[r * ncol + j];
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

struct Norm {  // standardisation: (v - mean) / sd
  float xmean[NF], xsd[NF], tmean[NT], tsd[NT];
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
  if (j < NT) tv = (traw[r * NT + j] - nm->tmean[j]) / nm->tsd[j];
  X[(int64_t)b * NPAD + j] = to_bf16(xv);
  XT[(int64_t)j * B + b] = to_bf16(xv);
  T[(int64_t)b * NPAD + j] = tv;
}

// ---- inference on a field
struct Thermo {  // what the gate needs to turn (T, Y) into the field's variables
  double tstr, hstr, gasc, wmstr, pback, beta1;
  double rwm[NSPS], hfo[NSPS], nasa[NSPS][14];
  double mel[3][NSPS];   // element (C, H, O) mass per unit mass of each species
  double proj[49][3];    // S A^T (A S A^T)^-1: minimum-norm C/H/O correction (species 1..49)
  double xlo[NF], xhi[NF];  // training range of the transformed features (with a margin)
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

// The gate.  Per listed cell: un-standardise the network's dT and dY; correct dY over
// species 1..49 to conserve C, H and O exactly; clip negative mass fractions; N2 by
// closure (solvesp.f:2839-2914 form).  Accept when every transformed input lies in the
// training range, the C/H/O residual of the applied change is <= tol_el and
// |sum Y + STMF - 1| <= tol_sum; then write Y, TK, HT (the sensible enthalpy of the new
// T and Y, inverse of solveh.f:1170-1230) and rho (main:1056-1061).  Otherwise set
// reject[c] = 1 and leave the cell.  stats: [0] accepted, [1] out of range,
// [2] residual, [3] sum; [4..6] max |residual| before the correction (bits of a double)
__global__ void k_gate(const float* __restrict__ out, const Norm* __restrict__ nm, const Thermo* __restrict__ th,
                       double* __restrict__ y, double* __restrict__ tk, double* __restrict__ ht,
                       double* __restrict__ rho, const double* __restrict__ stmf, const double* __restrict__ stnd,
                       int64_t n, const int32_t* __restrict__ cells, int32_t ncells, double log_dt,
                       double tol_el, double tol_sum, int32_t* __restrict__ reject,
                       unsigned long long* __restrict__ stats) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncells) return;
  const int64_t c = cells[q];
  const float* o = out + q * NPAD;
  double y0[NSPS], dy[NSPS];
  for (int k = 0; k < NSPS; ++k) y0[k] = y[k * n + c];
  const double T0 = tk[c] * th->tstr;
  // input range (the hull of the training features, per feature)
  bool inside = stmf[c] <= th->stmf_max && stnd[c] <= th->stnd_max;
  for (int j = 0; j < NF && inside; ++j) {
    const double v = j == 0 ? T0 : (j == 1 ? log_dt : boxcox(y0[j - 2]));
    inside = v >= th->xlo[j] && v <= th->xhi[j];
  }
  const double dT = (double)o[0] * nm->tsd[0] + nm->tmean[0];
  for (int k = 0; k < NSPS; ++k) dy[k] = (double)o[1 + k] * nm->tsd[1 + k] + nm->tmean[1 + k];
  // C/H/O residual of the raw prediction, then the minimum-norm correction over species 1..49
  double r[3];
  for (int e = 0; e < 3; ++e) {
    double s = 0.0;
    for (int k = 0; k < 49; ++k) s += th->mel[e][k] * dy[k];
    r[e] = s;
  }
  for (int e = 0; e < 3; ++e) atomicMax(&stats[4 + e], (unsigned long long)__double_as_longlong(fabs(r[e])));
  for (int k = 0; k < 49; ++k) dy[k] -= th->proj[k][0] * r[0] + th->proj[k][1] * r[1] + th->proj[k][2] * r[2];
  double y1[NSPS], tot = 0.0;
  for (int k = 0; k < 51; ++k) {
    const double v = y0[k] + dy[k];
    y1[k] = v < 0.0 ? 0.0 : v;
    tot += y1[k];
  }
  tot += stmf[c];
  y1[51] = 1.0 - tot;
  if (y1[51] < 0.0) y1[51] = 0.0;
  double res = 0.0, sum = stmf[c];
  for (int e = 0; e < 3; ++e) {
    double s = 0.0;
    for (int k = 0; k < NSPS; ++k) s += th->mel[e][k] * (y1[k] - y0[k]);
    res = fmax(res, fabs(s));
  }
  for (int k = 0; k < NSPS; ++k) sum += y1[k];
  const bool ok_res = res <= tol_el, ok_sum = fabs(sum - 1.0) <= tol_sum;
  if (!inside || !ok_res || !ok_sum) {
    reject[c] = 1;
    atomicAdd(&stats[!inside ? 1 : (!ok_res ? 2 : 3)], 1ull);
    return;
  }
  reject[c] = 0;
  atomicAdd(&stats[0], 1ull);
  const double T = T0 + dT;
  const int ip = T > 1000.0 ? 7 : 0;
  double a[6] = {0, 0, 0, 0, 0, 0}, hf = 0.0, rbar = 0.0;
  for (int k = 0; k < NSPS; ++k) {
    const double sw = y1[k] * th->rwm[k];
    for (int i = 0; i < 6; ++i) a[i] += th->nasa[k][ip + i] * sw / (i < 5 ? (double)(i + 1) : 1.0);
    hf += y1[k] * th->hfo[k] * th->hstr;
    rbar += sw;
    y[k * n + c] = y1[k];
  }
  const double H = ((a[0] + (a[1] + (a[2] + (a[3] + a[4] * T) * T) * T) * T) * T + a[5]) * th->gasc / th->wmstr;
  tk[c] = T / th->tstr;
  ht[c] = (H - hf) / th->hstr;
  rho[c] = th->pback / (th->beta1 * (T / th->tstr) * rbar);
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
  double log_dt, tol_el, tol_sum;
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
  __shared__ float sh[4][32][65];
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64, h = l / 32, cl = l % 32;
  const int64_t q = ((int64_t)blockIdx.x * 4 + wv) * 32 + cl;
  const bool live = q < a.ncells;
  const int64_t c = live ? a.cells[q] : 0;
  const Norm* nm = a.nm;
  const Thermo* th = a.th;
  const double T0 = live ? a.tk[c] * th->tstr : 0.0;
  // features: this lane supplies k = 16 s + 8 h + j of its cell
  bf16x8 x[4];
  int inside = 1;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int f = 16 * s + 8 * h + j;
      float v = 0.f;
      if (live && f < NF) {
        const double raw = f == 0 ? T0 : (f == 1 ? a.log_dt : boxcox(a.y[(int64_t)(f - 2) * a.n + c]));
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
  // the gate (k_gate's rules), one lane per cell, in two passes over the species
  int why = -1;  // -1 no cell, 0 accepted, 1 out of range, 2 residual, 3 sum
  double rr[3] = {0.0, 0.0, 0.0};
  if (live && h == 0) {
    const float* o = sh[wv][cl];
    for (int k = 0; k < 49; ++k) {
      const double dy = (double)o[1 + k] * nm->tsd[1 + k] + nm->tmean[1 + k];
      for (int e = 0; e < 3; ++e) rr[e] += th->mel[e][k] * dy;
    }
    double tot = 0.0, ee[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 51; ++k) {
      double dy = (double)o[1 + k] * nm->tsd[1 + k] + nm->tmean[1 + k];
      if (k < 49) dy -= th->proj[k][0] * rr[0] + th->proj[k][1] * rr[1] + th->proj[k][2] * rr[2];
      const double y0 = a.y[(int64_t)k * a.n + c];
      double y1 = y0 + dy;
      y1 = y1 < 0.0 ? 0.0 : y1;
      tot += y1;
      for (int e = 0; e < 3; ++e) ee[e] += th->mel[e][k] * (y1 - y0);
    }
    const double sm = a.stmf[c];
    const double y0n = a.y[(int64_t)51 * a.n + c];
    double yn = 1.0 - (tot + sm);
    yn = yn < 0.0 ? 0.0 : yn;
    double res = 0.0;
    for (int e = 0; e < 3; ++e) res = fmax(res, fabs(ee[e] + th->mel[e][51] * (yn - y0n)));
    const double sum = tot + sm + yn;
    const bool in = inside && sm <= th->stmf_max && a.stnd[c] <= th->stnd_max;
    why = !in ? 1 : (!(res <= a.tol_el) ? 2 : (!(fabs(sum - 1.0) <= a.tol_sum) ? 3 : 0));
    a.mask[c] = why == 0 ? 10 : a.iskip[c];
    if (why == 0) {
      const double T = T0 + ((double)o[0] * nm->tsd[0] + nm->tmean[0]);
      const int ip = T > 1000.0 ? 7 : 0;
      double p[6] = {0, 0, 0, 0, 0, 0}, hf = 0.0;
      for (int k = 0; k < NSPS; ++k) {
        double y1 = yn;
        if (k < 51) {  // the same arithmetic as the first pass
          double dy = (double)o[1 + k] * nm->tsd[1 + k] + nm->tmean[1 + k];
          if (k < 49) dy -= th->proj[k][0] * rr[0] + th->proj[k][1] * rr[1] + th->proj[k][2] * rr[2];
          y1 = a.y[(int64_t)k * a.n + c] + dy;
          y1 = y1 < 0.0 ? 0.0 : y1;
        }
        const double sw = y1 * th->rwm[k];
        for (int i = 0; i < 6; ++i) p[i] += th->nasa[k][ip + i] * sw / (i < 5 ? (double)(i + 1) : 1.0);
        hf += y1 * th->hfo[k] * th->hstr;
        a.y[(int64_t)k * a.n + c] = y1;
      }
      const double H = ((p[0] + (p[1] + (p[2] + (p[3] + p[4] * T) * T) * T) * T) * T + p[5]) * th->gasc / th->wmstr;
      a.tk[c] = T / th->tstr;
      a.ht[c] = (H - hf) / th->hstr;
    }
  }
  // statistics: one atomic per wave and counter
  for (int k = 0; k < 4; ++k) {
    const unsigned long long m = __ballot(why == k);
    if (l == 0 && m) atomicAdd(&a.stats[k], (unsigned long long)__popcll(m));
  }
  for (int e = 0; e < 3; ++e) {
    double v = fabs(rr[e]);
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if (l == 0 && v > 0.0) atomicMax(&a.stats[4 + e], (unsigned long long)__double_as_longlong(v));
  }
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
int us_colstats(const float* raw, int64_t rows, int ncol, int kind, int64_t chunk, double* part, void* stream) {
  const int64_t nchunk = (rows + chunk - 1) / chunk;
  k_colstats<<<dim3(ncol, nchunk), 256, 0, (hipStream_t)stream>>>(raw, rows, ncol, kind, chunk, part);
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

int us_gate(const float* out, const void* norm, const void* thermo, double* y, double* tk, double* ht, double* rho,
            const double* stmf, const double* stnd, int64_t n, const int32_t* cells, int32_t ncells, double log_dt,
            double tol_el, double tol_sum, int32_t* reject, unsigned long long* stats, void* stream) {
  k_gate<<<blocks(ncells, 128), 128, 0, (hipStream_t)stream>>>(out, (const Norm*)norm, (const Thermo*)thermo, y, tk, ht,
                                                               rho, stmf, stnd, n, cells, ncells, log_dt, tol_el,
                                                               tol_sum, reject, stats);
  HIPCHK(hipGetLastError());
  return 0;
}

// the fused step (k_infer); w and b: five device pointers each, as packed by
// um_pack_infer and the network's fp32 biases
int us_infer(const uint16_t* const* w, const float* const* b, const void* norm, const void* thermo, double* y,
             double* tk, double* ht, const double* stmf, const double* stnd, const int32_t* iskip,
             int32_t* mask, float* out, int64_t n, const int32_t* cells, int32_t ncells, double log_dt, double tol_el,
             double tol_sum, unsigned long long* stats, void* stream) {
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
  a.log_dt = log_dt; a.tol_el = tol_el; a.tol_sum = tol_sum; a.stats = stats;
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
