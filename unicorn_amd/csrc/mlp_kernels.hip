// MI355X (gfx950) kernels of the chemistry surrogate: a 54 -> 4 x 256 (GELU) -> 53
// MLP with BF16 weights and activations, FP32 accumulation (MFMA), trained with Adam
// on FP32 master weights.  No counterpart in the reference.
//
// Shapes are padded: the input layer takes 64 features (54 used), the output layer
// gives 64 values (53 used); every M and N is a multiple of 64 and every K of 16.
// All GEMMs are "NT": C[M,N] = A[M,K] * B[N,K]^T with both operands K-contiguous, so
// the forward pass keeps every activation in both orientations (H and H^T) and the
// weights in both (W [out][in] and W^T [in][out]).
//
// The MFMA is v_mfma_f32_32x32x16_bf16: lane l holds A[l%32][(l/32)*8 + 0..7] and
// B[(l/32)*8 + 0..7][l%32]; its 16 results are D[8*(r/4) + (l/32)*4 + r%4][l%32].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16_t;  // bit pattern in memory

thread_local std::string g_err;
int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}
#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return fail(-2, std::string(#x ": ") + hipGetErrorString(e_));  \
  } while (0)

__device__ __forceinline__ bf16_t to_bf16(float x) {
  return __builtin_bit_cast(bf16_t, (__bf16)x);  // round to nearest even
}
__device__ __forceinline__ float from_bf16(bf16_t x) { return (float)__builtin_bit_cast(__bf16, x); }

// ---------------------------------------------------------------- GEMM (NT)
// Block: 4 waves, a 64 x 64 tile of C; wave (w/2, w%2) owns a 32 x 32 quarter.  K is
// walked in steps of 32 (two MFMAs), loads for the next step issued before the MFMAs.
// blockIdx.z: split of K (k0 .. k0+kchunk); with splits the partial goes to
// C + z*M*ldc and a second kernel reduces.
__global__ void __launch_bounds__(256) k_gemm_nt(const bf16_t* __restrict__ A, int lda, const bf16_t* __restrict__ B,
                                                 int ldb, float* __restrict__ C, int ldc, int M, int K, int kchunk) {
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  const int row0 = blockIdx.y * 64 + (w / 2) * 32, col0 = blockIdx.x * 64 + (w % 2) * 32;
  const int k0 = blockIdx.z * kchunk, k1 = min(K, k0 + kchunk);
  const bf16_t* pa = A + (int64_t)(row0 + l % 32) * lda + (l / 32) * 8;
  const bf16_t* pb = B + (int64_t)(col0 + l % 32) * ldb + (l / 32) * 8;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  for (int k = k0; k < k1; k += 32) {
    const bf16x8 a0 = *(const bf16x8*)(pa + k);
    const bf16x8 b0 = *(const bf16x8*)(pb + k);
    bf16x8 a1, b1;
    const bool two = k + 16 < k1;
    if (two) {
      a1 = *(const bf16x8*)(pa + k + 16);
      b1 = *(const bf16x8*)(pb + k + 16);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
    if (two) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
  }
  float* c = C + (int64_t)blockIdx.z * M * ldc;
#pragma unroll
  for (int r = 0; r < 16; ++r) c[(int64_t)(row0 + 8 * (r / 4) + (l / 32) * 4 + r % 4) * ldc + col0 + l % 32] = acc[r];
}

__global__ void k_reduce_splits(const float* __restrict__ part, int splits, int64_t n, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += part[z * n + i];  // fixed order
  out[i] = s;
}

// ---------------------------------------------------------------- activations
// erf by Abramowitz & Stegun, Handbook of Mathematical Functions, 7.1.28:
// erf(t) = 1 - (1 + a1 t + ... + a6 t^6)^-16 for t >= 0, |error| <= 3e-7 (about 2e-6 in
// float arithmetic), odd in t.  Branch-free, so a wave evaluates it without divergence;
// the activations it feeds are rounded to bf16 (relative 4e-3).
__device__ __forceinline__ float erf_as(float x) {
  const float t = fabsf(x);
  const float p = 1.f + t * (0.0705230784f + t * (0.0422820123f + t * (0.0092705272f +
                  t * (0.0001520143f + t * (0.0002765672f + t * 0.0000430638f)))));
  const float p2 = p * p, p4 = p2 * p2, p8 = p4 * p4;
  return copysignf(1.f - __builtin_amdgcn_rcpf(p8 * p8), x);
}
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erf_as(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.f + erf_as(x * 0.70710678118654752f)) + x * 0.39894228040143268f * __expf(-0.5f * x * x);
}

// Z + bias, optionally GELU; bf16 to H [M][ldh] and H^T [N][ldht] (either may be null),
// fp32 to out [M][ldo] (may be null)
__global__ void k_bias_act(const float* __restrict__ Z, int ldz, const float* __restrict__ bias, int M, int N,
                           int act, bf16_t* __restrict__ H, int ldh, bf16_t* __restrict__ HT, int ldht,
                           float* __restrict__ out, int ldo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * N) return;
  const int m = (int)(i / N), n = (int)(i % N);
  float v = Z[(int64_t)m * ldz + n] + bias[n];
  if (act) v = gelu(v);
  if (H) H[(int64_t)m * ldh + n] = to_bf16(v);
  if (HT) HT[(int64_t)n * ldht + m] = to_bf16(v);
  if (out) out[(int64_t)m * ldo + n] = v;
}

// dZ = dH * gelu'(Z + bias), to bf16 dZ [M][ldd] and dZ^T [N][lddt]
__global__ void k_act_backward(const float* __restrict__ dH, int ldh, const float* __restrict__ Z, int ldz,
                               const float* __restrict__ bias, int M, int N, bf16_t* __restrict__ dZ, int ldd,
                               bf16_t* __restrict__ dZT, int lddt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * N) return;
  const int m = (int)(i / N), n = (int)(i % N);
  const float g = dH[(int64_t)m * ldh + n] * gelu_grad(Z[(int64_t)m * ldz + n] + bias[n]);
  dZ[(int64_t)m * ldd + n] = to_bf16(g);
  dZT[(int64_t)n * lddt + m] = to_bf16(g);
}

// mean squared error over the first nused of N columns: loss partials per block and
// dOut = 2 (out - target) / (M * nused) (0 in padding columns), bf16 both orientations
__global__ void k_mse(const float* __restrict__ out, int ldo, const float* __restrict__ tgt, int ldt, int M, int N,
                      int nused, bf16_t* __restrict__ dZ, int ldd, bf16_t* __restrict__ dZT, int lddt,
                      float* __restrict__ loss_part) {
  __shared__ float sh[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (i < (int64_t)M * N) {
    const int m = (int)(i / N), n = (int)(i % N);
    float g = 0.f;
    if (n < nused) {
      const float d = out[(int64_t)m * ldo + n] - tgt[(int64_t)m * ldt + n];
      e2 = d * d;
      g = 2.f * d / ((float)M * nused);
    }
    dZ[(int64_t)m * ldd + n] = to_bf16(g);
    dZT[(int64_t)n * lddt + m] = to_bf16(g);
  }
  sh[threadIdx.x] = e2;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = sh[0];
}

// db[n] = sum over m of dZT[n][m] (one block per n, fixed order)
__global__ void k_rowsum_bf16(const bf16_t* __restrict__ X, int ld, int M, float* __restrict__ out) {
  __shared__ float sh[256];
  const bf16_t* x = X + (int64_t)blockIdx.x * ld;
  float s = 0.f;
  for (int m = threadIdx.x; m < M; m += blockDim.x) s += from_bf16(x[m]);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// Adam (Kingma & Ba 2015) on fp32 master parameters, bias-corrected
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// fp32 W [out][in] (row stride ldw) -> bf16 W [out][ldb] and W^T [in][ldt]
__global__ void k_pack_w(const float* __restrict__ W, int out, int in, int ldw, bf16_t* __restrict__ Wb, int ldb,
                         bf16_t* __restrict__ WT, int ldt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)out * in) return;
  const int o = (int)(i / in), k = (int)(i % in);
  const bf16_t b = to_bf16(W[(int64_t)o * ldw + k]);
  Wb[(int64_t)o * ldb + k] = b;
  WT[(int64_t)k * ldt + o] = b;
}

// fp32 W [out][in] -> bf16 [out][in] for the fused inference kernel.  perm = 1 orders
// each aligned 32-column block as a 16x16x32 MFMA result fragment hands the previous
// layer's activations over: position 8g + j holds input 16(j/4) + 4g + j%4.
__global__ void k_pack_infer(const float* __restrict__ W, int out, int in, int perm, bf16_t* __restrict__ Wp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)out * in) return;
  const int o = (int)(i / in), k = (int)(i % in);
  int src = k;
  if (perm) {
    const int q = k % 32, g = q / 8, j = q % 8;
    src = k - q + 16 * (j / 4) + 4 * g + j % 4;
  }
  Wp[i] = to_bf16(W[(int64_t)o * in + src]);
}

__global__ void k_cast_bf16(const float* __restrict__ x, int64_t n, bf16_t* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = to_bf16(x[i]);
}

int64_t blocks(int64_t n, int bs) { return (n + bs - 1) / bs; }

}  // namespace

extern "C" {

const char* um_last_error(void) { return g_err.c_str(); }

// C[M,N] = A[M,K] B[N,K]^T (bf16 in, fp32 out).  M, N multiples of 64, K of 16, lda and
// ldb multiples of 8.  splits > 1: K in that many chunks (multiples of 32) reduced in a
// fixed order through `partial` (splits*M*ldc floats).
int um_gemm_nt(const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C, int ldc, int M, int N, int K,
               int splits, float* partial, void* stream) {
  if (M % 64 || N % 64 || K % 16 || lda % 8 || ldb % 8 || ldc < N || splits < 1 || (splits > 1 && !partial))
    return fail(-1, "um_gemm_nt: M, N must be multiples of 64, K of 16, lda/ldb of 8");
  hipStream_t s = (hipStream_t)stream;
  int kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + 31) / 32 * 32;
  const int z = (K + kchunk - 1) / kchunk;
  float* dst = z > 1 ? partial : C;
  k_gemm_nt<<<dim3(N / 64, M / 64, z), 256, 0, s>>>((const bf16_t*)A, lda, (const bf16_t*)B, ldb, dst, ldc, M, K,
                                                     kchunk);
  HIPCHK(hipGetLastError());
  if (z > 1) {
    const int64_t n = (int64_t)M * ldc;
    k_reduce_splits<<<blocks(n, 256), 256, 0, s>>>(partial, z, n, C);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

int um_bias_act(const float* Z, int ldz, const float* bias, int M, int N, int act, uint16_t* H, int ldh,
                uint16_t* HT, int ldht, float* out, int ldo, void* stream) {
  k_bias_act<<<blocks((int64_t)M * N, 256), 256, 0, (hipStream_t)stream>>>(Z, ldz, bias, M, N, act, (bf16_t*)H, ldh,
                                                                          (bf16_t*)HT, ldht, out, ldo);
  HIPCHK(hipGetLastError());
  return 0;
}

int um_act_backward(const float* dH, int ldh, const float* Z, int ldz, const float* bias, int M, int N,
                    uint16_t* dZ, int ldd, uint16_t* dZT, int lddt, void* stream) {
  k_act_backward<<<blocks((int64_t)M * N, 256), 256, 0, (hipStream_t)stream>>>(dH, ldh, Z, ldz, bias, M, N,
                                                                              (bf16_t*)dZ, ldd, (bf16_t*)dZT, lddt);
  HIPCHK(hipGetLastError());
  return 0;
}

// returns the number of loss partials written (one per block of 256)
int um_mse(const float* out, int ldo, const float* tgt, int ldt, int M, int N, int nused, uint16_t* dZ, int ldd,
           uint16_t* dZT, int lddt, float* loss_part, void* stream) {
  const int64_t nb = blocks((int64_t)M * N, 256);
  k_mse<<<nb, 256, 0, (hipStream_t)stream>>>(out, ldo, tgt, ldt, M, N, nused, (bf16_t*)dZ, ldd, (bf16_t*)dZT, lddt,
                                             loss_part);
  HIPCHK(hipGetLastError());
  return (int)nb;
}

int um_rowsum_bf16(const uint16_t* X, int ld, int rows, int M, float* out, void* stream) {
  k_rowsum_bf16<<<rows, 256, 0, (hipStream_t)stream>>>((const bf16_t*)X, ld, M, out);
  HIPCHK(hipGetLastError());
  return 0;
}

int um_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
            int t, void* stream) {
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  k_adam<<<blocks(n, 256), 256, 0, (hipStream_t)stream>>>(p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  HIPCHK(hipGetLastError());
  return 0;
}

int um_pack_w(const float* W, int out, int in, int ldw, uint16_t* Wb, int ldb, uint16_t* WT, int ldt, void* stream) {
  k_pack_w<<<blocks((int64_t)out * in, 256), 256, 0, (hipStream_t)stream>>>(W, out, in, ldw, (bf16_t*)Wb, ldb,
                                                                           (bf16_t*)WT, ldt);
  HIPCHK(hipGetLastError());
  return 0;
}

int um_pack_infer(const float* W, int out, int in, int perm, uint16_t* Wp, void* stream) {
  if (in % 32) return fail(-1, "um_pack_infer: in must be a multiple of 32");
  k_pack_infer<<<blocks((int64_t)out * in, 256), 256, 0, (hipStream_t)stream>>>(W, out, in, perm, (bf16_t*)Wp);
  HIPCHK(hipGetLastError());
  return 0;
}

int um_cast_bf16(const float* x, int64_t n, uint16_t* y, void* stream) {
  k_cast_bf16<<<blocks(n, 256), 256, 0, (hipStream_t)stream>>>(x, n, (bf16_t*)y);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // extern "C"
