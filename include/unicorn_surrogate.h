/* Learned surrogate of the chemistry step on MI355X (gfx950): libunicorn_surrogate.so.
 *
 * No counterpart in the reference: the network, its training and its gate are this
 * library's own, and the FP64 step of libunicorn_chem.so (unicorn_chem.h) is both the
 * source of its training pairs and its fallback.  The host layer that drives these
 * entry points is unicorn_amd/surrogate.py.
 *
 * Conventions.  All pointers are device pointers unless stated; `stream` is a
 * hipStream_t.  bf16 values are passed as their 16-bit patterns (uint16_t).  Field
 * arrays use the reference layout of unicorn_chem.h (cell c, species [52][n]); `cells`
 * lists the cells to work on (the active cells, ISKIP < 3).  Every function returns 0,
 * or a negative code with a message in um_last_error().  `norm` and `thermo` are the
 * parameter blocks whose layouts unicorn_amd/surrogate.py defines (Norm, Thermo);
 * us_sizes() returns their sizes so that the host can check them.
 *
 * Network: 54 -> 256 -> 256 -> 256 -> 256 -> 53, GELU, bf16 weights and activations,
 * fp32 accumulation, padded to 64 inputs and 64 outputs.
 */
#ifndef UNICORN_SURROGATE_H
#define UNICORN_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* um_last_error(void);

/* ---- MLP building blocks (training) */
/* C[M,N] = A[M,K] B[N,K]^T, bf16 in, fp32 out (v_mfma_f32_32x32x16_bf16).  M, N multiples
 * of 64, K of 16; splits > 1 splits K, reduced in a fixed order through `partial`. */
int um_gemm_nt(const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C, int ldc, int M, int N, int K,
               int splits, float* partial, void* stream);
/* Z + bias, then GELU when act; bf16 to H and/or H^T, fp32 to out (each may be null). */
int um_bias_act(const float* Z, int ldz, const float* bias, int M, int N, int act, uint16_t* H, int ldh,
                uint16_t* HT, int ldht, float* out, int ldo, void* stream);
/* dZ = dH * GELU'(Z + bias), bf16 both orientations. */
int um_act_backward(const float* dH, int ldh, const float* Z, int ldz, const float* bias, int M, int N,
                    uint16_t* dZ, int ldd, uint16_t* dZT, int lddt, void* stream);
/* Mean squared error over the first nused columns: loss partials and dOut; returns the
 * number of partials. */
int um_mse(const float* out, int ldo, const float* tgt, int ldt, int M, int N, int nused, uint16_t* dZ, int ldd,
           uint16_t* dZT, int lddt, float* loss_part, void* stream);
int um_rowsum_bf16(const uint16_t* X, int ld, int rows, int M, float* out, void* stream);
/* Adam (bias-corrected, step t) on fp32 master parameters. */
int um_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
            int t, void* stream);
/* fp32 W [out][in] -> bf16 W and W^T. */
int um_pack_w(const float* W, int out, int in, int ldw, uint16_t* Wb, int ldb, uint16_t* WT, int ldt, void* stream);
/* fp32 W [out][in] -> bf16 in the fused inference kernel's column order (perm = 1: each
 * 32-column block in 16x16x32 result-fragment order; 0: as is). */
int um_pack_infer(const float* W, int out, int in, int perm, uint16_t* Wp, void* stream);
int um_cast_bf16(const float* x, int64_t n, uint16_t* y, void* stream);

/* ---- training data */
/* Raw features (T, ln dt, Y) and targets (dT, change of Box-Cox(Y)) of one FP64 step of
 * the listed cells, rows row0 .. row0+ncells-1 of xraw [.][54] and traw [.][53]. */
int us_collect(const double* y0, const double* tk0, const double* y1, const double* tk1, int64_t n,
               const int32_t* cells, int32_t ncells, double tstr, double log_dt, float* xraw, float* traw,
               int64_t row0, void* stream);
/* Per-column partial sums, sums of squares, min and max of the transformed features
 * (kind 0) or the compressed targets asinh(d / tscale) (kind 1); returns the chunk count. */
int us_colstats(const float* raw, int64_t rows, int ncol, int kind, const float* tscale, int64_t chunk, double* part,
                void* stream);
/* A training batch from rows idx[0..B): bf16 X, X^T and fp32 targets. */
int us_batch(const float* xraw, const float* traw, const int64_t* idx, int B, const void* norm, uint16_t* X,
             uint16_t* XT, float* T, void* stream);

/* ---- inference */
/* Features of the listed cells as bf16 X [bpad][64] (the separate-kernel network path). */
int us_features(const double* y, const double* tk, int64_t n, const int32_t* cells, int32_t ncells, int32_t bpad,
                double tstr, double log_dt, const void* norm, uint16_t* X, void* stream);
/* The gate on network outputs out [ncells][64]: accepted cells get Y, TK and HT;
 * reject[c] = 1 where rejected, 0 where accepted; stats [8]: accepted, out of range,
 * residual, sum, active, and the largest raw C/H/O residuals (bits of doubles).  The
 * tolerances: C/H/O residual, |sum Y + STMF - 1|, |dT| (K), |dY|, |change of Box-Cox(Y)|. */
int us_gate(const float* out, const void* norm, const void* thermo, double* y, double* tk, double* ht,
            const double* stmf, const double* stnd, int64_t n, const int32_t* cells, int32_t ncells, double log_dt,
            double tol_el, double tol_sum, double tol_dt, double tol_dy, double tol_bc, int32_t* reject,
            unsigned long long* stats, void* stream);
/* The surrogate's prediction on a field: the network on the listed cells in one kernel
 * (w, b: five pointers each, packed by um_pack_infer, and the fp32 biases; outputs to
 * out [64][ldo], Y^0.1 to pbuf [52][ldo], ldo >= ncells), then the gate, writing
 * mask[c] = 10 where accepted and iskip[c] where rejected.  stats as us_gate. */
int us_infer(const uint16_t* const* w, const float* const* b, const void* norm, const void* thermo, double* y,
             double* tk, double* ht, const double* stmf, const double* stnd, const int32_t* iskip, int32_t* mask,
             float* out, double* pbuf, int64_t ldo, int64_t n, const int32_t* cells, int32_t ncells, double tstr,
             double log_dt, double tol_el, double tol_sum, double tol_dt, double tol_dy, double tol_bc,
             unsigned long long* stats, void* stream);
/* Density of every cell from (Y, T) and the grid's edges (main:1056-1068), for a step in
 * which the gate accepted every cell. */
int us_density(const void* thermo, const double* y, const double* tk, double* rho, int64_t li, int64_t lj,
               void* stream);
/* out[0] = sizeof(Norm), out[1] = sizeof(Thermo) (host pointer). */
int us_sizes(int64_t* out);

#ifdef __cplusplus
}
#endif

#endif
