/* swno.h -- C ABI of the spectral neural operator kernels (MI355X, gfx950).
 *
 * The learned surrogate has no counterpart in the reference's code: its
 * design document (TFNO_Full_Ring_Plan.tex:200-209, 252-254) describes a
 * Fourier neural operator that maps measurements to slowness squared. This
 * library holds the device kernels the host-side operator
 * (soundwavesimulation_amd/surrogate/operator.py) is built from: a batched
 * strided GEMM on bf16 MFMA with fp32 accumulation, which can carry every
 * contraction (truncated DFTs, per-mode complex mixing, pointwise linears and
 * all their gradients), the elementwise pieces around it, and fused kernels
 * for width 32 / hidden 128 (layer tails, projection head, the x transform,
 * weight and bias gradients in one pass) that move each activation once.
 *
 * All pointers are device pointers to float32 unless stated; every entry
 * point enqueues on `stream` (a hipStream_t) and returns SWNO_OK or an error
 * code with a message in swno_last_error(). Nothing here allocates.
 */
#ifndef SWNO_H
#define SWNO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWNO_OK 0
#define SWNO_EINVAL 1
#define SWNO_EHIP 2

/* C[b] = alpha * A[b] B[b] + beta * C[b] for b < batch, with
 *   A[b](m, k) = A[b * a_bs + m * a_rs + k * a_cs]   (M x K)
 *   B[b](k, n) = B[b * b_bs + k * b_rs + n * b_cs]   (K x N)
 *   C[b](m, n) = C[b * c_bs + m * c_rs + n * c_cs]   (M x N)
 * Operands are rounded to bf16 (nearest even) as they are staged; products
 * accumulate in fp32. k_split > 1 splits K into that many slices whose fp32
 * partials go to `workspace` (swno_gemm_workspace_floats) and are summed in a
 * fixed order, so results do not depend on scheduling. beta == 0 ignores the
 * old contents of C (which may then be uninitialised). */
typedef struct swno_gemm_args {
    int32_t M, N, K, batch, k_split;
    const float *A;
    int64_t a_rs, a_cs, a_bs;
    const float *B;
    int64_t b_rs, b_cs, b_bs;
    float *C;
    int64_t c_rs, c_cs, c_bs;
    float alpha, beta;
} swno_gemm_args;

const char *swno_last_error(void);

size_t swno_gemm_workspace_floats(const swno_gemm_args *g);

int swno_gemm(const swno_gemm_args *g, float *workspace, void *stream);

/* pre[r, c] = x[r, c] + bias[c] (bias may be NULL); y = GELU(pre) when
 * gelu != 0, else y = pre. GELU is the erf form 0.5 x (1 + erf(x / sqrt 2)).
 * y is float32, or bf16 codes (rounded to nearest even) when y_bf16 != 0.
 * pre may equal x; y may equal pre only when gelu == 0 and y is float32. */
int swno_bias_act(int64_t rows, int32_t cols, const float *x, const float *bias, float *pre,
                  void *y, int32_t gelu, int32_t y_bf16, void *stream);

/* dx[i] = dy[i] * GELU'(pre[i]) (dx may equal dy). */
int swno_gelu_backward(int64_t n, const float *pre, const float *dy, float *dx, void *stream);

/* Weight and bias gradients of a pointwise linear layer applied to `rows`
 * rows: dW (M x N, row-major) = A^T B and, when db is not NULL, db[c] =
 * sum_r B[r, c], for A (rows x M, row stride lda) and B (rows x N, row stride
 * ldb). Operands are rounded to bf16 for the products (MFMA, fp32
 * accumulation); the column sums are fp32. Both are summed in a fixed order
 * through `workspace` (swno_weight_grad_workspace_floats floats), so results
 * are bitwise reproducible. 1 <= M, N <= 128; an operand wider than 32
 * columns needs its width and row stride to be multiples of 4 and a 16-byte
 * aligned base; column sums need N (N / 4 for such an operand) to divide 256.
 * Replaces, for the operator's backward pass, a split-K GEMM plus a column
 * sum over the same B (one pass over A and B instead of two over B).
 * a_bf16 != 0: A holds bf16 codes (M a multiple of 8 above 16, lda a multiple
 * of 8, 16-byte aligned); the products are the same, since A is rounded to
 * bf16 for them anyway. */
size_t swno_weight_grad_workspace_floats(int64_t rows, int32_t M, int32_t N);
int swno_weight_grad(int64_t rows, int32_t M, int32_t N, const void *A, int64_t lda,
                     const float *B, int64_t ldb, float *dW, float *db, float *workspace,
                     int32_t a_bf16, void *stream);

/* Truncated transform along x over `slices` slices (one sample row each:
 * n pixels x ch channels, channel-last; ch = 32, n <= 128):
 * out[s] (rows_out x ch) = A X[s], with A (rows_out x n, rows_out <= 32)
 * given by element strides a_rs, a_cs (the forward transform Fx, or Gx^T in
 * the backward pass). bf16 MFMA with fp32 accumulation; replaces the batched
 * GEMM for this shape. X is float32, or bf16 codes when x_bf16 != 0 (the
 * same products: X is rounded to bf16 for them anyway). */
int swno_x_transform(int32_t slices, int32_t rows_out, int32_t n, int32_t ch, const float *A,
                     int64_t a_rs, int64_t a_cs, const void *X, float *out, int32_t x_bf16,
                     void *stream);

/* Tail of a spectral layer, fused, over `slices` slices (one sample row
 * each: n pixels x ch channels, channel-last; ch = 32, n a multiple of 16 up
 * to 128, an even number of modes up to 16):
 *   pre = Gx Z + v W + bias and vnext = GELU(pre) (gelu_act != 0) or pre,
 * with Gx (n x 2 modes) the inverse x transform, Z (slices, 2 modes, ch) the
 * spectra after the y inverse transform, v the layer input, W (ch x ch).
 * Replaces the inverse-x GEMM, the pointwise GEMM (beta = 1) and bias + GELU.
 * Gx and v must be 16-byte aligned. bf16_flags: bit 0, v holds bf16 codes
 * (rounded to bf16 for the product anyway: the same results); bit 1, vnext
 * is written as bf16 codes (16-byte aligned) -- activations stored at the
 * precision their consumers use, half the bytes. pre stays float32. */
int swno_layer_forward(int32_t slices, int32_t n, int32_t modes, int32_t ch, const float *Gx,
                       const float *Z, const void *v, const float *W, const float *bias,
                       float *pre, void *vnext, int32_t gelu_act, int32_t bf16_flags,
                       void *stream);

/* Backward counterpart: dv = dpre W^T + Fx^T dT1, with Fx (2 modes x n) the
 * forward x transform and dT1 (slices, 2 modes, ch); out = dv * GELU'(pre_prev)
 * when pre_prev is not NULL (the previous layer's dpre), else out = dv.
 * out must not alias dpre or pre_prev; dpre and W must be 16-byte aligned. */
int swno_layer_backward(int32_t slices, int32_t n, int32_t modes, int32_t ch, const float *dpre,
                        const float *W, const float *Fx, const float *dT1,
                        const float *pre_prev, float *out, void *stream);

/* Projection head, fused (ch = 32 channels, hidden = 128 units only):
 * hpre = v Q1 + bq1 (rows x hidden, kept for the backward pass) and
 * out[r] = GELU(hpre[r]) . Q2 + bq2, for v (rows x ch, 16-byte aligned),
 * Q1 (ch x hidden), Q2 (hidden). v Q1 is a bf16 MFMA product (fp32
 * accumulation); bias, GELU and the dot with Q2 are fp32. Replaces a GEMM,
 * bias + GELU and a second GEMM over the hidden units. v holds bf16 codes
 * when v_bf16 != 0 (the same results). */
int swno_head_forward(int64_t rows, int32_t ch, int32_t hidden, const void *v, const float *Q1,
                      const float *bq1, const float *Q2, const float *bq2, float *hpre,
                      float *out, int32_t v_bf16, void *stream);

/* Backward pass of swno_head_forward for d loss / d out = dout (rows):
 * dv = dh Q1^T with dh = dout Q2^T * GELU'(hpre), dQ1 = v^T dh,
 * dbq1 = sum_r dh, dQ2 = GELU(hpre)^T dout, dbq2 = sum_r dout. dv and dQ1
 * are bf16 MFMA products (fp32 accumulation); the parameter gradients are
 * summed in a fixed order through `workspace`
 * (swno_head_backward_workspace_floats floats). v holds bf16 codes when
 * v_bf16 != 0 (the same results). */
size_t swno_head_backward_workspace_floats(int64_t rows);
int swno_head_backward(int64_t rows, int32_t ch, int32_t hidden, const float *dout,
                       const float *hpre, const void *v, const float *Q1, const float *Q2,
                       float *dv, float *dQ1, float *dbq1, float *dQ2, float *dbq2,
                       float *workspace, int32_t v_bf16, void *stream);

/* out[c] = beta * out[c] + sum_r x[r, c], summed in a fixed order through
 * `workspace` (swno_col_sum_workspace_floats floats). */
size_t swno_col_sum_workspace_floats(int64_t rows, int32_t cols);
int swno_col_sum(int64_t rows, int32_t cols, const float *x, float beta, float *out,
                 float *workspace, void *stream);

/* Spectral repacks between the truncated transforms and the mode mixing, for
 * `batch` samples, `modes` retained modes per axis and `ch` channels:
 *   op 0 (combine):   P (batch, 2m, 2m, ch) -> X (batch, m, m, 2 ch)
 *        X_re = P[q][j] + P[m+q][m+j], X_im = P[q][m+j] - P[m+q][j]
 *   op 1 (combine^T): dX -> dP (the transpose of op 0)
 *   op 2 (expand):    Y (batch, m, m, 2 ch) -> U (batch, 2m, 2m, ch)
 *        U[q][j] = Y_re, U[q][m+j] = Y_im, U[m+q][j] = Y_im, U[m+q][m+j] = -Y_re
 *   op 3 (expand^T):  dU -> dY
 * where re / im are the first / second `ch` entries of the last axis. */
int swno_spectral_pack(int32_t op, int32_t batch, int32_t modes, int32_t ch, const float *in,
                       float *out, void *stream);

/* Mode-mixing weights: rm[mode] = [[re, im], [-im, re]] (2ch x 2ch) from
 * re/im (modes2, ch, ch), and the gradient back: d_re = dR11 + dR22,
 * d_im = dR12 - dR21 (accumulated into d_re / d_im when accumulate != 0). */
int swno_mix_assemble(int32_t modes2, int32_t ch, const float *re, const float *im, float *rm,
                      void *stream);
int swno_mix_assemble_backward(int32_t modes2, int32_t ch, const float *drm, float *d_re,
                               float *d_im, int32_t accumulate, void *stream);

/* Masked mean squared error: loss[0] = sum_i w_i (p_i - t_i)^2 / sum_i w_i
 * with w = mask (NULL: all ones), loss[1] = 1 / sum_i w_i (loss holds two
 * floats), grad = d loss[0] / d p. Deterministic two-pass sum through
 * `workspace` (swno_mse_workspace_floats floats). */
size_t swno_mse_workspace_floats(int64_t n);
int swno_mse(int64_t n, const float *pred, const float *target, const float *mask, float *loss,
             float *grad, float *workspace, void *stream);

/* One AdamW update of n parameters (decoupled weight decay, bias-corrected
 * moments; step >= 1). */
int swno_adamw(int64_t n, float *param, const float *grad, float *m, float *v, float lr,
               float beta1, float beta2, float eps, float weight_decay, int32_t step,
               void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SWNO_H */
