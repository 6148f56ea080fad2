/* cqgpu_guess.h -- kernels of the learned initial guess (a BF16 multilayer perceptron that
 * predicts atom-pair density blocks).  No reference counterpart: the reference's guesses are
 * SAD, CORE, RANDOM and READ (include/singleslater/singleslater_guess.h); SAD stays the
 * fallback.  Row-major matrices; bf16 stored as uint16_t; every call enqueues on `stream`
 * (a hipStream_t, NULL = the default stream) and returns 0, or a negative code with the
 * message in cq_mlp_last_error(). */
#ifndef CQGPU_GUESS_H
#define CQGPU_GUESS_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

const char* cq_mlp_last_error(void);

/* C[M x N] (FP32, leading dimension ldc) = op(A) op(B) (+ bias[N]) (ReLU) (+ C if
 * accumulate); op(A) = A (M x K, lda) or A^T (A stored K x M); op(B) = B (K x N, ldb) or B^T
 * (B stored N x K).  BF16 inputs, FP32 accumulation on the matrix cores; Cbf (optional)
 * receives the bf16 rounding of C. */
int cq_mlp_gemm_bf16(const uint16_t* A, int lda, int transA, const uint16_t* B, int ldb, int transB, float* C,
                     int ldc, uint16_t* Cbf, int M, int N, int K, const float* bias, int relu, int accumulate,
                     void* stream);
/* dZ = dH * (Z > 0), dZbf (optional) its bf16 copy. */
int cq_mlp_relu_backward(const float* dH, const float* Z, float* dZ, uint16_t* dZbf, long n, void* stream);
/* db[j] = sum_i dZ[i][j]. */
int cq_mlp_column_sum(const float* dZ, int M, int N, float* db, void* stream);
/* y = bf16(x). */
int cq_mlp_to_bf16(const float* x, uint16_t* y, long n, void* stream);
/* masked squared error: loss += sum m_ij w_i (y_ij - t_ij)^2 / norm, dY = 2 m w (y - t) / norm
 * (w optional; loss must be zeroed by the caller). */
int cq_mlp_mse(const float* Y, const float* T, const float* mask, const float* w, int M, int N, float norm,
               float* dY, float* loss, void* stream);
/* Adam step `step` (1-based) on FP32 weights p with moments m, v; pbf (optional) receives
 * bf16(p). */
int cq_mlp_adam(float* p, const float* g, float* m, float* v, uint16_t* pbf, long n, float lr, float b1, float b2,
                float eps, int step, float weight_decay, void* stream);

#ifdef __cplusplus
}
#endif
#endif
