// Kernels of the learned initial guess: a multilayer perceptron with BF16 weights and
// activations, FP32 accumulation on the matrix cores (v_mfma_f32_16x16x16_bf16), FP32 master
// weights and Adam moments.  No reference counterpart: the reference's guesses are SAD, CORE,
// RANDOM and READ (include/singleslater/singleslater_guess.h).
//
// Matrices are row-major.  bf16 values are stored as uint16_t (the upper half of an IEEE
// float, rounded to nearest even).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <stdexcept>
#include <string>

#include "../../include/cqgpu_guess.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);   // inf / nan kept
  u += 0x7fffu + ((u >> 16) & 1u);                                     // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// C[M x N] = op(A)[M x K] op(B)[K x N] (+ bias[N]) (ReLU) with FP32 accumulation.
// op(A) = A (A stored M x K, lda) or A^T (A stored K x M); op(B) = B (K x N, ldb) or B^T
// (B stored N x K).  Each wavefront computes one 16 x 16 tile with the 16x16x16 BF16 MFMA:
// lane l holds A[i][k] with i = l % 16, k = 4 (l / 16) + v, B[k][j] with j = l % 16,
// k = 4 (l / 16) + v, and C[i][j] with j = l % 16, i = 4 (l / 16) + v (v = 0..3).
// Out-of-range rows, columns and k are read as zero and not written.
__global__ void __launch_bounds__(256) gemm_bf16_kernel(const uint16_t* __restrict__ A, int lda, int transA,
                                                        const uint16_t* __restrict__ B, int ldb, int transB,
                                                        float* __restrict__ C, int ldc, uint16_t* __restrict__ Cbf,
                                                        int M, int N, int K, const float* __restrict__ bias,
                                                        int relu, int accumulate) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ti = blockIdx.y * 2 + (wave >> 1), tj = blockIdx.x * 2 + (wave & 1);
  const int i0 = ti * 16, j0 = tj * 16;
  if (i0 >= M || j0 >= N) return;
  const int r = lane & 15, q = lane >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = k0 + 4 * q + v;
      const int ai = i0 + r, bj = j0 + r;
      uint16_t av = 0, bv = 0;
      if (k < K && ai < M) av = transA ? A[(long)k * lda + ai] : A[(long)ai * lda + k];
      if (k < K && bj < N) bv = transB ? B[(long)bj * ldb + k] : B[(long)k * ldb + bj];
      a[v] = (short)av;
      b[v] = (short)bv;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  const int j = j0 + r;
  if (j >= N) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = i0 + 4 * q + v;
    if (i >= M) continue;
    float x = acc[v];
    if (accumulate) x += C[(long)i * ldc + j];
    if (bias) x += bias[j];
    if (relu) x = fmaxf(x, 0.f);
    C[(long)i * ldc + j] = x;
    if (Cbf) Cbf[(long)i * ldc + j] = f2bf(x);
  }
}

// dZ = dH * (Z > 0); also its bf16 copy
__global__ void relu_backward_kernel(const float* __restrict__ dH, const float* __restrict__ Z, float* __restrict__ dZ,
                                     uint16_t* __restrict__ dZbf, long n) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float g = Z[k] > 0.f ? dH[k] : 0.f;
  dZ[k] = g;
  if (dZbf) dZbf[k] = f2bf(g);
}

// db[j] = sum_i dZ[i][j]
__global__ void column_sum_kernel(const float* __restrict__ dZ, int M, int N, float* __restrict__ db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (int i = 0; i < M; ++i) s += dZ[(long)i * N + j];
  db[j] = s;
}

__global__ void to_bf16_kernel(const float* __restrict__ x, uint16_t* __restrict__ y, long n) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[k] = f2bf(x[k]);
}

// masked mean squared error: loss_i = sum_j m_ij (y_ij - t_ij)^2 w_i; dY = 2 m (y - t) w / norm
__global__ void mse_kernel(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ mask,
                           const float* __restrict__ w, int M, int N, float norm, float* __restrict__ dY,
                           float* __restrict__ loss) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;   // every lane reaches the wavefront reduction
  if (k < (long)M * N) {
    const int i = (int)(k / N);
    const float m = mask[k] * (w ? w[i] : 1.f);
    const float d = Y[k] - T[k];
    dY[k] = 2.f * m * d / norm;
    l = m * d * d / norm;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

// Adam on FP32 master weights; refreshes the bf16 copy
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, uint16_t* __restrict__ pbf, long n, float lr, float b1,
                            float b2, float eps, float c1, float c2, float wd) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = g[k] + wd * p[k];
  const float mk = b1 * m[k] + (1.f - b1) * gk;
  const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  const float upd = (mk / c1) / (sqrtf(vk / c2) + eps);
  const float pk = p[k] - lr * upd;
  p[k] = pk;
  if (pbf) pbf[k] = f2bf(pk);
}

thread_local std::string g_err;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
    g_err.clear();
    return 0;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return -1;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -3;
  }
}

inline int blocks(long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

extern "C" {

const char* cq_mlp_last_error(void) { return g_err.c_str(); }

int cq_mlp_gemm_bf16(const uint16_t* A, int lda, int transA, const uint16_t* B, int ldb, int transB, float* C,
                     int ldc, uint16_t* Cbf, int M, int N, int K, const float* bias, int relu, int accumulate,
                     void* stream) {
  return guarded([&] {
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) throw std::invalid_argument("gemm: bad sizes or pointers");
    if (lda < (transA ? M : K) || ldb < (transB ? K : N) || ldc < N)
      throw std::invalid_argument("gemm: leading dimension too small");
    const dim3 grid((N + 31) / 32, (M + 31) / 32);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, (hipStream_t)stream, A, lda, transA, B, ldb, transB,
                       C, ldc, Cbf, M, N, K, bias, relu, accumulate);
  });
}

int cq_mlp_relu_backward(const float* dH, const float* Z, float* dZ, uint16_t* dZbf, long n, void* stream) {
  return guarded([&] {
    if (n <= 0 || !dH || !Z || !dZ) throw std::invalid_argument("relu_backward: bad arguments");
    hipLaunchKernelGGL(relu_backward_kernel, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, dH, Z, dZ,
                       dZbf, n);
  });
}

int cq_mlp_column_sum(const float* dZ, int M, int N, float* db, void* stream) {
  return guarded([&] {
    if (M <= 0 || N <= 0 || !dZ || !db) throw std::invalid_argument("column_sum: bad arguments");
    hipLaunchKernelGGL(column_sum_kernel, dim3(blocks(N, 256)), dim3(256), 0, (hipStream_t)stream, dZ, M, N, db);
  });
}

int cq_mlp_to_bf16(const float* x, uint16_t* y, long n, void* stream) {
  return guarded([&] {
    if (n <= 0 || !x || !y) throw std::invalid_argument("to_bf16: bad arguments");
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, x, y, n);
  });
}

int cq_mlp_mse(const float* Y, const float* T, const float* mask, const float* w, int M, int N, float norm,
               float* dY, float* loss, void* stream) {
  return guarded([&] {
    if (M <= 0 || N <= 0 || !Y || !T || !mask || !dY || !loss || !(norm > 0.f))
      throw std::invalid_argument("mse: bad arguments");
    hipLaunchKernelGGL(mse_kernel, dim3(blocks((long)M * N, 256)), dim3(256), 0, (hipStream_t)stream, Y, T, mask,
                       w, M, N, norm, dY, loss);
  });
}

int cq_mlp_adam(float* p, const float* g, float* m, float* v, uint16_t* pbf, long n, float lr, float b1, float b2,
                float eps, int step, float weight_decay, void* stream) {
  return guarded([&] {
    if (n <= 0 || !p || !g || !m || !v || step < 1) throw std::invalid_argument("adam: bad arguments");
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, p, g, m, v, pbf, n,
                       lr, b1, b2, eps, c1, c2, weight_decay);
  });
}

}  // extern "C"
