// Kernel wrappers shared by the generated per-class translation units and the host code.
#pragma once
#include "gdirect_device.h"

#define WG 256

// ---------------------------------------------------------------------------------------
// ERI kernel wrappers, instantiated once per class pair by the generated units
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void make_quartet(const uint2* __restrict__ tasks, long t, long count, long wbase,
                                             const PairInfo* __restrict__ pairs,
                                             const PrimPair* __restrict__ prims, QuartetArgs& q) {
  const bool active = t < count;
  const uint2 tk = tasks[active ? t : wbase];
  const PairInfo B = pairs[tk.x];
  const PairInfo K = pairs[tk.y];
  q.prim = prims;
  q.bra_p0 = B.p0;
  q.bra_p1 = active ? B.p1 : B.p0;
  q.ket_p0 = K.p0;
  q.ket_p1 = K.p1;
  q.a0 = B.a0; q.b0 = B.b0; q.c0 = K.a0; q.d0 = K.b0;
  const double deg = (B.a == B.b ? 1.0 : 2.0) * (K.a == K.b ? 1.0 : 2.0) * (tk.x == tk.y ? 1.0 : 2.0);
  q.deg = active ? deg : 0.0;
  const unsigned int b0 = __shfl((int)tk.x, 0);
  q.bra_uniform = (__ballot(tk.x == b0) == ~0ull);
  q.bra_swap = B.a != B.s1;
  q.ket_swap = K.a != K.s1;
}

#define DEFINE_CLASS_KERNELS(I, J, BI, KI, NAME, LTOT)                                          \
  template <int NSLOT>                                                                          \
  __global__ void __launch_bounds__(WG) eri_jk_##NAME(                                          \
      const uint2* __restrict__ tasks, const unsigned long long* __restrict__ count_ptr,        \
      const PairInfo* __restrict__ pairs, const PrimPair* __restrict__ prims,                   \
      const double* __restrict__ boys, JKArgs jk) {                                             \
    __shared__ double boys_lds[BOYS_NGRID * BOYS_LDS_W];                                        \
    const long count = (long)*count_ptr;                                                        \
    if ((long)blockIdx.x * WG >= count) return;                                                 \
    stage_boys(boys, LTOT, boys_lds);                                                           \
    for (long base = (long)blockIdx.x * WG; base < count; base += (long)gridDim.x * WG) {      \
      const long wbase = base + (threadIdx.x & ~63);                                            \
      if (wbase >= count) break;                                                                \
      QuartetArgs q;                                                                            \
      make_quartet(tasks, base + threadIdx.x, count, wbase, pairs, prims, q);                   \
      quartet_jk_##NAME<NSLOT>(q, jk, boys_lds);                                                \
    }                                                                                           \
  }                                                                                             \
  __global__ void __launch_bounds__(WG) eri_block_##NAME(                                       \
      const uint2* __restrict__ tasks, long count, const PairInfo* __restrict__ pairs,          \
      const PrimPair* __restrict__ prims, const double* __restrict__ boys, double* out,          \
      int blk) {                                                                                \
    __shared__ double boys_lds[BOYS_NGRID * BOYS_LDS_W];                                        \
    if ((long)blockIdx.x * WG >= count) return;                                                 \
    stage_boys(boys, LTOT, boys_lds);                                                           \
    for (long base = (long)blockIdx.x * WG; base < count; base += (long)gridDim.x * WG) {      \
      const long wbase = base + (threadIdx.x & ~63);                                            \
      if (wbase >= count) break;                                                                \
      const long t = base + threadIdx.x;                                                        \
      QuartetArgs q;                                                                            \
      make_quartet(tasks, t, count, wbase, pairs, prims, q);                                    \
      if (t < count) quartet_block_##NAME(q, out + t * (long)blk, boys_lds);                   \
    }                                                                                           \
  }

typedef void (*JKKernel)(const uint2*, const unsigned long long*, const PairInfo*, const PrimPair*,
                         const double*, JKArgs);
typedef void (*BlockKernel)(const uint2*, long, const PairInfo*, const PrimPair*, const double*, double*, int);

struct ClassKernels {
  int i, j, bi, ki, L;
  const char* name;
  JKKernel jk1, jk2;
  BlockKernel block;
};


