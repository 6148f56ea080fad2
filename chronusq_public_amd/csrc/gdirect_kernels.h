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
  const unsigned bx = pair_of(tk.x), ky = pair_of(tk.y);
  const PairInfo B = pairs[bx];
  const PairInfo K = pairs[ky];
  int bp0, bp1, kp0, kp1;
  prim_range(B, block_of(tk.x), bp0, bp1);
  prim_range(K, block_of(tk.y), kp0, kp1);
  q.prim = prims;
  q.bra_p0 = bp0;
  q.bra_p1 = active ? bp1 : bp0;
  q.ket_p0 = kp0;
  q.ket_p1 = kp1;
  q.a0 = B.a0; q.b0 = B.b0; q.c0 = K.a0; q.d0 = K.b0;
  const double deg = (B.a == B.b ? 1.0 : 2.0) * (K.a == K.b ? 1.0 : 2.0) * (bx == ky ? 1.0 : 2.0);
  q.deg = active ? deg : 0.0;
  const unsigned int k0 = __shfl((int)tk.y, 0);
  q.ket_uniform = (__ballot(tk.y == k0) == ~0ull);
  q.bra_swap = B.a != B.s1;
  q.ket_swap = K.a != K.s1;
  q.loff = K.loff;
}

// Quartet of one lane's bra (pair B, task code bra_code) with the wavefront-uniform ket pair
// kk and its primitive block kblk (both from readfirstlane, so the ket data is loaded through
// the scalar path); lanes that do not own a task of this ket get an empty bra range and
// degeneracy 0.
__device__ __forceinline__ void make_quartet_uniform_ket(const PairInfo& B, unsigned bra_code, unsigned kk,
                                                         unsigned kblk, bool mine,
                                                         const PairInfo* __restrict__ pairs,
                                                         const PrimPair* __restrict__ prims, QuartetArgs& q) {
  const PairInfo K = pairs[kk];
  int bp0, bp1, kp0, kp1;
  prim_range(B, block_of(bra_code), bp0, bp1);
  prim_range<UNI_BLOCK>(K, kblk, kp0, kp1);
  q.prim = prims;
  q.bra_p0 = bp0;
  q.bra_p1 = mine ? bp1 : bp0;
  q.ket_p0 = kp0;
  q.ket_p1 = kp1;
  q.a0 = B.a0; q.b0 = B.b0; q.c0 = K.a0; q.d0 = K.b0;
  const double deg = (B.a == B.b ? 1.0 : 2.0) * (K.a == K.b ? 1.0 : 2.0) * (pair_of(bra_code) == kk ? 1.0 : 2.0);
  q.deg = mine ? deg : 0.0;
  q.ket_uniform = true;
  q.bra_swap = B.a != B.s1;
  q.ket_swap = K.a != K.s1;
  q.loff = K.loff;   // the uniform ket's molecule (every quartet of a segment lies in it)
}

// Quartet of the wavefront-uniform bra pair bb and its primitive block bblk (from
// readfirstlane) with one lane's ket (pair K, task code ket_code); lanes that do not own a task
// get an empty ket range and degeneracy 0.
__device__ __forceinline__ void make_quartet_uniform_bra(const PairInfo& K, unsigned ket_code, unsigned bb,
                                                         unsigned bblk, bool mine,
                                                         const PairInfo* __restrict__ pairs,
                                                         const PrimPair* __restrict__ prims, QuartetArgs& q) {
  const PairInfo B = pairs[bb];
  int bp0, bp1, kp0, kp1;
  prim_range<UNI_BLOCK>(B, bblk, bp0, bp1);
  prim_range(K, block_of(ket_code), kp0, kp1);
  q.prim = prims;
  q.bra_p0 = bp0;
  q.bra_p1 = bp1;
  q.ket_p0 = kp0;
  q.ket_p1 = mine ? kp1 : kp0;
  q.a0 = B.a0; q.b0 = B.b0; q.c0 = K.a0; q.d0 = K.b0;
  const double deg = (B.a == B.b ? 1.0 : 2.0) * (K.a == K.b ? 1.0 : 2.0) * (bb == pair_of(ket_code) ? 1.0 : 2.0);
  q.deg = mine ? deg : 0.0;
  q.ket_uniform = false;
  q.bra_swap = B.a != B.s1;
  q.ket_swap = K.a != K.s1;
  q.loff = B.loff;   // the uniform bra's molecule
}

// JK kernel: work items are (chunk, 64-task block) pairs, one chunk per wavefront, so classes
// whose ket components are split into independent chunks run them in parallel.
#define DEFINE_CLASS_KERNELS(I, J, BI, KI, NAME, LTOT)                                          \
  template <int NSLOT>                                                                          \
  __global__ void __launch_bounds__(WG) eri_jk_##NAME(                                          \
      const uint2* __restrict__ tasks, const unsigned long long* __restrict__ count_ptr,        \
      const PairInfo* __restrict__ pairs, const PrimPair* __restrict__ prims,                   \
      const double* __restrict__ boys, JKArgs jk) {                                             \
    constexpr int NCH = kChunks_##NAME;                                                         \
    __shared__ double boys_lds[BOYS_NGRID * BOYS_LDS_W];                                        \
    const long count = (long)*count_ptr;                                                        \
    const long nitems = ((count + 63) / 64) * 64 * NCH;                                         \
    if ((long)blockIdx.x * WG >= nitems) return;                                                \
    stage_boys(boys, LTOT, boys_lds);                                                           \
    for (long base = (long)blockIdx.x * WG; base < nitems; base += (long)gridDim.x * WG) {     \
      const long witem = (base + (threadIdx.x & ~63)) / 64;                                     \
      const int chunk = (int)(witem % NCH);                                                     \
      const long tbase = (witem / NCH) * 64;                                                    \
      if (tbase >= count) break;                                                                \
      const long t = tbase + (threadIdx.x & 63);                                                \
      const bool active = t < count;                                                            \
      const uint2 tk = tasks[active ? t : tbase];                                               \
      const PairInfo B = pairs[pair_of(tk.x)];                                                  \
      bool pending = active;                                                                    \
      /* waterfall over the distinct ket pairs of the wavefront: each pass has one uniform ket \
         (scalar loads of its primitive data), lanes of other kets contribute nothing */        \
      while (true) {                                                                            \
        const unsigned long long pm = __ballot(pending);                                        \
        if (pm == 0ull) break;                                                                  \
        const int leader = __ffsll((long long)pm) - 1;                                          \
        const unsigned kk = __builtin_amdgcn_readfirstlane(__shfl((int)tk.y, leader));          \
        const bool mine = pending && tk.y == kk;                                                \
        QuartetArgs q;                                                                          \
        make_quartet_uniform_ket(B, tk.x, kk, 0u, mine, pairs, prims, q);                       \
        quartet_jk_##NAME<NSLOT, 0>(q, jk, boys_lds, chunk, nullptr, nullptr);                  \
        pending = pending && !mine;                                                             \
      }                                                                                         \
    }                                                                                           \
  }                                                                                             \
  /* Segment kernel (one density slot, WGS threads per workgroup): each workgroup takes     \
     (segment, chunk) items from a queue; the segment's ket is uniform, K columns c/d and J(3,4) accumulate in LDS and are \
     flushed to HBM once per item. */                                                        \
  template <int WGS, typename R>                                                              \
  __global__ void __launch_bounds__(WGS) eri_jkseg_##NAME(                                     \
      const uint2* __restrict__ tasks, const Segment* __restrict__ segs,                      \
      const unsigned* __restrict__ nseg_ptr, unsigned* __restrict__ work,                     \
      const PairInfo* __restrict__ pairs, const PrimPair* __restrict__ prims,                 \
      const double* __restrict__ boys, JKArgs jk) {                                           \
    constexpr int NCH = kChunks_##NAME, NC = kNC_##NAME, ND = kND_##NAME, NCOL = NC + ND;     \
    __shared__ double boys_lds[BOYS_NGRID * BOYS_LDS_W];                                      \
    __shared__ int s_item;                                                                    \
    extern __shared__ double dyn_lds[];                                                       \
    const long nitems = (long)(*nseg_ptr) * NCH;                                              \
    if (nitems == 0) return;                                                                  \
    stage_boys(boys, LTOT, boys_lds);                                                         \
    const int N = jk.ncart, LN = jk.lds_n;                                                    \
    double* colK = dyn_lds;                                                                   \
    double* jcdL = dyn_lds + NCOL * LN;                                                       \
    for (int k = threadIdx.x; k < NCOL * LN + NC * ND; k += WGS) dyn_lds[k] = 0.0;            \
    const int wofs = threadIdx.x & ~63;                                                       \
    while (true) {                                                                            \
      __syncthreads();                                                                        \
      if (threadIdx.x == 0) s_item = (int)atomicAdd(work, 1u);                                \
      __syncthreads();                                                                        \
      const long item = s_item;                                                               \
      if (item >= nitems) break;                                                              \
      const Segment sg = segs[item / NCH];                                                    \
      const int chunk = (int)(item % NCH);                                                    \
      const unsigned kk = __builtin_amdgcn_readfirstlane((unsigned)sg.ket);                   \
      const unsigned kblk = __builtin_amdgcn_readfirstlane((unsigned)sg.blk);                 \
      for (int t0 = 0; t0 < sg.count; t0 += WGS) {                                             \
        if (t0 + wofs >= sg.count) continue;                                                  \
        const int t = t0 + threadIdx.x;                                                       \
        const bool active = t < sg.count;                                                     \
        const uint2 tk = tasks[sg.off + (active ? t : t0 + wofs)];                            \
        const PairInfo B = pairs[pair_of(tk.x)];                                              \
        QuartetArgs q;                                                                        \
        make_quartet_uniform_ket(B, tk.x, kk, kblk, active, pairs, prims, q);                 \
        quartet_jk_##NAME<1, 1, R>(q, jk, boys_lds, chunk, colK, jcdL);                       \
      }                                                                                       \
      __syncthreads();                                                                        \
      const PairInfo K = pairs[kk];                                                           \
      for (int k = threadIdx.x; k < NCOL * LN; k += WGS) {                                     \
        const double v = colK[k];                                                             \
        if (v != 0.0) {                                                                       \
          const int col = k / LN, row = k - col * LN + K.loff;                                \
          const int gcol = col < NC ? K.a0 + col : K.b0 + (col - NC);                         \
          atomicAdd(&jk.GK[0][(long)row * N + gcol], v);                                      \
          colK[k] = 0.0;                                                                      \
        }                                                                                     \
      }                                                                                       \
      if (threadIdx.x < NC * ND) {                                                            \
        const double v = jcdL[threadIdx.x];                                                   \
        if (v != 0.0) {                                                                       \
          const bool ks = K.a != K.s1;                                                        \
          const int ic = threadIdx.x / ND, idd = threadIdx.x - (threadIdx.x / ND) * ND;       \
          const int tcd = ks ? K.b0 * N + K.a0 : K.a0 * N + K.b0;                             \
          atomicAdd(&jk.GJ[0][tcd + ic * (ks ? 1 : N) + idd * (ks ? N : 1)], v);              \
          jcdL[threadIdx.x] = 0.0;                                                            \
        }                                                                                     \
      }                                                                                       \
    }                                                                                         \
  }                                                                                           \
  /* Block kernel: task t adds its (primitive-block) contribution to the contracted Cartesian \
     block of output slot slot[t] (t when slot is null), zeroed by the caller. */              \
  __global__ void __launch_bounds__(WG) eri_block_##NAME(                                       \
      const uint2* __restrict__ tasks, long count, const PairInfo* __restrict__ pairs,          \
      const PrimPair* __restrict__ prims, const double* __restrict__ boys, double* out,          \
      int blk, const int* __restrict__ slot) {                                                  \
    __shared__ double boys_lds[BOYS_NGRID * BOYS_LDS_W];                                        \
    if ((long)blockIdx.x * WG >= count) return;                                                 \
    stage_boys(boys, LTOT, boys_lds);                                                           \
    for (long base = (long)blockIdx.x * WG; base < count; base += (long)gridDim.x * WG) {      \
      const long wbase = base + (threadIdx.x & ~63);                                            \
      if (wbase >= count) break;                                                                \
      const long t = base + threadIdx.x;                                                        \
      QuartetArgs q;                                                                            \
      make_quartet(tasks, t, count, wbase, pairs, prims, q);                                    \
      if (t < count) quartet_block_##NAME(q, out + (slot ? (long)slot[t] : t) * (long)blk, boys_lds); \
    }                                                                                           \
  }

// Bra-uniform segment kernel (one density slot), for class pairs whose bra class has few
// pairs: a segment is one bra pair with the ket pairs that survive against it.  K rows a/b
// (NA + NB rows of ncart) and J(1,2) accumulate in LDS and are flushed once per item; J(3,4)
// goes to HBM per lane.
#define DEFINE_CLASS_SEGB(NAME, LTOT)                                                           \
  template <typename R>                                                                         \
  __global__ void __launch_bounds__(WG) eri_jksegb_##NAME(                                      \
      const uint2* __restrict__ tasks, const Segment* __restrict__ segs,                        \
      const unsigned* __restrict__ nseg_ptr, unsigned* __restrict__ work,                       \
      const PairInfo* __restrict__ pairs, const PrimPair* __restrict__ prims,                   \
      const double* __restrict__ boys, JKArgs jk) {                                             \
    constexpr int NCH = kChunks_##NAME, NA = kNA_##NAME, NB = kNB_##NAME, NROW = NA + NB;       \
    __shared__ double boys_lds[BOYS_NGRID * BOYS_LDS_W];                                        \
    __shared__ int s_item;                                                                      \
    extern __shared__ double dyn_lds[];                                                         \
    const long nitems = (long)(*nseg_ptr) * NCH;                                                \
    if (nitems == 0) return;                                                                    \
    stage_boys(boys, LTOT, boys_lds);                                                           \
    const int N = jk.ncart, LN = jk.lds_n;                                                      \
    double* rowK = dyn_lds;                                                                     \
    double* jabL = dyn_lds + NROW * LN;                                                         \
    for (int k = threadIdx.x; k < NROW * LN + NA * NB; k += WG) dyn_lds[k] = 0.0;              \
    const int wofs = threadIdx.x & ~63;                                                         \
    while (true) {                                                                              \
      __syncthreads();                                                                          \
      if (threadIdx.x == 0) s_item = (int)atomicAdd(work, 1u);                                  \
      __syncthreads();                                                                          \
      const long item = s_item;                                                                 \
      if (item >= nitems) break;                                                                \
      const Segment sg = segs[item / NCH];                                                      \
      const int chunk = (int)(item % NCH);                                                      \
      const unsigned bb = __builtin_amdgcn_readfirstlane((unsigned)sg.ket);                     \
      const unsigned bblk = __builtin_amdgcn_readfirstlane((unsigned)sg.blk);                   \
      for (int t0 = 0; t0 < sg.count; t0 += WG) {                                               \
        if (t0 + wofs >= sg.count) continue;                                                    \
        const int t = t0 + threadIdx.x;                                                         \
        const bool active = t < sg.count;                                                       \
        const uint2 tk = tasks[sg.off + (active ? t : t0 + wofs)];                              \
        const PairInfo K = pairs[pair_of(tk.y)];                                                \
        QuartetArgs q;                                                                          \
        make_quartet_uniform_bra(K, tk.y, bb, bblk, active, pairs, prims, q);                   \
        quartet_jk_##NAME<1, 2, R>(q, jk, boys_lds, chunk, rowK, jabL);                         \
      }                                                                                         \
      __syncthreads();                                                                          \
      const PairInfo A = pairs[bb];                                                             \
      for (int k = threadIdx.x; k < NROW * LN; k += WG) {                                       \
        const double v = rowK[k];                                                               \
        if (v != 0.0) {                                                                         \
          const int lrow = k / LN, col = k - lrow * LN + A.loff;                                \
          const int grow = lrow < NA ? A.a0 + lrow : A.b0 + (lrow - NA);                        \
          atomicAdd(&jk.GK[0][(long)grow * N + col], v);                                        \
          rowK[k] = 0.0;                                                                        \
        }                                                                                       \
      }                                                                                         \
      if (threadIdx.x < NA * NB) {                                                              \
        const double v = jabL[threadIdx.x];                                                     \
        if (v != 0.0) {                                                                         \
          const bool bs = A.a != A.s1;                                                          \
          const int ia = threadIdx.x / NB, ib = threadIdx.x - (threadIdx.x / NB) * NB;          \
          const int tab = bs ? A.b0 * N + A.a0 : A.a0 * N + A.b0;                               \
          atomicAdd(&jk.GJ[0][tab + ia * (bs ? 1 : N) + ib * (bs ? N : 1)], v);                 \
          jabL[threadIdx.x] = 0.0;                                                              \
        }                                                                                       \
      }                                                                                         \
    }                                                                                           \
  }

typedef void (*JKKernel)(const uint2*, const unsigned long long*, const PairInfo*, const PrimPair*,
                         const double*, JKArgs);
typedef void (*BlockKernel)(const uint2*, long, const PairInfo*, const PrimPair*, const double*, double*, int,
                            const int*);
typedef void (*SegKernel)(const uint2*, const Segment*, const unsigned*, unsigned*, const PairInfo*, const PrimPair*,
                          const double*, JKArgs);

struct ClassKernels {
  int i, j, bi, ki, L, nchunk;
  const char* name;
  JKKernel jk1, jk2;
  BlockKernel block;
  SegKernel jkseg;
  SegKernel jksegb;   // bra-uniform segment kernel, or null
  int nc, nd, na, nb;
  int seg_wg;         // workgroup size of jkseg
  SegKernel jkseg32, jksegb32;   // the same with FP32 integrals (mixed precision)
};


