// Kernel wrappers shared by the generated per-class translation units and the host code.
#pragma once
#include "gdirect_device.h"

#define WG 256

// ---------------------------------------------------------------------------------------
// Uniform groups and their task regions
// ---------------------------------------------------------------------------------------
// A group is up to MAX_GROUP shell pairs of one class, one molecule and one primitive-block
// count that share their first shell (kernel orientation): kets (c, d_0) .. (c, d_{n-1}) of a
// ket-keyed class pair, bras (a, b_0) .. of a bra-keyed one.  Its pairs are
// uorder[first .. first + n).  The screening kernel sweeps the other class once for the whole
// group; every swept pair that survives against at least one member joins the group's union
// (in sweep order, at most the class pair's union bound per task region).  One task region
// (a run of uint2 entries in the class pair's task list) holds
//   [0]            {n, U}                      members, union size
//   [1 + m]        {count_m, offset_m}         member m's quartet list (offsets from the region)
//   [offset_m + t] {swept pair | block << 24, slot}   one entry per primitive block of every
//                  swept pair whose quartet with member m survives (slot: its union index)
// so the quartets keep every lane busy (the members' lists run as one dense list), while the J
// of a swept pair collects in its LDS slot across the members and primitive blocks and leaves
// the work item once.
#define MAX_GROUP 16
struct GroupRec {
  int first;       // into uorder
  int n;           // members
  int ublk;        // primitive blocks of the uniform side (equal for every member)
  int nslot;       // segment slots reserved for the group's first task region (ublk)
  int task_base;   // the group's task entries, from the start of its batch's task list
  int seg_base;    // the group's segment slots, from the start of its batch's segment list
  int pad0, pad1;
};
// Every group owns a fixed part of its batch's task list, sized by the plan's worst case, and
// ublk reserved segment slots for its first task region, so the screening places its output
// without atomics; further task regions of a group (a union over the bound) take slots of the
// batch's overflow pool (one atomic each).  A group without survivors writes its reserved
// slots with count 0; the kernels skip them.

// ---------------------------------------------------------------------------------------
// ERI kernel wrappers, instantiated once per class pair by the generated units
// ---------------------------------------------------------------------------------------

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
  q.ar = B.ar; q.br = B.br;   // LDS rows of the swept bra's shells
  q.cr = 0; q.dr = 0;
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
  q.ar = 0; q.br = 0;
  q.cr = K.ar; q.dr = K.br;   // LDS rows of the swept ket's shells
}

// Take the next work item (whole workgroup).  Up to eight queues, one per XCD (workgroups are
// dispatched round robin over the XCDs): queue x hands out items x, x + nq, x + 2 nq, ... to the
// workgroups with blockIdx % nq == x, so the counters do not contend across the chip and every
// queue has a consumer (nq <= gridDim).
#define WORK_QUEUES 8
__device__ __forceinline__ long next_item(unsigned* __restrict__ work, int* s_item) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned nq = gridDim.x < WORK_QUEUES ? gridDim.x : WORK_QUEUES;
    const unsigned x = blockIdx.x % nq;
    *s_item = (int)(x + nq * atomicAdd(&work[x], 1u));
  }
  __syncthreads();
  return (long)__builtin_amdgcn_readfirstlane(*s_item);
}

// Coulomb block (1,2) of a pair held in kernel orientation: address of element (i, j), i a
// component of its a shell, j of its b shell, in the canonical order (s1, s2).
__device__ __forceinline__ long jaddr(const PairInfo& P, int N, int i, int j) {
  const bool sw = P.a != P.s1;
  return sw ? (long)(P.b0 + j) * N + P.a0 + i : (long)(P.a0 + i) * N + P.b0 + j;
}
// The same address kept in one int while a swept pair's J slot fills: element (0, 0) of the
// block, bitwise negated when the pair is held swapped (the context keeps N^2 < 2^31).
__device__ __forceinline__ int jcode(const PairInfo& P, int N) {
  const bool sw = P.a != P.s1;
  const int a00 = sw ? P.b0 * N + P.a0 : P.a0 * N + P.b0;
  return sw ? ~a00 : a00;
}
__device__ __forceinline__ long jcode_addr(int code, int N, int i, int j) {
  return code < 0 ? (long)(~code) + (long)j * N + i : (long)code + (long)i * N + j;
}

// Group kernels.  Work items are (task region, group of cpi chunks) pairs from an atomic queue
// (cpi divides the class pair's chunk count); each workgroup spreads the item's (chunk, member,
// block of 64 of that member's entries) units over its wavefronts -- the member wavefront-
// uniform, one quartet per lane -- then flushes its LDS accumulators to HBM once:
//  * ket-keyed (eri_grp, DIG 0 or 1): K columns of the shared c shell and of every member's d
//    shell (DIG 1; with DIG 0 K goes to HBM atomics directly), J(c, d_m) of every member, and J of
//    every union bra;
//  * bra-keyed (eri_grpb, DIG 2): K rows of the shared a shell and of every member's b shell,
//    J(a, b_m) of every member, and J of every union ket.
// Dynamic LDS, ket-keyed: [(NC + mmax ND) LN K columns (DIG 1)][mmax NC ND J(c,d)][ub NA NB J slots]
//                         [ub ints: the J address code of each slot's swept pair]
//              bra-keyed: [(NA + mmax NB) LN K rows][mmax NA NB J(a,b)][ub NC ND J slots][ub ints]
#define DEFINE_CLASS_KERNELS(NAME, LTOT)                                                        \
  template <int WGS, int DIG, typename R>                                                     \
  __global__ void __launch_bounds__(WGS) eri_grp_##NAME(                                       \
      const uint2* __restrict__ tasks, const Segment* __restrict__ segs,                      \
      int nres, const unsigned* __restrict__ ovf, unsigned* __restrict__ work,                \
      const PairInfo* __restrict__ pairs, const PrimPair* __restrict__ prims,                 \
      const double* __restrict__ boys, JKArgs jk, const GroupRec* __restrict__ groups,        \
      const int* __restrict__ uorder, int mmax, int ub, int cpi) {                            \
    constexpr int NCH = kChunks_##NAME, NC = kNC_##NAME, ND = kND_##NAME;                     \
    constexpr int NA = kNA_##NAME, NB = kNB_##NAME, NCD = NC * ND, NAB = NA * NB;             \
    __shared__ double boys_lds[BOYS_NGRID * BOYS_LDS_W];                                      \
    __shared__ int s_item;                                                                    \
    __shared__ int s_dcol[MAX_GROUP], s_moff[MAX_GROUP], s_mcnt[MAX_GROUP];                   \
    extern __shared__ double dyn_lds[];                                                       \
    const int nit = NCH / cpi;   /* items per task region: groups of cpi chunks */            \
    const long nitems = ((long)nres + *ovf) * nit;                                            \
    if (nitems == 0) return;                                                                  \
    stage_boys(boys, LTOT, boys_lds);                                                         \
    const int N = jk.ncart, LN = jk.lds_n;                                                    \
    const int ncolmax = DIG == 1 ? (NC + mmax * ND) * LN : 0;                                 \
    double* colK = dyn_lds;                                                                   \
    double* jcdL = dyn_lds + ncolmax;                                                         \
    double* jsl = jcdL + mmax * NCD;                                                          \
    int* spair = (int*)(jsl + ub * NAB);                                                      \
    for (int k = threadIdx.x; k < ncolmax + mmax * NCD + ub * NAB; k += WGS) dyn_lds[k] = 0.0; \
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;                                \
    while (true) {                                                                            \
      const long item = next_item(work, &s_item);                                             \
      if (item >= nitems) break;                                                              \
      const Segment sg = segs[item / nit];                                                    \
      if (sg.count == 0) continue;   /* a slot its group did not fill */                      \
      const int ch0 = (int)(item % nit) * cpi;                                                \
      const GroupRec gr = groups[sg.ket];                                                     \
      const int n = gr.n;                                                                     \
      const unsigned kblk = __builtin_amdgcn_readfirstlane((unsigned)sg.blk);                 \
      const uint2* reg = tasks + sg.off;                                                      \
      const int U = (int)reg[0].y;                                                            \
      if (threadIdx.x < n) {                                                                  \
        s_dcol[threadIdx.x] = pairs[uorder[gr.first + threadIdx.x]].b0;                       \
        s_moff[threadIdx.x] = (int)reg[1 + threadIdx.x].y;                                    \
        s_mcnt[threadIdx.x] = (int)reg[1 + threadIdx.x].x;                                    \
      }                                                                                       \
      __syncthreads();                                                                        \
      /* units (chunk, member, 64-entry block of the member's list), chunk-major, over the    \
         workgroup's wavefronts: the member is wavefront-uniform */                           \
      int nbt = 0;                                                                            \
      for (int mm = 0; mm < n; ++mm) nbt += (s_mcnt[mm] + 63) >> 6;                           \
      for (int u = wid; u < nbt * cpi; u += WGS / 64) {                                       \
        const int chunk = ch0 + u / nbt;                                                      \
        int v = u % nbt, mu = 0;                                                              \
        while (v >= ((s_mcnt[mu] + 63) >> 6)) { v -= (s_mcnt[mu] + 63) >> 6; ++mu; }          \
        mu = __builtin_amdgcn_readfirstlane(mu);                                              \
        const int b0 = v << 6, t = b0 + lane;                                                 \
        const bool active = t < s_mcnt[mu];                                                   \
        const uint2 e = reg[s_moff[mu] + (active ? t : b0)];                                  \
        const PairInfo B = pairs[pair_of(e.x)];                                               \
        if (active) spair[e.y] = jcode(B, N);                                                 \
        const unsigned kk = __builtin_amdgcn_readfirstlane((unsigned)uorder[gr.first + mu]);  \
        QuartetArgs q;                                                                        \
        make_quartet_uniform_ket(B, e.x, kk, kblk, active, pairs, prims, q);                  \
        quartet_jk_##NAME<1, DIG, R>(q, jk, boys_lds, chunk, colK, colK + (NC + mu * ND) * LN, \
                                     jcdL + mu * NCD, jsl + (long)e.y * NAB);                  \
      }                                                                                       \
      __syncthreads();                                                                        \
      const PairInfo K0 = pairs[uorder[gr.first]];                                            \
      if constexpr (DIG == 1) {                                                               \
        for (int k = threadIdx.x; k < (NC + n * ND) * LN; k += WGS) {                          \
          const double v = colK[k];                                                           \
          if (v != 0.0) {                                                                     \
            const int col = k / LN, row = jk.rowfn[K0.loff + k - col * LN];                   \
            const int md = (col - NC) / ND;                                                   \
            const int gcol = col < NC ? K0.a0 + col : s_dcol[md] + (col - NC - md * ND);       \
            atomicAdd(&jk.GK[0][(long)row * N + gcol], v);                                    \
            colK[k] = 0.0;                                                                    \
          }                                                                                   \
        }                                                                                     \
      }                                                                                       \
      for (int k = threadIdx.x; k < n * NCD; k += WGS) {                                      \
        const double v = jcdL[k];                                                             \
        if (v != 0.0) {                                                                       \
          const int m = k / NCD, ic = (k - m * NCD) / ND, idd = k - m * NCD - ic * ND;        \
          atomicAdd(&jk.GJ[0][jaddr(pairs[uorder[gr.first + m]], N, ic, idd)], v);            \
          jcdL[k] = 0.0;                                                                      \
        }                                                                                     \
      }                                                                                       \
      for (int k = threadIdx.x; k < U * NAB; k += WGS) {                                      \
        const double v = jsl[k];                                                              \
        if (v != 0.0) {                                                                       \
          const int u = k / NAB, ia = (k - u * NAB) / NB, ib = k - u * NAB - ia * NB;         \
          atomicAdd(&jk.GJ[0][jcode_addr(spair[u], N, ia, ib)], v);                           \
          jsl[k] = 0.0;                                                                       \
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
      const bool active = t < count;                                                            \
      const uint2 tk = tasks[active ? t : wbase];                                               \
      const PairInfo B = pairs[pair_of(tk.x)];                                                  \
      const PairInfo K = pairs[pair_of(tk.y)];                                                  \
      QuartetArgs q;                                                                            \
      q.prim = prims;                                                                           \
      prim_range(B, block_of(tk.x), q.bra_p0, q.bra_p1);                                        \
      prim_range(K, block_of(tk.y), q.ket_p0, q.ket_p1);                                        \
      if (active) quartet_block_##NAME(q, out + (slot ? (long)slot[t] : t) * (long)blk, boys_lds); \
    }                                                                                           \
  }

// Bra-keyed group kernel (DIG 2), for class pairs whose bra class has few pairs: the uniform
// members are bras (a, b_m) sharing shell a, the lanes hold the kets that survive against them.
#define DEFINE_CLASS_SEGB(NAME, LTOT)                                                           \
  template <typename R>                                                                         \
  __global__ void __launch_bounds__(WG) eri_grpb_##NAME(                                        \
      const uint2* __restrict__ tasks, const Segment* __restrict__ segs,                        \
      int nres, const unsigned* __restrict__ ovf, unsigned* __restrict__ work,                  \
      const PairInfo* __restrict__ pairs, const PrimPair* __restrict__ prims,                   \
      const double* __restrict__ boys, JKArgs jk, const GroupRec* __restrict__ groups,          \
      const int* __restrict__ uorder, int mmax, int ub, int cpi) {                              \
    constexpr int NCH = kChunks_##NAME, NC = kNC_##NAME, ND = kND_##NAME;                       \
    constexpr int NA = kNA_##NAME, NB = kNB_##NAME, NCD = NC * ND, NAB = NA * NB;               \
    __shared__ double boys_lds[BOYS_NGRID * BOYS_LDS_W];                                        \
    __shared__ int s_item;                                                                      \
    __shared__ int s_brow[MAX_GROUP], s_moff[MAX_GROUP], s_mcnt[MAX_GROUP];                     \
    extern __shared__ double dyn_lds[];                                                         \
    const int nit = NCH / cpi;   /* items per task region: groups of cpi chunks */              \
    const long nitems = ((long)nres + *ovf) * nit;                                              \
    if (nitems == 0) return;                                                                    \
    stage_boys(boys, LTOT, boys_lds);                                                           \
    const int N = jk.ncart, LN = jk.lds_n;                                                      \
    const int nrowmax = (NA + mmax * NB) * LN;                                                  \
    double* rowK = dyn_lds;                                                                     \
    double* jabL = dyn_lds + nrowmax;                                                           \
    double* jsl = jabL + mmax * NAB;                                                            \
    int* spair = (int*)(jsl + ub * NCD);                                                        \
    for (int k = threadIdx.x; k < nrowmax + mmax * NAB + ub * NCD; k += WG) dyn_lds[k] = 0.0;   \
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;                                  \
    while (true) {                                                                              \
      const long item = next_item(work, &s_item);                                               \
      if (item >= nitems) break;                                                                \
      const Segment sg = segs[item / nit];                                                      \
      if (sg.count == 0) continue;   /* a slot its group did not fill */                        \
      const int ch0 = (int)(item % nit) * cpi;                                                  \
      const GroupRec gr = groups[sg.ket];                                                       \
      const int n = gr.n;                                                                       \
      const unsigned bblk = __builtin_amdgcn_readfirstlane((unsigned)sg.blk);                   \
      const uint2* reg = tasks + sg.off;                                                        \
      const int U = (int)reg[0].y;                                                              \
      if (threadIdx.x < n) {                                                                    \
        s_brow[threadIdx.x] = pairs[uorder[gr.first + threadIdx.x]].b0;                         \
        s_moff[threadIdx.x] = (int)reg[1 + threadIdx.x].y;                                      \
        s_mcnt[threadIdx.x] = (int)reg[1 + threadIdx.x].x;                                      \
      }                                                                                         \
      __syncthreads();                                                                          \
      int nbt = 0;                                                                              \
      for (int mm = 0; mm < n; ++mm) nbt += (s_mcnt[mm] + 63) >> 6;                             \
      for (int u = wid; u < nbt * cpi; u += WG / 64) {                                          \
        const int chunk = ch0 + u / nbt;                                                        \
        int v = u % nbt, mu = 0;                                                                \
        while (v >= ((s_mcnt[mu] + 63) >> 6)) { v -= (s_mcnt[mu] + 63) >> 6; ++mu; }            \
        mu = __builtin_amdgcn_readfirstlane(mu);                                                \
        const int b0 = v << 6, t = b0 + lane;                                                   \
        const bool active = t < s_mcnt[mu];                                                     \
        const uint2 e = reg[s_moff[mu] + (active ? t : b0)];                                    \
        const PairInfo K = pairs[pair_of(e.x)];                                                 \
        if (active) spair[e.y] = jcode(K, N);                                                   \
        const unsigned bb = __builtin_amdgcn_readfirstlane((unsigned)uorder[gr.first + mu]);    \
        QuartetArgs q;                                                                          \
        make_quartet_uniform_bra(K, e.x, bb, bblk, active, pairs, prims, q);                    \
        quartet_jk_##NAME<1, 2, R>(q, jk, boys_lds, chunk, rowK, rowK + (NA + mu * NB) * LN,     \
                                   jabL + mu * NAB, jsl + (long)e.y * NCD);                     \
      }                                                                                         \
      __syncthreads();                                                                          \
      const PairInfo A0 = pairs[uorder[gr.first]];                                              \
      for (int k = threadIdx.x; k < (NA + n * NB) * LN; k += WG) {                              \
        const double v = rowK[k];                                                               \
        if (v != 0.0) {                                                                         \
          const int lrow = k / LN, col = jk.rowfn[A0.loff + k - lrow * LN];                     \
          const int mb = (lrow - NA) / NB;                                                      \
          const int grow = lrow < NA ? A0.a0 + lrow : s_brow[mb] + (lrow - NA - mb * NB);        \
          atomicAdd(&jk.GK[0][(long)grow * N + col], v);                                        \
          rowK[k] = 0.0;                                                                        \
        }                                                                                       \
      }                                                                                         \
      for (int k = threadIdx.x; k < n * NAB; k += WG) {                                         \
        const double v = jabL[k];                                                               \
        if (v != 0.0) {                                                                         \
          const int m = k / NAB, ia = (k - m * NAB) / NB, ib = k - m * NAB - ia * NB;           \
          atomicAdd(&jk.GJ[0][jaddr(pairs[uorder[gr.first + m]], N, ia, ib)], v);               \
          jabL[k] = 0.0;                                                                        \
        }                                                                                       \
      }                                                                                         \
      for (int k = threadIdx.x; k < U * NCD; k += WG) {                                         \
        const double v = jsl[k];                                                                \
        if (v != 0.0) {                                                                         \
          const int u = k / NCD, ic = (k - u * NCD) / ND, idd = k - u * NCD - ic * ND;          \
          atomicAdd(&jk.GJ[0][jcode_addr(spair[u], N, ic, idd)], v);                            \
          jsl[k] = 0.0;                                                                         \
        }                                                                                       \
      }                                                                                         \
    }                                                                                           \
  }

typedef void (*BlockKernel)(const uint2*, long, const PairInfo*, const PrimPair*, const double*, double*, int,
                            const int*);
typedef void (*GroupKernel)(const uint2*, const Segment*, int, const unsigned*, unsigned*, const PairInfo*,
                            const PrimPair*, const double*, JKArgs, const GroupRec*, const int*, int, int, int);

struct ClassKernels {
  int i, j, bi, ki, L, nchunk;
  const char* name;
  BlockKernel block;
  // grp[precision][variant]: precision 0 FP64 / 1 FP32 integrals; variant 0 ket-keyed with HBM K
  // atomics, 1 ket-keyed with LDS K columns, 2 bra-keyed with LDS K rows (null where the class
  // pair has no bra-keyed form)
  GroupKernel grp[2][3];
  int nc, nd, na, nb;
  int seg_wg;         // workgroup size of the ket-keyed kernels
};
