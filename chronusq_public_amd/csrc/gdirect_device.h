// Device-side types and helpers shared by the generated per-class ERI cores and the
// G[D] kernels.  gfx950 only (64-wide wavefronts, FP64 VALU, FP64 global atomics).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Boys-function table: F_m(T_i) on T_i = i * BOYS_STEP, i = 0..BOYS_NGRID-1, for
// m = 0..BOYS_MTAB-1, plus exp(-T_i) in the last column.  Each kernel stages the
// slice m = L..L+7 and the exp column (9 doubles per grid point) into LDS.
#define BOYS_STEP 0.1
#define BOYS_INV_STEP 10.0
#define BOYS_TMAX 36.0
#define BOYS_NGRID 361
#define BOYS_MTAB 24
#define BOYS_LDS_W 9

// One primitive pair of a shell pair (a, b):
//   p = alpha_a + alpha_b, K = sqrt(2 pi^(5/2)) exp(-alpha_a alpha_b / p |AB|^2) c_a c_b / p
//   (a primitive quartet's prefactor 2 pi^(5/2) K_ab K_cd / (p q sqrt(p + q)) is then
//   K_ab K_cd / sqrt(p + q), split evenly over its two pairs),
//   P = (alpha_a A + alpha_b B) / p, PA = P - A, PB = P - B, o2p = 1 / (2p).
struct alignas(16) PrimPair {
  double p;
  double K;
  double P[3];
  double PA[3];
  double PB[3];
  double o2p;
};

// Per shell pair, in kernel orientation (l_a >= l_b).
struct PairInfo {
  int a, b;        // shells in kernel orientation
  int s1, s2;      // canonical shells (s1 >= s2), as the reference loop visits them
  int p0, p1;      // primitive-pair range in the PrimPair array
  int a0, b0;      // first Cartesian function of a and b
  int loff;        // first Cartesian function of the pair's molecule (0 for one molecule)
  int ar, br;      // first LDS row of a and b: their place among the functions of the
                   // molecule's shells whose angular momentum is la or lb (the pair's class)
};

struct QuartetArgs {
  const PrimPair* __restrict__ prim;
  int bra_p0, bra_p1, ket_p0, ket_p1;
  int a0, b0, c0, d0;
  double deg;
  bool ket_uniform;          // every lane of the wavefront holds the same ket pair
  bool bra_swap, ket_swap;   // pair held reversed relative to its canonical (s1 >= s2) order
  int ar, br, cr, dr;        // LDS rows of the swept pair's shells (PairInfo::ar, br)
};

// One density slot per kernel launch (a call with several densities runs one pass per slot).
#define MAX_SLOTS 1
struct JKArgs {
  int ncart;
  int lds_n;                     // rows of an LDS accumulator column (or row): the largest
                                 // molecule's functions of the swept pair class's shells
  const int* rowfn;              // Cartesian function of LDS row r: rowfn[loff + r]
  const double* D[MAX_SLOTS];    // density in the Cartesian basis, row-major ncart x ncart
  double* GJ[MAX_SLOTS];         // forward Coulomb updates
  double* GJ2[MAX_SLOTS];        // forward Coulomb updates from the transposed density
  double* GK[MAX_SLOTS];         // forward exchange updates
  double* GK2[MAX_SLOTS];        // forward exchange updates from the transposed density
  int do_j[MAX_SLOTS];
  int do_k[MAX_SLOTS];
  const int* nonsym;             // device flags: density slot is not symmetric
};

// a * b + c rounded once, in the integral type of a kernel (FP64, or FP32 for the mixed
// window's kernels)
__device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// 1/sqrt(x) for x > 0: hardware estimate y0 (about 2^-23 relative), then one third-order
// correction y = y0 (1 + e/2 + 3e^2/8), e = 1 - x y0^2 (the binomial series of (1 - e)^-1/2;
// truncation 2.5 d^3 for a starting error d, far below rounding): five operations where two
// Newton steps take eight, the same accuracy (max 1.4e-16 relative in a CPU emulation with
// |d| <= 2^-22).
__device__ __forceinline__ double rsqrt_f64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-x * y, y, 1.0);
  const double c = __builtin_fma(e, 0.375, 0.5);
  return __builtin_fma(y * e, c, y);
}

// F_0..F_L at T.  T < 36: 8-term Taylor expansion of F_L about the nearest grid point
// (|dT| <= 0.05, truncation below 1e-15 relative), F(T0 - dT) = sum_k F_{L+k}(T0) dT^k / k!;
// for L >= 2 exp(-T) = exp(-T0) exp(dT) by an 8-term series and the downward recursion
// F_m = (2T F_{m+1} + e^-T) / (2m+1).  For L <= 1 the staged coefficients are already divided
// by k! (one FMA per term), and for L = 1 the slice holds F_0..F_8 / k!: F_0 is the polynomial
// and F_1 = -dF_0/dT its derivative in dT, both from one Horner pass (8 terms each).
// T >= 36: F_0 = sqrt(pi/T)/2, upward recursion F_{m+1} = (2m+1)/(2T) F_m (the dropped
// e^-T term is below 2.4e-16).
template <int L>
__device__ __forceinline__ void boys_eval(double T, const double* __restrict__ lds, double* F) {
  // Both branches are evaluated and the result selected, so that the primitive loop has no
  // divergent branch (unrolled iterations then interleave): the table path at min(T, T_max)
  // (index always in range), the asymptotic path at max(T, T_max).
  const bool big = T >= BOYS_TMAX;
  const double Tt = fmin(T, BOYS_TMAX);   // (the table's last point for the asymptotic range)
  const int i = (int)(Tt * BOYS_INV_STEP + 0.5);
  const double dT = (double)i * BOYS_STEP - Tt;
  const double* f = lds + i * BOYS_LDS_W;
  if constexpr (L >= 2) {
    // unscaled coefficients F_L..F_{L+7} (measured faster for these kernels than the scaled form)
    double s = f[7];
    s = f[6] + s * dT * (1.0 / 7.0);
    s = f[5] + s * dT * (1.0 / 6.0);
    s = f[4] + s * dT * (1.0 / 5.0);
    s = f[3] + s * dT * (1.0 / 4.0);
    s = f[2] + s * dT * (1.0 / 3.0);
    s = f[1] + s * dT * (1.0 / 2.0);
    s = f[0] + s * dT;
    const double Ta = big ? T : BOYS_TMAX;
    const double rs = rsqrt_f64(Ta);
    double fa = 0.88622692545275801365 * rs;   // sqrt(pi)/2 T^-1/2
    // table: F_L by Taylor, then downward recursion with e^{-T} (Taylor about T_i)
    double e = 1.0 + dT * (1.0 / 8.0);
    e = 1.0 + dT * (1.0 / 7.0) * e;
    e = 1.0 + dT * (1.0 / 6.0) * e;
    e = 1.0 + dT * (1.0 / 5.0) * e;
    e = 1.0 + dT * (1.0 / 4.0) * e;
    e = 1.0 + dT * (1.0 / 3.0) * e;
    e = 1.0 + dT * (1.0 / 2.0) * e;
    e = 1.0 + dT * e;
    e *= f[8];
    const double T2 = 2.0 * Tt;
    double Ft[L + 1];
    Ft[L] = s;
#pragma unroll
    for (int m = L - 1; m >= 0; --m) Ft[m] = (T2 * Ft[m + 1] + e) * (1.0 / (2 * m + 1));
    // asymptotic: F_0 = sqrt(pi/T)/2, upward F_{m+1} = F_m (2m+1) / (2T)
    const double hr = 0.5 * rs * rs;
    F[0] = big ? fa : Ft[0];
#pragma unroll
    for (int m = 0; m < L; ++m) {
      fa = fa * (double)(2 * m + 1) * hr;
      F[m + 1] = big ? fa : Ft[m + 1];
    }
  } else {
    const double Ta = big ? T : BOYS_TMAX;
    const double rs = rsqrt_f64(Ta);
    const double fa = 0.88622692545275801365 * rs;   // sqrt(pi)/2 T^-1/2
    if constexpr (L == 1) {
      // (the first step of the derivative's recurrence, dp = 0 dT + f[8], taken out: the
      // compiler cannot fold the product with zero itself)
      double dp = f[8], p = __builtin_fma(f[8], dT, f[7]);
#pragma unroll
      for (int k = 6; k >= 0; --k) {
        dp = __builtin_fma(dp, dT, p);
        p = __builtin_fma(p, dT, f[k]);
      }
      F[0] = big ? fa : p;
      F[1] = big ? fa * (0.5 * rs * rs) : dp;
    } else {
      double s = f[7];
#pragma unroll
      for (int k = 6; k >= 0; --k) s = __builtin_fma(s, dT, f[k]);
      F[0] = big ? fa : s;
    }
  }
}

// FP32 integrals take the Boys values from the FP64 evaluation, rounded.
template <int L>
__device__ __forceinline__ void boys_eval(double T, const double* __restrict__ lds, float* F) {
  double Fd[L + 1];
  boys_eval<L>(T, lds, Fd);
#pragma unroll
  for (int m = 0; m <= L; ++m) F[m] = (float)Fd[m];
}

// Stage the Boys slice for total angular momentum L into LDS (whole workgroup): per grid point
// F_{L+c}, c = 0..7, and e^{-T_i}; divided by c! for L = 0, and F_c / c!, c = 0..8, for L = 1.
__device__ __forceinline__ void stage_boys(const double* __restrict__ tab, int L, double* lds) {
  constexpr double ifact[9] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040,
                               1.0 / 40320};
  for (int k = threadIdx.x; k < BOYS_NGRID * BOYS_LDS_W; k += blockDim.x) {
    const int i = k / BOYS_LDS_W, c = k - i * BOYS_LDS_W;
    lds[k] = L == 1 ? tab[i * (BOYS_MTAB + 1) + c] * ifact[c]
             : L == 0 ? (c < 8 ? tab[i * (BOYS_MTAB + 1) + c] * ifact[c] : tab[i * (BOYS_MTAB + 1) + BOYS_MTAB])
                      : (c < 8) ? tab[i * (BOYS_MTAB + 1) + L + c] : tab[i * (BOYS_MTAB + 1) + BOYS_MTAB];
  }
  __syncthreads();
}

// Work unit of the segment kernels: `count` tasks of one ket pair starting at `off` in the
// class pair's task list.
struct Segment {
  int off, count, ket, blk;   // ket: the uniform pair; blk: its primitive-block code
};

// Primitive blocks.  A pair with many primitive pairs is split into blocks so that one heavily
// contracted quartet becomes several shorter lane tasks (the digestion is linear in the
// integrals, so partial sums digest independently): blocks of PRIM_BLOCK on the per-lane side
// of a task list, blocks of UNI_BLOCK on the wavefront-uniform side (each uniform block is a
// separate work item that repeats the per-lane J atomics and the LDS flush, so it is kept
// coarse).  Task entries and segments carry a block code in bits 24..31 of the pair index:
// 0 = the whole primitive range, k >= 1 = primitives [p0 + (k-1) B, min(p1, p0 + k B)).
#define PRIM_BLOCK 8
#define UNI_BLOCK 32
#define PAIR_MASK 0xFFFFFFu
__host__ __device__ __forceinline__ unsigned pair_of(unsigned code) { return code & PAIR_MASK; }
__host__ __device__ __forceinline__ unsigned block_of(unsigned code) { return code >> 24; }
template <int BLK = PRIM_BLOCK>
__host__ __device__ __forceinline__ void prim_range(const PairInfo& P, unsigned blk, int& p0, int& p1) {
  if (blk == 0) {
    p0 = P.p0;
    p1 = P.p1;
  } else {
    p0 = P.p0 + (int)(blk - 1) * BLK;
    p1 = P.p1 < p0 + BLK ? P.p1 : p0 + BLK;
  }
}

// Wavefront-reduced add into LDS (every lane targets the same LDS word).
__device__ __forceinline__ void wave_add_lds(double* p, double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(p, v);
}

// Atomic add of a value whose target is the same on every lane when `uniform` (wavefront-
// uniform flag): the 64 values are summed across lanes first and one lane issues the atomic.
__device__ __forceinline__ void wave_add(double* p, double v, bool uniform) {
  if (uniform) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(p, v);
  } else {
    atomicAdd(p, v);
  }
}
