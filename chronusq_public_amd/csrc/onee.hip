// One-electron integrals on the GPU: overlap S, kinetic T, nuclear attraction V, and the
// Cartesian multipoles up to third order (replaces AOIntegrals::computeAOOneE / OneEDriver,
// src/aointegrals/aointegrals_builders.cpp:27-132,337-467).  McMurchie-Davidson: Hermite expansion coefficients by their closed form
//   E^{ij}_t = sum_{k<=i,l<=j} C(i,k) C(j,l) PA^{i-k} PB^{j-l} n!/(t! m!) (1/4p)^m (1/2p)^t,
//   n = k + l, m = (n - t)/2,
// S_1D(i,j) = E^{ij}_0 sqrt(pi/p),
// T_1D(i,j) = -1/2 [ j(j-1) S(i,j-2) - 2b(2j+1) S(i,j) + 4b^2 S(i,j+2) ],
// V = -(2 pi / p) sum_C Z_C sum_tuv E_t E_u E_v R_tuv(p, P - C).
// One thread per (shell pair, slice), every loop unrolled at compile time: S and T in one pass
// whose slices are the primitives of shell a, V in a second pass whose slices are groups of
// ONEE_ATOMS nuclei (summing R_tuv over them first).  Partial blocks are added atomically to
// zeroed row-major ncart x ncart matrices (and their transposes for a != b).
#include <algorithm>

#include "gdirect_kernels.h"
#include "onee.h"

namespace {

__host__ __device__ constexpr int comp_row(int k) {
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= k) ++i;
  return i;
}
__host__ __device__ constexpr int comp_x(int l, int k) { return l - comp_row(k); }
__host__ __device__ constexpr int comp_y(int l, int k) { return comp_row(k) - (k - comp_row(k) * (comp_row(k) + 1) / 2); }
__host__ __device__ constexpr int comp_z(int l, int k) { return l - comp_x(l, k) - comp_y(l, k); }
// n! for n <= 12 from a table (indices are compile-time constants once the loops unroll, so the
// coefficients fold; a recursive definition would leave run-time calls)
__host__ __device__ constexpr double cfact(int n) {
  constexpr double f[13] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0, 362880.0,
                            3628800.0, 39916800.0, 479001600.0};
  return f[n < 0 ? 0 : (n > 12 ? 12 : n)];
}
__host__ __device__ constexpr double cbinom(int n, int k) { return cfact(n) / (cfact(k) * cfact(n - k)); }

// position of (t, u, v), t + u + v = s, among all Hermite triples with t + u + v <= L:
// the triples of lower sums first, then by descending t and ascending v
__host__ __device__ constexpr int hidx(int t, int u, int v) {
  const int s = t + u + v;
  return s * (s + 1) * (s + 2) / 6 + (s - t) * (s - t + 1) / 2 + v;
}

__device__ __forceinline__ double ipow(double x, int n) {
  double r = 1.0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (k < n) r *= x;
  return r;
}

// E^{ij}_t for one Cartesian direction (without the exp(-mu AB^2) factor).
__device__ __forceinline__ double hermE(int i, int j, int t, double PA, double PB, double o2p) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k <= 6; ++k) {
    if (k > i) break;
#pragma unroll
    for (int l = 0; l <= 6; ++l) {
      if (l > j) break;
      const int n = k + l;
      if (n < t || ((n - t) & 1)) continue;
      const int m = (n - t) / 2;
      const double c = cbinom(i, k) * cbinom(j, l) * cfact(n) / (cfact(t) * cfact(m));
      s += c * ipow(PA, i - k) * ipow(PB, j - l) * ipow(0.5 * o2p, m) * ipow(o2p, t);
    }
  }
  return s;
}

// nuclei per thread of the nuclear-attraction kernels (more slices: more threads in flight)
#define ONEE_ATOMS 2

template <int LA, int LB, bool DO_V>
__global__ void __launch_bounds__(WG) onee_kernel(const PairInfo* __restrict__ pairs, int p_begin, int p_end,
                                                  const OneEShellData sd, const double* __restrict__ Zc,
                                                  const double* __restrict__ xyz, int natom,
                                                  const double* __restrict__ boys, int ncart,
                                                  double* __restrict__ S, double* __restrict__ T,
                                                  double* __restrict__ V) {
  constexpr int NA = (LA + 1) * (LA + 2) / 2, NB = (LB + 1) * (LB + 2) / 2, L = LA + LB;
  __shared__ double boys_lds[BOYS_NGRID * BOYS_LDS_W];
  if (DO_V) stage_boys(boys, L, boys_lds);
  const int P = p_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (P >= p_end) return;
  const PairInfo pi = pairs[P];
  const int a = pi.a, b = pi.b;
  // slice: one primitive of a (S, T) or nuclei [c0, c1) (V)
  const int ip0 = DO_V ? sd.prim_off[a] : sd.prim_off[a] + (int)blockIdx.y;
  const int ip1 = DO_V ? sd.prim_off[a + 1] : ip0 + 1;
  if (ip0 >= sd.prim_off[a + 1]) return;
  // nuclei of this slice: those of the pair's molecule when several molecules share the context
  int at0 = 0, at1 = natom;
  if (DO_V && sd.mol_atom0) {
    const int m = sd.shell_mol[a];
    at0 = sd.mol_atom0[m];
    at1 = sd.mol_atom0[m + 1];
  }
  const int c0 = DO_V ? at0 + (int)blockIdx.y * ONEE_ATOMS : 0;
  const int c1 = DO_V ? min(at1, c0 + ONEE_ATOMS) : 0;
  if (DO_V && c0 >= at1) return;
  const double A[3] = {sd.origin[3 * a], sd.origin[3 * a + 1], sd.origin[3 * a + 2]};
  const double B[3] = {sd.origin[3 * b], sd.origin[3 * b + 1], sd.origin[3 * b + 2]};
  const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
  double s[NA][NB], t[NA][NB];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) { s[i][j] = 0.0; t[i][j] = 0.0; }
  for (int ip = ip0; ip < ip1; ++ip)
    for (int jp = sd.prim_off[b]; jp < sd.prim_off[b + 1]; ++jp) {
      const double al = sd.alpha[ip], be = sd.alpha[jp];
      const double p = al + be, o2p = 0.5 / p;
      const double K = exp(-al * be / p * AB2) * sd.coef[ip] * sd.coef[jp];
      const double PA0 = be / p * (B[0] - A[0]), PA1 = be / p * (B[1] - A[1]), PA2 = be / p * (B[2] - A[2]);
      const double PB0 = al / p * (A[0] - B[0]), PB1 = al / p * (A[1] - B[1]), PB2 = al / p * (A[2] - B[2]);
      if (!DO_V) {
        const double sq = sqrt(3.14159265358979323846 / p);
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
          for (int ib = 0; ib < NB; ++ib) {
            const int ax = comp_x(LA, ia), ay = comp_y(LA, ia), az = comp_z(LA, ia);
            const int bx = comp_x(LB, ib), by = comp_y(LB, ib), bz = comp_z(LB, ib);
            const double sx = hermE(ax, bx, 0, PA0, PB0, o2p) * sq;
            const double sy = hermE(ay, by, 0, PA1, PB1, o2p) * sq;
            const double sz = hermE(az, bz, 0, PA2, PB2, o2p) * sq;
            auto kin = [&](int i, int j, double PAd, double PBd, double sij) {
              double d2 = 4.0 * be * be * hermE(i, j + 2, 0, PAd, PBd, o2p) * sq - 2.0 * be * (2 * j + 1) * sij;
              if (j >= 2) d2 += (double)(j * (j - 1)) * hermE(i, j - 2, 0, PAd, PBd, o2p) * sq;
              return -0.5 * d2;
            };
            s[ia][ib] += K * sx * sy * sz;
            t[ia][ib] += K * (kin(ax, bx, PA0, PB0, sx) * sy * sz + sx * kin(ay, by, PA1, PB1, sy) * sz +
                              sx * sy * kin(az, bz, PA2, PB2, sz));
          }
      } else {
        const double Px = A[0] + PA0, Py = A[1] + PA1, Pz = A[2] + PA2;
        // Rsum_tuv = sum_C Z_C R_tuv(p, P - C), t + u + w <= L, in the compact index hidx;
        // the recursion over n keeps two levels
        constexpr int NH = (L + 1) * (L + 2) * (L + 3) / 6;
        double Rs[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h) Rs[h] = 0.0;
        for (int c = c0; c < c1; ++c) {
          const double X = Px - xyz[3 * c], Y = Py - xyz[3 * c + 1], Z = Pz - xyz[3 * c + 2];
          double F[L + 1];
          boys_eval<L>(p * (X * X + Y * Y + Z * Z), boys_lds, F);
          double fac[L + 1];
          fac[0] = Zc[c];
#pragma unroll
          for (int n = 1; n <= L; ++n) fac[n] = fac[n - 1] * (-2.0 * p);
          double Rc[NH], Rn[NH];
          Rc[0] = fac[L] * F[L];
#pragma unroll
          for (int n = L - 1; n >= 0; --n) {
            Rn[0] = fac[n] * F[n];
#pragma unroll
            for (int sl = 1; sl <= L; ++sl) {
              if (sl > L - n) break;
#pragma unroll
              for (int tt = sl; tt >= 0; --tt)
#pragma unroll
                for (int w = 0; w <= sl - tt; ++w) {
                  const int u = sl - tt - w;
                  double r;
                  if (tt > 0) r = X * Rc[hidx(tt - 1, u, w)] + (tt > 1 ? (tt - 1) * Rc[hidx(tt - 2, u, w)] : 0.0);
                  else if (u > 0) r = Y * Rc[hidx(tt, u - 1, w)] + (u > 1 ? (u - 1) * Rc[hidx(tt, u - 2, w)] : 0.0);
                  else r = Z * Rc[hidx(tt, u, w - 1)] + (w > 1 ? (w - 1) * Rc[hidx(tt, u, w - 2)] : 0.0);
                  Rn[hidx(tt, u, w)] = r;
                }
            }
#pragma unroll
            for (int h = 0; h < NH; ++h) Rc[h] = Rn[h];
          }
#pragma unroll
          for (int h = 0; h < NH; ++h) Rs[h] += Rc[h];
        }
        const double pref = -2.0 * 3.14159265358979323846 / p * K;
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
          for (int ib = 0; ib < NB; ++ib) {
            const int ax = comp_x(LA, ia), ay = comp_y(LA, ia), az = comp_z(LA, ia);
            const int bx = comp_x(LB, ib), by = comp_y(LB, ib), bz = comp_z(LB, ib);
            double acc = 0.0;
#pragma unroll
            for (int tt = 0; tt <= ax + bx; ++tt) {
              const double ex = hermE(ax, bx, tt, PA0, PB0, o2p);
#pragma unroll
              for (int u = 0; u <= ay + by; ++u) {
                const double ey = hermE(ay, by, u, PA1, PB1, o2p);
#pragma unroll
                for (int w = 0; w <= az + bz; ++w) acc += ex * ey * hermE(az, bz, w, PA2, PB2, o2p) * Rs[hidx(tt, u, w)];
              }
            }
            s[ia][ib] += pref * acc;   // V accumulates in s in this pass
          }
      }
    }
  double* M0 = DO_V ? V : S;
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) {
      const long r = pi.a0 + ia, cidx = pi.b0 + ib;
      atomicAdd(&M0[r * ncart + cidx], s[ia][ib]);
      if (a != b) atomicAdd(&M0[cidx * ncart + r], s[ia][ib]);
      if (!DO_V) {
        atomicAdd(&T[r * ncart + cidx], t[ia][ib]);
        if (a != b) atomicAdd(&T[cidx * ncart + r], t[ia][ib]);
      }
    }
}

// Cartesian multipole x_C^EX y_C^EY z_C^EZ about C: per direction
//   <i| x_C^e |j> = sum_k C(e,k) (B - C)^(e-k) <i| j + k>     (x - C = (x - B) + (B - C)),
// the overlaps with raised angular momentum on b from the Hermite coefficients E^{i,j+k}_0.
template <int E>
__device__ __forceinline__ double moment_1d(int i, int j, double PA, double PB, double o2p, double BC) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k <= E; ++k) s += cbinom(E, k) * ipow(BC, E - k) * hermE(i, j + k, 0, PA, PB, o2p);
  return s;
}

template <int LA, int LB, int EX, int EY, int EZ>
__global__ void __launch_bounds__(WG) multipole_kernel(const PairInfo* __restrict__ pairs, int p_begin, int p_end,
                                                       const OneEShellData sd, double Cx, double Cy, double Cz,
                                                       int ncart, double* __restrict__ M) {
  constexpr int NA = (LA + 1) * (LA + 2) / 2, NB = (LB + 1) * (LB + 2) / 2;
  const int P = p_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (P >= p_end) return;
  const PairInfo pi = pairs[P];
  const int a = pi.a, b = pi.b;
  const int ip = sd.prim_off[a] + (int)blockIdx.y;   // slice: one primitive of a
  if (ip >= sd.prim_off[a + 1]) return;
  const double A[3] = {sd.origin[3 * a], sd.origin[3 * a + 1], sd.origin[3 * a + 2]};
  const double B[3] = {sd.origin[3 * b], sd.origin[3 * b + 1], sd.origin[3 * b + 2]};
  const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
  const double BCx = B[0] - Cx, BCy = B[1] - Cy, BCz = B[2] - Cz;
  double m[NA][NB];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) m[i][j] = 0.0;
  for (int jp = sd.prim_off[b]; jp < sd.prim_off[b + 1]; ++jp) {
    const double al = sd.alpha[ip], be = sd.alpha[jp];
    const double p = al + be, o2p = 0.5 / p;
    const double sq = sqrt(3.14159265358979323846 / p);
    const double K = exp(-al * be / p * AB2) * sd.coef[ip] * sd.coef[jp] * sq * sq * sq;
    const double PA0 = be / p * (B[0] - A[0]), PA1 = be / p * (B[1] - A[1]), PA2 = be / p * (B[2] - A[2]);
    const double PB0 = al / p * (A[0] - B[0]), PB1 = al / p * (A[1] - B[1]), PB2 = al / p * (A[2] - B[2]);
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) {
        const int ax = comp_x(LA, ia), ay = comp_y(LA, ia), az = comp_z(LA, ia);
        const int bx = comp_x(LB, ib), by = comp_y(LB, ib), bz = comp_z(LB, ib);
        m[ia][ib] += K * moment_1d<EX>(ax, bx, PA0, PB0, o2p, BCx) * moment_1d<EY>(ay, by, PA1, PB1, o2p, BCy) *
                     moment_1d<EZ>(az, bz, PA2, PB2, o2p, BCz);
      }
  }
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) {
      const long r = pi.a0 + ia, cidx = pi.b0 + ib;
      atomicAdd(&M[r * ncart + cidx], m[ia][ib]);
      if (a != b) atomicAdd(&M[cidx * ncart + r], m[ia][ib]);
    }
}

template <int LA, int LB>
void launch_multipole_class(const PairInfo* pairs, int p_begin, int p_end, const OneEShellData& sd, int max_prim,
                            const double* C, int order, int ncart, double* M, hipStream_t stream) {
  const int n = p_end - p_begin;
  const dim3 grid((n + WG - 1) / WG, std::max(1, max_prim));
  const long nc2 = (long)ncart * ncart;
  int k = 0;
#define CQ_MP(EX, EY, EZ)                                                                                   \
  hipLaunchKernelGGL((multipole_kernel<LA, LB, EX, EY, EZ>), grid, dim3(WG), 0, stream, pairs, p_begin, p_end, \
                     sd, C[0], C[1], C[2], ncart, M + (k++) * nc2)
  if (order >= 1) { CQ_MP(1, 0, 0); CQ_MP(0, 1, 0); CQ_MP(0, 0, 1); }
  if (order >= 2) { CQ_MP(2, 0, 0); CQ_MP(1, 1, 0); CQ_MP(1, 0, 1); CQ_MP(0, 2, 0); CQ_MP(0, 1, 1); CQ_MP(0, 0, 2); }
  if (order >= 3) {
    CQ_MP(3, 0, 0); CQ_MP(2, 1, 0); CQ_MP(2, 0, 1); CQ_MP(1, 2, 0); CQ_MP(1, 1, 1);
    CQ_MP(1, 0, 2); CQ_MP(0, 3, 0); CQ_MP(0, 2, 1); CQ_MP(0, 1, 2); CQ_MP(0, 0, 3);
  }
#undef CQ_MP
}

}  // namespace

void launch_multipole(int pair_class, const PairInfo* pairs, int p_begin, int p_end, const OneEShellData& sd,
                      int max_prim, const double* C, int order, int ncart, double* M, hipStream_t stream) {
  if (p_end <= p_begin) return;
  switch (pair_class) {
    case 0: launch_multipole_class<0, 0>(pairs, p_begin, p_end, sd, max_prim, C, order, ncart, M, stream); break;
    case 1: launch_multipole_class<1, 0>(pairs, p_begin, p_end, sd, max_prim, C, order, ncart, M, stream); break;
    case 2: launch_multipole_class<1, 1>(pairs, p_begin, p_end, sd, max_prim, C, order, ncart, M, stream); break;
    case 3: launch_multipole_class<2, 0>(pairs, p_begin, p_end, sd, max_prim, C, order, ncart, M, stream); break;
    case 4: launch_multipole_class<2, 1>(pairs, p_begin, p_end, sd, max_prim, C, order, ncart, M, stream); break;
    case 5: launch_multipole_class<2, 2>(pairs, p_begin, p_end, sd, max_prim, C, order, ncart, M, stream); break;
    default: break;
  }
}

void launch_onee(int pair_class, const PairInfo* pairs, int p_begin, int p_end, const OneEShellData& sd,
                 int max_prim, const double* Z, const double* xyz, int natom, int max_atoms, const double* boys,
                 int ncart, double* S, double* T, double* V, hipStream_t stream) {
  const int n = p_end - p_begin;
  if (n <= 0) return;
  const dim3 grid_st((n + WG - 1) / WG, std::max(1, max_prim));
  const dim3 grid_v((n + WG - 1) / WG, std::max(1, (max_atoms + ONEE_ATOMS - 1) / ONEE_ATOMS));
  const dim3 block(WG);
#define CQ_ONEE(LA, LB)                                                                                 \
  do {                                                                                                   \
    hipLaunchKernelGGL((onee_kernel<LA, LB, false>), grid_st, block, 0, stream, pairs, p_begin, p_end, sd, Z, \
                       xyz, natom, boys, ncart, S, T, V);                                                \
    if (natom > 0)                                                                                       \
      hipLaunchKernelGGL((onee_kernel<LA, LB, true>), grid_v, block, 0, stream, pairs, p_begin, p_end, sd, Z,  \
                         xyz, natom, boys, ncart, S, T, V);                                              \
  } while (0)
  switch (pair_class) {
    case 0: CQ_ONEE(0, 0); break;
    case 1: CQ_ONEE(1, 0); break;
    case 2: CQ_ONEE(1, 1); break;
    case 3: CQ_ONEE(2, 0); break;
    case 4: CQ_ONEE(2, 1); break;
    case 5: CQ_ONEE(2, 2); break;
    default: break;
  }
}
