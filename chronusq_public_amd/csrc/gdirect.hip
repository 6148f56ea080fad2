// MI355X (gfx950) direct-SCF two-electron contraction G[D].
//
// Drop-in replacement for ChronusQ's AOIntegrals::newTwoEContractDirect<Op,T>
// (include/aointegrals/aointegrals_contract.h:24-194, declared include/aointegrals.h:474-480),
// exposed through the C ABI declared in include/cqgpu.h.
//
// Semantics reproduced (per input matrix X_i with type t_i and scale s_i):
//   COULOMB : AX_i += 2 s_i J[X_i],  J_{mu nu} = sum_{lam sig} (mu nu|lam sig) X_{sig lam}
//   EXCHANGE: AX_i += 2 s_i K[X_i],  K_{mu nu} = sum_{lam sig} (mu lam|nu sig) X_{lam sig}
// which is the reference's  AX += 1/2 s G  with its 4 Coulomb / 8 exchange scatter updates
// per integral (aointegrals_contract.h:145-193).  Quartets are screened exactly as the
// reference does (aointegrals_contract.h:111-131): skip when
//   max(N(s1,s4),N(s2,s4),N(s3,s4),N(s1,s3),N(s2,s3),N(s1,s2)) * Q(s1,s2) * Q(s3,s4) < tau
// evaluated left to right in FP64 on the canonical (s1 >= s2, (s3,s4) <= (s1,s2)) order,
// with N the block infinity norm maximised over all input matrices.
//
// Layout in HBM (all device-resident for the life of a context):
//   * shell pairs sorted by class (la lb), la >= lb, then by Schwarz bound descending;
//   * primitive pairs (PrimPair, 96 B AoS) referenced by pair ranges;
//   * Boys table F_m(T_i) (361 x 25 doubles);
//   * per call: Cartesian densities (row-major, ncart^2), forward J/K accumulators,
//     (plus transposed-density accumulators used only for a non-symmetric X),
//     compacted quartet task lists per class pair (uint2 {bra pair, ket pair}).
//
// The whole G[D] runs in the Cartesian AO basis: D_cart = T^T X T once per call, all
// quartets digest Cartesian integrals, G_sph = T G_cart T^T once at the end.  T is the
// block-diagonal real-solid-harmonic transform for pure shells (identity otherwise).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>
#include <exception>
#include <thread>

#include "../../include/cqgpu.h"

#include "gdirect_kernels.h"
#include "onee.h"
#include "_gen/eri_registry.inc"

static ClassKernels kClassKernels[CQ_NUM_CLASS_PAIRS];
#define MAX_MATS 8   // matrices of one contraction
// dynamic LDS of the group kernels: what the 64 KiB per workgroup leaves beside the static
// LDS (Boys table, 25992 B, the work-item slot and the members' column offsets)
static const size_t kStaticLds = BOYS_NGRID * BOYS_LDS_W * sizeof(double) + 256;
static const size_t kMaxDynLds = 64 * 1024 - kStaticLds;

// Dynamic LDS the group kernels may use on this device: the device's per-workgroup limit
// (160 KiB on gfx950) less their static LDS, once every group kernel has been allowed that
// much (hipFuncSetAttribute); 64 KiB less the static LDS if that fails.  Computed once.
static size_t device_dyn_lds() {
  static size_t v = 0;
  static std::once_flag once;
  std::call_once(once, [] {
    v = kMaxDynLds;
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      return;
    if ((size_t)lim <= 64 * 1024) return;
    const size_t want = (size_t)lim - kStaticLds;
    for (int k = 0; k < CQ_NUM_CLASS_PAIRS; ++k) {
      const ClassKernels& ck = kClassKernels[k];
      for (GroupKernel f : {ck.grp[0][0], ck.grp[0][1], ck.grp[0][2], ck.grp[1][0], ck.grp[1][1], ck.grp[1][2]}) {
        if (!f) continue;
        if (hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)want) != hipSuccess) {
          (void)hipGetLastError();
          return;
        }
      }
    }
    v = want;
  });
  return v;
}
static const int kNumClassPairs = CQ_NUM_CLASS_PAIRS;
static const bool kClassKernelsReady = (cq_register_all(kClassKernels), true);

// Per-call workspace of cq_contract_direct_device_noalloc: the Cartesian density and the four
// J/K accumulators of every density slot, then the slots' symmetry flags.  One copy per device,
// held until the process ends; every call on a device shares it, so calls that overlap on one
// device race.  Host code never uses the array's name as an address: the device address comes
// from hipGetSymbolAddress, resolved once per device at context creation (scratch_address).
#define CQ_NOALLOC_SCRATCH_BYTES (64ull << 20)
static __device__ __align__(256) double scratch[CQ_NOALLOC_SCRATCH_BYTES / sizeof(double)];

static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_ptr;   // by device id; null until resolved

// Device address of `scratch` on the current device (resolved on first use per device).
static double* scratch_address() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// ---------------------------------------------------------------------------------------
// Screening: one wavefront per uniform group; task regions (layout in gdirect_kernels.h).
// ---------------------------------------------------------------------------------------
struct CPMeta {
  int bra_start, nbra, ket_start, nket, same, bra_class, ket_class;
  int flip;        // 1: the uniform side is the bra class (bra-keyed groups)
  int grp_off;     // first group of this class pair in its table's group array
  int ngrp;
  int ub;          // union bound of a task region (swept pairs), 64 <= ub <= UNION_MAX
  int mmax;        // largest group
  long long task_off;
  long long task_cap;   // task entries of the largest batch of groups
  long long seg_off;
  long long seg_cap;
};

#define UNION_MAX 1024
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Each wavefront takes one group of the class pair (groups item0 .. item0 + nitem - 1 of it: one
// batch) and sweeps the other class's pairs of the group's molecule, sorted by Schwarz bound
// inside each primitive-count group, so a sweep stops once no member can pass; SCREEN_SW pairs
// per lane and step, their loads issued together.  Each lane tests its swept pairs against
// every member with the reference's test
//   max(N(s1,s4), N(s2,s4), N(s3,s4), N(s1,s3), N(s2,s3), N(s1,s2)) * Q(s1,s2) * Q(s3,s4) >= tau
// in the canonical order of the quartet (aointegrals_contract.h:111-131); the members it passes
// form a bit mask.  Swept pairs with a non-empty mask join the union (LDS); a full union (or the
// end of the sweep) is written out as one task region into the group's own part of the task
// list, with one segment per primitive block of the members (the group's reserved slots, then
// the batch's overflow pool).  [tau, tau_hi) is the bound window of this launch (mixed precision).
// (Measured alternatives: a whole 1024-thread workgroup per group, sharing the union, screens
// (H2O)16/cc-pVDZ in 3.4 ms instead of 1.6 ms; one swept pair per lane and step, 1.7 ms.)
#define SCREEN_SW 4
// The (class pair, batch of groups) entries one screening launch covers: entry e takes waves
// cum[e] .. cum[e + 1] - 1, one per group item0[e] + (wave - cum[e]) of class pair cp[e], whose
// plan is in table tb[e] (cps[tb], groups[tb]) and whose batch reserves nres[e] segment slots.
// One launch may hold every class pair that screens once per call (all in one batch): the
// waves of all of them share the device instead of 21 launches of a few hundred waves each.
struct ScreenList {
  const CPMeta* cps[2];
  const GroupRec* groups[2];
  int n;
  int cp[CQ_NUM_CLASS_PAIRS], tb[CQ_NUM_CLASS_PAIRS], item0[CQ_NUM_CLASS_PAIRS], nres[CQ_NUM_CLASS_PAIRS];
  int cum[CQ_NUM_CLASS_PAIRS + 1];
};

__global__ void __launch_bounds__(WG) screen_kernel(
    ScreenList sl, const int* __restrict__ uorder, const int2* __restrict__ s12, const double* __restrict__ pairQ,
    const int* __restrict__ grp, const int* __restrict__ cls_grp_off, const int* __restrict__ pair_mol,
    const int* __restrict__ cmg, int nmol, const double* __restrict__ Nmax, const double* __restrict__ Npair,
    const unsigned long long* __restrict__ nglob_bits, int nshell, double tau, double tau_hi,
    const unsigned char* __restrict__ pair_nblk, uint2* __restrict__ tasks,
    unsigned long long* __restrict__ qcounts, Segment* __restrict__ segs, unsigned* __restrict__ ovf_all,
    int count_quartets) {
  __shared__ int ucode[WG / 64][UNION_MAX];
  __shared__ unsigned short umask[WG / 64][UNION_MAX];
  __shared__ double mQ[WG / 64][MAX_GROUP], mN[WG / 64][MAX_GROUP];
  __shared__ int ms1[WG / 64][MAX_GROUP], ms2[WG / 64][MAX_GROUP], mloc[WG / 64][MAX_GROUP];
  __shared__ unsigned mcnt[WG / 64][MAX_GROUP];
  const int wid = threadIdx.x >> 6;
  const int wave = blockIdx.x * (WG / 64) + wid;
  const int lane = threadIdx.x & 63;
  if (wave >= sl.cum[sl.n]) return;
  int e = 0;
  while (wave >= sl.cum[e + 1]) ++e;
  const int cp = sl.cp[e];
  const CPMeta m = sl.cps[sl.tb[e]][cp];
  const int g = m.grp_off + sl.item0[e] + (wave - sl.cum[e]);
  const GroupRec gr = sl.groups[sl.tb[e]][g];
  const int nres = sl.nres[e];
  unsigned* const ovf = ovf_all + cp;
  const int n = gr.n, ublk = gr.ublk;
  const int uni_start = m.flip ? m.bra_start : m.ket_start;
  const int sw_start = m.flip ? m.ket_start : m.bra_start;
  const int nsw = m.flip ? m.nket : m.nbra;
  const int sw_class = m.flip ? m.ket_class : m.bra_class;
  if (lane < n) {
    const int K = uorder[gr.first + lane];
    const int2 sk = s12[K];
    mQ[wid][lane] = pairQ[K];
    mN[wid][lane] = Npair[K];
    ms1[wid][lane] = sk.x;
    ms2[wid][lane] = sk.y;
    mloc[wid][lane] = K - uni_start;
  }
  wave_sync_lds();
  // the shell every member shares (members are pairs (a, b_m) with one shell a in common):
  // its norms against a swept pair are read once per swept pair, not once per member
  int ash = ms1[wid][0];
  if (n >= 2 && ms1[wid][0] != ms1[wid][1] && ms1[wid][0] != ms2[wid][1]) ash = ms2[wid][0];
  double Qkmax = 0.0;
  int lmin = 0x7fffffff;
  for (int mm = 0; mm < n; ++mm) {
    Qkmax = fmax(Qkmax, mQ[wid][mm]);
    lmin = min(lmin, mloc[wid][mm]);
  }
  const int i_begin = m.same ? lmin : 0;
  const double nglob = __longlong_as_double((long long)*nglob_bits);
  // exactly symmetric norms: the six norms of a quartet are the two pairs' own norms and the
  // four cross norms N(k, p), k a shell of the uniform pair, p of the swept one -- read along
  // the uniform pair's rows (cache-local for the whole wavefront)
  const bool symN = nglob_bits[1 + MAX_MATS] == 0ull;
  const unsigned long long lt = (1ull << lane) - 1ull;
  // the group's own part of the batch's task list (no atomics) and its segment slots
  uint2* const tout = tasks + m.task_off + gr.task_base;
  Segment* const sout = segs + m.seg_off + gr.seg_base;
  unsigned tcur = 0;
  int scur = 0;
  int U = 0;
  unsigned long long nq = 0;
  auto flush = [&]() {
    wave_sync_lds();
    // entries per member: one per primitive block of each union pair it passed
    unsigned cl[MAX_GROUP];
#pragma unroll
    for (int mm = 0; mm < MAX_GROUP; ++mm) cl[mm] = 0u;
    for (int u = lane; u < U; u += 64) {
      const unsigned nb = pair_nblk[ucode[wid][u]];
      const unsigned msk = umask[wid][u];
#pragma unroll
      for (int mm = 0; mm < MAX_GROUP; ++mm) cl[mm] += ((msk >> mm) & 1u) * nb;
    }
    unsigned tot = 0;
#pragma unroll
    for (int mm = 0; mm < MAX_GROUP; ++mm) {
      if (mm >= n) break;
      const unsigned c = wave_sum(cl[mm]);
      if (lane == 0) mcnt[wid][mm] = c;
      tot += c;
    }
    wave_sync_lds();
    const unsigned hdr = 1u + (unsigned)n;
    const unsigned total = hdr + tot;
    // the first task region takes the group's reserved segment slots, later ones slots of the
    // batch's overflow pool (after the reserved slots)
    int sbase = 0;
    if (scur > 0) {
      unsigned o = 0;
      if (lane == 0) o = atomicAdd(ovf, (unsigned)ublk);
      sbase = nres + (int)__shfl(o, 0) - gr.seg_base;
    }
    if (lane < ublk) sout[sbase + lane] = Segment{gr.task_base + (int)tcur, (int)total, g, ublk > 1 ? lane + 1 : 0};
    uint2* out = tout + tcur;
    tcur += total;
    scur += ublk;
    if (lane == 0) out[0] = make_uint2((unsigned)n, (unsigned)U);
    if (lane < n) {
      unsigned off = hdr;
      for (int mm = 0; mm < lane; ++mm) off += mcnt[wid][mm];
      out[1 + lane] = make_uint2(mcnt[wid][lane], off);
    }
    unsigned run = hdr;
    for (int mm = 0; mm < n; ++mm) {
      for (int u0 = 0; u0 < U; u0 += 64) {
        const int u = u0 + lane;
        const unsigned code = u < U ? (unsigned)ucode[wid][u] : 0u;
        const int nb = (u < U && ((umask[wid][u] >> mm) & 1u)) ? (int)pair_nblk[code] : 0;
        for (int j = 0;; ++j) {
          const bool e = j < nb;
          const unsigned long long bal = __ballot(e);
          if (bal == 0ull) break;
          if (e) out[run + __popcll(bal & lt)] = make_uint2(code | (nb > 1 ? (unsigned)(j + 1) << 24 : 0u), (unsigned)u);
          run += (unsigned)__popcll(bal);
        }
      }
    }
    wave_sync_lds();
    U = 0;
  };
  // the swept pairs of the group's own molecule (one molecule: the whole class)
  const int g_hi = cls_grp_off[sw_class + 1];
  const int mK = pair_mol[uorder[gr.first]];
  const int gm_lo = cmg[sw_class * (nmol + 1) + mK], gm_hi = cmg[sw_class * (nmol + 1) + mK + 1];
  for (int gq = gm_lo; gq < gm_hi; ++gq) {
    const int grp_end = (gq + 1 < g_hi) ? grp[gq + 1] : nsw;
    if (grp_end <= i_begin) continue;
    for (int i0 = max(grp[gq], i_begin); i0 < grp_end; i0 += 64 * SCREEN_SW) {
      const double Qi0 = pairQ[sw_start + i0];
      if (nglob * Qkmax * Qi0 * (1.0 + 1e-12) < tau) break;
      double Qp[SCREEN_SW], Np[SCREEN_SW];
      int2 sp[SCREEN_SW];
#pragma unroll
      for (int r = 0; r < SCREEN_SW; ++r) {
        const int i = i0 + r * 64 + lane;
        const int P = sw_start + (i < grp_end ? i : i0);
        Qp[r] = pairQ[P];
        sp[r] = s12[P];
        Np[r] = symN ? Npair[P] : 0.0;
      }
      unsigned mask[SCREEN_SW];
#pragma unroll
      for (int r = 0; r < SCREEN_SW; ++r) {
        const int i = i0 + r * 64 + lane;
        mask[r] = 0u;
        if (i >= grp_end) continue;
        const long canonP = (long)sp[r].x * (sp[r].x + 1) / 2 + sp[r].y;
        // symmetric norms: N(P) and N(a, p) for the shared shell a, the same for every member
        double nbase = 0.0;
        if (symN) {
          const double* ra = Nmax + (long)ash * nshell;
          nbase = fmax(Np[r], fmax(ra[sp[r].x], ra[sp[r].y]));
        }
        for (int mm = 0; mm < n; ++mm) {
          const double Qk = mQ[wid][mm];
          if (nglob * Qk * Qi0 * (1.0 + 1e-12) < tau) continue;   // no swept pair from here passes mm
          if (m.same && i < mloc[wid][mm]) continue;               // each unordered pair once
          const int kx = ms1[wid][mm], ky = ms2[wid][mm];
          const long canonK = (long)kx * (kx + 1) / 2 + ky;
          const bool pfirst = canonP >= canonK;
          const double Q12 = pfirst ? Qp[r] : Qk, Q34 = pfirst ? Qk : Qp[r];
          double mx;
          if (symN) {
            // the six norms: N(P), N(K), and N(k, p) for k in {a, b_m}, p in P (N symmetric)
            const double* ro = Nmax + (long)(kx == ash ? ky : kx) * nshell;
            mx = fmax(fmax(nbase, mN[wid][mm]), fmax(ro[sp[r].x], ro[sp[r].y]));
          } else {
            int s1, s2, s3, s4;
            if (pfirst) { s1 = sp[r].x; s2 = sp[r].y; s3 = kx; s4 = ky; }
            else        { s1 = kx; s2 = ky; s3 = sp[r].x; s4 = sp[r].y; }
            const double n14 = Nmax[s1 * nshell + s4], n24 = Nmax[s2 * nshell + s4];
            const double n34 = Nmax[s3 * nshell + s4], n13 = Nmax[s1 * nshell + s3];
            const double n23 = Nmax[s2 * nshell + s3], n12 = Nmax[s1 * nshell + s2];
            mx = fmax(n14, fmax(n24, fmax(n34, fmax(n13, fmax(n23, n12)))));
          }
          // the reference's test, and the upper end of this launch's bound window
          const double bound = __dmul_rn(__dmul_rn(mx, Q12), Q34);
          if (!(bound < tau) && bound < tau_hi) mask[r] |= 1u << mm;
        }
      }
#pragma unroll
      for (int r = 0; r < SCREEN_SW; ++r) {
        nq += (unsigned long long)__popc(mask[r]);
        const bool has = mask[r] != 0u;
        const unsigned long long bal = __ballot(has);
        if (has) {
          const int slot = U + __popcll(bal & lt);
          ucode[wid][slot] = sw_start + i0 + r * 64 + lane;
          umask[wid][slot] = (unsigned short)mask[r];
        }
        U += __popcll(bal);
        if (U > m.ub - 64) flush();
      }
    }
  }
  if (U > 0) flush();
  // a group without survivors leaves its reserved slots empty
  if (scur == 0 && lane < ublk) sout[lane] = Segment{0, 0, g, 0};
  // (a list screened again for a later slot or batch is not counted twice)
  nq = wave_sum(nq);
  if (lane == 0 && nq > 0 && count_quartets) atomicAdd(&qcounts[cp], nq);
}

// ---------------------------------------------------------------------------------------
// Shell metadata and the solid-harmonic transform on the device
// ---------------------------------------------------------------------------------------
struct ShellMeta {
  int l, pure, sph_off, cart_off;
};

// Coefficient T(mu, a) between spherical function mu (index within shell) and Cartesian
// component a (index within shell) of a shell.
__device__ __forceinline__ double tcoef(const ShellMeta& s, int mu, int a, const double* __restrict__ tco,
                                        const int* __restrict__ tco_off) {
  if (!s.pure) return mu == a ? 1.0 : 0.0;
  const int nc = (s.l + 1) * (s.l + 2) / 2;
  return tco[tco_off[s.l] + mu * nc + a];
}

struct MatList {
  const double* X[MAX_MATS];
  int n;
};

// Block infinity norms of every (s,t) shell block, maximised over the matrices; also the
// largest |X| of the call (nglob_bits[0]) and of each matrix m (nglob_bits[1 + m]), as the bit
// patterns of non-negative doubles, so that atomicMax orders them.  Reduced across the
// wavefront first: one atomic per wavefront and value.
__device__ __forceinline__ void wave_max_bits(unsigned long long* p, double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(p, (unsigned long long)__double_as_longlong(v));
}
// nglob_bits[1 + MAX_MATS] is set when the norm matrix is not exactly symmetric (some
// N(s,t) != N(t,s)): the screening then evaluates the reference's six norms in its order.
__global__ void norm_kernel(MatList ml, int ld, const ShellMeta* __restrict__ sh, int nshell,
                            double* __restrict__ Nmax, unsigned long long* __restrict__ nglob_bits) {
  const int st = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = st < nshell * nshell;
  const int s = in ? st / nshell : 0, t = in ? st - s * nshell : 0;
  double mx = 0.0, mxT = 0.0;
  for (int m = 0; m < ml.n; ++m) {
    double mxm = 0.0;
    if (in) {
      const ShellMeta A = sh[s], B = sh[t];
      const int na = A.pure ? 2 * A.l + 1 : (A.l + 1) * (A.l + 2) / 2;
      const int nb = B.pure ? 2 * B.l + 1 : (B.l + 1) * (B.l + 2) / 2;
      const double* X = ml.X[m];
      for (int j = 0; j < nb; ++j)
        for (int i = 0; i < na; ++i) mxm = fmax(mxm, fabs(X[(long)(B.sph_off + j) * ld + A.sph_off + i]));
      if (s < t) {   // the transposed block, N(t, s)
        for (int j = 0; j < nb; ++j)
          for (int i = 0; i < na; ++i) mxT = fmax(mxT, fabs(X[(long)(A.sph_off + i) * ld + B.sph_off + j]));
      }
    }
    mx = fmax(mx, mxm);
    wave_max_bits(&nglob_bits[1 + m], mxm);
  }
  if (in) Nmax[st] = mx;
  wave_max_bits(&nglob_bits[0], mx);
  if (in && s < t && mx != mxT) atomicOr(&nglob_bits[1 + MAX_MATS], 1ull);
}

// Npair[P] = N(s1, s2) of every shell pair (the pair's own block norm, read coalesced by the
// screening).
__global__ void pair_norm_kernel(const int2* __restrict__ s12, int npair, const double* __restrict__ Nmax,
                                 int nshell, double* __restrict__ Npair) {
  const int P = blockIdx.x * blockDim.x + threadIdx.x;
  if (P >= npair) return;
  const int2 sp = s12[P];
  Npair[P] = Nmax[sp.x * nshell + sp.y];
}


// Cartesian density (row-major ncart x ncart): Dc = T^T X T.
// D_cart = T^T X T for one density; threads below n*n also flag a non-symmetric X: one whose
// X(i,j) and X(j,i) differ by more than 1e-12 of this X's own largest |X| (xmax_bits, from
// norm_kernel).  Differences below that are GEMM rounding (C C^T need not round (i,j) and
// (j,i) alike).  Digesting such an X as symmetric replaces it by (X + X^T)/2, an entrywise
// change of at most 5e-13 max|X|, so each G element moves by at most 5e-13 max|X| times the
// sum of |ERI| over the quartets that element collects; the non-symmetric path would digest
// the transposed density as well (about four times the work).
__global__ void cart_density_kernel(const double* __restrict__ X, int n, int ld, const ShellMeta* __restrict__ sh,
                                    const int* __restrict__ cart2shell, int ncart,
                                    const double* __restrict__ tco, const int* __restrict__ tco_off,
                                    const unsigned long long* __restrict__ xmax_bits,
                                    double* __restrict__ Dc, int* __restrict__ nonsym) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < (long)n * n) {
    const int i = k % n, j = k / n;
    const double tol = 1e-12 * __longlong_as_double((long long)*xmax_bits);
    if (i < j && !(fabs(X[(long)j * ld + i] - X[(long)i * ld + j]) <= tol)) atomicOr(nonsym, 1);
  }
  if (k >= (long)ncart * ncart) return;
  const int a = k / ncart, b = k - (long)a * ncart;
  const int sa = cart2shell[a], sb = cart2shell[b];
  const ShellMeta A = sh[sa], B = sh[sb];
  const int ia = a - A.cart_off, ib = b - B.cart_off;
  const int na = A.pure ? 2 * A.l + 1 : (A.l + 1) * (A.l + 2) / 2;
  const int nb = B.pure ? 2 * B.l + 1 : (B.l + 1) * (B.l + 2) / 2;
  double sk = 0.0;
  for (int mu = 0; mu < na; ++mu) {
    const double ta = tcoef(A, mu, ia, tco, tco_off);
    if (ta == 0.0) continue;
    for (int nu = 0; nu < nb; ++nu) {
      const double tb = tcoef(B, nu, ib, tco, tco_off);
      if (tb == 0.0) continue;
      sk += ta * tb * X[(long)(B.sph_off + nu) * ld + A.sph_off + mu];
    }
  }
  Dc[k] = sk;
}

// AX_i[mu + nu*ld] += f_i * sum_{a,c} T(mu,a) T(nu,c) (G1_i[a][c] + G2_i[c][a]) for every
// matrix i of the call, G2_i = G1_i unless the density was flagged non-symmetric.
struct FinList {
  double* AX[MAX_MATS];
  const double* G1[MAX_MATS];
  const double* G2alt[MAX_MATS];
  const int* nonsym[MAX_MATS];
  double f[MAX_MATS];
  int n;
};
__global__ void finalize_kernel(FinList fl, int n, int ld, const ShellMeta* __restrict__ sh,
                                const int* __restrict__ sph2shell, int ncart,
                                const double* __restrict__ tco, const int* __restrict__ tco_off) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)n * n) return;
  const int mu = k % n, nu = k / n;
  const ShellMeta A = sh[sph2shell[mu]], B = sh[sph2shell[nu]];
  const int im = mu - A.sph_off, in = nu - B.sph_off;
  const int nca = (A.l + 1) * (A.l + 2) / 2, ncb = (B.l + 1) * (B.l + 2) / 2;
  for (int i = 0; i < fl.n; ++i) {
    const double* G1 = fl.G1[i];
    const double* g2 = *fl.nonsym[i] ? fl.G2alt[i] : G1;
    double s = 0.0;
    for (int a = 0; a < nca; ++a) {
      const double ta = tcoef(A, im, a, tco, tco_off);
      if (ta == 0.0) continue;
      for (int c = 0; c < ncb; ++c) {
        const double tc = tcoef(B, in, c, tco, tco_off);
        if (tc == 0.0) continue;
        const long ia = A.cart_off + a, ic = B.cart_off + c;
        s += ta * tc * (G1[ia * ncart + ic] + g2[ic * ncart + ia]);
      }
    }
    fl.AX[i][(long)nu * ld + mu] += fl.f[i] * s;
  }
}

// Zero a list of word ranges in one launch (accumulators, counters, flags).
#define MAX_ZERO 12
struct ZeroList {
  unsigned* p[MAX_ZERO];
  long n[MAX_ZERO];   // 32-bit words
  int count;
};
__global__ void zero_kernel(ZeroList zl) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (int r = 0; r < zl.count; ++r) {
    unsigned* p = zl.p[r];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < zl.n[r]; k += stride) p[k] = 0u;
  }
}

// out[mu + nu*ld] = sum_{a,c} T(mu,a) T(nu,c) Mc[a][c]   (Cartesian row-major -> shells' basis)
__global__ void sph_transform_kernel(const double* __restrict__ Mc, double* __restrict__ out, int n, int ld,
                                     const ShellMeta* __restrict__ sh, const int* __restrict__ sph2shell, int ncart,
                                     const double* __restrict__ tco, const int* __restrict__ tco_off) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)n * n) return;
  const int mu = k % n, nu = k / n;
  const ShellMeta A = sh[sph2shell[mu]], B = sh[sph2shell[nu]];
  const int im = mu - A.sph_off, in = nu - B.sph_off;
  const int nca = (A.l + 1) * (A.l + 2) / 2, ncb = (B.l + 1) * (B.l + 2) / 2;
  double s = 0.0;
  for (int a = 0; a < nca; ++a) {
    const double ta = tcoef(A, im, a, tco, tco_off);
    if (ta == 0.0) continue;
    for (int c = 0; c < ncb; ++c) {
      const double tc = tcoef(B, in, c, tco, tco_off);
      if (tc == 0.0) continue;
      s += ta * tc * Mc[(long)(A.cart_off + a) * ncart + B.cart_off + c];
    }
  }
  out[(long)nu * ld + mu] = s;
}

// Schwarz: Q(P) = sqrt(max_{ij} |(ij|ij)|) in the shells' own (pure or Cartesian) basis,
// from the Cartesian diagonal block (ab|ab) of each pair.
__global__ void schwarz_kernel(const double* __restrict__ blocks, const int* __restrict__ pair_ids, int npair,
                               const PairInfo* __restrict__ pairs, const ShellMeta* __restrict__ sh,
                               const double* __restrict__ tco, const int* __restrict__ tco_off,
                               double* __restrict__ pairQ) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npair) return;
  const int P = pair_ids[k];
  const PairInfo pi = pairs[P];
  const ShellMeta A = sh[pi.a], B = sh[pi.b];
  const int nca = (A.l + 1) * (A.l + 2) / 2, ncb = (B.l + 1) * (B.l + 2) / 2;
  const int nab = nca * ncb;
  const int na = A.pure ? 2 * A.l + 1 : nca;
  const int nb = B.pure ? 2 * B.l + 1 : ncb;
  const double* blk = blocks + (long)k * nab * nab;
  double mx = 0.0;
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j) {
      double s = 0.0;
      for (int a = 0; a < nca; ++a) {
        const double ta = tcoef(A, i, a, tco, tco_off);
        if (ta == 0.0) continue;
        for (int b = 0; b < ncb; ++b) {
          const double tb = tcoef(B, j, b, tco, tco_off);
          if (tb == 0.0) continue;
          for (int c = 0; c < nca; ++c) {
            const double tc = tcoef(A, i, c, tco, tco_off);
            if (tc == 0.0) continue;
            for (int d = 0; d < ncb; ++d) {
              const double td = tcoef(B, j, d, tco, tco_off);
              if (td == 0.0) continue;
              s += ta * tb * tc * td * blk[(long)(a * ncb + b) * nab + c * ncb + d];
            }
          }
        }
      }
      mx = fmax(mx, fabs(s));
    }
  pairQ[P] = sqrt(mx);
}

// Work sums over one class pair's task regions, for the algorithmic FLOP model (kernel
// orientation: the bra is the (la lb) side): acc[0] += K_ab K_cd (primitive quartets),
// acc[1] += K_ab, acc[2] += sequential Cartesian->spherical transform cost, acc[3] += output
// integrals in the shells' basis.  One wavefront per segment.
__global__ void work_stats_kernel(const uint2* __restrict__ tasks, const Segment* __restrict__ segs,
                                  int nres, const unsigned* __restrict__ ovf, const GroupRec* __restrict__ groups,
                                  const int* __restrict__ uorder, int flip, const PairInfo* __restrict__ pairs,
                                  const ShellMeta* __restrict__ sh, double* __restrict__ acc) {
  const long nseg = (long)nres + *ovf;
  const int lane = threadIdx.x & 63;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const long nwaves = (long)gridDim.x * blockDim.x / 64;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (long si = wave; si < nseg; si += nwaves) {
    const Segment sg = segs[si];
    if (sg.count == 0) continue;
    const GroupRec gr = groups[sg.ket];
    const uint2* reg = tasks + sg.off;
    for (int m = 0; m < gr.n; ++m) {
      const PairInfo Up = pairs[uorder[gr.first + m]];
      int up0, up1;
      prim_range<UNI_BLOCK>(Up, (unsigned)sg.blk, up0, up1);
      const uint2 hm = reg[1 + m];
      for (int t = lane; t < (int)hm.x; t += 64) {
        const uint2 e = reg[hm.y + t];
        const PairInfo Sp = pairs[pair_of(e.x)];
        int sp0, sp1;
        prim_range(Sp, block_of(e.x), sp0, sp1);
        const PairInfo& B = flip ? Up : Sp;
        const PairInfo& K = flip ? Sp : Up;
        const double kab = flip ? up1 - up0 : sp1 - sp0, kcd = flip ? sp1 - sp0 : up1 - up0;
        s0 += kab * kcd;
        s1 += kab;
        // per-quartet terms once per quartet: entries of a split pair other than its first
        // block are further primitive blocks of the same quartet
        if (block_of(e.x) > 1 || sg.blk > 1) continue;
        const int shl[4] = {B.a, B.b, K.a, K.b};
        double dims[4];
        int l4[4], p4[4];
        for (int k = 0; k < 4; ++k) {
          const ShellMeta sm = sh[shl[k]];
          l4[k] = sm.l; p4[k] = sm.pure;
          dims[k] = (sm.l + 1) * (sm.l + 2) / 2;
        }
        double cost = 0.0;
        for (int k = 0; k < 4; ++k) {
          if (!p4[k] || l4[k] < 2) continue;
          double other = 1.0;
          for (int q = 0; q < 4; ++q) if (q != k) other *= dims[q];
          cost += 2.0 * (2 * l4[k] + 1) * dims[k] * other;
          dims[k] = 2 * l4[k] + 1;
        }
        s2 += cost;
        double nsph = 1.0;
        for (int k = 0; k < 4; ++k) nsph *= p4[k] ? (2 * l4[k] + 1) : (l4[k] + 1) * (l4[k] + 2) / 2;
        s3 += nsph;
      }
    }
  }
  s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3);
  if (lane == 0 && s0 + s1 + s2 + s3 > 0.0) {
    atomicAdd(&acc[0], s0);
    atomicAdd(&acc[1], s1);
    atomicAdd(&acc[2], s2);
    atomicAdd(&acc[3], s3);
  }
}

// =======================================================================================
// Host side
// =======================================================================================
namespace {

int ncart_l(int l) { return (l + 1) * (l + 2) / 2; }

// Real solid harmonics in the uniformly normalised Cartesian basis of a shell
// (every Cartesian component carries the x^l normalisation), m = -l..l, unit norm,
// Helgaker-Jorgensen-Olsen expansion (Molecular Electronic-Structure Theory, eq. 6.4.47-52).
std::vector<double> solid_harmonics(int l) {
  const int nc = ncart_l(l);
  std::vector<int> cx, cy, cz;
  for (int i = 0; i <= l; ++i)
    for (int j = 0; j <= i; ++j) { cx.push_back(l - i); cy.push_back(i - j); cz.push_back(j); }
  auto fac = [](int n) { double r = 1; for (int k = 2; k <= n; ++k) r *= k; return r; };
  auto bin = [&](int n, int k) { return (k < 0 || k > n) ? 0.0 : fac(n) / (fac(k) * fac(n - k)); };
  auto dfac = [](int n) { double r = 1; while (n > 1) { r *= n; n -= 2; } return r; };
  auto comp = [&](int x, int y, int z) {
    for (int k = 0; k < nc; ++k) if (cx[k] == x && cy[k] == y && cz[k] == z) return k;
    return -1;
  };
  std::vector<double> M((size_t)(2 * l + 1) * nc, 0.0);
  for (int m = -l; m <= l; ++m) {
    const int am = std::abs(m);
    const int twovm = m < 0 ? 1 : 0;   // 2 * v_m
    std::vector<double> c(nc, 0.0);
    for (int t = 0; t <= (l - am) / 2; ++t)
      for (int u = 0; u <= t; ++u)
        for (int twov = twovm; twov <= am; twov += 2) {
          // v = twov/2 ; sign (-1)^(t + v - v_m)
          const int sgn_exp = t + (twov - twovm) / 2;
          const double C = ((sgn_exp & 1) ? -1.0 : 1.0) * std::pow(0.25, t) * bin(l, t) * bin(l - t, am + t) *
                           bin(t, u) * bin(am, twov);
          const int ex = 2 * t + am - 2 * u - twov, ey = 2 * u + twov, ez = l - 2 * t - am;
          if (ex < 0) continue;
          const int k = comp(ex, ey, ez);
          if (k >= 0) c[k] += C;
        }
    // normalise in the metric <g_n|g_n'> = prod_d (n_d + n'_d - 1)!! / (2l-1)!!
    double nrm = 0.0;
    for (int a = 0; a < nc; ++a)
      for (int b = 0; b < nc; ++b) {
        const int sx = cx[a] + cx[b], sy = cy[a] + cy[b], sz = cz[a] + cz[b];
        if ((sx | sy | sz) & 1) continue;
        nrm += c[a] * c[b] * dfac(sx - 1) * dfac(sy - 1) * dfac(sz - 1) / dfac(2 * l - 1);
      }
    const double sc = 1.0 / std::sqrt(nrm);
    for (int a = 0; a < nc; ++a) M[(size_t)(m + l) * nc + a] = c[a] * sc;
  }
  return M;
}

// Boys table F_m(i * step), m < BOYS_MTAB, plus exp(-T) (long-double series + downward
// recursion).
std::vector<double> boys_table() {
  std::vector<double> tab((size_t)BOYS_NGRID * (BOYS_MTAB + 1));
  for (int i = 0; i < BOYS_NGRID; ++i) {
    const long double T = (long double)i * (long double)BOYS_STEP;
    long double F[BOYS_MTAB];
    const int mm = BOYS_MTAB - 1;
    const long double eT = std::exp(-T);
    if (T == 0.0L) {
      for (int m = 0; m < BOYS_MTAB; ++m) F[m] = 1.0L / (2 * m + 1);
    } else {
      long double term = 1.0L / (2 * mm + 1), sum = term;
      for (int k = 1; k < 4000; ++k) {
        term *= 2.0L * T / (2 * mm + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-21L) break;
      }
      F[mm] = eT * sum;
      for (int m = mm - 1; m >= 0; --m) F[m] = (2.0L * T * F[m + 1] + eT) / (2 * m + 1);
    }
    for (int m = 0; m < BOYS_MTAB; ++m) tab[(size_t)i * (BOYS_MTAB + 1) + m] = (double)F[m];
    tab[(size_t)i * (BOYS_MTAB + 1) + BOYS_MTAB] = (double)eT;
  }
  return tab;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_));           \
    }                                                                                       \
  } while (0)

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void ensure(size_t need) {
    if (need <= n && p) return;
    if (p) HIPCHK(hipFree(p));
    p = nullptr;
    n = 0;
    if (need == 0) return;
    HIPCHK(hipMalloc(&p, need * sizeof(T)));
    n = need;
  }
  void upload(const std::vector<T>& v, hipStream_t s) {
    ensure(v.size());
    if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace

struct cq_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  int nshell = 0, nsph = 0, ncart = 0, maxl = 0, npc = 0;
  std::vector<ShellMeta> shells;
  std::vector<int> cart2shell, sph2shell;
  std::vector<PairInfo> pairs;
  std::vector<double> pairQ;
  std::vector<int> class_start;      // npc + 1 (two-electron pair list)
  std::vector<int> class_start_1e;   // npc + 1 (every shell pair, one-electron integrals)
  std::vector<PrimPair> prims;
  std::vector<double> Q;             // nshell x nshell
  double tau = 1e-14;                // include/aointegrals.h:334
  int precision_mode = 0;
  double mixed_threshold = 1e-5;
  double prim_threshold = 1e-17;
  // class-pair tables: [0] every class pair ket-keyed, [1] with bra-keyed (flipped) class pairs
  std::vector<CPMeta> cps[2];
  long long task_capacity = 0;
  // task lists per class pair are at most task_budget entries (8 B each); a class pair whose
  // worst case is larger screens and digests its groups in batches:
  // batch_start[t][k] holds the class-pair-local group boundaries (nbatch + 1 of them)
  long long task_budget = 1LL << 28;
  // uniform groups: uorder lists every class's pairs (global indices, at class_start) in the
  // order groups are cut from (molecule, primitive blocks, first shell, second shell);
  // groups_h[t] the groups of every class pair of table t (CPMeta::grp_off)
  std::vector<int> uorder_h;
  std::vector<GroupRec> groups_h[2];
  DevBuf<int> d_uorder;
  DevBuf<GroupRec> d_groups[2];
  // per class pair: dynamic LDS budget of the group kernels (bytes) and an optional cap on
  // the group size and on the union bound (0: from the budget)
  std::vector<long long> tile_lds = std::vector<long long>(CQ_NUM_CLASS_PAIRS, 0);
  std::vector<int> tile_mmax = std::vector<int>(CQ_NUM_CLASS_PAIRS, 0);
  std::vector<int> tile_ub = std::vector<int>(CQ_NUM_CLASS_PAIRS, 0);
  // per class pair: the digestion table of the last contraction (work statistics)
  std::vector<int> last_table = std::vector<int>(CQ_NUM_CLASS_PAIRS, 0);
  bool plan_ok = false;   // task and segment buffers are allocated for the current plan
  double* scratch = nullptr;   // this device's address of the no-allocation workspace
  std::vector<std::vector<int>> batch_start[2];
  std::vector<std::vector<int>> batch_slots[2];   // reserved segment slots of each batch (one set per group)
  std::vector<unsigned char> nblk_h, nublk_h;
  std::vector<int> pair_mol_h;
  std::vector<unsigned long long> last_counts;
  // profiling: events around every class-pair kernel of every contraction
  bool profiling = false;
  int screen_mode = CQ_SCREEN_PER_STREAM;   // how the class pairs screened once per call are launched
  std::vector<int> launch_order;
  std::vector<std::vector<std::pair<hipEvent_t, hipEvent_t>>> prof_events, prof_screen_events;
  std::vector<double> prof_ms, prof_screen_ms;
  std::vector<long long> prof_launches, prof_screen_launches;
  std::vector<hipEvent_t> event_pool;
  size_t event_next = 0;
  hipEvent_t get_event() {
    if (event_next == event_pool.size()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      event_pool.push_back(e);
    }
    return event_pool[event_next++];
  }
  std::vector<double> tco;
  std::vector<int> tco_off;
  std::vector<double> boys_host;
  std::vector<int> prim_off_host;
  std::vector<double> alpha_host, coef_host, origin_host;
  DevBuf<int> d_prim_off;
  DevBuf<double> d_alpha, d_coef, d_origin, d_onee;
  // device
  DevBuf<ShellMeta> d_shells;
  DevBuf<int> d_cart2shell, d_sph2shell, d_tco_off, d_flags;
  DevBuf<double> d_tco, d_boys, d_pairQ, d_Nmax, d_Npair, d_work, d_mats;
  DevBuf<PairInfo> d_pairs, d_pairs_1e;
  std::vector<PairInfo> pairs_1e_h;   // host copies kept for the asynchronous uploads
  std::vector<int> rowfn_h;
  DevBuf<int2> d_s12;
  std::vector<int> grp, cls_grp_off;
  DevBuf<int> d_grp, d_cls_grp_off;
  // several molecules in one context: shell pairs and quartets only within a molecule
  int nmol = 1;
  // LDS accumulator rows: a K column (ket-keyed) or K row (bra-keyed) holds one entry per
  // Cartesian function of the swept pairs' shells, i.e. of the shells whose angular momentum
  // belongs to the swept pair class {la, lb}; lds_n_set[pair class] is the largest molecule's
  // count of such functions, and d_rowfn[pair class * ncart + loff + r] the Cartesian function
  // of row r of the molecule whose functions start at loff
  int lds_n_set[6] = {};
  DevBuf<int> d_rowfn;
  size_t max_dyn_lds = 0;             // dynamic LDS bytes a segment kernel may use
  std::vector<int> shell_mol;
  std::vector<int> cmg;               // npc x (nmol + 1): first screening group of each molecule
  DevBuf<int> d_pair_mol, d_cmg;
  // class-pair kernels run on the contraction's own stream and up to kAux more: four streams
  // in all, one per hardware queue of the process (GPU_MAX_HW_QUEUES defaults to 4; a fifth
  // stream would share a queue and serialise behind another's kernels)
  int n_streams = 4;
  static constexpr int kAux = 3;
  hipStream_t aux[kAux] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kAux] = {};
  DevBuf<PrimPair> d_prims;
  DevBuf<CPMeta> d_cps[2];
  DevBuf<uint2> d_tasks;
  DevBuf<Segment> d_segs;
  DevBuf<unsigned> d_work_q, d_ovf;
  int lds_mode = 2;   // 0 HBM atomics, 1 ket-keyed LDS segments, 2 also bra-keyed segments
  // per class pair: the digestion variant (0..2 as lds_mode, capped by it), -1 = lds_mode
  std::vector<int> class_variant = std::vector<int>(CQ_NUM_CLASS_PAIRS, -1);
  // per class pair: the auxiliary stream its kernels go to (-1: round robin in launch order)
  std::vector<int> class_stream = std::vector<int>(CQ_NUM_CLASS_PAIRS, -1);
  // per class pair: chunks per work item of the group kernels (0: the built-in choice)
  std::vector<int> class_cpi = std::vector<int>(CQ_NUM_CLASS_PAIRS, 0);
  DevBuf<unsigned long long> d_qcounts, d_nglob;
  DevBuf<unsigned char> d_pair_nblk, d_pair_nublk;
  std::string err;

  ~cq_ctx() {
    d_shells.release(); d_cart2shell.release(); d_sph2shell.release(); d_tco_off.release();
    d_uorder.release(); d_groups[0].release(); d_groups[1].release(); d_flags.release(); d_tco.release(); d_boys.release(); d_pairQ.release();
    d_Nmax.release(); d_Npair.release(); d_work.release(); d_mats.release(); d_pairs.release(); d_pairs_1e.release(); d_prims.release();
    d_cps[0].release(); d_cps[1].release(); d_tasks.release(); d_segs.release(); d_work_q.release(); d_ovf.release(); d_qcounts.release(); d_nglob.release(); d_pair_nblk.release(); d_pair_nublk.release();
    d_prim_off.release(); d_alpha.release(); d_coef.release(); d_origin.release(); d_onee.release();
    for (auto e : event_pool) (void)hipEventDestroy(e);
    for (int k = 0; k < kAux; ++k) {
      if (aux[k]) (void)hipStreamDestroy(aux[k]);
      if (ev_join[k]) (void)hipEventDestroy(ev_join[k]);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    d_s12.release(); d_grp.release(); d_cls_grp_off.release(); d_pair_mol.release(); d_cmg.release();
    if (stream && own_stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }

int grid_for(long n, int block) { return (int)std::max<long>(1, (n + block - 1) / block); }

int persistent_grid() {
  static int g = 0;
  if (!g) {
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    g = std::max(1, ncu) * 8;
  }
  return g;
}

// Rows of a class pair's LDS K accumulators: the functions of the swept side's shells (the bra
// pair class for the ket-keyed kernels, the ket pair class for the bra-keyed ones).
int lds_rows(const cq_ctx* c, const ClassKernels& ck, int variant) {
  return c->lds_n_set[variant == 2 ? ck.ki : ck.bi];
}

// Dynamic LDS bytes of a group kernel: ket-keyed with LDS K columns (variant 1), with HBM K
// atomics (variant 0), or bra-keyed (variant 2), for groups of up to mmax members and unions of
// up to ub swept pairs.
size_t group_lds(const cq_ctx* c, const ClassKernels& ck, int variant, int mmax, int ub) {
  const size_t ncd = (size_t)ck.nc * ck.nd, nab = (size_t)ck.na * ck.nb;
  const size_t lds_n = (size_t)lds_rows(c, ck, variant);
  // (the last term: one int per union slot, its swept pair)
  if (variant == 2)
    return (((size_t)ck.na + (size_t)mmax * ck.nb) * lds_n + mmax * nab + ub * ncd) * sizeof(double) + ub * sizeof(int);
  const size_t cols = variant == 1 ? ((size_t)ck.nc + (size_t)mmax * ck.nd) * lds_n : 0;
  return (cols + mmax * ncd + ub * nab) * sizeof(double) + ub * sizeof(int);
}

// Tile of a class pair for one digestion variant: the union bound (J slots of the swept side,
// 64 .. 1024 pairs) and the largest group (at most the cap, default 1) that keeps the kernel's
// dynamic LDS within the budget (default 48 KiB).  mmax = 0: the variant does not fit the
// device even with single-pair groups.
// Default union bound per class pair (kernel orientation), from a sweep of 256..1024 with
// single-pair groups on (H2O)16/cc-pVDZ, tau = 1e-12 (per-class kernel + screening time): the
// bound trades fewer K-column flushes (large) against more work items to balance (small).
// Built-in tile of each class pair: the largest group (uniform pairs sharing a shell) and the
// union bound (swept pairs per task region), the group size then lowered until the LDS budget
// holds it.  One molecule (measured on (H2O)16/cc-pVDZ with per-member wavefront units and
// LDS rows for the swept shells only): groups of 4-16 members pay where a swept pair's J slot
// would otherwise be flushed to HBM once per uniform pair ((ps|ps) kernel 0.82 -> 0.37 ms from
// 1 to 14 members, (ss|ps) 0.95 -> 0.61 ms at 4; G[D] 4.23 -> 4.14 ms for the whole table,
// although the screening launch's one wavefront per group then runs longer: 0.48 -> 0.62 ms
// serialised, hidden behind the other streams' kernels); larger groups cost where a class
// pair has few work items ((dp|pp) 0.21 ms with one member, 0.29 with 8); elsewhere
// single-pair groups are as fast.  Several molecules in one context (measured on 256 alkanes and water clusters,
// 6-31G*, groups of 4096 functions): a uniform pair has few partners inside its small
// molecule, so large groups fill the wavefronts (serialised kernel time 306 -> 273 ms).
// cpi: chunks of the split (high angular momentum) class pairs per work item, 0 = all of them.
// A work item's chunks share its wavefronts and its one LDS flush; more chunks per item means
// fewer items, which starves the device when a class pair has few task regions.  (H2O)16:
// (dp|dp) 0.29 -> 0.17 ms at 3, (dd|dd) 0.14 -> 0.08 ms; the batch: G[D] 152 -> 118 ms with all.
struct TileDefault { const char* name; int max_group, ub, cpi; };
static const TileDefault kTileDefaults[] = {
    {"ssss", 4, 256, 1}, {"ssps", 4, 256, 1}, {"psps", 16, 256, 1}, {"sspp", 4, 512, 1}, {"ppps", 12, 512, 1},
    {"pppp", 3, 256, 1}, {"ssds", 4, 256, 1}, {"psds", 8, 256, 1}, {"ppds", 1, 256, 1}, {"dsds", 1, 256, 1},
    {"ssdp", 1, 512, 1}, {"dpps", 1, 256, 1}, {"dppp", 1, 256, 3}, {"dpds", 1, 256, 1}, {"dpdp", 1, 512, 3},
    {"ddss", 1, 256, 1}, {"ddps", 1, 512, 1}, {"ddpp", 1, 512, 1}, {"ddds", 1, 512, 1}, {"dddp", 1, 512, 2},
    {"dddd", 1, 256, 3}};
static const TileDefault kTileDefaultsMulti[] = {
    {"ssss", 4, 512, 0}, {"ssps", 16, 512, 0}, {"psps", 16, 1024, 0}, {"sspp", 4, 512, 0}, {"ppps", 16, 1024, 0},
    {"pppp", 16, 512, 0}, {"ssds", 16, 256, 0}, {"psds", 16, 512, 0}, {"ppds", 8, 256, 0}, {"dsds", 16, 512, 0},
    {"ssdp", 16, 1024, 0}, {"dpps", 2, 1024, 0}, {"dppp", 16, 256, 0}, {"dpds", 16, 256, 0}, {"dpdp", 16, 256, 0},
    {"ddss", 2, 1024, 0}, {"ddps", 2, 1024, 0}, {"ddpp", 16, 512, 0}, {"ddds", 16, 256, 0}, {"dddp", 4, 1024, 0},
    {"dddd", 4, 512, 0}};

static_assert(sizeof(kTileDefaults) / sizeof(TileDefault) == CQ_NUM_CLASS_PAIRS, "one tile per class pair");
static_assert(sizeof(kTileDefaultsMulti) / sizeof(TileDefault) == CQ_NUM_CLASS_PAIRS, "one tile per class pair");

const TileDefault* tile_default(const cq_ctx* c, const ClassKernels& ck) {
  const TileDefault* tab = c->nmol > 1 ? kTileDefaultsMulti : kTileDefaults;
  for (int i = 0; i < CQ_NUM_CLASS_PAIRS; ++i)
    if (std::strcmp(tab[i].name, ck.name) == 0) return &tab[i];
  return nullptr;
}

int default_union_bound(const cq_ctx* c, const ClassKernels& ck, size_t jside) {
  if (const TileDefault* t = tile_default(c, ck)) return t->ub;
  return (int)std::min<size_t>(UNION_MAX, std::max<size_t>(64, (16 * 1024) / (jside * sizeof(double)) / 64 * 64));
}

// Chunks per work item of class pair k: the caller's choice, else the built-in table's; always
// a divisor of the chunk count.
int class_chunks_per_item(const cq_ctx* c, int k) {
  const int nch = kClassKernels[k].nchunk;
  if (c->class_cpi[k] > 0) return c->class_cpi[k];
  const TileDefault* t = tile_default(c, kClassKernels[k]);
  int d = (!t || t->cpi == 0) ? nch : std::min(t->cpi, nch);
  while (d > 1 && nch % d != 0) --d;
  return std::max(1, d);
}

void class_tile(const cq_ctx* c, int k, int variant, int& mmax, int& ub) {
  const ClassKernels& ck = kClassKernels[k];
  const size_t jside = variant == 2 ? (size_t)ck.nc * ck.nd : (size_t)ck.na * ck.nb;
  ub = default_union_bound(c, ck, jside);
  if (c->tile_ub[k] > 0) ub = std::max(64, std::min(c->tile_ub[k], UNION_MAX));
  const size_t budget = c->tile_lds[k] > 0 ? (size_t)c->tile_lds[k] : (size_t)48 * 1024;
  const size_t limit = std::min(budget, c->max_dyn_lds);
  // a union bound whose J slots alone overflow the device's LDS is lowered (multiples of 64)
  while (ub > 64 && group_lds(c, ck, variant, 1, ub) > c->max_dyn_lds) ub -= 64;
  mmax = 0;
  const TileDefault* t = tile_default(c, ck);
  const int cap = c->tile_mmax[k] > 0 ? std::min(c->tile_mmax[k], MAX_GROUP) : (t ? t->max_group : 1);
  for (int mm = cap; mm >= 1; --mm)
    if (group_lds(c, ck, variant, mm, ub) <= limit) { mmax = mm; break; }
  if (mmax == 0 && group_lds(c, ck, variant, 1, ub) <= c->max_dyn_lds) mmax = 1;
}

// Class-pair metadata, uniform groups, task-list capacities and batches of groups.  Table 1 keys
// the groups of a class pair on the bra pairs when the bra class has at most half as many pairs
// as the ket class (few bras survive per ket, many kets per bra) and the bra-keyed LDS rows fit.
// A group's worst case is every primitive block of every swept pair of its molecule (same
// class: from its first member on) in every member's list, and a header per flushed task
// region.  Consecutive groups form one batch while their worst cases sum to at most
// task_budget, so a class pair's task list never exceeds max(task_budget, one group's worst
// case) whatever the density; a batch is screened and digested before the next one.
void plan_tasks(cq_ctx* c) {
  c->plan_ok = false;
  const int nm = c->nmol;
  const std::vector<unsigned char>& nblk = c->nblk_h;
  const std::vector<unsigned char>& nublk = c->nublk_h;
  const std::vector<int>& pair_mol = c->pair_mol_h;
  for (int t = 0; t < 2; ++t) {
    c->cps[t].clear();
    c->groups_h[t].clear();
    c->batch_start[t].assign(kNumClassPairs, std::vector<int>());
    c->batch_slots[t].assign(kNumClassPairs, std::vector<int>());
  }
  const long long budget = c->task_budget;
  long long off = 0, seg_total[2] = {0, 0};
  for (int k = 0; k < kNumClassPairs; ++k) {
    const ClassKernels& ck = kClassKernels[k];
    CPMeta m{};
    m.bra_start = c->class_start[ck.bi];
    m.nbra = c->class_start[ck.bi + 1] - c->class_start[ck.bi];
    m.ket_start = c->class_start[ck.ki];
    m.nket = c->class_start[ck.ki + 1] - c->class_start[ck.ki];
    m.same = ck.i == ck.j;
    m.bra_class = ck.bi;
    m.ket_class = ck.ki;
    // swept-side pairs and primitive blocks per molecule
    std::vector<long long> bp_m(nm, 0), kp_m(nm, 0), bb_m(nm, 0), kb_m(nm, 0);
    for (int i = 0; i < m.nbra; ++i) { bp_m[pair_mol[m.bra_start + i]] += 1; bb_m[pair_mol[m.bra_start + i]] += nblk[m.bra_start + i]; }
    for (int i = 0; i < m.nket; ++i) { kp_m[pair_mol[m.ket_start + i]] += 1; kb_m[pair_mol[m.ket_start + i]] += nblk[m.ket_start + i]; }
    int mm2 = 0, ub2 = 64;
    if (ck.grp[0][2]) class_tile(c, k, 2, mm2, ub2);
    const int flip = (!m.same && ck.grp[0][2] && mm2 > 0 && 2LL * m.nbra <= m.nket) ? 1 : 0;
    CPMeta mt[2] = {m, m};
    long long region = 0;
    for (int t = 0; t < 2; ++t) {
      CPMeta& mc = mt[t];
      mc.flip = t == 1 ? flip : 0;
      int mmax = 0, ub = 64;
      if (mc.flip) {
        mmax = mm2;
        ub = ub2;
      } else {
        class_tile(c, k, 1, mmax, ub);
        if (mmax == 0) class_tile(c, k, 0, mmax, ub);   // K columns do not fit: HBM K atomics
        if (mmax == 0) throw std::invalid_argument(std::string("class pair ") + ck.name + ": no tile fits the device's LDS");
      }
      mc.mmax = mmax;
      mc.ub = ub;
      const int ucls = mc.flip ? m.bra_class : m.ket_class;
      const int uni_start = mc.flip ? m.bra_start : m.ket_start;
      // groups: runs of the class's uniform order with one molecule, one primitive-block count
      // and one first shell, at most mmax long
      mc.grp_off = (int)c->groups_h[t].size();
      const int u0 = c->class_start[ucls], u1 = c->class_start[ucls + 1];
      for (int x = u0; x < u1;) {
        const int P0 = c->uorder_h[x];
        int y = x + 1;
        while (y < u1 && y - x < mmax) {
          const int P = c->uorder_h[y];
          if (pair_mol[P] != pair_mol[P0] || nublk[P] != nublk[P0] || c->pairs[P].a != c->pairs[P0].a) break;
          ++y;
        }
        c->groups_h[t].push_back(GroupRec{x, y - x, nublk[P0], 0, 0, 0, 0, 0});
        x = y;
      }
      mc.ngrp = (int)c->groups_h[t].size() - mc.grp_off;
      // suffix counts of the swept class (same class: from the group's first member on)
      std::vector<long long> suf_p, suf_b;
      if (m.same) {
        suf_p.assign(m.nbra + 1, 0);
        suf_b.assign(m.nbra + 1, 0);
        for (int i = m.nbra - 1; i >= 0; --i) {
          const bool brk = i + 1 < m.nbra && pair_mol[m.bra_start + i] != pair_mol[m.bra_start + i + 1];
          suf_p[i] = 1 + (brk ? 0 : suf_p[i + 1]);
          suf_b[i] = nblk[m.bra_start + i] + (brk ? 0 : suf_b[i + 1]);
        }
      }
      std::vector<int>& bs = c->batch_start[t][k];
      std::vector<int>& bslots = c->batch_slots[t][k];
      bs.assign(1, 0);
      bslots.clear();
      // segments of a batch: each group's reserved slots (its first task region), then the
      // overflow pool for the groups' further task regions
      long long cap = 0, cur = 0, seg_cap = 0, seg_cur = 0, seg_ovf = 0;
      auto close = [&](int gi) {
        bs.push_back(gi);
        bslots.push_back((int)seg_cur);
        cap = std::max(cap, cur);
        seg_cap = std::max(seg_cap, seg_cur + seg_ovf);
        cur = 0;
        seg_cur = 0;
        seg_ovf = 0;
      };
      for (int gi = 0; gi < mc.ngrp; ++gi) {
        GroupRec& gr = c->groups_h[t][mc.grp_off + gi];
        const int mol = pair_mol[c->uorder_h[gr.first]];
        long long S, Bk;
        if (m.same) {
          int lmin = 0x7fffffff;
          for (int q = 0; q < gr.n; ++q) lmin = std::min(lmin, c->uorder_h[gr.first + q] - uni_start);
          S = suf_p[lmin];
          Bk = suf_b[lmin];
        } else {
          S = mc.flip ? kp_m[mol] : bp_m[mol];
          Bk = mc.flip ? kb_m[mol] : bb_m[mol];
        }
        const long long nflush = S / (ub - 63) + 1;
        const long long w = (long long)gr.n * Bk + nflush * (1 + gr.n);
        if (gi > bs.back() && cur + w > budget) close(gi);
        // task offsets within a list are 32-bit (Segment::off, GroupRec::task_base)
        if (cur + w > 0x7fffffffLL || seg_cur + seg_ovf + nflush * gr.ublk > 0x7fffffffLL)
          throw std::invalid_argument("one batch's task list exceeds 2^31 entries");
        gr.task_base = (int)cur;
        gr.seg_base = (int)seg_cur;
        gr.nslot = gr.ublk;
        cur += w;
        seg_cur += gr.ublk;
        seg_ovf += (nflush - 1) * gr.ublk;
      }
      close(mc.ngrp);
      mc.task_cap = cap;
      mc.seg_cap = seg_cap;
      mc.seg_off = seg_total[t];
      seg_total[t] += seg_cap;
      region = std::max(region, cap);
    }
    for (int t = 0; t < 2; ++t) {
      mt[t].task_off = off;
      c->cps[t].push_back(mt[t]);
    }
    off += region;
  }
  c->task_capacity = off;
  // one call may take table 0 for some class pairs and table 1 for others (per-class digestion
  // variants): table 1's segments live after table 0's
  for (auto& mt : c->cps[1]) mt.seg_off += seg_total[0];
  for (int t = 0; t < 2; ++t) {
    c->d_cps[t].upload(c->cps[t], c->stream);
    c->d_groups[t].upload(c->groups_h[t], c->stream);
  }
  c->d_tasks.ensure((size_t)std::max<long long>(1, off));
  c->d_segs.ensure((size_t)std::max<long long>(1, seg_total[0] + seg_total[1]));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->plan_ok = true;
}

void build_context(cq_ctx* c, int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                   const double* coef, const double* origin, const int* mol = nullptr) {
  HIPCHK(hipGetDevice(&c->device));
  HIPCHK(hipStreamCreate(&c->stream));
  for (int k = 0; k < cq_ctx::kAux; ++k) {
    HIPCHK(hipStreamCreate(&c->aux[k]));
    HIPCHK(hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming));
  }
  HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  c->nshell = nshell;
  c->shells.resize(nshell);
  std::vector<int> poff(nshell + 1, 0);
  int sph = 0, cart = 0;
  for (int s = 0; s < nshell; ++s) {
    if (l[s] < 0 || l[s] > 2) throw std::invalid_argument("shell angular momentum above d is not supported");
    if (nprim[s] < 1) throw std::invalid_argument("shell with no primitives");
    c->shells[s] = ShellMeta{l[s], pure[s] ? 1 : 0, sph, cart};
    const int nsz = pure[s] ? 2 * l[s] + 1 : ncart_l(l[s]);
    for (int k = 0; k < nsz; ++k) c->sph2shell.push_back(s);
    for (int k = 0; k < ncart_l(l[s]); ++k) c->cart2shell.push_back(s);
    sph += nsz;
    cart += ncart_l(l[s]);
    poff[s + 1] = poff[s] + nprim[s];
    c->maxl = std::max(c->maxl, l[s]);
  }
  c->nsph = sph;
  c->ncart = cart;
  // Cartesian J/K accumulators of ncart^2 elements, addressed with 32-bit offsets in the kernels
  if (cart > 46340) throw std::invalid_argument("a context holds at most 46340 Cartesian functions");
  c->max_dyn_lds = device_dyn_lds();
  // molecules: consecutive shells, ids 0, 1, ... in order (nullptr: one molecule)
  c->shell_mol.assign(nshell, 0);
  if (mol) {
    if (nshell > 0 && mol[0] != 0) throw std::invalid_argument("molecule ids must start at 0");
    for (int s = 0; s < nshell; ++s) {
      if (s > 0 && mol[s] != mol[s - 1] && mol[s] != mol[s - 1] + 1)
        throw std::invalid_argument("molecule ids must be consecutive and nondecreasing over the shells");
      c->shell_mol[s] = mol[s];
    }
  }
  c->nmol = nshell > 0 ? c->shell_mol[nshell - 1] + 1 : 1;
  std::vector<int> mol_cart0(c->nmol + 1, cart);
  for (int s = nshell - 1; s >= 0; --s) mol_cart0[c->shell_mol[s]] = c->shells[s].cart_off;
  c->npc = 6;   // ss ps pp ds dp dd
  // LDS rows per pair class: shell s's first row among the functions of its molecule's shells
  // with angular momentum in {la, lb} (crow, -1 outside the set), and the inverse map
  std::vector<int> crow((size_t)c->npc * nshell, -1);
  {
    std::vector<int>& rowfn = c->rowfn_h;
    rowfn.assign((size_t)c->npc * std::max(cart, 1), 0);
    for (int la = 0; la <= 2; ++la)
      for (int lb = 0; lb <= la; ++lb) {
        const int pc = pair_class(la, lb);
        c->lds_n_set[pc] = 0;
        int r = 0;
        for (int sh = 0; sh < nshell; ++sh) {
          if (sh > 0 && c->shell_mol[sh] != c->shell_mol[sh - 1]) r = 0;
          if (l[sh] != la && l[sh] != lb) continue;
          crow[(size_t)pc * nshell + sh] = r;
          const int nf = (l[sh] + 1) * (l[sh] + 2) / 2;
          for (int i = 0; i < nf; ++i)
            rowfn[(size_t)pc * cart + mol_cart0[c->shell_mol[sh]] + r + i] = c->shells[sh].cart_off + i;
          r += nf;
          c->lds_n_set[pc] = std::max(c->lds_n_set[pc], r);
        }
      }
    c->d_rowfn.upload(rowfn, c->stream);
  }
  // solid-harmonic tables for l = 0..2
  c->tco_off.assign(5, 0);
  for (int L = 0; L <= 4; ++L) {
    c->tco_off[L] = (int)c->tco.size();
    if (L <= 2) {
      auto M = solid_harmonics(L);
      c->tco.insert(c->tco.end(), M.begin(), M.end());
    }
  }
  // shell pairs (canonical s1 >= s2) and their primitive pairs, built by up to 16 host threads
  // over ranges of s1 (equal shares of the canonical pairs) and joined in s1 order, so the
  // list is the same as a serial build's; a pair's primitive pairs are allprim[off, off + n)
  struct HostPair { PairInfo info; int cls; long off; int n; };
  std::vector<HostPair> hp;
  std::vector<PrimPair> allprim;
  const double pi = 3.14159265358979323846;
  const double q_num = 2.0 * std::pow(pi, 2.5), q_half = std::sqrt(q_num);
  auto build_range = [&](int s1_lo, int s1_hi, std::vector<HostPair>& out, std::vector<PrimPair>& prims) {
    for (int s1 = s1_lo; s1 < s1_hi; ++s1)
      for (int s2 = 0; s2 <= s1; ++s2) {
        if (c->shell_mol[s2] != c->shell_mol[s1]) continue;   // pairs only within a molecule
        int a = s1, b = s2;
        if (l[s2] > l[s1]) std::swap(a, b);
        HostPair h;
        h.info.a = a; h.info.b = b; h.info.s1 = s1; h.info.s2 = s2;
        h.info.a0 = c->shells[a].cart_off; h.info.b0 = c->shells[b].cart_off;
        h.info.loff = mol_cart0[c->shell_mol[s1]];
        h.cls = pair_class(l[a], l[b]);
        h.info.ar = crow[(size_t)h.cls * nshell + a];
        h.info.br = crow[(size_t)h.cls * nshell + b];
        const double* A = origin + 3 * a;
        const double* B = origin + 3 * b;
        const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
        h.off = (long)prims.size();
        for (int i = 0; i < nprim[a]; ++i)
          for (int j = 0; j < nprim[b]; ++j) {
            const double ea = alpha[poff[a] + i], eb = alpha[poff[b] + j];
            const double p = ea + eb;
            const double Kab = std::exp(-ea * eb / p * AB2) * coef[poff[a] + i] * coef[poff[b] + j];
            // primitive-pair bound sqrt((pp|pp)) of the s-type part
            const double qp = std::fabs(Kab) * std::sqrt(q_num / (p * p * std::sqrt(2.0 * p)));
            if (qp < c->prim_threshold) continue;
            PrimPair pp;
            pp.p = p;
            pp.K = q_half * Kab / p;   // sqrt(2 pi^(5/2)) of the quartet prefactor per pair
            for (int k = 0; k < 3; ++k) {
              pp.P[k] = (ea * A[k] + eb * B[k]) / p;
              pp.PA[k] = pp.P[k] - A[k];
              pp.PB[k] = pp.P[k] - B[k];
            }
            pp.o2p = 0.5 / p;
            prims.push_back(pp);
          }
        h.n = (int)((long)prims.size() - h.off);
        out.push_back(h);
      }
  };
  {
    const int nthr = nshell < 64 ? 1 : (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<int> cut(nthr + 1, nshell);
    cut[0] = 0;
    for (int t = 1; t < nthr; ++t)   // s1 < cut[t] holds about t / nthr of the canonical pairs
      cut[t] = std::max(cut[t - 1], std::min(nshell, (int)std::lround(nshell * std::sqrt((double)t / nthr))));
    std::vector<std::vector<HostPair>> parts(nthr);
    std::vector<std::vector<PrimPair>> pparts(nthr);
    std::vector<std::exception_ptr> errs(nthr);
    std::vector<std::thread> pool;
    for (int t = 1; t < nthr; ++t)
      pool.emplace_back([&, t] {
        try { build_range(cut[t], cut[t + 1], parts[t], pparts[t]); } catch (...) { errs[t] = std::current_exception(); }
      });
    try { build_range(cut[0], cut[1], parts[0], pparts[0]); } catch (...) { errs[0] = std::current_exception(); }
    for (auto& th : pool) th.join();
    for (auto& e : errs)
      if (e) std::rethrow_exception(e);
    size_t total = 0, ptotal = 0;
    for (int t = 0; t < nthr; ++t) { total += parts[t].size(); ptotal += pparts[t].size(); }
    hp.reserve(total);
    allprim.reserve(ptotal);
    for (int t = 0; t < nthr; ++t) {
      const long base = (long)allprim.size();
      allprim.insert(allprim.end(), pparts[t].begin(), pparts[t].end());
      for (HostPair h : parts[t]) { h.off += base; hp.push_back(h); }
    }
  }
  // the one-electron integrals cover every shell pair; the two-electron lists drop the pairs
  // without a primitive pair above the threshold (Q = 0: none of their quartets can pass)
  {
    std::vector<PairInfo> all;
    std::vector<int> cls_all;
    for (auto& h : hp) { all.push_back(h.info); cls_all.push_back(h.cls); }
    std::vector<int> order(all.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cls_all[x] < cls_all[y]; });
    std::vector<PairInfo>& sorted_all = c->pairs_1e_h;
    sorted_all.clear();
    sorted_all.reserve(all.size());
    c->class_start_1e.assign(c->npc + 1, 0);
    for (int k : order) { sorted_all.push_back(all[k]); c->class_start_1e[cls_all[k] + 1]++; }
    for (int k = 0; k < c->npc; ++k) c->class_start_1e[k + 1] += c->class_start_1e[k];
    c->d_pairs_1e.upload(sorted_all, c->stream);
  }
  hp.erase(std::remove_if(hp.begin(), hp.end(), [](const HostPair& h) { return h.n == 0; }), hp.end());
  // group by class (stable)
  std::stable_sort(hp.begin(), hp.end(), [](const HostPair& x, const HostPair& y) { return x.cls < y.cls; });
  c->class_start.assign(c->npc + 1, 0);
  for (auto& h : hp) c->class_start[h.cls + 1]++;
  for (int k = 0; k < c->npc; ++k) c->class_start[k + 1] += c->class_start[k];
  // the pair list in hp's order; every pair's primitive pairs stay where the build put them
  // (allprim, uploaded once), the pair pointing at its range
  auto flatten = [&]() {
    c->pairs.resize(hp.size());
    for (size_t k = 0; k < hp.size(); ++k) {
      PairInfo pi_ = hp[k].info;
      pi_.p0 = (int)hp[k].off;
      pi_.p1 = (int)(hp[k].off + hp[k].n);
      c->pairs[k] = pi_;
    }
  };
  flatten();
  const int npair = (int)c->pairs.size();
  // device uploads
  c->prim_off_host = poff;
  c->alpha_host.assign(alpha, alpha + poff[nshell]);
  c->coef_host.assign(coef, coef + poff[nshell]);
  c->origin_host.assign(origin, origin + 3 * nshell);
  c->d_prim_off.upload(c->prim_off_host, c->stream);
  c->d_alpha.upload(c->alpha_host, c->stream);
  c->d_coef.upload(c->coef_host, c->stream);
  c->d_origin.upload(c->origin_host, c->stream);
  c->d_shells.upload(c->shells, c->stream);
  c->d_cart2shell.upload(c->cart2shell, c->stream);
  c->d_sph2shell.upload(c->sph2shell, c->stream);
  c->d_tco.upload(c->tco, c->stream);
  c->d_tco_off.upload(c->tco_off, c->stream);
  c->boys_host = boys_table();   // stays alive until the stream is synchronised
  c->d_boys.upload(c->boys_host, c->stream);
  c->d_pairs.upload(c->pairs, c->stream);
  if (allprim.size() > (size_t)INT32_MAX) throw std::invalid_argument("too many primitive pairs");
  c->prims.swap(allprim);
  c->d_prims.upload(c->prims, c->stream);
  c->d_pairQ.ensure(npair);
  // Schwarz bounds: diagonal quartets (P|P) per class on the device, the classes on the
  // auxiliary streams; a pair with more than PRIM_BLOCK primitive pairs is split into
  // (block i | block j) tasks that add into the pair's output slot
  {
    std::vector<uint2> tasks;
    std::vector<int> slots, ids;
    struct Launch { int k; long t0, nt; int s0, ns, nab; size_t out0; };
    std::vector<Launch> launches;
    size_t out_total = 0;
    for (int k = 0; k < kNumClassPairs; ++k) {
      const ClassKernels& ck = kClassKernels[k];
      if (ck.i != ck.j) continue;
      const int X = ck.i;
      const int n = c->class_start[X + 1] - c->class_start[X];
      if (n == 0) continue;
      int la = 0;
      while ((la + 1) * (la + 2) / 2 <= X) ++la;
      const int lb = X - la * (la + 1) / 2;
      const int nab = ncart_l(la) * ncart_l(lb);
      Launch L{k, (long)tasks.size(), 0, (int)ids.size(), n, nab, out_total};
      for (int q = 0; q < n; ++q) {
        const int P = c->class_start[X] + q;
        const int np_ = c->pairs[P].p1 - c->pairs[P].p0;
        if (np_ > 64 * PRIM_BLOCK) throw std::invalid_argument("more than 512 primitive pairs in one shell pair");
        const int nb = np_ > PRIM_BLOCK ? (np_ + PRIM_BLOCK - 1) / PRIM_BLOCK : 1;
        for (int i = 0; i < nb; ++i)
          for (int j = 0; j < nb; ++j) {
            const unsigned ci = (unsigned)P | (nb > 1 ? (unsigned)(i + 1) << 24 : 0u);
            const unsigned cj = (unsigned)P | (nb > 1 ? (unsigned)(j + 1) << 24 : 0u);
            tasks.push_back(make_uint2(ci, cj));
            slots.push_back(q);
          }
        ids.push_back(P);
      }
      L.nt = (long)tasks.size() - L.t0;
      out_total += (size_t)n * nab * nab;
      launches.push_back(L);
    }
    DevBuf<uint2> dt;
    DevBuf<int> dslot, di;
    DevBuf<double> dblk;
    dt.upload(tasks, c->stream);
    dslot.upload(slots, c->stream);
    di.upload(ids, c->stream);
    dblk.ensure(std::max<size_t>(1, out_total));
    HIPCHK(hipMemsetAsync(dblk.p, 0, std::max<size_t>(1, out_total) * sizeof(double), c->stream));
    HIPCHK(hipEventRecord(c->ev_fork, c->stream));
    for (int a = 0; a < cq_ctx::kAux; ++a) HIPCHK(hipStreamWaitEvent(c->aux[a], c->ev_fork, 0));
    int li = 0;
    for (const Launch& L : launches) {
      const ClassKernels& ck = kClassKernels[L.k];
      hipStream_t sa = c->aux[li++ % cq_ctx::kAux];
      // slots are per class: the slot array entries count from the class's first pair
      hipLaunchKernelGGL(ck.block, dim3(grid_for(L.nt, WG)), dim3(WG), 0, sa, dt.p + L.t0, L.nt, c->d_pairs.p,
                         c->d_prims.p, c->d_boys.p, dblk.p + L.out0, L.nab * L.nab, dslot.p + L.t0);
      HIPCHK(hipGetLastError());
      hipLaunchKernelGGL(schwarz_kernel, dim3(grid_for(L.ns, 128)), dim3(128), 0, sa, dblk.p + L.out0, di.p + L.s0,
                         L.ns, c->d_pairs.p, c->d_shells.p, c->d_tco.p, c->d_tco_off.p, c->d_pairQ.p);
      HIPCHK(hipGetLastError());
    }
    for (int a = 0; a < cq_ctx::kAux; ++a) HIPCHK(hipStreamSynchronize(c->aux[a]));
    HIPCHK(hipStreamSynchronize(c->stream));
    dt.release(); dslot.release(); di.release(); dblk.release();
  }
  c->pairQ.resize(npair);
  HIPCHK(hipMemcpy(c->pairQ.data(), c->d_pairQ.p, npair * sizeof(double), hipMemcpyDeviceToHost));
  // sort within class by primitive-pair count (descending), then Schwarz bound (descending):
  // equal counts side by side keep the lanes of a wavefront in step, and inside each count
  // group the bound order lets screening stop early
  std::vector<int> order(npair);
  std::iota(order.begin(), order.end(), 0);
  for (int X = 0; X < c->npc; ++X)
    std::stable_sort(order.begin() + c->class_start[X], order.begin() + c->class_start[X + 1], [&](int x, int y) {
      const int mx = c->shell_mol[hp[x].info.s1], my = c->shell_mol[hp[y].info.s1];
      if (mx != my) return mx < my;
      const int nx = hp[x].n, ny = hp[y].n;
      if (nx != ny) return nx > ny;
      return c->pairQ[x] > c->pairQ[y];
    });
  std::vector<HostPair> hp2;
  std::vector<double> q2;
  hp2.reserve(npair);
  for (int k : order) { hp2.push_back(hp[k]); q2.push_back(c->pairQ[k]); }
  hp.swap(hp2);
  c->pairQ.swap(q2);
  flatten();
  c->d_pairs.upload(c->pairs, c->stream);
  c->d_pairQ.upload(c->pairQ, c->stream);
  // primitive-count groups per class: grp[cls_grp_off[X] .. cls_grp_off[X+1]] are the group
  // starts (class-local), closed by the class size
  c->cls_grp_off.assign(c->npc + 1, 0);
  c->grp.clear();
  const int nm = c->nmol;
  c->cmg.assign((size_t)c->npc * (nm + 1), 0);
  std::vector<int> pair_mol(npair);
  for (int k = 0; k < npair; ++k) pair_mol[k] = c->shell_mol[c->pairs[k].s1];
  for (int X = 0; X < c->npc; ++X) {
    c->cls_grp_off[X] = (int)c->grp.size();
    int* cm = &c->cmg[(size_t)X * (nm + 1)];
    int next_m = 0;   // molecules whose first group is not yet known
    for (int k = c->class_start[X]; k < c->class_start[X + 1]; ++k) {
      const int np_ = c->pairs[k].p1 - c->pairs[k].p0;
      const bool new_mol = k == c->class_start[X] || pair_mol[k] != pair_mol[k - 1];
      if (new_mol || np_ != c->pairs[k - 1].p1 - c->pairs[k - 1].p0) {
        while (new_mol && next_m <= pair_mol[k]) cm[next_m++] = (int)c->grp.size();
        c->grp.push_back(k - c->class_start[X]);
      }
    }
    while (next_m <= nm) cm[next_m++] = (int)c->grp.size();
  }
  c->cls_grp_off[c->npc] = (int)c->grp.size();
  c->d_grp.upload(c->grp, c->stream);
  c->d_cls_grp_off.upload(c->cls_grp_off, c->stream);
  c->d_cmg.upload(c->cmg, c->stream);
  c->d_pair_mol.upload(pair_mol, c->stream);
  c->pair_mol_h = pair_mol;
  {
    std::vector<int2> s12(npair);
    for (int k = 0; k < npair; ++k) s12[k] = make_int2(c->pairs[k].s1, c->pairs[k].s2);
    c->d_s12.upload(s12, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->Q.assign((size_t)nshell * nshell, 0.0);
  for (int k = 0; k < npair; ++k) {
    const PairInfo& pi_ = c->pairs[k];
    c->Q[(size_t)pi_.s1 * nshell + pi_.s2] = c->pairQ[k];
    c->Q[(size_t)pi_.s2 * nshell + pi_.s1] = c->pairQ[k];
  }
  // primitive blocks per pair (1: unsplit) as the per-lane side and as the uniform side
  std::vector<unsigned char> nblk(npair), nublk(npair);
  for (int k = 0; k < npair; ++k) {
    const int np = c->pairs[k].p1 - c->pairs[k].p0;
    if (np > 64 * PRIM_BLOCK) throw std::invalid_argument("more than 512 primitive pairs in one shell pair");
    nblk[k] = (unsigned char)(np > PRIM_BLOCK ? (np + PRIM_BLOCK - 1) / PRIM_BLOCK : 1);
    nublk[k] = (unsigned char)(np > UNI_BLOCK ? (np + UNI_BLOCK - 1) / UNI_BLOCK : 1);
  }
  if ((unsigned)npair > PAIR_MASK) throw std::invalid_argument("too many shell pairs");
  c->d_pair_nblk.upload(nblk, c->stream);
  c->d_pair_nublk.upload(nublk, c->stream);
  c->nblk_h.swap(nblk);
  c->nublk_h.swap(nublk);
  // uniform order of every class: by molecule, primitive blocks, first and second shell (groups
  // are cut from runs sharing the first three)
  c->uorder_h.resize(npair);
  std::iota(c->uorder_h.begin(), c->uorder_h.end(), 0);
  for (int X = 0; X < c->npc; ++X)
    std::stable_sort(c->uorder_h.begin() + c->class_start[X], c->uorder_h.begin() + c->class_start[X + 1],
                     [&](int x, int y) {
                       if (c->pair_mol_h[x] != c->pair_mol_h[y]) return c->pair_mol_h[x] < c->pair_mol_h[y];
                       if (c->nublk_h[x] != c->nublk_h[y]) return c->nublk_h[x] < c->nublk_h[y];
                       if (c->pairs[x].a != c->pairs[y].a) return c->pairs[x].a < c->pairs[y].a;
                       return c->pairs[x].b < c->pairs[y].b;
                     });
  c->d_uorder.upload(c->uorder_h, c->stream);
  plan_tasks(c);
  c->scratch = scratch_address();
  // WORK_QUEUES work queues per class pair and slot pass; one overflow-slot counter per class pair
  c->d_work_q.ensure((size_t)kNumClassPairs * MAX_MATS * WORK_QUEUES);
  c->d_ovf.ensure(kNumClassPairs);
  c->d_qcounts.ensure(2 * kNumClassPairs);   // surviving quartets: FP64 window, then FP32 window
  c->d_nglob.ensure(2 + MAX_MATS);   // largest |X| of the call, of each matrix, norm asymmetry flag
  c->d_Nmax.ensure((size_t)nshell * nshell);
  c->d_Npair.ensure(std::max(1, npair));
  c->last_counts.assign(kNumClassPairs, 0);
  c->prof_events.assign(kNumClassPairs, {});
  c->prof_screen_events.assign(kNumClassPairs + 1, {});   // the last: the fused screening launch
  c->prof_screen_ms.assign(kNumClassPairs + 1, 0.0);
  c->prof_screen_launches.assign(kNumClassPairs + 1, 0);
  c->launch_order.resize(kNumClassPairs);
  std::iota(c->launch_order.begin(), c->launch_order.end(), 0);
  std::stable_sort(c->launch_order.begin(), c->launch_order.end(),
                   [](int x, int y) { return kClassKernels[x].L < kClassKernels[y].L; });
  c->prof_ms.assign(kNumClassPairs, 0.0);
  c->prof_launches.assign(kNumClassPairs, 0);
  HIPCHK(hipStreamSynchronize(c->stream));
}

struct Slot {
  const double* X;
  int do_j = 0, do_k = 0;
};

// The contraction on device-resident X/AX (column-major, leading dimension ld).
// With noalloc, the per-call workspace is the device's static scratch array (nothing is
// allocated: a call whose workspace exceeds it is rejected) and no profiling events are recorded.
void contract_device(cq_ctx* c, int nmat, const double* const* X, double* const* AX, const int* types,
                     const double* scales, int n, int ld, hipStream_t s, bool noalloc = false) {
  if (nmat < 1 || nmat > MAX_MATS) throw std::invalid_argument("nmat must be in 1..8");
  if (!c->plan_ok) throw std::runtime_error("no task buffers: the last cq_set_task_budget could not allocate a plan");
  if (n != c->nsph) throw std::invalid_argument("matrix dimension does not match the basis");
  if (ld < n) throw std::invalid_argument("leading dimension smaller than the matrix dimension");
  for (int i = 0; i < nmat; ++i) {
    if (types[i] != CQ_COULOMB && types[i] != CQ_EXCHANGE)
      throw std::invalid_argument("only COULOMB and EXCHANGE contractions are supported by the direct path");
    if (!X[i] || !AX[i]) throw std::invalid_argument("null matrix pointer");
  }
  // unique densities (screening uses all of them, scale == 0 included)
  std::vector<const double*> uniq;
  for (int i = 0; i < nmat; ++i)
    if (std::find(uniq.begin(), uniq.end(), X[i]) == uniq.end()) uniq.push_back(X[i]);
  if ((int)uniq.size() > MAX_MATS) throw std::invalid_argument("too many distinct input matrices");
  std::vector<Slot> slots;
  std::vector<int> mat_slot(nmat, -1);
  for (int i = 0; i < nmat; ++i) {
    if (scales[i] == 0.0) continue;
    int k = 0;
    while (k < (int)slots.size() && slots[k].X != X[i]) ++k;
    if (k == (int)slots.size()) slots.push_back(Slot{X[i]});
    if (types[i] == CQ_COULOMB) slots[k].do_j = 1; else slots[k].do_k = 1;
    mat_slot[i] = k;
  }
  const int ns = c->nshell;
  const long nc2 = (long)c->ncart * c->ncart;
  const int nslot = (int)slots.size();
  // workspace: per slot Dc, GJ, GJ2, GK, GK2, then one flag per slot
  const size_t work_words = (size_t)std::max(1, nslot) * 5 * nc2;
  double* work = nullptr;
  int* flags = nullptr;
  if (noalloc) {
    const size_t need = work_words * sizeof(double) + (size_t)std::max(1, nslot) * sizeof(int);
    if (need > CQ_NOALLOC_SCRATCH_BYTES)
      throw std::invalid_argument("the no-allocation workspace (64 MiB) is too small for this call: it needs " +
                                  std::to_string(need) + " bytes (5 ncart^2 doubles per distinct density)");
    if (!c->scratch) throw std::runtime_error("no-allocation workspace not resolved on this device");
    work = c->scratch;
    flags = (int*)(c->scratch + work_words);
  } else {
    c->d_work.ensure(work_words);
    c->d_flags.ensure(std::max(1, nslot));
    work = c->d_work.p;
    flags = c->d_flags.p;
  }
  const bool profiling = c->profiling && !noalloc;
  // 1. one launch zeroes the accumulators, counters and flags
  ZeroList zl{};
  auto zero = [&](void* p, size_t bytes) {
    zl.p[zl.count] = (unsigned*)p;
    zl.n[zl.count] = (long)(bytes / 4);
    ++zl.count;
  };
  zero(c->d_nglob.p, (2 + MAX_MATS) * sizeof(unsigned long long));
  zero(flags, std::max(1, nslot) * sizeof(int));
  zero(c->d_qcounts.p, 2 * kNumClassPairs * sizeof(unsigned long long));
  zero(c->d_work_q.p, (size_t)kNumClassPairs * MAX_MATS * WORK_QUEUES * sizeof(unsigned));
  zero(c->d_ovf.p, kNumClassPairs * sizeof(unsigned));
  // the whole workspace (the Cartesian densities in it are rewritten below)
  zero(work, (size_t)std::max(1, nslot) * 5 * nc2 * sizeof(double));
  hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(WG), 0, s, zl);
  HIPCHK(hipGetLastError());
  // 2. block norms over every input matrix (scale 0 included)
  MatList ml{};
  ml.n = (int)uniq.size();
  for (int i = 0; i < ml.n; ++i) ml.X[i] = uniq[i];
  hipLaunchKernelGGL(norm_kernel, dim3(grid_for((long)ns * ns, WG)), dim3(WG), 0, s, ml, ld, c->d_shells.p, ns,
                     c->d_Nmax.p, c->d_nglob.p);
  HIPCHK(hipGetLastError());
  {
    const int np = (int)c->pairs.size();
    hipLaunchKernelGGL(pair_norm_kernel, dim3(grid_for(np, WG)), dim3(WG), 0, s, c->d_s12.p, np, c->d_Nmax.p, ns,
                       c->d_Npair.p);
    HIPCHK(hipGetLastError());
  }
  if (nslot == 0) return;
  // 3. Cartesian densities and symmetry flags; per slot: [Dc | GJ | GJ2 | GK | GK2], ncart^2 each
  for (int k = 0; k < nslot; ++k) {
    double* base = work + (size_t)k * 5 * nc2;
    hipLaunchKernelGGL(cart_density_kernel, dim3(grid_for(std::max<long>(nc2, (long)n * n), WG)), dim3(WG), 0, s,
                       slots[k].X, n, ld, c->d_shells.p, c->d_cart2shell.p, c->ncart, c->d_tco.p, c->d_tco_off.p,
                       c->d_nglob.p + 1 + (std::find(uniq.begin(), uniq.end(), slots[k].X) - uniq.begin()), base,
                       flags + k);
    HIPCHK(hipGetLastError());
  }
  // digestion variant per class pair (the lds_mode, or the class's own choice below it)
  auto variant_of = [&](int k) {
    const int v = c->class_variant[k] < 0 ? c->lds_mode : std::min(c->class_variant[k], c->lds_mode);
    return v;
  };
  // mixed precision: a second round screens the bound window [tau, threshold) into the same
  // task buffers and runs it with FP32 integrals (one density slot)
  const bool mixed = c->precision_mode == CQ_PREC_MIXED && nslot == 1 && c->mixed_threshold > c->tau;
  const int nround = mixed ? 2 : 1;
  for (int round = 0; round < nround; ++round) {
    const bool fp32 = round == 1;
    const double lo = mixed && !fp32 ? std::max(c->tau, c->mixed_threshold) : c->tau;
    const double hi = fp32 ? c->mixed_threshold : INFINITY;
    if (round > 0) {
      // the task lists, segments and work queues are rebuilt; quartet counts accumulate
      ZeroList zr{};
      auto zero2 = [&](void* p, size_t bytes) {
        zr.p[zr.count] = (unsigned*)p;
        zr.n[zr.count] = (long)(bytes / 4);
        ++zr.count;
      };
      zero2(c->d_work_q.p, (size_t)kNumClassPairs * MAX_MATS * WORK_QUEUES * sizeof(unsigned));
      zero2(c->d_ovf.p, kNumClassPairs * sizeof(unsigned));
      hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(WG), 0, s, zr);
      HIPCHK(hipGetLastError());
    }
  // screening launches: the class pairs' entries over groups (their plans' tables), into the
  // task regions of this round's bound window [lo, hi)
  auto screen_list_base = [&]() {
    ScreenList sl{};
    sl.cps[0] = c->d_cps[0].p;
    sl.cps[1] = c->d_cps[1].p;
    sl.groups[0] = c->d_groups[0].p;
    sl.groups[1] = c->d_groups[1].p;
    return sl;
  };
  auto add_screen_entry = [&](ScreenList& sl, int k, int tb, int g0, int ng, int nres) {
    sl.cp[sl.n] = k;
    sl.tb[sl.n] = tb;
    sl.item0[sl.n] = g0;
    sl.nres[sl.n] = nres;
    sl.cum[sl.n + 1] = sl.cum[sl.n] + ng;
    ++sl.n;
  };
  auto launch_screen = [&](const ScreenList& sl, hipStream_t st, std::vector<std::pair<hipEvent_t, hipEvent_t>>& evs,
                           bool count) {
    const int waves = sl.cum[sl.n];
    if (waves == 0) return;
    hipEvent_t q0 = nullptr, q1 = nullptr;
    if (profiling) {
      q0 = c->get_event();
      q1 = c->get_event();
      HIPCHK(hipEventRecord(q0, st));
    }
    hipLaunchKernelGGL(screen_kernel, dim3(grid_for(waves, WG / 64)), dim3(WG), 0, st, sl, c->d_uorder.p,
                       c->d_s12.p, c->d_pairQ.p, c->d_grp.p, c->d_cls_grp_off.p, c->d_pair_mol.p, c->d_cmg.p,
                       c->nmol, c->d_Nmax.p, c->d_Npair.p, c->d_nglob.p, ns, lo, hi, c->d_pair_nblk.p,
                       c->d_tasks.p, c->d_qcounts.p + (fp32 ? kNumClassPairs : 0), c->d_segs.p, c->d_ovf.p,
                       count ? 1 : 0);
    HIPCHK(hipGetLastError());
    if (profiling) {
      HIPCHK(hipEventRecord(q1, st));
      evs.push_back({q0, q1});
    }
  };
  // stream of each class pair's kernels (0: s, i > 0: aux[i - 1]): its explicit choice or
  // round robin in launch order (highest angular momentum first: the latency-bound kernels
  // start early) over the class pairs that have work; the last, incomplete round runs from the
  // last stream down, so its (lowest angular momentum, longest) kernels do not all land on the
  // streams that already hold one kernel more ((H2O)16/cc-pVDZ: 21 class pairs, 4.11 -> 4.03 ms)
  int stream_of[kNumClassPairs];
  {
    int nwork = 0;
    for (int k = 0; k < kNumClassPairs; ++k) {
      const CPMeta& m = c->cps[variant_of(k) >= 2 ? 1 : 0][k];
      stream_of[k] = -1;
      if (m.nbra == 0 || m.nket == 0 || m.ngrp == 0) continue;
      if (c->class_stream[k] < 0) ++nwork;
    }
    const int full = nwork - nwork % c->n_streams;   // kernels in complete rounds
    int launched = 0;
    for (int kk = kNumClassPairs - 1; kk >= 0; --kk) {
      const int k = c->launch_order[kk];
      const CPMeta& m = c->cps[variant_of(k) >= 2 ? 1 : 0][k];
      if (m.nbra == 0 || m.nket == 0 || m.ngrp == 0) continue;
      if (c->class_stream[k] >= 0) {
        stream_of[k] = c->class_stream[k] % c->n_streams;
      } else {
        const int i = launched++;
        stream_of[k] = i < full ? i % c->n_streams : c->n_streams - 1 - (i - full);
      }
    }
  }
  auto stream_at = [&](int si) { return si == 0 ? s : c->aux[si - 1]; };
  // class pairs screened once per call (one batch) share screening launches: with
  // CQ_SCREEN_PER_STREAM each stream screens its first such class pair alone (high angular
  // momentum: few quartets, a short launch), runs that pair's kernel, and screens the rest of
  // its class pairs behind it, so the kernels start early and the bulk of the screening runs
  // beside other streams' kernels; with CQ_SCREEN_ONE one launch on s before the fork screens
  // them all
  std::vector<char> fused(kNumClassPairs, 0);
  std::vector<int> once[cq_ctx::kAux + 1];   // per stream, in launch order
  if (c->screen_mode != CQ_SCREEN_PER_CLASS)
    for (int kk = kNumClassPairs - 1; kk >= 0; --kk) {
      const int k = c->launch_order[kk];
      if (stream_of[k] < 0 || c->batch_start[variant_of(k) >= 2 ? 1 : 0][k].size() != 2) continue;
      once[stream_of[k]].push_back(k);
      fused[k] = 1;
    }
  auto screen_of = [&](const int* ks, int n) {
    ScreenList sl = screen_list_base();
    for (int i = 0; i < n; ++i) {
      const int k = ks[i], tb = variant_of(k) >= 2 ? 1 : 0;
      const std::vector<int>& bs = c->batch_start[tb][k];
      add_screen_entry(sl, k, tb, bs[0], bs[1] - bs[0], c->batch_slots[tb][k][0]);
    }
    return sl;
  };
  if (c->screen_mode == CQ_SCREEN_ONE) {
    std::vector<int> all;
    for (int si = 0; si < c->n_streams; ++si) all.insert(all.end(), once[si].begin(), once[si].end());
    launch_screen(screen_of(all.data(), (int)all.size()), s, c->prof_screen_events[kNumClassPairs], true);
  }
  HIPCHK(hipEventRecord(c->ev_fork, s));
  for (int a = 0; a < cq_ctx::kAux; ++a) HIPCHK(hipStreamWaitEvent(c->aux[a], c->ev_fork, 0));
  if (c->screen_mode == CQ_SCREEN_PER_STREAM)
    for (int si = 0; si < c->n_streams; ++si)
      if (!once[si].empty())
        launch_screen(screen_of(once[si].data(), 1), stream_at(si), c->prof_screen_events[kNumClassPairs], true);
  // 4. per class pair, on its stream: screening into its task regions (unless screened above),
  //    then ERI + J/K digestion, one density slot per pass (the lists are screened once and kept
  //    for every slot when the class pair is one batch); joined back into s before the
  //    finalisation
  for (int s0 = 0; s0 < nslot; ++s0) {
    JKArgs jk{};
    jk.ncart = c->ncart;
    double* base = work + (size_t)s0 * 5 * nc2;
    jk.D[0] = base;
    jk.GJ[0] = base + nc2;
    jk.GJ2[0] = base + 2 * nc2;
    jk.GK[0] = base + 3 * nc2;
    jk.GK2[0] = base + 4 * nc2;
    jk.do_j[0] = slots[s0].do_j;
    jk.do_k[0] = slots[s0].do_k;
    jk.nonsym = flags + s0;
    for (int kk = kNumClassPairs - 1; kk >= 0; --kk) {
      const int k = c->launch_order[kk];
      const ClassKernels& ck = kClassKernels[k];
      const int var = variant_of(k);
      const int tb = var >= 2 ? 1 : 0;
      const CPMeta& m = c->cps[tb][k];
      c->last_table[k] = tb;
      if (stream_of[k] < 0) continue;
      hipStream_t sk = stream_at(stream_of[k]);
      // kernel: bra-keyed groups, ket-keyed with LDS K columns where they fit, else HBM K atomics
      const int kv = m.flip ? 2 : (var >= 1 && group_lds(c, ck, 1, m.mmax, m.ub) <= c->max_dyn_lds) ? 1 : 0;
      const size_t lds_bytes = group_lds(c, ck, kv, m.mmax, m.ub);
      if (m.mmax < 1 || lds_bytes > c->max_dyn_lds)
        throw std::runtime_error(std::string("class pair ") + ck.name + ": the planned tile does not fit the device's LDS");
      GroupKernel kern = ck.grp[fp32 ? 1 : 0][kv];
      const int wgs = kv == 2 ? WG : ck.seg_wg;
      JKArgs jkk = jk;   // the LDS rows of this class pair's swept side
      jkk.lds_n = lds_rows(c, ck, kv);
      jkk.rowfn = c->d_rowfn.p + (size_t)(kv == 2 ? ck.ki : ck.bi) * c->ncart;
      const int cpi = class_chunks_per_item(c, k);
      unsigned* wq = c->d_work_q.p + ((size_t)s0 * kNumClassPairs + k) * WORK_QUEUES;
      // batches of groups, in order on this class pair's stream: a class pair in one batch
      // screens once and keeps its lists for every slot pass; a batched one clears its counters
      // and screens again before each batch
      const std::vector<int>& bs = c->batch_start[tb][k];
      const int nbatch = (int)bs.size() - 1;
      for (int b = 0; b < nbatch; ++b) {
        const int g0 = bs[b], ng = bs[b + 1] - bs[b];
        const int nslots = c->batch_slots[tb][k][b];
        if (nbatch > 1) {
          ZeroList zb{};
          zb.p[0] = wq;
          zb.n[0] = WORK_QUEUES;
          zb.p[1] = c->d_ovf.p + k;
          zb.n[1] = 1;
          zb.count = 2;
          hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, sk, zb);
          HIPCHK(hipGetLastError());
        }
        if ((s0 == 0 && !fused[k]) || nbatch > 1) {
          ScreenList sl = screen_list_base();
          add_screen_entry(sl, k, tb, g0, ng, nslots);
          launch_screen(sl, sk, c->prof_screen_events[k], s0 == 0);
        }
        // profiling times the ERI + digestion kernel of each batch (not the screen)
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (profiling) {
          e0 = c->get_event();
          e1 = c->get_event();
          HIPCHK(hipEventRecord(e0, sk));
        }
        // (the overflow regions are known on the device only: the grid covers the segment capacity)
        const int sgrid = (int)std::min<long long>((long long)persistent_grid() * WG / wgs,
                                                   m.seg_cap * (ck.nchunk / cpi));
        hipLaunchKernelGGL(kern, dim3(std::max(1, sgrid)), dim3(wgs), lds_bytes, sk, c->d_tasks.p + m.task_off,
                           c->d_segs.p + m.seg_off, nslots, c->d_ovf.p + k, wq, c->d_pairs.p, c->d_prims.p,
                           c->d_boys.p, jkk,
                           c->d_groups[tb].p, c->d_uorder.p, m.mmax, m.ub, cpi);
        HIPCHK(hipGetLastError());
        if (profiling) {
          HIPCHK(hipEventRecord(e1, sk));
          c->prof_events[k].push_back({e0, e1});
        }
      }
      // behind the first class pair of its stream, the screening of the stream's other
      // once-per-call class pairs
      const std::vector<int>& ol = once[stream_of[k]];
      if (s0 == 0 && c->screen_mode == CQ_SCREEN_PER_STREAM && ol.size() > 1 && ol[0] == k)
        launch_screen(screen_of(ol.data() + 1, (int)ol.size() - 1), sk, c->prof_screen_events[kNumClassPairs], true);
    }
  }
  for (int a = 0; a < cq_ctx::kAux; ++a) {
    HIPCHK(hipEventRecord(c->ev_join[a], c->aux[a]));
    HIPCHK(hipStreamWaitEvent(s, c->ev_join[a], 0));
  }
  }
  // 5. back to the shells' basis and accumulate, every matrix in one launch:
  //    AX += 1/2 scale (G + G'^T); COULOMB G = GJ, G' = (nonsym ? GJ2 : GJ); EXCHANGE with GK/GK2
  FinList fl{};
  for (int i = 0; i < nmat; ++i) {
    const int k = mat_slot[i];
    if (k < 0) continue;
    double* base = work + (size_t)k * 5 * nc2;
    fl.AX[fl.n] = AX[i];
    fl.G1[fl.n] = base + (types[i] == CQ_COULOMB ? 1 : 3) * nc2;
    fl.G2alt[fl.n] = base + (types[i] == CQ_COULOMB ? 2 : 4) * nc2;
    fl.nonsym[fl.n] = flags + k;
    fl.f[fl.n] = 0.5 * scales[i];
    ++fl.n;
  }
  if (fl.n > 0) {
    hipLaunchKernelGGL(finalize_kernel, dim3(grid_for((long)n * n, WG)), dim3(WG), 0, s, fl, n, ld, c->d_shells.p,
                       c->d_sph2shell.p, c->ncart, c->d_tco.p, c->d_tco_off.p);
    HIPCHK(hipGetLastError());
  }
}

template <typename F>
int guarded(cq_ctx* c, F&& f) {
  if (!c) return CQ_ERR_INVALID;
  try {
    f();
    c->err.clear();
    return CQ_OK;
  } catch (const std::invalid_argument& e) {
    c->err = e.what();
    return CQ_ERR_INVALID;
  } catch (const std::bad_alloc& e) {
    c->err = "out of memory";
    return CQ_ERR_NOMEM;
  } catch (const std::exception& e) {
    c->err = e.what();
    return CQ_ERR_DEVICE;
  }
}

thread_local std::string g_create_error;

// S, T, V of the context's shell pairs into ld-strided n x n outputs (column-major, symmetric);
// with atom_mol (nuclei grouped by molecule, ids nondecreasing) each molecule's pairs feel only
// its own nuclei.  Synchronous.
void one_electron(cq_ctx* c, int natom, const double* Z, const double* xyz, const int* atom_mol, double* S,
                  double* T, double* V, int ld, void* stream) {
  if (natom < 0 || (natom > 0 && (!Z || !xyz)) || !S || !T || !V || ld < c->nsph)
    throw std::invalid_argument("invalid one-electron arguments");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  // nuclei per molecule
  std::vector<int> mol_atom0(c->nmol + 1, 0);
  int max_atoms = natom;
  if (atom_mol) {
    for (int a = 0; a < natom; ++a) {
      if (atom_mol[a] < 0 || atom_mol[a] >= c->nmol || (a > 0 && atom_mol[a] < atom_mol[a - 1]))
        throw std::invalid_argument("atom_molecule must be nondecreasing molecule ids of the context");
      mol_atom0[atom_mol[a] + 1]++;
    }
    max_atoms = 0;
    for (int m = 0; m < c->nmol; ++m) {
      max_atoms = std::max(max_atoms, mol_atom0[m + 1]);
      mol_atom0[m + 1] += mol_atom0[m];
    }
  }
  const long nc2 = (long)c->ncart * c->ncart;
  const int na1 = std::max(1, natom);
  c->d_onee.ensure((size_t)3 * nc2 + 4 * (size_t)na1);
  double* Sc = c->d_onee.p;
  double* Tc = Sc + nc2;
  double* Vc = Tc + nc2;
  double* dZ = Vc + nc2;
  double* dxyz = dZ + na1;
  std::vector<double> zx((size_t)4 * na1, 0.0);
  for (int a = 0; a < natom; ++a) {
    zx[a] = Z[a];
    for (int k = 0; k < 3; ++k) zx[na1 + 3 * a + k] = xyz[3 * a + k];
  }
  HIPCHK(hipMemcpyAsync(dZ, zx.data(), zx.size() * sizeof(double), hipMemcpyHostToDevice, s));
  OneEShellData sd{c->d_prim_off.p, c->d_alpha.p, c->d_coef.p, c->d_origin.p};
  DevBuf<int> d_smol, d_mat0;
  if (atom_mol) {
    d_smol.upload(c->shell_mol, s);
    d_mat0.upload(mol_atom0, s);
    sd.shell_mol = d_smol.p;
    sd.mol_atom0 = d_mat0.p;
  }
  ZeroList zl{};
  zl.p[0] = (unsigned*)Sc;
  zl.n[0] = 3 * nc2 * 2;
  zl.count = 1;
  hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(WG), 0, s, zl);
  int max_prim = 1;
  for (size_t k = 0; k + 1 < c->prim_off_host.size(); ++k)
    max_prim = std::max(max_prim, c->prim_off_host[k + 1] - c->prim_off_host[k]);
  // the pair classes' kernels (small grids, latency-bound) side by side on s and the context's
  // auxiliary streams; every class adds into its own blocks
  HIPCHK(hipEventRecord(c->ev_fork, s));
  for (int a = 0; a < cq_ctx::kAux; ++a) HIPCHK(hipStreamWaitEvent(c->aux[a], c->ev_fork, 0));
  for (int X = 0; X < c->npc; ++X) {
    const int si = X % (cq_ctx::kAux + 1);
    launch_onee(X, c->d_pairs_1e.p, c->class_start_1e[X], c->class_start_1e[X + 1], sd, max_prim, dZ, dxyz, natom,
                max_atoms, c->d_boys.p, c->ncart, Sc, Tc, Vc, si == 0 ? s : c->aux[si - 1]);
  }
  HIPCHK(hipGetLastError());
  for (int a = 0; a < cq_ctx::kAux; ++a) {
    HIPCHK(hipEventRecord(c->ev_join[a], c->aux[a]));
    HIPCHK(hipStreamWaitEvent(s, c->ev_join[a], 0));
  }
  double* outs[3] = {S, T, V};
  for (int k = 0; k < 3; ++k)
    hipLaunchKernelGGL(sph_transform_kernel, dim3(grid_for((long)c->nsph * c->nsph, WG)), dim3(WG), 0, s,
                       c->d_onee.p + k * nc2, outs[k], c->nsph, ld, c->d_shells.p, c->d_sph2shell.p, c->ncart,
                       c->d_tco.p, c->d_tco_off.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));   // the host staging vectors and device tables are released on return
  d_smol.release();
  d_mat0.release();
}

// Cartesian multipole integrals (orders 1 .. order) about `origin`, transformed to the shells'
// basis: 3, 9 or 19 column-major n x n matrices (leading dimension ld) back to back in `out`.
// Synchronous.
void multipole(cq_ctx* c, int order, const double* origin, double* out, int ld, void* stream) {
  if (order < 1 || order > 3 || !out || ld < c->nsph)
    throw std::invalid_argument("multipole: order must be 1, 2 or 3, with an output of leading dimension >= n");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const int nops = order == 1 ? 3 : (order == 2 ? 9 : 19);
  const double C[3] = {origin ? origin[0] : 0.0, origin ? origin[1] : 0.0, origin ? origin[2] : 0.0};
  const long nc2 = (long)c->ncart * c->ncart;
  c->d_onee.ensure((size_t)nops * nc2);
  double* Mc = c->d_onee.p;
  ZeroList zl{};
  zl.p[0] = (unsigned*)Mc;
  zl.n[0] = nops * nc2 * 2;
  zl.count = 1;
  hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(WG), 0, s, zl);
  OneEShellData sd{c->d_prim_off.p, c->d_alpha.p, c->d_coef.p, c->d_origin.p};
  int max_prim = 1;
  for (size_t k = 0; k + 1 < c->prim_off_host.size(); ++k)
    max_prim = std::max(max_prim, c->prim_off_host[k + 1] - c->prim_off_host[k]);
  for (int X = 0; X < c->npc; ++X)
    launch_multipole(X, c->d_pairs_1e.p, c->class_start_1e[X], c->class_start_1e[X + 1], sd, max_prim, C, order,
                     c->ncart, Mc, s);
  HIPCHK(hipGetLastError());
  for (int k = 0; k < nops; ++k)
    hipLaunchKernelGGL(sph_transform_kernel, dim3(grid_for((long)c->nsph * c->nsph, WG)), dim3(WG), 0, s,
                       Mc + k * nc2, out + (long)k * ld * c->nsph, c->nsph, ld, c->d_shells.p, c->d_sph2shell.p,
                       c->ncart, c->d_tco.p, c->d_tco_off.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
}

}  // namespace

// =======================================================================================
// C ABI
// =======================================================================================
extern "C" {

int cq_ctx_create_batch(cq_ctx_t** out, int nshell, const int* l, const int* pure, const int* nprim,
                        const double* alpha, const double* coef, const double* origin, const int* molecule) {
  if (!out) return CQ_ERR_INVALID;
  *out = nullptr;
  if (nshell < 1 || !l || !pure || !nprim || !alpha || !coef || !origin) {
    g_create_error = "invalid shell description";
    return CQ_ERR_INVALID;
  }
  cq_ctx* c = new (std::nothrow) cq_ctx();
  if (!c) return CQ_ERR_NOMEM;
  try {
    build_context(c, nshell, l, pure, nprim, alpha, coef, origin, molecule);
  } catch (const std::invalid_argument& e) {
    g_create_error = e.what();
    delete c;
    return CQ_ERR_INVALID;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    delete c;
    return CQ_ERR_DEVICE;
  }
  *out = c;
  return CQ_OK;
}

int cq_ctx_destroy(cq_ctx_t* c) {
  if (!c) return CQ_ERR_INVALID;
  (void)hipStreamSynchronize(c->stream);
  delete c;
  return CQ_OK;
}

int cq_ctx_create(cq_ctx_t** out, int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                  const double* coef, const double* origin) {
  return cq_ctx_create_batch(out, nshell, l, pure, nprim, alpha, coef, origin, nullptr);
}

const char* cq_last_error(const cq_ctx_t* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int cq_set_schwarz_threshold(cq_ctx_t* c, double tau) {
  return guarded(c, [&] {
    if (!(tau >= 0.0)) throw std::invalid_argument("threshold must be non-negative");
    c->tau = tau;
  });
}

int cq_get_schwarz_threshold(const cq_ctx_t* c, double* tau) {
  if (!c || !tau) return CQ_ERR_INVALID;
  *tau = c->tau;
  return CQ_OK;
}

int cq_set_precision_mode(cq_ctx_t* c, int mode, double threshold) {
  return guarded(c, [&] {
    if (mode != CQ_PREC_FP64 && mode != CQ_PREC_MIXED) throw std::invalid_argument("unknown precision mode");
    if (mode == CQ_PREC_MIXED && !(threshold > 0.0)) throw std::invalid_argument("mixed threshold must be > 0");
    c->precision_mode = mode;
    if (mode == CQ_PREC_MIXED) c->mixed_threshold = threshold;
  });
}

int cq_get_dims(const cq_ctx_t* c, int* nshell, int* nbf, int* ncart, int* npair, int* nprim_pair) {
  if (!c) return CQ_ERR_INVALID;
  if (nshell) *nshell = c->nshell;
  if (nbf) *nbf = c->nsph;
  if (ncart) *ncart = c->ncart;
  if (npair) *npair = (int)c->pairs.size();
  if (nprim_pair) *nprim_pair = (int)c->prims.size();
  return CQ_OK;
}

int cq_get_schwarz(const cq_ctx_t* c, double* Q) {
  if (!c || !Q) return CQ_ERR_INVALID;
  std::memcpy(Q, c->Q.data(), c->Q.size() * sizeof(double));
  return CQ_OK;
}

int cq_contract_direct(cq_ctx_t* c, int nmat, const double* const* X, double* const* AX, const int* types,
                       const double* scales, int n, int ld) {
  return guarded(c, [&] {
    if (nmat < 1 || nmat > MAX_MATS) throw std::invalid_argument("nmat out of range");
    if (ld < n) throw std::invalid_argument("leading dimension smaller than the matrix dimension");
    const size_t bytes = (size_t)ld * n * sizeof(double);
    c->d_mats.ensure((size_t)2 * nmat * ld * n);
    std::vector<const double*> dX(nmat);
    std::vector<double*> dAX(nmat);
    for (int i = 0; i < nmat; ++i) {
      if (!X[i] || !AX[i]) throw std::invalid_argument("null matrix pointer");
      double* px = c->d_mats.p + (size_t)(2 * i) * ld * n;
      double* pa = c->d_mats.p + (size_t)(2 * i + 1) * ld * n;
      // identical host buffers map to one device buffer: the same density is screened once,
      // and an output listed twice (the reference's formPT passes AX = {PT, PT, Mz},
      // include/singleslater/singleslater_fock.h:37-41) receives both contributions, added in
      // list order by the finalisation, and is copied in and back once
      int dupx = -1, dupa = -1;
      for (int k = 0; k < i; ++k) {
        if (X[k] == X[i]) dupx = k;
        if (AX[k] == AX[i]) dupa = k;
      }
      if (dupx >= 0) px = const_cast<double*>(dX[dupx]);
      else HIPCHK(hipMemcpyAsync(px, X[i], bytes, hipMemcpyHostToDevice, c->stream));
      if (dupa >= 0) pa = dAX[dupa];
      else HIPCHK(hipMemcpyAsync(pa, AX[i], bytes, hipMemcpyHostToDevice, c->stream));
      dX[i] = px;
      dAX[i] = pa;
    }
    contract_device(c, nmat, dX.data(), dAX.data(), types, scales, n, ld, c->stream);
    for (int i = 0; i < nmat; ++i) {
      bool first = true;
      for (int k = 0; k < i; ++k) if (AX[k] == AX[i]) first = false;
      if (first) HIPCHK(hipMemcpyAsync(AX[i], dAX[i], bytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int cq_contract_direct_device(cq_ctx_t* c, int nmat, const double* const* dX, double* const* dAX,
                              const int* types, const double* scales, int n, int ld, void* stream) {
  return guarded(c, [&] {
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    contract_device(c, nmat, dX, dAX, types, scales, n, ld, s);
  });
}

int cq_contract_direct_device_noalloc(cq_ctx_t* c, int nmat, const double* const* dX, double* const* dAX,
                                      const int* types, const double* scales, int n, int ld, void* stream) {
  return guarded(c, [&] {
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    contract_device(c, nmat, dX, dAX, types, scales, n, ld, s, true);
  });
}

int cq_get_quartet_counts(cq_ctx_t* c, long long* counts, int max_counts, int* n_out) {
  return guarded(c, [&] {
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<unsigned long long> h(2 * kNumClassPairs);
    HIPCHK(hipMemcpy(h.data(), c->d_qcounts.p, 2 * kNumClassPairs * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int k = 0; k < kNumClassPairs && k < max_counts; ++k) counts[k] = (long long)(h[k] + h[kNumClassPairs + k]);
    if (n_out) *n_out = kNumClassPairs;
  });
}

// Surviving quartets of the last contraction by arithmetic: FP64 integrals and (mixed precision
// only) FP32 integrals, summed over class pairs.
int cq_get_precision_counts(cq_ctx_t* c, long long* fp64, long long* fp32) {
  return guarded(c, [&] {
    if (!fp64 || !fp32) throw std::invalid_argument("null output");
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<unsigned long long> h(2 * kNumClassPairs);
    HIPCHK(hipMemcpy(h.data(), c->d_qcounts.p, 2 * kNumClassPairs * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    long long a = 0, b = 0;
    for (int k = 0; k < kNumClassPairs; ++k) { a += (long long)h[k]; b += (long long)h[kNumClassPairs + k]; }
    *fp64 = a;
    *fp32 = b;
  });
}

int cq_class_pair_info(int k, int* la, int* lb, int* lc, int* ld_) {
  if (k < 0 || k >= kNumClassPairs) return CQ_ERR_INVALID;
  const ClassKernels& ck = kClassKernels[k];
  auto split = [](int X, int* a, int* b) {
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= X) ++l;
    *a = l;
    *b = X - l * (l + 1) / 2;
  };
  split(ck.bi, la, lb);
  split(ck.ki, lc, ld_);
  return CQ_OK;
}

int cq_num_class_pairs(void) { return kNumClassPairs; }

// Coefficients of the real solid harmonics of angular momentum l (m = -l..l, row-major
// (2l+1) x ncart(l)) in the uniformly normalised Cartesian functions of a shell: the transform
// every pure shell uses (the reference's cart2sphCoeff / makeCar2Sph,
// src/basisset/basisset_read.cpp:871-990).
int cq_solid_harmonics(int l, double* out) {
  if (l < 0 || l > 2 || !out) return CQ_ERR_INVALID;
  const std::vector<double> M = solid_harmonics(l);
  std::copy(M.begin(), M.end(), out);
  return CQ_OK;
}

// S, T, V in the shells' basis (column-major n x n, leading dimension ld, device memory),
// replacing AOIntegrals::computeAOOneE (src/aointegrals/aointegrals_builders.cpp:337-467).
// Z[natom] nuclear charges and xyz[3*natom] positions (bohr) are host arrays.
int cq_one_electron(cq_ctx_t* c, int natom, const double* Z, const double* xyz, double* S, double* T, double* V,
                    int ld, void* stream) {
  return guarded(c, [&] {
    if (c->nmol > 1)
      throw std::invalid_argument("a context of several molecules needs cq_one_electron_batch (which nuclei "
                                  "belong to which molecule)");
    one_electron(c, natom, Z, xyz, nullptr, S, T, V, ld, stream);
  });
}

int cq_one_electron_batch(cq_ctx_t* c, int natom, const double* Z, const double* xyz, const int* atom_molecule,
                          double* S, double* T, double* V, int ld, void* stream) {
  return guarded(c, [&] {
    if (!atom_molecule) throw std::invalid_argument("atom_molecule is required");
    one_electron(c, natom, Z, xyz, atom_molecule, S, T, V, ld, stream);
  });
}

int cq_multipole(cq_ctx_t* c, int order, const double* origin, double* out, int ld, void* stream) {
  return guarded(c, [&] { multipole(c, order, origin, out, ld, stream); });
}

// Selects the J/K digestion: 1 (default) accumulates K columns and J(3,4) of each ket pair in
// LDS and flushes once per work item; 0 issues every update as an HBM atomic.
int cq_set_lds_digestion(cq_ctx_t* c, int mode) {
  return guarded(c, [&] {
    if (mode < 0 || mode > 2) throw std::invalid_argument("digestion mode must be 0, 1 or 2");
    c->lds_mode = mode;
  });
}

int cq_set_class_digestion(cq_ctx_t* c, int k, int mode) {
  return guarded(c, [&] {
    if (k < 0 || k >= kNumClassPairs) throw std::invalid_argument("class pair index out of range");
    if (mode < -1 || mode > 2) throw std::invalid_argument("digestion mode must be -1, 0, 1 or 2");
    c->class_variant[k] = mode;
  });
}

int cq_set_class_stream(cq_ctx_t* c, int k, int stream) {
  return guarded(c, [&] {
    if (k < 0 || k >= kNumClassPairs) throw std::invalid_argument("class pair index out of range");
    if (stream < -1 || stream > cq_ctx::kAux) throw std::invalid_argument("stream index must be -1 .. 3");
    c->class_stream[k] = stream;
  });
}

// At most cpi chunks of a class pair's ket (or bra) components per work item of its group
// kernels (the largest divisor of the class pair's chunk count not above cpi); 0 for the
// built-in choice.
int cq_set_class_chunk_group(cq_ctx_t* c, int k, int cpi) {
  return guarded(c, [&] {
    if (k < 0 || k >= kNumClassPairs) throw std::invalid_argument("class pair index out of range");
    if (cpi < 0) throw std::invalid_argument("chunks per work item must be >= 0");
    const int nch = kClassKernels[k].nchunk;
    int d = std::min(cpi, nch);
    while (d > 1 && nch % d != 0) --d;
    c->class_cpi[k] = cpi == 0 ? 0 : d;
  });
}

int cq_get_class_digestion(const cq_ctx_t* c, int k, int* mode) {
  if (!c || !mode || k < 0 || k >= kNumClassPairs) return CQ_ERR_INVALID;
  *mode = c->class_variant[k];
  return CQ_OK;
}

// Largest task list of one class pair, in 8-byte entries (default 2^28, 2 GiB): class pairs
// with a larger worst case are screened and digested in batches of uniform pairs.  Waits for
// the context's work, then re-plans the batches and reallocates the task buffers.
int cq_set_task_budget(cq_ctx_t* c, long long entries) {
  return guarded(c, [&] {
    if (entries < 1 || entries > 0x7fffffffLL) throw std::invalid_argument("task budget must be in [1, 2^31 - 1] entries");
    for (int a = 0; a < cq_ctx::kAux; ++a)
      if (c->aux[a]) HIPCHK(hipStreamSynchronize(c->aux[a]));
    HIPCHK(hipStreamSynchronize(c->stream));
    const long long old = c->task_budget;
    c->task_budget = entries;
    c->d_tasks.release();
    c->d_segs.release();
    try {
      plan_tasks(c);
    } catch (...) {
      // (out of device memory for the new plan) keep the context usable with the old one; if
      // even that cannot be allocated, plan_ok stays false and contractions raise an error
      // instead of screening into released task buffers
      c->task_budget = old;
      c->d_tasks.release();
      c->d_segs.release();
      try {
        plan_tasks(c);
      } catch (...) {
      }
      throw;
    }
  });
}

// Tile of class pair k's group kernels: the dynamic LDS budget in bytes (0: the default 48 KiB),
// the largest group (members of one work item, at most 16; 0: the built-in table's) and the
// union bound of a task region (swept pairs whose J collects in LDS slots; 0: the built-in
// table's).  Waits for the context's work and re-plans the task lists.
int cq_set_class_tile(cq_ctx_t* c, int k, long long lds_bytes, int max_group, int union_bound) {
  return guarded(c, [&] {
    if (k < 0 || k >= kNumClassPairs) throw std::invalid_argument("class pair index out of range");
    if (lds_bytes < 0 || max_group < 0 || max_group > MAX_GROUP || union_bound < 0 || union_bound > UNION_MAX ||
        (union_bound > 0 && union_bound < 64))
      throw std::invalid_argument("tile out of range: lds_bytes >= 0, max_group 0..16, union_bound 0 or 64..1024");
    for (int a = 0; a < cq_ctx::kAux; ++a)
      if (c->aux[a]) HIPCHK(hipStreamSynchronize(c->aux[a]));
    HIPCHK(hipStreamSynchronize(c->stream));
    const long long o0 = c->tile_lds[k];
    const int o1 = c->tile_mmax[k], o2 = c->tile_ub[k];
    c->tile_lds[k] = lds_bytes;
    c->tile_mmax[k] = max_group;
    c->tile_ub[k] = union_bound;
    try {
      plan_tasks(c);
    } catch (...) {
      c->tile_lds[k] = o0;
      c->tile_mmax[k] = o1;
      c->tile_ub[k] = o2;
      try {
        plan_tasks(c);
      } catch (...) {
      }
      throw;
    }
  });
}

// The planned tile of class pair k in the digestion table of the given variant (0/1: ket-keyed,
// 2: bra-keyed where the class pair has that form): largest group, union bound, number of
// groups, and the kernel's dynamic LDS bytes.
int cq_get_class_tile(const cq_ctx_t* c, int k, int variant, int* max_group, int* union_bound, int* ngroups,
                      long long* lds_bytes) {
  if (!c || k < 0 || k >= kNumClassPairs || variant < 0 || variant > 2 || !max_group || !union_bound || !ngroups ||
      !lds_bytes)
    return CQ_ERR_INVALID;
  const CPMeta& m = c->cps[variant == 2 ? 1 : 0][k];
  const ClassKernels& ck = kClassKernels[k];
  *max_group = m.mmax;
  *union_bound = m.ub;
  *ngroups = m.ngrp;
  const int kv = m.flip ? 2 : (variant >= 1 && group_lds(c, ck, 1, m.mmax, m.ub) <= c->max_dyn_lds) ? 1 : 0;
  *lds_bytes = (long long)group_lds(c, ck, kv, m.mmax, m.ub);
  return CQ_OK;
}

// Task entries allocated over all class pairs, and the largest number of batches any class
// pair takes (over both digestion tables).
int cq_get_task_plan(const cq_ctx_t* c, long long* entries, int* max_batches) {
  if (!c || !entries || !max_batches) return CQ_ERR_INVALID;
  *entries = c->task_capacity;
  int mb = 0;
  for (int t = 0; t < 2; ++t)
    for (const auto& bs : c->batch_start[t]) mb = std::max(mb, (int)bs.size() - 1);
  *max_batches = mb;
  return CQ_OK;
}

// Number of streams the class-pair kernels are spread over, the contraction's own stream
// first (1 serialises them on it, for per-kernel profiling; default 4).
int cq_set_launch_streams(cq_ctx_t* c, int n) {
  return guarded(c, [&] {
    if (n < 1 || n > cq_ctx::kAux + 1) throw std::invalid_argument("stream count out of range");
    c->n_streams = n;
  });
}

// How the class pairs that screen once per call are screened: CQ_SCREEN_PER_CLASS (one launch
// per class pair on its own stream), CQ_SCREEN_PER_STREAM (one launch per stream for the class
// pairs on it; the default) or CQ_SCREEN_ONE (one launch for all, before the fork).  The
// surviving quartets are the same in every mode.
int cq_set_screen_fusion(cq_ctx_t* c, int on) {
  return guarded(c, [&] {
    if (on < CQ_SCREEN_PER_CLASS || on > CQ_SCREEN_ONE) throw std::invalid_argument("unknown screening mode");
    c->screen_mode = on;
  });
}

// Per-class-pair kernel timing with HIP events recorded on the launch stream.
int cq_set_profiling(cq_ctx_t* c, int on) {
  return guarded(c, [&] {
    c->profiling = on != 0;
    for (auto& v : c->prof_events) v.clear();
    for (auto& v : c->prof_screen_events) v.clear();
    std::fill(c->prof_ms.begin(), c->prof_ms.end(), 0.0);
    std::fill(c->prof_launches.begin(), c->prof_launches.end(), 0);
    std::fill(c->prof_screen_ms.begin(), c->prof_screen_ms.end(), 0.0);
    std::fill(c->prof_screen_launches.begin(), c->prof_screen_launches.end(), 0);
    c->event_next = 0;
  });
}

// Accumulated kernel milliseconds and launch counts per class pair since profiling was
// switched on (synchronises every recorded event).
namespace {
// Fold the recorded event pairs of every class pair into the accumulated times (synchronises
// every recorded event); the event pool is reused afterwards.
void collect_times(cq_ctx* c) {
  auto fold = [](std::vector<std::pair<hipEvent_t, hipEvent_t>>& evs, double& ms, long long& n) {
    for (auto& pr : evs) {
      HIPCHK(hipEventSynchronize(pr.second));
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, pr.first, pr.second));
      ms += t;
      n += 1;
    }
    evs.clear();
  };
  for (int k = 0; k < kNumClassPairs; ++k) fold(c->prof_events[k], c->prof_ms[k], c->prof_launches[k]);
  for (int k = 0; k <= kNumClassPairs; ++k)
    fold(c->prof_screen_events[k], c->prof_screen_ms[k], c->prof_screen_launches[k]);
  c->event_next = 0;
}
}  // namespace

int cq_get_kernel_times(cq_ctx_t* c, double* ms, long long* launches, int max_n) {
  return guarded(c, [&] {
    collect_times(c);
    for (int k = 0; k < kNumClassPairs && k < max_n; ++k) {
      ms[k] = c->prof_ms[k];
      launches[k] = c->prof_launches[k];
    }
  });
}

// Accumulated screening-kernel milliseconds and launch counts per class pair since profiling
// was switched on (HIP events around each screen_kernel launch on its stream); entry
// CQ_NUM_CLASS_PAIRS, when max_n leaves room for it, is the fused launch of all class pairs
// that screen once per call.
int cq_get_screen_times(cq_ctx_t* c, double* ms, long long* launches, int max_n) {
  return guarded(c, [&] {
    collect_times(c);
    for (int k = 0; k <= kNumClassPairs && k < max_n; ++k) {
      ms[k] = c->prof_screen_ms[k];
      launches[k] = c->prof_screen_launches[k];
    }
  });
}

// Work sums of the last contraction per class pair (4 doubles each): primitive quartets
// sum K_ab K_cd, sum K_ab, Cartesian->spherical transform FLOPs, output integrals.
int cq_get_work_stats(cq_ctx_t* c, double* out, int max_n) {
  return guarded(c, [&] {
    // the task lists hold the last batch only when a class pair is batched
    for (int t = 0; t < 2; ++t)
      for (const auto& bs : c->batch_start[t])
        if (bs.size() > 2) throw std::invalid_argument("work statistics need every class pair in one batch (raise the task budget)");
    DevBuf<double> acc;
    acc.ensure((size_t)4 * kNumClassPairs);
    HIPCHK(hipMemsetAsync(acc.p, 0, 4 * kNumClassPairs * sizeof(double), c->stream));
    for (int k = 0; k < kNumClassPairs; ++k) {
      const int tb = c->last_table[k];
      const CPMeta& m = c->cps[tb][k];
      if (m.nbra == 0 || m.nket == 0 || m.ngrp == 0) continue;
      hipLaunchKernelGGL(work_stats_kernel, dim3(persistent_grid()), dim3(WG), 0, c->stream,
                         c->d_tasks.p + m.task_off, c->d_segs.p + m.seg_off, c->batch_slots[tb][k][0], c->d_ovf.p + k,
                         c->d_groups[tb].p,
                         c->d_uorder.p, m.flip, c->d_pairs.p, c->d_shells.p, acc.p + 4 * k);
      HIPCHK(hipGetLastError());
    }
    std::vector<double> h((size_t)4 * kNumClassPairs);
    HIPCHK(hipMemcpyAsync(h.data(), acc.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    acc.release();
    for (int k = 0; k < kNumClassPairs && k < max_n; ++k)
      for (int q = 0; q < 4; ++q) out[4 * k + q] = h[4 * k + q];
  });
}

// Contracted quartets (s1 s2|s3 s4) in the shells' own basis, row-major per quartet,
// concatenated; for per-quartet parity tests.
int cq_debug_quartets(cq_ctx_t* c, int nq, const int* shells4, double* out) {
  return guarded(c, [&] {
    // map canonical shell pair -> pair index
    std::vector<int> pidx((size_t)c->nshell * (c->nshell + 1) / 2, -1);
    for (int k = 0; k < (int)c->pairs.size(); ++k) {
      const PairInfo& p = c->pairs[k];
      pidx[(size_t)p.s1 * (p.s1 + 1) / 2 + p.s2] = k;
    }
    size_t out_off = 0;
    for (int q = 0; q < nq; ++q) {
      int s[4] = {shells4[4 * q], shells4[4 * q + 1], shells4[4 * q + 2], shells4[4 * q + 3]};
      for (int k = 0; k < 4; ++k)
        if (s[k] < 0 || s[k] >= c->nshell) throw std::invalid_argument("shell index out of range");
      auto pid = [&](int x, int y) { return pidx[(size_t)std::max(x, y) * (std::max(x, y) + 1) / 2 + std::min(x, y)]; };
      const int P = pid(s[0], s[1]), Qi = pid(s[2], s[3]);
      if (P < 0 || Qi < 0) {
        // a shell pair without any primitive pair above the threshold: the block is zero
        size_t n = 1;
        for (int r = 0; r < 4; ++r) {
          const ShellMeta& sm = c->shells[s[r]];
          n *= sm.pure ? 2 * sm.l + 1 : ncart_l(sm.l);
        }
        std::memset(out + out_off, 0, n * sizeof(double));
        out_off += n;
        continue;
      }
      const PairInfo& pp = c->pairs[P];
      const PairInfo& pq = c->pairs[Qi];
      const int clsP = pair_class(c->shells[pp.a].l, c->shells[pp.b].l);
      const int clsQ = pair_class(c->shells[pq.a].l, c->shells[pq.b].l);
      int kk = -1;
      bool swap_bk = false;
      for (int k = 0; k < kNumClassPairs; ++k) {
        if (kClassKernels[k].bi == clsP && kClassKernels[k].ki == clsQ) { kk = k; swap_bk = false; break; }
        if (kClassKernels[k].bi == clsQ && kClassKernels[k].ki == clsP) { kk = k; swap_bk = true; break; }
      }
      if (kk < 0) throw std::runtime_error("no kernel for quartet class");
      const int bra = swap_bk ? Qi : P, ket = swap_bk ? P : Qi;
      const PairInfo& B = c->pairs[bra];
      const PairInfo& K = c->pairs[ket];
      const int na = ncart_l(c->shells[B.a].l), nb = ncart_l(c->shells[B.b].l);
      const int nc = ncart_l(c->shells[K.a].l), nd = ncart_l(c->shells[K.b].l);
      const int blk = na * nb * nc * nd;
      std::vector<uint2> t1{make_uint2(bra, ket)};
      DevBuf<uint2> dt;
      DevBuf<double> dout;
      dt.upload(t1, c->stream);
      dout.ensure(blk);
      HIPCHK(hipMemsetAsync(dout.p, 0, blk * sizeof(double), c->stream));
      hipLaunchKernelGGL(kClassKernels[kk].block, dim3(1), dim3(WG), 0, c->stream, dt.p, 1L, c->d_pairs.p,
                         c->d_prims.p, c->d_boys.p, dout.p, blk, (const int*)nullptr);
      HIPCHK(hipGetLastError());
      std::vector<double> h(blk);
      HIPCHK(hipMemcpyAsync(h.data(), dout.p, blk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      dt.release();
      dout.release();
      // kernel order (B.a B.b | K.a K.b) -> requested (s0 s1 | s2 s3), Cartesian
      int kshell[4] = {B.a, B.b, K.a, K.b};
      int kdim[4] = {na, nb, nc, nd};
      int perm[4];   // requested position r -> kernel position
      bool used[4] = {false, false, false, false};
      // bra/ket of the request map to the pair that holds them
      const int rb0 = swap_bk ? 2 : 0;   // kernel positions holding the request's bra pair
      const int rk0 = swap_bk ? 0 : 2;
      auto assign = [&](int r, int base) {
        for (int k2 = base; k2 < base + 2; ++k2)
          if (!used[k2] && kshell[k2] == s[r]) { perm[r] = k2; used[k2] = true; return; }
      };
      assign(0, rb0); assign(1, rb0); assign(2, rk0); assign(3, rk0);
      int rdim[4];
      for (int r = 0; r < 4; ++r) rdim[r] = kdim[perm[r]];
      std::vector<double> cartq((size_t)blk);
      int idx[4];
      for (idx[0] = 0; idx[0] < rdim[0]; ++idx[0])
        for (idx[1] = 0; idx[1] < rdim[1]; ++idx[1])
          for (idx[2] = 0; idx[2] < rdim[2]; ++idx[2])
            for (idx[3] = 0; idx[3] < rdim[3]; ++idx[3]) {
              int kidx[4];
              for (int r = 0; r < 4; ++r) kidx[perm[r]] = idx[r];
              const size_t ksrc = (((size_t)kidx[0] * nb + kidx[1]) * nc + kidx[2]) * nd + kidx[3];
              const size_t rdst = (((size_t)idx[0] * rdim[1] + idx[1]) * rdim[2] + idx[2]) * rdim[3] + idx[3];
              cartq[rdst] = h[ksrc];
            }
      // Cartesian -> shells' basis, index by index
      std::vector<double> cur = cartq;
      int dims[4] = {rdim[0], rdim[1], rdim[2], rdim[3]};
      for (int r = 0; r < 4; ++r) {
        const ShellMeta& sm = c->shells[s[r]];
        if (!sm.pure) continue;
        const int ncs = ncart_l(sm.l), nss = 2 * sm.l + 1;
        const double* M = c->tco.data() + c->tco_off[sm.l];
        size_t pre = 1, post = 1;
        for (int k2 = 0; k2 < r; ++k2) pre *= dims[k2];
        for (int k2 = r + 1; k2 < 4; ++k2) post *= dims[k2];
        std::vector<double> nx(pre * nss * post, 0.0);
        for (size_t p0 = 0; p0 < pre; ++p0)
          for (int i = 0; i < nss; ++i)
            for (int a = 0; a < ncs; ++a) {
              const double w = M[i * ncs + a];
              if (w == 0.0) continue;
              for (size_t p1 = 0; p1 < post; ++p1)
                nx[(p0 * nss + i) * post + p1] += w * cur[(p0 * ncs + a) * post + p1];
            }
        cur.swap(nx);
        dims[r] = nss;
      }
      std::memcpy(out + out_off, cur.data(), cur.size() * sizeof(double));
      out_off += cur.size();
    }
  });
}

}  // extern "C"
