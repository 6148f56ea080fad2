// CPU oracle for the direct-SCF G[D] path.  TEST INFRASTRUCTURE ONLY: it is loaded by
// tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker
// and the CPU baseline.  The product library never links or calls it.
//
// Contents (each restates the reference behaviour named beside it):
//   * Boys function F_m(T)  -- series / erf + upward recursion in long double
//     (method reference: src/aointegrals/aointegrals.cpp:26-141; pinned by scipy gammainc
//     vectors in tests/golden/boys.json).
//   * Contracted shell-quartet ERIs (ab|cd) by Obara-Saika VRR + Head-Gordon-Pople HRR,
//     templated on the real type (stands in for libint2::Engine::compute,
//     include/aointegrals/aointegrals_contract.h:138-143).
//   * An independent McMurchie-Davidson engine in long double (cross-check only).
//   * Cartesian -> spherical coefficients restating BasisSet::cart2sphCoeff /
//     makeCar2Sph (src/basisset/basisset_read.cpp:871-990).
//   * Schwarz bounds (src/aointegrals/aointegrals_builders.cpp:469-518).
//   * The direct contraction, loop for loop (include/aointegrals/aointegrals_contract.h:24-194).
//   * The N^4 brute-force contraction (include/aointegrals/aointegrals_contract.h:197-262).
//   * One-electron S, T, V (src/aointegrals/aointegrals_builders.cpp:27-132,337-467).
//   * Multipole integrals to third order (libint2 emultipole1..3 in OneEDriver,
//     src/aointegrals/aointegrals_builders.cpp:38-55), by Hermite moments.
//
// Parity status: every function is pinned by a test in tests/test_oracle.py (closed forms,
// the second algorithm, finite differences, permutational symmetry, the reference's golden
// SCF energies).  None is "parity unpinned".
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr int kMaxL = 4;               // per shell
constexpr int kMaxLTot = 4 * kMaxL;    // per quartet

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian triples of total l in x^l ... z^l lexicographic order.
void cart_list(int l, std::vector<int>& lx, std::vector<int>& ly, std::vector<int>& lz) {
  lx.clear(); ly.clear(); lz.clear();
  for (int i = 0; i <= l; ++i) {
    int x = l - i;
    for (int j = 0; j <= i; ++j) {
      int y = i - j;
      lx.push_back(x); ly.push_back(y); lz.push_back(l - x - y);
    }
  }
}

// ---------------------------------------------------------------- Boys function
void boys_ld(int mmax, long double T, long double* F) {
  if (T < 1e-30L) {
    for (int m = 0; m <= mmax; ++m) F[m] = 1.0L / (2 * m + 1);
    return;
  }
  const long double eT = std::exp(-T);
  if (T < 40.0L) {
    // F_mmax by the series e^-T sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)), then downward.
    long double term = 1.0L / (2 * mmax + 1);
    long double sum = term;
    for (int k = 1; k < 2000; ++k) {
      term *= 2.0L * T / (2 * mmax + 2 * k + 1);
      sum += term;
      if (term < sum * 1e-22L) break;
    }
    F[mmax] = eT * sum;
    for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0L * T * F[m + 1] + eT) / (2 * m + 1);
  } else {
    const long double sT = std::sqrt(T);
    F[0] = 0.5L * std::sqrt(3.14159265358979323846264338327950288L / T) * std::erf(sT);
    for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - eT) / (2.0L * T);
  }
}

// ---------------------------------------------------------------- basis
struct Shell {
  int l, pure, nprim, size, off;
  std::vector<double> a, c;
  double O[3];
};

struct Basis {
  std::vector<Shell> sh;
  int nbf = 0, maxl = 0, maxprim = 0;
};

// ---------------------------------------------------------------- cart -> sph
// Restates BasisSet::cart2sphCoeff (src/basisset/basisset_read.cpp:871-958).
double fact(int n) { double r = 1; for (int i = 2; i <= n; ++i) r *= i; return r; }
double dfact(int n) { double r = 1; while (n > 1) { r *= n; n -= 2; } return r; }
double binom(int n, int k) { if (n == 0) return 1.0; return fact(n) / fact(k) / fact(n - k); }

std::pair<double, double> cart2sph_coeff(int l, int m, int x, int y, int z) {
  using cd = std::complex<double>;
  cd tmp = fact(2 * x) * fact(2 * y) * fact(2 * z) * fact(l) * fact(l - m);
  tmp /= (fact(2 * l) * fact(x) * fact(y) * fact(z) * fact(l + m));
  tmp = std::sqrt(tmp) / (std::pow(2.0, l) * fact(l));
  int j = x + y - m;
  if (j % 2 != 0) return {0.0, 0.0};
  j /= 2;
  cd s(0, 0);
  for (int i = 0; i <= (l - m) / 2; ++i) {
    if (i > l || i < 0) continue;
    if (j > i || j < 0) continue;
    double t = binom(l, i) * binom(i, j);
    t *= std::pow(-1.0, i) * fact(2 * l - 2 * i);
    t /= fact(l - m - 2 * i);
    s += t;
  }
  tmp *= s;
  cd sp(0, 0), sm(0, 0);
  for (int k = 0; k <= j; ++k) {
    int xk = x - 2 * k;
    if (xk > m || xk < 0) continue;
    double base = binom(j, k) * binom(m, xk);
    double e = double(m - x + 2 * k) / 2.0;
    sp += base * std::pow(cd(-1.0, 0.0), e);
    sm += base * std::pow(cd(-1.0, 0.0), -e);
  }
  cd Rp = tmp, Rm = tmp;
  if (j != 0 || m != 0) { Rp *= sp; Rm *= sm; }
  int L = x + y + z;
  double f = dfact(2 * L - 1);
  if (x > 0) f /= dfact(2 * x - 1);
  if (y > 0) f /= dfact(2 * y - 1);
  if (z > 0) f /= dfact(2 * z - 1);
  f = std::sqrt(f);
  Rp *= f; Rm *= f;
  if (m == 0) return {std::real(Rp), 0.0};
  return {std::real((Rp + Rm) / std::sqrt(2.0)), std::real((Rp - Rm) / std::sqrt(cd(-2.0, 0.0)))};
}

// (2l+1) x ncart(l) matrix, row-major; row l+m cos-type, row l-m sin-type (makeCar2Sph).
std::vector<double> car2sph(int l) {
  int nc = ncart(l);
  std::vector<double> M((2 * l + 1) * nc, 0.0);
  int I = 0;
  for (int i = 0; i <= l; ++i) {
    int lx = l - i;
    for (int j = 0; j <= i; ++j, ++I) {
      int ly = i - j, lz = l - lx - ly;
      M[l * nc + I] = cart2sph_coeff(l, 0, lx, ly, lz).first;
      for (int m = 1; m <= l; ++m) {
        auto t = cart2sph_coeff(l, m, lx, ly, lz);
        M[(l + m) * nc + I] = t.first;
        M[(l - m) * nc + I] = t.second;
      }
    }
  }
  return M;
}

// Initialised once, thread-safely (C++11 function-local static), before any OpenMP region
// can read it.
const std::vector<double>& car2sph_cached(int l) {
  static const std::vector<std::vector<double>> cache = [] {
    std::vector<std::vector<double>> c(kMaxL + 1);
    for (int i = 0; i <= kMaxL; ++i) c[i] = car2sph(i);
    return c;
  }();
  return cache[l];
}

// Transform a Cartesian quartet block (row-major a,b,c,d) to the shells' own basis.
void cart_to_shell_basis(const Shell* s[4], std::vector<double>& buf) {
  int dims[4];
  for (int k = 0; k < 4; ++k) dims[k] = ncart(s[k]->l);
  for (int k = 0; k < 4; ++k) {
    if (!s[k]->pure) continue;
    const auto& M = car2sph_cached(s[k]->l);
    int nc = dims[k], ns = 2 * s[k]->l + 1;
    int pre = 1, post = 1;
    for (int q = 0; q < k; ++q) pre *= dims[q];
    for (int q = k + 1; q < 4; ++q) post *= dims[q];
    std::vector<double> out((size_t)pre * ns * post, 0.0);
    for (int p = 0; p < pre; ++p)
      for (int i = 0; i < ns; ++i)
        for (int c = 0; c < nc; ++c) {
          double w = M[i * nc + c];
          if (w == 0.0) continue;
          const double* src = &buf[((size_t)p * nc + c) * post];
          double* dst = &out[((size_t)p * ns + i) * post];
          for (int r = 0; r < post; ++r) dst[r] += w * src[r];
        }
    buf.swap(out);
    dims[k] = ns;
  }
}

// ---------------------------------------------------------------- OS/HGP ERI
// Index of a Cartesian triple among all triples with total <= some L, ordered by total
// then lexicographic: idx(l, x, y) = H(l-1) + position within level l.
inline int tri_index(int x, int y, int z) {
  int l = x + y + z;
  int base = l * (l + 1) * (l + 2) / 6;
  int i = l - x;          // y+z
  int pos = i * (i + 1) / 2 + (i - y);
  return base + pos;
}
inline int ntri(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Double-precision Boys: 7-term Taylor expansion about grid points (step 0.05) of a table
// filled once from the long-double series, then downward recursion; erf + upward
// recursion for T >= 40.
struct BoysTable {
  static constexpr int kM = 4 * kMaxL + 8;
  static constexpr double kStep = 0.05;
  static constexpr int kN = 801;   // T in [0, 40]
  std::vector<double> tab;         // tab[i * kM + m]
  BoysTable() : tab((size_t)kN * kM) {
    long double F[kM];
    for (int i = 0; i < kN; ++i) {
      boys_ld(kM - 1, (long double)(i * kStep), F);
      for (int m = 0; m < kM; ++m) tab[(size_t)i * kM + m] = (double)F[m];
    }
  }
};

void boys_d(int mmax, double T, double* F) {
  static const BoysTable bt;
  if (T < 40.0) {
    int i = (int)(T / BoysTable::kStep + 0.5);
    double dT = i * BoysTable::kStep - T;   // F(T) = sum_k F_{m+k}(T0) (T0-T)^k / k!
    const double* f = &bt.tab[(size_t)i * BoysTable::kM + mmax];
    double s = f[6];
    for (int k = 5; k >= 0; --k) s = f[k] + s * dT / (k + 1);
    F[mmax] = s;
    const double eT = std::exp(-T);
    for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + eT) / (2 * m + 1);
  } else {
    const double eT = std::exp(-T);
    F[0] = 0.5 * std::sqrt(3.14159265358979323846 / T) * std::erf(std::sqrt(T));
    for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - eT) / (2.0 * T);
  }
}

template <typename R>
void boysR(int mmax, R T, R* F) {
  long double Fl[4 * kMaxL + 2];
  boys_ld(mmax, (long double)T, Fl);
  for (int m = 0; m <= mmax; ++m) F[m] = (R)Fl[m];
}
template <>
void boysR<double>(int mmax, double T, double* F) { boys_d(mmax, T, F); }

// Contracted Cartesian quartet by OS VRR on primitives + HGP HRR on contracted [e0|f0].
template <typename R>
void eri_cart_os(const Shell& A, const Shell& B, const Shell& C, const Shell& D, std::vector<double>& out) {
  const int la = A.l, lb = B.l, lc = C.l, ld = D.l;
  const int Lab = la + lb, Lcd = lc + ld, L = Lab + Lcd;
  const int nE = ntri(Lab), nF = ntri(Lcd);
  const R pi = (R)3.14159265358979323846264338327950288L;
  // contracted [e0|f0] for all e <= Lab, f <= Lcd (only la<=|e|, lc<=|f| used)
  std::vector<R> ef((size_t)nE * nF, (R)0);
  std::vector<R> V((size_t)(L + 1) * nE * nF);
  auto Vat = [&](int m, int e, int f) -> R& { return V[((size_t)m * nE + e) * nF + f]; };
  R AB[3], CD[3];
  for (int k = 0; k < 3; ++k) { AB[k] = (R)A.O[k] - (R)B.O[k]; CD[k] = (R)C.O[k] - (R)D.O[k]; }
  R AB2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  R CD2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
  std::vector<int> ex(nE), ey(nE), ez(nE), fx(nF), fy(nF), fz(nF);
  for (int l = 0; l <= Lab; ++l) for (int i = 0; i <= l; ++i) for (int j = 0; j <= i; ++j) {
    int x = l - i, y = i - j, z = l - x - y, id = tri_index(x, y, z);
    ex[id] = x; ey[id] = y; ez[id] = z;
  }
  for (int l = 0; l <= Lcd; ++l) for (int i = 0; i <= l; ++i) for (int j = 0; j <= i; ++j) {
    int x = l - i, y = i - j, z = l - x - y, id = tri_index(x, y, z);
    fx[id] = x; fy[id] = y; fz[id] = z;
  }
  R F[4 * kMaxL + 2];
  for (int pa = 0; pa < A.nprim; ++pa)
  for (int pb = 0; pb < B.nprim; ++pb) {
    R a = A.a[pa], b = B.a[pb], p = a + b;
    R P[3], PA[3];
    for (int k = 0; k < 3; ++k) { P[k] = (a * (R)A.O[k] + b * (R)B.O[k]) / p; PA[k] = P[k] - (R)A.O[k]; }
    R Kab = std::exp(-a * b / p * AB2) * (R)A.c[pa] * (R)B.c[pb];
    for (int pc = 0; pc < C.nprim; ++pc)
    for (int pd = 0; pd < D.nprim; ++pd) {
      R c = C.a[pc], d = D.a[pd], q = c + d;
      R Q[3], QC[3], W[3], WP[3], WQ[3];
      for (int k = 0; k < 3; ++k) { Q[k] = (c * (R)C.O[k] + d * (R)D.O[k]) / q; QC[k] = Q[k] - (R)C.O[k]; }
      R Kcd = std::exp(-c * d / q * CD2) * (R)C.c[pc] * (R)D.c[pd];
      R pq = p + q, rho = p * q / pq;
      R PQ2 = 0;
      for (int k = 0; k < 3; ++k) {
        W[k] = (p * P[k] + q * Q[k]) / pq;
        WP[k] = W[k] - P[k]; WQ[k] = W[k] - Q[k];
        R d_ = P[k] - Q[k]; PQ2 += d_ * d_;
      }
      R pref = (R)2 * std::pow(pi, (R)2.5) / (p * q * std::sqrt(pq)) * Kab * Kcd;
      boysR<R>(L, rho * PQ2, F);
      for (int m = 0; m <= L; ++m) Vat(m, 0, 0) = pref * F[m];
      // bra VRR: [e0|00]^(m)
      for (int le = 1; le <= Lab; ++le)
        for (int i = 0; i <= le; ++i) for (int j = 0; j <= i; ++j) {
          int x = le - i, y = i - j, z = le - x - y;
          int e = tri_index(x, y, z);
          int dir = x > 0 ? 0 : (y > 0 ? 1 : 2);
          int e1 = tri_index(x - (dir == 0), y - (dir == 1), z - (dir == 2));
          int ne1 = (dir == 0 ? x : dir == 1 ? y : z) - 1;   // count in dir of e1
          int e2 = ne1 > 0 ? tri_index(x - 2 * (dir == 0), y - 2 * (dir == 1), z - 2 * (dir == 2)) : -1;
          for (int m = 0; m <= L - le; ++m) {
            R v = PA[dir] * Vat(m, e1, 0) + WP[dir] * Vat(m + 1, e1, 0);
            if (e2 >= 0) v += (R)ne1 / (2 * p) * (Vat(m, e2, 0) - rho / p * Vat(m + 1, e2, 0));
            Vat(m, e, 0) = v;
          }
        }
      // ket VRR: [e0|f0]^(m)
      for (int lf = 1; lf <= Lcd; ++lf)
        for (int i = 0; i <= lf; ++i) for (int j = 0; j <= i; ++j) {
          int x = lf - i, y = i - j, z = lf - x - y;
          int f = tri_index(x, y, z);
          int dir = x > 0 ? 0 : (y > 0 ? 1 : 2);
          int f1 = tri_index(x - (dir == 0), y - (dir == 1), z - (dir == 2));
          int nf1 = (dir == 0 ? x : dir == 1 ? y : z) - 1;
          int f2 = nf1 > 0 ? tri_index(x - 2 * (dir == 0), y - 2 * (dir == 1), z - 2 * (dir == 2)) : -1;
          for (int e = 0; e < nE; ++e) {
            int le = ex[e] + ey[e] + ez[e];
            int nedir = dir == 0 ? ex[e] : dir == 1 ? ey[e] : ez[e];
            int em = nedir > 0 ? tri_index(ex[e] - (dir == 0), ey[e] - (dir == 1), ez[e] - (dir == 2)) : -1;
            for (int m = 0; m <= L - le - lf; ++m) {
              R v = QC[dir] * Vat(m, e, f1) + WQ[dir] * Vat(m + 1, e, f1);
              if (f2 >= 0) v += (R)nf1 / (2 * q) * (Vat(m, e, f2) - rho / q * Vat(m + 1, e, f2));
              if (em >= 0) v += (R)nedir / (2 * pq) * Vat(m + 1, em, f1);
              Vat(m, e, f) = v;
            }
          }
        }
      for (int e = 0; e < nE; ++e) for (int f = 0; f < nF; ++f) ef[(size_t)e * nF + f] += Vat(0, e, f);
    }
  }
  // HRR on the bra: (a b| with b built up from 0.  Store as map over (a-triple, b-triple).
  // Work arrays indexed by tri_index over totals up to Lab for a, up to lb for b.
  const int nB = ntri(lb), nD = ntri(ld);
  // bra: X[a][b][f], a total in [la, Lab - |b|]
  std::vector<R> X((size_t)nE * nB * nF, (R)0);
  auto Xat = [&](int a_, int b_, int f_) -> R& { return X[((size_t)a_ * nB + b_) * nF + f_]; };
  for (int e = 0; e < nE; ++e) for (int f = 0; f < nF; ++f) Xat(e, 0, f) = ef[(size_t)e * nF + f];
  for (int lbb = 1; lbb <= lb; ++lbb)
    for (int i = 0; i <= lbb; ++i) for (int j = 0; j <= i; ++j) {
      int x = lbb - i, y = i - j, z = lbb - x - y, bb = tri_index(x, y, z);
      int dir = x > 0 ? 0 : (y > 0 ? 1 : 2);
      int b1 = tri_index(x - (dir == 0), y - (dir == 1), z - (dir == 2));
      for (int la_ = la; la_ <= Lab - lbb; ++la_)
        for (int i2 = 0; i2 <= la_; ++i2) for (int j2 = 0; j2 <= i2; ++j2) {
          int ax = la_ - i2, ay = i2 - j2, az = la_ - ax - ay, aa = tri_index(ax, ay, az);
          int ap = tri_index(ax + (dir == 0), ay + (dir == 1), az + (dir == 2));
          for (int f = 0; f < nF; ++f) Xat(aa, bb, f) = Xat(ap, b1, f) + AB[dir] * Xat(aa, b1, f);
        }
    }
  // ket HRR: Y[ab][c][d]
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
  std::vector<int> ax, ay, az, bx, by, bz, cx, cy, cz, dx, dy, dz;
  cart_list(la, ax, ay, az); cart_list(lb, bx, by, bz); cart_list(lc, cx, cy, cz); cart_list(ld, dx, dy, dz);
  out.assign((size_t)na * nb * nc * nd, 0.0);
  std::vector<R> Y((size_t)nF * nD);
  auto Yat = [&](int c_, int d_) -> R& { return Y[(size_t)c_ * nD + d_]; };
  for (int ia = 0; ia < na; ++ia) for (int ib = 0; ib < nb; ++ib) {
    int aa = tri_index(ax[ia], ay[ia], az[ia]), bb = tri_index(bx[ib], by[ib], bz[ib]);
    for (int f = 0; f < nF; ++f) Yat(f, 0) = Xat(aa, bb, f);
    for (int ldd = 1; ldd <= ld; ++ldd)
      for (int i = 0; i <= ldd; ++i) for (int j = 0; j <= i; ++j) {
        int x = ldd - i, y = i - j, z = ldd - x - y, dd = tri_index(x, y, z);
        int dir = x > 0 ? 0 : (y > 0 ? 1 : 2);
        int d1 = tri_index(x - (dir == 0), y - (dir == 1), z - (dir == 2));
        for (int lcc = lc; lcc <= Lcd - ldd; ++lcc)
          for (int i2 = 0; i2 <= lcc; ++i2) for (int j2 = 0; j2 <= i2; ++j2) {
            int qx = lcc - i2, qy = i2 - j2, qz = lcc - qx - qy, cc = tri_index(qx, qy, qz);
            int cp = tri_index(qx + (dir == 0), qy + (dir == 1), qz + (dir == 2));
            Yat(cc, dd) = Yat(cp, d1) + CD[dir] * Yat(cc, d1);
          }
      }
    for (int ic = 0; ic < nc; ++ic) for (int id = 0; id < nd; ++id)
      out[(((size_t)ia * nb + ib) * nc + ic) * nd + id] =
          (double)Yat(tri_index(cx[ic], cy[ic], cz[ic]), tri_index(dx[id], dy[id], dz[id]));
  }
}

// ---------------------------------------------------------------- MD ERI (long double)
typedef long double LD;
// Hermite expansion coefficients E^{ij}_t for one Cartesian direction.
void hermite_E(int imax, int jmax, LD a, LD b, LD Ax, LD Bx, std::vector<LD>& E) {
  // E[(i*(jmax+1)+j)*(imax+jmax+1)+t]
  int T = imax + jmax + 1;
  E.assign((size_t)(imax + 1) * (jmax + 1) * T, 0.0L);
  LD p = a + b, mu = a * b / p, X = Ax - Bx;
  LD Px = (a * Ax + b * Bx) / p, PA = Px - Ax, PB = Px - Bx;
  auto at = [&](int i, int j, int t) -> LD& { return E[((size_t)i * (jmax + 1) + j) * T + t]; };
  at(0, 0, 0) = std::exp(-mu * X * X);
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        LD v;
        if (i > 0) {
          v = PA * at(i - 1, j, t);
          if (t > 0) v += at(i - 1, j, t - 1) / (2 * p);
          if (t + 1 <= i - 1 + j) v += (t + 1) * at(i - 1, j, t + 1);
        } else {
          v = PB * at(i, j - 1, t);
          if (t > 0) v += at(i, j - 1, t - 1) / (2 * p);
          if (t + 1 <= i + j - 1) v += (t + 1) * at(i, j - 1, t + 1);
        }
        at(i, j, t) = v;
      }
    }
}

// Hermite Coulomb integrals R_{tuv} for t+u+v <= L; R[(t*(L+1)+u)*(L+1)+v].
void hermite_R(int L, LD alpha, LD X, LD Y, LD Z, LD pref, std::vector<LD>& Rout) {
  int n1 = L + 1;
  LD Fb[4 * kMaxL + 2];
  boys_ld(L, alpha * (X * X + Y * Y + Z * Z), Fb);
  std::vector<LD> Rn((size_t)(L + 1) * n1 * n1 * n1, 0.0L);
  auto at = [&](int n, int t, int u, int v) -> LD& { return Rn[(((size_t)n * n1 + t) * n1 + u) * n1 + v]; };
  LD f = 1.0L;
  for (int n = 0; n <= L; ++n) { at(n, 0, 0, 0) = pref * f * Fb[n]; f *= -2.0L * alpha; }
  for (int s = 1; s <= L; ++s)
    for (int n = 0; n <= L - s; ++n)
      for (int t = 0; t <= s; ++t)
        for (int u = 0; u <= s - t; ++u) {
          int v = s - t - u;
          LD r;
          if (t > 0) { r = X * at(n + 1, t - 1, u, v); if (t > 1) r += (t - 1) * at(n + 1, t - 2, u, v); }
          else if (u > 0) { r = Y * at(n + 1, t, u - 1, v); if (u > 1) r += (u - 1) * at(n + 1, t, u - 2, v); }
          else { r = Z * at(n + 1, t, u, v - 1); if (v > 1) r += (v - 1) * at(n + 1, t, u, v - 2); }
          at(n, t, u, v) = r;
        }
  Rout.assign((size_t)n1 * n1 * n1, 0.0L);
  for (int t = 0; t <= L; ++t) for (int u = 0; u <= L - t; ++u) for (int v = 0; v <= L - t - u; ++v)
    Rout[((size_t)t * n1 + u) * n1 + v] = at(0, t, u, v);
}

void eri_cart_md(const Shell& A, const Shell& B, const Shell& C, const Shell& D, std::vector<double>& out) {
  const int la = A.l, lb = B.l, lc = C.l, ld = D.l, L = la + lb + lc + ld, n1 = L + 1;
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
  std::vector<int> ax, ay, az, bx, by, bz, cx, cy, cz, dx, dy, dz;
  cart_list(la, ax, ay, az); cart_list(lb, bx, by, bz); cart_list(lc, cx, cy, cz); cart_list(ld, dx, dy, dz);
  std::vector<LD> acc((size_t)na * nb * nc * nd, 0.0L);
  const LD pi = 3.14159265358979323846264338327950288L;
  std::vector<LD> Eab[3], Ecd[3], R;
  const int Tab = la + lb + 1, Tcd = lc + ld + 1;
  for (int pa = 0; pa < A.nprim; ++pa) for (int pb = 0; pb < B.nprim; ++pb) {
    LD a = A.a[pa], b = B.a[pb], p = a + b;
    for (int k = 0; k < 3; ++k) hermite_E(la, lb, a, b, A.O[k], B.O[k], Eab[k]);
    LD P[3]; for (int k = 0; k < 3; ++k) P[k] = (a * (LD)A.O[k] + b * (LD)B.O[k]) / p;
    LD cab = (LD)A.c[pa] * (LD)B.c[pb];
    for (int pc = 0; pc < C.nprim; ++pc) for (int pd = 0; pd < D.nprim; ++pd) {
      LD c = C.a[pc], d = D.a[pd], q = c + d;
      for (int k = 0; k < 3; ++k) hermite_E(lc, ld, c, d, C.O[k], D.O[k], Ecd[k]);
      LD Q[3]; for (int k = 0; k < 3; ++k) Q[k] = (c * (LD)C.O[k] + d * (LD)D.O[k]) / q;
      LD alpha = p * q / (p + q);
      LD pref = 2.0L * std::pow(pi, 2.5L) / (p * q * std::sqrt(p + q)) * cab * (LD)C.c[pc] * (LD)D.c[pd];
      hermite_R(L, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], pref, R);
      for (int ia = 0; ia < na; ++ia) for (int ib = 0; ib < nb; ++ib)
      for (int ic = 0; ic < nc; ++ic) for (int id = 0; id < nd; ++id) {
        LD s = 0;
        const int ix = ax[ia], jx = bx[ib], kx = cx[ic], lx = dx[id];
        const int iy = ay[ia], jy = by[ib], ky = cy[ic], ly = dy[id];
        const int iz = az[ia], jz = bz[ib], kz = cz[ic], lz = dz[id];
        for (int t = 0; t <= ix + jx; ++t) {
          LD e1 = Eab[0][((size_t)ix * (lb + 1) + jx) * Tab + t]; if (e1 == 0) continue;
        for (int u = 0; u <= iy + jy; ++u) {
          LD e2 = e1 * Eab[1][((size_t)iy * (lb + 1) + jy) * Tab + u]; if (e2 == 0) continue;
        for (int v = 0; v <= iz + jz; ++v) {
          LD e3 = e2 * Eab[2][((size_t)iz * (lb + 1) + jz) * Tab + v]; if (e3 == 0) continue;
          for (int tau = 0; tau <= kx + lx; ++tau) {
            LD f1 = Ecd[0][((size_t)kx * (ld + 1) + lx) * Tcd + tau];
          for (int nu = 0; nu <= ky + ly; ++nu) {
            LD f2 = f1 * Ecd[1][((size_t)ky * (ld + 1) + ly) * Tcd + nu];
          for (int ph = 0; ph <= kz + lz; ++ph) {
            LD f3 = f2 * Ecd[2][((size_t)kz * (ld + 1) + lz) * Tcd + ph];
            LD sg = ((tau + nu + ph) & 1) ? -1.0L : 1.0L;
            s += e3 * sg * f3 * R[((size_t)(t + tau) * n1 + (u + nu)) * n1 + (v + ph)];
          }}}
        }}}
        acc[(((size_t)ia * nb + ib) * nc + ic) * nd + id] += s;
      }
    }
  }
  out.resize(acc.size());
  for (size_t i = 0; i < acc.size(); ++i) out[i] = (double)acc[i];
}

void quartet(const Basis& bs, int s1, int s2, int s3, int s4, std::vector<double>& buf, int algo) {
  const Shell* s[4] = {&bs.sh[s1], &bs.sh[s2], &bs.sh[s3], &bs.sh[s4]};
  if (algo == 0) eri_cart_os<double>(*s[0], *s[1], *s[2], *s[3], buf);
  else if (algo == 1) eri_cart_os<long double>(*s[0], *s[1], *s[2], *s[3], buf);
  else eri_cart_md(*s[0], *s[1], *s[2], *s[3], buf);
  cart_to_shell_basis(s, buf);
}

// ---------------------------------------------------------------- one-electron (MD)
// A Cartesian block of shells A, B (row-major na x nb) in the shells' basis, written into the
// column-major N x N matrix dst.
void scatter_block(const Shell& A, const Shell& B, const std::vector<LD>& src, double* dst, int N) {
  if (!dst) return;
  const int la = A.l, lb = B.l, na = ncart(la), nb = ncart(lb);
  std::vector<double> blk(src.size());
  for (size_t i = 0; i < src.size(); ++i) blk[i] = (double)src[i];
  int dA = na, dB = nb;
  if (A.pure) {
    const auto& M = car2sph_cached(la); int ns = 2 * la + 1;
    std::vector<double> o((size_t)ns * dB, 0.0);
    for (int i = 0; i < ns; ++i) for (int c = 0; c < na; ++c) for (int j = 0; j < dB; ++j)
      o[i * dB + j] += M[i * na + c] * blk[c * dB + j];
    blk.swap(o); dA = ns;
  }
  if (B.pure) {
    const auto& M = car2sph_cached(lb); int ns = 2 * lb + 1;
    std::vector<double> o((size_t)dA * ns, 0.0);
    for (int i = 0; i < dA; ++i) for (int j = 0; j < ns; ++j) for (int c = 0; c < nb; ++c)
      o[i * ns + j] += M[j * nb + c] * blk[i * nb + c];
    blk.swap(o); dB = ns;
  }
  for (int i = 0; i < dA; ++i) for (int j = 0; j < dB; ++j)
    dst[(size_t)(B.off + j) * N + (A.off + i)] = blk[i * dB + j];
}

void overlap_kinetic_nuclear(const Basis& bs, int natom, const double* Z, const double* xyz,
                             double* S, double* T, double* V) {
  const int N = bs.nbf;
  const LD pi = 3.14159265358979323846264338327950288L;
  for (int s1 = 0; s1 < (int)bs.sh.size(); ++s1)
  for (int s2 = 0; s2 < (int)bs.sh.size(); ++s2) {
    const Shell& A = bs.sh[s1]; const Shell& B = bs.sh[s2];
    int la = A.l, lb = B.l, na = ncart(la), nb = ncart(lb);
    std::vector<int> ax, ay, az, bx, by, bz;
    cart_list(la, ax, ay, az); cart_list(lb, bx, by, bz);
    std::vector<LD> s_((size_t)na * nb, 0), t_((size_t)na * nb, 0), v_((size_t)na * nb, 0);
    std::vector<LD> E[3], R;
    const int TT = la + lb + 3;   // E built with lb+2 for kinetic
    for (int pa = 0; pa < A.nprim; ++pa) for (int pb = 0; pb < B.nprim; ++pb) {
      LD a = A.a[pa], b = B.a[pb], p = a + b;
      LD cab = (LD)A.c[pa] * (LD)B.c[pb];
      for (int k = 0; k < 3; ++k) hermite_E(la, lb + 2, a, b, A.O[k], B.O[k], E[k]);
      LD P[3]; for (int k = 0; k < 3; ++k) P[k] = (a * (LD)A.O[k] + b * (LD)B.O[k]) / p;
      auto E0 = [&](int k, int i, int j) { return E[k][((size_t)i * (lb + 3) + j) * TT + 0]; };
      LD sq = std::sqrt(pi / p);
      // nuclear Hermite integrals per atom
      std::vector<std::vector<LD>> Rn(natom);
      for (int c = 0; c < natom; ++c)
        hermite_R(la + lb, p, P[0] - xyz[3 * c], P[1] - xyz[3 * c + 1], P[2] - xyz[3 * c + 2], 1.0L, Rn[c]);
      int n1 = la + lb + 1;
      for (int ia = 0; ia < na; ++ia) for (int ib = 0; ib < nb; ++ib) {
        int i3[3] = {ax[ia], ay[ia], az[ia]}, j3[3] = {bx[ib], by[ib], bz[ib]};
        LD s1d[3], t1d[3];
        for (int k = 0; k < 3; ++k) {
          int i = i3[k], j = j3[k];
          s1d[k] = E0(k, i, j) * sq;
          LD d2 = 4 * b * b * E0(k, i, j + 2) * sq - 2 * b * (2 * j + 1) * E0(k, i, j) * sq;
          if (j >= 2) d2 += (LD)j * (j - 1) * E0(k, i, j - 2) * sq;
          t1d[k] = -0.5L * d2;
        }
        s_[ia * nb + ib] += cab * s1d[0] * s1d[1] * s1d[2];
        t_[ia * nb + ib] += cab * (t1d[0] * s1d[1] * s1d[2] + s1d[0] * t1d[1] * s1d[2] + s1d[0] * s1d[1] * t1d[2]);
        LD vv = 0;
        for (int c = 0; c < natom; ++c) {
          LD sum = 0;
          for (int t = 0; t <= i3[0] + j3[0]; ++t)
          for (int u = 0; u <= i3[1] + j3[1]; ++u)
          for (int v = 0; v <= i3[2] + j3[2]; ++v)
            sum += E[0][((size_t)i3[0] * (lb + 3) + j3[0]) * TT + t] *
                   E[1][((size_t)i3[1] * (lb + 3) + j3[1]) * TT + u] *
                   E[2][((size_t)i3[2] * (lb + 3) + j3[2]) * TT + v] *
                   Rn[c][((size_t)t * n1 + u) * n1 + v];
          vv += -(LD)Z[c] * sum;
        }
        v_[ia * nb + ib] += cab * 2.0L * pi / p * vv;
      }
    }
    // to shell basis and scatter (column-major N x N)
    scatter_block(A, B, s_, S, N);
    scatter_block(A, B, t_, T, N);
    scatter_block(A, B, v_, V, N);
  }
}

// Cartesian multipole integrals <a| x_C^ex y_C^ey z_C^ez |b> (restating what libint2's
// emultipole engines give AOIntegrals::OneEDriver, src/aointegrals/aointegrals_builders.cpp:
// 38-55) by Hermite moments, a different route from the device's: per direction
//   <i| x_C^e |j> = sum_t E^{ij}_t M^e_t,   M^0_0 = sqrt(pi / p),
//   M^{e+1}_t = t M^e_{t-1} + X_PC M^e_t + M^e_{t+1} / (2p)
// (Helgaker, Jorgensen, Olsen, Molecular Electronic-Structure Theory, eq. 9.5.x).  Operators of
// orders 1 .. order in the order x, y, z; xx, xy, xz, yy, yz, zz; xxx, xxy, ..., zzz; out holds
// one column-major N x N matrix per operator.
void multipole_ints(const Basis& bs, int order, const double* Cin, double* out) {
  const int N = bs.nbf;
  const LD pi = 3.14159265358979323846264338327950288L;
  std::vector<std::array<int, 3>> ops;
  for (int e = 1; e <= order; ++e)
    for (int i = e; i >= 0; --i)
      for (int j = e - i; j >= 0; --j) ops.push_back({i, j, e - i - j});
  const LD C[3] = {Cin[0], Cin[1], Cin[2]};
  for (int s1 = 0; s1 < (int)bs.sh.size(); ++s1)
  for (int s2 = 0; s2 < (int)bs.sh.size(); ++s2) {
    const Shell& A = bs.sh[s1]; const Shell& B = bs.sh[s2];
    const int la = A.l, lb = B.l, na = ncart(la), nb = ncart(lb);
    std::vector<int> ax, ay, az, bx, by, bz;
    cart_list(la, ax, ay, az); cart_list(lb, bx, by, bz);
    std::vector<std::vector<LD>> m_(ops.size(), std::vector<LD>((size_t)na * nb, 0));
    std::vector<LD> E[3];
    const int TT = la + lb + 1;
    for (int pa = 0; pa < A.nprim; ++pa) for (int pb = 0; pb < B.nprim; ++pb) {
      const LD a = A.a[pa], b = B.a[pb], p = a + b;
      const LD cab = (LD)A.c[pa] * (LD)B.c[pb];
      // Hermite moments M[k][e][t], t <= e + la + lb
      LD M[3][4][8] = {};
      for (int k = 0; k < 3; ++k) {
        hermite_E(la, lb, a, b, A.O[k], B.O[k], E[k]);
        const LD XPC = (a * (LD)A.O[k] + b * (LD)B.O[k]) / p - C[k];
        M[k][0][0] = std::sqrt(pi / p);
        for (int e = 0; e < 3; ++e)
          for (int t = 0; t <= e + 1; ++t) {
            LD v = XPC * M[k][e][t] + M[k][e][t + 1] / (2 * p);
            if (t > 0) v += t * M[k][e][t - 1];
            M[k][e + 1][t] = v;
          }
      }
      auto mom = [&](int k, int i, int j, int e) {
        LD sum = 0;
        for (int t = 0; t <= std::min(i + j, e); ++t) sum += E[k][((size_t)i * (lb + 1) + j) * TT + t] * M[k][e][t];
        return sum;
      };
      for (size_t o = 0; o < ops.size(); ++o)
        for (int ia = 0; ia < na; ++ia) for (int ib = 0; ib < nb; ++ib)
          m_[o][ia * nb + ib] += cab * mom(0, ax[ia], bx[ib], ops[o][0]) * mom(1, ay[ia], by[ib], ops[o][1]) *
                                 mom(2, az[ia], bz[ib], ops[o][2]);
    }
    for (size_t o = 0; o < ops.size(); ++o) scatter_block(A, B, m_[o], out + o * (size_t)N * N, N);
  }
}

}  // namespace

// =====================================================================================
// C ABI (ctypes)
// =====================================================================================
extern "C" {

void* orc_basis_create(int nshell, const int* l, const int* pure, const int* nprim,
                       const double* alpha, const double* coef, const double* origin) {
  Basis* b = new Basis();
  int off = 0, po = 0;
  for (int i = 0; i < nshell; ++i) {
    Shell s;
    s.l = l[i]; s.pure = pure[i]; s.nprim = nprim[i];
    s.size = s.pure ? 2 * s.l + 1 : ncart(s.l);
    s.off = off; off += s.size;
    s.a.assign(alpha + po, alpha + po + s.nprim);
    s.c.assign(coef + po, coef + po + s.nprim);
    po += s.nprim;
    for (int k = 0; k < 3; ++k) s.O[k] = origin[3 * i + k];
    b->maxl = std::max(b->maxl, s.l);
    b->maxprim = std::max(b->maxprim, s.nprim);
    b->sh.push_back(s);
  }
  b->nbf = off;
  return b;
}

void orc_basis_destroy(void* h) { delete static_cast<Basis*>(h); }

int orc_nbf(void* h) { return static_cast<Basis*>(h)->nbf; }

void orc_boys(int mmax, double T, double* F) {
  long double Fl[64];
  boys_ld(mmax, (long double)T, Fl);
  for (int m = 0; m <= mmax; ++m) F[m] = (double)Fl[m];
}

// Contracted quartet in the shells' own basis, row-major (i,j,k,l).
// algo: 0 = OS/HGP double, 1 = OS/HGP long double, 2 = MD long double.
int orc_quartet(void* h, int s1, int s2, int s3, int s4, int algo, double* out) {
  const Basis& bs = *static_cast<Basis*>(h);
  std::vector<double> buf;
  quartet(bs, s1, s2, s3, s4, buf, algo);
  std::memcpy(out, buf.data(), buf.size() * sizeof(double));
  return (int)buf.size();
}

// Cartesian-to-spherical matrix (2l+1) x ncart(l), row-major.
void orc_car2sph(int l, double* out) {
  auto M = car2sph(l);
  std::memcpy(out, M.data(), M.size() * sizeof(double));
}

// Q(s1,s2) = sqrt(max_ij |(ij|ij)|), symmetric nshell x nshell.
void orc_schwarz(void* h, double* Q, int nthreads) {
  const Basis& bs = *static_cast<Basis*>(h);
  const int ns = (int)bs.sh.size();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
  for (int s1 = 0; s1 < ns; ++s1) {
    std::vector<double> buf;
    for (int s2 = 0; s2 <= s1; ++s2) {
      quartet(bs, s1, s2, s1, s2, buf, 0);
      int n1 = bs.sh[s1].size, n2 = bs.sh[s2].size;
      double mx = 0.0;
      for (int ij = 0; ij < n1 * n2; ++ij) mx = std::max(mx, std::fabs(buf[(size_t)ij * n1 * n2 + ij]));
      Q[(size_t)s1 * ns + s2] = std::sqrt(mx);
      Q[(size_t)s2 * ns + s1] = std::sqrt(mx);
    }
  }
}

// Block infinity norms, nshell x nshell column-major, of a column-major N x N matrix.
void orc_shblk_norm(void* h, const double* X, double* out) {
  const Basis& bs = *static_cast<Basis*>(h);
  const int ns = (int)bs.sh.size(), N = bs.nbf;
  for (int s1 = 0; s1 < ns; ++s1) for (int s2 = 0; s2 < ns; ++s2) {
    double mx = 0.0;
    for (int i = 0; i < bs.sh[s1].size; ++i) for (int j = 0; j < bs.sh[s2].size; ++j)
      mx = std::max(mx, std::fabs(X[(size_t)(bs.sh[s2].off + j) * N + bs.sh[s1].off + i]));
    out[(size_t)s2 * ns + s1] = mx;
  }
}

// Restatement of AOIntegrals::newTwoEContractDirect (aointegrals_contract.h:24-194).
// types: 0 = COULOMB, 1 = EXCHANGE.  X, AX column-major N x N (ld = N).  AX accumulates.
// Q: Schwarz matrix (nshell^2); tau: threshold.  Returns the number of surviving quartets.
long orc_contract_direct(void* h, const double* Q, double tau, int nmat, const double* const* X,
                         double* const* AX, const int* types, const double* scales, int nthreads) {
  const Basis& bs = *static_cast<Basis*>(h);
  const int ns = (int)bs.sh.size(), N = bs.nbf;
  std::vector<std::vector<double>> norms(nmat, std::vector<double>((size_t)ns * ns));
  for (int i = 0; i < nmat; ++i) orc_shblk_norm(h, X[i], norms[i].data());
  auto nrm = [&](int i, int a, int b) { return norms[i][(size_t)b * ns + a]; };
  auto q = [&](int a, int b) { return Q[(size_t)b * ns + a]; };
  std::vector<std::vector<double>> G((size_t)nthreads * nmat, std::vector<double>((size_t)N * N, 0.0));
  long nsurv_total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : nsurv_total)
  {
    const int tid = omp_get_thread_num();
    std::vector<double> buf;
    long nsurv = 0;
    size_t s1234 = 0;
    for (int s1 = 0; s1 < ns; ++s1) {
      const int bf1s = bs.sh[s1].off, n1 = bs.sh[s1].size;
      for (int s2 = 0; s2 <= s1; ++s2) {
        const int bf2s = bs.sh[s2].off, n2 = bs.sh[s2].size;
        for (int s3 = 0; s3 <= s1; ++s3) {
          const int bf3s = bs.sh[s3].off, n3 = bs.sh[s3].size;
          const int s4max = (s1 == s3) ? s2 : s3;
          for (int s4 = 0; s4 <= s4max; ++s4, ++s1234) {
            if ((int)(s1234 % (size_t)nthreads) != tid) continue;
            const int bf4s = bs.sh[s4].off, n4 = bs.sh[s4].size;
            double shMax = std::max(nrm(0, s1, s4), std::max(nrm(0, s2, s4), std::max(nrm(0, s3, s4),
                           std::max(nrm(0, s1, s3), std::max(nrm(0, s2, s3), nrm(0, s1, s2)))))) *
                           q(s1, s2) * q(s3, s4);
            for (int m = 1; m < nmat; ++m)
              shMax = std::max(shMax, std::max(nrm(m, s1, s4), std::max(nrm(m, s2, s4), std::max(nrm(m, s3, s4),
                           std::max(nrm(m, s1, s3), std::max(nrm(m, s2, s3), nrm(m, s1, s2)))))) *
                           q(s1, s2) * q(s3, s4));
            if (shMax < tau) continue;
            ++nsurv;
            const double s12 = (s1 == s2) ? 1.0 : 2.0;
            const double s34 = (s3 == s4) ? 1.0 : 2.0;
            const double s1234d = (s1 == s3) ? (s2 == s4 ? 1.0 : 2.0) : 2.0;
            const double deg = s12 * s34 * s1234d;
            quartet(bs, s1, s2, s3, s4, buf, 0);
            for (int m = 0; m < nmat; ++m) {
              if (scales[m] == 0.0) continue;
              double* g = G[(size_t)tid * nmat + m].data();
              const double* x = X[m];
              auto Xa = [&](int r, int c) { return x[(size_t)c * N + r]; };
              auto Ga = [&](int r, int c) -> double& { return g[(size_t)c * N + r]; };
              size_t ijkl = 0;
              for (int i = 0, bf1 = bf1s; i < n1; ++i, ++bf1)
              for (int j = 0, bf2 = bf2s; j < n2; ++j, ++bf2)
              for (int k = 0, bf3 = bf3s; k < n3; ++k, ++bf3)
              for (int l = 0, bf4 = bf4s; l < n4; ++l, ++bf4, ++ijkl) {
                const double v = buf[ijkl] * deg;
                if (types[m] == 0) {
                  Ga(bf1, bf2) += Xa(bf4, bf3) * v;
                  Ga(bf3, bf4) += Xa(bf2, bf1) * v;
                  Ga(bf2, bf1) += Xa(bf3, bf4) * v;
                  Ga(bf4, bf3) += Xa(bf1, bf2) * v;
                } else {
                  Ga(bf1, bf3) += 0.5 * Xa(bf2, bf4) * v;
                  Ga(bf2, bf4) += 0.5 * Xa(bf1, bf3) * v;
                  Ga(bf1, bf4) += 0.5 * Xa(bf2, bf3) * v;
                  Ga(bf2, bf3) += 0.5 * Xa(bf1, bf4) * v;
                  Ga(bf3, bf1) += 0.5 * Xa(bf4, bf2) * v;
                  Ga(bf4, bf2) += 0.5 * Xa(bf3, bf1) * v;
                  Ga(bf4, bf1) += 0.5 * Xa(bf3, bf2) * v;
                  Ga(bf3, bf2) += 0.5 * Xa(bf4, bf1) * v;
                }
              }
            }
          }
        }
      }
    }
    nsurv_total += nsurv;
  }
  for (int m = 0; m < nmat; ++m)
    for (int t = 0; t < nthreads; ++t) {
      const double* g = G[(size_t)t * nmat + m].data();
      for (size_t i = 0; i < (size_t)N * N; ++i) AX[m][i] += 0.5 * scales[m] * g[i];
    }
  return nsurv_total;
}

// Full N^4 tensor (ij|kl), row-major, for the brute-force contraction.
void orc_eri_tensor(void* h, double* eri, int nthreads) {
  const Basis& bs = *static_cast<Basis*>(h);
  const int ns = (int)bs.sh.size();
  const size_t N = bs.nbf;
#pragma omp parallel for schedule(dynamic) collapse(2) num_threads(nthreads)
  for (int s1 = 0; s1 < ns; ++s1)
  for (int s2 = 0; s2 < ns; ++s2) {
    std::vector<double> buf;
    for (int s3 = 0; s3 < ns; ++s3)
    for (int s4 = 0; s4 < ns; ++s4) {
      quartet(bs, s1, s2, s3, s4, buf, 0);
      const Shell &A = bs.sh[s1], &B = bs.sh[s2], &C = bs.sh[s3], &D = bs.sh[s4];
      size_t ijkl = 0;
      for (int i = 0; i < A.size; ++i) for (int j = 0; j < B.size; ++j)
      for (int k = 0; k < C.size; ++k) for (int l = 0; l < D.size; ++l, ++ijkl)
        eri[(((A.off + i) * N + (B.off + j)) * N + (C.off + k)) * N + (D.off + l)] = buf[ijkl];
    }
  }
}

// S, T, V column-major N x N.  Z, xyz: nuclear charges and positions (bohr).
void orc_one_electron(void* h, int natom, const double* Z, const double* xyz, double* S, double* T, double* V) {
  overlap_kinetic_nuclear(*static_cast<Basis*>(h), natom, Z, xyz, S, T, V);
}

// Multipole integrals of orders 1 .. order (3, 9 or 19 column-major N x N matrices) about C.
void orc_multipole(void* h, int order, const double* C, double* out) {
  multipole_ints(*static_cast<Basis*>(h), order, C, out);
}

// Screening decisions only (the test of aointegrals_contract.h:111-131): number of canonical
// quartets that survive, for a combined norm matrix Nmax (nshell^2, column-major) and Q.
long orc_count_screened(int ns, const double* Nmax, const double* Q, double tau) {
  auto nrm = [&](int a, int b) { return Nmax[(size_t)b * ns + a]; };
  auto q = [&](int a, int b) { return Q[(size_t)b * ns + a]; };
  long n = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : n)
  for (int s1 = 0; s1 < ns; ++s1)
    for (int s2 = 0; s2 <= s1; ++s2)
      for (int s3 = 0; s3 <= s1; ++s3) {
        const int s4max = (s1 == s3) ? s2 : s3;
        for (int s4 = 0; s4 <= s4max; ++s4) {
          const double shMax = std::max(nrm(s1, s4), std::max(nrm(s2, s4), std::max(nrm(s3, s4),
                               std::max(nrm(s1, s3), std::max(nrm(s2, s3), nrm(s1, s2)))))) *
                               q(s1, s2) * q(s3, s4);
          if (!(shMax < tau)) ++n;
        }
      }
  return n;
}

}  // extern "C"

