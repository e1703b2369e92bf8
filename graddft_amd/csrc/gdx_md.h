// McMurchie-Davidson machinery shared by the integral kernels on the grid (gdx_hf.hip) and the
// electron-repulsion integrals (gdx_eri.hip).
//
// For a primitive pair (exponents a, b on centres A, B; p = a + b, P = (a A + b B) / p) the
// Gaussian product is a sum of Hermite Gaussians, x_A^i x_B^j e^{..} = sum_t E^{ij}_t Lambda_t,
// and a Coulomb-type integral against it is a sum over the Hermite Coulomb integrals
// R_tuv(alpha, X, Y, Z), built from Boys functions F_n(alpha |XYZ|^2).  The point-independent
// part of a primitive pair -- E coefficients, contraction coefficients, 2 pi / p and the
// Cartesian-to-AO component matrices -- is folded on the host into one matrix
// M[a][b][tuv] per primitive pair (pair_matrix), so device code only evaluates R.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace gdx {
namespace md {
namespace {

constexpr int LMAX = 3;           // f shells
constexpr int MAXF = 10;          // AO components of a shell (Cartesian f)
constexpr double PI = 3.14159265358979323846;

// Boys function table: F_n(T_k) for T_k = k h, k < BOYS_ROWS, n < NCOL; Taylor expansion about
// the nearest T_k with BOYS_TAYLOR terms (|T - T_k| <= h / 2: remainder below
// (h/2)^7 / 7! ~ 1.5e-13 F_{n+7} <= 1e-14); T >= BOYS_TMAX: the asymptotic F_0 = sqrt(pi/T)/2
// (erfc(sqrt(40)) < 1e-18) and upward recursion.
constexpr int BOYS_TAYLOR = 7;
constexpr double BOYS_H = 0.1;
constexpr int BOYS_ROWS = 400;
constexpr double BOYS_TMAX = 39.9;

__host__ __device__ constexpr int n_tuv(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
// flat index of (t, u, v): shells of k = t + u + v in increasing k; inside a shell t descending,
// then u descending
__host__ __device__ constexpr int tuv_index(int t, int u, int v) {
  return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}

// F_n(T_k) for every table row: series for the highest order, then downward recursion
// F_{n-1} = (2 T F_n + e^-T) / (2n - 1) (stable).  One thread per row.
template <int NCOL>
__global__ void boys_table_kernel(double* __restrict__ tab) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= BOYS_ROWS) return;
  const double T = k * BOYS_H;
  const int nmax = NCOL - 1;
  // F_n(T) = e^-T sum_i (2T)^i / ((2n+1)(2n+3)...(2n+2i+1))
  double term = 1.0 / (2 * nmax + 1), sum = term;
  for (int i = 1; i < 2000; ++i) {
    term *= 2.0 * T / (2 * nmax + 2 * i + 1);
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  const double e = exp(-T);
  double F = e * sum;
  tab[k * NCOL + nmax] = F;
  for (int n = nmax; n > 0; --n) {
    F = (2.0 * T * F + e) / (2 * n - 1);
    tab[k * NCOL + n - 1] = F;
  }
}

template <int L, int NCOL>
__device__ __forceinline__ void boys_eval(double T, const double* __restrict__ tab, double (&F)[L + 1]) {
  static_assert(L + BOYS_TAYLOR <= NCOL, "Boys table too narrow for this order");
  if (T < BOYS_TMAX) {
    const int k = (int)(T * (1.0 / BOYS_H) + 0.5);
    const double d = k * BOYS_H - T;  // dF_n/dT = -F_{n+1}: F_n(T) = sum_j F_{n+j}(T_k) d^j / j!
    const double* row = tab + k * NCOL + L;
    double f = row[BOYS_TAYLOR - 1];
#pragma unroll
    for (int j = BOYS_TAYLOR - 2; j >= 0; --j) f = row[j] + f * d / (j + 1);
    F[L] = f;
    const double e = exp(-T);
#pragma unroll
    for (int n = L; n > 0; --n) F[n - 1] = (2.0 * T * F[n] + e) / (2 * n - 1);
  } else {
    F[0] = 0.5 * sqrt(PI / T);
    const double e = exp(-T), i2t = 0.5 / T;
#pragma unroll
    for (int n = 0; n < L; ++n) F[n + 1] = ((2 * n + 1) * F[n] - e) * i2t;
  }
}

// Hermite Coulomb integrals R_tuv = R^{(0)}_tuv for t + u + v <= L, in place over decreasing
// order n:  R^{(n)}_000 = (-2p)^n F_n,  R^{(n)}_{t+1,u,v} = t R^{(n+1)}_{t-1,u,v} + X R^{(n+1)}_{t,u,v}
// (likewise u, v).  Entries of sum k at level n read level n+1 entries of sum k-1 and k-2, which
// are still in place when a level is filled in decreasing k.
template <int L>
__device__ __forceinline__ void hermite_coulomb(double p, double X, double Y, double Z, const double (&F)[L + 1],
                                                double (&R)[n_tuv(L)]) {
  double m2p[L + 1];
  m2p[0] = 1.0;
#pragma unroll
  for (int n = 1; n <= L; ++n) m2p[n] = m2p[n - 1] * (-2.0 * p);
  R[0] = m2p[L] * F[L];
#pragma unroll
  for (int n = L - 1; n >= 0; --n) {
#pragma unroll
    for (int k = L - n; k >= 1; --k) {
#pragma unroll
      for (int t = k; t >= 0; --t) {
#pragma unroll
        for (int u = k - t; u >= 0; --u) {
          const int v = k - t - u;
          double val;
          if (t > 0) {
            val = X * R[tuv_index(t - 1, u, v)];
            if (t > 1) val += (t - 1) * R[tuv_index(t - 2, u, v)];
          } else if (u > 0) {
            val = Y * R[tuv_index(t, u - 1, v)];
            if (u > 1) val += (u - 1) * R[tuv_index(t, u - 2, v)];
          } else {
            val = Z * R[tuv_index(t, u, v - 1)];
            if (v > 1) val += (v - 1) * R[tuv_index(t, u, v - 2)];
          }
          R[tuv_index(t, u, v)] = val;
        }
      }
    }
    R[0] = m2p[n] * F[n];
  }
}

// ------------------------------------------------------------------------ host side ----
inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

struct Shell {
  int l, nprim, poff, aoff, nfunc, toff;
  double c[3];
};

inline Shell shell_view(const int* si, const double* sf, int s) {
  Shell x;
  x.l = si[6 * s];
  x.nprim = si[6 * s + 1];
  x.poff = si[6 * s + 2];
  x.aoff = si[6 * s + 3];
  x.nfunc = si[6 * s + 4];
  x.toff = si[6 * s + 5];
  for (int k = 0; k < 3; ++k) x.c[k] = sf[3 * s + k];
  return x;
}

// Hermite expansion coefficients of one Cartesian direction: E[i][j][t], i <= la, j <= lb.
struct Herm {
  double E[LMAX + 1][LMAX + 1][2 * LMAX + 1];
};

inline void hermite_1d(int la, int lb, double a, double b, double XAB, Herm& H) {
  std::memset(&H, 0, sizeof(H));
  const double p = a + b, mu = a * b / p;
  const double XPA = -b * XAB / p, XPB = a * XAB / p, o2p = 0.5 / p;
  H.E[0][0][0] = std::exp(-mu * XAB * XAB);
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0) {  // E^{i,j}_t = E^{i-1,j}_{t-1}/(2p) + X_PA E^{i-1,j}_t + (t+1) E^{i-1,j}_{t+1}
          v = XPA * H.E[i - 1][j][t];
          if (t > 0) v += o2p * H.E[i - 1][j][t - 1];
          if (t + 1 <= i - 1 + j) v += (t + 1) * H.E[i - 1][j][t + 1];
        } else {
          v = XPB * H.E[i][j - 1][t];
          if (t > 0) v += o2p * H.E[i][j - 1][t - 1];
          if (t + 1 <= i + j - 1) v += (t + 1) * H.E[i][j - 1][t + 1];
        }
        H.E[i][j][t] = v;
      }
    }
}

inline void cart_exponents(int l, int (*e)[3]) {
  int k = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y) {
      e[k][0] = x;
      e[k][1] = y;
      e[k][2] = l - x - y;
      ++k;
    }
}

// One primitive pair (a, ca) on A and (b, cb) on B: appends M[fa][fb][tuv] (nfA x nfB x
// n_tuv(la + lb)) = (2 pi / p) ca cb sum_{ia, jb} TA[fa][ia] TB[fb][jb] E^x E^y E^z to Mall and
// returns p and P.
inline void pair_matrix(const Shell& A, const Shell& B, double a, double ca, double b, double cb,
                        const double* tmat, std::vector<double>& Mall, double& p, double P[3]) {
  const int L = A.l + B.l, nA = ncart(A.l), nB = ncart(B.l), NT = n_tuv(L);
  int ea[MAXF][3], eb[MAXF][3];
  cart_exponents(A.l, ea);
  cart_exponents(B.l, eb);
  const double AB[3] = {A.c[0] - B.c[0], A.c[1] - B.c[1], A.c[2] - B.c[2]};
  p = a + b;
  for (int d = 0; d < 3; ++d) P[d] = (a * A.c[d] + b * B.c[d]) / p;
  const double pref = 2.0 * PI / p * ca * cb;
  Herm H[3];
  for (int d = 0; d < 3; ++d) hermite_1d(A.l, B.l, a, b, AB[d], H[d]);
  std::vector<double> mc((size_t)nA * nB * NT, 0.0);
  for (int ia = 0; ia < nA; ++ia)
    for (int jb = 0; jb < nB; ++jb) {
      const int tx = ea[ia][0] + eb[jb][0], ty = ea[ia][1] + eb[jb][1], tz = ea[ia][2] + eb[jb][2];
      for (int t = 0; t <= tx; ++t)
        for (int uu = 0; uu <= ty; ++uu)
          for (int v = 0; v <= tz; ++v)
            mc[((size_t)ia * nB + jb) * NT + tuv_index(t, uu, v)] =
                H[0].E[ea[ia][0]][eb[jb][0]][t] * H[1].E[ea[ia][1]][eb[jb][1]][uu] * H[2].E[ea[ia][2]][eb[jb][2]][v];
    }
  const double* TA = tmat + A.toff;
  const double* TB = tmat + B.toff;
  const size_t base = Mall.size();
  Mall.resize(base + (size_t)A.nfunc * B.nfunc * NT, 0.0);
  double* Mo = Mall.data() + base;
  for (int fa = 0; fa < A.nfunc; ++fa)
    for (int fb = 0; fb < B.nfunc; ++fb)
      for (int ia = 0; ia < nA; ++ia) {
        const double ta = TA[fa * nA + ia];
        if (ta == 0.0) continue;
        for (int jb = 0; jb < nB; ++jb) {
          const double tb = TB[fb * nB + jb];
          if (tb == 0.0) continue;
          const double w = pref * ta * tb;
          const double* src = mc.data() + ((size_t)ia * nB + jb) * NT;
          double* dst = Mo + ((size_t)fa * B.nfunc + fb) * NT;
          for (int q = 0; q < NT; ++q) dst[q] += w * src[q];
        }
      }
}

}  // namespace
}  // namespace md
}  // namespace gdx
