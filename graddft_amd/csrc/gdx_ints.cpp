// Host-side Gaussian integrals (overlap, kinetic, nuclear attraction, electron repulsion)
// by the McMurchie-Davidson scheme.  This is the build's stand-in for the PySCF calls the
// reference makes when it builds a Molecule: mf.get_ovlp / mf.get_hcore
// (grad_dft/interface/pyscf.py:828-829) and mol.intor("int2e") (pyscf.py:797).
//
// Shells are passed in the packed form produced by graddft_amd/interface/basis.py
// (pack_shells): per shell {l, nprim, prim_offset, ao_offset, nfunc, tmat_offset}, the
// centre, primitive (exponent, normalised coefficient) pairs and the component matrix
// T (nfunc x ncart) mapping Cartesian monomials to AO components.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

constexpr int LMAX = 3;
constexpr int NCART_MAX = 10;
constexpr double PI = 3.14159265358979323846;

struct ShellView {
  int l, nprim, poff, aoff, nfunc, toff;
  double c[3];
};

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

void cart_exps(int l, int (*out)[3]) {
  int k = 0;
  for (int a = l; a >= 0; --a)
    for (int b = l - a; b >= 0; --b) {
      out[k][0] = a; out[k][1] = b; out[k][2] = l - a - b; ++k;
    }
}

// Boys function F_n(T) for n = 0..nmax.
void boys(int nmax, double T, double* F) {
  if (T < 1e-14) {
    for (int n = 0; n <= nmax; ++n) F[n] = 1.0 / (2 * n + 1) - T / (2 * n + 3);
    return;
  }
  double eT = std::exp(-T);
  if (T < 30.0) {
    // series for F_nmax then downward recursion (stable)
    double term = 1.0 / (2 * nmax + 1);
    double sum = term;
    for (int k = 1; k < 500; ++k) {
      term *= 2.0 * T / (2 * nmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    F[nmax] = eT * sum;
    for (int n = nmax; n > 0; --n) F[n - 1] = (2.0 * T * F[n] + eT) / (2 * n - 1);
  } else {
    F[0] = 0.5 * std::sqrt(PI / T) * std::erf(std::sqrt(T));
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - eT) / (2.0 * T);
  }
}

// Hermite expansion coefficients E[i][j][t] for one Cartesian direction.
struct Herm1D {
  double E[LMAX + 3][LMAX + 3][2 * LMAX + 6];
};

void hermite(int imax, int jmax, double a, double b, double XAB, Herm1D& H) {
  const double p = a + b, mu = a * b / p;
  const double XPA = -b * XAB / p, XPB = a * XAB / p;
  const double o2p = 0.5 / p;
  std::memset(&H, 0, sizeof(H));
  H.E[0][0][0] = std::exp(-mu * XAB * XAB);
  for (int i = 0; i <= imax; ++i) {
    for (int j = 0; j <= jmax; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0) {
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
}

// Hermite Coulomb integrals R_{tuv}^0 for t+u+v <= L.
struct RTab {
  int L;
  std::vector<double> R;  // (L+1)^3 indexed [t][u][v]
  double& at(int t, int u, int v) { return R[(t * (L + 1) + u) * (L + 1) + v]; }
};

void hermite_coulomb(int L, double p, double X, double Y, double Z, RTab& out) {
  const int n1 = L + 1;
  std::vector<double> buf((size_t)(L + 1) * n1 * n1 * n1, 0.0);
  auto idx = [&](int n, int t, int u, int v) { return ((n * n1 + t) * n1 + u) * n1 + v; };
  double F[4 * LMAX + 2];
  boys(L, p * (X * X + Y * Y + Z * Z), F);
  double m2p = 1.0;
  for (int n = 0; n <= L; ++n) { buf[idx(n, 0, 0, 0)] = m2p * F[n]; m2p *= -2.0 * p; }
  for (int n = L - 1; n >= 0; --n) {
    const int lmax = L - n;
    for (int t = 0; t <= lmax; ++t)
      for (int u = 0; u <= lmax - t; ++u)
        for (int v = 0; v <= lmax - t - u; ++v) {
          if (t == 0 && u == 0 && v == 0) continue;
          double val;
          if (t > 0) {
            val = X * buf[idx(n + 1, t - 1, u, v)];
            if (t > 1) val += (t - 1) * buf[idx(n + 1, t - 2, u, v)];
          } else if (u > 0) {
            val = Y * buf[idx(n + 1, t, u - 1, v)];
            if (u > 1) val += (u - 1) * buf[idx(n + 1, t, u - 2, v)];
          } else {
            val = Z * buf[idx(n + 1, t, u, v - 1)];
            if (v > 1) val += (v - 1) * buf[idx(n + 1, t, u, v - 2)];
          }
          buf[idx(n, t, u, v)] = val;
        }
  }
  out.L = L;
  out.R.assign((size_t)n1 * n1 * n1, 0.0);
  for (int t = 0; t <= L; ++t)
    for (int u = 0; u <= L - t; ++u)
      for (int v = 0; v <= L - t - u; ++v) out.at(t, u, v) = buf[idx(0, t, u, v)];
}

ShellView view(const int* si, const double* sf, int s) {
  ShellView v;
  v.l = si[6 * s + 0]; v.nprim = si[6 * s + 1]; v.poff = si[6 * s + 2];
  v.aoff = si[6 * s + 3]; v.nfunc = si[6 * s + 4]; v.toff = si[6 * s + 5];
  for (int k = 0; k < 3; ++k) v.c[k] = sf[3 * s + k];
  return v;
}

// Transform a Cartesian block (ncA x ncB) to AO components with T matrices.
void to_ao_2(const ShellView& A, const ShellView& B, const double* tmat, const double* cart,
             double* out /* nfA x nfB */) {
  const int nA = ncart(A.l), nB = ncart(B.l);
  const double* TA = tmat + A.toff;
  const double* TB = tmat + B.toff;
  double tmp[NCART_MAX * NCART_MAX];
  for (int i = 0; i < A.nfunc; ++i)
    for (int q = 0; q < nB; ++q) {
      double s = 0;
      for (int p = 0; p < nA; ++p) s += TA[i * nA + p] * cart[p * nB + q];
      tmp[i * nB + q] = s;
    }
  for (int i = 0; i < A.nfunc; ++i)
    for (int j = 0; j < B.nfunc; ++j) {
      double s = 0;
      for (int q = 0; q < nB; ++q) s += tmp[i * nB + q] * TB[j * nB + q];
      out[i * B.nfunc + j] = s;
    }
}

}  // namespace

extern "C" {

// One-electron integrals S, T (kinetic) and V (nuclear attraction), each (nao, nao).
int gdx_int1e(int nshell, const int* shell_i32, const double* shell_f64,
              const double* prim_f64, const double* tmat_f64, int natm,
              const double* charges, const double* atom_xyz, int nao,
              double* S, double* Tk, double* V) {
  if (nshell <= 0 || nao <= 0) return -1;
#pragma omp parallel for schedule(dynamic)
  for (int sa = 0; sa < nshell; ++sa) {
    ShellView A = view(shell_i32, shell_f64, sa);
    int ea[NCART_MAX][3];
    cart_exps(A.l, ea);
    for (int sb = 0; sb <= sa; ++sb) {
      ShellView B = view(shell_i32, shell_f64, sb);
      int eb[NCART_MAX][3];
      cart_exps(B.l, eb);
      const int nA = ncart(A.l), nB = ncart(B.l);
      double cs[NCART_MAX * NCART_MAX] = {0}, ct[NCART_MAX * NCART_MAX] = {0},
             cv[NCART_MAX * NCART_MAX] = {0};
      const double AB[3] = {A.c[0] - B.c[0], A.c[1] - B.c[1], A.c[2] - B.c[2]};
      for (int ip = 0; ip < A.nprim; ++ip) {
        const double a = prim_f64[2 * (A.poff + ip)], ca = prim_f64[2 * (A.poff + ip) + 1];
        for (int jp = 0; jp < B.nprim; ++jp) {
          const double b = prim_f64[2 * (B.poff + jp)], cb = prim_f64[2 * (B.poff + jp) + 1];
          const double p = a + b;
          Herm1D H[3];
          for (int d = 0; d < 3; ++d) hermite(A.l, B.l + 2, a, b, AB[d], H[d]);
          const double pref = std::pow(PI / p, 1.5);
          double P[3];
          for (int d = 0; d < 3; ++d) P[d] = (a * A.c[d] + b * B.c[d]) / p;
          RTab R;
          std::vector<RTab> Rn(natm);
          for (int c = 0; c < natm; ++c)
            hermite_coulomb(A.l + B.l, p, P[0] - atom_xyz[3 * c], P[1] - atom_xyz[3 * c + 1],
                            P[2] - atom_xyz[3 * c + 2], Rn[c]);
          for (int i = 0; i < nA; ++i)
            for (int j = 0; j < nB; ++j) {
              double s1[3], k1[3];
              for (int d = 0; d < 3; ++d) {
                const int li = ea[i][d], lj = eb[j][d];
                s1[d] = H[d].E[li][lj][0];
                double kk = -2.0 * b * b * H[d].E[li][lj + 2][0] + b * (2 * lj + 1) * H[d].E[li][lj][0];
                if (lj >= 2) kk -= 0.5 * lj * (lj - 1) * H[d].E[li][lj - 2][0];
                k1[d] = kk;
              }
              const double w = ca * cb * pref;
              cs[i * nB + j] += w * s1[0] * s1[1] * s1[2];
              ct[i * nB + j] += w * (k1[0] * s1[1] * s1[2] + s1[0] * k1[1] * s1[2] + s1[0] * s1[1] * k1[2]);
              double vsum = 0;
              for (int c = 0; c < natm; ++c) {
                double acc = 0;
                for (int t = 0; t <= ea[i][0] + eb[j][0]; ++t)
                  for (int u = 0; u <= ea[i][1] + eb[j][1]; ++u)
                    for (int v = 0; v <= ea[i][2] + eb[j][2]; ++v)
                      acc += H[0].E[ea[i][0]][eb[j][0]][t] * H[1].E[ea[i][1]][eb[j][1]][u] *
                             H[2].E[ea[i][2]][eb[j][2]][v] * Rn[c].at(t, u, v);
                vsum -= charges[c] * acc;
              }
              cv[i * nB + j] += ca * cb * 2.0 * PI / p * vsum;
            }
        }
      }
      double os[NCART_MAX * NCART_MAX], ot[NCART_MAX * NCART_MAX], ov[NCART_MAX * NCART_MAX];
      to_ao_2(A, B, tmat_f64, cs, os);
      to_ao_2(A, B, tmat_f64, ct, ot);
      to_ao_2(A, B, tmat_f64, cv, ov);
      for (int i = 0; i < A.nfunc; ++i)
        for (int j = 0; j < B.nfunc; ++j) {
          const int I = A.aoff + i, J = B.aoff + j;
          const int k = i * B.nfunc + j;
          S[(size_t)I * nao + J] = S[(size_t)J * nao + I] = os[k];
          Tk[(size_t)I * nao + J] = Tk[(size_t)J * nao + I] = ot[k];
          V[(size_t)I * nao + J] = V[(size_t)J * nao + I] = ov[k];
        }
    }
  }
  return 0;
}

// Electron repulsion integrals (ij|kl) in 8-fold packed order:
//   ij = i(i+1)/2 + j (i >= j), kl likewise, index = ij(ij+1)/2 + kl (ij >= kl).
// Output length npair(npair+1)/2 with npair = nao(nao+1)/2.
int gdx_int2e_packed(int nshell, const int* shell_i32, const double* shell_f64,
                     const double* prim_f64, const double* tmat_f64, int nao, double* eri) {
  if (nshell <= 0 || nao <= 0) return -1;
  // shell-pair list (sa >= sb)
  std::vector<std::pair<int, int>> pairs;
  for (int a = 0; a < nshell; ++a)
    for (int b = 0; b <= a; ++b) pairs.emplace_back(a, b);
  const long np = (long)pairs.size();
#pragma omp parallel for schedule(dynamic, 1)
  for (long pab = np - 1; pab >= 0; --pab) {
    const int sa = pairs[pab].first, sb = pairs[pab].second;
    ShellView A = view(shell_i32, shell_f64, sa), B = view(shell_i32, shell_f64, sb);
    int ea[NCART_MAX][3], eb[NCART_MAX][3];
    cart_exps(A.l, ea); cart_exps(B.l, eb);
    const int nA = ncart(A.l), nB = ncart(B.l);
    // Hermite coefficients for the bra primitive pairs
    struct PrimPair { double p, P[3], w; Herm1D H[3]; };
    std::vector<PrimPair> bra;
    for (int ip = 0; ip < A.nprim; ++ip)
      for (int jp = 0; jp < B.nprim; ++jp) {
        PrimPair pp;
        const double a = prim_f64[2 * (A.poff + ip)], b = prim_f64[2 * (B.poff + jp)];
        pp.p = a + b;
        for (int d = 0; d < 3; ++d) {
          pp.P[d] = (a * A.c[d] + b * B.c[d]) / pp.p;
          hermite(A.l, B.l, a, b, A.c[d] - B.c[d], pp.H[d]);
        }
        pp.w = prim_f64[2 * (A.poff + ip) + 1] * prim_f64[2 * (B.poff + jp) + 1];
        bra.push_back(pp);
      }
    for (long pcd = 0; pcd <= pab; ++pcd) {
      const int sc = pairs[pcd].first, sd = pairs[pcd].second;
      ShellView C = view(shell_i32, shell_f64, sc), D = view(shell_i32, shell_f64, sd);
      int ec[NCART_MAX][3], ed[NCART_MAX][3];
      cart_exps(C.l, ec); cart_exps(D.l, ed);
      const int nC = ncart(C.l), nD = ncart(D.l);
      const int L = A.l + B.l + C.l + D.l;
      std::vector<double> cart((size_t)nA * nB * nC * nD, 0.0);
      for (int kp = 0; kp < C.nprim; ++kp)
        for (int lp = 0; lp < D.nprim; ++lp) {
          const double c = prim_f64[2 * (C.poff + kp)], d = prim_f64[2 * (D.poff + lp)];
          const double q = c + d;
          double Q[3];
          Herm1D HK[3];
          for (int k = 0; k < 3; ++k) {
            Q[k] = (c * C.c[k] + d * D.c[k]) / q;
            hermite(C.l, D.l, c, d, C.c[k] - D.c[k], HK[k]);
          }
          const double wk = prim_f64[2 * (C.poff + kp) + 1] * prim_f64[2 * (D.poff + lp) + 1];
          for (const PrimPair& pp : bra) {
            const double alpha = pp.p * q / (pp.p + q);
            RTab R;
            hermite_coulomb(L, alpha, pp.P[0] - Q[0], pp.P[1] - Q[1], pp.P[2] - Q[2], R);
            const double pref = pp.w * wk * 2.0 * std::pow(PI, 2.5) / (pp.p * q * std::sqrt(pp.p + q));
            for (int i = 0; i < nA; ++i)
              for (int j = 0; j < nB; ++j) {
                const int tx = ea[i][0] + eb[j][0], ty = ea[i][1] + eb[j][1], tz = ea[i][2] + eb[j][2];
                for (int k = 0; k < nC; ++k)
                  for (int l = 0; l < nD; ++l) {
                    const int sx = ec[k][0] + ed[l][0], sy = ec[k][1] + ed[l][1], sz = ec[k][2] + ed[l][2];
                    double acc = 0;
                    for (int t = 0; t <= tx; ++t) {
                      const double e1 = pp.H[0].E[ea[i][0]][eb[j][0]][t];
                      if (e1 == 0) continue;
                      for (int u = 0; u <= ty; ++u) {
                        const double e2 = e1 * pp.H[1].E[ea[i][1]][eb[j][1]][u];
                        if (e2 == 0) continue;
                        for (int v = 0; v <= tz; ++v) {
                          const double e3 = e2 * pp.H[2].E[ea[i][2]][eb[j][2]][v];
                          if (e3 == 0) continue;
                          double inner = 0;
                          for (int tau = 0; tau <= sx; ++tau) {
                            const double f1 = HK[0].E[ec[k][0]][ed[l][0]][tau];
                            for (int nu = 0; nu <= sy; ++nu) {
                              const double f2 = f1 * HK[1].E[ec[k][1]][ed[l][1]][nu];
                              for (int phi = 0; phi <= sz; ++phi) {
                                const double f3 = f2 * HK[2].E[ec[k][2]][ed[l][2]][phi];
                                const double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                                inner += sg * f3 * R.at(t + tau, u + nu, v + phi);
                              }
                            }
                          }
                          acc += e3 * inner;
                        }
                      }
                    }
                    cart[((i * nB + j) * nC + k) * nD + l] += pref * acc;
                  }
              }
          }
        }
      // transform the four indices to AO components
      const double *TA = tmat_f64 + A.toff, *TB = tmat_f64 + B.toff,
                   *TC = tmat_f64 + C.toff, *TD = tmat_f64 + D.toff;
      std::vector<double> t1((size_t)A.nfunc * nB * nC * nD, 0.0);
      for (int I = 0; I < A.nfunc; ++I)
        for (int i = 0; i < nA; ++i) {
          const double tv = TA[I * nA + i];
          if (tv == 0) continue;
          for (size_t r = 0; r < (size_t)nB * nC * nD; ++r) t1[I * nB * nC * nD + r] += tv * cart[i * nB * nC * nD + r];
        }
      std::vector<double> t2((size_t)A.nfunc * B.nfunc * nC * nD, 0.0);
      for (int I = 0; I < A.nfunc; ++I)
        for (int J = 0; J < B.nfunc; ++J)
          for (int j = 0; j < nB; ++j) {
            const double tv = TB[J * nB + j];
            if (tv == 0) continue;
            for (int r = 0; r < nC * nD; ++r)
              t2[((size_t)I * B.nfunc + J) * nC * nD + r] += tv * t1[((size_t)I * nB + j) * nC * nD + r];
          }
      std::vector<double> t3((size_t)A.nfunc * B.nfunc * C.nfunc * nD, 0.0);
      for (int IJ = 0; IJ < A.nfunc * B.nfunc; ++IJ)
        for (int K = 0; K < C.nfunc; ++K)
          for (int k = 0; k < nC; ++k) {
            const double tv = TC[K * nC + k];
            if (tv == 0) continue;
            for (int l = 0; l < nD; ++l)
              t3[((size_t)IJ * C.nfunc + K) * nD + l] += tv * t2[((size_t)IJ * nC + k) * nD + l];
          }
      for (int I = 0; I < A.nfunc; ++I)
        for (int J = 0; J < B.nfunc; ++J)
          for (int K = 0; K < C.nfunc; ++K)
            for (int Lf = 0; Lf < D.nfunc; ++Lf) {
              double s = 0;
              for (int l = 0; l < nD; ++l)
                s += TD[Lf * nD + l] * t3[(((size_t)I * B.nfunc + J) * C.nfunc + K) * nD + l];
              long i = A.aoff + I, j = B.aoff + J, k = C.aoff + K, l = D.aoff + Lf;
              if (i < j) std::swap(i, j);
              if (k < l) std::swap(k, l);
              long ij = i * (i + 1) / 2 + j, kl = k * (k + 1) / 2 + l;
              if (ij < kl) std::swap(ij, kl);
              eri[ij * (ij + 1) / 2 + kl] = s;
            }
    }
  }
  return 0;
}

}  // extern "C"
