// Per-grid-point B3LYP energy densities, written once as templates over a scalar type so
// the same expression tree yields values (T = double), exact first derivatives
// (T = Dual<K>, forward mode) and directional second derivatives (T = Dual<K, Dual<1>>).
// Every expression follows the reference term by term, including its clips and
// where-selects, so the derivatives match what jax.grad of the reference produces:
//   lsda_x_e        grad_dft/popular_functionals.py:29-50
//   b88_x_e         grad_dft/popular_functionals.py:52-103
//   pw92_c_e        grad_dft/popular_functionals.py:105-139
//   vwn_c_e         grad_dft/popular_functionals.py:141-195
//   lyp_c_e         grad_dft/popular_functionals.py:197-269
//   exchange_polarization_correction     grad_dft/functional.py:950-979
//   correlation_polarization_correction  grad_dft/functional.py:982-1045
// Inputs per point: rho_up, rho_dn, |grad rho_up|^2, |grad rho_dn|^2, lapl_up, lapl_dn.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace gdx {

#define GDX_HD __host__ __device__ __forceinline__

// Dual<K, S>: value and K partial derivatives over the scalar type S.  S = double gives
// first derivatives; S = Dual<1> (a value with a tangent along a fixed direction) gives,
// in d[i].d[0], the directional derivative of the i-th partial (a Hessian-vector row).
template <int K, typename S = double>
struct Dual {
  S v;
  S d[K];
  GDX_HD Dual() {}
  GDX_HD explicit Dual(double c) : v(c) {
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = S(0.0);
  }
};

template <int K, typename S = double>
GDX_HD Dual<K, S> dconst(double c) { return Dual<K, S>(c); }
// independent variable idx with value x
template <int K, typename S = double>
GDX_HD Dual<K, S> dvar(const S& x, int idx) {
  Dual<K, S> r(0.0);
  r.v = x;
  r.d[idx] = S(1.0);
  return r;
}
// chain rule helper: r = f(a), r' = fp * a'
template <int K, typename S, typename FP>
GDX_HD Dual<K, S> dchain(const Dual<K, S>& a, const S& f, const FP& fp) {
  Dual<K, S> r; r.v = f;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = fp * a.d[i];
  return r;
}
template <int K, typename S> GDX_HD Dual<K, S> operator+(const Dual<K, S>& a, const Dual<K, S>& b) {
  Dual<K, S> r; r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int K, typename S> GDX_HD Dual<K, S> operator-(const Dual<K, S>& a, const Dual<K, S>& b) {
  Dual<K, S> r; r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int K, typename S> GDX_HD Dual<K, S> operator-(const Dual<K, S>& a) {
  Dual<K, S> r; r.v = -a.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int K, typename S> GDX_HD Dual<K, S> operator*(const Dual<K, S>& a, const Dual<K, S>& b) {
  Dual<K, S> r; r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int K, typename S> GDX_HD Dual<K, S> operator/(const Dual<K, S>& a, const Dual<K, S>& b) {
  Dual<K, S> r; r.v = a.v / b.v;
  const S ib = 1.0 / b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int K, typename S> GDX_HD Dual<K, S> operator+(const Dual<K, S>& a, double c) {
  Dual<K, S> r = a; r.v = r.v + c; return r;
}
template <int K, typename S> GDX_HD Dual<K, S> operator+(double c, const Dual<K, S>& a) { return a + c; }
template <int K, typename S> GDX_HD Dual<K, S> operator-(const Dual<K, S>& a, double c) {
  Dual<K, S> r = a; r.v = r.v - c; return r;
}
template <int K, typename S> GDX_HD Dual<K, S> operator-(double c, const Dual<K, S>& a) { return (-a) + c; }
template <int K, typename S> GDX_HD Dual<K, S> operator*(const Dual<K, S>& a, double c) {
  return dchain(a, a.v * c, c);
}
template <int K, typename S> GDX_HD Dual<K, S> operator*(double c, const Dual<K, S>& a) { return a * c; }
template <int K, typename S> GDX_HD Dual<K, S> operator/(const Dual<K, S>& a, double c) { return a * (1.0 / c); }
template <int K, typename S> GDX_HD Dual<K, S> operator/(double c, const Dual<K, S>& a) {
  const S r = c / a.v; return dchain(a, r, -r / a.v);
}

// ---- elementary functions (value overloads for double, chain rule for Dual) ----
GDX_HD double val(double x) { return x; }
template <int K, typename S> GDX_HD double val(const Dual<K, S>& x) { return val(x.v); }

GDX_HD double flog(double x) { return log(x); }
GDX_HD double flog2(double x) { return log2(x); }
GDX_HD double fexp(double x) { return exp(x); }
GDX_HD double fexp2(double x) { return exp2(x); }
GDX_HD double fsqrt(double x) { return sqrt(x); }
GDX_HD double fasinh(double x) { return asinh(x); }
GDX_HD double fatan(double x) { return atan(x); }
GDX_HD double fpow(double x, double p) { return pow(x, p); }

template <int K, typename S> GDX_HD Dual<K, S> flog(const Dual<K, S>& a) {
  return dchain(a, flog(a.v), 1.0 / a.v);
}
template <int K, typename S> GDX_HD Dual<K, S> flog2(const Dual<K, S>& a) {
  return dchain(a, flog2(a.v), 1.4426950408889634 / a.v);
}
template <int K, typename S> GDX_HD Dual<K, S> fexp(const Dual<K, S>& a) {
  const S e = fexp(a.v); return dchain(a, e, e);
}
template <int K, typename S> GDX_HD Dual<K, S> fexp2(const Dual<K, S>& a) {
  const S e = fexp2(a.v); return dchain(a, e, e * 0.6931471805599453);
}
template <int K, typename S> GDX_HD Dual<K, S> fsqrt(const Dual<K, S>& a) {
  const S s = fsqrt(a.v); return dchain(a, s, 0.5 / s);
}
template <int K, typename S> GDX_HD Dual<K, S> fasinh(const Dual<K, S>& a) {
  return dchain(a, fasinh(a.v), 1.0 / fsqrt(a.v * a.v + 1.0));
}
template <int K, typename S> GDX_HD Dual<K, S> fatan(const Dual<K, S>& a) {
  return dchain(a, fatan(a.v), 1.0 / (1.0 + a.v * a.v));
}
template <int K, typename S> GDX_HD Dual<K, S> fpow(const Dual<K, S>& a, double p) {
  const S r = fpow(a.v, p);
  return dchain(a, r, p * fpow(a.v, p - 1.0));
}

// jnp.clip(x, a_min=c): max(x, c) with gradient 1 where x > c
GDX_HD double clipmin(double x, double c) { return x > c ? x : c; }
template <int K, typename S> GDX_HD Dual<K, S> clipmin(const Dual<K, S>& x, double c) {
  return val(x) > c ? x : Dual<K, S>(c);
}
// jnp.where(cond, a, b): selects whole value (and tangent)
template <typename T> GDX_HD T sel(bool c, const T& a, const T& b) { return c ? a : b; }

template <typename T> GDX_HD T cst(double c) { return T(c); }

constexpr double CLIP = 1e-30;
constexpr double PI = 3.14159265358979323846;

// functional.py:950-979
template <typename T>
GDX_HD T exchange_pol(const T& eP, const T& eF, const T& ru, const T& rd) {
  const T z = (ru - rd) / (ru + rd);
  const double den = 2.0 * (cbrt(2.0) - 1.0);
  const T fz = (fpow(1.0 - z, 4.0 / 3.0) + fpow(1.0 + z, 4.0 / 3.0) - 2.0) / den;
  return eP + (eF - eP) * fz;
}

// functional.py:982-1045; e_tilde_PF given as (eP, eF)
template <typename T>
GDX_HD T correlation_pol(const T& eP, const T& eF, const T& ru, const T& rd) {
  const T s = ru + rd;
  const T log_rho = flog2(clipmin(s, CLIP));
  const double l0 = log2(cbrt(3.0 / (4.0 * PI)));
  const T log_rs = l0 - log_rho / 3.0;
  const T zeta = sel(val(s) > CLIP, (ru - rd) / s, cst<T>(0.0));
  const double den = 2.0 * (cbrt(2.0) - 1.0);
  const T zm = fexp2(4.0 * flog2(1.0 - zeta) / 3.0);
  const T zp = fexp2(4.0 * flog2(1.0 + zeta) / 3.0);
  const T fz = (zm + zp - 2.0) / den;
  const double A_ = 0.016887, alpha1 = 0.11125;
  const double beta1 = 10.357, beta2 = 3.6231, beta3 = 0.88026, beta4 = 0.49671;
  const T ars = fexp2(log2(alpha1) + log_rs);
  const T brs_1_2 = fexp2(log2(beta1) + log_rs / 2.0);
  const T brs = fexp2(log2(beta2) + log_rs);
  const T brs_3_2 = fexp2(log2(beta3) + 3.0 * log_rs / 2.0);
  const T brs2 = fexp2(log2(beta4) + 2.0 * log_rs);
  const T alphac = 2.0 * A_ * (1.0 + ars) * flog(1.0 + (1.0 / (2.0 * A_)) / (brs_1_2 + brs + brs_3_2 + brs2));
  const double fpp0 = 4.0 / (9.0 * (cbrt(2.0) - 1.0));
  const T z2 = zeta * zeta;
  const T z4 = z2 * z2;
  return eP + alphac * (fz / fpp0) * (1.0 - z4) + (eF - eP) * fz * z4;
}

// popular_functionals.py:29-50
template <typename T>
GDX_HD T lsda_x(const T& ru0, const T& rd0) {
  const T ru = clipmin(ru0, CLIP), rd = clipmin(rd0, CLIP);
  const T s43 = fpow(ru + rd, 4.0 / 3.0);
  const T eP = (-0.75 * cbrt(3.0 / PI)) * s43;
  const T eF = (-0.75 * cbrt(6.0 / PI)) * s43;
  return exchange_pol(eP, eF, ru, rd);
}

// popular_functionals.py:52-103 (gradient correction only)
template <typename T>
GDX_HD T b88_spin(const T& r0, const T& g) {
  const double beta = 0.0042;
  const T r = clipmin(r0, CLIP);
  const T log_rho = flog2(clipmin(r, CLIP));
  const T log_g = flog2(clipmin(g, CLIP)) / 2.0;
  const T log_x = log_g - (4.0 / 3.0) * log_rho;
  const T x = fexp2(log_x);
  return beta * fexp2(4.0 * log_rho / 3.0 + 2.0 * log_x - flog2(1.0 + 6.0 * beta * x * fasinh(x)));
}
template <typename T>
GDX_HD T b88_x(const T& ru, const T& rd, const T& gu, const T& gd) {
  return -(b88_spin(ru, gu) + b88_spin(rd, gd));
}

// popular_functionals.py:141-195 (VWN5 parameters, paramagnetic / ferromagnetic)
template <typename T>
GDX_HD T vwn_eps(const T& x, const T& log_x, double A, double b, double c, double x0) {
  const T X = fexp2(2.0 * log_x) + fexp2(log_x + log2(b)) + c;
  const double X0 = x0 * x0 + b * x0 + c;
  const double Q = sqrt(4.0 * c - b * b);
  const T at = fatan(Q / (2.0 * x + b));
  const T xm = x - x0;
  return (A / 2.0) * (2.0 * flog(x) - flog(X) + (2.0 * b / Q) * at
                      - (b * x0 / X0) * (flog(xm * xm / X) + (2.0 * (2.0 * x0 + b) / Q) * at));
}
template <typename T>
GDX_HD T vwn_c(const T& ru0, const T& rd0) {
  const T ru = sel(val(ru0) > CLIP, ru0, cst<T>(0.0));
  const T rd = sel(val(rd0) > CLIP, rd0, cst<T>(0.0));
  const T log_rho = flog2(clipmin(ru + rd, CLIP));
  const double l0 = log2(cbrt(3.0 / (4.0 * PI)));
  const T log_rs = l0 - log_rho / 3.0;
  const T log_x = log_rs / 2.0;
  const T x = fexp2(log_x);
  const T eP = vwn_eps(x, log_x, 0.0621814, 3.72744, 12.9352, -0.10498);
  const T eF = vwn_eps(x, log_x, 0.0621814 / 2.0, 7.06042, 18.0578, -0.325);
  return correlation_pol(eP, eF, ru, rd) * (ru + rd);
}

// popular_functionals.py:105-139: Perdew-Wang 1992 fit G(rs) for the paramagnetic and the
// ferromagnetic gas, combined by the polarisation correction, times the total density
template <typename T>
GDX_HD T pw92_eps(const T& log_rs, double A, double alpha1, double b1, double b2, double b3, double b4) {
  const T brs_1_2 = fexp2(log_rs / 2.0 + log2(b1));
  const T ars = fexp2(log_rs + log2(alpha1));
  const T brs = fexp2(log_rs + log2(b2));
  const T brs_3_2 = fexp2(3.0 * log_rs / 2.0 + log2(b3));
  const T brs2 = fexp2(2.0 * log_rs + log2(b4));
  return -2.0 * A * (1.0 + ars) * flog(1.0 + (1.0 / (2.0 * A)) / (brs_1_2 + brs + brs_3_2 + brs2));
}
template <typename T>
GDX_HD T pw92_c(const T& ru, const T& rd) {
  const T s = ru + rd;
  const T log_rho = flog2(clipmin(s, CLIP));
  const double l0 = log2(cbrt(3.0 / (4.0 * PI)));
  const T log_rs = l0 - log_rho / 3.0;
  const T eP = pw92_eps(log_rs, 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294);
  const T eF = pw92_eps(log_rs, 0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517);
  return correlation_pol(eP, eF, ru, rd) * s;
}

// popular_functionals.py:197-269 (original LYP)
template <typename T>
GDX_HD T lyp_c(const T& ru0, const T& rd0, const T& gu, const T& gd, const T& lu, const T& ld) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double CF = 0.3 * pow(3.0 * PI * PI, 2.0 / 3.0);
  const T ru = clipmin(ru0, CLIP), rd = clipmin(rd0, CLIP);
  const T tu = (sel(val(ru) > CLIP, gu / ru, cst<T>(0.0)) - lu) / 8.0;
  const T td = (sel(val(rd) > CLIP, gd / rd, cst<T>(0.0)) - ld) / 8.0;
  const T s = ru + rd;
  const bool big = val(s) > CLIP;
  const T frac = sel(big, (ru * ru + rd * rd) / (s * s), cst<T>(1.0));
  const T gamma = 2.0 * (1.0 - frac);
  const T rhos_ts = s * (tu + td);
  const T rho_t = ru * tu + rd * td;
  const T rho_l = ru * lu + rd * ld;
  const T rm13 = fpow(s, -1.0 / 3.0);
  const T r83 = fpow(ru, 8.0 / 3.0) + fpow(rd, 8.0 / 3.0);
  const T rm53 = fpow(s, -5.0 / 3.0);
  const T ef = sel(val(s) > 0.0, fexp(-c * rm13), cst<T>(0.0));
  const T par = (pow(2.0, 2.0 / 3.0) * CF) * r83 - rhos_ts + rho_t / 9.0 + rho_l / 18.0;
  const T brk = sel(big, 2.0 * b * rm53 * par * ef, cst<T>(0.0));
  return -a * sel(big, gamma / (1.0 + d * rm13) * (s + brk), cst<T>(0.0));
}

// abs_clip, grad_dft/molecule.py:687-689
template <typename T>
GDX_HD T abs_clip(const T& x) { return fabs(val(x)) > CLIP ? x : cst<T>(0.0); }

// b3lyp_exhf_densities (popular_functionals.py:306-326) followed by the clip of
// compute_densities (functional.py:184): e[0..3] = lda, b88, vwn, lyp.  corr = 1 puts PW92
// correlation (pw92_densities, :297-301) in the correlation slot e[2] in place of VWN.
template <typename T>
GDX_HD void b3lyp_parts(const T& ru, const T& rd, const T& gu, const T& gd, const T& lu,
                        const T& ld, int corr, T* e) {
  e[0] = abs_clip(lsda_x(ru, rd));
  e[1] = abs_clip(b88_x(ru, rd, gu, gd));
  e[2] = abs_clip(corr ? pw92_c(ru, rd) : vwn_c(ru, rd));
  e[3] = abs_clip(lyp_c(ru, rd, gu, gd, lu, ld));
}

}  // namespace gdx
