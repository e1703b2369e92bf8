// Symmetric eigensolver for n in (64, 224]: Householder tridiagonalisation, multisection
// bisection for the eigenvalues, inverse iteration for the eigenvectors of the tridiagonal
// matrix (re-orthogonalised inside clusters of close eigenvalues), back-transformation.
//
//   A = Q T Q^T, Q = H_0 H_1 ... H_{n-3}, H_k = I - tau_k v_k v_k^T   (tri_reduce_kernel)
//   eigenvalues of T ascending                                        (tri_bisect_kernel)
//   T x_j = w_j x_j                                                   (tri_invit_kernel,
//                                                                      tri_orth_kernel)
//   X <- 1.5 X - 0.5 X (X^T X)   (first-order Loewdin polish, two GEMMs in the caller)
//   eigenvectors of A: Z = Q X                                        (tri_back_kernel)
//
// Replaces the one-workgroup Jacobi sweeps where the matrix no longer fits LDS: the reduction
// keeps the lower triangle of A in the registers of one 1024-thread workgroup (n - 2 steps of a
// few barriers each), and the other phases spread over many workgroups.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace gdx {
namespace tri {
namespace {

constexpr int NMAX = 224;        // largest n (7 row blocks of 32)
constexpr int RED_THREADS = 1024;

// Workgroup barrier that orders LDS only: the global stores issued before it (reflectors,
// d, e, tau -- read by later kernels only) need not complete, where __syncthreads() would wait
// for them.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Cross-lane sums without the LDS crossbar: DPP moves inside each row of 16 lanes
// (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror: every lane ends with its row's
// sum), then the four row sums read into scalar registers.  Every lane (and every wave summing
// the same data) gets bitwise the same value.  Call with all 64 lanes active.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double row16_sum(double v) {
  v += dpp_mov<0xB1>(v);    // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);    // quad_perm [2,3,0,1]
  v += dpp_mov<0x141>(v);   // row_half_mirror
  v += dpp_mov<0x140>(v);   // row_mirror
  return v;
}
__device__ __forceinline__ double lane_value(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ __forceinline__ double wave_sum(double v) {
  v = row16_sum(v);
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}
// sum over the 32 lanes of each half-wave
__device__ __forceinline__ double half_sum(double v) {
  v = row16_sum(v);
  const double h0 = lane_value(v, 0) + lane_value(v, 16), h1 = lane_value(v, 32) + lane_value(v, 48);
  return (threadIdx.x & 32) ? h1 : h0;
}
// the value of lane l ^ 32
__device__ __forceinline__ double swap_halves(double v) {
  const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
  const bool upper = threadIdx.x & 32;
  return __hiloint2double(upper ? hi[0] : hi[1], upper ? lo[0] : lo[1]);
}

// ---------------------------------------------------------------- reduction to T ----
// Thread (rg, cg) = (tid / 32, tid % 32) owns the lower-triangle elements (rg + 32 a,
// cg + 32 c), c <= a < RA (on the diagonal blocks a == c only rg >= cg): row blocks a < RR in
// registers, the rest in LDS (one slot per thread and element).
// Step k (trailing set S = {k+1..n-1}):
//   x = A[S, k]; v = x / (x_1 - beta) with v_1 = 1, beta = -sign(x_1) ||x||,
//   tau = (beta - x_1) / beta;   p = tau A[S,S] v;   w = p - (tau/2)(p.v) v;
//   A[S,S] -= v w^T + w v^T.
// p is formed from row sums (reduced across the 32 lanes of a half-wave) plus column sums of
// the strictly lower part (reduced across the 16 waves through LDS).  Each wave recomputes
// ||x|| and p.v itself from LDS (identical operations, identical results), which saves a
// barrier per reduction.  The input is symmetrised on load: (A + A^T) / 2.
// Outputs (per batch entry): d[n], e[n] (e[n-1] = 0), tau[n] (0 for k >= n-2), reflector k
// in row k of Vr (entries 0..k zero, k+1 one).
template <int RA, int RR>
__global__ __launch_bounds__(RED_THREADS) void tri_reduce_kernel(int n, const double* __restrict__ Ain, long sA,
                                                                double* __restrict__ d_all,
                                                                double* __restrict__ e_all,
                                                                double* __restrict__ tau_all,
                                                                double* __restrict__ Vr_all) {
  constexpr int NR = RR * (RR + 1) / 2;        // slots in registers (row blocks < RR)
  constexpr int NX = RA * (RA + 1) / 2 - NR;   // slots in LDS (row blocks >= RR)
  constexpr int NW = RED_THREADS / 64;
  // one LDS block: xs (2 x NMAX, double-buffered by step parity), ps, rowp, colp (NW x NMAX),
  // then the LDS slots [slot][thread] (conflict-free)
  constexpr int XS = 0, PS = 2 * NMAX, ROWP = 3 * NMAX, COLP = 4 * NMAX, LX = (4 + NW) * NMAX;
  __shared__ double sm[LX + (NX > 0 ? NX : 1) * RED_THREADS];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int rg = tid >> 5, cg = tid & 31, lane = tid & 63, wv = tid >> 6;
  const bool dg = rg == cg, up = rg < cg;   // on a diagonal block: diagonal / not owned
  const double* A = Ain + (size_t)b * sA;
  double* d = d_all + (size_t)b * n;
  double* e = e_all + (size_t)b * n;
  double* tau_o = tau_all + (size_t)b * n;
  double* Vr = Vr_all + (size_t)b * n * n;

  double L[NR > 0 ? NR : 1];
  int orow = rg, ocol = cg, othr = tid;   // per-thread LDS bases; constant offsets fold into
                                          // the LDS instructions (see the start of each step)
  auto get = [&](int a, int c) -> double {
    const int q = a * (a + 1) / 2 + c;
    return q < NR ? L[q < NR ? q : 0] : sm[LX + (q >= NR ? q - NR : 0) * RED_THREADS + othr];
  };
  auto put = [&](int a, int c, double v) {
    const int q = a * (a + 1) / 2 + c;
    if (q < NR) L[q < NR ? q : 0] = v;
    else sm[LX + (q >= NR ? q - NR : 0) * RED_THREADS + othr] = v;
  };
#pragma unroll
  for (int a = 0; a < RA; ++a)
#pragma unroll
    for (int c = 0; c <= a; ++c) {
      const int i = rg + 32 * a, j = cg + 32 * c;
      double v = 0.0;
      if (i < n && (c < a || !up)) v = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
      put(a, c, v);
    }

  for (int k = 0; k < n - 2; ++k) {
    // keep the bases opaque to loop-invariant code motion: otherwise every (base + constant)
    // LDS address is hoisted out of the loop into a register of its own
    asm volatile("" : "+v"(orow), "+v"(ocol), "+v"(othr));
    const int par = k & 1, kc = k >> 5, kl = k & 31;
    const int xsb = XS + par * NMAX;
    // column k below the diagonal to LDS; d_k
    if (cg == kl) {
#pragma unroll
      for (int a = 0; a < RA; ++a)
#pragma unroll
        for (int c = 0; c <= a; ++c)
          if (c == kc) {
            const int i = rg + 32 * a;
            const double v = get(a, c);
            if (i > k && i < n) sm[xsb + orow + 32 * a] = v;
            if (i == k) d[k] = v;
          }
    }
    lds_barrier();
    // Householder vector (every wave)
    const double alpha = sm[xsb + k + 1];
    double sg = 0.0;
    for (int i = k + 2 + lane; i < n; i += 64) {
      const double x = sm[xsb + i];
      sg += x * x;
    }
    sg = wave_sum(sg);
    double tau = 0.0, beta = alpha, scal = 0.0;
    if (sg != 0.0) {
      const double mu = sqrt(alpha * alpha + sg);
      beta = alpha <= 0.0 ? mu : -mu;
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    if (tid == 0) {
      e[k] = beta;
      tau_o[k] = tau;
    }
    for (int i = tid; i < n; i += RED_THREADS)
      Vr[(size_t)k * n + i] = i <= k ? 0.0 : (i == k + 1 ? 1.0 : sm[xsb + i] * scal);
    // v at row index base + 32 blk (base = orow or ocol, x the true index)
    auto vat = [&](int base, int blk, int x) -> double {
      return x == k + 1 ? 1.0 : ((x > k + 1 && x < n) ? sm[xsb + base + 32 * blk] * scal : 0.0);
    };
    // p = A v: row sums and strictly-lower column sums.  Blocks of column block c < kc (every
    // row and column index <= k) no longer take part: v and p vanish there, the update leaves
    // them unchanged, and p is read only at indices > k -- the step skips them (uniform branches).
    double vcol[RA], wcol[RA];
#pragma unroll
    for (int c = 0; c < RA; ++c) {
      vcol[c] = vat(ocol, c, cg + 32 * c);
      wcol[c] = 0.0;   // column partial sums here, w later
    }
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      if (a < kc) continue;
      const int i = rg + 32 * a;
      const double vr = vat(orow, a, i);
      double s = 0.0;
#pragma unroll
      for (int c = 0; c <= a; ++c) {
        if (c < kc) continue;
        const double l = get(a, c);
        s += l * vcol[c];
        if (c < a || !dg) wcol[c] += l * vr;   // unowned (up) entries hold 0
      }
      s = half_sum(s);
      if (cg == 0 && i < n) sm[ROWP + orow + 32 * a] = s;
    }
#pragma unroll
    for (int c = 0; c < RA; ++c) {
      if (c < kc) continue;
      const double t = wcol[c] + swap_halves(wcol[c]);
      const int j = cg + 32 * c;
      if (lane < 32 && j < n) sm[COLP + wv * NMAX + ocol + 32 * c] = t;
    }
    lds_barrier();
    if (tid < n) {
      double p = 0.0;
      if (tid > k) {
        p = sm[ROWP + tid];
#pragma unroll
        for (int w2 = 0; w2 < NW; ++w2) p += sm[COLP + w2 * NMAX + tid];
        p *= tau;
      }
      sm[PS + tid] = p;
    }
    lds_barrier();
    double dot = 0.0;
    for (int i = k + 1 + lane; i < n; i += 64)
      dot += sm[PS + i] * (i == k + 1 ? 1.0 : sm[xsb + i] * scal);
    const double K = 0.5 * tau * wave_sum(dot);
#pragma unroll
    for (int c = 0; c < RA; ++c) {
      if (c < kc) continue;
      const int j = cg + 32 * c;
      wcol[c] = (j < n ? sm[PS + ocol + 32 * c] : 0.0) - K * vcol[c];
    }
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      if (a < kc) continue;
      const int i = rg + 32 * a;
      const double vr = vat(orow, a, i);
      const double wr = (i < n ? sm[PS + orow + 32 * a] : 0.0) - K * vr;
#pragma unroll
      for (int c = 0; c <= a; ++c)
        if (c >= kc && (c < a || !up)) put(a, c, get(a, c) - (vr * wcol[c] + wr * vcol[c]));
    }
  }
  // the last 2 x 2 block (or the single element)
#pragma unroll
  for (int a = 0; a < RA; ++a)
#pragma unroll
    for (int c = 0; c <= a; ++c) {
      const int i = rg + 32 * a, j = cg + 32 * c;
      const double v = get(a, c);
      if (c == a && dg && i < n && i >= n - 2) d[i] = v;
      if (n >= 2 && i == n - 1 && j == n - 2) e[n - 2] = v;
    }
  if (tid == 0) {
    e[n - 1] = 0.0;
    tau_o[n - 1] = 0.0;
    if (n >= 2) tau_o[n - 2] = 0.0;
  }
}

// ---------------------------------------------------------------- eigenvalues -------
// Sturm count: the number of eigenvalues of T below x (negative pivots of T - x I = L D L^T),
// pivots smaller than pivmin in magnitude replaced by -pivmin.
__device__ __forceinline__ int sturm_count(int n, const double* __restrict__ d, const double* __restrict__ e2,
                                           double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int cnt = q < 0.0;
#pragma unroll 8
  for (int i = 1; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// Gershgorin interval, ||T|| bound and pivmin of T, reduced over one wave (all lanes get them).
__device__ __forceinline__ void tri_bounds(int n, const double* d, const double* e, const double* e2, double& gl,
                                           double& gu, double& tnorm, double& pivmin) {
  const int lane = threadIdx.x & 63;
  double lo = DBL_MAX, hi = -DBL_MAX, em = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    if (i < n - 1) em = fmax(em, e2[i]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
    em = fmax(em, __shfl_xor(em, o, 64));
  }
  tnorm = fmax(fabs(lo), fabs(hi));
  pivmin = DBL_MIN * fmax(1.0, em);
  const double fudge = 2.0 * DBL_EPSILON * tnorm * n + 4.0 * pivmin;
  gl = lo - fudge;
  gu = hi + fudge;
}

// One wave per eigenvalue j (ascending order): each round the 64 lanes count at 64 interior
// points of [lo, hi] and the bracket shrinks 65-fold to [x_{g-1}, x_g], g the first lane whose
// count exceeds j; stops at width max(eps ||T||, 2 eps |x|).  Writes w[j], and for j == 0 the
// norm bound tn[b] (used for the cluster test) and sweeps[b] = 0.
__global__ __launch_bounds__(256) void tri_bisect_kernel(int n, const double* __restrict__ d_all,
                                                         const double* __restrict__ e_all, double* __restrict__ w_all,
                                                         double* __restrict__ tn_all, int* __restrict__ sweeps) {
  __shared__ double sd[NMAX], se[NMAX], se2[NMAX];
  const int b = blockIdx.y, tid = threadIdx.x, lane = tid & 63;
  const double* d = d_all + (size_t)b * n;
  const double* e = e_all + (size_t)b * n;
  for (int i = tid; i < n; i += 256) {
    sd[i] = d[i];
    se[i] = e[i];
    se2[i] = e[i] * e[i];
  }
  __syncthreads();
  const int j = blockIdx.x * 4 + (tid >> 6);
  if (j >= n) return;
  double lo, hi, tnorm, pivmin;
  tri_bounds(n, sd, se, se2, lo, hi, tnorm, pivmin);
  for (int round = 0; round < 24; ++round) {
    const double tol = fmax(DBL_EPSILON * tnorm, 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)));
    if (!(hi - lo > tol)) break;
    const double h = (hi - lo) / 65.0;
    const double x = lo + h * (lane + 1);
    const int cnt = sturm_count(n, sd, se2, x, pivmin);
    const unsigned long long m = __ballot(cnt > j);
    const int g = m ? __ffsll((long long)m) - 1 : 64;
    const double nhi = g < 64 ? lo + h * (g + 1) : hi;
    const double nlo = g > 0 ? lo + h * g : lo;
    lo = nlo;
    hi = nhi;
  }
  if (lane == 0) {
    w_all[(size_t)b * n + j] = 0.5 * (lo + hi);
    if (j == 0) {
      tn_all[b] = tnorm;
      if (sweeps) sweeps[b] = 0;
    }
  }
}

// ---------------------------------------------------------------- eigenvectors ------
__device__ __forceinline__ double start_value(int i, int j) {
  uint32_t h = (uint32_t)i * 0x9E3779B1u ^ ((uint32_t)j * 0x85EBCA77u + 0x165667B1u);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return (double)(h >> 8) * (2.0 / 16777216.0) - 1.0;
}

// Gaussian elimination with partial pivoting of T - lam I (U with two superdiagonals,
// multipliers, row swaps; pivots below ptol in magnitude replaced by +-ptol), and the solve
// (T - lam I) y = y with those factors.  Element i of the arrays sits at i * stride.
struct TriLU {
  double *u0, *u1, *u2, *lm;
  unsigned char* pv;
  int stride;
  __device__ double& U0(int i) const { return u0[i * stride]; }
  __device__ double& U1(int i) const { return u1[i * stride]; }
  __device__ double& U2(int i) const { return u2[i * stride]; }
  __device__ double& LM(int i) const { return lm[i * stride]; }
  __device__ unsigned char& PV(int i) const { return pv[i * stride]; }
};

__device__ __forceinline__ void tri_factor(int n, const double* d, const double* e, double lam, double ptol,
                                           const TriLU& f) {
  auto clamp_piv = [&](double u) { return fabs(u) < ptol ? (u < 0.0 ? -ptol : ptol) : u; };
  double alpha = d[0] - lam, beta = n > 1 ? e[0] : 0.0;
  for (int i = 0; i < n - 1; ++i) {
    const double bi = e[i], a1 = d[i + 1] - lam, c1 = i + 1 < n - 1 ? e[i + 1] : 0.0;
    if (fabs(alpha) >= fabs(bi)) {
      const double u0 = clamp_piv(alpha);
      const double l = bi / u0;
      f.U0(i) = u0;
      f.U1(i) = beta;
      f.U2(i) = 0.0;
      f.LM(i) = l;
      f.PV(i) = 0;
      alpha = a1 - l * beta;
      beta = c1;
    } else {
      const double l = alpha / bi;
      f.U0(i) = bi;
      f.U1(i) = a1;
      f.U2(i) = c1;
      f.LM(i) = l;
      f.PV(i) = 1;
      alpha = beta - l * a1;
      beta = -l * c1;
    }
  }
  f.U0(n - 1) = clamp_piv(alpha);
}

// y <- (T - lam I)^{-1} (scale y) in place (y[i * ys]); returns ||y||^2 of the result.
__device__ __forceinline__ double tri_solve(int n, const TriLU& f, double* y, int ys, double scale) {
  double cur = y[0] * scale;
#pragma unroll 4
  for (int i = 0; i < n - 1; ++i) {
    double nxt = y[(i + 1) * ys] * scale;
    if (f.PV(i)) {
      const double t = cur;
      cur = nxt;
      nxt = t;
    }
    y[i * ys] = cur;
    cur = nxt - f.LM(i) * cur;
  }
  y[(n - 1) * ys] = cur;
  double y1 = 0.0, y2 = 0.0, ss = 0.0;
#pragma unroll 4
  for (int i = n - 1; i >= 0; --i) {
    const double v = (y[i * ys] - f.U1(i) * y1 - f.U2(i) * y2) / f.U0(i);
    y[i * ys] = v;
    ss += v * v;
    y2 = y1;
    y1 = v;
  }
  return ss;
}

// One thread per eigenvalue, INV_TPB per workgroup with the factors and the iterate in LDS
// ([i][thread]: conflict-free): factorise T - w_j I, then three solves from a pseudo-random
// start, normalising after each.  X[i][j] = x_j(i).
constexpr int INV_TPB = 16;
__global__ __launch_bounds__(INV_TPB) void tri_invit_kernel(int n, const double* __restrict__ d_all,
                                                            const double* __restrict__ e_all,
                                                            const double* __restrict__ w_all,
                                                            const double* __restrict__ tn_all,
                                                            double* __restrict__ X_all) {
  __shared__ double sd[NMAX], se[NMAX];
  __shared__ double fu0[NMAX][INV_TPB], fu1[NMAX][INV_TPB], fu2[NMAX][INV_TPB], flm[NMAX][INV_TPB],
      fy[NMAX][INV_TPB];
  __shared__ unsigned char fpv[NMAX][INV_TPB];
  const int b = blockIdx.y, t = threadIdx.x;
  for (int i = t; i < n; i += INV_TPB) {
    sd[i] = d_all[(size_t)b * n + i];
    se[i] = e_all[(size_t)b * n + i];
  }
  __syncthreads();
  const int j = blockIdx.x * INV_TPB + t;
  if (j >= n) return;
  const double lam = w_all[(size_t)b * n + j];
  const double ptol = DBL_EPSILON * fmax(tn_all[b], DBL_MIN);
  const TriLU f{&fu0[0][t], &fu1[0][t], &fu2[0][t], &flm[0][t], &fpv[0][t], INV_TPB};
  tri_factor(n, sd, se, lam, ptol, f);
  double* y = &fy[0][t];
  for (int i = 0; i < n; ++i) y[i * INV_TPB] = start_value(i, j);
  double scale = 1.0;
  for (int it = 0; it < 3; ++it) {
    const double ss = tri_solve(n, f, y, INV_TPB, scale);
    scale = ss > 0.0 ? 1.0 / sqrt(ss) : 1.0;
  }
  double* X = X_all + (size_t)b * n * n;
  for (int i = 0; i < n; ++i) X[(size_t)i * n + j] = y[i * INV_TPB] * scale;
}

// Clusters: runs of eigenvalues with consecutive gaps <= 1e-9 ||T|| (numerically double
// eigenvalues).  Independent inverse iteration can return nearly parallel vectors for members
// of such a cluster (both iterates may settle on the same direction), and orthogonalising them
// afterwards would leave mostly rounding noise.  Wider gaps need nothing here: the vectors are
// then orthogonal to ~eps ||T|| / gap <= 1e-7, which the first-order Loewdin polish that
// follows reduces to its square, without raising the residuals (eps ||T||).  So, one wave per cluster (its first index):
// each later member m is orthogonalised against the members before it (modified Gram-Schmidt,
// twice) and solved again with its own factorisation of T - w_m I, twice, then orthogonalised
// and normalised once more.  The solve amplifies whatever survives the projection, so the
// result is an accurate eigenvector orthogonal to the others.  The vector being refined and
// the factors live in LDS; lane 0 runs the (sequential) factorisation and solves.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__global__ __launch_bounds__(256) void tri_orth_kernel(int n, const double* __restrict__ d_all,
                                                       const double* __restrict__ e_all,
                                                       const double* __restrict__ w_all,
                                                       const double* __restrict__ tn_all, double* __restrict__ X_all) {
  __shared__ double sd[NMAX], se[NMAX];
  __shared__ double ybuf[4][NMAX], fu0[4][NMAX], fu1[4][NMAX], fu2[4][NMAX], flm[4][NMAX];
  __shared__ unsigned char fpv[4][NMAX];
  const int b = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < n; i += 256) {
    sd[i] = d_all[(size_t)b * n + i];
    se[i] = e_all[(size_t)b * n + i];
  }
  __syncthreads();
  const int j = blockIdx.x * 4 + wv;
  if (j >= n) return;
  const double* w = w_all + (size_t)b * n;
  const double ctol = 1e-9 * tn_all[b];
  if (j > 0 && !(w[j] - w[j - 1] > ctol)) return;   // not the first of its cluster
  int end = j + 1;
  while (end < n && !(w[end] - w[end - 1] > ctol)) ++end;
  if (end - j < 2) return;
  const double ptol = DBL_EPSILON * fmax(tn_all[b], DBL_MIN);
  double* X = X_all + (size_t)b * n * n;
  double* y = ybuf[wv];
  const TriLU f{fu0[wv], fu1[wv], fu2[wv], flm[wv], fpv[wv], 1};
  auto project = [&](int m) {   // y <- y minus its components on members j..m-1, normalised
    for (int pass = 0; pass < 2; ++pass)
      for (int p = j; p < m; ++p) {
        double dt = 0.0;
        for (int i = lane; i < n; i += 64) dt += y[i] * X[(size_t)i * n + p];
        dt = wave_sum(dt);
        for (int i = lane; i < n; i += 64) y[i] -= dt * X[(size_t)i * n + p];
      }
    double ss = 0.0;
    for (int i = lane; i < n; i += 64) ss += y[i] * y[i];
    ss = wave_sum(ss);
    const double sc = ss > 0.0 ? 1.0 / sqrt(ss) : 0.0;
    for (int i = lane; i < n; i += 64) y[i] *= sc;
  };
  for (int m = j + 1; m < end; ++m) {
    for (int i = lane; i < n; i += 64) y[i] = X[(size_t)i * n + m];
    if (lane == 0) tri_factor(n, sd, se, w[m], ptol, f);
    for (int it = 0; it < 2; ++it) {
      project(m);
      wave_sync();
      if (lane == 0) (void)tri_solve(n, f, y, 1, 1.0);
      wave_sync();
    }
    project(m);
    for (int i = lane; i < n; i += 64) X[(size_t)i * n + m] = y[i];
  }
}

// Z = Q X = H_0 (H_1 (... (H_{n-3} X))): 32 columns of X per workgroup, held in LDS; thread
// (rgp, col) = (tid / 32, tid % 32) owns rows rgp, rgp + 8, ... of its column; the column dot
// products v_k . X[:, col] are reduced across the eight row groups through LDS.  The
// reflectors are staged BT_KB at a time in LDS with one coalesced pass.
constexpr int BT_COLS = 32, BT_KB = 16;
__global__ __launch_bounds__(256) void tri_back_kernel(int n, const double* __restrict__ Vr_all,
                                                       const double* __restrict__ tau_all,
                                                       const double* __restrict__ X_all, double* __restrict__ Z_all,
                                                       long sZ) {
  __shared__ double Xs[NMAX][BT_COLS + 1];
  __shared__ double Vs[BT_KB][NMAX];
  __shared__ double red[2][8][BT_COLS];
  const int b = blockIdx.y, tid = threadIdx.x, col = tid & 31, rgp = tid >> 5;
  const int c0 = blockIdx.x * BT_COLS, cj = c0 + col;
  const size_t nn = (size_t)n * n;
  const double* Vr = Vr_all + (size_t)b * nn;
  const double* tau = tau_all + (size_t)b * n;
  const double* X = X_all + (size_t)b * nn;
  for (int i = rgp; i < n; i += 8) Xs[i][col] = cj < n ? X[(size_t)i * n + cj] : 0.0;
  for (int khi = n - 3; khi >= 0; khi -= BT_KB) {
    const int klo = khi - BT_KB + 1 > 0 ? khi - BT_KB + 1 : 0;
    __syncthreads();   // the previous chunk's reflectors are no longer read
    for (int q = tid; q < (khi - klo + 1) * n; q += 256) {
      const int r = q / n, i = q - r * n;
      Vs[r][i] = Vr[(size_t)(klo + r) * n + i];
    }
    __syncthreads();
    for (int k = khi; k >= klo; --k) {
      const double tk = tau[k];
      if (tk == 0.0) continue;
      const double* v = Vs[k - klo];
      const int i0 = k + 1 + ((rgp - (k + 1)) % 8 + 8) % 8;   // first row >= k+1 of this group
      double part = 0.0;
      for (int i = i0; i < n; i += 8) part += v[i] * Xs[i][col];
      red[k & 1][rgp][col] = part;
      __syncthreads();
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < 8; ++r) s += red[k & 1][r][col];
      s *= tk;
      for (int i = i0; i < n; i += 8) Xs[i][col] -= v[i] * s;
    }
  }
  __syncthreads();
  double* Z = Z_all + (size_t)b * sZ;
  if (cj < n)
    for (int i = rgp; i < n; i += 8) Z[(size_t)i * n + cj] = Xs[i][col];
}

// Workspace of one matrix, in doubles: Vr, X, two n x n products of the orthogonality
// polish, d, e, tau, tn.
inline size_t eig_ws_doubles(int n) { return (size_t)4 * n * n + 4 * (size_t)n + 8; }

}  // namespace
}  // namespace tri
}  // namespace gdx
