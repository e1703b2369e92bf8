// The exchange-correlation grid path on gfx950: densities, energy densities, coefficient
// MLP, FP64 grid integration, and the hand-written adjoints dE/dD (V_xc) and dE/dtheta.
//
// Reference semantics followed (grad_dft/...):
//   densities        molecule.py:388-502   rho = sum D phi phi, grad rho = 2 sum D phi dphi,
//                                          tau = 1/2 sum D dphi.dphi,
//                                          lapl = 2 sum D dphi.dphi + 2 sum D phi lapl(phi)
//   energy densities popular_functionals.py:29-326 (gdx_functionals.h)
//   coefficients     functional.py:143-158, 396-419 (dense layers, 2 sigmoid(z/2) head)
//   integration      functional.py:219-253, 316-342 (abs_clip of densities, of the
//                    per-point product and of the weights; FP64 sum)
//   dE/dD            train.py:86-121 (value_and_grad w.r.t. rdm1, unsymmetrised)
//
// Precision: the matrix products (density X = A D, V_xc, MLP layers) run on MFMA in the
// element type T of the path -- double (GDX_FP64) or float (GDX_FP32); the per-point
// functional math, the density row contractions, every reduction over grid points or
// split partials and the E_xc sum are FP64 in both paths.
//
// Kernel sequence, forward:  pad_dm -> density_ao -> point_fwd -> [layer0 -> gemm_tile
//   (bias+ELU) x (L-1) -> head_fwd (skinny MFMA + coefficient/energy epilogue)] -> reduce.
// Backward:  head_seed -> [head_delta -> skinny_wgrad (last layer) -> per hidden layer
//   gemm_tile wgrad (split-K) + gemm_tile bwd-data (ELU') -> layer0_bwd (skinny) +
//   skinny_wgrad (first layer)] -> point_bwd (duals) -> vxc_ao -> reduce_vxc.
#include "gdx_common.h"
#include "gdx_functionals.h"
#include "gdx_profile.h"
#include "gdx_gemm.h"
#include "gdx_gemm_bf16.h"
#include "gdx_skinny.h"
#include "gdx_aogemm.h"
#include "gdx_rows.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace gdx {
namespace {

constexpr int NT = 256;  // threads per block in every kernel here
constexpr double SQUASH = 1e-4;
constexpr int MAXND = 8;

// Constant coefficients of a non-neural functional over the four energy-density slots
// (lda, b88, vwn or pw92, lyp) and which correlation fills the third slot (0: VWN5, 1: PW92).
struct LinCoef {
  double c[4];
  int corr;
};
inline bool is_neural(const gdx_xc_desc& d) {
  return d.kind == GDX_FUNC_NEURAL_B3LYP || d.kind == GDX_FUNC_B3LYP_NEURAL_CORR;
}
inline LinCoef lin_coef(const gdx_xc_desc& d) {
  LinCoef lc;
  if (d.kind == GDX_FUNC_LINEAR) {
    for (int f = 0; f < 4; ++f) lc.c[f] = d.coef[f];
  } else {  // b3lyp_coefficients, grad_dft/popular_functionals.py:340-347 (without exact exchange)
    lc.c[0] = 1.0 - 0.2;
    lc.c[1] = 0.72;
    lc.c[2] = 1.0 - 0.81;
    lc.c[3] = 0.81;
  }
  lc.corr = d.correlation;
  return lc;
}

__device__ __forceinline__ double clip30(double x) { return fabs(x) > 1e-30 ? x : 0.0; }
__device__ __forceinline__ double elu_grad_from_out(double y) { return y > 0.0 ? 1.0 : y + 1.0; }

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = warp_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
  return t;
}

// ---------------------------------------------------------------- small kernels ----

template <typename T>
__global__ void pad_dm_kernel(const double* __restrict__ rdm1, int n, int nP, T* __restrict__ Dp) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = 2L * nP * nP;
  if (i >= tot) return;
  const int s = (int)(i / ((long)nP * nP));
  const int r = (int)((i / nP) % nP);
  const int c = (int)(i % nP);
  Dp[i] = (T)((r < n && c < n) ? rdm1[((long)s * n + r) * n + c] : 0.0);
}

template <typename T>
__global__ void convert_kernel(const double* __restrict__ src, long n, T* __restrict__ dst) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (T)src[i];
}

// single-block deterministic sum of n block partials -> *out
__global__ void reduce_scalar_kernel(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double red[NT / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];
  acc = warp_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; ++w) t += red[w];
    *out = t;
  }
}

// V partials (nchunk, nsv, nP, nP) -> d_rdm1 (2, n, n); nsv = 1: spin 0 copied to spin 1
__global__ void reduce_vxc_kernel(const double* __restrict__ part, int nchunk, int nsv, int nP, int n,
                                  double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = 2L * n * n;
  if (i >= tot) return;
  const int s = nsv == 1 ? 0 : (int)(i / ((long)n * n));
  const int a = (int)((i / n) % n);
  const int b = (int)(i % n);
  const long off = ((long)s * nP + a) * nP + b;
  const long stride = (long)nsv * nP * nP;
  double acc = 0.0;
  for (int c = 0; c < nchunk; ++c) acc += part[c * stride + off];
  out[i] = acc;
}

// First stage of a two-stage sum over many slices: out[y][i] = sum over slices s in
// [y per, (y+1) per) of part[s * stride + i]; a block covers 64 consecutive i with 4 slice
// subgroups combined in LDS (fixed order: deterministic).
template <typename TP>
__device__ __forceinline__ void reduce_stage1_block(const TP* __restrict__ part, int nslices, long stride, long len,
                                                    int per, double* __restrict__ out, int bx, int by) {
  __shared__ double red[4][64];
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const long i = (long)bx * 64 + c;
  const int s0 = by * per, s1 = min(nslices, s0 + per);
  double acc = 0.0;
  if (i < len) {
    int s = s0 + g;
    for (; s + 12 < s1; s += 16) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = (double)part[(long)(s + 4 * u) * stride + i];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc += v[u];
    }
    for (; s < s1; s += 4) acc += (double)part[(long)s * stride + i];
  }
  red[g][c] = acc;
  __syncthreads();
  if (g == 0 && i < len) out[(long)by * len + i] = red[0][c] + red[1][c] + red[2][c] + red[3][c];
}

template <typename TP>
__global__ __launch_bounds__(256) void reduce_stage1_kernel(const TP* __restrict__ part, int nslices, long stride,
                                                            long len, int per, double* __restrict__ out) {
  reduce_stage1_block<TP>(part, nslices, stride, len, per, out, blockIdx.x, blockIdx.y);
}

// the first stage of two independent reductions in one launch (grid.y = the larger group
// count; blocks past a reduction's own count exit)
template <typename TP>
__global__ __launch_bounds__(256) void reduce_stage1_2_kernel(const TP* __restrict__ p0, int ns0, long st0, long len0,
                                                              int per0, int G0, double* __restrict__ t0, int nb0,
                                                              const TP* __restrict__ p1, int ns1, long st1, long len1,
                                                              int per1, int G1, double* __restrict__ t1) {
  if ((int)blockIdx.x < nb0) {
    if ((int)blockIdx.y < G0) reduce_stage1_block<TP>(p0, ns0, st0, len0, per0, t0, blockIdx.x, blockIdx.y);
  } else if ((int)blockIdx.y < G1) {
    reduce_stage1_block<TP>(p1, ns1, st1, len1, per1, t1, blockIdx.x - nb0, blockIdx.y);
  }
}

// ---------------------------------------------------------------- per-point fwd ----
//
// dens_fin (npts_pad, 2, 6): rho, grad rho (3), tau, lapl per spin (final)
// E (npts_pad, 4): clipped energy densities
// X16 (npts_pad, 16) (neural): six squashed MLP inputs log(|x| + 1e-4), a 1 (the bias
//   column of the first layer's gradient), zeros
template <typename T>
__global__ __launch_bounds__(NT) void point_fwd_kernel(
    const double* __restrict__ dens_part, int nsd, int npts, int npts_pad,
    const double* __restrict__ weights, int neural, LinCoef lc, double* __restrict__ dens_fin,
    double* __restrict__ E, T* __restrict__ X, double* __restrict__ eps_part) {
  __shared__ double red[NT / 64];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  if (r < npts_pad) {
    double d[2][DENS_STRIDE];
    for (int s = 0; s < 2; ++s) {
      const int ds = nsd == 1 ? 0 : s;
      const double* raw = dens_part + ((size_t)ds * npts_pad + r) * DENS_STRIDE;
      d[s][0] = raw[0];
      d[s][1] = 2.0 * raw[1];
      d[s][2] = 2.0 * raw[2];
      d[s][3] = 2.0 * raw[3];
      d[s][4] = 0.5 * raw[4];
      d[s][5] = 2.0 * raw[4] + 2.0 * raw[5];
#pragma unroll
      for (int q = 0; q < DENS_STRIDE; ++q) dens_fin[((size_t)r * 2 + s) * DENS_STRIDE + q] = d[s][q];
    }
    const double gu = d[0][1] * d[0][1] + d[0][2] * d[0][2] + d[0][3] * d[0][3];
    const double gd = d[1][1] * d[1][1] + d[1][2] * d[1][2] + d[1][3] * d[1][3];
    double e[4];
    b3lyp_parts<double>(d[0][0], d[1][0], gu, gd, d[0][5], d[1][5], lc.corr, e);
#pragma unroll
    for (int f = 0; f < 4; ++f) E[(size_t)r * 4 + f] = e[f];
    if (neural) {
      const double feat[6] = {d[0][0], d[1][0], gu, gd, d[0][4], d[1][4]};
#pragma unroll
      for (int j = 0; j < 6; ++j) X[(size_t)r * 16 + j] = (T)log(fabs(feat[j]) + SQUASH);
      X[(size_t)r * 16 + 6] = (T)1;
#pragma unroll
      for (int j = 7; j < 16; ++j) X[(size_t)r * 16 + j] = (T)0;
    } else if (weights != nullptr && r < npts) {
      const double* c = lc.c;
      double eps = c[0] * e[0] + c[1] * e[1] + c[2] * e[2] + c[3] * e[3];
      eps = clip30(clip30(eps));
      contrib = clip30(weights[r]) * eps;
    }
  }
  const double t = block_sum(contrib, red);
  if (!neural && threadIdx.x == 0) eps_part[blockIdx.x] = t;
}

// Row-wise products against at most 16 weight rows, for the MLP's narrow edges:
//   v[r][f] = sum_{j < kj} A[r][j] Wt[j][f]   (A with row stride 16, Wt (16 x h) in LDS)
//   ER_LAYER0:    out = ELU(v + b[f])             (first layer, forward)
//   ER_TANGENT0:  out = v * elu'(Y[r][f])         (first layer, tangent)
//   ER_DELTA:     out = v * elu'(Y[r][f])         (last layer's input adjoint)
//   ER_RAW:       out = v
// elu'(z) is taken from Y = ELU(z).  Each thread forms 8 consecutive features of one row
// (16- / 32-byte stores); a block covers ER_ROWS rows.
enum { ER_LAYER0 = 0, ER_TANGENT0 = 1, ER_DELTA = 2, ER_RAW = 3 };
constexpr int ER_ROWS = 64;

template <typename T, int MODE>
__global__ __launch_bounds__(NT) void edge_rows_kernel(const T* __restrict__ A, int kj, const T* __restrict__ Wt,
                                                       const double* __restrict__ bias, int h, long rows,
                                                       const T* __restrict__ Y, T* __restrict__ out) {
  typedef T v8 __attribute__((ext_vector_type(8)));
  __shared__ __align__(16) T Ws[16][256];
  __shared__ __align__(16) T bs[256];
  {
    // stage the 16 x h weight rows with 16-byte loads, all issued before the LDS stores
    // (rows >= kj are the buffer's zero padding)
    constexpr int VEC = 16 / sizeof(T), NBV = 16 * 256 / (NT * VEC);
    typedef T vec_t __attribute__((ext_vector_type(VEC)));
    vec_t wv[NBV];
#pragma unroll
    for (int i = 0; i < NBV; ++i) {
      const int e = (threadIdx.x + NT * i) * VEC;
      wv[i] = e < 16 * h ? *(const vec_t*)(Wt + e) : (vec_t)(T)0;
    }
#pragma unroll
    for (int i = 0; i < NBV; ++i) {
      const int e = (threadIdx.x + NT * i) * VEC;
      if (e < 16 * h) *(vec_t*)&Ws[e / h][e % h] = wv[i];
    }
  }
  if (MODE == ER_LAYER0)
    for (int e = threadIdx.x; e < h; e += NT) bs[e] = (T)bias[e];
  __syncthreads();
  const int tpr = h / 8;            // threads per row
  const int rpp = NT / tpr;         // rows per pass
  const int t = threadIdx.x;
  if (t >= rpp * tpr) return;
  const int f0 = (t % tpr) * 8;
  const long rbase = (long)blockIdx.x * ER_ROWS;
  for (int rr = t / tpr; rr < ER_ROWS; rr += rpp) {
    const long r = rbase + rr;
    if (r >= rows) break;
    T a[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) a[j] = j < kj ? A[r * 16 + j] : (T)0;
    v8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (T)0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      if (j < kj) {
        const v8 w = *(const v8*)&Ws[j][f0];
        v += a[j] * w;
      }
    }
    const size_t o = (size_t)r * h + f0;
    if (MODE == ER_LAYER0) {
      const v8 b = *(const v8*)&bs[f0];
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = elu_any(v[e] + b[e]);
    } else if (MODE == ER_TANGENT0 || MODE == ER_DELTA) {
      const v8 y = *(const v8*)(Y + o);
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] *= (y[e] > (T)0 ? (T)1 : y[e] + (T)1);
    }
    *(v8*)(out + o) = v;
  }
}

// ------------------------------------------------- guarded GEMM (any shape, small) ----
// Used only when a shape is not whole 128 x 128 tiles (narrow test MLPs).
constexpr int GBK = 16;

template <typename T, int BM, int BN, bool TA, bool TB, int EPI>
__global__ __launch_bounds__(NT) void gemm_any_kernel(
    int M, int N, int K, const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb,
    T* __restrict__ C, int ldc, const T* __restrict__ bias, const T* __restrict__ Y, int ldy, int kchunk,
    T* __restrict__ colsum, int kseg, int sps) {
  typedef typename MfmaT<T>::acc_t acc_t;
  constexpr int WM = BM / 2, WN = BN / 2, MB = WM / 16, NB = WN / 16;
  __shared__ T As[GBK][BM + 2];
  __shared__ T Bs[GBK][BN + 2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  int kbeg, kend;
  split_range(blockIdx.z, kchunk, kseg, sps, kbeg, kend);
  kend = min(K, kend);
  T csum = 0;
  acc_t acc[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = (acc_t){0, 0, 0, 0};
  for (int k0 = kbeg; k0 < kend; k0 += GBK) {
    for (int e = tid; e < BM * GBK; e += NT) {
      int m, k;
      if (TA) { k = e / BM; m = e % BM; } else { m = e / GBK; k = e % GBK; }
      const int gm = m0 + m, gk = k0 + k;
      T v = 0;
      if (gm < M && gk < kend) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      As[k][m] = v;
    }
    for (int e = tid; e < BN * GBK; e += NT) {
      int n, k;
      if (TB) { n = e / GBK; k = e % GBK; } else { k = e / BN; n = e % BN; }
      const int gn = n0 + n, gk = k0 + k;
      T v = 0;
      if (gn < N && gk < kend) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
      Bs[k][n] = v;
    }
    __syncthreads();
    if (colsum != nullptr && blockIdx.x == 0 && tid < BN)
#pragma unroll
      for (int k = 0; k < GBK; ++k) csum += Bs[k][tid];
#pragma unroll
    for (int kk = 0; kk < GBK / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      T af[MB], bf[NB];
#pragma unroll
      for (int i = 0; i < MB; ++i) af[i] = As[kr][wm + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < NB; ++j) bf[j] = Bs[kr][wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = MfmaT<T>::op(af[i], bf[j], acc[i][j]);
    }
    __syncthreads();
  }
  if (colsum != nullptr && blockIdx.x == 0 && tid < BN && n0 + tid < N)
    colsum[(size_t)blockIdx.z * N + n0 + tid] = csum;
  T* Cz = C + (size_t)blockIdx.z * M * ldc;
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int col = n0 + wn + j * 16 + (lane & 15);
      if (col >= N) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + i * 16 + MfmaT<T>::row(lane, q);
        if (row >= M) continue;
        T v = acc[i][j][q];
        if (EPI == EPI_BIAS_ELU) v = elu_any(v + bias[col]);
        if (EPI == EPI_ELU_BWD) {
          const T y = Y[(size_t)row * ldy + col];
          v *= (y > (T)0 ? (T)1 : y + (T)1);
        }
        Cz[(size_t)row * ldc + col] = v;
      }
    }
}

// ----------------------------------------------------------------- MLP head ----

// Padded weight tiles for the skinny kernels: WT16 (16 x h) rows o < nd = W_last[:, o];
// W016 (16 x h) rows j < nc = W0[j, :]; every other row zero.
template <typename T>
__global__ void prep_edge_weights_kernel(const double* __restrict__ W0, int nc,
                                         const double* __restrict__ Wl, int nd, int h,
                                         T* __restrict__ W016, T* __restrict__ WT16) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 16 * h) return;
  const int row = i / h, f = i % h;
  W016[i] = (T)(row < nc ? W0[row * h + f] : 0.0);
  WT16[i] = (T)(row < nd ? Wl[f * nd + row] : 0.0);
}

// Adjoint seeds, one thread per point.  g = g_exc clip(w) [|eps| > 1e-30];
// GE = dE/de_f = g c_f; neural: DZ16 = dE/dz = g e_f dc_f/dz zero-padded to 16,
// and block partials of the last layer's bias gradient sum_r DZ[r][o] -> dbpart[block][o].
template <typename T>
__global__ __launch_bounds__(NT) void head_seed_kernel(
    const T* __restrict__ Cc, const double* __restrict__ E, const double* __restrict__ weights,
    int nd, int npts, int npts_pad, const double* __restrict__ g_exc, int neural,
    double* __restrict__ GE, T* __restrict__ DZ16, double* __restrict__ dbpart, int head_mode, LinCoef lc) {
  __shared__ double red[NT / 64];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const double* bc = lc.c;
  double dz[MAXND];
  for (int o = 0; o < MAXND; ++o) dz[o] = 0.0;
  if (r < npts_pad) {
    double c[MAXND], e[MAXND];
    double eps = 0.0;
    // padded rows (r >= npts) have no coefficients stored: they take c = 0, seed 0
    for (int o = 0; o < nd; ++o) {
      c[o] = r >= npts ? 0.0 : (neural ? (double)Cc[(size_t)r * nd + o] : bc[o]);
      e[o] = E[(size_t)r * nd + o];
      eps += c[o] * e[o];
    }
    const double g = (r < npts && fabs(eps) > 1e-30) ? (*g_exc) * clip30(weights[r]) : 0.0;
    for (int o = 0; o < nd; ++o) {
      GE[(size_t)r * nd + o] = g * c[o];
      dz[o] = g * e[o] * head_d1(c[o], head_mode);
    }
    if (neural)
      for (int o = 0; o < 16; ++o) DZ16[(size_t)r * 16 + o] = (T)(o < nd ? dz[o] : 0.0);
  }
  if (neural)
    for (int o = 0; o < nd; ++o) {
      const double t = block_sum(dz[o], red);
      if (threadIdx.x == 0) dbpart[(size_t)blockIdx.x * nd + o] = t;
    }
}

// d_params of the last layer: W (h, nd) from the reduced transposed (16 x h) tile
__global__ void unpack_last_kernel(const double* __restrict__ dWt16, int h, int nd, double* __restrict__ dW) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= h * nd) return;
  const int f = i / nd, o = i % nd;
  dW[i] = dWt16[(size_t)o * h + f];
}

// Per-point potentials for the V_xc build from dE/de_f (GE) and dE/dx (DX, neural):
// forward-mode duals over (rho_up, rho_dn, |grad rho_up|^2, |grad rho_dn|^2, lapl_up,
// lapl_dn) give sum_f GE_f de_f/dv; the MLP input gradient adds through the log squash.
template <typename T>
__global__ __launch_bounds__(NT) void point_bwd_kernel(const double* __restrict__ dens_fin,
                                                       const double* __restrict__ GE, const T* __restrict__ DX,
                                                       int nc, int npts, int npts_pad, int corr,
                                                       double* __restrict__ vpt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npts_pad) return;
  double out[2][VPT_STRIDE];
  for (int s = 0; s < 2; ++s)
    for (int q = 0; q < VPT_STRIDE; ++q) out[s][q] = 0.0;
  if (r < npts) {
    const double* d0 = dens_fin + ((size_t)r * 2 + 0) * DENS_STRIDE;
    const double* d1 = dens_fin + ((size_t)r * 2 + 1) * DENS_STRIDE;
    const double gu = d0[1] * d0[1] + d0[2] * d0[2] + d0[3] * d0[3];
    const double gd = d1[1] * d1[1] + d1[2] * d1[2] + d1[3] * d1[3];
    typedef Dual<6> D6;
    D6 e[4];
    b3lyp_parts<D6>(dvar<6>(d0[0], 0), dvar<6>(d1[0], 1), dvar<6>(gu, 2), dvar<6>(gd, 3),
                    dvar<6>(d0[5], 4), dvar<6>(d1[5], 5), corr, e);
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int f = 0; f < 4; ++f) {
      const double ge = GE[(size_t)r * 4 + f];
      for (int i = 0; i < 6; ++i) v[i] += ge * e[f].d[i];
    }
    double vtau[2] = {0.0, 0.0};
    if (DX != nullptr) {
      const double feat[6] = {d0[0], d1[0], gu, gd, d0[4], d1[4]};
      double df[6];
      for (int j = 0; j < 6; ++j) {
        const double sg = feat[j] > 0.0 ? 1.0 : (feat[j] < 0.0 ? -1.0 : 0.0);
        df[j] = (double)DX[(size_t)r * nc + j] * sg / (fabs(feat[j]) + SQUASH);
      }
      v[0] += df[0]; v[1] += df[1]; v[2] += df[2]; v[3] += df[3];
      vtau[0] = df[4]; vtau[1] = df[5];
    }
    const double* ds[2] = {d0, d1};
    for (int s = 0; s < 2; ++s) {
      const double vrho = v[s], vg = v[2 + s], vl = v[4 + s];
      out[s][0] = vrho;
      out[s][1] = 4.0 * vg * ds[s][1];
      out[s][2] = 4.0 * vg * ds[s][2];
      out[s][3] = 4.0 * vg * ds[s][3];
      out[s][4] = 2.0 * vl;
      out[s][5] = 0.5 * vtau[s] + 2.0 * vl;
    }
  }
  for (int s = 0; s < 2; ++s)
    for (int q = 0; q < VPT_STRIDE; ++q) vpt[((size_t)s * npts_pad + r) * VPT_STRIDE + q] = out[s][q];
}

// ------------------------------------------------- second order (HVP) kernels ----
//
// For a direction dD (2, n, n) the HVP pass differentiates
//   Phi(theta, D) = g E(theta, D) + <dD, dE/dD>
// where <dD, dE/dD> = Edot is the forward-mode derivative of E along dD.  Tangents (dot)
// are carried through densities, energy densities, features and the MLP; the reverse sweep
// then runs on stacked [value; tangent] rows (2 N_pad rows per activation buffer).

__device__ __forceinline__ Dual<1> tan1(double v, double t) {
  Dual<1> x(0.0);
  x.v = v;
  x.d[0] = t;
  return x;
}

// Final tangent densities from the raw sums of dD; energy-density tangents (forward mode
// along the tangent of (rho, |grad rho|^2, lapl) per spin); MLP input tangents (Xdot16,
// zero bias column).
template <typename T>
__global__ __launch_bounds__(NT) void point_tangent_kernel(const double* __restrict__ dpart_t,
                                                           const double* __restrict__ dens_fin, int npts_pad,
                                                           int neural, int corr, double* __restrict__ dfin_t,
                                                           double* __restrict__ Edot, T* __restrict__ Xdot) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npts_pad) return;
  double u[2][DENS_STRIDE], t[2][DENS_STRIDE];
  for (int s = 0; s < 2; ++s) {
    const double* raw = dpart_t + ((size_t)s * npts_pad + r) * DENS_STRIDE;
    t[s][0] = raw[0];
    t[s][1] = 2.0 * raw[1];
    t[s][2] = 2.0 * raw[2];
    t[s][3] = 2.0 * raw[3];
    t[s][4] = 0.5 * raw[4];
    t[s][5] = 2.0 * raw[4] + 2.0 * raw[5];
#pragma unroll
    for (int q = 0; q < DENS_STRIDE; ++q) {
      dfin_t[((size_t)r * 2 + s) * DENS_STRIDE + q] = t[s][q];
      u[s][q] = dens_fin[((size_t)r * 2 + s) * DENS_STRIDE + q];
    }
  }
  double gam[2], gamt[2];
  for (int s = 0; s < 2; ++s) {
    gam[s] = u[s][1] * u[s][1] + u[s][2] * u[s][2] + u[s][3] * u[s][3];
    gamt[s] = 2.0 * (u[s][1] * t[s][1] + u[s][2] * t[s][2] + u[s][3] * t[s][3]);
  }
  Dual<1> e[4];
  b3lyp_parts<Dual<1>>(tan1(u[0][0], t[0][0]), tan1(u[1][0], t[1][0]), tan1(gam[0], gamt[0]),
                       tan1(gam[1], gamt[1]), tan1(u[0][5], t[0][5]), tan1(u[1][5], t[1][5]), corr, e);
#pragma unroll
  for (int f = 0; f < 4; ++f) Edot[(size_t)r * 4 + f] = e[f].d[0];
  if (neural) {
    const double feat[6] = {u[0][0], u[1][0], gam[0], gam[1], u[0][4], u[1][4]};
    const double featt[6] = {t[0][0], t[1][0], gamt[0], gamt[1], t[0][4], t[1][4]};
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double sg = feat[j] > 0.0 ? 1.0 : (feat[j] < 0.0 ? -1.0 : 0.0);
      Xdot[(size_t)r * 16 + j] = (T)(sg * featt[j] / (fabs(feat[j]) + SQUASH));
    }
#pragma unroll
    for (int j = 6; j < 16; ++j) Xdot[(size_t)r * 16 + j] = (T)0;
  }
}

// Seeds of the HVP reverse sweep, one thread per point.  With m = clip(w) [|eps| > 1e-30],
// c = head_coef(z), c' and c'' its derivatives (sigmoid head: c' = (c/2)(1 - c/2),
// c'' = c'(1 - c)/2; linear correction: 1, 0), cdot = c' zdot:
//   Edot    += m sum_o (cdot_o e_o + c_o edot_o)                     (block partials)
//   GE_o     = dPhi/de_o    = m (g c_o + cdot_o),   GEd_o = dPhi/dedot_o = m c_o
//   zbar_o   = m (g e_o + edot_o) c'_o + m e_o c''_o zdot_o,  zdotbar_o = m e_o c'_o
// DZ2 rows [0, N_pad) hold zbar, rows [N_pad, 2 N_pad) zdotbar (zero-padded to 16).
template <typename T>
__global__ __launch_bounds__(NT) void hvp_seed_kernel(
    const T* __restrict__ Cc, const T* __restrict__ ZT, const double* __restrict__ E,
    const double* __restrict__ Edot, const double* __restrict__ weights, int nd, int npts, int npts_pad,
    const double* __restrict__ g_exc, int neural, double* __restrict__ GE, double* __restrict__ GEd,
    T* __restrict__ DZ2, double* __restrict__ dbpart, double* __restrict__ edot_part, int head_mode,
    LinCoef lc, const double* __restrict__ g_pt) {
  __shared__ double red[NT / 64];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const double* bc = lc.c;
  double zb[MAXND];
  for (int o = 0; o < MAXND; ++o) zb[o] = 0.0;
  double ed_contrib = 0.0;
  if (r < npts_pad) {
    double c[MAXND], e[MAXND], et[MAXND], cp[MAXND], zt[MAXND];
    double eps = 0.0;
    for (int o = 0; o < nd; ++o) {
      c[o] = r >= npts ? 0.0 : (neural ? (double)Cc[(size_t)r * nd + o] : bc[o]);
      e[o] = E[(size_t)r * nd + o];
      et[o] = Edot[(size_t)r * nd + o];
      cp[o] = neural ? head_d1(c[o], head_mode) : 0.0;
      zt[o] = (neural && r < npts) ? (double)ZT[(size_t)r * nd + o] : 0.0;
      eps += c[o] * e[o];
    }
    const double m = (r < npts && fabs(eps) > 1e-30) ? clip30(weights[r]) : 0.0;
    const double g = g_pt ? g_pt[r] : *g_exc;
    double zdb[MAXND];
    for (int o = 0; o < nd; ++o) {
      const double cd = cp[o] * zt[o];
      ed_contrib += m * (cd * e[o] + c[o] * et[o]);
      GE[(size_t)r * nd + o] = m * (g * c[o] + cd);
      GEd[(size_t)r * nd + o] = m * c[o];
      const double cpp = neural ? head_d2(c[o], head_mode) : 0.0;
      zb[o] = m * (g * e[o] + et[o]) * cp[o] + m * e[o] * cpp * zt[o];
      zdb[o] = m * e[o] * cp[o];
    }
    if (neural)
      for (int o = 0; o < 16; ++o) {
        DZ2[(size_t)r * 16 + o] = (T)(o < nd ? zb[o] : 0.0);
        DZ2[((size_t)npts_pad + r) * 16 + o] = (T)(o < nd ? zdb[o] : 0.0);
      }
  }
  if (neural)
    for (int o = 0; o < nd; ++o) {
      const double t = block_sum(zb[o], red);
      if (threadIdx.x == 0) dbpart[(size_t)blockIdx.x * nd + o] = t;
    }
  const double t = block_sum(ed_contrib, red);
  if (threadIdx.x == 0) edot_part[blockIdx.x] = t;
}

// Stacked activation adjoints -> pre-activation adjoints, in place.  With a = elu(z),
// adot = elu'(z) zdot, G2 = [abar; adotbar] (raw products with W^T), H2 = [a; adot]:
//   zbar    = abar elu'(z) + adotbar elu''(z) zdot = abar elu'(z) + [a <= 0] adotbar adot
//   zdotbar = adotbar elu'(z)
template <typename T>
__global__ __launch_bounds__(NT) void hvp_combine_kernel(T* __restrict__ G2, const T* __restrict__ H2, long half) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half) return;
  const T a = H2[i], ad = H2[half + i];
  const T gb = G2[i], gd = G2[half + i];
  const bool pos = a > (T)0;
  const T ep = pos ? (T)1 : a + (T)1;
  G2[i] = gb * ep + (pos ? (T)0 : gd * ad);
  G2[half + i] = gd * ep;
}

// Per-point potentials of the HVP: the gradient of
//   Phi_r(u) = sum_o [GE_o e_o + GEd_o edot_o] + sum_j [xbar_j x_j + xdotbar_j xdot_j]
// with respect to the final densities u of the point (tangents udot fixed), mapped to the
// raw-sum layout consumed by vxc_ao_kernel (as point_bwd_kernel).  Second derivatives of the
// energy densities come from Dual<6, Dual<1>>: d[i].v = de/dq_i, d[i].d[0] = (H qdot)_i.
template <typename T>
__global__ __launch_bounds__(NT) void point_hvp_bwd_kernel(const double* __restrict__ dens_fin,
                                                           const double* __restrict__ dfin_t,
                                                           const double* __restrict__ GE,
                                                           const double* __restrict__ GEd, const T* __restrict__ DX2,
                                                           int nc, int npts, int npts_pad, int corr,
                                                           double* __restrict__ vpt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npts_pad) return;
  double out[2][VPT_STRIDE];
  for (int s = 0; s < 2; ++s)
    for (int q = 0; q < VPT_STRIDE; ++q) out[s][q] = 0.0;
  if (r < npts) {
    double u[2][DENS_STRIDE], t[2][DENS_STRIDE];
    for (int s = 0; s < 2; ++s)
      for (int q = 0; q < DENS_STRIDE; ++q) {
        u[s][q] = dens_fin[((size_t)r * 2 + s) * DENS_STRIDE + q];
        t[s][q] = dfin_t[((size_t)r * 2 + s) * DENS_STRIDE + q];
      }
    double gam[2], gamt[2];
    for (int s = 0; s < 2; ++s) {
      gam[s] = u[s][1] * u[s][1] + u[s][2] * u[s][2] + u[s][3] * u[s][3];
      gamt[s] = 2.0 * (u[s][1] * t[s][1] + u[s][2] * t[s][2] + u[s][3] * t[s][3]);
    }
    typedef Dual<6, Dual<1>> H6;
    H6 e[4];
    b3lyp_parts<H6>(dvar<6, Dual<1>>(tan1(u[0][0], t[0][0]), 0), dvar<6, Dual<1>>(tan1(u[1][0], t[1][0]), 1),
                    dvar<6, Dual<1>>(tan1(gam[0], gamt[0]), 2), dvar<6, Dual<1>>(tan1(gam[1], gamt[1]), 3),
                    dvar<6, Dual<1>>(tan1(u[0][5], t[0][5]), 4), dvar<6, Dual<1>>(tan1(u[1][5], t[1][5]), 5), corr,
                    e);
    // P_i = dPhi/dq_i, Pd_i = dPhi/dqdot_i over q = (rho_u, rho_d, gam_u, gam_d, lapl_u, lapl_d)
    double P[6] = {0, 0, 0, 0, 0, 0}, Pd[6] = {0, 0, 0, 0, 0, 0};
    for (int f = 0; f < 4; ++f) {
      const double ge = GE[(size_t)r * 4 + f], ged = GEd[(size_t)r * 4 + f];
      for (int i = 0; i < 6; ++i) {
        P[i] += ge * e[f].d[i].v + ged * e[f].d[i].d[0];
        Pd[i] += ged * e[f].d[i].v;
      }
    }
    // Q_j = dPhi/df_j, Qd_j = dPhi/dfdot_j over f = (rho_u, rho_d, gam_u, gam_d, tau_u, tau_d)
    double Q[6] = {0, 0, 0, 0, 0, 0}, Qd[6] = {0, 0, 0, 0, 0, 0};
    if (DX2 != nullptr) {
      const double feat[6] = {u[0][0], u[1][0], gam[0], gam[1], u[0][4], u[1][4]};
      const double featt[6] = {t[0][0], t[1][0], gamt[0], gamt[1], t[0][4], t[1][4]};
      for (int j = 0; j < 6; ++j) {
        const double xb = (double)DX2[(size_t)r * nc + j];
        const double xdb = (double)DX2[((size_t)npts_pad + r) * nc + j];
        const double sg = feat[j] > 0.0 ? 1.0 : (feat[j] < 0.0 ? -1.0 : 0.0);
        const double inv = 1.0 / (fabs(feat[j]) + SQUASH);
        Q[j] = xb * sg * inv - xdb * sg * sg * featt[j] * inv * inv;
        Qd[j] = xdb * sg * inv;
      }
    }
    for (int s = 0; s < 2; ++s) {
      const double vr = P[s] + Q[s];
      const double vg = P[2 + s] + Q[2 + s], vgd = Pd[2 + s] + Qd[2 + s];
      const double vl = P[4 + s], vt = Q[4 + s];
      out[s][0] = vr;
      for (int j = 0; j < 3; ++j) out[s][1 + j] = 4.0 * (vg * u[s][1 + j] + vgd * t[s][1 + j]);
      out[s][4] = 2.0 * vl;
      out[s][5] = 0.5 * vt + 2.0 * vl;
    }
  }
  for (int s = 0; s < 2; ++s)
    for (int q = 0; q < VPT_STRIDE; ++q) vpt[((size_t)s * npts_pad + r) * VPT_STRIDE + q] = out[s][q];
}

// ------------------------------------------------------------ workspace layout ----
constexpr int NUM_CU = 256;    // MI355X compute units
constexpr int RED_GROUPS = 32; // slice groups of the first reduction stage

// Point chunking of the V_xc build: the grid (tiles x chunks x spins) is sized to whole
// rounds of co-resident blocks (resident blocks per CU from the kernels' register and LDS
// use), so no CU is left with an extra block at the end.
struct VxcSplit {
  int vt, ntile, nchunk, chunk;
  bool flush;   // FP32 accumulators flushed into FP64 (long chunks)
};
// FP32 V_xc: every build flushes its FP32 partial sums into FP64 (VX_FLUSH_MIN = 0): the
// SCF adjoint through many cycles is sensitive to the forward Fock matrices, and 500-point
// FP32 sums measurably degrade it (relative gradient error 7e-5 against 5e-6 flushed)
constexpr int VX_FLUSH_MIN = 0;
// Output tile edge of the V_xc build chosen by gdx_xc_autotune (measured on the device for a
// shape), keyed by (N_pad, n_pad, nsv, precision, force_flush); absent: the default rule.
std::mutex vxc_tune_mu;
std::map<std::tuple<int, int, int, int, int>, int> vxc_tune;

int tuned_vt(int Np, int nP, int nsv, int prec, bool force_flush) {
  std::lock_guard<std::mutex> g(vxc_tune_mu);
  const auto it = vxc_tune.find(std::make_tuple(Np, nP, nsv, prec, (int)force_flush));
  return it == vxc_tune.end() ? 0 : it->second;
}

// prec: GDX_FP64 / GDX_FP32 / GDX_BF16X6 (bf16 MFMA for tiles of 128).  force_flush: the
// FP32 build always flushes its accumulators into FP64 (used by the second-order pass, whose
// potentials cancel more strongly than the first-order ones).  vt: output tile edge (64 or
// 128); 0 takes the tuned value for this shape, else 64 for n_pad <= 64 and 128 above.
VxcSplit vxc_split(int Np, int nP, int nsv, int prec, bool force_flush = false, int vt = 0) {
  VxcSplit v{};
  const bool fp64 = prec == GDX_FP64;
  if (vt == 0) vt = tuned_vt(Np, nP, nsv, prec, force_flush);
  if (vt == 0) vt = nP <= 64 ? 64 : 128;
  v.vt = vt;
  const int nt = (nP + v.vt - 1) / v.vt;
  v.ntile = nt * nt;
  int resident = v.vt == 64 ? (fp64 ? 4 : 7) : (fp64 ? 2 : 3);
  if (prec == GDX_BF16X6 && v.vt == 128) resident = 1;
  const int per = v.ntile * nsv;
  // point chunks per (tile, spin): at least one full wave of blocks over the CUs, and among
  // up to 4x that count the one whose last wave of blocks is fullest (fewest idle slots)
  auto split = [&](int res) {
    const int slots = NUM_CU * res;
    const int cap = std::max(1, Np / 128);   // at least 128 points per chunk
    const int n0 = std::min(cap, std::max(1, (slots + per - 1) / per));
    int best = n0;
    double best_eff = 0.0;
    for (int n = n0; n <= std::min(cap, 4 * n0); ++n) {
      const long blocks = (long)per * n;
      const double eff = (double)blocks / ((double)((blocks + slots - 1) / slots) * slots);
      if (eff > best_eff + 1e-9) { best_eff = eff; best = n; }
    }
    v.chunk = (Np + best - 1) / best;
    v.chunk = (v.chunk + VX_P - 1) / VX_P * VX_P;
    v.nchunk = (Np + v.chunk - 1) / v.chunk;
  };
  split(resident);
  v.flush = false;
  if (!fp64 && !(prec == GDX_BF16X6 && v.vt == 128) && (force_flush || v.chunk > VX_FLUSH_MIN)) {
    // the flushing kernel holds 4 WB^2 FP64 accumulators a lane: for tiles of 128, two
    // blocks per CU for the single-tile case (its launch bounds), one otherwise
    v.flush = true;
    split(v.vt == 64 ? 4 : (v.ntile == 1 ? 2 : 1));
  }
  return v;
}
constexpr int SKW_ROWS = 512;  // grid points per block of the skinny weight-gradient kernel

// hvp = true: activation-side buffers hold stacked [value; tangent] rows (R = 2 N_pad), plus
// the tangent buffers of the HVP pass.
struct Layout {
  size_t Dp, dens_part, dens_fin, E, X, C, H, eps_part, GE, DZ, W16, Wlp, dbpart, dWt16, G, DX, vpt, vpart,
      wpart, cpart;
  size_t Dp2, dpart_t, dfin_t, Edot, ZT, GEd, edot_part, scal, rtmp;
  size_t total;
  size_t hstride;  // elements between the activation buffers of consecutive layers
  int nsd, nchunk, chunk, nsplit, kchunk, kseg, sps, npt_blocks, nwblk, wrows, rows;
  size_t nparam;
};

size_t elem_size(int precision) { return precision == GDX_FP64 ? 8 : 4; }
bool mlp_bf16(const gdx_xc_desc& d) { return d.precision == GDX_BF16X6; }

Layout make_layout(const gdx_xc_desc& d, bool hvp = false, bool batch = false) {
  Layout L{};
  const size_t Np = d.npts_pad, nP = d.nao_pad;
  const size_t R = hvp ? 2 * Np : Np;  // rows of the activation-side buffers
  const size_t es = elem_size(d.precision);
  const bool neural = is_neural(d);
  const int hmode = d.kind == GDX_FUNC_B3LYP_NEURAL_CORR ? 1 : 0;
  L.nsd = d.spin_restricted ? 1 : 2;
  L.rows = (int)R;
  {
    // partial slots nchunk x nsv, for every tile edge the tuner may pick
    L.nchunk = 0;
    for (int vt : {64, 128})
      for (bool ff : {false, true}) {
        const VxcSplit v1 = vxc_split((int)Np, (int)nP, 1, d.precision, ff, vt),
                       v2 = vxc_split((int)Np, (int)nP, 2, d.precision, ff, vt);
        L.nchunk = std::max(L.nchunk, std::max(v1.nchunk, 2 * v2.nchunk));
      }
  }
  // split-K for the hidden weight-gradient GEMMs: 128 splits of whole 32-row stages, the
  // last one of a segment shorter; for hvp the value and tangent rows are two segments of
  // 64 splits each, so the first half of the splits covers the value rows exactly
  {
    const int segs = hvp ? 2 : 1;
    L.kseg = (int)Np;
    L.sps = 128 / segs;
    const size_t kc = (Np + L.sps - 1) / L.sps;
    L.kchunk = (int)std::max<size_t>(GB_K, (kc + GB_K - 1) / GB_K * GB_K);
    L.nsplit = segs * L.sps;
  }
  L.npt_blocks = (int)((Np + NT - 1) / NT);
  L.wrows = SKW_ROWS;
  L.nwblk = (int)((R + SKW_ROWS - 1) / SKW_ROWS);
  const size_t h = neural ? d.mlp_h : 0, nh = neural ? d.mlp_nhidden : 0;
  L.hstride = R * h;
  size_t np = 0;
  if (neural) {
    int din = d.mlp_nc;
    for (int l = 0; l <= d.mlp_nhidden; ++l) {
      const int dout = l == d.mlp_nhidden ? d.mlp_nd : d.mlp_h;
      np += (size_t)din * dout + dout;
      din = dout;
    }
  }
  L.nparam = np;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += round_up(bytes, 256); return o; };
  L.Dp = take(batch ? 0 : 2 * nP * nP * es);
  L.dens_part = take((size_t)L.nsd * Np * DENS_STRIDE * 8);
  L.dens_fin = take(Np * 2 * DENS_STRIDE * 8);
  L.E = take(Np * 4 * 8);
  L.X = take(neural ? R * 16 * es : 0);
  L.C = take(neural ? Np * d.mlp_nd * es : 0);
  L.H = take(R * h * nh * es);
  L.eps_part = take(std::max<size_t>(Np / std::min(64, RD_ROWS) + 1, 4096) * 8);
  L.GE = take(Np * 4 * 8);
  L.DZ = take(neural ? R * 16 * es : 0);
  L.W16 = take(neural ? 2 * 16 * h * es : 0);
  L.Wlp = take(neural && es != 8 ? np * es : 0);
  L.dbpart = take(neural ? (size_t)L.npt_blocks * d.mlp_nd * 8 : 0);
  L.dWt16 = take(neural ? 16 * h * 8 : 0);
  L.G = take(neural ? 2 * R * h * es : 0);
  L.DX = take(neural ? R * d.mlp_nc * es : 0);
  L.vpt = take(2 * Np * VPT_STRIDE * 8);
  L.vpart = take(batch ? 0 : (size_t)L.nchunk * nP * nP * 8);
  size_t wbytes = 0;
  if (neural) {
    wbytes = (size_t)L.nsplit * h * h * es;
    wbytes = std::max(wbytes, (size_t)L.nwblk * 16 * h * 8);
    wbytes = std::max(wbytes, (size_t)((R + RW_ROWS - 1) / RW_ROWS) * 7 * h * 8);
    wbytes = std::max(wbytes, (size_t)(2 * (R / GT_M)) * (d.mlp_nc + 1) * h * 8);   // EPI_ELU_BWD_W0 slices
  }
  L.wpart = take(wbytes);
  L.cpart = take(neural ? (size_t)L.nsplit * h * es : 0);
  // two-stage reductions: up to two at once (reduce_slices2: the largest one plus the last layer's bias)
  L.rtmp = take((size_t)RED_GROUPS * (std::max(std::max(h * h, 16 * h), 2 * nP * nP) + 64) * 8);
  if (hvp) {
    L.Dp2 = take(batch ? 0 : 2 * nP * nP * es);
    L.dpart_t = take(2 * Np * DENS_STRIDE * 8);
    L.dfin_t = take(Np * 2 * DENS_STRIDE * 8);
    L.Edot = take(Np * 4 * 8);
    L.ZT = take(neural ? Np * d.mlp_nd * es : 0);
    L.GEd = take(Np * 4 * 8);
    L.edot_part = take((size_t)L.npt_blocks * 8);
    L.scal = take(2 * 8);
  }
  L.total = off;
  return L;
}

int validate(const gdx_xc_desc* d) {
  GDX_REQUIRE(d != nullptr, "null gdx_xc_desc");
  GDX_REQUIRE(d->npts > 0 && d->nao > 0, "empty grid or basis");
  GDX_REQUIRE(d->npts_pad >= d->npts && d->npts_pad % 64 == 0, "npts_pad must be a multiple of 64 >= npts");
  GDX_REQUIRE(d->nao_pad >= d->nao && d->nao_pad % 16 == 0, "nao_pad must be a multiple of 16 >= nao");
  GDX_REQUIRE(d->kind == GDX_FUNC_B3LYP || d->kind == GDX_FUNC_NEURAL_B3LYP || d->kind == GDX_FUNC_B3LYP_NEURAL_CORR ||
                  d->kind == GDX_FUNC_LINEAR,
              "unknown functional kind");
  GDX_REQUIRE(d->correlation == GDX_CORR_VWN5 || d->correlation == GDX_CORR_PW92,
              "correlation must be GDX_CORR_VWN5 or GDX_CORR_PW92");
  if (d->precision != GDX_FP64 && d->precision != GDX_FP32 && d->precision != GDX_BF16X6) {
    set_error("unknown precision (GDX_FP64, GDX_FP32, GDX_BF16X6)");
    return GDX_ERR_UNSUPPORTED;
  }
  if (is_neural(*d)) {
    GDX_REQUIRE(d->mlp_nc == 6, "coefficient inputs must be 6 (rho, |grad rho|^2, tau per spin)");
    GDX_REQUIRE(d->mlp_nd == 4, "MLP outputs must equal the 4 B3LYP energy densities");
    GDX_REQUIRE(d->mlp_h >= 16 && d->mlp_h % 16 == 0 && d->mlp_h <= 256,
                "MLP width must be a multiple of 16 in [16, 256]");
    GDX_REQUIRE(d->mlp_nhidden >= 1 && d->mlp_nhidden <= 16, "MLP hidden layers must be 1..16");
  }
  return GDX_OK;
}

inline unsigned nblocks(long n) { return (unsigned)((n + NT - 1) / NT); }
// the streaming edge kernels split a row over h / 4 lanes in xor groups: h a power of two
inline bool rows_path(int h) { return h >= 16 && (h & (h - 1)) == 0; }
inline unsigned nblocks64(long n) { return (unsigned)((n + 63) / 64); }

template <typename T> const char* tname(const char* f64, const char* f32) { return sizeof(T) == 8 ? f64 : f32; }

template <typename T, int BM, int BN, bool TA, bool TB, int EPI>
void launch_any(hipStream_t st, const char* name, int M, int N, int K, const T* A, int lda, const T* B, int ldb,
                T* C, int ldc, const T* bias, const T* Y, int ldy, int nsplit, int kchunk, T* colsum,
                int kseg = 0, int sps = 0) {
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, nsplit);
  if (sps <= 0) sps = nsplit;
  if (kseg <= 0) kseg = nsplit * kchunk;
  GDX_LAUNCH(name, (gemm_any_kernel<T, BM, BN, TA, TB, EPI>), grid, dim3(NT), 0, st, M, N, K, A, lda, B, ldb, C,
             ldc, bias, Y, ldy, kchunk, colsum, kseg, sps);
}

template <typename T, bool TA, bool TB, int EPI>
void launch_tile(hipStream_t st, const char* name, int M, int N, int nsplit, int kchunk, const T* A, int lda,
                 const T* B, int ldb, T* C, int ldc, const T* bias, const T* Y, int ldy, T* colsum,
                 bool bf16 = false, int kseg = 0, int sps = 0, const T* X16 = nullptr, double* w0part = nullptr,
                 int ncp = 0) {
  dim3 grid((M / GT_M) * (N / GT_N), nsplit);
  if (sps <= 0) sps = nsplit;
  if (kseg <= 0) kseg = nsplit * kchunk;
  if constexpr (sizeof(T) == 4) {
    if (bf16) {
      char nm[64];
      snprintf(nm, sizeof(nm), "%s", name);
      // "gemm_f32_*" -> "gemm_bf16x6_*" in the timing table
      const char* tag = TA ? "gemm_bf16x6_wgrad" : (EPI == EPI_BIAS_ELU ? "gemm_bf16x6_fwd" : "gemm_bf16x6_bwd_data");
      GDX_LAUNCH(tag, (gemm_bf16x6_kernel<TA, TB, EPI>), grid, dim3(NT), 0, st, M, N, A, lda, B, ldb, C, ldc, bias,
                 Y, ldy, kchunk, colsum, kseg, sps);
      (void)nm;
      return;
    }
  }
  GDX_LAUNCH(name, (gemm_tile_kernel<T, TA, TB, EPI>), grid, dim3(NT), 0, st, M, N, A, lda, B, ldb, C, ldc,
             bias, Y, ldy, kchunk, colsum, kseg, sps, X16, w0part, ncp);
}

// parameter offsets (elements) of W_l and b_l in the flat parameter vector
struct ParamOffsets {
  std::vector<size_t> w, b;
};
ParamOffsets param_offsets(const gdx_xc_desc& d) {
  ParamOffsets o;
  size_t off = 0;
  int din = d.mlp_nc;
  for (int l = 0; l <= d.mlp_nhidden; ++l) {
    const int dout = l == d.mlp_nhidden ? d.mlp_nd : d.mlp_h;
    o.w.push_back(off); off += (size_t)din * dout;
    o.b.push_back(off); off += dout;
    din = dout;
  }
  return o;
}

// ---------------------------------------------------------------- molecule batch ----
//
// A batch of molecules (one functional, one precision, different grids and bases) runs as one
// evaluation: every per-point buffer is the concatenation of the molecules' rows (each
// molecule's npts_pad a multiple of 256, so every 256-, 128- and 64-row block belongs to one
// molecule), the MLP GEMMs run once over all rows, and the AO-side kernels (padding of D, the
// density kernel, V_xc and its reduction) are single launches over a ragged queue of
// (molecule, tile) work items.  Per-molecule scalars (E_xc, edot) are segment sums of the
// block partials; per-molecule cotangents enter as per-point scales.

struct PadBatchMol {   // rdm1_cat[dm_off ...] (2, n, n) -> Dp region [dp_off ...] (2, nP, nP)
  long dm_off, dp_off;
  int n, nP, blk0, nblk;
};
struct RedBatchMol {   // V partials of one molecule -> d_rdm1_cat[out_off ...] (2, n, n)
  const double* vpart;
  long out_off;
  int nchunk, nP, n, blk0;
};

template <typename T>
__global__ void pad_dm_batch_kernel(const PadBatchMol* __restrict__ mols, int nmol, const double* __restrict__ rdm1,
                                    T* __restrict__ Dp) {
  int lo = 0, hi = nmol - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mols[mid].blk0 <= (int)blockIdx.x) lo = mid; else hi = mid - 1;
  }
  const PadBatchMol& M = mols[lo];
  const long i = (long)(blockIdx.x - M.blk0) * blockDim.x + threadIdx.x;
  const long tot = 2L * M.nP * M.nP;
  if (i >= tot) return;
  const int sp = (int)(i / ((long)M.nP * M.nP));
  const int r = (int)((i / M.nP) % M.nP);
  const int c = (int)(i % M.nP);
  Dp[M.dp_off + i] = (T)((r < M.n && c < M.n) ? rdm1[M.dm_off + ((long)sp * M.n + r) * M.n + c] : 0.0);
}

// out[m] = sum of part[seg[m] .. seg[m+1]) in the order of reduce_scalar_kernel
__global__ void reduce_segments_kernel(const double* __restrict__ part, const int* __restrict__ seg,
                                       double* __restrict__ out) {
  __shared__ double red[NT / 64];
  const int m = blockIdx.x;
  const int b = seg[m], n = seg[m + 1] - b;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += part[b + i];
  acc = warp_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; ++w) t += red[w];
    out[m] = t;
  }
}

// per-point g (g_pt[r] = g[m(r)]) and weights scaled by it (wg[r] = g[m(r)] w[r]); rows[m]
// is molecule m's first row, rows[nmol] the total
__global__ void segment_scale_kernel(const int* __restrict__ rows, int nmol, const double* __restrict__ g,
                                     const double* __restrict__ w, double* __restrict__ g_pt, double* __restrict__ wg) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows[nmol]) return;
  int lo = 0, hi = nmol - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rows[mid] <= r) lo = mid; else hi = mid - 1;
  }
  const double gm = g[lo];
  if (g_pt) g_pt[r] = gm;
  if (wg) wg[r] = gm * w[r];
}

// V partials (nchunk, nsv, nP, nP) of each molecule -> d_rdm1_cat (2, n, n) blocks, as
// reduce_vxc_kernel (nsv = 1: spin 0 copied to spin 1)
__global__ void reduce_vxc_batch_kernel(const RedBatchMol* __restrict__ mols, int nmol, int nsv,
                                        double* __restrict__ out) {
  int lo = 0, hi = nmol - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mols[mid].blk0 <= (int)blockIdx.x) lo = mid; else hi = mid - 1;
  }
  const RedBatchMol& M = mols[lo];
  const long i = (long)(blockIdx.x - M.blk0) * blockDim.x + threadIdx.x;
  const long n = M.n, nP = M.nP;
  if (i >= 2L * n * n) return;
  const int s = nsv == 1 ? 0 : (int)(i / (n * n));
  const long a = (i / n) % n, b = i % n;
  const long off = ((long)s * nP + a) * nP + b;
  const long stride = (long)nsv * nP * nP;
  double acc = 0.0;
  for (int c = 0; c < M.nchunk; ++c) acc += M.vpart[c * stride + off];
  out[M.out_off + i] = acc;
}

constexpr int VXB_CHUNK = 2048;   // points per V_xc block of a batch (multiple of VX_P)
// V_xc group of a molecule in a batch by its padded basis size (see BatchCtx::d_vx)
inline int vxb_group(int nP) { return nP <= 64 ? 0 : (nP <= 128 ? 1 : 2); }

struct BatchCtx {
  int nmol = 0;
  int ntile_total = 0;                       // 64-point tiles of the density launch
  const AoBatchMol* d_ao = nullptr;          // Dp -> the densities' padded matrices
  const AoBatchMol* d_ao_t = nullptr;        // Dp -> the HVP direction's padded matrices
  const PadBatchMol* d_pad = nullptr;
  int pad_blocks = 0;
  void* Dp = nullptr;                        // padded matrices (sum of 2 nP^2 elements)
  void* Dp_t = nullptr;
  // V_xc groups by padded basis size: <= 64 (one tile of 64), <= 128 (one tile of 128),
  // larger (tiles of 64: no more padding than 128 tiles up to 256, four blocks per CU)
  const VxcBatchMol* d_vx[3] = {nullptr, nullptr, nullptr};
  int nvx[3] = {0, 0, 0}, vx_blocks[3] = {0, 0, 0};
  const RedBatchMol* d_red = nullptr;
  int red_blocks = 0;
  const int* seg256 = nullptr;               // (nmol + 1) segment starts in 256-row blocks
  const int* seg64 = nullptr;                // in 64-row blocks
  const int* rows = nullptr;                 // (nmol + 1) first rows
  const double* weights = nullptr;           // concatenated grid weights
  double* wg = nullptr;                      // weights scaled by the molecules' cotangents
  double* g_pt = nullptr;                    // per-point cotangent
  double* one = nullptr;                     // device 1.0
  double* scal = nullptr;                    // nmol doubles of scratch
  long dm_total = 0;                         // elements of a concatenated (2, n, n) array
};

struct Ctx {
  const gdx_xc_desc& d;
  Layout L;
  char* ws;
  hipStream_t st;
  const BatchCtx* b = nullptr;
  template <typename U> U* P(size_t off) const { return (U*)(ws + off); }
};

template <typename T>
void launch_density_batch(const Ctx& c, const double* mats, bool tangent, int nspin, double* out) {
  const BatchCtx& B = *c.b;
  hipStream_t st = c.st;
  GDX_LAUNCH("pad_dm", pad_dm_batch_kernel<T>, dim3(B.pad_blocks), dim3(NT), 0, st, B.d_pad, B.nmol, mats,
             (T*)(tangent ? B.Dp_t : B.Dp));
  GDX_LAUNCH(tname<T>("density_f64", "density_f32"), density_ao_batch_kernel<T>, dim3(B.ntile_total, nspin),
             dim3(NT), 0, st, tangent ? B.d_ao_t : B.d_ao, B.nmol, (size_t)c.d.npts_pad, out);
}

template <typename T>
void launch_vxc_batch(const Ctx& c, int nsv, double* d_rdm1) {
  const BatchCtx& B = *c.b;
  hipStream_t st = c.st;
  const char* nm = tname<T>("vxc_f64", "vxc_f32");
  const double* vpt = c.P<double>(c.L.vpt);
  const size_t rows = (size_t)c.d.npts_pad;
  for (int g = 0; g < 3; ++g) {
    if (!B.nvx[g]) continue;
    const dim3 grid(B.vx_blocks[g], nsv);
    const VxcBatchMol* t = B.d_vx[g];
    const int nv = B.nvx[g];
    if constexpr (sizeof(T) == 8) {
      if (g == 0)
        GDX_LAUNCH(nm, (vxc_ao_batch_kernel<double, true, 64>), grid, dim3(NT), 0, st, t, nv, vpt, rows, nsv);
      else if (g == 1)
        GDX_LAUNCH(nm, (vxc_ao_batch_kernel<double, true, 128>), grid, dim3(NT), 0, st, t, nv, vpt, rows, nsv);
      else
        GDX_LAUNCH(nm, (vxc_ao_batch_kernel<double, false, 64>), grid, dim3(NT), 0, st, t, nv, vpt, rows, nsv);
    } else {
      if (g == 0)
        GDX_LAUNCH(nm, (vxc_ao_batch_kernel<float, true, 64, true>), grid, dim3(NT), 0, st, t, nv, vpt, rows, nsv);
      else if (g == 1)
        GDX_LAUNCH(nm, (vxc_ao_batch_kernel<float, true, 128, true>), grid, dim3(NT), 0, st, t, nv, vpt, rows, nsv);
      else
        GDX_LAUNCH(nm, (vxc_ao_batch_kernel<float, false, 64, true>), grid, dim3(NT), 0, st, t, nv, vpt, rows, nsv);
    }
  }
  GDX_LAUNCH("reduce_vxc", reduce_vxc_batch_kernel, dim3(B.red_blocks), dim3(NT), 0, st, B.d_red, B.nmol, nsv,
             d_rdm1);
}

// densities: pad D -> density_ao -> point_fwd
template <typename T>
void run_densities(const Ctx& c, const T* ao, const double* rdm1, const double* weights, double* dens_fin) {
  const gdx_xc_desc& d = c.d;
  const int Np = d.npts_pad, nP = d.nao_pad;
  const bool neural = is_neural(d);
  const int hmode = d.kind == GDX_FUNC_B3LYP_NEURAL_CORR ? 1 : 0;
  hipStream_t st = c.st;
  if (c.b) {
    launch_density_batch<T>(c, rdm1, false, c.L.nsd, c.P<double>(c.L.dens_part));
  } else {
    GDX_LAUNCH("pad_dm", pad_dm_kernel<T>, dim3(nblocks(2L * nP * nP)), dim3(NT), 0, st, rdm1, d.nao, nP,
               c.P<T>(c.L.Dp));
    GDX_LAUNCH(tname<T>("density_f64", "density_f32"), density_ao_kernel<T>, dim3(Np / DN_PT, c.L.nsd), dim3(NT), 0,
               st, ao, Np, nP, c.P<T>(c.L.Dp), c.P<double>(c.L.dens_part));
  }
  GDX_LAUNCH("point_fwd", point_fwd_kernel<T>, dim3(c.L.npt_blocks), dim3(NT), 0, st, c.P<double>(c.L.dens_part),
             c.L.nsd, d.npts, Np, weights, neural ? 1 : 0, lin_coef(d), dens_fin, c.P<double>(c.L.E), c.P<T>(c.L.X),
             c.P<double>(c.L.eps_part));
}

// Custom coefficient functions: the energy densities E (npts_pad, 4) and the coefficient
// inputs X (npts_pad, 16, first six columns) of the first npts points out to the caller, and the
// caller's cotangents of both (dE/de_f, dE/dx_j) into the buffers point_bwd reads (zero rows
// past npts).
template <typename T>
__global__ void custom_out_kernel(int npts, const double* __restrict__ E, const T* __restrict__ X,
                                  double* __restrict__ e_out, double* __restrict__ x_out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npts) return;
#pragma unroll
  for (int f = 0; f < 4; ++f) e_out[(size_t)r * 4 + f] = E[(size_t)r * 4 + f];
#pragma unroll
  for (int j = 0; j < 6; ++j) x_out[(size_t)r * 6 + j] = (double)X[(size_t)r * 16 + j];
}

template <typename T>
__global__ void custom_in_kernel(int npts, int npts_pad, int nc, const double* __restrict__ ge,
                                 const double* __restrict__ dx, double* __restrict__ GE, T* __restrict__ DX) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npts_pad) return;
  const bool in = r < npts;
#pragma unroll
  for (int f = 0; f < 4; ++f) GE[(size_t)r * 4 + f] = in ? ge[(size_t)r * 4 + f] : 0.0;
  for (int j = 0; j < nc; ++j) DX[(size_t)r * nc + j] = (T)(in && j < 6 ? dx[(size_t)r * 6 + j] : 0.0);
}

// hidden-layer weights in the path's element type (params are FP64 masters)
template <typename T>
const T* lowp_params(const Ctx& c, const double* params) {
  if (sizeof(T) == 8) return (const T*)params;
  T* dst = c.P<T>(c.L.Wlp);
  GDX_LAUNCH("convert_params", convert_kernel<T>, dim3(nblocks((long)c.L.nparam)), dim3(NT), 0, c.st, params,
             (long)c.L.nparam, dst);
  return dst;
}

// out[i] = sum_s part[s * stride + i] for i < len: one pass for few slices, else two
// stages through L.rtmp (RED_GROUPS groups), so long slice lists run at full width.
template <typename TP>
void reduce_slices(const Ctx& c, const TP* part, int nslices, long stride, long len, double* out) {
  hipStream_t st = c.st;
  if (nslices <= 16 || (nslices <= 128 && len >= 16384)) {
    GDX_LAUNCH("reduce_cols", reduce_cols_kernel<TP>, dim3(nblocks64(len)), dim3(NT), 0, st, part, nslices, stride,
               len, out);
    return;
  }
  const int per = (nslices + RED_GROUPS - 1) / RED_GROUPS;
  const int G = (nslices + per - 1) / per;
  double* tmp = c.P<double>(c.L.rtmp);
  GDX_LAUNCH("reduce_cols", reduce_stage1_kernel<TP>, dim3(nblocks64(len), G), dim3(NT), 0, st, part, nslices, stride,
             len, per, tmp);
  GDX_LAUNCH("reduce_cols", reduce_cols_kernel<double>, dim3(nblocks64(len)), dim3(NT), 0, st, tmp, G, len, len, out);
}

// Two reductions at once (the weight and bias partials of one split-K GEMM, the last layer's
// weight and bias partials): one launch when both fit one pass -- the second then also runs
// in one pass -- two when both need two stages, else as two reduce_slices calls.
template <typename TP>
void reduce_slices2(const Ctx& c, const TP* p0, int ns0, long st0, long len0, double* o0, const TP* p1, int ns1,
                    long st1, long len1, double* o1) {
  hipStream_t st = c.st;
  auto one_pass = [](int ns, long len) { return ns <= 16 || (ns <= 128 && len >= 16384); };
  if (one_pass(ns0, len0) && ns1 <= 128) {
    const int nb0 = (int)nblocks64(len0);
    GDX_LAUNCH("reduce_cols", reduce_cols2_kernel<TP>, dim3(nb0 + nblocks64(len1)), dim3(NT), 0, st, p0, ns0, st0,
               len0, o0, nb0, p1, ns1, st1, len1, o1);
    return;
  }
  if (one_pass(ns0, len0) || one_pass(ns1, len1)) {
    reduce_slices<TP>(c, p0, ns0, st0, len0, o0);
    reduce_slices<TP>(c, p1, ns1, st1, len1, o1);
    return;
  }
  const int per0 = (ns0 + RED_GROUPS - 1) / RED_GROUPS, G0 = (ns0 + per0 - 1) / per0;
  const int per1 = (ns1 + RED_GROUPS - 1) / RED_GROUPS, G1 = (ns1 + per1 - 1) / per1;
  double* t0 = c.P<double>(c.L.rtmp);
  double* t1 = t0 + (size_t)G0 * len0;
  const int nb0 = (int)nblocks64(len0);
  GDX_LAUNCH("reduce_cols", reduce_stage1_2_kernel<TP>, dim3(nb0 + nblocks64(len1), std::max(G0, G1)), dim3(NT), 0,
             st, p0, ns0, st0, len0, per0, G0, t0, nb0, p1, ns1, st1, len1, per1, G1, t1);
  GDX_LAUNCH("reduce_cols", reduce_cols2_kernel<double>, dim3(nb0 + nblocks64(len1)), dim3(NT), 0, st, t0, G0, len0,
             len0, o0, nb0, t1, G1, len1, len1, o1);
}

// The V_xc kernel of a split: partial sums of the nsv spins into L.vpart.
template <typename T>
void launch_vxc_kernel(const Ctx& c, const T* ao, int nsv, const VxcSplit& v) {
  const gdx_xc_desc& d = c.d;
  const Layout& L = c.L;
  hipStream_t st = c.st;
  const int Np = d.npts_pad, nP = d.nao_pad;
  const char* nm = tname<T>("vxc_f64", "vxc_f32");
  const dim3 grid(v.ntile, v.nchunk, nsv);
  const double* vpt = c.P<double>(L.vpt);
  double* vpart = c.P<double>(L.vpart);
  const bool same = v.ntile == 1;
  if constexpr (sizeof(T) == 4) {
    if (d.precision == GDX_BF16X6 && v.vt == 128) {
      if (same)
        GDX_LAUNCH("vxc_bf16x6", vxc_bf16x6_kernel<true>, grid, dim3(NT), 0, st, ao, Np, nP, vpt, v.chunk, vpart);
      else
        GDX_LAUNCH("vxc_bf16x6", vxc_bf16x6_kernel<false>, grid, dim3(NT), 0, st, ao, Np, nP, vpt, v.chunk, vpart);
      return;
    }
    if (v.flush) {
      if (v.vt == 64 && same)
        GDX_LAUNCH(nm, (vxc_ao_kernel<T, true, 64, true>), grid, dim3(NT), 0, st, ao, Np, nP, vpt, v.chunk, vpart);
      else if (v.vt == 64)
        GDX_LAUNCH(nm, (vxc_ao_kernel<T, false, 64, true>), grid, dim3(NT), 0, st, ao, Np, nP, vpt, v.chunk, vpart);
      else if (same)
        GDX_LAUNCH(nm, (vxc_ao_kernel<T, true, 128, true>), grid, dim3(NT), 0, st, ao, Np, nP, vpt, v.chunk, vpart);
      else
        GDX_LAUNCH(nm, (vxc_ao_kernel<T, false, 128, true>), grid, dim3(NT), 0, st, ao, Np, nP, vpt, v.chunk, vpart);
      return;
    }
  }
  if (v.vt == 64 && same)
    GDX_LAUNCH(nm, (vxc_ao_kernel<T, true, 64>), grid, dim3(NT), 0, st, ao, Np, nP, vpt, v.chunk, vpart);
  else if (v.vt == 64)
    GDX_LAUNCH(nm, (vxc_ao_kernel<T, false, 64>), grid, dim3(NT), 0, st, ao, Np, nP, vpt, v.chunk, vpart);
  else if (same)
    GDX_LAUNCH(nm, (vxc_ao_kernel<T, true, 128>), grid, dim3(NT), 0, st, ao, Np, nP, vpt, v.chunk, vpart);
  else
    GDX_LAUNCH(nm, (vxc_ao_kernel<T, false, 128>), grid, dim3(NT), 0, st, ao, Np, nP, vpt, v.chunk, vpart);
}

// V_s (nsv spins) from the per-point potentials in L.vpt, reduced into d_rdm1 (2, n, n)
template <typename T>
void launch_vxc(const Ctx& c, const T* ao, int nsv, double* d_rdm1, bool force_flush = false) {
  const gdx_xc_desc& d = c.d;
  const Layout& L = c.L;
  hipStream_t st = c.st;
  const int Np = d.npts_pad, nP = d.nao_pad;
  const VxcSplit v = vxc_split(Np, nP, nsv, d.precision, force_flush);
  launch_vxc_kernel<T>(c, ao, nsv, v);
  // chunk partials -> RED_GROUPS partials (first stage), then the (2, n, n) output
  const long len = (long)nsv * nP * nP;
  const int per = (v.nchunk + RED_GROUPS - 1) / RED_GROUPS;
  const int G = (v.nchunk + per - 1) / per;
  GDX_LAUNCH("reduce_vxc", reduce_stage1_kernel<double>, dim3(nblocks64(len), G), dim3(NT), 0, st,
             c.P<double>(L.vpart), v.nchunk, len, len, per, c.P<double>(L.rtmp));
  GDX_LAUNCH("reduce_vxc", reduce_vxc_kernel, dim3(nblocks(2L * d.nao * d.nao)), dim3(NT), 0, st,
             c.P<double>(L.rtmp), G, nsv, nP, d.nao, d_rdm1);
}

template <typename T>
int forward_impl(const Ctx& c, const T* ao, const double* weights, const double* rdm1, const double* params,
                 double* exc) {
  const gdx_xc_desc& d = c.d;
  const Layout& L = c.L;
  hipStream_t st = c.st;
  const bool neural = is_neural(d);
  const int hmode = d.kind == GDX_FUNC_B3LYP_NEURAL_CORR ? 1 : 0;
  const int Np = d.npts_pad;
  run_densities<T>(c, ao, rdm1, weights, c.P<double>(L.dens_fin));
  int nred = L.npt_blocks;
  if (neural) {
    const int h = d.mlp_h;
    const bool tiles = Np % GT_M == 0 && h % GT_N == 0;
    const ParamOffsets po = param_offsets(d);
    const T* plp = lowp_params<T>(c, params);
    T* H = c.P<T>(L.H);
    const int nl = d.mlp_nhidden;
    T* W016 = c.P<T>(L.W16);
    T* WT16 = W016 + 16 * h;
    GDX_LAUNCH("prep_edge_weights", prep_edge_weights_kernel<T>, dim3(nblocks(16L * h)), dim3(NT), 0, st,
               params + po.w[0], d.mlp_nc, params + po.w[nl], d.mlp_nd, h, W016, WT16);
    GDX_LAUNCH("layer0", (edge_rows_kernel<T, ER_LAYER0>), dim3((Np + ER_ROWS - 1) / ER_ROWS), dim3(NT), 0, st,
               c.P<T>(L.X), d.mlp_nc, W016, params + po.b[0], h, (long)Np, nullptr, H);
    for (int l = 1; l < d.mlp_nhidden; ++l) {
      const T* Wl = plp + po.w[l];
      const T* bl = plp + po.b[l];
      const T* Hin = H + (size_t)(l - 1) * L.hstride;
      T* Hout = H + (size_t)l * L.hstride;
      if (tiles)
        launch_tile<T, false, false, EPI_BIAS_ELU>(st, tname<T>("gemm_f64_fwd", "gemm_f32_fwd"), Np, h, 1, h, Hin,
                                                   h, Wl, h, Hout, h, bl, nullptr, 0, nullptr, mlp_bf16(d));
      else
        launch_any<T, 128, 128, false, false, EPI_BIAS_ELU>(st, "gemm_any", Np, h, h, Hin, h, Wl, h, Hout, h, bl,
                                                            nullptr, 0, 1, h, nullptr);
    }
    if (rows_path(h)) {
      nred = (Np + RD_ROWS - 1) / RD_ROWS;
      GDX_LAUNCH("head_fwd", (rowdot_mfma_kernel<T, true>), dim3(nred), dim3(NT), 0, st, Np, h,
                 H + (size_t)(nl - 1) * L.hstride, WT16, c.P<T>(L.C), d.mlp_nd, plp + po.b[nl], c.P<double>(L.E),
                 weights, d.npts, c.P<double>(L.eps_part), hmode);
    } else {
      GDX_LAUNCH("head_fwd", (skinny_out16_kernel<T, true>), dim3(Np / 64), dim3(NT), 0, st, Np, h,
                 H + (size_t)(nl - 1) * L.hstride, h, WT16, c.P<T>(L.C), d.mlp_nd, d.mlp_nd, plp + po.b[nl],
                 d.mlp_nd, c.P<double>(L.E), weights, d.npts, c.P<double>(L.eps_part), hmode);
      nred = Np / 64;
    }
  }
  if (c.b) {
    // block partials of 256 rows (point_fwd) or 64 rows (head kernels)
    const int* seg = nred == L.npt_blocks ? c.b->seg256 : c.b->seg64;
    GDX_LAUNCH("reduce_scalar", reduce_segments_kernel, dim3(c.b->nmol), dim3(NT), 0, st, c.P<double>(L.eps_part), seg,
               exc);
  } else {
    GDX_LAUNCH("reduce_scalar", reduce_scalar_kernel, dim3(1), dim3(NT), 0, st, c.P<double>(L.eps_part), nred, exc);
  }
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

template <typename T>
int backward_impl(const Ctx& c, const T* ao, const double* weights, const double* params, const double* g_exc,
                  double* d_rdm1, double* d_params) {
  const gdx_xc_desc& d = c.d;
  const Layout& L = c.L;
  hipStream_t st = c.st;
  const bool neural = is_neural(d);
  const int hmode = d.kind == GDX_FUNC_B3LYP_NEURAL_CORR ? 1 : 0;
  const int Np = d.npts_pad, nP = d.nao_pad;
  const int nd = 4;
  if (neural) {
    const int h = d.mlp_h;
    const bool tiles = Np % GT_M == 0 && h % GT_N == 0;
    const ParamOffsets po = param_offsets(d);
    const int nl = d.mlp_nhidden;
    // the low-precision parameter copy and edge tiles were written by the forward call
    const T* plp = sizeof(T) == 8 ? (const T*)params : c.P<T>(L.Wlp);
    T* H = c.P<T>(L.H);
    T* G0 = c.P<T>(L.G);
    T* G1 = G0 + (size_t)Np * h;
    const T* Hlast = H + (size_t)(nl - 1) * L.hstride;
    const T* W016 = c.P<T>(L.W16);
    GDX_LAUNCH("head_seed", head_seed_kernel<T>, dim3(L.npt_blocks), dim3(NT), 0, st, c.P<T>(L.C),
               c.P<double>(L.E), weights, nd, d.npts, Np, g_exc, 1, c.P<double>(L.GE), c.P<T>(L.DZ),
               c.P<double>(L.dbpart), hmode, lin_coef(d));
    const int rwblk = (Np + RW_ROWS - 1) / RW_ROWS;
    if (rows_path(h)) {
      GDX_LAUNCH("head_delta", delta_wgrad_rows_kernel<T>, dim3(rwblk), dim3(NT), 0, st, Hlast, h, Np, c.P<T>(L.DZ),
                 W016 + 16 * h, nd, G0, c.P<double>(L.wpart));
    } else {
      GDX_LAUNCH("head_delta", (edge_rows_kernel<T, ER_DELTA>), dim3((Np + ER_ROWS - 1) / ER_ROWS), dim3(NT), 0, st,
                 c.P<T>(L.DZ), nd, W016 + 16 * h, nullptr, h, (long)Np, Hlast, G0);
      if (d_params)
        GDX_LAUNCH("skinny_wgrad", skinny_wgrad_kernel<T>, dim3(L.nwblk), dim3(NT), 0, st, Np, L.wrows, c.P<T>(L.DZ),
                   Hlast, h, c.P<double>(L.wpart));
    }
    if (d_params) {
      if (rows_path(h))
        reduce_slices2<double>(c, c.P<double>(L.wpart), rwblk, (long)nd * h, (long)nd * h, c.P<double>(L.dWt16),
                               c.P<double>(L.dbpart), L.npt_blocks, (long)nd, (long)nd, d_params + po.b[nl]);
      else
        reduce_slices2<double>(c, c.P<double>(L.wpart), L.nwblk, 16L * h, 16L * h, c.P<double>(L.dWt16),
                               c.P<double>(L.dbpart), L.npt_blocks, (long)nd, (long)nd, d_params + po.b[nl]);
      GDX_LAUNCH("unpack_last", unpack_last_kernel, dim3(nblocks((long)h * nd)), dim3(NT), 0, st,
                 c.P<double>(L.dWt16), h, nd, d_params + po.w[nl]);
    }
    T* Gcur = G0;
    T* Gnext = G1;
    const bool w0_fused = d_params && tiles && rows_path(h) && !mlp_bf16(d) && nl >= 2;
    for (int l = nl - 1; l >= 1; --l) {
      const T* Hin = H + (size_t)(l - 1) * L.hstride;  // input activations of layer l
      const T* Wl = plp + po.w[l];
      if (d_params) {
        T* wp = c.P<T>(L.wpart);
        T* cp = c.P<T>(L.cpart);
        if (tiles)
          launch_tile<T, true, false, EPI_STORE>(st, tname<T>("gemm_f64_wgrad", "gemm_f32_wgrad"), h, h, L.nsplit,
                                                 L.kchunk, Hin, h, Gcur, h, wp, h, nullptr, nullptr, 0, cp,
                                                 mlp_bf16(d), L.kseg, L.sps);
        else
          launch_any<T, 64, 64, true, false, EPI_STORE>(st, "gemm_any", h, h, Np, Hin, h, Gcur, h, wp, h, nullptr,
                                                        nullptr, 0, L.nsplit, L.kchunk, cp, L.kseg, L.sps);
        reduce_slices2<T>(c, wp, L.nsplit, (long)h * h, (long)h * h, d_params + po.w[l], cp, L.nsplit, (long)h,
                          (long)h, d_params + po.b[l]);
      }
      // delta for layer l's input: (Gcur W_l^T) * elu'(Hin); for l = 1 that delta's first-layer
      // weight-gradient partials come out of the same epilogue (w0_fused)
      if (tiles && l == 1 && w0_fused)
        launch_tile<T, false, true, EPI_ELU_BWD_W0>(st, tname<T>("gemm_f64_bwd_data_w0", "gemm_f32_bwd_data_w0"), Np, h,
                                                    1, h, Gcur, h, Wl, h, Gnext, h, nullptr, Hin, h, nullptr, false,
                                                    0, 0, c.P<T>(L.X), c.P<double>(L.wpart), d.mlp_nc + 1);
      else if (tiles)
        launch_tile<T, false, true, EPI_ELU_BWD>(st, tname<T>("gemm_f64_bwd_data", "gemm_f32_bwd_data"), Np, h, 1,
                                                 h, Gcur, h, Wl, h, Gnext, h, nullptr, Hin, h, nullptr, mlp_bf16(d));
      else
        launch_any<T, 128, 128, false, true, EPI_ELU_BWD>(st, "gemm_any", Np, h, h, Gcur, h, Wl, h, Gnext, h,
                                                          nullptr, Hin, h, 1, h, nullptr);
      std::swap(Gcur, Gnext);
    }
    // first layer: DX = G W0^T, dW0 and db0 = X16^T G (rows 0..nc-1 and the ones row nc);
    // W0 (nc x h) followed by b0 is contiguous in params
    const long l0len = (long)(d.mlp_nc + 1) * h;
    if (rows_path(h)) {
      GDX_LAUNCH("layer0_bwd", (rowdot_mfma_kernel<T, false>), dim3((Np + RD_ROWS - 1) / RD_ROWS), dim3(NT), 0, st, Np, h, Gcur, W016,
                 c.P<T>(L.DX), d.mlp_nc, nullptr, nullptr, nullptr, 0, nullptr, 0);
      if (w0_fused) {
        reduce_slices<double>(c, c.P<double>(L.wpart), 2 * (Np / GT_M), l0len, l0len, d_params + po.w[0]);
      } else if (d_params) {
        GDX_LAUNCH("wgrad0", wgrad0_rows_kernel<T>, dim3(rwblk), dim3(NT), 0, st, Gcur, h, Np, c.P<T>(L.X), d.mlp_nc,
                   c.P<double>(L.wpart));
        reduce_slices<double>(c, c.P<double>(L.wpart), rwblk, l0len, l0len, d_params + po.w[0]);
      }
    } else {
      GDX_LAUNCH("layer0_bwd", (skinny_out16_kernel<T, false>), dim3(Np / 64), dim3(NT), 0, st, Np, h, Gcur, h,
                 W016, c.P<T>(L.DX), d.mlp_nc, d.mlp_nc, nullptr, 0, nullptr, nullptr, 0, nullptr);
      if (d_params) {
        GDX_LAUNCH("skinny_wgrad", skinny_wgrad_kernel<T>, dim3(L.nwblk), dim3(NT), 0, st, Np, L.wrows, c.P<T>(L.X),
                   Gcur, h, c.P<double>(L.wpart));
        reduce_slices<double>(c, c.P<double>(L.wpart), L.nwblk, 16L * h, l0len, d_params + po.w[0]);
      }
    }
  } else {
    GDX_LAUNCH("head_seed", head_seed_kernel<T>, dim3(L.npt_blocks), dim3(NT), 0, st, nullptr, c.P<double>(L.E),
               weights, nd, d.npts, Np, g_exc, 0, c.P<double>(L.GE), nullptr, nullptr, 0, lin_coef(d));
  }
  if (d_rdm1) {
    GDX_LAUNCH("point_bwd", point_bwd_kernel<T>, dim3(L.npt_blocks), dim3(NT), 0, st, c.P<double>(L.dens_fin),
               c.P<double>(L.GE), neural ? c.P<T>(L.DX) : nullptr, d.mlp_nc, d.npts, Np, d.correlation,
               c.P<double>(L.vpt));
    // B3LYP on equal spin densities has equal spin potentials: form V once
    if (c.b)
      launch_vxc_batch<T>(c, (!neural && d.spin_restricted) ? 1 : 2, d_rdm1);
    else
      launch_vxc<T>(c, ao, (!neural && d.spin_restricted) ? 1 : 2, d_rdm1);
  }
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

// Gradient of Phi = g E + <dD, dE/dD> w.r.t. D and theta (see the HVP kernels above).
template <typename T>
int hvp_impl(const Ctx& c, const T* ao, const double* weights, const double* rdm1, const double* params,
             const double* dir, const double* g_exc, double* exc, double* edot, double* d_rdm1, double* d_params) {
  const gdx_xc_desc& d = c.d;
  const Layout& L = c.L;
  hipStream_t st = c.st;
  const bool neural = is_neural(d);
  const int hmode = d.kind == GDX_FUNC_B3LYP_NEURAL_CORR ? 1 : 0;
  const int Np = d.npts_pad, nP = d.nao_pad;
  const int nd = 4;
  double* scal = c.b ? c.b->scal : c.P<double>(L.scal);
  int rc = forward_impl<T>(c, ao, weights, rdm1, params, exc ? exc : scal);
  if (rc != GDX_OK) return rc;
  // tangent densities along dir (both spins, general direction)
  if (c.b) {
    launch_density_batch<T>(c, dir, true, 2, c.P<double>(L.dpart_t));
  } else {
    GDX_LAUNCH("pad_dm", pad_dm_kernel<T>, dim3(nblocks(2L * nP * nP)), dim3(NT), 0, st, dir, d.nao, nP,
               c.P<T>(L.Dp2));
    GDX_LAUNCH(tname<T>("density_f64", "density_f32"), density_ao_kernel<T>, dim3(Np / DN_PT, 2), dim3(NT), 0, st,
               ao, Np, nP, c.P<T>(L.Dp2), c.P<double>(L.dpart_t));
  }
  T* X2 = c.P<T>(L.X);
  GDX_LAUNCH("point_tangent", point_tangent_kernel<T>, dim3(L.npt_blocks), dim3(NT), 0, st,
             c.P<double>(L.dpart_t), c.P<double>(L.dens_fin), Np, neural ? 1 : 0, d.correlation, c.P<double>(L.dfin_t),
             c.P<double>(L.Edot), X2 + (size_t)Np * 16);
  const ParamOffsets po = neural ? param_offsets(d) : ParamOffsets{};
  const int h = neural ? d.mlp_h : 0;
  const int nl = neural ? d.mlp_nhidden : 0;
  const bool tiles = neural && Np % GT_M == 0 && h % GT_N == 0;
  const T* plp = neural ? (sizeof(T) == 8 ? (const T*)params : c.P<T>(L.Wlp)) : nullptr;
  T* H = c.P<T>(L.H);
  const T* W016 = c.P<T>(L.W16);
  const T* WT16 = W016 + 16 * h;
  if (neural) {
    // tangent MLP: Hdot_l = (Hdot_{l-1} W_l) * elu'(z_l); zdot of the head
    GDX_LAUNCH("layer0_tangent", (edge_rows_kernel<T, ER_TANGENT0>), dim3((Np + ER_ROWS - 1) / ER_ROWS), dim3(NT),
               0, st, X2 + (size_t)Np * 16, d.mlp_nc, W016, nullptr, h, (long)Np, H, H + (size_t)Np * h);
    for (int l = 1; l < nl; ++l) {
      const T* Wl = plp + po.w[l];
      const T* Hd_in = H + (size_t)(l - 1) * L.hstride + (size_t)Np * h;
      const T* Hl = H + (size_t)l * L.hstride;
      T* Hd_out = H + (size_t)l * L.hstride + (size_t)Np * h;
      if (tiles)
        launch_tile<T, false, false, EPI_ELU_BWD>(st, tname<T>("gemm_f64_tangent", "gemm_f32_tangent"), Np, h, 1,
                                                  h, Hd_in, h, Wl, h, Hd_out, h, nullptr, Hl, h, nullptr,
                                                  mlp_bf16(d));
      else
        launch_any<T, 128, 128, false, false, EPI_ELU_BWD>(st, "gemm_any", Np, h, h, Hd_in, h, Wl, h, Hd_out, h,
                                                           nullptr, Hl, h, 1, h, nullptr);
    }
    GDX_LAUNCH("head_tangent", (skinny_out16_kernel<T, false>), dim3(Np / 64), dim3(NT), 0, st, Np, h,
               H + (size_t)(nl - 1) * L.hstride + (size_t)Np * h, h, WT16, c.P<T>(L.ZT), nd, nd, nullptr, 0,
               nullptr, nullptr, 0, nullptr);
  }
  GDX_LAUNCH("hvp_seed", hvp_seed_kernel<T>, dim3(L.npt_blocks), dim3(NT), 0, st, neural ? c.P<T>(L.C) : nullptr,
             neural ? c.P<T>(L.ZT) : nullptr, c.P<double>(L.E), c.P<double>(L.Edot), weights, nd, d.npts, Np, g_exc,
             neural ? 1 : 0, c.P<double>(L.GE), c.P<double>(L.GEd), neural ? c.P<T>(L.DZ) : nullptr,
             c.P<double>(L.dbpart), c.P<double>(L.edot_part), hmode, lin_coef(d), c.b ? c.b->g_pt : nullptr);
  if (edot && c.b)
    GDX_LAUNCH("reduce_scalar", reduce_segments_kernel, dim3(c.b->nmol), dim3(NT), 0, st, c.P<double>(L.edot_part),
               c.b->seg256, edot);
  else if (edot)
    GDX_LAUNCH("reduce_scalar", reduce_scalar_kernel, dim3(1), dim3(NT), 0, st, c.P<double>(L.edot_part),
               L.npt_blocks, edot);
  if (neural) {
    const int R = 2 * Np;
    const long half = (long)Np * h;
    T* G0 = c.P<T>(L.G);
    T* G1 = G0 + (size_t)R * h;
    const T* Hlast2 = H + (size_t)(nl - 1) * L.hstride;
    T* DZ2 = c.P<T>(L.DZ);
    GDX_LAUNCH("head_delta", (edge_rows_kernel<T, ER_RAW>), dim3((R + ER_ROWS - 1) / ER_ROWS), dim3(NT), 0, st,
               DZ2, nd, WT16, nullptr, h, (long)R, nullptr, G0);
    GDX_LAUNCH("hvp_combine", hvp_combine_kernel<T>, dim3(nblocks(half)), dim3(NT), 0, st, G0, Hlast2, half);
    if (d_params) {
      GDX_LAUNCH("skinny_wgrad", skinny_wgrad_kernel<T>, dim3(L.nwblk), dim3(NT), 0, st, R, L.wrows, DZ2, Hlast2, h,
                 c.P<double>(L.wpart));
      reduce_slices<double>(c, c.P<double>(L.wpart), L.nwblk, 16L * h, 16L * h, c.P<double>(L.dWt16));
      GDX_LAUNCH("unpack_last", unpack_last_kernel, dim3(nblocks((long)h * nd)), dim3(NT), 0, st,
                 c.P<double>(L.dWt16), h, nd, d_params + po.w[nl]);
      reduce_slices<double>(c, c.P<double>(L.dbpart), L.npt_blocks, (long)nd, (long)nd, d_params + po.b[nl]);
    }
    T* Gcur = G0;
    T* Gnext = G1;
    for (int l = nl - 1; l >= 1; --l) {
      const T* Hin2 = H + (size_t)(l - 1) * L.hstride;  // [a_{l-1}; adot_{l-1}]
      const T* Wl = plp + po.w[l];
      if (d_params) {
        T* wp = c.P<T>(L.wpart);
        T* cp = c.P<T>(L.cpart);
        if (tiles)
          launch_tile<T, true, false, EPI_STORE>(st, tname<T>("gemm_f64_wgrad", "gemm_f32_wgrad"), h, h, L.nsplit,
                                                 L.kchunk, Hin2, h, Gcur, h, wp, h, nullptr, nullptr, 0, cp,
                                                 mlp_bf16(d), L.kseg, L.sps);
        else
          launch_any<T, 64, 64, true, false, EPI_STORE>(st, "gemm_any", h, h, R, Hin2, h, Gcur, h, wp, h, nullptr,
                                                        nullptr, 0, L.nsplit, L.kchunk, cp, L.kseg, L.sps);
        reduce_slices<T>(c, wp, L.nsplit, (long)h * h, (long)h * h, d_params + po.w[l]);
        // bias: column sums over the value rows only (the first half of the splits)
        reduce_slices<T>(c, cp, L.nsplit / 2, (long)h, (long)h, d_params + po.b[l]);
      }
      if (tiles)
        launch_tile<T, false, true, EPI_STORE>(st, tname<T>("gemm_f64_bwd_data", "gemm_f32_bwd_data"), R, h, 1, h,
                                               Gcur, h, Wl, h, Gnext, h, nullptr, nullptr, 0, nullptr, mlp_bf16(d));
      else
        launch_any<T, 128, 128, false, true, EPI_STORE>(st, "gemm_any", R, h, h, Gcur, h, Wl, h, Gnext, h, nullptr,
                                                        nullptr, 0, 1, h, nullptr);
      GDX_LAUNCH("hvp_combine", hvp_combine_kernel<T>, dim3(nblocks(half)), dim3(NT), 0, st, Gnext, Hin2, half);
      std::swap(Gcur, Gnext);
    }
    // first layer: [xbar; xdotbar] = G W0^T; dW0, db0 = [X16; Xdot16]^T G
    GDX_LAUNCH("layer0_bwd", (skinny_out16_kernel<T, false>), dim3(R / 64), dim3(NT), 0, st, R, h, Gcur, h, W016,
               c.P<T>(L.DX), d.mlp_nc, d.mlp_nc, nullptr, 0, nullptr, nullptr, 0, nullptr);
    if (d_params) {
      GDX_LAUNCH("skinny_wgrad", skinny_wgrad_kernel<T>, dim3(L.nwblk), dim3(NT), 0, st, R, L.wrows, X2, Gcur, h,
                 c.P<double>(L.wpart));
      reduce_slices<double>(c, c.P<double>(L.wpart), L.nwblk, 16L * h, (long)(d.mlp_nc + 1) * h, d_params + po.w[0]);
    }
  }
  if (d_rdm1) {
    GDX_LAUNCH("point_hvp_bwd", point_hvp_bwd_kernel<T>, dim3(L.npt_blocks), dim3(NT), 0, st,
               c.P<double>(L.dens_fin), c.P<double>(L.dfin_t), c.P<double>(L.GE), c.P<double>(L.GEd),
               neural ? c.P<T>(L.DX) : nullptr, d.mlp_nc, d.npts, Np, d.correlation, c.P<double>(L.vpt));
    if (c.b)
      launch_vxc_batch<T>(c, 2, d_rdm1);
    else
      launch_vxc<T>(c, ao, 2, d_rdm1, true);
  }
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

}  // namespace
}  // namespace gdx

using namespace gdx;

extern "C" {

int gdx_xc_workspace_size(const gdx_xc_desc* desc, size_t* bytes) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  GDX_REQUIRE(bytes != nullptr, "null output pointer");
  *bytes = make_layout(*desc).total;
  return GDX_OK;
}

int gdx_xc_forward(void* stream, const gdx_xc_desc* desc, const void* ao, const double* weights,
                   const double* rdm1, const double* params, double* exc, void* workspace) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  const bool neural = is_neural(*desc);
  GDX_REQUIRE(ao && weights && rdm1 && exc && workspace, "gdx_xc_forward: null pointer");
  GDX_REQUIRE(!neural || params, "gdx_xc_forward: neural functional needs params");
  Ctx c{*desc, make_layout(*desc), (char*)workspace, (hipStream_t)stream};
  if (desc->precision == GDX_FP64) return forward_impl<double>(c, (const double*)ao, weights, rdm1, params, exc);
  return forward_impl<float>(c, (const float*)ao, weights, rdm1, params, exc);
}

int gdx_xc_autotune(void* stream, const gdx_xc_desc* desc, const void* ao, void* workspace, int32_t* vt_inout) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  GDX_REQUIRE(ao && workspace, "gdx_xc_autotune: null pointer");
  const int pin = vt_inout ? *vt_inout : 0;
  GDX_REQUIRE(pin == 0 || pin == 64 || pin == 128, "gdx_xc_autotune: the tile may be pinned to 64 or 128 only");
  if (pin != 0) {   // pinned by the caller: record without measuring
    const bool nrl = is_neural(*desc);
    const int nsv_fwd = (!nrl && desc->spin_restricted) ? 1 : 2;
    std::lock_guard<std::mutex> g(vxc_tune_mu);
    vxc_tune[std::make_tuple(desc->npts_pad, desc->nao_pad, nsv_fwd, desc->precision, 0)] = pin;
    vxc_tune[std::make_tuple(desc->npts_pad, desc->nao_pad, 2, desc->precision, 1)] = pin;
    return GDX_OK;
  }
  Ctx c{*desc, make_layout(*desc), (char*)workspace, (hipStream_t)stream};
  const int Np = desc->npts_pad, nP = desc->nao_pad;
  const bool neural = is_neural(*desc);
  // the builds a plan runs: the first-order V_xc (nsv spins) and the second-order one (two
  // spins, FP32 accumulators always flushed)
  const int nsv_fwd = (!neural && desc->spin_restricted) ? 1 : 2;
  const std::pair<int, bool> builds[2] = {{nsv_fwd, false}, {2, true}};
  hipEvent_t e0, e1;
  GDX_HIP_CHECK(hipEventCreate(&e0));
  GDX_HIP_CHECK(hipEventCreate(&e1));
  // the timing launches read the potentials region: clear it first (contents are overwritten
  // by every real build)
  GDX_HIP_CHECK(hipMemsetAsync(c.P<double>(c.L.vpt), 0, (size_t)2 * Np * VPT_STRIDE * sizeof(double), c.st));
  int first = 0;
  for (int b = 0; b < 2; ++b) {
    const int nsv = builds[b].first;
    const bool ff = builds[b].second;
    const int known = tuned_vt(Np, nP, nsv, desc->precision, ff);
    if (known != 0) {   // measured before for this shape: keep it (plans of one shape agree)
      if (b == 0) first = known;
      continue;
    }
    if (desc->precision == GDX_BF16X6) {
      // split-BF16 keeps its own kernel (tiles of 128): the FP32 kernel of the 64 tiles would
      // change its accuracy, not only its speed
      const int vt = nP <= 64 ? 64 : 128;
      {
        std::lock_guard<std::mutex> g(vxc_tune_mu);
        vxc_tune[std::make_tuple(Np, nP, nsv, desc->precision, (int)ff)] = vt;
      }
      if (b == 0) first = vt;
      continue;
    }
    int best_vt = 0;
    float best_ms = 0.0f;
    for (int vt : {64, 128}) {
      const VxcSplit v = vxc_split(Np, nP, nsv, desc->precision, ff, vt);
      auto run = [&]() {
        if (desc->precision == GDX_FP64) launch_vxc_kernel<double>(c, (const double*)ao, nsv, v);
        else launch_vxc_kernel<float>(c, (const float*)ao, nsv, v);
      };
      run();   // warm-up
      GDX_HIP_CHECK(hipEventRecord(e0, c.st));
      for (int r = 0; r < 3; ++r) run();
      GDX_HIP_CHECK(hipEventRecord(e1, c.st));
      GDX_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      GDX_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (best_vt == 0 || ms < best_ms) {
        best_vt = vt;
        best_ms = ms;
      }
    }
    {
      std::lock_guard<std::mutex> g(vxc_tune_mu);
      vxc_tune[std::make_tuple(Np, nP, nsv, desc->precision, (int)ff)] = best_vt;
    }
    if (b == 0) first = best_vt;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  GDX_HIP_CHECK(hipGetLastError());
  if (vt_inout) *vt_inout = first;
  return GDX_OK;
}

int gdx_xc_densities(void* stream, const gdx_xc_desc* desc, const void* ao, const double* rdm1, double* dens,
                     void* workspace) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  GDX_REQUIRE(ao && rdm1 && dens && workspace, "gdx_xc_densities: null pointer");
  Ctx c{*desc, make_layout(*desc), (char*)workspace, (hipStream_t)stream};
  // weights == nullptr: the per-point kernel forms densities and energy densities only
  if (desc->precision == GDX_FP64) run_densities<double>(c, (const double*)ao, rdm1, nullptr, dens);
  else run_densities<float>(c, (const float*)ao, rdm1, nullptr, dens);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_xc_custom_inputs(void* stream, const gdx_xc_desc* desc, const void* ao, const double* weights,
                         const double* rdm1, void* workspace, double* e_out, double* x_out) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  GDX_REQUIRE(desc->kind == GDX_FUNC_NEURAL_B3LYP && desc->mlp_nc == 6 && desc->precision == GDX_FP64,
              "gdx_xc_custom_inputs: needs kind GDX_FUNC_NEURAL_B3LYP (the six coefficient inputs), mlp_nc 6, FP64");
  GDX_REQUIRE(ao && weights && rdm1 && workspace && e_out && x_out, "gdx_xc_custom_inputs: null pointer");
  Ctx c{*desc, make_layout(*desc), (char*)workspace, (hipStream_t)stream};
  run_densities<double>(c, (const double*)ao, rdm1, weights, c.P<double>(c.L.dens_fin));
  GDX_LAUNCH("custom_out", custom_out_kernel<double>, dim3(nblocks(desc->npts)), dim3(NT), 0, c.st, desc->npts,
             c.P<double>(c.L.E), c.P<double>(c.L.X), e_out, x_out);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_xc_custom_vxc(void* stream, const gdx_xc_desc* desc, const void* ao, const double* g_e, const double* g_x,
                      void* workspace, double* d_rdm1) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  GDX_REQUIRE(desc->kind == GDX_FUNC_NEURAL_B3LYP && desc->mlp_nc == 6 && desc->precision == GDX_FP64,
              "gdx_xc_custom_vxc: needs kind GDX_FUNC_NEURAL_B3LYP (the six coefficient inputs), mlp_nc 6, FP64");
  GDX_REQUIRE(ao && g_e && g_x && workspace && d_rdm1, "gdx_xc_custom_vxc: null pointer");
  Ctx c{*desc, make_layout(*desc), (char*)workspace, (hipStream_t)stream};
  const gdx_xc_desc& d = *desc;
  GDX_LAUNCH("custom_in", custom_in_kernel<double>, dim3(nblocks(d.npts_pad)), dim3(NT), 0, c.st, d.npts,
             d.npts_pad, d.mlp_nc, g_e, g_x, c.P<double>(c.L.GE), c.P<double>(c.L.DX));
  GDX_LAUNCH("point_bwd", point_bwd_kernel<double>, dim3(c.L.npt_blocks), dim3(NT), 0, c.st,
             c.P<double>(c.L.dens_fin), c.P<double>(c.L.GE), c.P<double>(c.L.DX), d.mlp_nc, d.npts, d.npts_pad,
             d.correlation, c.P<double>(c.L.vpt));
  launch_vxc<double>(c, (const double*)ao, 2, d_rdm1);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_xc_hvp_workspace_size(const gdx_xc_desc* desc, size_t* bytes) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  GDX_REQUIRE(bytes != nullptr, "null output pointer");
  *bytes = make_layout(*desc, true).total;
  return GDX_OK;
}

int gdx_xc_hvp(void* stream, const gdx_xc_desc* desc, const void* ao, const double* weights, const double* rdm1,
               const double* params, const double* dir, const double* g_exc, double* exc, double* edot,
               double* d_rdm1, double* d_params, void* workspace) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  const bool neural = is_neural(*desc);
  GDX_REQUIRE(ao && weights && rdm1 && dir && g_exc && workspace, "gdx_xc_hvp: null pointer");
  GDX_REQUIRE(!neural || params, "gdx_xc_hvp: neural functional needs params");
  Ctx c{*desc, make_layout(*desc, true), (char*)workspace, (hipStream_t)stream};
  if (desc->precision == GDX_FP64)
    return hvp_impl<double>(c, (const double*)ao, weights, rdm1, params, dir, g_exc, exc, edot, d_rdm1, d_params);
  return hvp_impl<float>(c, (const float*)ao, weights, rdm1, params, dir, g_exc, exc, edot, d_rdm1, d_params);
}

}  // extern "C"

namespace gdx {
namespace {

struct XcBatchPlan {
  gdx_xc_desc d{};
  Layout L{}, Lh{};
  char* ws = nullptr;    // forward / backward workspace
  char* wsh = nullptr;   // second-order workspace
  char* tab = nullptr;   // tables and batch buffers
  BatchCtx B;
  std::vector<long> dm_off;
  bool hvp = false;
};

void batch_free(XcBatchPlan* P) {
  if (!P) return;
  if (P->ws) (void)hipFree(P->ws);
  if (P->wsh) (void)hipFree(P->wsh);
  if (P->tab) (void)hipFree(P->tab);
  delete P;
}

}  // namespace
}  // namespace gdx

extern "C" {

int gdx_xc_batch_create(const gdx_xc_desc* desc, int32_t nmol, const gdx_xc_batch_mol* mols, int32_t with_hvp,
                        void** plan) {
  GDX_REQUIRE(desc && mols && plan && nmol > 0, "gdx_xc_batch_create: bad arguments");
  *plan = nullptr;
  long rows = 0;
  int maxP = 16;
  for (int m = 0; m < nmol; ++m) {
    const gdx_xc_batch_mol& M = mols[m];
    GDX_REQUIRE(M.ao && M.weights, "gdx_xc_batch_create: null AO planes or weights");
    GDX_REQUIRE(M.npts > 0 && M.npts_pad >= M.npts && M.npts_pad % 256 == 0,
                "gdx_xc_batch_create: every molecule's npts_pad must be a multiple of 256 >= npts");
    GDX_REQUIRE(M.nao > 0 && M.nao_pad >= M.nao && M.nao_pad % 16 == 0,
                "gdx_xc_batch_create: nao_pad must be a multiple of 16 >= nao");
    rows += M.npts_pad;
    maxP = std::max(maxP, (int)M.nao_pad);
  }
  GDX_REQUIRE(rows < (1L << 31), "gdx_xc_batch_create: more than 2^31 grid rows in the batch");
  XcBatchPlan* P = new XcBatchPlan();
  P->d = *desc;
  P->d.npts = (int)rows;
  P->d.npts_pad = (int)rows;
  P->d.nao = maxP;
  P->d.nao_pad = maxP;
  int rc = validate(&P->d);
  if (rc != GDX_OK) {
    delete P;
    return rc;
  }
  const gdx_xc_desc& d = P->d;
  const size_t es = elem_size(d.precision);
  const int nsv_max = 2;
  P->hvp = with_hvp != 0;
  P->L = make_layout(d, false, true);
  if (P->hvp) P->Lh = make_layout(d, true, true);
  // host tables
  std::vector<AoBatchMol> ao(nmol), ao_t(nmol);
  std::vector<PadBatchMol> pad(nmol);
  std::vector<VxcBatchMol> vx[3];
  std::vector<RedBatchMol> red(nmol);
  std::vector<int> seg256(nmol + 1), seg64(nmol + 1), rowv(nmol + 1);
  long row = 0, dm = 0, dp = 0, vp = 0;
  int tile = 0, pblk = 0, rblk = 0, vblk[3] = {0, 0, 0};
  P->dm_off.resize(nmol + 1);
  std::vector<long> vpart_off(nmol);
  for (int m = 0; m < nmol; ++m) {
    const gdx_xc_batch_mol& M = mols[m];
    const int nP = M.nao_pad, n = M.nao;
    P->dm_off[m] = dm;
    seg256[m] = (int)(row / 256);
    seg64[m] = (int)(row / 64);
    rowv[m] = (int)row;
    ao[m].ao = M.ao;
    ao[m].row0 = row;
    ao[m].npts_pad = M.npts_pad;
    ao[m].nP = nP;
    ao[m].tile0 = tile;
    ao[m].ntile = M.npts_pad / DN_PT;
    pad[m].dm_off = dm;
    pad[m].dp_off = dp;
    pad[m].n = n;
    pad[m].nP = nP;
    pad[m].blk0 = pblk;
    pad[m].nblk = (int)((2L * nP * nP + NT - 1) / NT);
    const int g = vxb_group(nP);
    const int vt = g == 1 ? 128 : 64;
    const int nt = (nP + vt - 1) / vt;
    VxcBatchMol v{};
    v.ao = M.ao;
    v.row0 = row;
    v.npts_pad = M.npts_pad;
    v.nP = nP;
    v.ntile = nt * nt;
    v.chunk = VXB_CHUNK;
    v.nchunk = (M.npts_pad + VXB_CHUNK - 1) / VXB_CHUNK;
    v.blk0 = vblk[g];
    vblk[g] += v.ntile * v.nchunk;
    vpart_off[m] = vp;
    vx[g].push_back(v);
    red[m].out_off = dm;
    red[m].nchunk = v.nchunk;
    red[m].nP = nP;
    red[m].n = n;
    red[m].blk0 = rblk;
    rblk += (int)((2L * n * n + NT - 1) / NT);
    tile += ao[m].ntile;
    pblk += pad[m].nblk;
    row += M.npts_pad;
    dm += 2L * n * n;
    dp += 2L * nP * nP;
    vp += (long)v.nchunk * nsv_max * nP * nP;
  }
  seg256[nmol] = (int)(row / 256);
  seg64[nmol] = (int)(row / 64);
  rowv[nmol] = (int)row;
  P->dm_off[nmol] = dm;
  // device memory: tables, per-point buffers, padded matrices, V partials
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += round_up(std::max<size_t>(bytes, 8), 256); return o; };
  const size_t o_ao = take(nmol * sizeof(AoBatchMol)), o_aot = take(nmol * sizeof(AoBatchMol));
  const size_t o_pad = take(nmol * sizeof(PadBatchMol));
  const size_t o_vx0 = take(vx[0].size() * sizeof(VxcBatchMol)), o_vx1 = take(vx[1].size() * sizeof(VxcBatchMol)),
               o_vx2 = take(vx[2].size() * sizeof(VxcBatchMol));
  const size_t o_red = take(nmol * sizeof(RedBatchMol));
  const size_t o_s256 = take((nmol + 1) * sizeof(int)), o_s64 = take((nmol + 1) * sizeof(int));
  const size_t o_rows = take((nmol + 1) * sizeof(int));
  const size_t o_w = take(rows * 8), o_wg = take(rows * 8), o_gpt = take(rows * 8);
  const size_t o_one = take(8), o_scal = take(nmol * 8);
  const size_t o_dp = take(dp * es), o_dpt = take(P->hvp ? dp * es : 0);
  const size_t o_vp = take(vp * 8);
  auto fail = [&](hipError_t e, const char* what) {
    set_error(std::string("gdx_xc_batch_create: ") + what + ": " + hipGetErrorString(e));
    batch_free(P);
    return GDX_ERR_HIP;
  };
  hipError_t e = hipMalloc((void**)&P->tab, off);
  if (e != hipSuccess) return fail(e, "hipMalloc");
  e = hipMalloc((void**)&P->ws, P->L.total);
  if (e != hipSuccess) return fail(e, "hipMalloc");
  if (P->hvp) {
    e = hipMalloc((void**)&P->wsh, P->Lh.total);
    if (e != hipSuccess) return fail(e, "hipMalloc");
  }
  char* T0 = P->tab;
  for (int m = 0; m < nmol; ++m) {
    ao[m].Dp = T0 + o_dp + pad[m].dp_off * es;
    ao_t[m] = ao[m];
    ao_t[m].Dp = T0 + o_dpt + pad[m].dp_off * es;
    red[m].vpart = (const double*)(T0 + o_vp) + vpart_off[m];
  }
  {
    int k[3] = {0, 0, 0};
    for (int m = 0; m < nmol; ++m) {
      const int g = vxb_group(mols[m].nao_pad);
      vx[g][k[g]++].vpart = (double*)(T0 + o_vp) + vpart_off[m];
    }
  }
  const double one = 1.0;
  struct Up { size_t off; const void* src; size_t bytes; };
  const Up ups[] = {{o_ao, ao.data(), nmol * sizeof(AoBatchMol)}, {o_aot, ao_t.data(), nmol * sizeof(AoBatchMol)},
                    {o_pad, pad.data(), nmol * sizeof(PadBatchMol)},
                    {o_vx0, vx[0].data(), vx[0].size() * sizeof(VxcBatchMol)},
                    {o_vx1, vx[1].data(), vx[1].size() * sizeof(VxcBatchMol)},
                    {o_vx2, vx[2].data(), vx[2].size() * sizeof(VxcBatchMol)},
                    {o_red, red.data(), nmol * sizeof(RedBatchMol)},
                    {o_s256, seg256.data(), (nmol + 1) * sizeof(int)}, {o_s64, seg64.data(), (nmol + 1) * sizeof(int)},
                    {o_rows, rowv.data(), (nmol + 1) * sizeof(int)}, {o_one, &one, 8}};
  for (const Up& u : ups) {
    if (!u.bytes) continue;
    e = hipMemcpy(T0 + u.off, u.src, u.bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(e, "hipMemcpy");
  }
  for (int m = 0; m < nmol; ++m) {   // concatenated weights (zero past each molecule's own rows)
    e = hipMemcpy(T0 + o_w + (size_t)rowv[m] * 8, mols[m].weights, (size_t)mols[m].npts_pad * 8,
                  hipMemcpyDeviceToDevice);
    if (e != hipSuccess) return fail(e, "hipMemcpy");
  }
  BatchCtx& B = P->B;
  B.nmol = nmol;
  B.ntile_total = tile;
  B.d_ao = (const AoBatchMol*)(T0 + o_ao);
  B.d_ao_t = (const AoBatchMol*)(T0 + o_aot);
  B.d_pad = (const PadBatchMol*)(T0 + o_pad);
  B.pad_blocks = pblk;
  B.Dp = T0 + o_dp;
  B.Dp_t = T0 + o_dpt;
  const size_t o_vx[3] = {o_vx0, o_vx1, o_vx2};
  for (int g = 0; g < 3; ++g) {
    B.d_vx[g] = (const VxcBatchMol*)(T0 + o_vx[g]);
    B.nvx[g] = (int)vx[g].size();
    B.vx_blocks[g] = vblk[g];
  }
  B.d_red = (const RedBatchMol*)(T0 + o_red);
  B.red_blocks = rblk;
  B.seg256 = (const int*)(T0 + o_s256);
  B.seg64 = (const int*)(T0 + o_s64);
  B.rows = (const int*)(T0 + o_rows);
  B.weights = (const double*)(T0 + o_w);
  B.wg = (double*)(T0 + o_wg);
  B.g_pt = (double*)(T0 + o_gpt);
  B.one = (double*)(T0 + o_one);
  B.scal = (double*)(T0 + o_scal);
  B.dm_total = dm;
  *plan = P;
  return GDX_OK;
}

int gdx_xc_batch_destroy(void* plan) {
  batch_free((XcBatchPlan*)plan);
  return GDX_OK;
}

int gdx_xc_batch_offsets(const void* plan, int64_t* dm_off) {
  GDX_REQUIRE(plan && dm_off, "gdx_xc_batch_offsets: null pointer");
  const XcBatchPlan* P = (const XcBatchPlan*)plan;
  for (size_t m = 0; m < P->dm_off.size(); ++m) dm_off[m] = P->dm_off[m];
  return GDX_OK;
}

int gdx_xc_batch_forward(void* stream, void* plan, const double* rdm1, const double* params, double* exc) {
  GDX_REQUIRE(plan && rdm1 && exc, "gdx_xc_batch_forward: null pointer");
  XcBatchPlan* P = (XcBatchPlan*)plan;
  GDX_REQUIRE(!is_neural(P->d) || params, "gdx_xc_batch_forward: neural functional needs params");
  Ctx c{P->d, P->L, P->ws, (hipStream_t)stream, &P->B};
  if (P->d.precision == GDX_FP64) return forward_impl<double>(c, nullptr, P->B.weights, rdm1, params, exc);
  return forward_impl<float>(c, nullptr, P->B.weights, rdm1, params, exc);
}

int gdx_xc_batch_backward(void* stream, void* plan, const double* params, const double* g_exc, double* d_rdm1,
                          double* d_params) {
  GDX_REQUIRE(plan && g_exc, "gdx_xc_batch_backward: null pointer");
  XcBatchPlan* P = (XcBatchPlan*)plan;
  GDX_REQUIRE(!is_neural(P->d) || params, "gdx_xc_batch_backward: neural functional needs params");
  hipStream_t st = (hipStream_t)stream;
  const BatchCtx& B = P->B;
  GDX_LAUNCH("segment_scale", segment_scale_kernel, dim3(nblocks(P->d.npts_pad)), dim3(NT), 0, st, B.rows, B.nmol,
             g_exc, B.weights, (double*)nullptr, B.wg);
  Ctx c{P->d, P->L, P->ws, st, &P->B};
  if (P->d.precision == GDX_FP64) return backward_impl<double>(c, nullptr, B.wg, params, B.one, d_rdm1, d_params);
  return backward_impl<float>(c, nullptr, B.wg, params, B.one, d_rdm1, d_params);
}

int gdx_xc_batch_hvp(void* stream, void* plan, const double* rdm1, const double* params, const double* dir,
                     const double* g_exc, double* exc, double* edot, double* d_rdm1, double* d_params) {
  GDX_REQUIRE(plan && rdm1 && dir && g_exc, "gdx_xc_batch_hvp: null pointer");
  XcBatchPlan* P = (XcBatchPlan*)plan;
  GDX_REQUIRE(P->hvp, "gdx_xc_batch_hvp: plan created without the second-order workspace");
  GDX_REQUIRE(!is_neural(P->d) || params, "gdx_xc_batch_hvp: neural functional needs params");
  hipStream_t st = (hipStream_t)stream;
  const BatchCtx& B = P->B;
  GDX_LAUNCH("segment_scale", segment_scale_kernel, dim3(nblocks(P->d.npts_pad)), dim3(NT), 0, st, B.rows, B.nmol,
             g_exc, B.weights, B.g_pt, (double*)nullptr);
  Ctx c{P->d, P->Lh, P->wsh, st, &P->B};
  if (P->d.precision == GDX_FP64)
    return hvp_impl<double>(c, nullptr, B.weights, rdm1, params, dir, B.one, exc, edot, d_rdm1, d_params);
  return hvp_impl<float>(c, nullptr, B.weights, rdm1, params, dir, B.one, exc, edot, d_rdm1, d_params);
}

int gdx_xc_backward(void* stream, const gdx_xc_desc* desc, const void* ao, const double* weights,
                    const double* rdm1, const double* params, const double* g_exc, double* d_rdm1,
                    double* d_params, void* workspace) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  const bool neural = is_neural(*desc);
  GDX_REQUIRE(ao && weights && rdm1 && g_exc && workspace, "gdx_xc_backward: null pointer");
  GDX_REQUIRE(!neural || params, "gdx_xc_backward: neural functional needs params");
  Ctx c{*desc, make_layout(*desc), (char*)workspace, (hipStream_t)stream};
  if (desc->precision == GDX_FP64)
    return backward_impl<double>(c, (const double*)ao, weights, params, g_exc, d_rdm1, d_params);
  return backward_impl<float>(c, (const float*)ao, weights, params, g_exc, d_rdm1, d_params);
}

}  // extern "C"
