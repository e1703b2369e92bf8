// The exchange-correlation grid path on gfx950: densities, energy densities, coefficient
// MLP, FP64 grid integration, and the hand-written adjoints dE/dD (V_xc) and dE/dtheta.
//
// Reference semantics followed (grad_dft/...):
//   densities        molecule.py:388-502   rho = sum D phi phi, grad rho = 2 sum D phi dphi,
//                                          tau = 1/2 sum D dphi.dphi,
//                                          lapl = 2 sum D dphi.dphi + 2 sum D phi lapl(phi)
//   energy densities popular_functionals.py:29-326 (gdx_functionals.h)
//   integration      functional.py:219-253, 316-342 (abs_clip of densities, of the
//                    per-point product and of the weights; FP64 sum)
//   dE/dD            train.py:86-121 (value_and_grad w.r.t. rdm1, unsymmetrised)
//
// Kernels (FP64 path; all matrix products on v_mfma_f64_16x16x4f64):
//   pad_dm            rdm1 (2,n,n) -> zero-padded (2,nP,nP)
//   density_f64       per 32-point tile: X_c = A_c D for c in {phi, dx, dy, dz} on MFMA,
//                     contracted row-wise against the AO planes in the epilogue
//   point_fwd         partial sums -> rho, grad rho, tau, lapl per spin; B3LYP energy
//                     densities (clipped); MLP inputs log(|x| + 1e-4) or the B3LYP sum
//   layer0            first dense layer (K = nc) + ELU on VALU
//   gemm_f64          tiled MFMA GEMM with bias+ELU / ELU-backward epilogues and split-K
//   head_fwd          last dense layer, 2 sigmoid(z/2), sum_f c_f e_f, clip, w-weighted sum
//   head_bwd          seeds of the adjoint: dE/de_f, dE/dz and the last-layer delta
//   point_bwd         dE/d(rho, |grad rho|^2, lapl) by forward-mode duals + MLP input grads
//                     -> per-point potentials
//   vxc_f64           V_s = phi^T Z_s + sum_j dphi_j^T (u_s dphi_j) on MFMA, split over
//                     point chunks, Z built while staging tiles in LDS
//   reduce_*          deterministic FP64 reductions of block / chunk partials
#include "gdx_common.h"
#include "gdx_functionals.h"
#include "gdx_profile.h"
#include "gdx_gemm.h"
#include "gdx_skinny.h"
#include "gdx_aogemm.h"

#include <algorithm>
#include <vector>

namespace gdx {
namespace {

constexpr int NT = 256;     // threads per block in every kernel here
constexpr double SQUASH = 1e-4;

// ---------------------------------------------------------------- small kernels ----

__global__ void pad_dm_kernel(const double* __restrict__ rdm1, int n, int nP, double* __restrict__ Dp) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = 2L * nP * nP;
  if (i >= tot) return;
  const int s = (int)(i / ((long)nP * nP));
  const int r = (int)((i / nP) % nP);
  const int c = (int)(i % nP);
  Dp[i] = (r < n && c < n) ? rdm1[((long)s * n + r) * n + c] : 0.0;
}

// out[i] = sum_s part[s * stride + i], i < len (deterministic order)
__global__ void reduce_partials_kernel(const double* __restrict__ part, int nsplit, long stride,
                                       long len, double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double acc = 0.0;
  for (int s = 0; s < nsplit; ++s) acc += part[(long)s * stride + i];
  out[i] = acc;
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

// V partials (nchunk, 2, nP, nP) -> d_rdm1 (2, n, n)
// nsv = spins computed (1: both spin blocks equal, spin 0 copied)
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

__device__ __forceinline__ double elu(double x) { return x > 0.0 ? x : expm1(x); }
__device__ __forceinline__ double elu_grad_from_out(double y) { return y > 0.0 ? 1.0 : y + 1.0; }
__device__ __forceinline__ double clip30(double x) { return fabs(x) > 1e-30 ? x : 0.0; }

// ---------------------------------------------------------------- per-point fwd ----
//
// dens_fin (npts_pad, 2, 6): rho, grad rho (3), tau, lapl per spin (final)
// E (npts_pad, 4): clipped energy densities; X (npts_pad, 6): MLP inputs (neural only)
__global__ __launch_bounds__(NT) void point_fwd_kernel(
    const double* __restrict__ dens_part, int ncb, int nsd, int npts, int npts_pad,
    const double* __restrict__ weights, int neural, double* __restrict__ dens_fin,
    double* __restrict__ E, double* __restrict__ X, double* __restrict__ eps_part) {
  __shared__ double red[NT / 64];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  if (r < npts_pad) {
    double d[2][DENS_STRIDE];
    for (int s = 0; s < 2; ++s) {
      const int ds = nsd == 1 ? 0 : s;
      double raw[DENS_STRIDE] = {0, 0, 0, 0, 0, 0};
      for (int cb = 0; cb < ncb; ++cb) {
        const double* src = dens_part + (((size_t)cb * nsd + ds) * npts_pad + r) * DENS_STRIDE;
#pragma unroll
        for (int q = 0; q < DENS_STRIDE; ++q) raw[q] += src[q];
      }
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
    b3lyp_parts<double>(d[0][0], d[1][0], gu, gd, d[0][5], d[1][5], e);
#pragma unroll
    for (int f = 0; f < 4; ++f) E[(size_t)r * 4 + f] = e[f];
    if (neural) {
      // X16 row: six squashed features, a 1 (bias column of the first layer's gradient), zeros
      const double feat[6] = {d[0][0], d[1][0], gu, gd, d[0][4], d[1][4]};
#pragma unroll
      for (int j = 0; j < 6; ++j) X[(size_t)r * 16 + j] = log(fabs(feat[j]) + SQUASH);
      X[(size_t)r * 16 + 6] = 1.0;
#pragma unroll
      for (int j = 7; j < 16; ++j) X[(size_t)r * 16 + j] = 0.0;
    } else if (weights != nullptr && r < npts) {
      const double c[4] = {1.0 - 0.2, 0.72, 1.0 - 0.81, 0.81};
      double eps = c[0] * e[0] + c[1] * e[1] + c[2] * e[2] + c[3] * e[3];
      eps = clip30(clip30(eps));
      contrib = clip30(weights[r]) * eps;
    }
  }
  const double t = block_sum(contrib, red);
  if (!neural && threadIdx.x == 0) eps_part[blockIdx.x] = t;
}

// H (npts_pad, h) = ELU(X (npts_pad, nc) W0 (nc, h) + b0)
__global__ __launch_bounds__(NT) void layer0_kernel(const double* __restrict__ X, int nc, int h,
                                                    long total, const double* __restrict__ W,
                                                    const double* __restrict__ b,
                                                    double* __restrict__ H) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long r = i / h;
  const int f = (int)(i % h);
  double z = b[f];
  for (int j = 0; j < nc; ++j) z += X[r * 16 + j] * W[j * h + f];
  H[i] = elu(z);
}

// ------------------------------------------------------------------ GEMM (f64) ----
constexpr int GBK = 16;

// C (+ z * M * ldc for split-K) = epi(op(A) op(B)); op(A)[m][k] = TA ? A[k*lda+m] : A[m*lda+k],
// op(B)[k][n] = TB ? B[n*ldb+k] : B[k*ldb+n].  colsum (optional, split-K): column sums of
// op(B) over this block's k range, written by blockIdx.x == 0 to colsum[z * N + n].
template <int BM, int BN, bool TA, bool TB, int EPI>
__global__ __launch_bounds__(NT) void gemm_f64_kernel(
    int M, int N, int K, const double* __restrict__ A, int lda, const double* __restrict__ B,
    int ldb, double* __restrict__ C, int ldc, const double* __restrict__ bias,
    const double* __restrict__ Y, int ldy, int kchunk, double* __restrict__ colsum) {
  constexpr int WM = BM / 2, WN = BN / 2, MB = WM / 16, NB = WN / 16;
  __shared__ double As[GBK][BM + 2];
  __shared__ double Bs[GBK][BN + 2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  double csum = 0.0;

  d4 acc[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

  for (int k0 = kbeg; k0 < kend; k0 += GBK) {
    for (int e = tid; e < BM * GBK; e += NT) {
      int m, k;
      if (TA) { k = e / BM; m = e % BM; } else { m = e / GBK; k = e % GBK; }
      const int gm = m0 + m, gk = k0 + k;
      double v = 0.0;
      if (gm < M && gk < kend) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      As[k][m] = v;
    }
    for (int e = tid; e < BN * GBK; e += NT) {
      int n, k;
      if (TB) { n = e / GBK; k = e % GBK; } else { k = e / BN; n = e % BN; }
      const int gn = n0 + n, gk = k0 + k;
      double v = 0.0;
      if (gn < N && gk < kend) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
      Bs[k][n] = v;
    }
    __syncthreads();
    if (colsum != nullptr && blockIdx.x == 0 && tid < BN) {
#pragma unroll
      for (int k = 0; k < GBK; ++k) csum += Bs[k][tid];
    }
#pragma unroll
    for (int kk = 0; kk < GBK / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      double af[MB], bf[NB];
#pragma unroll
      for (int i = 0; i < MB; ++i) af[i] = As[kr][wm + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < NB; ++j) bf[j] = Bs[kr][wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  if (colsum != nullptr && blockIdx.x == 0 && tid < BN && n0 + tid < N)
    colsum[(size_t)blockIdx.z * N + n0 + tid] = csum;

  double* Cz = C + (size_t)blockIdx.z * M * ldc;
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int col = n0 + wn + j * 16 + (lane & 15);
      if (col >= N) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + i * 16 + mfma64_row(lane, q);
        if (row >= M) continue;
        double v = acc[i][j][q];
        if (EPI == EPI_BIAS_ELU) v = elu(v + bias[col]);
        if (EPI == EPI_ELU_BWD) v *= elu_grad_from_out(Y[(size_t)row * ldy + col]);
        Cz[(size_t)row * ldc + col] = v;
      }
    }
}

// ----------------------------------------------------------------- MLP head ----
constexpr int MAXND = 8;

// Padded weight tiles for the skinny kernels: WT16 (16 x h) rows o < nd = W_last[:, o];
// W016 (16 x h) rows j < nc = W0[j, :]; every other row zero.
__global__ void prep_edge_weights_kernel(const double* __restrict__ W0, int nc,
                                         const double* __restrict__ Wl, int nd, int h,
                                         double* __restrict__ W016, double* __restrict__ WT16) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 16 * h) return;
  const int row = i / h, f = i % h;
  W016[i] = row < nc ? W0[row * h + f] : 0.0;
  WT16[i] = row < nd ? Wl[f * nd + row] : 0.0;
}

// Adjoint seeds, one thread per point.  g = g_exc clip(w) [|eps| > 1e-30];
// GE = dE/de_f = g c_f; neural: DZ16 = dE/dz = g e_f (c_f/2)(1 - c_f/2) zero-padded to 16,
// and block partials of the last layer's bias gradient sum_r DZ[r][o] -> dbpart[block][o].
__global__ __launch_bounds__(NT) void head_seed_kernel(
    const double* __restrict__ Cc, const double* __restrict__ E, const double* __restrict__ weights,
    int nd, int npts, int npts_pad, const double* __restrict__ g_exc, int neural,
    double* __restrict__ GE, double* __restrict__ DZ16, double* __restrict__ dbpart) {
  __shared__ double red[NT / 64];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const double bc[4] = {1.0 - 0.2, 0.72, 1.0 - 0.81, 0.81};
  double dz[MAXND];
  for (int o = 0; o < MAXND; ++o) dz[o] = 0.0;
  if (r < npts_pad) {
    double c[MAXND], e[MAXND];
    double eps = 0.0;
    // padded rows (r >= npts) have no coefficients stored: they take c = 0, seed 0
    for (int o = 0; o < nd; ++o) {
      c[o] = r >= npts ? 0.0 : (neural ? Cc[(size_t)r * nd + o] : bc[o]);
      e[o] = E[(size_t)r * nd + o];
      eps += c[o] * e[o];
    }
    const double g = (r < npts && fabs(eps) > 1e-30) ? (*g_exc) * clip30(weights[r]) : 0.0;
    for (int o = 0; o < nd; ++o) {
      GE[(size_t)r * nd + o] = g * c[o];
      dz[o] = g * e[o] * (c[o] * 0.5) * (1.0 - c[o] * 0.5);
    }
    if (neural)
      for (int o = 0; o < 16; ++o) DZ16[(size_t)r * 16 + o] = o < nd ? dz[o] : 0.0;
  }
  if (neural)
    for (int o = 0; o < nd; ++o) {
      const double t = block_sum(dz[o], red);
      if (threadIdx.x == 0) dbpart[(size_t)blockIdx.x * nd + o] = t;
    }
}

// G[r][f] = (sum_o DZ16[r][o] W_last[f][o]) * elu'(H[r][f]), one thread per element
__global__ __launch_bounds__(NT) void head_delta_kernel(const double* __restrict__ DZ16,
                                                        const double* __restrict__ Wl, int nd, int h,
                                                        const double* __restrict__ H, long total,
                                                        double* __restrict__ G) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long r = i / h;
  const int f = (int)(i % h);
  double v = 0.0;
  for (int o = 0; o < nd; ++o) v += DZ16[r * 16 + o] * Wl[f * nd + o];
  G[i] = v * elu_grad_from_out(H[i]);
}

// d_params last layer: W (h, nd) from the reduced (16 x h) transpose, b from block partials
__global__ void unpack_last_kernel(const double* __restrict__ dWt16, int h, int nd,
                                   double* __restrict__ dW) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= h * nd) return;
  const int f = i / nd, o = i % nd;
  dW[i] = dWt16[(size_t)o * h + f];
}

// Per-point potentials for the V_xc build from dE/de_f (GE) and dE/dx (DX, neural).
__global__ __launch_bounds__(NT) void point_bwd_kernel(
    const double* __restrict__ dens_fin, const double* __restrict__ GE,
    const double* __restrict__ DX, const double* __restrict__ X, int npts, int npts_pad,
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
                    dvar<6>(d0[5], 4), dvar<6>(d1[5], 5), e);
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
        df[j] = DX[(size_t)r * 6 + j] * sg / (fabs(feat[j]) + SQUASH);
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

// ------------------------------------------------------------ workspace layout ----
struct Layout {
  size_t Dp, dens_part, dens_fin, E, X, C, H, eps_part, GE, DZ, G, DX, vpt, vpart, wpart, cpart;
  size_t W16, dbpart, dWt16;
  int nwblk, wrows;
  size_t total;
  int ncb, wcb, nsd, nchunk, chunk, nsplit, kchunk, nparam, npt_blocks;
};

constexpr int VCHUNK = 512;   // points per V_xc split
constexpr int SKW_ROWS = 512;  // grid points per block of the skinny weight-gradient kernel

Layout make_layout(const gdx_xc_desc& d) {
  Layout L{};
  const size_t Np = d.npts_pad, nP = d.nao_pad;
  const bool neural = d.kind == GDX_FUNC_NEURAL_B3LYP;
  L.wcb = 1;
  L.ncb = 1;
  L.nsd = d.spin_restricted ? 1 : 2;
  L.chunk = VCHUNK;
  L.nchunk = (int)((Np + VCHUNK - 1) / VCHUNK);
  // split-K for the weight-gradient GEMMs over grid points: whole 16-row stages per split
  L.nsplit = 1;
  while (L.nsplit < 128 && (Np / GT_K) % (2 * L.nsplit) == 0) L.nsplit *= 2;
  L.kchunk = (int)(Np / L.nsplit);
  L.npt_blocks = (int)((Np + NT - 1) / NT);
  const size_t h = neural ? d.mlp_h : 0, nh = neural ? d.mlp_nhidden : 0;
  size_t np = 0;
  if (neural) {
    int din = d.mlp_nc;
    for (int l = 0; l <= d.mlp_nhidden; ++l) {
      const int dout = l == d.mlp_nhidden ? d.mlp_nd : d.mlp_h;
      np += (size_t)din * dout + dout;
      din = dout;
    }
  }
  L.nparam = (int)np;
  size_t off = 0;
  auto take = [&](size_t ndoubles) { size_t o = off; off += round_up(ndoubles * 8, 256); return o; };
  L.Dp = take(2 * nP * nP);
  L.dens_part = take((size_t)L.ncb * L.nsd * Np * DENS_STRIDE);
  L.dens_fin = take(Np * 2 * DENS_STRIDE);
  L.E = take(Np * 4);
  L.X = take(neural ? Np * 16 : 0);
  L.C = take(neural ? Np * d.mlp_nd : 0);
  L.H = take(Np * h * nh);
  L.eps_part = take(std::max<size_t>(Np / 64 + 1, 4096));
  L.GE = take(Np * 4);
  L.DZ = take(neural ? Np * 16 : 0);
  L.W16 = take(neural ? 2 * 16 * (size_t)d.mlp_h : 0);
  L.dbpart = take(neural ? (size_t)L.npt_blocks * d.mlp_nd : 0);
  L.dWt16 = take(neural ? 16 * (size_t)d.mlp_h : 0);
  L.G = take(neural ? 2 * Np * h : 0);
  L.DX = take(neural ? Np * d.mlp_nc : 0);
  L.vpt = take(2 * Np * VPT_STRIDE);
  L.vpart = take((size_t)L.nchunk * 2 * nP * nP);
  // skinny weight-gradient splits: row chunks of SKW_ROWS points
  L.wrows = SKW_ROWS;
  L.nwblk = (int)((Np + SKW_ROWS - 1) / SKW_ROWS);
  size_t wmax = 0;
  if (neural) {
    wmax = (size_t)L.nsplit * d.mlp_h * d.mlp_h;
    wmax = std::max(wmax, (size_t)L.nwblk * 16 * d.mlp_h);
  }
  L.wpart = take(wmax);
  L.cpart = take(neural ? (size_t)L.nsplit * d.mlp_h : 0);
  L.total = off;
  return L;
}

int validate(const gdx_xc_desc* d) {
  GDX_REQUIRE(d != nullptr, "null gdx_xc_desc");
  GDX_REQUIRE(d->npts > 0 && d->nao > 0, "empty grid or basis");
  GDX_REQUIRE(d->npts_pad >= d->npts && d->npts_pad % 64 == 0, "npts_pad must be a multiple of 64 >= npts");
  GDX_REQUIRE(d->nao_pad >= d->nao && d->nao_pad % 16 == 0, "nao_pad must be a multiple of 16 >= nao");
  GDX_REQUIRE(d->kind == GDX_FUNC_B3LYP || d->kind == GDX_FUNC_NEURAL_B3LYP, "unknown functional kind");
  if (d->precision != GDX_FP64) {
    set_error("this build implements the FP64 grid path only");
    return GDX_ERR_UNSUPPORTED;
  }
  if (d->kind == GDX_FUNC_NEURAL_B3LYP) {
    GDX_REQUIRE(d->mlp_nc == 6, "coefficient inputs must be 6 (rho, |grad rho|^2, tau per spin)");
    GDX_REQUIRE(d->mlp_nd == 4, "MLP outputs must equal the 4 B3LYP energy densities");
    GDX_REQUIRE(d->mlp_h >= 16 && d->mlp_h % 16 == 0 && d->mlp_h <= 256,
                "MLP width must be a multiple of 16 in [16, 256]");
    GDX_REQUIRE(d->mlp_nhidden >= 1 && d->mlp_nhidden <= 16, "MLP hidden layers must be 1..16");
  }
  return GDX_OK;
}

inline unsigned nblocks(long n) { return (unsigned)((n + NT - 1) / NT); }
inline unsigned nblocks64(long n) { return (unsigned)((n + 63) / 64); }

template <int BM, int BN, bool TA, bool TB, int EPI>
void launch_gemm(hipStream_t st, int M, int N, int K, const double* A, int lda, const double* B,
                 int ldb, double* C, int ldc, const double* bias, const double* Y, int ldy,
                 int nsplit, int kchunk, double* colsum) {
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, nsplit);
  const char* name = TA ? (M == N ? "gemm_f64_wgrad" : "gemm_f64_wgrad_edge")
                        : (EPI == EPI_BIAS_ELU ? "gemm_f64_fwd" : "gemm_f64_bwd_data");
  GDX_LAUNCH(name, (gemm_f64_kernel<BM, BN, TA, TB, EPI>), grid, dim3(NT), 0, st, M, N, K, A, lda,
             B, ldb, C, ldc, bias, Y, ldy, kchunk, colsum);
}

template <typename T, bool TA, bool TB, int EPI>
void launch_tile(hipStream_t st, const char* name, int M, int N, int nsplit, int kchunk, const T* A,
                 int lda, const T* B, int ldb, T* C, int ldc, const T* bias, const T* Y, int ldy,
                 T* colsum) {
  dim3 grid((M / GT_M) * (N / GT_N), nsplit);
  GDX_LAUNCH(name, (gemm_tile_kernel<T, TA, TB, EPI>), grid, dim3(NT), 0, st, M, N, A, lda, B, ldb, C,
             ldc, bias, Y, ldy, kchunk, colsum);
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

int gdx_xc_forward(void* stream, const gdx_xc_desc* desc, const double* ao, const double* weights,
                   const double* rdm1, const double* params, double* exc, void* workspace) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  const gdx_xc_desc& d = *desc;
  const bool neural = d.kind == GDX_FUNC_NEURAL_B3LYP;
  GDX_REQUIRE(ao && weights && rdm1 && exc && workspace, "gdx_xc_forward: null pointer");
  GDX_REQUIRE(!neural || params, "gdx_xc_forward: neural functional needs params");
  hipStream_t st = (hipStream_t)stream;
  const Layout L = make_layout(d);
  char* ws = (char*)workspace;
  auto P = [&](size_t off) { return (double*)(ws + off); };
  const int Np = d.npts_pad, nP = d.nao_pad;

  GDX_LAUNCH("pad_dm", pad_dm_kernel, dim3(nblocks(2L * nP * nP)), dim3(NT), 0, st, rdm1, d.nao, nP, P(L.Dp));
  GDX_LAUNCH("density_f64", density_ao_kernel<double>, dim3(Np / DN_PT, L.nsd), dim3(NT), 0, st, ao, Np, nP,
             P(L.Dp), P(L.dens_part));
  GDX_LAUNCH("point_fwd", point_fwd_kernel, dim3(L.npt_blocks), dim3(NT), 0, st, P(L.dens_part), L.ncb,
                     L.nsd, d.npts, Np, weights, neural ? 1 : 0, P(L.dens_fin), P(L.E), P(L.X),
                     P(L.eps_part));
  int nred = L.npt_blocks;
  if (neural) {
    const int h = d.mlp_h;
    const bool tiles = Np % GT_M == 0 && h % GT_N == 0;
    const double* W = params;
    const double* b = W + (size_t)d.mlp_nc * h;
    double* H = P(L.H);
    GDX_LAUNCH("layer0", layer0_kernel, dim3(nblocks((long)Np * h)), dim3(NT), 0, st, P(L.X), d.mlp_nc, h,
                       (long)Np * h, W, b, H);
    const double* pw = b + h;
    for (int l = 1; l < d.mlp_nhidden; ++l) {
      const double* Wl = pw;
      const double* bl = Wl + (size_t)h * h;
      if (tiles)
        launch_tile<double, false, false, EPI_BIAS_ELU>(st, "gemm_f64_fwd", Np, h, 1, h,
                                                        H + (size_t)(l - 1) * Np * h, h, Wl, h,
                                                        H + (size_t)l * Np * h, h, bl, nullptr, 0, nullptr);
      else
        launch_gemm<128, 128, false, false, EPI_BIAS_ELU>(st, Np, h, h, H + (size_t)(l - 1) * Np * h, h,
                                                          Wl, h, H + (size_t)l * Np * h, h, bl, nullptr,
                                                          0, 1, h, nullptr);
      pw = bl + h;
    }
    const double* Wlast = pw;
    const double* blast = Wlast + (size_t)h * d.mlp_nd;
    double* W016 = P(L.W16);
    double* WT16 = W016 + 16 * h;
    GDX_LAUNCH("prep_edge_weights", prep_edge_weights_kernel, dim3(nblocks(16L * h)), dim3(NT), 0, st, W, d.mlp_nc,
               Wlast, d.mlp_nd, h, W016, WT16);
    GDX_LAUNCH("head_fwd", (skinny_out16_kernel<double, true>), dim3(Np / 64), dim3(NT), 0, st, Np, h,
               H + (size_t)(d.mlp_nhidden - 1) * Np * h, h, WT16, P(L.C), d.mlp_nd, d.mlp_nd, blast, d.mlp_nd,
               P(L.E), weights, d.npts, P(L.eps_part));
    nred = Np / 64;
  }
  GDX_LAUNCH("reduce_scalar", reduce_scalar_kernel, dim3(1), dim3(NT), 0, st, P(L.eps_part), nred, exc);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_xc_densities(void* stream, const gdx_xc_desc* desc, const double* ao, const double* rdm1,
                     double* dens, void* workspace) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  const gdx_xc_desc& d = *desc;
  GDX_REQUIRE(ao && rdm1 && dens && workspace, "gdx_xc_densities: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const Layout L = make_layout(d);
  char* ws = (char*)workspace;
  auto P = [&](size_t off) { return (double*)(ws + off); };
  const int Np = d.npts_pad, nP = d.nao_pad;
  GDX_LAUNCH("pad_dm", pad_dm_kernel, dim3(nblocks(2L * nP * nP)), dim3(NT), 0, st, rdm1, d.nao, nP, P(L.Dp));
  GDX_LAUNCH("density_f64", density_ao_kernel<double>, dim3(Np / DN_PT, L.nsd), dim3(NT), 0, st, ao, Np, nP,
             P(L.Dp), P(L.dens_part));
  // weights == nullptr: the per-point kernel forms densities and energy densities only
  const bool neural = d.kind == GDX_FUNC_NEURAL_B3LYP;
  GDX_LAUNCH("point_fwd", point_fwd_kernel, dim3(L.npt_blocks), dim3(NT), 0, st, P(L.dens_part), L.ncb,
                     L.nsd, d.npts, Np, nullptr, neural ? 1 : 0, dens, P(L.E), P(L.X), P(L.eps_part));
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_xc_backward(void* stream, const gdx_xc_desc* desc, const double* ao,
                    const double* weights, const double* rdm1, const double* params,
                    const double* g_exc, double* d_rdm1, double* d_params, void* workspace) {
  int rc = validate(desc);
  if (rc != GDX_OK) return rc;
  const gdx_xc_desc& d = *desc;
  const bool neural = d.kind == GDX_FUNC_NEURAL_B3LYP;
  GDX_REQUIRE(ao && weights && rdm1 && g_exc && workspace, "gdx_xc_backward: null pointer");
  GDX_REQUIRE(!neural || params, "gdx_xc_backward: neural functional needs params");
  hipStream_t st = (hipStream_t)stream;
  const Layout L = make_layout(d);
  char* ws = (char*)workspace;
  auto P = [&](size_t off) { return (double*)(ws + off); };
  const int Np = d.npts_pad, nP = d.nao_pad;
  const int h = neural ? d.mlp_h : 0;
  const int nd = 4;

  const double* Wlast = nullptr;
  if (neural) {
    // parameter offsets
    std::vector<size_t> woff, boff;
    size_t off = 0;
    int din = d.mlp_nc;
    for (int l = 0; l <= d.mlp_nhidden; ++l) {
      const int dout = l == d.mlp_nhidden ? d.mlp_nd : h;
      woff.push_back(off); off += (size_t)din * dout;
      boff.push_back(off); off += dout;
      din = dout;
    }
    Wlast = params + woff.back();
    const bool tiles = Np % GT_M == 0 && h % GT_N == 0;
    double* H = P(L.H);
    double* G0 = P(L.G);
    double* G1 = P(L.G) + (size_t)Np * h;
    const double* Hlast = H + (size_t)(d.mlp_nhidden - 1) * Np * h;
    const double* W016 = P(L.W16);
    const int nl = d.mlp_nhidden;
    GDX_LAUNCH("head_seed", head_seed_kernel, dim3(L.npt_blocks), dim3(NT), 0, st, P(L.C), P(L.E), weights, nd,
               d.npts, Np, g_exc, 1, P(L.GE), P(L.DZ), P(L.dbpart));
    GDX_LAUNCH("head_delta", head_delta_kernel, dim3(nblocks((long)Np * h)), dim3(NT), 0, st, P(L.DZ), Wlast, nd,
               h, Hlast, (long)Np * h, G0);
    if (d_params) {
      GDX_LAUNCH("skinny_wgrad", skinny_wgrad_kernel<double>, dim3(L.nwblk), dim3(NT), 0, st, Np, L.wrows, P(L.DZ),
                 Hlast, h, P(L.wpart));
      GDX_LAUNCH("reduce_cols", reduce_cols_kernel, dim3(nblocks64(16L * h)), dim3(NT), 0, st, P(L.wpart),
                 L.nwblk, 16L * h, 16L * h, P(L.dWt16));
      GDX_LAUNCH("unpack_last", unpack_last_kernel, dim3(nblocks((long)h * nd)), dim3(NT), 0, st, P(L.dWt16), h,
                 nd, d_params + woff[nl]);
      GDX_LAUNCH("reduce_cols", reduce_cols_kernel, dim3(1), dim3(NT), 0, st, P(L.dbpart), L.npt_blocks,
                 (long)nd, (long)nd, d_params + boff[nl]);
    }
    // hidden h x h layers, from the top
    double* Gcur = G0;
    double* Gnext = G1;
    for (int l = nl - 1; l >= 1; --l) {
      const double* Hin = H + (size_t)(l - 1) * Np * h;  // input activations of layer l
      const double* Wl = params + woff[l];
      if (d_params) {
        if (tiles)
          launch_tile<double, true, false, EPI_STORE>(st, "gemm_f64_wgrad", h, h, L.nsplit, L.kchunk, Hin, h,
                                                      Gcur, h, P(L.wpart), h, nullptr, nullptr, 0, P(L.cpart));
        else
          launch_gemm<64, 64, true, false, EPI_STORE>(st, h, h, Np, Hin, h, Gcur, h, P(L.wpart), h, nullptr,
                                                      nullptr, 0, L.nsplit, L.kchunk, P(L.cpart));
        GDX_LAUNCH("reduce_cols", reduce_cols_kernel, dim3(nblocks64((long)h * h)), dim3(NT), 0, st,
                   P(L.wpart), L.nsplit, (long)h * h, (long)h * h, d_params + woff[l]);
        GDX_LAUNCH("reduce_cols", reduce_cols_kernel, dim3(nblocks64(h)), dim3(NT), 0, st, P(L.cpart),
                   L.nsplit, (long)h, (long)h, d_params + boff[l]);
      }
      // delta for layer l's input: (Gcur W_l^T) * elu'(Hin)
      if (tiles)
        launch_tile<double, false, true, EPI_ELU_BWD>(st, "gemm_f64_bwd_data", Np, h, 1, h, Gcur, h, Wl, h, Gnext,
                                                      h, nullptr, Hin, h, nullptr);
      else
        launch_gemm<128, 128, false, true, EPI_ELU_BWD>(st, Np, h, h, Gcur, h, Wl, h, Gnext, h, nullptr, Hin,
                                                        h, 1, h, nullptr);
      std::swap(Gcur, Gnext);
    }
    // first layer: DX = G W0^T (skinny), dW0 and db0 = X16^T G (rows 0..nc-1 and the ones row nc)
    GDX_LAUNCH("layer0_bwd", (skinny_out16_kernel<double, false>), dim3(Np / 64), dim3(NT), 0, st, Np, h, Gcur,
               h, W016, P(L.DX), d.mlp_nc, d.mlp_nc, nullptr, 0, nullptr, nullptr, 0, nullptr);
    if (d_params) {
      GDX_LAUNCH("skinny_wgrad", skinny_wgrad_kernel<double>, dim3(L.nwblk), dim3(NT), 0, st, Np, L.wrows, P(L.X),
                 Gcur, h, P(L.wpart));
      // rows 0..nc of the 16 x h result are W0 (nc x h) followed by b0: contiguous in params
      GDX_LAUNCH("reduce_cols", reduce_cols_kernel, dim3(nblocks64((long)(d.mlp_nc + 1) * h)), dim3(NT), 0, st,
                 P(L.wpart), L.nwblk, 16L * h, (long)(d.mlp_nc + 1) * h, d_params + woff[0]);
    }
  } else {
    GDX_LAUNCH("head_seed", head_seed_kernel, dim3(L.npt_blocks), dim3(NT), 0, st, nullptr, P(L.E), weights, nd,
               d.npts, Np, g_exc, 0, P(L.GE), nullptr, nullptr);
  }
  if (d_rdm1) {
    GDX_LAUNCH("point_bwd", point_bwd_kernel, dim3(L.npt_blocks), dim3(NT), 0, st, P(L.dens_fin), P(L.GE),
                       neural ? P(L.DX) : nullptr, P(L.X), d.npts, Np, P(L.vpt));
    const int nt = (nP + VX_T - 1) / VX_T;
    // B3LYP on equal spin densities has equal spin potentials: form V once
    const int nsv = (!neural && d.spin_restricted) ? 1 : 2;
    if (nt == 1)
      GDX_LAUNCH("vxc_f64", (vxc_ao_kernel<double, true>), dim3(1, L.nchunk, nsv), dim3(NT), 0, st, ao, Np, nP,
                 P(L.vpt), L.chunk, P(L.vpart));
    else
      GDX_LAUNCH("vxc_f64", (vxc_ao_kernel<double, false>), dim3(nt * nt, L.nchunk, nsv), dim3(NT), 0, st, ao, Np,
                 nP, P(L.vpt), L.chunk, P(L.vpart));
    GDX_LAUNCH("reduce_vxc", reduce_vxc_kernel, dim3(nblocks(2L * d.nao * d.nao)), dim3(NT), 0, st, P(L.vpart),
               L.nchunk, nsv, nP, d.nao, d_rdm1);
  }
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

}  // extern "C"
