// Streaming kernels for the MLP's narrow edges (fan-in nc = 6 or fan-out nd = 4 against a
// hidden width h): HBM-bound passes over an (N x h) activation, one read per pass.
//
// A row of h values is spread over LPR = h / 4 lanes, 4 consecutive features per lane
// (one 16- or 32-byte load); a wave holds 64 / LPR rows at a time.  Dot products over the
// row are finished with xor shuffles inside each LPR-lane group.  Weight-gradient sums over
// a block's rows accumulate per lane and go out as per-block partials (reduced afterwards in
// FP64, fixed order).
//
//   delta_wgrad_rows_kernel G = (DZ W_last^T) * elu'(H) and dW_last partials sum_r DZ^T H
//   wgrad0_rows_kernel      [dW0; db0] partials sum_r [X; 1]^T G
// and, with each lane loading 16 bytes of one row and the row dot products done on MFMA
// (no cross-lane reductions):
//   rowdot_mfma_kernel      HEAD: z = H W_last + b, c = head_coef(z), eps = sum_o c_o e_o,
//                           clipped weighted E_xc block partial (grad_dft/functional.py:219-253,
//                           407-419); otherwise out = A Bt^T for at most 16 output columns
#pragma once
#include "gdx_common.h"
#include "gdx_gemm.h"

namespace gdx {

constexpr int RW_ROWS = 256;  // rows per block of the delta / weight-gradient row kernels
constexpr int RD_ROWS = 64;   // rows per block of rowdot_mfma_kernel (one 16-row group a wave)
constexpr int RW_UNROLL = 8;  // rows per lane loaded before any is consumed (FP32; FP64 half)

template <typename T>
struct Vec4T;
template <> struct Vec4T<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Vec4T<double> { typedef double type __attribute__((ext_vector_type(4))); };

// sum over the aligned group of `width` lanes (width a power of two <= 64)
template <typename V>
__device__ __forceinline__ V group_sum(V v, int width) {
  for (int off = width >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// G[r][f] = (sum_o DZ16[r][o] WT16[o][f]) * elu'(H[r][f]);  part[blk][o][f] = sum_r DZ16[r][o] H[r][f]
template <typename T>
__global__ __launch_bounds__(256) void delta_wgrad_rows_kernel(const T* __restrict__ H, int h, int rows,
                                                               const T* __restrict__ DZ16,
                                                               const T* __restrict__ WT16, int nd,
                                                               T* __restrict__ G, double* __restrict__ part) {
  typedef typename Vec4T<T>::type v4;
  __shared__ double red[4][4][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lpr = h >> 2, rpw = 64 / lpr;
  const int sub = lane / lpr, li = lane % lpr;
  const int f0 = li * 4;
  T w[4][4];
#pragma unroll
  for (int o = 0; o < 4; ++o)
#pragma unroll
    for (int e = 0; e < 4; ++e) w[o][e] = o < nd ? WT16[(size_t)o * h + f0 + e] : (T)0;
  T acc[4][4];
#pragma unroll
  for (int o = 0; o < 4; ++o)
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[o][e] = (T)0;
  const int r0 = blockIdx.x * RW_ROWS + wave * (RW_ROWS / 4);
  constexpr int U = sizeof(T) == 8 ? RW_UNROLL / 2 : RW_UNROLL;
  for (int rb = sub; rb < RW_ROWS / 4; rb += U * rpw) {
    v4 y[U], dzv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + rb + u * rpw;
      const bool ok = rb + u * rpw < RW_ROWS / 4 && r < rows;
      y[u] = ok ? *(const v4*)(H + (size_t)r * h + f0) : (v4)(T)0;
      dzv[u] = ok ? *(const v4*)(DZ16 + (size_t)r * 16) : (v4)(T)0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + rb + u * rpw;
      v4 g;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const T v = dzv[u][0] * w[0][e] + dzv[u][1] * w[1][e] + dzv[u][2] * w[2][e] + dzv[u][3] * w[3][e];
        g[e] = v * (y[u][e] > (T)0 ? (T)1 : y[u][e] + (T)1);
#pragma unroll
        for (int o = 0; o < 4; ++o) acc[o][e] += dzv[u][o] * y[u][e];
      }
      if (rb + u * rpw < RW_ROWS / 4 && r < rows) *(v4*)(G + (size_t)r * h + f0) = g;
    }
  }
  // sum the rpw row groups of each wave, then the 4 waves (fixed order)
#pragma unroll
  for (int o = 0; o < 4; ++o)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double v = (double)acc[o][e];
      for (int off = lpr; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
      if (sub == 0) red[wave][o][f0 + e] = v;
    }
  __syncthreads();
  for (int i = threadIdx.x; i < nd * h; i += 256) {
    const int o = i / h, f = i % h;
    part[(size_t)blockIdx.x * nd * h + i] = red[0][o][f] + red[1][o][f] + red[2][o][f] + red[3][o][f];
  }
}

// out (rows x nout) = A (rows x K) Bt^T, Bt (16 x K) in LDS; one 16-row group per wave.  Lane
// l loads VEC = 16 / sizeof(T) consecutive elements of row l % 16 at column
// t (4 VEC) + (l / 16) VEC for every t; element q of chunk t is the A operand of MFMA
// (t, q), whose k-slice l / 16 then stands for that column -- B is read with the same
// column map, so the products sum over every k exactly once.  The row loads are issued
// before Bt is staged (the two latencies overlap), and Bt's LDS rows are padded by 16 bytes
// so the 16 rows' B vectors fall in different banks.
template <typename T, bool HEAD>
__global__ __launch_bounds__(256) void rowdot_mfma_kernel(int rows, int K, const T* __restrict__ A,
                                                          const T* __restrict__ Bt, T* __restrict__ out, int nout,
                                                          const T* __restrict__ bias, const double* __restrict__ E,
                                                          const double* __restrict__ weights, int npts,
                                                          double* __restrict__ eps_part, int head_mode) {
  static_assert(RD_ROWS == 64, "one 16-row group per wave of a 256-thread block");
  typedef typename MfmaT<T>::acc_t acc_t;
  constexpr int VEC = 16 / sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  constexpr int MAXC = 256 / (4 * VEC);  // chunks for K <= 256
  constexpr int LDB = 256 + VEC;
  __shared__ __align__(16) T Bs[16 * LDB];
  __shared__ double zs[4][16][17];
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nch = K / (4 * VEC);
  const int rl = lane & 15, ks = lane >> 4;
  const int r0 = blockIdx.x * RD_ROWS + wave * 16;
  const int r = r0 + rl;
  vec_t x[MAXC];
#pragma unroll
  for (int t = 0; t < MAXC; ++t)
    if (t < nch) x[t] = r < rows ? *(const vec_t*)(A + (size_t)r * K + t * 4 * VEC + ks * VEC) : (vec_t)(T)0;
  {
    // stage Bt (16 x K, K <= 256) with 16-byte loads, all issued before the LDS stores
    constexpr int NBV = 16 * 256 / (256 * VEC);
    vec_t bv[NBV];
#pragma unroll
    for (int i = 0; i < NBV; ++i) {
      const int e = (tid + 256 * i) * VEC;
      bv[i] = e < 16 * K ? *(const vec_t*)(Bt + e) : (vec_t)(T)0;
    }
#pragma unroll
    for (int i = 0; i < NBV; ++i) {
      const int e = (tid + 256 * i) * VEC;
      if (e < 16 * K) *(vec_t*)&Bs[(e / K) * LDB + e % K] = bv[i];
    }
  }
  __syncthreads();
  double contrib = 0.0;
  if (r0 < rows) {
    acc_t acc = (acc_t){0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < MAXC; ++t) {
      if (t < nch) {
        const vec_t b = *(const vec_t*)&Bs[rl * LDB + t * 4 * VEC + ks * VEC];
#pragma unroll
        for (int q = 0; q < VEC; ++q) acc = MfmaT<T>::op(x[t][q], b[q], acc);
      }
    }
    if (!HEAD) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rr = r0 + MfmaT<T>::row(lane, q);
        if (rr < rows && rl < nout) out[(size_t)rr * nout + rl] = acc[q];
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) zs[wave][MfmaT<T>::row(lane, q)][rl] = (double)acc[q];
      // zs is private to this wave: lanes 0..15 finish one row each (wave-scope ordering of
      // the LDS writes above before the reads below)
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (nout <= 4) {
        // all 64 lanes: row lane / 4, output lane % 4; eps of a row summed over its four lanes
        // as (c0 e0 + c1 e1) + (c2 e2 + c3 e3)
        const int hr = lane >> 2, ho = lane & 3;
        const int rr = r0 + hr;
        double prod = 0.0;
        if (rr < npts && ho < nout) {
          const double zz = zs[wave][hr][ho] + (double)bias[ho];
          const double c = head_coef(zz, head_mode, ho);
          out[(size_t)rr * nout + ho] = (T)c;
          prod = c * E[(size_t)rr * nout + ho];
        }
        const double pr = prod + __shfl_xor(prod, 1, 64);
        double eps = pr + __shfl_xor(pr, 2, 64);
        if (ho == 0 && rr < npts) {
          eps = fabs(eps) > 1e-30 ? eps : 0.0;
          const double wt = weights[rr];
          contrib += (fabs(wt) > 1e-30 ? wt : 0.0) * eps;
        }
      } else if (lane < 16) {
        const int rr = r0 + lane;
        if (rr < npts) {
          double eps = 0.0;
          for (int o = 0; o < nout; ++o) {
            const double zz = zs[wave][lane][o] + (double)bias[o];
            const double c = head_coef(zz, head_mode, o);
            out[(size_t)rr * nout + o] = (T)c;
            eps += c * E[(size_t)rr * nout + o];
          }
          eps = fabs(eps) > 1e-30 ? eps : 0.0;
          const double wt = weights[rr];
          contrib += (fabs(wt) > 1e-30 ? wt : 0.0) * eps;
        }
      }
    }
  }
  if (HEAD) {
    contrib = warp_sum(contrib);
    if (lane == 0) red[wave] = contrib;
    __syncthreads();
    if (tid == 0) eps_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
  }
}

// part[blk][j][f] = sum_r X16[r][j] G[r][f] for j <= nc (X16's column nc is the ones column)
template <typename T>
__global__ __launch_bounds__(256) void wgrad0_rows_kernel(const T* __restrict__ G, int h, int rows,
                                                          const T* __restrict__ X16, int nc,
                                                          double* __restrict__ part) {
  typedef typename Vec4T<T>::type v4;
  constexpr int NJ = 7;
  __shared__ double red[4][NJ][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lpr = h >> 2, rpw = 64 / lpr;
  const int sub = lane / lpr, li = lane % lpr;
  const int f0 = li * 4;
  T acc[NJ][4];
#pragma unroll
  for (int j = 0; j < NJ; ++j)
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[j][e] = (T)0;
  const int r0 = blockIdx.x * RW_ROWS + wave * (RW_ROWS / 4);
  constexpr int U = sizeof(T) == 8 ? RW_UNROLL / 2 : RW_UNROLL;
  for (int rb = sub; rb < RW_ROWS / 4; rb += U * rpw) {
    v4 g[U], xa[U], xb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + rb + u * rpw;
      const bool ok = rb + u * rpw < RW_ROWS / 4 && r < rows;
      g[u] = ok ? *(const v4*)(G + (size_t)r * h + f0) : (v4)(T)0;
      xa[u] = ok ? *(const v4*)(X16 + (size_t)r * 16) : (v4)(T)0;
      xb[u] = ok ? *(const v4*)(X16 + (size_t)r * 16 + 4) : (v4)(T)0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T x[NJ] = {xa[u][0], xa[u][1], xa[u][2], xa[u][3], xb[u][0], xb[u][1], xb[u][2]};
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[j][e] += x[j] * g[u][e];
    }
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double v = (double)acc[j][e];
      for (int off = lpr; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
      if (sub == 0) red[wave][j][f0 + e] = v;
    }
  __syncthreads();
  for (int i = threadIdx.x; i < (nc + 1) * h; i += 256) {
    const int j = i / h, f = i % h;
    part[(size_t)blockIdx.x * (nc + 1) * h + i] = red[0][j][f] + red[1][j][f] + red[2][j][f] + red[3][j][f];
  }
}

}  // namespace gdx
