// MFMA kernels for the MLP's narrow edges (input width nc = 6, output width nd = 4, both
// zero-padded to one 16-wide MFMA tile) and a parallel column reduction of split partials.
//
//   skinny_out16   C16 (M x 16) = A (M x K) . Bt^T, Bt (16 x K): per-row products against
//                  at most 16 weight rows (the last layer's z = H W_last, and dE/dx = G W0^T).
//                  Block = 4 waves x 16 rows; A staged through LDS in 64-wide K chunks.
//                  Optional head epilogue: c = head_coef(z + b), eps = sum_o c_o e_o,
//                  clipped, weighted FP64 block sum (grad_dft/functional.py:219-253, 407-419).
//   skinny_wgrad   P[blk] (16 x N) = sum over the block's rows r of A16[r][m] B[r][n]: weight
//                  gradients whose fan-in or fan-out is <= 16, split over row chunks; the 4
//                  waves of a block take interleaved 4-row MFMA steps and are summed in LDS.
//   reduce_cols    out[i] = sum_s part[s * stride + i], 64 columns x 4 split groups per block.
#pragma once
#include "gdx_common.h"
#include "gdx_gemm.h"

namespace gdx {

constexpr int SK_KC = 64;  // K chunk staged per LDS pass

template <typename T, bool HEAD>
__global__ __launch_bounds__(256) void skinny_out16_kernel(
    int M, int K, const T* __restrict__ A, int lda, const T* __restrict__ Bt,
    T* __restrict__ C16, int ldc, int ncols,
    // head epilogue
    const T* __restrict__ bias, int nd, const double* __restrict__ E,
    const double* __restrict__ weights, int npts, double* __restrict__ eps_part, int head_mode = 0) {
  typedef typename MfmaT<T>::acc_t acc_t;
  __shared__ T As[64][SK_KC + 2];
  __shared__ double zs[64][17];
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r0 = blockIdx.x * 64;
  acc_t acc = (acc_t){0, 0, 0, 0};
  constexpr int VEC = 16 / sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  const bool whole = (K % SK_KC == 0) && (r0 + 64 <= M) && (lda % VEC == 0);
  for (int k0 = 0; k0 < K; k0 += SK_KC) {
    // 64 rows x 64 k staged in LDS
    if (whole) {
#pragma unroll
      for (int i = 0; i < 64 * SK_KC / VEC / 256; ++i) {
        const int idx = tid + 256 * i;
        const int r = idx / (SK_KC / VEC), cv = idx % (SK_KC / VEC);
        const vec_t v = *(const vec_t*)(A + (size_t)(r0 + r) * lda + k0 + cv * VEC);
#pragma unroll
        for (int e = 0; e < VEC; ++e) As[r][cv * VEC + e] = v[e];
      }
    } else {
      for (int e = tid; e < 64 * SK_KC; e += 256) {
        const int r = e / SK_KC, k = e % SK_KC;
        As[r][k] = (r0 + r < M && k0 + k < K) ? A[(size_t)(r0 + r) * lda + k0 + k] : (T)0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < SK_KC / 4; ++kk) {
      const int k = kk * 4 + (lane >> 4);
      const T af = As[wave * 16 + (lane & 15)][k];
      const T bf = (k0 + k < K) ? Bt[(size_t)(lane & 15) * K + k0 + k] : (T)0;
      acc = MfmaT<T>::op(af, bf, acc);
    }
    __syncthreads();
  }
  if (!HEAD) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + wave * 16 + MfmaT<T>::row(lane, q);
      const int c = lane & 15;
      if (r < M && c < ncols) C16[(size_t)r * ldc + c] = acc[q];
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) zs[wave * 16 + MfmaT<T>::row(lane, q)][lane & 15] = (double)acc[q];
  __syncthreads();
  double contrib = 0.0;
  if (tid < 64) {
    const int r = r0 + tid;
    if (r < npts) {
      double eps = 0.0;
      for (int o = 0; o < nd; ++o) {
        const double z = zs[tid][o] + (double)bias[o];
        const double c = head_coef(z, head_mode, o);
        C16[(size_t)r * ldc + o] = (T)c;
        eps += c * E[(size_t)r * nd + o];
      }
      eps = fabs(eps) > 1e-30 ? eps : 0.0;
      const double w = weights[r];
      contrib = (fabs(w) > 1e-30 ? w : 0.0) * eps;
    }
  }
  contrib = warp_sum(contrib);
  if (tid == 0) eps_part[blockIdx.x] = contrib;
  (void)red;
}

// P[blockIdx.x] (16 x N) = sum over rows r in [blockIdx.x*rows, +rows) of A16[r][m] B[r][n].
// N <= 256 and a multiple of 16; wave w owns columns [64 w, 64 w + 64), so the four waves
// never share an output.  32-row stages of A16 and B go through double-buffered LDS with
// 16-byte global loads prefetched one stage ahead.
constexpr int SW_KS = 32;

template <typename T>
__global__ __launch_bounds__(256) void skinny_wgrad_kernel(
    int M, int rows, const T* __restrict__ A16, const T* __restrict__ B, int N,
    double* __restrict__ part) {
  typedef typename MfmaT<T>::acc_t acc_t;
  constexpr int VEC = 16 / sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  constexpr int SLOTS = 256 / VEC;               // 16-byte vectors per 256-column row
  constexpr int NVB = SW_KS * SLOTS / 256;       // B vectors per thread per stage
  constexpr int NVA = (SW_KS * 16 / VEC + 255) / 256;
  __shared__ T As[2][SW_KS][16 + 1];
  __shared__ T Bs[2][SW_KS][256 + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rbeg = blockIdx.x * rows;
  const int rend = min(M, rbeg + rows);
  const int nvrow = N / VEC;  // vectors per B row
  acc_t acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = (acc_t){0, 0, 0, 0};
  vec_t rb[NVB], ra[NVA];
  auto gload = [&](int r0) {
#pragma unroll
    for (int i = 0; i < NVB; ++i) {
      const int idx = tid + 256 * i;
      const int r = idx / SLOTS, cv = idx % SLOTS;   // vector slots of a 256-column row
      const int rr = r0 + r;
      rb[i] = (rr < rend && cv < nvrow) ? *(const vec_t*)(B + (size_t)rr * N + cv * VEC) : (vec_t)(T)0;
    }
#pragma unroll
    for (int i = 0; i < NVA; ++i) {
      const int idx = tid + 256 * i;
      const int r = idx / (16 / VEC), cv = idx % (16 / VEC);
      const int rr = r0 + r;
      if (idx < SW_KS * 16 / VEC)
        ra[i] = rr < rend ? *(const vec_t*)(A16 + (size_t)rr * 16 + cv * VEC) : (vec_t)(T)0;
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NVB; ++i) {
      const int idx = tid + 256 * i;
      const int r = idx / SLOTS, cv = idx % SLOTS;
#pragma unroll
      for (int e = 0; e < VEC; ++e) Bs[buf][r][cv * VEC + e] = rb[i][e];
    }
#pragma unroll
    for (int i = 0; i < NVA; ++i) {
      const int idx = tid + 256 * i;
      const int r = idx / (16 / VEC), cv = idx % (16 / VEC);
      if (idx < SW_KS * 16 / VEC)
#pragma unroll
        for (int e = 0; e < VEC; ++e) As[buf][r][cv * VEC + e] = ra[i][e];
    }
  };
  const int nst = (rend - rbeg + SW_KS - 1) / SW_KS;
  if (nst > 0) {
    gload(rbeg);
    lstore(0);
    __syncthreads();
  }
  for (int st = 0; st < nst; ++st) {
    const int cur = st & 1;
    if (st + 1 < nst) gload(rbeg + (st + 1) * SW_KS);
#pragma unroll
    for (int kk = 0; kk < SW_KS / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      const T af = As[cur][kr][lane & 15];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const T bf = Bs[cur][kr][wave * 64 + j * 16 + (lane & 15)];
        acc[j] = MfmaT<T>::op(af, bf, acc[j]);
      }
    }
    if (st + 1 < nst) lstore(cur ^ 1);
    __syncthreads();
  }
  double* out = part + (size_t)blockIdx.x * 16 * N;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n = wave * 64 + j * 16 + (lane & 15);
    if (n < N)
#pragma unroll
      for (int q = 0; q < 4; ++q) out[(size_t)MfmaT<T>::row(lane, q) * N + n] = (double)acc[j][q];
  }
}

// out[i] = sum over the nsplit slices of part[s * stride + i], 64 consecutive i per block
// (block bx), 4 slice subgroups combined in LDS in a fixed order.
template <typename TP>
__device__ __forceinline__ void reduce_cols_block(const TP* __restrict__ part, int nsplit, long stride, long len,
                                                  double* __restrict__ out, int bx) {
  __shared__ double red[4][64];
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const long i = (long)bx * 64 + c;
  double acc = 0.0;
  if (i < len) {
    // 8 independent loads in flight per thread; the sum order stays fixed
    int s = g;
    for (; s + 28 < nsplit; s += 32) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = (double)part[(long)(s + 4 * u) * stride + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; s < nsplit; s += 4) acc += (double)part[(long)s * stride + i];
  }
  red[g][c] = acc;
  __syncthreads();
  if (g == 0 && i < len) out[i] = red[0][c] + red[1][c] + red[2][c] + red[3][c];
}

template <typename TP>
__global__ __launch_bounds__(256) void reduce_cols_kernel(const TP* __restrict__ part, int nsplit,
                                                          long stride, long len,
                                                          double* __restrict__ out) {
  reduce_cols_block<TP>(part, nsplit, stride, len, out, blockIdx.x);
}

// two independent column reductions in one launch: blocks [0, nb0) serve the first
template <typename TP>
__global__ __launch_bounds__(256) void reduce_cols2_kernel(const TP* __restrict__ p0, int ns0, long st0, long len0,
                                                           double* __restrict__ o0, int nb0,
                                                           const TP* __restrict__ p1, int ns1, long st1, long len1,
                                                           double* __restrict__ o1) {
  if ((int)blockIdx.x < nb0) reduce_cols_block<TP>(p0, ns0, st0, len0, o0, blockIdx.x);
  else reduce_cols_block<TP>(p1, ns1, st1, len1, o1, blockIdx.x - nb0);
}

}  // namespace gdx
