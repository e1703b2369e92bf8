// Tiled MFMA GEMM for the coefficient-MLP layers: C = epi(op(A) op(B)), FP64 or FP32.
//
// op(A)[m][k] = TA ? A[k*lda + m] : A[m*lda + k];  op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n]
// Block 256 threads = 4 waves (2 x 2), tile 128 x 128 x 16, each wave 64 x 64 = 4 x 4 MFMA
// 16x16 accumulators.  Global tiles are fetched with 16-byte vector loads into registers
// one stage ahead and stored into a double-buffered LDS tile, so HBM latency overlaps the
// MFMA work of the current stage (one barrier per stage).  Shapes must be whole tiles:
// M % 128 == N % 128 == 0 and the K range of every split a multiple of 16 (callers pad).
// Blocks are renumbered so consecutive tiles (which share the same A rows) are dispatched to
// the same XCD and its L2.
// Split-K: blockIdx.y selects K range [y*kchunk, (y+1)*kchunk); C is offset by y*M*ldc.
// colsum (optional): per split, the column sums of op(B) over its K range (bias gradients).
#pragma once
#include "gdx_common.h"

namespace gdx {

template <typename T> struct MfmaT;
template <> struct MfmaT<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int q) { return mfma64_row(lane, q); }
};
template <> struct MfmaT<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int q) { return mfma32_row(lane, q); }
};

enum { EPI_STORE = 0, EPI_BIAS_ELU = 1, EPI_ELU_BWD = 2 };

__device__ __forceinline__ double elu_d(double x) { return x > 0.0 ? x : expm1(x); }
__device__ __forceinline__ float elu_f(float x) { return x > 0.0f ? x : expm1f(x); }
__device__ __forceinline__ double elu_any(double x) { return elu_d(x); }
__device__ __forceinline__ float elu_any(float x) { return elu_f(x); }

constexpr int GT_M = 128, GT_N = 128, GT_K = 16;

template <typename T, bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_tile_kernel(
    int M, int N, const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb,
    T* __restrict__ C, int ldc, const T* __restrict__ bias, const T* __restrict__ Y, int ldy,
    int kchunk, T* __restrict__ colsum) {
  typedef typename MfmaT<T>::acc_t acc_t;
  constexpr int VEC = 16 / sizeof(T);                // elements per 16-byte load
  constexpr int NVA = GT_M * GT_K / VEC / 256;       // vector loads per thread per tile
  constexpr int NVB = GT_N * GT_K / VEC / 256;
  constexpr int PAD = 4;
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  __shared__ T As[2][GT_K][GT_M + PAD];
  __shared__ T Bs[2][GT_K][GT_N + PAD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  // XCD-aware renumbering: hardware dispatches block b to XCD b % 8
  const int nbn = N / GT_N, nblk = (M / GT_M) * nbn;
  int b = blockIdx.x;
  if (nblk % 8 == 0) b = (b % 8) * (nblk / 8) + b / 8;
  const int m0 = (b / nbn) * GT_M, n0 = (b % nbn) * GT_N;
  const int kbeg = blockIdx.y * kchunk;
  const int nk = kchunk / GT_K;

  vec_t ra[NVA], rb[NVB];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NVA; ++i) {
      const int idx = tid + 256 * i;
      if (TA) {
        const int k = idx / (GT_M / VEC), c = idx % (GT_M / VEC);
        ra[i] = *(const vec_t*)(A + (size_t)(k0 + k) * lda + m0 + c * VEC);
      } else {
        const int r = idx / (GT_K / VEC), c = idx % (GT_K / VEC);
        ra[i] = *(const vec_t*)(A + (size_t)(m0 + r) * lda + k0 + c * VEC);
      }
    }
#pragma unroll
    for (int i = 0; i < NVB; ++i) {
      const int idx = tid + 256 * i;
      if (!TB) {
        const int k = idx / (GT_N / VEC), c = idx % (GT_N / VEC);
        rb[i] = *(const vec_t*)(B + (size_t)(k0 + k) * ldb + n0 + c * VEC);
      } else {
        const int r = idx / (GT_K / VEC), c = idx % (GT_K / VEC);
        rb[i] = *(const vec_t*)(B + (size_t)(n0 + r) * ldb + k0 + c * VEC);
      }
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NVA; ++i) {
      const int idx = tid + 256 * i;
      if (TA) {
        const int k = idx / (GT_M / VEC), c = idx % (GT_M / VEC);
#pragma unroll
        for (int e = 0; e < VEC; ++e) As[buf][k][c * VEC + e] = ra[i][e];
      } else {
        const int r = idx / (GT_K / VEC), c = idx % (GT_K / VEC);
#pragma unroll
        for (int e = 0; e < VEC; ++e) As[buf][c * VEC + e][r] = ra[i][e];
      }
    }
#pragma unroll
    for (int i = 0; i < NVB; ++i) {
      const int idx = tid + 256 * i;
      if (!TB) {
        const int k = idx / (GT_N / VEC), c = idx % (GT_N / VEC);
#pragma unroll
        for (int e = 0; e < VEC; ++e) Bs[buf][k][c * VEC + e] = rb[i][e];
      } else {
        const int r = idx / (GT_K / VEC), c = idx % (GT_K / VEC);
#pragma unroll
        for (int e = 0; e < VEC; ++e) Bs[buf][c * VEC + e][r] = rb[i][e];
      }
    }
  };

  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (acc_t){0, 0, 0, 0};
  T csum = 0;

  if (nk > 0) {
    gload(kbeg);
    lstore(0);
    __syncthreads();
  }
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload(kbeg + (kt + 1) * GT_K);
    if (colsum != nullptr && m0 == 0 && tid < GT_N) {
#pragma unroll
      for (int k = 0; k < GT_K; ++k) csum += Bs[cur][k][tid];
    }
#pragma unroll
    for (int kk = 0; kk < GT_K / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      T af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = As[cur][kr][wm + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Bs[cur][kr][wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = MfmaT<T>::op(af[i], bf[j], acc[i][j]);
    }
    if (kt + 1 < nk) lstore(cur ^ 1);
    __syncthreads();
  }
  if (colsum != nullptr && m0 == 0 && tid < GT_N) colsum[(size_t)blockIdx.y * N + n0 + tid] = csum;

  T* Cz = C + (size_t)blockIdx.y * M * ldc;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn + j * 16 + (lane & 15);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + i * 16 + MfmaT<T>::row(lane, q);
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

}  // namespace gdx
