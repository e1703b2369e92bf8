// Tiled MFMA GEMM for the coefficient-MLP layers: C = epi(op(A) op(B)), FP64 or FP32.
//
// op(A)[m][k] = TA ? A[k*lda + m] : A[m*lda + k];  op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n]
// Block 256 threads = 4 waves (2 x 2), tile 128 x 128 x 16, each wave 64 x 64 = 4 x 4 MFMA
// 16x16 accumulators.  Global tiles are fetched with 16-byte vector loads into registers
// one stage ahead and stored, in their global orientation, into a double-buffered LDS tile
// (16-byte stores), so HBM latency overlaps the MFMA work of the current stage (one barrier
// per stage).  Shapes must be whole tiles:
// M % 128 == N % 128 == 0 and the K range of every split a multiple of 16 (callers pad).
// Blocks are renumbered so consecutive tiles (which share the same A rows) are dispatched to
// the same XCD and its L2.
// Split-K: split y (see xcd_tile_split) covers the K range split_range gives; C is offset by y*M*ldc.
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

// EPI_ELU_BWD_W0: EPI_ELU_BWD whose result G also feeds the first-layer weight gradient:
// per wave, the MFMA sums sum_r G[r][f] X16[r][c] over its 64 rows into
// w0part[slice][c][f] (c < ncp, slice = 2 (row tile) + row half), in the C layout, so G is
// contracted over rows without another pass over it.
enum { EPI_STORE = 0, EPI_BIAS_ELU = 1, EPI_ELU_BWD = 2, EPI_ELU_BWD_W0 = 3 };

__device__ __forceinline__ double elu_d(double x) { return x > 0.0 ? x : expm1(x); }
__device__ __forceinline__ float elu_f(float x) { return x > 0.0f ? x : expm1f(x); }
__device__ __forceinline__ double elu_any(double x) { return elu_d(x); }
__device__ __forceinline__ float elu_any(float x) { return elu_f(x); }

constexpr int GT_M = 128, GT_N = 128;
// k per pipeline stage (16: keeps three FP32 blocks resident per CU; 32 measured slower)
template <typename T> struct GemmK { static constexpr int value = 16; };

// Tile b and K-split of this block, with blocks sharing operands placed on one XCD (the
// hardware dispatches linear block L to XCD L % 8, and each XCD has its own L2):
//  * split-K grids (gridDim.y splits, a multiple of 8): the tiles of one split -- which read
//    the same K-range rows of both operands -- go to the same XCD, consecutively;
//  * otherwise, when the tile count is a multiple of 8, consecutive tiles (same A rows) go
//    to the same XCD.
__device__ __forceinline__ void xcd_tile_split(int nblk, int& b, int& ysplit) {
  b = blockIdx.x;
  ysplit = blockIdx.y;
  if (gridDim.y > 1 && gridDim.y % 8 == 0) {
    const int L = blockIdx.x + blockIdx.y * gridDim.x;
    const int xcd = L % 8, slot = L / 8;
    b = slot % gridDim.x;
    ysplit = xcd + 8 * (slot / gridDim.x);
  } else if (gridDim.y == 1 && nblk % 8 == 0) {
    b = (b % 8) * (nblk / 8) + b / 8;
  }
}

// K range of split y: the splits are sps per segment of kseg rows (segments in order), each
// kchunk rows long except the last of a segment, which stops at the segment's end (splits
// past it are empty).  One segment of nsplit * kchunk rows is plain split-K; two segments
// keep the value and tangent rows of the second-order pass in separate splits.
__device__ __forceinline__ void split_range(int y, int kchunk, int kseg, int sps, int& kbeg, int& kend) {
  const int seg = y / sps, js = y - seg * sps;
  kbeg = seg * kseg + js * kchunk;
  kend = min(kbeg + kchunk, (seg + 1) * kseg);
  if (kend < kbeg) kend = kbeg;
}

// Blocks per CU: FP32 3 (2 with the ELU-backward epilogues); FP64 2 -- its 128 x 128 FP64
// accumulators then live in VGPRs (256, no spills; 75-82 KB LDS a block), 28% faster for the
// forward GEMM than one block with AGPR accumulators.
template <typename T, bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? (EPI >= EPI_ELU_BWD ? 2 : 3) : 2) void gemm_tile_kernel(
    int M, int N, const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb,
    T* __restrict__ C, int ldc, const T* __restrict__ bias, const T* __restrict__ Y, int ldy,
    int kchunk, T* __restrict__ colsum, int kseg, int sps, const T* __restrict__ X16 = nullptr,
    double* __restrict__ w0part = nullptr, int ncp = 0) {
  typedef typename MfmaT<T>::acc_t acc_t;
  constexpr int GT_K = GemmK<T>::value;
  constexpr int VEC = 16 / sizeof(T);                // elements per 16-byte load
  constexpr int NVA = GT_M * GT_K / VEC / 256;       // vector loads per thread per tile
  constexpr int NVB = GT_N * GT_K / VEC / 256;
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  typedef T v4_t __attribute__((ext_vector_type(4)));
  // LDS tiles keep the global orientation (16-byte stores, no transposition):
  //   A: TA ? [k][m] (row GT_M + 4) : [m][k] (row GT_K + 4);  B: TB ? [n][k] : [k][n]
  constexpr int AKM = GT_M + 4, AMK = GT_K + 4, BKN = GT_N + 4, BNK = GT_K + 4;
  constexpr int ABUF = TA ? GT_K * AKM : GT_M * AMK;
  constexpr int BBUF = TB ? GT_N * BNK : GT_K * BKN;
  __shared__ __align__(16) T As[2 * ABUF];
  __shared__ __align__(16) T Bs[2 * BBUF];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int nbn = N / GT_N, nblk = (M / GT_M) * nbn;
  int b, ysplit;
  xcd_tile_split(nblk, b, ysplit);
  const int m0 = (b / nbn) * GT_M, n0 = (b % nbn) * GT_N;
  int kbeg, kend;
  split_range(ysplit, kchunk, kseg, sps, kbeg, kend);
  const int nk = (kend - kbeg) / GT_K;

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
        *(vec_t*)&As[buf * ABUF + k * AKM + c * VEC] = ra[i];
      } else {
        const int r = idx / (GT_K / VEC), c = idx % (GT_K / VEC);
        *(vec_t*)&As[buf * ABUF + r * AMK + c * VEC] = ra[i];
      }
    }
#pragma unroll
    for (int i = 0; i < NVB; ++i) {
      const int idx = tid + 256 * i;
      if (!TB) {
        const int k = idx / (GT_N / VEC), c = idx % (GT_N / VEC);
        *(vec_t*)&Bs[buf * BBUF + k * BKN + c * VEC] = rb[i];
      } else {
        const int r = idx / (GT_K / VEC), c = idx % (GT_K / VEC);
        *(vec_t*)&Bs[buf * BBUF + r * BNK + c * VEC] = rb[i];
      }
    }
  };
  // in each 16-k half of the stage, k-group lane / 16 supplies k = 4 g + kk to MFMA step kk
  // (a permutation of those 16 k, the same for both operands), so a [m][k] or [n][k] tile is
  // read with one 4-element vector per fragment
  auto afrag = [&](int buf, int m, int g, T* out) {
    if (TA) {
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) out[kk] = As[buf * ABUF + (4 * g + kk) * AKM + m];
    } else {
      const v4_t v = *(const v4_t*)&As[buf * ABUF + m * AMK + 4 * g];
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) out[kk] = v[kk];
    }
  };
  auto bfrag = [&](int buf, int n, int g, T* out) {
    if (TB) {
      const v4_t v = *(const v4_t*)&Bs[buf * BBUF + n * BNK + 4 * g];
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) out[kk] = v[kk];
    } else {
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) out[kk] = Bs[buf * BBUF + (4 * g + kk) * BKN + n];
    }
  };

  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (acc_t){0, 0, 0, 0};
  T csum = 0;
  constexpr bool ELUB = EPI == EPI_ELU_BWD || EPI == EPI_ELU_BWD_W0, W0 = EPI == EPI_ELU_BWD_W0;
  acc_t wacc[W0 ? 4 : 1];
#pragma unroll
  for (int j = 0; j < (W0 ? 4 : 1); ++j) wacc[j] = (acc_t){0, 0, 0, 0};
  // ELU backward, FP32: this lane's 64 elements of Y are fetched two stages before the
  // epilogue, so their latency overlaps the last MFMAs instead of following them (FP64 keeps
  // the loads in the epilogue: 64 more doubles a lane would spill)
  constexpr bool YPF = ELUB && sizeof(T) == 4;
  T yv[YPF ? 4 : 1][YPF ? 4 : 1][4];
  auto yload = [&]() {
    if constexpr (YPF) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int q = 0; q < 4; ++q)
            yv[i][j][q] = Y[(size_t)(m0 + wm + i * 16 + MfmaT<T>::row(lane, q)) * ldy + n0 + wn + j * 16 + (lane & 15)];
    }
  };
  const int ykt = nk > 2 ? nk - 2 : 0;

  if (nk > 0) {
    gload(kbeg);
    lstore(0);
    __syncthreads();
  } else {
    yload();
  }
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload(kbeg + (kt + 1) * GT_K);
    if (kt == ykt) yload();
    if (colsum != nullptr && m0 == 0 && tid < GT_N && !TB) {
#pragma unroll
      for (int k = 0; k < GT_K; ++k) csum += Bs[cur * BBUF + k * BKN + tid];
    }
#pragma unroll
    for (int kb = 0; kb < GT_K / 16; ++kb) {
      const int g = kb * 4 + (lane >> 4), r = lane & 15;
      T af[4][4], bf[4][4];
#pragma unroll
      for (int i = 0; i < 4; ++i) afrag(cur, wm + i * 16 + r, g, af[i]);
#pragma unroll
      for (int j = 0; j < 4; ++j) bfrag(cur, wn + j * 16 + r, g, bf[j]);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = MfmaT<T>::op(af[i][kk], bf[j][kk], acc[i][j]);
    }
    if (kt + 1 < nk) lstore(cur ^ 1);
    __syncthreads();
  }
  if (colsum != nullptr && m0 == 0 && tid < GT_N) colsum[(size_t)ysplit * N + n0 + tid] = csum;

  T* Cz = C + (size_t)ysplit * M * ldc;
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
        if constexpr (ELUB) {
          const T y = YPF ? yv[YPF ? i : 0][YPF ? j : 0][q] : Y[(size_t)row * ldy + col];
          v *= (y > (T)0 ? (T)1 : y + (T)1);
        }
        if constexpr (W0) {
          // A = G^T (f = lane % 16, row slot lane / 16), B = X16 rows of the same slots
          const T xr = X16[(size_t)row * 16 + (lane & 15)];
          wacc[W0 ? j : 0] = MfmaT<T>::op(v, xr, wacc[W0 ? j : 0]);
        }
        Cz[(size_t)row * ldc + col] = v;
      }
    }
  if constexpr (W0) {
    const int slice = (m0 / GT_M) * 2 + (wm >> 6);
    const int c = lane & 15;
    if (c < ncp) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int f = n0 + wn + j * 16 + MfmaT<T>::row(lane, q);
          w0part[((size_t)slice * ncp + c) * N + f] = (double)wacc[W0 ? j : 0][q];
        }
    }
  }
}

}  // namespace gdx
