// Split-BF16 GEMM for the coefficient-MLP layers: FP32 operands in HBM, FP32 accumulation,
// products on v_mfma_f32_16x16x32_bf16.
//
// Each FP32 value x is split while it is staged into LDS into three bf16 terms
//   h = bf16(x), m = bf16(x - h), l = bf16(x - h - m)      (|x - h - m - l| <~ 2^-27 |x|)
// and a product a.b is formed from the six terms whose order is <= 2^-16:
//   a_h b_h + a_h b_m + a_m b_h + a_h b_l + a_m b_m + a_l b_h.
// This keeps FP32-level accuracy (a two-term split with three products does not: on the
// benchmark MLP it moves E_xc by ~5e-6 Ha).
//
// C = epi(op(A) op(B)) with the conventions of gemm_tile_kernel (gdx_gemm.h); tile
// 128 x 128 x 32, 4 waves of 64 x 64, LDS tiles stored k-contiguous ([m][k], [n][k]) so a
// lane's 8-element bf16 fragment is one 16-byte read; global tiles are prefetched one stage
// ahead into registers and double-buffered in LDS.
#pragma once
#include "gdx_common.h"
#include "gdx_gemm.h"

namespace gdx {

constexpr int GB_M = 128, GB_N = 128, GB_K = 32, GB_KP = GB_K + 8;

__device__ __forceinline__ void split3_bf16(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  const float r1 = x - (float)h;
  m = (__bf16)r1;
  const float r2 = r1 - (float)m;
  l = (__bf16)r2;
}

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16x6_kernel(
    int M, int N, const float* __restrict__ A, int lda, const float* __restrict__ B, int ldb,
    float* __restrict__ C, int ldc, const float* __restrict__ bias, const float* __restrict__ Y, int ldy,
    int kchunk, float* __restrict__ colsum, int kseg, int sps) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  __shared__ __bf16 As[2][3][GB_M][GB_KP];
  __shared__ __bf16 Bs[2][3][GB_N][GB_KP];
  __shared__ float cs_red[8][GB_N];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int nbn = N / GB_N, nblk = (M / GB_M) * nbn;
  int b, ysplit;
  xcd_tile_split(nblk, b, ysplit);
  const int m0 = (b / nbn) * GB_M, n0 = (b % nbn) * GB_N;
  int kbeg, kend;
  split_range(ysplit, kchunk, kseg, sps, kbeg, kend);
  const int nk = (kend - kbeg) / GB_K;
  const bool do_cs = colsum != nullptr && m0 == 0;

  // 128 x 32 floats per operand tile = 1024 float4, 4 per thread
  f4v ra[4], rb[4];
  float csum[4] = {0.f, 0.f, 0.f, 0.f};
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int idx = tid + 256 * i;
      if (TA) {  // A stored [k][m]: 32 rows of 32 float4
        const int k = idx / 32, mv = idx % 32;
        ra[i] = *(const f4v*)(A + (size_t)(k0 + k) * lda + m0 + mv * 4);
      } else {   // A stored [m][k]: 128 rows of 8 float4
        const int m = idx / 8, kv = idx % 8;
        ra[i] = *(const f4v*)(A + (size_t)(m0 + m) * lda + k0 + kv * 4);
      }
      if (TB) {  // B stored [n][k]
        const int n = idx / 8, kv = idx % 8;
        rb[i] = *(const f4v*)(B + (size_t)(n0 + n) * ldb + k0 + kv * 4);
      } else {   // B stored [k][n]
        const int k = idx / 32, nv = idx % 32;
        rb[i] = *(const f4v*)(B + (size_t)(k0 + k) * ldb + n0 + nv * 4);
      }
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int idx = tid + 256 * i;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        __bf16 h, m, l;
        split3_bf16(ra[i][e], h, m, l);
        int mm, kk;
        if (TA) { kk = idx / 32; mm = (idx % 32) * 4 + e; } else { mm = idx / 8; kk = (idx % 8) * 4 + e; }
        As[buf][0][mm][kk] = h; As[buf][1][mm][kk] = m; As[buf][2][mm][kk] = l;
        split3_bf16(rb[i][e], h, m, l);
        int nn;
        if (TB) { nn = idx / 8; kk = (idx % 8) * 4 + e; } else { kk = idx / 32; nn = (idx % 32) * 4 + e; }
        Bs[buf][0][nn][kk] = h; Bs[buf][1][nn][kk] = m; Bs[buf][2][nn][kk] = l;
      }
      if (!TB && do_cs)
#pragma unroll
        for (int e = 0; e < 4; ++e) csum[e] += rb[i][e];
    }
  };

  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};

  if (nk > 0) {
    gload(kbeg);
    lstore(0);
    __syncthreads();
  }
  const int g = lane >> 4, r = lane & 15;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload(kbeg + (kt + 1) * GB_K);
    bf8 af[4][3], bfr[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int p = 0; p < 3; ++p) af[i][p] = *(const bf8*)&As[cur][p][wm + i * 16 + r][8 * g];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int p = 0; p < 3; ++p) bfr[j][p] = *(const bf8*)&Bs[cur][p][wn + j * 16 + r][8 * g];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        f4 c = acc[i][j];
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][2], bfr[j][0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][1], bfr[j][1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][0], bfr[j][2], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][1], bfr[j][0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][0], bfr[j][1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][0], bfr[j][0], c, 0, 0, 0);
        acc[i][j] = c;
      }
    if (kt + 1 < nk) lstore(cur ^ 1);
    __syncthreads();
  }
  if (!TB && do_cs) {
    // thread t summed columns 4 (t % 32) .. +3 over rows k = t / 32 + 8 i of every stage
#pragma unroll
    for (int e = 0; e < 4; ++e) cs_red[tid / 32][(tid % 32) * 4 + e] = csum[e];
    __syncthreads();
    if (tid < GB_N) {
      float t = 0.f;
#pragma unroll
      for (int q = 0; q < 8; ++q) t += cs_red[q][tid];
      colsum[(size_t)ysplit * N + n0 + tid] = t;
    }
  }
  float* Cz = C + (size_t)ysplit * M * ldc;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn + j * 16 + r;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + i * 16 + mfma32_row(lane, q);
        float v = acc[i][j][q];
        if (EPI == EPI_BIAS_ELU) v = elu_f(v + bias[col]);
        if (EPI == EPI_ELU_BWD) {
          const float y = Y[(size_t)row * ldy + col];
          v *= (y > 0.f ? 1.f : y + 1.f);
        }
        Cz[(size_t)row * ldc + col] = v;
      }
    }
}

}  // namespace gdx
