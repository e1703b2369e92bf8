// AO-side MFMA kernels of the XC path, templated on the element type of the AO planes and
// the density matrix (double: FP64 path, float: FP32 path).  Both follow the staging scheme
// of gdx_gemm.h: 16-byte global loads into registers one stage ahead, double-buffered LDS,
// one barrier per stage.
//
// density_ao_kernel  (grad_dft/molecule.py:388-502)
//   Block = 64 grid points (16 per wave), all four AO components c in {phi, dx, dy, dz}.
//   For each 128-column (FP64: 64) chunk of orbitals: X_c = A_c D on MFMA (4 components x
//   8 column blocks per wave), then the AO values at the chunk's columns are staged through
//   LDS 16 columns at a time and the row contractions added to per-lane FP64 sums:
//     rho += X_phi . phi, grad_j += X_phi . d_j phi, lapl += X_phi . lapl(phi),
//     T += sum_j X_j . d_j phi.
//   After the last chunk the 16 lanes sharing a row are summed; raw sums are written as
//   (rho, gx, gy, gz, T, L) and point_fwd applies grad = 2 g, tau = T / 2, lapl = 2 T + 2 L.
//
// vxc_ao_kernel  (adjoint of the above, grad_dft/train.py:86-121)
//   V_s[a][b] = sum_r phi_a Z_b + sum_j d_j phi_a (u d_j phi_b), Z = v_rho phi + sum_j
//   2 v_grad_j d_j phi + 2 v_lapl lapl(phi), u = v_tau / 2 + 2 v_lapl.  One 128 x 128 output
//   tile per block (4 waves x 64 x 64), K' = 4 components x 4 points per stage; B rows are
//   formed from the AO values and the per-point potentials while being stored to LDS.
//   Blocks split the points into chunks; per-chunk tiles are summed by reduce_vxc.
#pragma once
#include "gdx_common.h"
#include "gdx_gemm.h"
#include "gdx_gemm_bf16.h"

namespace gdx {

constexpr int DN_PT = 64, DN_K = 16;

// Column chunk of the density kernel: 64 output columns per pass (accumulator registers:
// 4 components x CB/16 blocks of 16 x 16 per wave).  With 64 columns the FP32 kernel fits
// two blocks per CU (228 VGPRs, no spills; 15% faster than one block of 128-column chunks on
// benzene 6-31G*); a 128-column chunk at two blocks per CU spills.
template <typename T> struct DensCB { static constexpr int value = 64; };

// One 64-point tile (bx) of one spin (s): ao (5, npts_pad, nP) planes of one molecule, Dp its
// padded density (nsd, nP, nP); raw sums to dens_part[(s * out_rows + p) * 6 + v].
template <typename T, int CB = DensCB<T>::value>
__device__ __forceinline__ void density_ao_body(const T* __restrict__ ao, int npts_pad, int nP,
                                                const T* __restrict__ Dp, double* __restrict__ dens_part,
                                                size_t out_rows, int bx, int s) {
  typedef typename MfmaT<T>::acc_t acc_t;
  constexpr int NJ = CB / 16;
  constexpr int VEC = 16 / sizeof(T);
  constexpr int KV = DN_K / VEC;                           // vectors per 16-k row segment
  constexpr int NVA = 4 * DN_PT * KV / 256;                // A vectors per thread
  constexpr int BVEC = DN_K * CB / VEC;                    // B vectors per stage
  constexpr int NVB = (BVEC + 255) / 256;
  constexpr int CVEC = 5 * DN_PT * 16 / VEC;               // contraction vectors per stage
  constexpr int NVC = CVEC / 256;
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  constexpr int AK = DN_K + 4;   // A tile row: 16 k of one (component, point), 16-byte aligned
  constexpr int AS_ELEMS = 2 * 4 * DN_PT * AK;
  constexpr int CS_ELEMS = 5 * DN_PT * 17;
  static_assert(CS_ELEMS <= AS_ELEMS, "contraction tile must fit in the A tile's LDS");
  __shared__ __align__(16) T Araw[AS_ELEMS];
  __shared__ __align__(16) T Bs[2][DN_K][CB + 4];
  T (*As)[4][DN_PT][AK] = (T (*)[4][DN_PT][AK])Araw;   // phase 1: [buf][c][p][k]
  T (*Cs)[DN_PT][17] = (T (*)[DN_PT][17])Araw;         // phase 2 (aliases As)
  typedef T v4_t __attribute__((ext_vector_type(4)));

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int p0 = bx * DN_PT;
  const T* D = Dp + (size_t)s * nP * nP;
  const size_t plane = (size_t)npts_pad * nP;
  const int nk = nP / DN_K;

  double part[4][6];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int v = 0; v < 6; ++v) part[q][v] = 0.0;

  vec_t ra[NVA], rb[NVB], rc[NVC];
  for (int c0 = 0; c0 < nP; c0 += CB) {
    // ---- phase 1: X_c = A_c D[:, c0:c0+CB] for the 4 AO components ----
    auto gload = [&](int k0) {
#pragma unroll
      for (int i = 0; i < NVA; ++i) {
        const int idx = tid + 256 * i;
        const int kv = idx % KV, rest = idx / KV;
        const int p = rest % DN_PT, c = rest / DN_PT;
        ra[i] = *(const vec_t*)(ao + c * plane + (size_t)(p0 + p) * nP + k0 + kv * VEC);
      }
#pragma unroll
      for (int i = 0; i < NVB; ++i) {
        const int idx = tid + 256 * i;
        if (idx < BVEC) {
          const int k = idx / (CB / VEC), cv = idx % (CB / VEC);
          const int col = c0 + cv * VEC;
          rb[i] = col < nP ? *(const vec_t*)(D + (size_t)(k0 + k) * nP + col) : (vec_t)(T)0;
        }
      }
    };
    auto lstore = [&](int buf) {
#pragma unroll
      for (int i = 0; i < NVA; ++i) {
        const int idx = tid + 256 * i;
        const int kv = idx % KV, rest = idx / KV;
        const int p = rest % DN_PT, c = rest / DN_PT;
        *(vec_t*)&As[buf][c][p][kv * VEC] = ra[i];
      }
#pragma unroll
      for (int i = 0; i < NVB; ++i) {
        const int idx = tid + 256 * i;
        if (idx < BVEC) {
          const int k = idx / (CB / VEC), cv = idx % (CB / VEC);
#pragma unroll
          for (int e = 0; e < VEC; ++e) Bs[buf][k][cv * VEC + e] = rb[i][e];
        }
      }
    };
    acc_t acc[4][NJ];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int j = 0; j < NJ; ++j) acc[c][j] = (acc_t){0, 0, 0, 0};
    // 16-column blocks of this chunk inside the matrix (the last chunk of an nP that is not a
    // multiple of CB has fewer): the MFMAs of the others would only multiply zero columns
    const int njv = min(NJ, (nP - c0) / 16);
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < nk) gload((kt + 1) * DN_K);
      {
        // k-group g = lane / 16 supplies k = 4 g + kk to MFMA step kk (one permutation of the
        // stage's 16 k for both operands): one 4-element LDS read per A fragment
        const int g = lane >> 4, r = lane & 15;
        v4_t af[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) af[c] = *(const v4_t*)&As[cur][c][wave * 16 + r][4 * g];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          T bf[NJ];
#pragma unroll
          for (int j = 0; j < NJ; ++j) bf[j] = Bs[cur][4 * g + kk][j * 16 + r];
#pragma unroll
          for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int j = 0; j < NJ; ++j)
              if (j < njv) acc[c][j] = MfmaT<T>::op(af[c][kk], bf[j], acc[c][j]);
        }
      }
      if (kt + 1 < nk) lstore(cur ^ 1);
      __syncthreads();
    }
    // ---- phase 2: contraction with the AO values at the chunk's columns, 16 columns per
    // stage staged through LDS (coalesced loads of the next stage overlap this one) ----
    const int nst = njv;
    auto cload = [&](int st) {
#pragma unroll
      for (int i = 0; i < NVC; ++i) {
        const int idx = tid + 256 * i;
        const int cv = idx % (16 / VEC), rest = idx / (16 / VEC);
        const int p = rest % DN_PT, pl = rest / DN_PT;
        rc[i] = *(const vec_t*)(ao + pl * plane + (size_t)(p0 + p) * nP + c0 + st * 16 + cv * VEC);
      }
    };
    if (nst > 0) cload(0);
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      if (j >= nst) break;
#pragma unroll
      for (int i = 0; i < NVC; ++i) {
        const int idx = tid + 256 * i;
        const int cv = idx % (16 / VEC), rest = idx / (16 / VEC);
        const int p = rest % DN_PT, pl = rest / DN_PT;
#pragma unroll
        for (int e = 0; e < VEC; ++e) Cs[pl][p][cv * VEC + e] = rc[i][e];
      }
      __syncthreads();
      if (j + 1 < nst) cload(j + 1);
      const int col = lane & 15;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int p = wave * 16 + MfmaT<T>::row(lane, q);
        const double phi = Cs[0][p][col], gx = Cs[1][p][col], gy = Cs[2][p][col], gz = Cs[3][p][col],
                     lp = Cs[4][p][col];
        const double x0 = acc[0][j][q];
        part[q][0] += x0 * phi;
        part[q][1] += x0 * gx;
        part[q][2] += x0 * gy;
        part[q][3] += x0 * gz;
        part[q][4] += (double)acc[1][j][q] * gx + (double)acc[2][j][q] * gy + (double)acc[3][j][q] * gz;
        part[q][5] += x0 * lp;
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int v = 0; v < 6; ++v) {
      const double t = sum16(part[q][v]);
      if ((lane & 15) == 0) {
        const int p = p0 + wave * 16 + MfmaT<T>::row(lane, q);
        dens_part[((size_t)s * out_rows + p) * DENS_STRIDE + v] = t;
      }
    }
}

template <typename T>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 2 : 1) void density_ao_kernel(const T* __restrict__ ao, int npts_pad,
                                                                                 int nP, const T* __restrict__ Dp,
                                                                                 double* __restrict__ dens_part) {
  density_ao_body<T>(ao, npts_pad, nP, Dp, dens_part, (size_t)npts_pad, blockIdx.x, blockIdx.y);
}

// A batch of molecules with different grids and bases, one launch: molecule m owns tiles
// [tile0_m, tile0_m + npts_pad_m / 64) of blockIdx.x and writes rows [row0_m, row0_m + npts_pad_m)
// of every spin plane of the batch's dens_part (out_rows rows a plane).
struct AoBatchMol {
  const void* ao;      // (5, npts_pad, nP) planes in the element type
  const void* Dp;      // padded density (or direction) of this molecule, (2, nP, nP)
  long row0;           // first row in the batch's per-point arrays
  int npts_pad, nP;
  int tile0, ntile;    // first 64-point tile in the launch, tiles
};

__device__ __forceinline__ int batch_find(const AoBatchMol* __restrict__ mols, int nmol, int tile) {
  int lo = 0, hi = nmol - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mols[mid].tile0 <= tile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <typename T>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 2 : 1) void density_ao_batch_kernel(
    const AoBatchMol* __restrict__ mols, int nmol, size_t out_rows, double* __restrict__ dens_part) {
  const int m = batch_find(mols, nmol, blockIdx.x);
  const AoBatchMol& M = mols[m];
  density_ao_body<T>((const T*)M.ao, M.npts_pad, M.nP, (const T*)M.Dp, dens_part + (size_t)M.row0 * DENS_STRIDE,
                     out_rows, blockIdx.x - M.tile0, blockIdx.y);
}

constexpr int VX_P = 4, VX_K = 4 * VX_P;
constexpr int VX_FLUSH = 8;   // FP32 path: stages (of VX_P points) per FP32 partial sum
constexpr int VB_FLUSH = 4;   // split-BF16 path: stages (of VB_P points) per FP32 partial sum

// VT: output tile edge (128, or 64 for nao_pad <= 64); each wave owns a (VT/2)^2 quarter.
// FLUSH (FP32 only): the MFMA accumulators are added into FP64 registers every VX_FLUSH
// stages, so no FP32 sum runs over more than VX_FLUSH * VX_K terms -- needed when a block's
// point chunk is long, because the V_xc integrand cancels strongly near the nuclei (large
// Laplacian terms); short chunks go without it (fewer registers, more blocks per CU).
// Body of one block: output tile bx, point chunk by, of one spin whose per-point potentials
// start at vp (rows of this molecule); the tile's partial sums go to out (nP x nP).
template <typename T, bool SAME, int VT, bool FLUSH>
__device__ __forceinline__ void vxc_ao_body(const T* __restrict__ ao, int npts_pad, int nP,
                                            const double* __restrict__ vp, int chunk, double* __restrict__ out,
                                            int bx, int by) {
  typedef typename MfmaT<T>::acc_t acc_t;
  constexpr int WB = VT / 32;          // 16-wide MFMA blocks per wave edge
  constexpr int CPT = VT / 64;         // columns staged per thread
  __shared__ T As[2][VX_K][VT + 4];
  __shared__ T Bs[2][VX_K][VT + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nt = (nP + VT - 1) / VT;
  const int a0 = (bx / nt) * VT, b0 = (bx % nt) * VT;
  const int rbeg = by * chunk;
  const int rend = min(npts_pad, rbeg + chunk);
  const size_t plane = (size_t)npts_pad * nP;
  const int wm = (wave >> 1) * (VT / 2), wn = (wave & 1) * (VT / 2);
  // 16 x 16 blocks of this wave's quarter inside the nP x nP matrix (edge tiles of an nP that
  // is not a multiple of VT have fewer; the others would only accumulate zeros)
  const int niv = max(0, min(WB, (nP - a0 - wm) / 16)), njv = max(0, min(WB, (nP - b0 - wn) / 16));
  // staging role of this thread: point p of the stage, CPT adjacent columns
  const int sp = tid >> 6, sc = (tid & 63) * CPT;
  typedef T v2 __attribute__((ext_vector_type(CPT)));

  v2 ra[4], rb[5];
  double sv[VPT_STRIDE];
  auto gload = [&](int r0) {
    const int r = r0 + sp;
    const bool ok = r < rend;
    const int ca = a0 + sc, cb = b0 + sc;
    if (!SAME) {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        ra[c] = (ok && ca < nP) ? *(const v2*)(ao + c * plane + (size_t)r * nP + ca) : (v2)(T)0;
    }
#pragma unroll
    for (int c = 0; c < 5; ++c)
      rb[c] = (ok && cb < nP) ? *(const v2*)(ao + c * plane + (size_t)r * nP + cb) : (v2)(T)0;
#pragma unroll
    for (int q = 0; q < VPT_STRIDE; ++q) sv[q] = ok ? vp[(size_t)r * VPT_STRIDE + q] : 0.0;
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int e = 0; e < CPT; ++e) {
#pragma unroll
      for (int c = 0; c < 4; ++c) As[buf][c * VX_P + sp][sc + e] = SAME ? rb[c][e] : ra[c][e];
      const double phi = rb[0][e], gx = rb[1][e], gy = rb[2][e], gz = rb[3][e], lp = rb[4][e];
      Bs[buf][sp][sc + e] = (T)(sv[0] * phi + sv[1] * gx + sv[2] * gy + sv[3] * gz + sv[4] * lp);
      Bs[buf][VX_P + sp][sc + e] = (T)(sv[5] * gx);
      Bs[buf][2 * VX_P + sp][sc + e] = (T)(sv[5] * gy);
      Bs[buf][3 * VX_P + sp][sc + e] = (T)(sv[5] * gz);
    }
  };

  acc_t acc[WB][WB];
#pragma unroll
  for (int i = 0; i < WB; ++i)
#pragma unroll
    for (int j = 0; j < WB; ++j) acc[i][j] = (acc_t){0, 0, 0, 0};
  static_assert(!FLUSH || sizeof(T) == 4, "FP64 accumulates in FP64 already");
  double dacc[FLUSH ? WB : 1][FLUSH ? WB : 1][4];
  if (FLUSH) {
#pragma unroll
    for (int i = 0; i < WB; ++i)
#pragma unroll
      for (int j = 0; j < WB; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) dacc[i][j][q] = 0.0;
  }
  const int nst = (rend - rbeg + VX_P - 1) / VX_P;
  if (nst > 0) {
    gload(rbeg);
    lstore(0);
    __syncthreads();
  }
  for (int st = 0; st < nst; ++st) {
    const int cur = st & 1;
    if (FLUSH && st > 0 && st % VX_FLUSH == 0) {
#pragma unroll
      for (int i = 0; i < WB; ++i)
#pragma unroll
        for (int j = 0; j < WB; ++j) {
#pragma unroll
          for (int q = 0; q < 4; ++q) dacc[i][j][q] += (double)acc[i][j][q];
          acc[i][j] = (acc_t){0, 0, 0, 0};
        }
    }
    if (st + 1 < nst) gload(rbeg + (st + 1) * VX_P);
#pragma unroll
    for (int kk = 0; kk < VX_K / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      T af[WB], bf[WB];
#pragma unroll
      for (int i = 0; i < WB; ++i) af[i] = As[cur][kr][wm + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < WB; ++j) bf[j] = Bs[cur][kr][wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < WB; ++i)
#pragma unroll
        for (int j = 0; j < WB; ++j)
          if (i < niv && j < njv) acc[i][j] = MfmaT<T>::op(af[i], bf[j], acc[i][j]);
    }
    if (st + 1 < nst) lstore(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < WB; ++i)
#pragma unroll
    for (int j = 0; j < WB; ++j) {
      const int col = b0 + wn + j * 16 + (lane & 15);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = a0 + wm + i * 16 + MfmaT<T>::row(lane, q);
        double v = (double)acc[i][j][q];
        if (FLUSH) v += dacc[FLUSH ? i : 0][FLUSH ? j : 0][q];
        if (row < nP && col < nP) out[(size_t)row * nP + col] = v;
      }
    }
}

template <typename T, bool SAME, int VT, bool FLUSH = false>
__global__ __launch_bounds__(256, VT == 64 ? 4 : ((FLUSH && SAME) ? 2 : 1)) void vxc_ao_kernel(const T* __restrict__ ao, int npts_pad,
                                                                             int nP, const double* __restrict__ vpt,
                                                                             int chunk, double* __restrict__ vpart) {
  const int s = blockIdx.z;
  vxc_ao_body<T, SAME, VT, FLUSH>(ao, npts_pad, nP, vpt + (size_t)s * npts_pad * VPT_STRIDE, chunk,
                                  vpart + ((size_t)blockIdx.y * gridDim.z + s) * nP * nP, blockIdx.x, blockIdx.y);
}

// V_xc of a batch of molecules in one launch per kernel variant: block (x, s) of entry e runs
// tile (x - blk0_e) % ntile_e, chunk (x - blk0_e) / ntile_e of molecule mol_e, with the batch's
// per-point potentials (rows row0 of every spin plane of vpt_rows rows) and partials to
// vpart_e[(chunk * nsv + s) * nP^2].
struct VxcBatchMol {
  const void* ao;
  double* vpart;
  long row0;
  int npts_pad, nP;
  int blk0, ntile, nchunk, chunk;
};

__device__ __forceinline__ int vxc_batch_find(const VxcBatchMol* __restrict__ mols, int nmol, int blk) {
  int lo = 0, hi = nmol - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mols[mid].blk0 <= blk) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <typename T, bool SAME, int VT, bool FLUSH = false>
__global__ __launch_bounds__(256, VT == 64 ? 4 : ((FLUSH && SAME) ? 2 : 1)) void vxc_ao_batch_kernel(
    const VxcBatchMol* __restrict__ mols, int nmol, const double* __restrict__ vpt, size_t vpt_rows, int nsv) {
  const int m = vxc_batch_find(mols, nmol, blockIdx.x);
  const VxcBatchMol& M = mols[m];
  const int local = blockIdx.x - M.blk0;
  const int tile = local % M.ntile, chunk = local / M.ntile;
  const int s = blockIdx.y;
  vxc_ao_body<T, SAME, VT, FLUSH>((const T*)M.ao, M.npts_pad, M.nP,
                                  vpt + ((size_t)s * vpt_rows + M.row0) * VPT_STRIDE, M.chunk,
                                  M.vpart + ((size_t)chunk * nsv + s) * M.nP * M.nP, tile, chunk);
}

// vxc_bf16x6_kernel: vxc_ao_kernel's product on bf16 MFMA (v_mfma_f32_16x16x32_bf16) with
// the three-term bf16 split of every FP32 operand and the six leading products (see
// gdx_gemm_bf16.h), for T = float AO planes.  Stage = 8 points x 4 components = 32 k;
// LDS tiles are k-contiguous per output row/column ([split][row][k], k = 8 c + p) so a
// lane's 8-k fragment is one 16-byte read.  Each thread stages 2 points x 2 columns, so the
// split terms of consecutive points pack into 4-byte LDS stores.
constexpr int VB_P = 8, VB_K = 4 * VB_P, VB_KP = VB_K + 8;

template <bool SAME>
__global__ __launch_bounds__(256) void vxc_bf16x6_kernel(const float* __restrict__ ao, int npts_pad, int nP,
                                                         const double* __restrict__ vpt, int chunk,
                                                         double* __restrict__ vpart) {
  constexpr int VT = 128;
  __shared__ __align__(16) __bf16 As[2][3][VT][VB_KP];
  __shared__ __align__(16) __bf16 Bs[2][3][VT][VB_KP];
  typedef float v2f __attribute__((ext_vector_type(2)));
  typedef __bf16 v2b __attribute__((ext_vector_type(2)));
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nt = (nP + VT - 1) / VT;
  const int a0 = (blockIdx.x / nt) * VT, b0 = (blockIdx.x % nt) * VT;
  const int s = blockIdx.z;
  const int rbeg = blockIdx.y * chunk;
  const int rend = min(npts_pad, rbeg + chunk);
  const size_t plane = (size_t)npts_pad * nP;
  const double* vp = vpt + (size_t)s * npts_pad * VPT_STRIDE;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  // 16 x 16 blocks of this wave's quarter inside the matrix (the rest multiply zero padding)
  const int niv = max(0, min(4, (nP - a0 - wm) / 16)), njv = max(0, min(4, (nP - b0 - wn) / 16));
  // staging role: points 2 pp, 2 pp + 1 of the stage; columns 2 cc, 2 cc + 1 of the tile
  const int pp = tid >> 6, cc = tid & 63;
  v2f ra[2][4], rb[2][5];
  double sv[2][VPT_STRIDE];
  auto gload = [&](int r0) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int r = r0 + 2 * pp + e;
      const bool ok = r < rend;
      const int ca = a0 + 2 * cc, cb = b0 + 2 * cc;
      if (!SAME) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          ra[e][c] = (ok && ca < nP) ? *(const v2f*)(ao + c * plane + (size_t)r * nP + ca) : (v2f)0.f;
      }
#pragma unroll
      for (int c = 0; c < 5; ++c)
        rb[e][c] = (ok && cb < nP) ? *(const v2f*)(ao + c * plane + (size_t)r * nP + cb) : (v2f)0.f;
#pragma unroll
      for (int q = 0; q < VPT_STRIDE; ++q) sv[e][q] = ok ? vp[(size_t)r * VPT_STRIDE + q] : 0.0;
    }
  };
  auto put = [&](__bf16 (*T)[VT][VB_KP], int col, int k, float x0, float x1) {
    __bf16 h0, m0, l0, h1, m1, l1;
    split3_bf16(x0, h0, m0, l0);
    split3_bf16(x1, h1, m1, l1);
    *(v2b*)&T[0][col][k] = (v2b){h0, h1};
    *(v2b*)&T[1][col][k] = (v2b){m0, m1};
    *(v2b*)&T[2][col][k] = (v2b){l0, l1};
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = 2 * cc + j;
      const int k0 = 2 * pp;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float x0 = SAME ? rb[0][c][j] : ra[0][c][j];
        const float x1 = SAME ? rb[1][c][j] : ra[1][c][j];
        put(As[buf], col, c * VB_P + k0, x0, x1);
      }
      float z[2], u[2][3];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const double phi = rb[e][0][j], gx = rb[e][1][j], gy = rb[e][2][j], gz = rb[e][3][j], lp = rb[e][4][j];
        z[e] = (float)(sv[e][0] * phi + sv[e][1] * gx + sv[e][2] * gy + sv[e][3] * gz + sv[e][4] * lp);
        u[e][0] = (float)(sv[e][5] * gx);
        u[e][1] = (float)(sv[e][5] * gy);
        u[e][2] = (float)(sv[e][5] * gz);
      }
      put(Bs[buf], col, k0, z[0], z[1]);
#pragma unroll
      for (int q = 0; q < 3; ++q) put(Bs[buf], col, (q + 1) * VB_P + k0, u[0][q], u[1][q]);
    }
  };

  f4 acc[4][4];
  double dacc[4][4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q = 0; q < 4; ++q) dacc[i][j][q] = 0.0;
    }
  const int nst = (rend - rbeg + VB_P - 1) / VB_P;
  if (nst > 0) {
    gload(rbeg);
    lstore(0);
    __syncthreads();
  }
  const int g = lane >> 4, r = lane & 15;
  for (int st = 0; st < nst; ++st) {
    const int cur = st & 1;
    if (st > 0 && st % VB_FLUSH == 0) {   // 32 points (128 k) per FP32 sum
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
          for (int q = 0; q < 4; ++q) dacc[i][j][q] += (double)acc[i][j][q];
          acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
        }
    }
    if (st + 1 < nst) gload(rbeg + (st + 1) * VB_P);
    bf8 af[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int p = 0; p < 3; ++p) af[i][p] = *(const bf8*)&As[cur][p][wm + i * 16 + r][8 * g];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bf8 bq[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) bq[p] = *(const bf8*)&Bs[cur][p][wn + j * 16 + r][8 * g];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i >= niv || j >= njv) continue;
        f4 cacc = acc[i][j];
        cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][2], bq[0], cacc, 0, 0, 0);
        cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][1], bq[1], cacc, 0, 0, 0);
        cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][0], bq[2], cacc, 0, 0, 0);
        cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][1], bq[0], cacc, 0, 0, 0);
        cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][0], bq[1], cacc, 0, 0, 0);
        cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][0], bq[0], cacc, 0, 0, 0);
        acc[i][j] = cacc;
      }
    }
    if (st + 1 < nst) lstore(cur ^ 1);
    __syncthreads();
  }
  double* out = vpart + ((size_t)blockIdx.y * gridDim.z + s) * nP * nP;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = b0 + wn + j * 16 + r;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = a0 + wm + i * 16 + mfma32_row(lane, q);
        if (row < nP && col < nP) out[(size_t)row * nP + col] = dacc[i][j][q] + (double)acc[i][j][q];
      }
    }
}

}  // namespace gdx
