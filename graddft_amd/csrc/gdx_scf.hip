// SCF-side kernels: Coulomb matrix from the 8-fold packed ERI, exchange (set-up), Cholesky
// of the overlap, batched symmetric eigensolver, DIIS linear solve, non-XC energy.
//
// Reference computations replaced:
//   coulomb_potential / coulomb_energy / nonXC   grad_dft/molecule.py:694-811
//   exchange of the initial density (the stale semi-numerical HF term of B3LYP,
//     molecule.py:507-613 with chi from interface/pyscf.py:152-160) taken from the ERI
//   safe_general_eigh: cholesky, inv, eigh        grad_dft/utils/eigenproblem.py:110-149
//   JittableDiis.cdiis_minimize: inv(B) @ e_0     grad_dft/evaluate.py:1162-1205
//
// Packed ERI order (shared with the host integral library): pair index I = i(i+1)/2 + j
// (i >= j), element index I(I+1)/2 + K (I >= K).
#include <algorithm>
#include <mutex>
#include <string>

#include "gdx_common.h"
#include "gdx_profile.h"
#include "gdx_tridiag.h"

namespace gdx {
namespace {

constexpr int NT = 256;

// ------------------------------------------------------------- Coulomb (J) ----------
// J_ij = sum_kl (ij|kl) D_kl = sum_{K} M_IK Dt_K with Dt_K = D_kl + D_lk (k > l), D_kk,
// M the symmetric P x P matrix whose lower triangle is the packed ERI.  Tiles of
// JT x JT elements of the lower triangle: each block forms row sums (M Dt over its
// columns) and column sums (M^T Dt over its strictly-lower rows); partial vectors are
// added by coulomb_finish.  Every packed element is read exactly once.  Tiles strictly below
// the diagonal and inside the matrix (all but O(P / JT) of them) take a path without
// per-element bounds checks, so a thread issues a row's loads back to back.
constexpr int JT = 256;   // large tile (P >= 4096)
constexpr int JTS = 64;   // small tile
constexpr long J_ROWS_MIN = 12000;   // P from which coulomb_rows_kernel reads the tiles

__global__ void pack_density_kernel(const double* __restrict__ rdm1, int n, double* __restrict__ Dt) {
  const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long P = (long)n * (n + 1) / 2;
  if (I >= P) return;
  const int i = (int)((sqrt(8.0 * (double)I + 1.0) - 1.0) / 2.0);
  int ii = i;
  while ((long)ii * (ii + 1) / 2 > I) --ii;
  while ((long)(ii + 1) * (ii + 2) / 2 <= I) ++ii;
  const int j = (int)(I - (long)ii * (ii + 1) / 2);
  const double* D0 = rdm1;
  const double* D1 = rdm1 + (size_t)n * n;
  const double dij = D0[(size_t)ii * n + j] + D1[(size_t)ii * n + j];
  const double dji = D0[(size_t)j * n + ii] + D1[(size_t)j * n + ii];
  Dt[I] = ii == j ? dij : dij + dji;
}

__device__ __forceinline__ void coulomb_tile_of(long x, long& bi, long& bk) {
  // triangular tile index -> (bi, bk), bk <= bi
  bi = (long)((sqrt(8.0 * (double)x + 1.0) - 1.0) / 2.0);
  while (bi * (bi + 1) / 2 > x) --bi;
  while ((bi + 1) * (bi + 2) / 2 <= x) ++bi;
  bk = x - bi * (bi + 1) / 2;
}

template <int TT>
__global__ __launch_bounds__(NT) void coulomb_tile_kernel(const double* __restrict__ eri, long P,
                                                          const double* __restrict__ Dt,
                                                          double* __restrict__ prow, double* __restrict__ pcol) {
  // TT x TT tile of the lower triangle; 256 threads as 16 (ty) x 16 (tx); rows I = I0 + ty + 16 a,
  // columns K = K0 + 16 j + tx
  constexpr int NA = TT / 16;
  __shared__ double cs[16][TT + 1];
  long bi, bk;
  coulomb_tile_of(blockIdx.x, bi, bk);
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  const long I0 = bi * TT, K0 = bk * TT;
  double dk[NA], csum[NA];
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    const long K = K0 + j * 16 + tx;
    dk[j] = K < P ? Dt[K] : 0.0;
    csum[j] = 0.0;
  }
  double rsum[NA];
  if (bk < bi && I0 + TT <= P) {
    // every K < every I of the tile, every I < P
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const long I = I0 + ty + 16 * a;
      const double di = Dt[I];
      const double* row = eri + I * (I + 1) / 2 + K0 + tx;
      double m[NA];
#pragma unroll
      for (int j = 0; j < NA; ++j) m[j] = __builtin_nontemporal_load(row + 16 * j);
      double r = 0.0;
#pragma unroll
      for (int j = 0; j < NA; ++j) {
        r += m[j] * dk[j];
        csum[j] += m[j] * di;
      }
      rsum[a] = r;
    }
  } else {
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const long I = I0 + ty + 16 * a;
      double r = 0.0;
      if (I < P) {
        const double di = Dt[I];
        const double* row = eri + I * (I + 1) / 2;
#pragma unroll
        for (int j = 0; j < NA; ++j) {
          const long K = K0 + j * 16 + tx;
          if (K <= I) {
            const double m = row[K];
            r += m * dk[j];
            if (K < I) csum[j] += m * di;
          }
        }
      }
      rsum[a] = r;
    }
  }
  // row sums over the 16 threads of this ty (consecutive lanes)
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const double r = sum16(rsum[a]);
    const long I = I0 + ty + 16 * a;
    if (tx == 0 && I < P) prow[bk * P + I] = r;
  }
#pragma unroll
  for (int j = 0; j < NA; ++j) cs[ty][j * 16 + tx] = csum[j];
  __syncthreads();
  for (int k = tid; k < TT; k += NT) {
    const long K = K0 + k;
    double c = 0.0;
#pragma unroll
    for (int t = 0; t < 16; ++t) c += cs[t][k];
    if (K < P) pcol[bi * P + K] = c;
  }
}

// Large P: the same JT x JT tiles and partial vectors, read with a wave per row: lane l takes
// columns K0 + l + 64 c (c < 4), so one load instruction covers 64 consecutive elements of one
// row.  Each wave takes groups of J_RG consecutive rows; the J_RG row partials of a group are
// summed across the wave by a halving butterfly (lane l ends with the row l >> 4), the column
// partials stay in registers until the four waves add them through LDS.
constexpr int J_RG = 4;

template <bool CHECK>
__device__ __forceinline__ void coulomb_row_group(const double* __restrict__ eri, long P, const double* __restrict__ Dt,
                                                  long Ib, long K0, int lane, const double (&dk)[4],
                                                  double (&csum)[4], double* __restrict__ prow_b) {
  double v[J_RG];
#pragma unroll
  for (int a = 0; a < J_RG; ++a) {
    const long I = Ib + a;
    double r = 0.0;
    if (!CHECK || I < P) {
      const double di = Dt[I];
      const double* row = eri + I * (I + 1) / 2 + K0 + lane;
      double m[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) m[c] = (!CHECK || K0 + lane + 64 * c <= I) ? __builtin_nontemporal_load(row + 64 * c) : 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        r += m[c] * dk[c];
        if (!CHECK || K0 + lane + 64 * c < I) csum[c] += m[c] * di;
      }
    }
    v[a] = r;
  }
  // halving butterfly over lane bits 5 and 4: lane l keeps row (l >> 4), summed over 4 lanes
  {
    const bool up = (lane & 32) != 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const double send = up ? v[i] : v[i + 2], keep = up ? v[i + 2] : v[i];
      v[i] = keep + __shfl_xor(send, 32, 64);
    }
    const bool up2 = (lane & 16) != 0;
    const double send = up2 ? v[0] : v[1], keep = up2 ? v[1] : v[0];
    v[0] = keep + __shfl_xor(send, 16, 64);
  }
  double r = v[0];
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
  const long I = Ib + (lane >> 4);
  if ((lane & 15) == 0 && (!CHECK || I < P)) prow_b[I] = r;
}

__global__ __launch_bounds__(NT) void coulomb_rows_kernel(const double* __restrict__ eri, long P,
                                                          const double* __restrict__ Dt,
                                                          double* __restrict__ prow, double* __restrict__ pcol) {
  static_assert(NT == 256 && JT == 256, "four waves, one column per thread in the epilogue");
  __shared__ double cs[4][JT];
  long bi, bk;
  coulomb_tile_of(blockIdx.x, bi, bk);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long I0 = bi * JT, K0 = bk * JT;
  double dk[4], csum[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const long K = K0 + lane + 64 * c;
    dk[c] = K < P ? Dt[K] : 0.0;
    csum[c] = 0.0;
  }
  const bool check = !(bk < bi && I0 + JT <= P);
  for (int g = 0; g < JT / (4 * J_RG); ++g) {
    const long Ib = I0 + (long)(g * 4 + wave) * J_RG;
    if (check)
      coulomb_row_group<true>(eri, P, Dt, Ib, K0, lane, dk, csum, prow + bk * P);
    else
      coulomb_row_group<false>(eri, P, Dt, Ib, K0, lane, dk, csum, prow + bk * P);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) cs[wave][lane + 64 * c] = csum[c];
  __syncthreads();
  const int k = threadIdx.x;
  const long K = K0 + k;
  const double c = (cs[0][k] + cs[1][k]) + (cs[2][k] + cs[3][k]);
  if (K < P) pcol[bi * P + K] = c;
}

// J~_I = sum_{bk <= I/JT} prow[bk][I] + sum_{bi >= I/JT} pcol[bi][I], unpacked to J (n x n)
__global__ void coulomb_finish_kernel(const double* __restrict__ prow, const double* __restrict__ pcol, long P,
                                      int nb, int tt, int n, double* __restrict__ J) {
  const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= P) return;
  const int t = (int)(I / tt);
  double acc = 0.0;
  for (int b = 0; b <= t; ++b) acc += prow[b * P + I];
  for (int b = t; b < nb; ++b) acc += pcol[b * P + I];
  int i = (int)((sqrt(8.0 * (double)I + 1.0) - 1.0) / 2.0);
  while ((long)i * (i + 1) / 2 > I) --i;
  while ((long)(i + 1) * (i + 2) / 2 <= I) ++i;
  const int j = (int)(I - (long)i * (i + 1) / 2);
  J[(size_t)i * n + j] = acc;
  J[(size_t)j * n + i] = acc;
}

// E = e_nuc + sum D h + 1/2 sum D J (D = D_up + D_dn), plus, when K0 is given, the stale
// exact-exchange energy -(a0/2) sum_s sum D_s K0_s (single block)
__global__ void nonxc_energy_kernel(const double* __restrict__ rdm1, const double* __restrict__ h1e,
                                    const double* __restrict__ J, const double* __restrict__ K0, double a0, int n,
                                    double e_nuc, double* __restrict__ out) {
  __shared__ double red[NT / 64];
  double acc = 0.0;
  const long nn = (long)n * n;
  for (long i = threadIdx.x; i < nn; i += blockDim.x) {
    const double D = rdm1[i] + rdm1[nn + i];
    acc += D * h1e[i] + 0.5 * D * J[i];
    if (K0 != nullptr) acc += -0.5 * a0 * (rdm1[i] * K0[i] + rdm1[nn + i] * K0[nn + i]);
  }
  acc = warp_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = e_nuc + red[0] + red[1] + red[2] + red[3];
}

// Fock assembly of predict (grad_dft/train.py:148-215), per spin s:
//   F = clip(h + J + V_s); F = clip((F + F^T)/2); [F = clip(F + X_s + X_s^T), X_s = -(a0/2) K0_s];
//   F = clip(F), clip(x) = x if |x| > 1e-30 else 0.  One thread per (s, i, j).
__device__ __forceinline__ double clip30s(double x) { return fabs(x) > 1e-30 ? x : 0.0; }

__global__ void fock_assemble_kernel(const double* __restrict__ h1e, const double* __restrict__ J,
                                     const double* __restrict__ V, const double* __restrict__ K0, double a0, int n,
                                     double* __restrict__ F) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nn = (long)n * n;
  if (e >= 2 * nn) return;
  const int s = (int)(e / nn);
  const int i = (int)((e % nn) / n), j = (int)(e % n);
  const long ij = (long)i * n + j, ji = (long)j * n + i;
  const double fij = clip30s(h1e[ij] + J[ij] + V[s * nn + ij]);
  const double fji = clip30s(h1e[ji] + J[ji] + V[s * nn + ji]);
  double f = clip30s(0.5 * (fij + fji));
  if (K0 != nullptr) {
    const double x = -0.5 * a0 * K0[s * nn + ij], xt = -0.5 * a0 * K0[s * nn + ji];
    f = clip30s(f + x + xt);
  }
  F[e] = clip30s(f);
}

// ------------------------------------------------------------- exchange (K) ---------
// K_s[a][c] = sum_bd (ab|cd) D_s[b][d] for general D (the reverse of K is K itself: the
// operator is self-adjoint under <A, B> = sum A_ac B_ac).  One wave per (s, a, c); its lanes
// run over d (for d <= c and pair(a,b) >= pair(c,d) the packed elements are contiguous), b
// sequentially; each lane sums its own d in a fixed order and the lanes are combined in a
// fixed pattern, so K is reproducible bit for bit.
__device__ __forceinline__ long pair_index(int i, int j) {
  return i >= j ? (long)i * (i + 1) / 2 + j : (long)j * (j + 1) / 2 + i;
}

__global__ __launch_bounds__(256) void exchange_gather_kernel(const double* __restrict__ eri, int n, int ns,
                                                              const double* __restrict__ D, double* __restrict__ K) {
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const long nn = (long)n * n;
  if (w >= ns * nn) return;
  const int s = (int)(w / nn);
  const int a = (int)((w % nn) / n), c = (int)(w % n);
  const double* Ds = D + (size_t)s * nn;
  double acc = 0.0;
  for (int b = 0; b < n; ++b) {
    const long P1 = pair_index(a, b);
    const long t1 = P1 * (P1 + 1) / 2;
    for (int d = lane; d < n; d += 64) {
      const long P2 = pair_index(c, d);
      const long idx = P1 >= P2 ? t1 + P2 : P2 * (P2 + 1) / 2 + P1;
      acc += eri[idx] * Ds[(size_t)b * n + d];
    }
  }
  acc = warp_sum(acc);
  if (lane == 0) K[w] = acc;
}

// Exchange-ordered packed ERI ("XO"): K as a symmetric matrix-vector product.  Over ordered
// pairs I = a n + c, M[I][J] = (ab|cd) for J = b n + d is symmetric (M[J][I] = (ba|dc)), and
// K_s = M vec(D_s) with vec(D)[b n + d] = D[b][d] -- so the Coulomb kernels' packed
// lower-triangle GEMV serves K with P' = n^2 "pairs" and the density matrix itself as the
// vector.  xo[I (I + 1) / 2 + J] = M[I][J] for J <= I (n^2 (n^2 + 1) / 2 doubles, 4x the
// 8-fold packed ERI).  One block per row I, its threads striding over J (coalesced stores,
// gathered loads: built once per molecule).
__global__ __launch_bounds__(NT) void xo_build_kernel(const double* __restrict__ eri, int n,
                                                      double* __restrict__ xo) {
  const long I = blockIdx.x;
  const int a = (int)(I / n), c = (int)(I % n);
  double* row = xo + I * (I + 1) / 2;
  for (long J = threadIdx.x; J <= I; J += NT) {
    const int b = (int)(J / n), d = (int)(J % n);
    const long P1 = pair_index(a, b), P2 = pair_index(c, d);
    row[J] = P1 >= P2 ? eri[P1 * (P1 + 1) / 2 + P2] : eri[P2 * (P2 + 1) / 2 + P1];
  }
}

// Both spins in one pass over the XO copy: coulomb_rows_kernel with two right-hand sides
// (row partials of spin s to prow_s, column partials to pcol_s).  Per spin the products, the
// butterfly and the LDS sums run in the single-spin kernel's order, so each spin's result is
// bit for bit the one-spin product's.
template <bool CHECK>
__device__ __forceinline__ void xo_row_group2(const double* __restrict__ xo, long P, const double* __restrict__ D0,
                                              const double* __restrict__ D1, long Ib, long K0, int lane,
                                              const double (&dk)[2][4], double (&csum)[2][4],
                                              double* __restrict__ prow0, double* __restrict__ prow1) {
  double v[2][J_RG];
#pragma unroll
  for (int a = 0; a < J_RG; ++a) {
    const long I = Ib + a;
    double r0 = 0.0, r1 = 0.0;
    if (!CHECK || I < P) {
      const double di0 = D0[I], di1 = D1[I];
      const double* row = xo + I * (I + 1) / 2 + K0 + lane;
      double m[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) m[c] = (!CHECK || K0 + lane + 64 * c <= I) ? __builtin_nontemporal_load(row + 64 * c) : 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        r0 += m[c] * dk[0][c];
        r1 += m[c] * dk[1][c];
        if (!CHECK || K0 + lane + 64 * c < I) {
          csum[0][c] += m[c] * di0;
          csum[1][c] += m[c] * di1;
        }
      }
    }
    v[0][a] = r0;
    v[1][a] = r1;
  }
#pragma unroll
  for (int sp = 0; sp < 2; ++sp) {
    double* w = v[sp];
    const bool up = (lane & 32) != 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const double send = up ? w[i] : w[i + 2], keep = up ? w[i + 2] : w[i];
      w[i] = keep + __shfl_xor(send, 32, 64);
    }
    const bool up2 = (lane & 16) != 0;
    const double send = up2 ? w[0] : w[1], keep = up2 ? w[1] : w[0];
    double r = keep + __shfl_xor(send, 16, 64);
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
    const long I = Ib + (lane >> 4);
    if ((lane & 15) == 0 && (!CHECK || I < P)) (sp == 0 ? prow0 : prow1)[I] = r;
  }
}

__global__ __launch_bounds__(NT) void xo_rows2_kernel(const double* __restrict__ xo, long P,
                                                      const double* __restrict__ D0, const double* __restrict__ D1,
                                                      double* __restrict__ prow, double* __restrict__ pcol,
                                                      long nbP) {
  static_assert(NT == 256 && JT == 256, "four waves, one column per thread in the epilogue");
  __shared__ double cs[2][4][JT];
  long bi, bk;
  coulomb_tile_of(blockIdx.x, bi, bk);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long I0 = bi * JT, K0 = bk * JT;
  double dk[2][4], csum[2][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const long K = K0 + lane + 64 * c;
    dk[0][c] = K < P ? D0[K] : 0.0;
    dk[1][c] = K < P ? D1[K] : 0.0;
    csum[0][c] = 0.0;
    csum[1][c] = 0.0;
  }
  const bool check = !(bk < bi && I0 + JT <= P);
  for (int g = 0; g < JT / (4 * J_RG); ++g) {
    const long Ib = I0 + (long)(g * 4 + wave) * J_RG;
    if (check)
      xo_row_group2<true>(xo, P, D0, D1, Ib, K0, lane, dk, csum, prow + bk * P, prow + nbP + bk * P);
    else
      xo_row_group2<false>(xo, P, D0, D1, Ib, K0, lane, dk, csum, prow + bk * P, prow + nbP + bk * P);
  }
#pragma unroll
  for (int sp = 0; sp < 2; ++sp)
#pragma unroll
    for (int c = 0; c < 4; ++c) cs[sp][wave][lane + 64 * c] = csum[sp][c];
  __syncthreads();
  const int k = threadIdx.x;
  const long K = K0 + k;
#pragma unroll
  for (int sp = 0; sp < 2; ++sp) {
    const double c = (cs[sp][0][k] + cs[sp][1][k]) + (cs[sp][2][k] + cs[sp][3][k]);
    if (K < P) pcol[sp * nbP + bi * P + K] = c;
  }
}

// K[I] = sum of the row and column partials of the XO product (coulomb_finish_kernel without
// the pair unpacking: I = a n + c is K's own row-major index).
__global__ void xo_finish_kernel(const double* __restrict__ prow, const double* __restrict__ pcol, long P, int nb,
                                 int tt, double* __restrict__ K) {
  const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= P) return;
  const int t = (int)(I / tt);
  double acc = 0.0;
  for (int b = 0; b <= t; ++b) acc += prow[b * P + I];
  for (int b = t; b < nb; ++b) acc += pcol[b * P + I];
  K[I] = acc;
}

// Status word of an SCF cycle, OR-ed into *status by one block (no atomics): bit 0 the
// eigensolver ran max_sweeps without converging (sweeps[s] >= max_sweeps for a spin), bit 1
// the energy is not finite, bit 2 |tr(D_s S) - sum_i occ_s,i| > tol for a spin (the orbitals
// the density was built from are not S-orthonormal).  NaN inputs set bits 1 / 2.
__global__ __launch_bounds__(NT) void scf_status_kernel(int nspin, int n, const double* __restrict__ rdm1,
                                                       const double* __restrict__ S, const double* __restrict__ occ,
                                                       const int* __restrict__ sweeps, int max_sweeps,
                                                       const double* __restrict__ energy, double tol,
                                                       int* __restrict__ status) {
  __shared__ double red[2][NT / 64];
  const int tid = threadIdx.x;
  int bits = 0;
  for (int s = 0; s < nspin; ++s) {
    double tr = 0.0, nocc = 0.0;
    const double* D = rdm1 + (size_t)s * n * n;
    for (long e = tid; e < (long)n * n; e += NT) {
      const int i = (int)(e / n), j = (int)(e % n);
      tr += D[e] * S[(size_t)j * n + i];
    }
    for (int i = tid; i < n; i += NT) nocc += occ[(size_t)s * n + i];
    tr = warp_sum(tr);
    nocc = warp_sum(nocc);
    if ((tid & 63) == 0) {
      red[0][tid >> 6] = tr;
      red[1][tid >> 6] = nocc;
    }
    __syncthreads();
    if (tid == 0) {
      double a = 0.0, b = 0.0;
      for (int w = 0; w < NT / 64; ++w) {
        a += red[0][w];
        b += red[1][w];
      }
      if (!(fabs(a - b) <= tol)) bits |= 4;
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (sweeps)
      for (int s = 0; s < nspin; ++s)
        if (sweeps[s] >= max_sweeps) bits |= 1;
    if (energy && !isfinite(*energy)) bits |= 2;
    *status |= bits;
  }
}

__global__ void zero_kernel(double* __restrict__ p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ------------------------------------------------------------- Cholesky ---------------
// S = L L^T (lower) and Linv = L^-1, one workgroup; n sequential column steps.
__global__ __launch_bounds__(1024) void cholesky_inv_kernel(const double* __restrict__ S, int n,
                                                            double* __restrict__ L, double* __restrict__ Linv,
                                                            int* __restrict__ status) {
  const int tid = threadIdx.x;
  for (long e = tid; e < (long)n * n; e += blockDim.x) L[e] = 0.0;
  __syncthreads();
  __shared__ double diag;
  __shared__ int bad;
  if (tid == 0) bad = 0;
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      double s = S[(size_t)j * n + j];
      for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
      if (!(s > 0.0)) bad = 1;
      diag = sqrt(s > 0.0 ? s : 1.0);
      L[(size_t)j * n + j] = diag;
    }
    __syncthreads();
    for (int i = j + 1 + tid; i < n; i += blockDim.x) {
      double s = S[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = s / diag;
    }
    __syncthreads();
  }
  // Linv column c by forward substitution (one thread per column)
  for (int c = tid; c < n; c += blockDim.x) {
    for (int i = 0; i < n; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) s -= L[(size_t)i * n + k] * Linv[(size_t)k * n + c];
      Linv[(size_t)i * n + c] = i < c ? 0.0 : s / L[(size_t)i * n + i];
    }
  }
  if (tid == 0) *status = bad;
}

// ------------------------------------------------------------- eigensolver --------------
// Parallel cyclic Jacobi (round-robin pair schedule), one workgroup per symmetric matrix.
// Each round rotates n/2 disjoint (p, q) pairs: rotation angles from the current 2x2
// blocks, then one pass that applies both sides of the rotation (A <- J^T A J) to every 2x2
// block (rows of pair k, columns of pair l) and the column rotation to V (V <- V J) -- the
// per-element arithmetic of a row pass followed by a column pass, with one barrier less.
// Sweeps stop when the off-diagonal Frobenius norm is below tol * ||A||_F.  Eigenvalues are
// returned ascending with eigenvectors as columns of V (the order of jnp.linalg.eigh).
// A, V and the warm-start product live in LDS when 3 n^2 doubles fit (SMEM),
// otherwise in the workspace; 1024 threads either way.
constexpr int EIG_THREADS_SMEM = 1024, EIG_THREADS_GLOBAL = 1024;
constexpr size_t EIG_SMEM_MAX = 64 * 1024;   // floor of the LDS budget (any device)

__device__ __forceinline__ int rr_partner(int r, int k, int m, int* p, int* q) {
  // round r of the round-robin tournament on m (even) players: pair k
  int a, b;
  if (k == 0) { a = m - 1; b = r; }
  else { a = (r + k) % (m - 1); b = (r - k + (m - 1)) % (m - 1); }
  *p = a < b ? a : b;
  *q = a < b ? b : a;
  return 0;
}

// Shared LDS bookkeeping of one Jacobi workgroup.
struct JacobiShared {
  double cs_c[512], cs_s[512];
  int pp[512], qq[512];
  double red[16];
  int done;
};

// Dynamic LDS a Jacobi workgroup may take on the current device: the per-workgroup LDS limit
// (160 KB on gfx950) less the static JacobiShared block and 1 KB of slack, never below
// EIG_SMEM_MAX.  Cached per device.
static size_t eig_smem_budget() {
  static size_t cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return EIG_SMEM_MAX;
  if (cached[dev] == 0) {
    int v = 0;
    size_t b = EIG_SMEM_MAX;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        (size_t)v > sizeof(JacobiShared) + 1024 + EIG_SMEM_MAX)
      b = (size_t)v - sizeof(JacobiShared) - 1024;
    cached[dev] = b;
  }
  return cached[dev];
}

// C = op(A) op(B) for n x n row-major matrices, op(X) = X or X^T, by one 1024-thread
// workgroup through 32 x 32 LDS tiles (ta/tb: take the transpose).  Used where the matrices
// live in the workspace: each element is fetched once per tile with coalesced loads instead
// of n dependent L2 reads per output.  C must not alias A or B.  Every thread must call it.
constexpr int MT = 32;
__device__ void mm_tiled(double* __restrict__ C, const double* __restrict__ A, bool ta,
                         const double* __restrict__ B, bool tb, int n, double (*As)[MT + 1],
                         double (*Bs)[MT + 1]) {
  const int tid = threadIdx.x, ty = tid / MT, tx = tid % MT;
  const int nb = (n + MT - 1) / MT;
  for (int bi = 0; bi < nb; ++bi)
    for (int bj = 0; bj < nb; ++bj) {
      double acc = 0.0;
      for (int k0 = 0; k0 < n; k0 += MT) {
        // As[r][k] = op(A)[bi*MT + r][k0 + k], Bs[k][c] = op(B)[k0 + k][bj*MT + c]
        if (!ta) {
          const int r = bi * MT + ty, k = k0 + tx;
          As[ty][tx] = (r < n && k < n) ? A[(size_t)r * n + k] : 0.0;
        } else {   // op(A)[r][k] = A[k][r]: read row k0+ty of A, columns bi*MT+tx
          const int k = k0 + ty, r = bi * MT + tx;
          As[tx][ty] = (r < n && k < n) ? A[(size_t)k * n + r] : 0.0;
        }
        if (!tb) {
          const int k = k0 + ty, c = bj * MT + tx;
          Bs[ty][tx] = (k < n && c < n) ? B[(size_t)k * n + c] : 0.0;
        } else {   // op(B)[k][c] = B[c][k]: read row bj*MT+ty of B, columns k0+tx
          const int c = bj * MT + ty, k = k0 + tx;
          Bs[tx][ty] = (k < n && c < n) ? B[(size_t)c * n + k] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < MT; ++kk) acc += As[ty][kk] * Bs[kk][tx];
        __syncthreads();
      }
      const int r = bi * MT + ty, c = bj * MT + tx;
      if (r < n && c < n) C[(size_t)r * n + c] = acc;
    }
  __syncthreads();
}

// Workspace path: dynamic LDS for the matrix being rotated (n^2 doubles) when it fits beside
// the static JacobiShared block and the product tiles, else 0 (A stays in the workspace).
static size_t a_lds_bytes(int n) {
  const size_t a = (size_t)n * n * sizeof(double), tiles = 2 * (size_t)MT * (MT + 1) * sizeof(double);
  return a + tiles <= eig_smem_budget() ? a : 0;
}

// Warm start (V0 != NULL, orthonormal columns): the sweeps run on V0^T A V0 starting from
// V = V0, so a matrix close to one already diagonalised by V0 needs few sweeps; the result
// is the same eigendecomposition (eigenvalues, and eigenvectors up to sign).  On entry T
// holds the matrix; A, V, T are n x n buffers (LDS or workspace) and the two products read
// no global memory in their inner loops.  On exit A holds the (rotated) matrix, V the start.
// V is kept transposed (V[j * n + i] = V_ij, column j contiguous) from here to jacobi_sort, so
// the rotation of a column pair touches two contiguous rows of the buffer (coalesced in the
// workspace path; consecutive banks in LDS).
template <int NTH>
__device__ void jacobi_start(double* A, double* V, double* T, int n, const double* __restrict__ V0,
                             double (*As)[MT + 1] = nullptr, double (*Bs)[MT + 1] = nullptr) {
  const int tid = threadIdx.x;
  const size_t nn = (size_t)n * n;
  if (V0 == nullptr) {
    for (size_t e = tid; e < nn; e += NTH) {
      A[e] = T[e];
      V[e] = (e / n == e % n) ? 1.0 : 0.0;
    }
    __syncthreads();
    return;
  }
  for (size_t e = tid; e < nn; e += NTH) V[e] = V0[(e % n) * n + e / n];   // V0 transposed
  __syncthreads();
  if (As != nullptr) {
    mm_tiled(A, T, false, V, true, n, As, Bs);    // A <- T V0
    mm_tiled(T, V, false, A, false, n, As, Bs);   // T <- V0^T A
  } else {
  // A <- T V0   (V0_kj = V[j * n + k])
  for (size_t e = tid; e < nn; e += NTH) {
    const int i = (int)(e / n), j = (int)(e % n);
    double t = 0.0;
    for (int k = 0; k < n; ++k) t += T[(size_t)i * n + k] * V[(size_t)j * n + k];
    A[e] = t;
  }
  __syncthreads();
  // T <- V0^T A   (V0_ki = V[i * n + k])
  for (size_t e = tid; e < nn; e += NTH) {
    const int i = (int)(e / n), j = (int)(e % n);
    double t = 0.0;
    for (int k = 0; k < n; ++k) t += V[(size_t)i * n + k] * A[(size_t)k * n + j];
    T[e] = t;
  }
  __syncthreads();
  }
  // A <- T symmetrised (rounding); V already holds V0
  for (size_t e = tid; e < nn; e += NTH) {
    const int i = (int)(e / n), j = (int)(e % n);
    if (i < j) {
      const double a = 0.5 * (T[(size_t)i * n + j] + T[(size_t)j * n + i]);
      A[(size_t)i * n + j] = a;
      A[(size_t)j * n + i] = a;
    } else if (i == j) {
      A[e] = T[e];
    }
  }
  __syncthreads();
}

// Cyclic Jacobi sweeps on A (symmetric, n x n) accumulating V <- V J (V stored transposed);
// returns the sweeps run.
template <int NTH>
__device__ int jacobi_sweeps(double* A, double* V, int n, int max_sweeps, double tol, JacobiShared& sh) {
  const int tid = threadIdx.x;
  const size_t nn = (size_t)n * n;
  const int m = n + (n & 1);  // players incl. a dummy when n is odd
  const int npair = m / 2;
  const int nA = npair * npair, nV = n * npair;
  int sweep = 0;
  for (; sweep <= max_sweeps; ++sweep) {
    // convergence test
    double off = 0.0, tot = 0.0;
    for (size_t e = tid; e < nn; e += NTH) {
      const double v = A[e] * A[e];
      tot += v;
      if (e / n != e % n) off += v;
    }
    off = warp_sum(off);
    tot = warp_sum(tot);
    if ((tid & 63) == 0) sh.red[tid >> 6] = off;
    __syncthreads();
    if (tid == 0) {
      double o = 0.0;
      for (int i = 0; i < NTH / 64; ++i) o += sh.red[i];
      sh.red[0] = o;
    }
    __syncthreads();
    const double offs = sh.red[0];
    __syncthreads();
    if ((tid & 63) == 0) sh.red[tid >> 6] = tot;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int i = 0; i < NTH / 64; ++i) t += sh.red[i];
      sh.done = (offs <= tol * tol * t) || sweep == max_sweeps;
    }
    __syncthreads();
    if (sh.done) break;
    for (int r = 0; r < m - 1; ++r) {
      // rotation parameters
      for (int k = tid; k < npair; k += NTH) {
        int p, q;
        rr_partner(r, k, m, &p, &q);
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = A[(size_t)p * n + q];
          if (apq != 0.0) {
            const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
          }
        }
        sh.cs_c[k] = c; sh.cs_s[k] = s; sh.pp[k] = p; sh.qq[k] = q;
      }
      __syncthreads();
      // items e = tid, tid + NTH, ...: first the 2x2 blocks (k, l) = (e / npair, e % npair),
      // then the V rows (l, i) = (ev / n, ev % n), ev = e - nA; the indices are stepped
      // instead of divided per item (NTH > npair and NTH > n - 1: one carry at most)
      int e = tid;
      {
        int k = e / npair, l = e % npair;
        const int dk = NTH / npair, dl = NTH % npair;
        for (; e < nA; e += NTH) {
          const int pk = sh.pp[k], qk = sh.qq[k], pl = sh.pp[l], ql = sh.qq[l];
          const bool rk = qk < n, rl = ql < n;   // false: the dummy player, no rotation
          const double ck = sh.cs_c[k], sk = sh.cs_s[k], cl = sh.cs_c[l], sl = sh.cs_s[l];
          double r_pp = A[(size_t)pk * n + pl];
          double r_pq = rl ? A[(size_t)pk * n + ql] : 0.0;
          double r_qp = rk ? A[(size_t)qk * n + pl] : 0.0;
          double r_qq = (rk && rl) ? A[(size_t)qk * n + ql] : 0.0;
          if (rk) {   // rows of pair k first
            const double apj = r_pp, aqj = r_qp;
            r_pp = ck * apj - sk * aqj;
            r_qp = sk * apj + ck * aqj;
            const double apj2 = r_pq, aqj2 = r_qq;
            r_pq = ck * apj2 - sk * aqj2;
            r_qq = sk * apj2 + ck * aqj2;
          }
          if (rl) {   // then columns of pair l
            A[(size_t)pk * n + pl] = cl * r_pp - sl * r_pq;
            A[(size_t)pk * n + ql] = sl * r_pp + cl * r_pq;
            if (rk) {
              A[(size_t)qk * n + pl] = cl * r_qp - sl * r_qq;
              A[(size_t)qk * n + ql] = sl * r_qp + cl * r_qq;
            }
          } else {
            A[(size_t)pk * n + pl] = r_pp;
            if (rk) A[(size_t)qk * n + pl] = r_qp;
          }
          l += dl;
          k += dk;
          if (l >= npair) { l -= npair; ++k; }
        }
      }
      {
        // V <- V J on the transposed storage: rows pl, ql of the buffer, element i
        int ev = e - nA;
        int l = ev / n, i = ev % n;
        const int dlv = NTH / n, di = NTH % n;
        for (; ev < nV; ev += NTH) {
          const int pl = sh.pp[l], ql = sh.qq[l];
          if (ql < n) {
            const double c = sh.cs_c[l], sv = sh.cs_s[l];
            const double vip = V[(size_t)pl * n + i], viq = V[(size_t)ql * n + i];
            V[(size_t)pl * n + i] = c * vip - sv * viq;
            V[(size_t)ql * n + i] = sv * vip + c * viq;
          }
          i += di;
          l += dlv;
          if (i >= n) { i -= n; ++l; }
        }
      }
      __syncthreads();
    }
  }
  return sweep;
}

// Ascending eigenvalues (rank of each diagonal element of A, ties by index) to w and the
// matching columns of V (stored transposed) to Vo, row-major (either may live anywhere).
template <int NTH>
__device__ void jacobi_sort(const double* A, const double* V, int n, double* __restrict__ w,
                            double* __restrict__ Vo) {
  for (int i = threadIdx.x; i < n; i += NTH) {
    const double di = A[(size_t)i * n + i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = A[(size_t)j * n + j];
      rank += (dj < di) || (dj == di && j < i);
    }
    w[rank] = di;
    for (int row = 0; row < n; ++row) Vo[(size_t)row * n + rank] = V[(size_t)i * n + row];
  }
}

template <bool SMEM, int NTH>
__global__ __launch_bounds__(NTH) void jacobi_eigh_kernel(const double* __restrict__ Ain, int n,
                                                          const double* __restrict__ V0in,
                                                          double* __restrict__ w_out,
                                                          double* __restrict__ V_out,
                                                          double* __restrict__ gws, int max_sweeps,
                                                          double tol, int* __restrict__ sweeps_out,
                                                          int a_lds) {
  extern __shared__ double smem[];
  __shared__ JacobiShared sh;
  const int b = blockIdx.x;
  const size_t nn = (size_t)n * n;
  double* A = SMEM ? smem : gws + (size_t)b * 3 * nn;
  double* V = A + nn;
  double* T = V + nn;
  const double* Asrc = Ain + (size_t)b * nn;
  for (size_t e = threadIdx.x; e < nn; e += NTH) T[e] = Asrc[e];
  __syncthreads();
  if constexpr (SMEM) {
    jacobi_start<NTH>(A, V, T, n, V0in == nullptr ? nullptr : V0in + (size_t)b * nn);
  } else {
    __shared__ double tA[MT][MT + 1], tB[MT][MT + 1];
    jacobi_start<NTH>(A, V, T, n, V0in == nullptr ? nullptr : V0in + (size_t)b * nn, tA, tB);
  }
  double* Ar = A;
  if (!SMEM && a_lds) {   // workspace path, A alone fits in LDS: rotate it there
    for (size_t e = threadIdx.x; e < nn; e += NTH) smem[e] = A[e];
    __syncthreads();
    Ar = smem;
  }
  const int sweep = jacobi_sweeps<NTH>(Ar, V, n, max_sweeps, tol, sh);
  jacobi_sort<NTH>(Ar, V, n, w_out + (size_t)b * n, V_out + (size_t)b * nn);
  if (threadIdx.x == 0 && sweeps_out) sweeps_out[b] = sweep;
}

// One generalised-eigenproblem step of the SCF loop per spin (forward only):
//   F = sum_i x_{1+i} F_i                           (DIIS extrapolation, evaluate.py:1198-1205)
//   C' = (Linv F Linv^T + (Linv F Linv^T)^T) / 2    (safe_general_eigh, eigenproblem.py:110-129)
//   C' = V diag(w) V^T (Jacobi, warm-started from V0), C = Linv^T V
//   D = C diag(occ) C^T                             (make_rdm1 with aufbau occupations)
// focks (m, nspin, n, n), x (nspin, m+1).  Buffers A, V, T, L (n x n each) in LDS when 4 n^2
// doubles fit (SMEM), otherwise in the workspace.
template <bool SMEM, int NTH>
__global__ __launch_bounds__(NTH) void scf_step_kernel(int n, int m, int nspin, const double* __restrict__ focks,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ Linv,
                                                       const double* __restrict__ V0in,
                                                       const double* __restrict__ occ, double* __restrict__ F_out,
                                                       double* __restrict__ w_out, double* __restrict__ V_out,
                                                       double* __restrict__ C_out, double* __restrict__ D_out,
                                                       double* __restrict__ gws, int max_sweeps, double tol,
                                                       int* __restrict__ sweeps_out, int a_lds) {
  extern __shared__ double smem[];
  __shared__ JacobiShared sh;
  const int s = blockIdx.x, tid = threadIdx.x;
  const size_t nn = (size_t)n * n;
  double* A = SMEM ? smem : gws + (size_t)s * 4 * nn;
  double* V = A + nn;
  double* T = V + nn;
  double* L = T + nn;
  const double* xs = x + (size_t)s * (m + 1) + 1;
  for (size_t e = tid; e < nn; e += NTH) {
    // the m history entries of this element are loaded before any is summed (m <= 16 at a
    // time), so their latencies overlap
    double f = 0.0;
    for (int i0 = 0; i0 < m; i0 += 16) {
      double fv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u)
        fv[u] = i0 + u < m ? focks[((size_t)(i0 + u) * nspin + s) * nn + e] : 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (i0 + u < m) f += xs[i0 + u] * fv[u];
    }
    T[e] = f;
    F_out[(size_t)s * nn + e] = f;
    L[e] = Linv[e];
  }
  __syncthreads();
  // workspace path: LDS tiles for the n^3 products (unused in the LDS variant)
  __shared__ double tA[SMEM ? 1 : MT][MT + 1], tB[SMEM ? 1 : MT][MT + 1];
  if constexpr (!SMEM) {
    mm_tiled(A, L, false, T, false, n, tA, tB);   // A <- L F
    mm_tiled(V, A, false, L, true, n, tA, tB);    // V <- A L^T
  } else {
  // A <- L F
  for (size_t e = tid; e < nn; e += NTH) {
    const int i = (int)(e / n), j = (int)(e % n);
    double t = 0.0;
    for (int k = 0; k < n; ++k) t += L[(size_t)i * n + k] * T[(size_t)k * n + j];
    A[e] = t;
  }
  __syncthreads();
  // V <- A L^T
  for (size_t e = tid; e < nn; e += NTH) {
    const int i = (int)(e / n), j = (int)(e % n);
    double t = 0.0;
    for (int k = 0; k < n; ++k) t += A[(size_t)i * n + k] * L[(size_t)j * n + k];
    V[e] = t;
  }
  __syncthreads();
  }
  // T <- (V + V^T) / 2
  for (size_t e = tid; e < nn; e += NTH) {
    const int i = (int)(e / n), j = (int)(e % n);
    T[e] = 0.5 * (V[e] + V[(size_t)j * n + i]);
  }
  __syncthreads();
  if constexpr (SMEM)
    jacobi_start<NTH>(A, V, T, n, V0in == nullptr ? nullptr : V0in + (size_t)s * nn);
  else
    jacobi_start<NTH>(A, V, T, n, V0in == nullptr ? nullptr : V0in + (size_t)s * nn, tA, tB);
  double* Ar = A;
  if (!SMEM && a_lds) {   // workspace path, A alone fits in LDS: rotate it there
    for (size_t e = tid; e < nn; e += NTH) smem[e] = A[e];
    __syncthreads();
    Ar = smem;
  }
  const int sweep = jacobi_sweeps<NTH>(Ar, V, n, max_sweeps, tol, sh);
  // sorted eigenpairs: w, V to the outputs and the sorted V to T
  jacobi_sort<NTH>(Ar, V, n, w_out + (size_t)s * n, T);
  __syncthreads();
  for (size_t e = tid; e < nn; e += NTH) V_out[(size_t)s * nn + e] = T[e];
  const double* os = occ + (size_t)s * n;
  if constexpr (!SMEM) {
    mm_tiled(A, L, true, T, false, n, tA, tB);    // A <- L^T T  (C = Linv^T V)
    for (size_t e = tid; e < nn; e += NTH) {
      C_out[(size_t)s * nn + e] = A[e];
      T[e] = A[e] * os[e % n];                    // T <- C diag(occ)
    }
    __syncthreads();
    mm_tiled(D_out + (size_t)s * nn, T, false, A, true, n, tA, tB);   // D = T C^T
  } else {
  // A <- L^T T  (C = Linv^T V)
  for (size_t e = tid; e < nn; e += NTH) {
    const int i = (int)(e / n), j = (int)(e % n);
    double t = 0.0;
    for (int k = 0; k < n; ++k) t += L[(size_t)k * n + i] * T[(size_t)k * n + j];
    A[e] = t;
    C_out[(size_t)s * nn + e] = t;
  }
  __syncthreads();
  // T <- C diag(occ), then D[i][k] = sum_j T[i][j] C[k][j]
  for (size_t e = tid; e < nn; e += NTH) T[e] = A[e] * os[e % n];
  __syncthreads();
  for (size_t e = tid; e < nn; e += NTH) {
    const int i = (int)(e / n), k = (int)(e % n);
    double t = 0.0;
    for (int j = 0; j < n; ++j) t += T[(size_t)i * n + j] * A[(size_t)k * n + j];
    D_out[(size_t)s * nn + e] = t;
  }
  }
  if (tid == 0 && sweeps_out) sweeps_out[s] = sweep;
}

// ------------------------------------------------------------- tridiagonal path ------
// Where the Jacobi buffers do not fit LDS (n > ~70) and n <= tri::NMAX the eigensolver is the
// Householder / bisection / inverse-iteration sequence of gdx_tridiag.h, and the SCF step's
// n^3 products run as many-workgroup tiled GEMMs instead of inside one workgroup.

// C[b] = alpha op(A[b]) diag(S[b]) op(B[b]) + beta C0[b] for n x n row-major matrices (op =
// transpose when TA / TB; S and C0 may be NULL; C may alias C0 but not A or B); batch strides
// sA, sB, sS, sC (0: shared by the batch; C0 uses sC).  32 x 32 output tile per 256-thread
// workgroup, 2 x 2 outputs per thread, K staged through LDS 32 at a time.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void small_gemm_kernel(int n, const double* __restrict__ A, long sA,
                                                         const double* __restrict__ B, long sB,
                                                         const double* __restrict__ S, long sS, double alpha,
                                                         const double* C0, double beta, double* C, long sC) {
  __shared__ double As[32][33], Bs[32][33];
  const int b = blockIdx.z, tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  A += (size_t)b * sA;
  B += (size_t)b * sB;
  C += (size_t)b * sC;
  if (C0) C0 += (size_t)b * sC;
  if (S) S += (size_t)b * sS;
  double acc00 = 0.0, acc01 = 0.0, acc10 = 0.0, acc11 = 0.0;
  for (int k0 = 0; k0 < n; k0 += 32) {
    for (int q = tid; q < 1024; q += 256) {
      const int u = q >> 5, v = q & 31;   // v runs along contiguous memory
      if (!TA) {
        const int r = r0 + u, kk = k0 + v;
        As[u][v] = (r < n && kk < n) ? A[(size_t)r * n + kk] : 0.0;
      } else {
        const int kk = k0 + u, r = r0 + v;
        As[v][u] = (r < n && kk < n) ? A[(size_t)kk * n + r] : 0.0;
      }
      if (!TB) {
        const int kk = k0 + u, c = c0 + v;
        Bs[u][v] = (kk < n && c < n) ? B[(size_t)kk * n + c] * (S ? S[kk] : 1.0) : 0.0;
      } else {
        const int c = c0 + u, kk = k0 + v;
        Bs[v][u] = (kk < n && c < n) ? B[(size_t)c * n + kk] * (S ? S[kk] : 1.0) : 0.0;
      }
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < 32; ++kk) {
      const double a0 = As[ty][kk], a1 = As[ty + 16][kk], b0 = Bs[kk][tx], b1 = Bs[kk][tx + 16];
      acc00 += a0 * b0;
      acc01 += a0 * b1;
      acc10 += a1 * b0;
      acc11 += a1 * b1;
    }
    __syncthreads();
  }
  const int ra = r0 + ty, rb = r0 + ty + 16, ca = c0 + tx, cb = c0 + tx + 16;
  auto out = [&](int r, int c, double acc) {
    if (r >= n || c >= n) return;
    const size_t o = (size_t)r * n + c;
    C[o] = C0 ? alpha * acc + beta * C0[o] : alpha * acc;
  };
  out(ra, ca, acc00);
  out(ra, cb, acc01);
  out(rb, ca, acc10);
  out(rb, cb, acc11);
}

template <bool TA, bool TB>
static void small_gemm(hipStream_t st, int nb, int n, const double* A, long sA, const double* B, long sB,
                       const double* S, long sS, double* C, long sC, double alpha = 1.0,
                       const double* C0 = nullptr, double beta = 0.0) {
  auto kern = small_gemm_kernel<TA, TB>;
  const unsigned t = (unsigned)((n + 31) / 32);
  GDX_LAUNCH("scf_gemm", kern, dim3(t, t, nb), dim3(256), 0, st, n, A, sA, B, sB, S, sS, alpha, C0, beta, C, sC);
}

// F = sum_i x_{1+i} F_i per spin, summed in history order (as scf_step_kernel does).
__global__ void fock_extrap_kernel(int n, int m, int nspin, const double* __restrict__ focks,
                                   const double* __restrict__ x, double* __restrict__ F) {
  const size_t nn = (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * NT + threadIdx.x;
  const int s = blockIdx.y;
  if (e >= nn) return;
  const double* xs = x + (size_t)s * (m + 1) + 1;
  double f = 0.0;
  for (int i = 0; i < m; ++i) f += xs[i] * focks[((size_t)i * nspin + s) * nn + e];
  F[(size_t)s * nn + e] = f;
}

// err = A - A^T per spin (A n x n, row-major).
__global__ void antisym_kernel(int n, const double* __restrict__ A, double* __restrict__ err) {
  const size_t nn = (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * NT + threadIdx.x;
  const int s = blockIdx.y;
  if (e >= nn) return;
  const int i = (int)(e / n), j = (int)(e % n);
  err[s * nn + e] = A[s * nn + e] - A[s * nn + (size_t)j * n + i];
}

static bool use_tridiag(int n, size_t lds_need) { return lds_need > eig_smem_budget() && n <= tri::NMAX; }

template <int RA, int RR>
static void tri_reduce_launch(hipStream_t st, int nb, int n, const double* A, long sA, double* d, double* e,
                              double* tau, double* Vr) {
  auto kern = tri::tri_reduce_kernel<RA, RR>;
  GDX_LAUNCH("tri_reduce", kern, dim3(nb), dim3(tri::RED_THREADS), 0, st, n, A, sA, d, e, tau, Vr);
}

// Eigen-decomposition of nb symmetric matrices A[b] (stride sA; symmetrised on load) into
// ascending w (nb x n) and eigenvector columns Z (nb x n x n); ws: nb * tri::eig_ws_doubles(n).
static int tri_eigh(hipStream_t st, int nb, int n, const double* A, long sA, double* w, double* Z, double* ws,
                    int* sweeps) {
  const size_t nn = (size_t)n * n;
  double* Vr = ws;
  double* X = Vr + nb * nn;
  double* scr = X + nb * nn;
  double* d = scr + 2 * nb * nn;
  double* e = d + (size_t)nb * n;
  double* tau = e + (size_t)nb * n;
  double* tn = tau + (size_t)nb * n;
  switch ((n + 31) / 32) {
    case 1: tri_reduce_launch<1, 1>(st, nb, n, A, sA, d, e, tau, Vr); break;
    case 2: tri_reduce_launch<2, 2>(st, nb, n, A, sA, d, e, tau, Vr); break;
    case 3: tri_reduce_launch<3, 3>(st, nb, n, A, sA, d, e, tau, Vr); break;
    case 4: tri_reduce_launch<4, 4>(st, nb, n, A, sA, d, e, tau, Vr); break;
    case 5: tri_reduce_launch<5, 5>(st, nb, n, A, sA, d, e, tau, Vr); break;
    case 6: tri_reduce_launch<6, 5>(st, nb, n, A, sA, d, e, tau, Vr); break;
    case 7: tri_reduce_launch<7, 5>(st, nb, n, A, sA, d, e, tau, Vr); break;
    default: return GDX_ERR_INVALID;
  }
  GDX_LAUNCH("tri_bisect", tri::tri_bisect_kernel, dim3((n + 3) / 4, nb), dim3(256), 0, st, n, d, e, w, tn, sweeps);
  GDX_LAUNCH("tri_invit", tri::tri_invit_kernel, dim3((n + tri::INV_TPB - 1) / tri::INV_TPB, nb),
             dim3(tri::INV_TPB), 0, st, n, d, e, w, tn, X);
  GDX_LAUNCH("tri_orth", tri::tri_orth_kernel, dim3((n + 3) / 4, nb), dim3(256), 0, st, n, d, e, w, tn, X);
  // first-order Loewdin polish of the vectors computed independently (orthogonality
  // ~ eps ||T|| / gap between clusters): X <- X (I - (X^T X - I) / 2) = 1.5 X - 0.5 X (X^T X)
  double* G = scr;
  double* Xp = scr + nb * nn;
  small_gemm<true, false>(st, nb, n, X, (long)nn, X, (long)nn, nullptr, 0, G, (long)nn);
  small_gemm<false, false>(st, nb, n, X, (long)nn, G, (long)nn, nullptr, 0, Xp, (long)nn, -0.5, X, 1.5);
  GDX_LAUNCH("tri_back", tri::tri_back_kernel, dim3((n + tri::BT_COLS - 1) / tri::BT_COLS, nb), dim3(256), 0, st, n,
             Vr, tau, Xp, Z, (long)nn);
  return GDX_OK;
}

// ------------------------------------------------------------- DIIS solve ----------
// x = B^-1 e_0 for one bordered DIIS matrix B (m1 x m1, m1 <= DIIS_MAXM1) by Gaussian
// elimination with partial pivoting (the first row of largest |pivot| wins), in one wave:
// lane t < m1 holds row t and the right-hand side in registers; the pivot search, the row
// swap and the pivot row are read across lanes with v_readlane (uniform lane indices), so
// the solve needs no LDS round trips.  An exactly singular B -- a loop at a bitwise fixed
// point stores identical error vectors, whose Gram rows coincide -- gets 0 for the unknowns
// of its zero pivots (those slots drop out of the extrapolation) where an inverse would give
// non-finite coefficients.
constexpr int DIIS_MAXM1 = 16;
__device__ __forceinline__ double lane_bcast(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ void diis_solve_wave(const double* __restrict__ Bs, int m1, double* __restrict__ xs) {
  const int t = threadIdx.x;
  double r[DIIS_MAXM1 + 1];
#pragma unroll
  for (int j = 0; j <= DIIS_MAXM1; ++j) r[j] = 0.0;
#pragma unroll
  for (int j = 0; j < DIIS_MAXM1; ++j)
    if (j < m1 && t < m1) r[j] = Bs[(size_t)t * m1 + j];
  r[DIIS_MAXM1] = t == 0 ? 1.0 : 0.0;   // right-hand side e_0
#pragma unroll
  for (int c = 0; c < DIIS_MAXM1; ++c) {
    if (c < m1) {
      // pivot: first row p >= c with the largest |r[c]|
      const double ac = fabs(r[c]);
      double best = lane_bcast(ac, c);
      int p = c;
#pragma unroll
      for (int i = c + 1; i < DIIS_MAXM1; ++i) {
        if (i < m1) {
          const double v = lane_bcast(ac, i);
          if (v > best) { best = v; p = i; }
        }
      }
      p = __builtin_amdgcn_readfirstlane(p);
      if (p != c) {
#pragma unroll
        for (int j = 0; j <= DIIS_MAXM1; ++j) {
          const double vc = lane_bcast(r[j], c), vp = lane_bcast(r[j], p);
          if (t == c) r[j] = vp;
          else if (t == p) r[j] = vc;
        }
      }
      // eliminate below the pivot with the pivot row (columns j >= c and the right-hand side);
      // an exactly zero pivot (the column is already zero below it) eliminates nothing
      const double pc = lane_bcast(r[c], c);
      const double f = (t > c && t < m1 && pc != 0.0) ? r[c] / pc : 0.0;
#pragma unroll
      for (int j = c; j <= DIIS_MAXM1; ++j) {
        if (j < m1 || j == DIIS_MAXM1) {
          const double pj = lane_bcast(r[j], c);
          if (t > c && t < m1) r[j] -= f * pj;
        }
      }
    }
  }
  // back substitution, last row first
  double x[DIIS_MAXM1];
#pragma unroll
  for (int j = 0; j < DIIS_MAXM1; ++j) x[j] = 0.0;
#pragma unroll
  for (int i = DIIS_MAXM1 - 1; i >= 0; --i) {
    if (i < m1) {
      double v = r[DIIS_MAXM1];
#pragma unroll
      for (int j = i + 1; j < DIIS_MAXM1; ++j)
        if (j < m1) v -= r[j] * x[j];
      x[i] = lane_bcast(r[i] != 0.0 ? v / r[i] : 0.0, i);
    }
  }
  if (t < m1) {
    double xv = 0.0;
#pragma unroll
    for (int j = 0; j < DIIS_MAXM1; ++j)
      if (j == t) xv = x[j];
    xs[t] = xv;
  }
}

// x_s = B_s^-1 e_0 for each spin: one wave per spin.
__global__ __launch_bounds__(64) void diis_solve_kernel(const double* __restrict__ B, int m1,
                                                        double* __restrict__ x) {
  const int s = blockIdx.x;
  diis_solve_wave(B + (size_t)s * m1 * m1, m1, x + (size_t)s * m1);
}

// JittableDiis coefficients (evaluate.py:1185-1205), two launches:
//  diis_gram_kernel: block (pair (i, j), spin s) = <err_i, err_j> over the n2 elements of
//    errs (m, nspin, n2), summed per wave then over the waves in a fixed order, written to
//    B_s[1 + i][1 + j] and B_s[1 + j][1 + i];
//  diis_border_solve_kernel: the borders (+1 for the slots i <= cycle, B[0][0] = 0) and 1 on
//    the diagonal of the inactive slots, B written back, then x = B^-1 e_0.
constexpr int DIIS_THREADS = 256;
__global__ __launch_bounds__(DIIS_THREADS) void diis_gram_kernel(const double* __restrict__ errs, int m, int nspin,
                                                                 long n2, double* __restrict__ B) {
  __shared__ double wsum[DIIS_THREADS / 64];
  const int pi = blockIdx.x, s = blockIdx.y, t = threadIdx.x;
  int i = 0, rem = pi;
  while (rem >= m - i) { rem -= m - i; ++i; }
  const int j = i + rem;
  const size_t stride = (size_t)nspin * n2;
  const double* Ei = errs + (size_t)i * stride + (size_t)s * n2;
  const double* Ej = errs + (size_t)j * stride + (size_t)s * n2;
  double v = 0.0;
  for (long e = t; e < n2; e += DIIS_THREADS) v += Ei[e] * Ej[e];
  v = warp_sum(v);
  if ((t & 63) == 0) wsum[t >> 6] = v;
  __syncthreads();
  if (t == 0) {
    double g = 0.0;
    for (int w = 0; w < DIIS_THREADS / 64; ++w) g += wsum[w];
    const int m1 = m + 1;
    double* Bs = B + (size_t)s * m1 * m1;
    Bs[(size_t)(1 + i) * m1 + 1 + j] = g;
    Bs[(size_t)(1 + j) * m1 + 1 + i] = g;
  }
}

__global__ __launch_bounds__(64) void diis_border_solve_kernel(int m, int cycle, double* __restrict__ B,
                                                               double* __restrict__ x) {
  const int s = blockIdx.x, t = threadIdx.x, m1 = m + 1;
  double* Bs = B + (size_t)s * m1 * m1;
  if (t == 0) Bs[0] = 0.0;
  if (t < m) {
    const double act = t <= cycle ? 1.0 : 0.0;
    Bs[1 + t] = act;
    Bs[(size_t)(1 + t) * m1] = act;
    if (t > cycle) Bs[(size_t)(1 + t) * m1 + 1 + t] = 1.0;
  }
  __syncthreads();
  diis_solve_wave(Bs, m1, x + (size_t)s * m1);
}

inline unsigned nblk(long n) { return (unsigned)((n + NT - 1) / NT); }

}  // namespace
}  // namespace gdx

using namespace gdx;

extern "C" {

inline int coulomb_tile(long P) { return P >= 4096 ? JT : JTS; }

int gdx_coulomb_workspace_size(int32_t n, size_t* bytes) {
  GDX_REQUIRE(n > 0 && bytes, "gdx_coulomb_workspace_size: bad arguments");
  const long P = (long)n * (n + 1) / 2;
  const long nb = (P + coulomb_tile(P) - 1) / coulomb_tile(P);
  *bytes = (size_t)(P + 2 * nb * P) * sizeof(double);
  return GDX_OK;
}

}  // extern "C"

namespace gdx {
namespace {

long coulomb_ws_doubles(long n) {
  const long P = n * (n + 1) / 2;
  const long nb = (P + coulomb_tile(P) - 1) / coulomb_tile(P);
  return P + 2 * nb * P;
}

int coulomb_launch(hipStream_t st, int32_t n, const double* eri, const double* rdm1, double* J, double* ws) {
  const long P = (long)n * (n + 1) / 2;
  const int tt = coulomb_tile(P);
  const long nb = (P + tt - 1) / tt;
  double* Dt = ws;
  double* prow = Dt + P;
  double* pcol = prow + nb * P;
  GDX_LAUNCH("pack_density", pack_density_kernel, dim3(nblk(P)), dim3(NT), 0, st, rdm1, n, Dt);
  const dim3 grid((unsigned)(nb * (nb + 1) / 2));
  if (tt == JT && P >= J_ROWS_MIN)
    GDX_LAUNCH("coulomb_tile", coulomb_rows_kernel, grid, dim3(NT), 0, st, eri, P, Dt, prow, pcol);
  else if (tt == JT)
    GDX_LAUNCH("coulomb_tile", coulomb_tile_kernel<JT>, grid, dim3(NT), 0, st, eri, P, Dt, prow, pcol);
  else
    GDX_LAUNCH("coulomb_tile", coulomb_tile_kernel<JTS>, grid, dim3(NT), 0, st, eri, P, Dt, prow, pcol);
  GDX_LAUNCH("coulomb_finish", coulomb_finish_kernel, dim3(nblk(P)), dim3(NT), 0, st, prow, pcol, P, (int)nb, tt, n,
             J);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

// Scratch of gdx_coulomb_packed_noalloc: a fixed device array, one copy per GPU, held until
// the process ends.  Host code never passes `coulomb_scratch` itself to a kernel (on the host
// the name is only a stand-in); it uses the per-device address hipGetSymbolAddress returns.
constexpr long COULOMB_SCRATCH_DOUBLES = 4L << 20;  // 32 MiB
static __device__ __align__(256) double coulomb_scratch[COULOMB_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// Address of this device's scratch array (looked up once per device, with it current).
int scratch_for_current_device(double** out) {
  int dev = -1;
  GDX_HIP_CHECK(hipGetDevice(&dev));
  GDX_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "gdx_coulomb_packed_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    GDX_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(coulomb_scratch)));
    GDX_REQUIRE(p != nullptr, "gdx_coulomb_packed_noalloc: scratch symbol has no device address");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return GDX_OK;
}

}  // namespace
}  // namespace gdx

extern "C" {

int gdx_coulomb_packed(void* stream, int32_t n, const double* eri, const double* rdm1, double* J, void* workspace) {
  GDX_REQUIRE(n > 0 && eri && rdm1 && J && workspace, "gdx_coulomb_packed: bad arguments");
  return coulomb_launch((hipStream_t)stream, n, eri, rdm1, J, (double*)workspace);
}

int gdx_coulomb_noalloc_max_n(int32_t* n_max) {
  GDX_REQUIRE(n_max, "gdx_coulomb_noalloc_max_n: bad arguments");
  int32_t n = 1;
  while (coulomb_ws_doubles(n + 1) <= COULOMB_SCRATCH_DOUBLES) ++n;
  *n_max = n;
  return GDX_OK;
}

int gdx_coulomb_packed_noalloc(void* stream, int32_t n, const double* eri, const double* rdm1, double* J) {
  GDX_REQUIRE(n > 0 && eri && rdm1 && J, "gdx_coulomb_packed_noalloc: bad arguments");
  if (coulomb_ws_doubles(n) > COULOMB_SCRATCH_DOUBLES) {
    int32_t n_max = 0;
    gdx_coulomb_noalloc_max_n(&n_max);
    set_error("gdx_coulomb_packed_noalloc: n = " + std::to_string(n) + " needs " +
              std::to_string(coulomb_ws_doubles(n) * 8) + " bytes of scratch; the fixed scratch holds " +
              std::to_string(COULOMB_SCRATCH_DOUBLES * 8) + " (n <= " + std::to_string(n_max) +
              "); use gdx_coulomb_packed with a workspace");
    return GDX_ERR_INVALID;
  }
  double* ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc != GDX_OK) return rc;
  return coulomb_launch((hipStream_t)stream, n, eri, rdm1, J, ws);
}

int gdx_nonxc_energy(void* stream, int32_t n, const double* rdm1, const double* h1e, const double* J,
                     const double* K0, double a0, double e_nuc, double* out) {
  GDX_REQUIRE(n > 0 && rdm1 && h1e && J && out, "gdx_nonxc_energy: bad arguments");
  GDX_LAUNCH("nonxc_energy", nonxc_energy_kernel, dim3(1), dim3(NT), 0, (hipStream_t)stream, rdm1, h1e, J, K0, a0, n,
             e_nuc, out);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_fock_assemble(void* stream, int32_t n, const double* h1e, const double* J, const double* vxc,
                      const double* K0, double a0, double* fock) {
  GDX_REQUIRE(n > 0 && h1e && J && vxc && fock, "gdx_fock_assemble: bad arguments");
  GDX_LAUNCH("fock_assemble", fock_assemble_kernel, dim3(nblk(2L * n * n)), dim3(NT), 0, (hipStream_t)stream, h1e, J,
             vxc, K0, a0, n, fock);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_exchange_packed(void* stream, int32_t n, int32_t nspin, const double* eri, const double* D, double* K) {
  GDX_REQUIRE(n > 0 && nspin > 0 && eri && D && K, "gdx_exchange_packed: bad arguments");
  GDX_REQUIRE(D != K, "gdx_exchange_packed: K must not alias D");
  hipStream_t st = (hipStream_t)stream;
  const long waves = (long)nspin * n * n;
  GDX_LAUNCH("exchange", exchange_gather_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, eri, n, nspin,
             D, K);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_exchange_xo_size(int32_t n, size_t* bytes) {
  GDX_REQUIRE(n > 0 && n <= 2048 && bytes, "gdx_exchange_xo_size: bad arguments");
  const long P = (long)n * n;
  *bytes = (size_t)(P * (P + 1) / 2) * sizeof(double);
  return GDX_OK;
}

int gdx_exchange_xo_build(void* stream, int32_t n, const double* eri, double* xo) {
  GDX_REQUIRE(n > 0 && n <= 2048 && eri && xo, "gdx_exchange_xo_build: bad arguments");
  GDX_LAUNCH("exchange_xo_build", xo_build_kernel, dim3((unsigned)((long)n * n)), dim3(NT), 0, (hipStream_t)stream,
             eri, n, xo);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_exchange_xo_workspace_size(int32_t n, size_t* bytes) {
  GDX_REQUIRE(n > 0 && n <= 2048 && bytes, "gdx_exchange_xo_workspace_size: bad arguments");
  const long P = (long)n * n;
  const long nb = (P + coulomb_tile(P) - 1) / coulomb_tile(P);
  *bytes = (size_t)(4 * nb * P) * sizeof(double);   // row and column partials of two spins
  return GDX_OK;
}

int gdx_exchange_xo(void* stream, int32_t n, int32_t nspin, const double* xo, const double* D, double* K,
                    void* workspace) {
  GDX_REQUIRE(n > 0 && n <= 2048 && nspin > 0 && xo && D && K && workspace, "gdx_exchange_xo: bad arguments");
  GDX_REQUIRE(D != K, "gdx_exchange_xo: K must not alias D");
  hipStream_t st = (hipStream_t)stream;
  const long P = (long)n * n;
  const int tt = coulomb_tile(P);
  const long nb = (P + tt - 1) / tt;
  double* prow = (double*)workspace;
  double* pcol = prow + 2 * nb * P;
  const dim3 grid((unsigned)(nb * (nb + 1) / 2));
  int s0 = 0;
  if (nspin >= 2 && tt == JT && P >= J_ROWS_MIN) {   // the large case: both spins in one pass
    GDX_LAUNCH("exchange_xo", xo_rows2_kernel, grid, dim3(NT), 0, st, xo, P, D, D + P, prow, pcol, nb * P);
    for (int s = 0; s < 2; ++s)
      GDX_LAUNCH("exchange_xo_finish", xo_finish_kernel, dim3(nblk(P)), dim3(NT), 0, st, prow + s * nb * P,
                 pcol + s * nb * P, P, (int)nb, tt, K + (size_t)s * P);
    s0 = 2;
  }
  for (int s = s0; s < nspin; ++s) {
    const double* Ds = D + (size_t)s * P;
    if (tt == JT && P >= J_ROWS_MIN)
      GDX_LAUNCH("exchange_xo", coulomb_rows_kernel, grid, dim3(NT), 0, st, xo, P, Ds, prow, pcol);
    else if (tt == JT)
      GDX_LAUNCH("exchange_xo", coulomb_tile_kernel<JT>, grid, dim3(NT), 0, st, xo, P, Ds, prow, pcol);
    else
      GDX_LAUNCH("exchange_xo", coulomb_tile_kernel<JTS>, grid, dim3(NT), 0, st, xo, P, Ds, prow, pcol);
    GDX_LAUNCH("exchange_xo_finish", xo_finish_kernel, dim3(nblk(P)), dim3(NT), 0, st, prow, pcol, P, (int)nb, tt,
               K + (size_t)s * P);
  }
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_cholesky_inv(void* stream, int32_t n, const double* S, double* L, double* Linv, int32_t* status) {
  GDX_REQUIRE(n > 0 && S && L && Linv && status, "gdx_cholesky_inv: bad arguments");
  GDX_LAUNCH("cholesky_inv", cholesky_inv_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, S, n, L, Linv, status);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_eigh_workspace_size(int32_t nbatch, int32_t n, size_t* bytes) {
  GDX_REQUIRE(nbatch > 0 && n > 0 && bytes, "gdx_eigh_workspace_size: bad arguments");
  size_t per = (size_t)3 * n * n;
  if (n <= tri::NMAX) per = std::max(per, tri::eig_ws_doubles(n));
  *bytes = (size_t)nbatch * per * sizeof(double);
  return GDX_OK;
}

int gdx_eigh_batched(void* stream, int32_t nbatch, int32_t n, const double* A, const double* V0, double* w,
                     double* V, void* workspace, int32_t max_sweeps, double tol, int32_t* sweeps) {
  GDX_REQUIRE(nbatch > 0 && n > 0 && n <= 1022 && A && w && V, "gdx_eigh_batched: bad arguments");
  GDX_REQUIRE(V0 == nullptr || (V0 != V && V0 != A), "gdx_eigh_batched: V0 must not alias A or V");
  hipStream_t st = (hipStream_t)stream;
  const size_t smem = (size_t)3 * n * n * sizeof(double);
  if (use_tridiag(n, smem)) {   // V0 is not used: the result does not depend on a start
    GDX_REQUIRE(workspace, "gdx_eigh_batched: workspace required when 3 n^2 doubles exceed the LDS budget");
    const int rc = tri_eigh(st, nbatch, n, A, (long)n * n, w, V, (double*)workspace, sweeps);
    if (rc != GDX_OK) return rc;
  } else if (smem <= eig_smem_budget()) {
    auto kern = jacobi_eigh_kernel<true, EIG_THREADS_SMEM>;
    if (smem > EIG_SMEM_MAX)
      GDX_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    GDX_LAUNCH("jacobi_eigh", kern, dim3(nbatch), dim3(EIG_THREADS_SMEM), smem, st, A, n, V0, w, V,
               (double*)workspace, max_sweeps, tol, sweeps, 0);
  } else {
    GDX_REQUIRE(workspace, "gdx_eigh_batched: workspace required when 3 n^2 doubles exceed the LDS budget");
    auto kern = jacobi_eigh_kernel<false, EIG_THREADS_GLOBAL>;
    const size_t dyn = a_lds_bytes(n);
    if (dyn > EIG_SMEM_MAX)
      GDX_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    GDX_LAUNCH("jacobi_eigh", kern, dim3(nbatch), dim3(EIG_THREADS_GLOBAL), dyn, st, A, n, V0, w, V,
               (double*)workspace, max_sweeps, tol, sweeps, dyn > 0 ? 1 : 0);
  }
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_scf_step_workspace_size(int32_t nspin, int32_t n, size_t* bytes) {
  GDX_REQUIRE(nspin > 0 && n > 0 && bytes, "gdx_scf_step_workspace_size: bad arguments");
  size_t per = (size_t)4 * n * n;
  if (n <= tri::NMAX) per = std::max(per, (size_t)2 * n * n + tri::eig_ws_doubles(n));
  *bytes = (size_t)nspin * per * sizeof(double);
  return GDX_OK;
}

int gdx_scf_step(void* stream, int32_t nspin, int32_t n, int32_t m, const double* focks, const double* x,
                 const double* Linv, const double* V0, const double* occ, double* F, double* w, double* V,
                 double* C, double* rdm1, void* workspace, int32_t max_sweeps, double tol, int32_t* sweeps) {
  GDX_REQUIRE(nspin > 0 && nspin <= 64 && n > 0 && n <= 1022 && m > 0 && focks && x && Linv && occ && F && w && V &&
                  C && rdm1,
              "gdx_scf_step: bad arguments");
  GDX_REQUIRE(V0 == nullptr || (V0 != V && V0 != C && V0 != rdm1), "gdx_scf_step: V0 must not alias an output");
  hipStream_t st = (hipStream_t)stream;
  const size_t smem = (size_t)4 * n * n * sizeof(double);
  if (use_tridiag(n, smem)) {
    // F = sum x_i F_i; B = Linv F Linv^T; B = V diag(w) V^T; C = Linv^T V; D = C diag(occ) C^T
    GDX_REQUIRE(workspace, "gdx_scf_step: workspace required when 4 n^2 doubles exceed the LDS budget");
    const long nn = (long)n * n;
    double* T1 = (double*)workspace;
    double* B = T1 + nspin * nn;
    double* ews = B + nspin * nn;
    GDX_LAUNCH("fock_extrap", fock_extrap_kernel, dim3((unsigned)((nn + NT - 1) / NT), nspin), dim3(NT), 0, st, n, m,
               nspin, focks, x, F);
    small_gemm<false, false>(st, nspin, n, Linv, 0, F, nn, nullptr, 0, T1, nn);
    small_gemm<false, true>(st, nspin, n, T1, nn, Linv, 0, nullptr, 0, B, nn);
    const int rc = tri_eigh(st, nspin, n, B, nn, w, V, ews, sweeps);
    if (rc != GDX_OK) return rc;
    small_gemm<true, false>(st, nspin, n, Linv, 0, V, nn, nullptr, 0, C, nn);
    small_gemm<false, true>(st, nspin, n, C, nn, C, nn, occ, n, rdm1, nn);
  } else if (smem <= eig_smem_budget()) {
    auto kern = scf_step_kernel<true, EIG_THREADS_SMEM>;
    if (smem > EIG_SMEM_MAX)
      GDX_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    GDX_LAUNCH("scf_step", kern, dim3(nspin), dim3(EIG_THREADS_SMEM), smem, st, n, m, nspin, focks, x, Linv, V0, occ,
               F, w, V, C, rdm1, (double*)workspace, max_sweeps, tol, sweeps, 0);
  } else {
    GDX_REQUIRE(workspace, "gdx_scf_step: workspace required when 4 n^2 doubles exceed the LDS budget");
    auto kern = scf_step_kernel<false, EIG_THREADS_GLOBAL>;
    const size_t dyn = a_lds_bytes(n);
    if (dyn > EIG_SMEM_MAX)
      GDX_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    GDX_LAUNCH("scf_step", kern, dim3(nspin), dim3(EIG_THREADS_GLOBAL), dyn, st, n, m, nspin, focks, x, Linv, V0,
               occ, F, w, V, C, rdm1, (double*)workspace, max_sweeps, tol, sweeps, dyn > 0 ? 1 : 0);
  }
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_diis_error_workspace_size(int32_t nspin, int32_t n, size_t* bytes) {
  GDX_REQUIRE(nspin > 0 && n > 0 && bytes, "gdx_diis_error_workspace_size: bad arguments");
  *bytes = (size_t)2 * nspin * n * n * sizeof(double);
  return GDX_OK;
}

int gdx_diis_error(void* stream, int32_t nspin, int32_t n, const double* F, const double* D, const double* S,
                   double* err, void* workspace) {
  GDX_REQUIRE(nspin > 0 && nspin <= 64 && n > 0 && F && D && S && err && workspace, "gdx_diis_error: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const long nn = (long)n * n;
  double* T = (double*)workspace;
  double* A = T + nspin * nn;
  small_gemm<false, false>(st, nspin, n, F, nn, D, nn, nullptr, 0, T, nn);   // T = F D
  small_gemm<false, false>(st, nspin, n, T, nn, S, 0, nullptr, 0, A, nn);    // A = T S
  GDX_LAUNCH("diis_error", antisym_kernel, dim3((unsigned)((nn + NT - 1) / NT), nspin), dim3(NT), 0, st, n, A, err);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_diis_coefficients(void* stream, int32_t nspin, int32_t m, int64_t n2, const double* errs, int32_t cycle,
                          double* B, double* x) {
  GDX_REQUIRE(nspin > 0 && nspin <= 64 && m > 0 && m + 1 <= DIIS_MAXM1 && n2 > 0 && errs && B && x,
              "gdx_diis_coefficients: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  GDX_LAUNCH("diis_gram", diis_gram_kernel, dim3(m * (m + 1) / 2, nspin), dim3(DIIS_THREADS), 0, st, errs, m, nspin,
             (long)n2, B);
  GDX_LAUNCH("diis_solve", diis_border_solve_kernel, dim3(nspin), dim3(64), 0, st, m, cycle, B, x);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_scf_status(void* stream, int32_t nspin, int32_t n, const double* rdm1, const double* S, const double* occ,
                   const int32_t* sweeps, int32_t max_sweeps, const double* energy, double tol, int32_t* status) {
  GDX_REQUIRE(nspin > 0 && n > 0 && rdm1 && S && occ && status, "gdx_scf_status: bad arguments");
  GDX_LAUNCH("scf_status", scf_status_kernel, dim3(1), dim3(NT), 0, (hipStream_t)stream, nspin, n, rdm1, S, occ,
             sweeps, max_sweeps, energy, tol, status);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_diis_solve(void* stream, int32_t nspin, int32_t m1, const double* B, double* x) {
  GDX_REQUIRE(nspin > 0 && nspin <= 64 && m1 > 0 && m1 <= DIIS_MAXM1 && B && x, "gdx_diis_solve: bad arguments");
  GDX_LAUNCH("diis_solve", diis_solve_kernel, dim3(nspin), dim3(64), 0, (hipStream_t)stream, B, m1, x);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

}  // extern "C"
