// Semi-numerical exact exchange on the grid.
//
// Reference computations replaced:
//   nu_bd(r) = int dr' phi_b(r') f(|r - r'|) phi_d(r')   PySCF int1e_grids_sph, as called by
//       grad_dft/external/_hf_density.py:18-67 (f = 1/|r-r'|, or erf(omega |r-r'|)/|r-r'| under
//       mol.with_range_coulomb(omega))
//   chi[r, omega, s, a] = sum_bd D_s[b][d] phi_b(r) nu_da(r)  generate_chi_tensor,
//       grad_dft/interface/pyscf.py:1067-1126 (chi_make)
//   e_HF[omega, s, r] = -1/2 sum_ac chi[r,omega,s,c] D_s[a][c] phi_a(r)   HF_energy_density,
//       grad_dft/molecule.py:507-541
//   V[omega, s, a, b] = -1/2 sum_r gr[omega,s,r] phi_a(r) chi[r,omega,s,b]  HF_density_grad_2_Fock,
//       grad_dft/molecule.py:545-613
//
// nu integrals by the McMurchie-Davidson scheme: for a primitive pair (exponents a, b on centres
// A, B; p = a + b, P = (a A + b B) / p)
//   nu_ij(C) = (2 pi / p) sum_tuv E^x_{i_x j_x t} E^y_{i_y j_y u} E^z_{i_z j_z v} R_tuv(p, P - C)
// with the Hermite Coulomb integrals R_tuv built from Boys functions F_n(p |P - C|^2).  The
// attenuated kernel erf(omega r)/r replaces p by p' = p omega^2 / (p + omega^2) inside R and
// scales by sqrt(p' / p) (the Boys integral's upper limit moves from 1 to sqrt(p' / p)).
//
// Everything that does not depend on the grid point -- the Hermite expansion coefficients,
// contraction coefficients, 2 pi / p and the Cartesian-to-AO component matrices -- is folded on
// the host, per primitive pair of an ordered shell pair (A, B), into one matrix
//   M[a][b][tuv]   (a: AO component of A, b: of B, tuv: Hermite index with t + u + v <= la + lb)
// so the kernel evaluates, per grid point and primitive pair, only R_tuv and the contraction
//   chi_b(r) += sum_a u_a(r) sum_tuv M[a][b][tuv] R_tuv(r),   u(r) = D0^T phi(r).
// Work is split by L = la + lb (one launch per L, register arrays sized for that L); a lane owns
// one grid point and one output shell B and loops over the primitive pairs of every shell A
// paired with B.  Each chi element is written by one lane per launch and the launches run in a
// fixed order, so the result is bitwise reproducible (no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gdx_common.h"
#include "gdx_md.h"
#include "gdx_profile.h"

namespace gdx {
namespace {

constexpr int HF_LMAX = 3;                 // f shells
constexpr int HF_LTOT = 2 * HF_LMAX;       // la + lb
constexpr int HF_MAXF = 10;                // AO components of a shell (Cartesian f)
constexpr int HF_NSPIN_MAX = 2;
using md::n_tuv;
using md::tuv_index;
constexpr int BOYS_NCOL = HF_LTOT + md::BOYS_TAYLOR;  // orders 0..12

// One launch per L.  grid.x: 256-point tiles; grid.y: output shells B that pair with some shell
// A at la + lb = L.  bl[y] = {B ao offset, B components, first entry, end entry}.  Entry e:
// ef[4e..] = {p, Px, Py, Pz}, ei[2e..] = {A ao offset, A components}, em[e] = offset of its
// M block (nfA x nfB x n_tuv(L)).  chi += (accumulated over launches; zeroed beforehand).
template <int L, int NS>
__global__ __launch_bounds__(256) void hf_chi_kernel(int npts, int npts_pad, const double* __restrict__ coords,
                                                     const int4* __restrict__ bl, const double* __restrict__ ef,
                                                     const int* __restrict__ ei, const long* __restrict__ em,
                                                     const double* __restrict__ M, double omega,
                                                     const double* __restrict__ boys, int nao,
                                                     const double* __restrict__ u, double* __restrict__ chi) {
  constexpr int NT = n_tuv(L);
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npts) return;
  const int4 b = bl[blockIdx.y];
  const int boff = b.x, nfB = b.y;
  const double cx = coords[3 * r], cy = coords[3 * r + 1], cz = coords[3 * r + 2];
  const size_t splane = (size_t)nao * npts_pad;
  double acc[NS][HF_MAXF];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int j = 0; j < HF_MAXF; ++j) acc[s][j] = 0.0;
  for (int e = b.z; e < b.w; ++e) {
    const double p = ef[4 * e];
    const double X = ef[4 * e + 1] - cx, Y = ef[4 * e + 2] - cy, Z = ef[4 * e + 3] - cz;
    double pe = p, scale = 1.0;
    if (omega > 0.0) {
      const double o2 = omega * omega;
      pe = p * o2 / (p + o2);
      scale = sqrt(o2 / (p + o2));
    }
    double F[L + 1];
    md::boys_eval<L, BOYS_NCOL>(pe * (X * X + Y * Y + Z * Z), boys, F);
    double R[NT];
    md::hermite_coulomb<L>(pe, X, Y, Z, F, R);
#pragma unroll
    for (int q = 0; q < NT; ++q) R[q] *= scale;
    const int aoff = ei[2 * e], nfA = ei[2 * e + 1];
    const double* Me = M + em[e];
    for (int a = 0; a < nfA; ++a) {
      double ua[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) ua[s] = u[s * splane + (size_t)(aoff + a) * npts_pad + r];
      const double* Ma = Me + (size_t)a * nfB * NT;
#pragma unroll
      for (int j = 0; j < HF_MAXF; ++j) {
        if (j < nfB) {
          const double* Mab = Ma + j * NT;
          double sab = 0.0;
#pragma unroll
          for (int q = 0; q < NT; ++q) sab += Mab[q] * R[q];
#pragma unroll
          for (int s = 0; s < NS; ++s) acc[s][j] += ua[s] * sab;
        }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int j = 0; j < HF_MAXF; ++j)
      if (j < nfB) chi[s * splane + (size_t)(boff + j) * npts_pad + r] += acc[s][j];
}

__global__ void hf_zero_kernel(double* __restrict__ p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// out[s][c][r] = sum_a phi[r][a] M_s[a][c] for r < npts (zero for npts <= r < npts_pad).
// 64 points x 64 columns a block (256 threads: 64 points x 4 column groups of 16); phi and
// M staged through LDS 16 orbitals at a time.
constexpr int MM_P = 64, MM_C = 64, MM_K = 16;
__global__ __launch_bounds__(256) void ao_matmul_kernel(int npts, int npts_pad, int nao, int nao_pad,
                                                        const double* __restrict__ phi, const double* __restrict__ Mm,
                                                        double* __restrict__ out) {
  __shared__ double ps[MM_K][MM_P + 1];
  __shared__ double ms[MM_K][MM_C];
  const int s = blockIdx.z;
  const int r0 = blockIdx.x * MM_P, c0 = blockIdx.y * MM_C;
  const int tid = threadIdx.x, pt = tid & 63, cg = tid >> 6;
  const double* Ms = Mm + (size_t)s * nao * nao;
  double acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0;
  for (int k0 = 0; k0 < nao; k0 += MM_K) {
    // phi tile: 64 points x 16 orbitals (each thread 4 elements)
    for (int i = tid; i < MM_P * MM_K; i += 256) {
      const int pr = i / MM_K, kk = i % MM_K;
      const int r = r0 + pr, a = k0 + kk;
      ps[kk][pr] = (r < npts && a < nao) ? phi[(size_t)r * nao_pad + a] : 0.0;
    }
    for (int i = tid; i < MM_K * MM_C; i += 256) {
      const int kk = i / MM_C, cc = i % MM_C;
      const int a = k0 + kk, c = c0 + cc;
      ms[kk][cc] = (a < nao && c < nao) ? Ms[(size_t)a * nao + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < MM_K; ++kk) {
      const double pv = ps[kk][pt];
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[j] += pv * ms[kk][cg * 16 + j];
    }
    __syncthreads();
  }
  const int r = r0 + pt;
  if (r < npts_pad) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int c = c0 + cg * 16 + j;
      if (c < nao) out[((size_t)s * nao + c) * npts_pad + r] = r < npts ? acc[j] : 0.0;
    }
  }
}

// e[s][r] = -1/2 sum_c chi[s][c][r] v[s][c][r]
__global__ void hf_rowdot_kernel(int npts, int npts_pad, int nao, int nspin, const double* __restrict__ chi,
                                 const double* __restrict__ v, double* __restrict__ e) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nspin * npts_pad) return;
  const int s = (int)(i / npts_pad), r = (int)(i % npts_pad);
  double acc = 0.0;
  if (r < npts) {
    const size_t base = (size_t)s * nao * npts_pad + r;
    for (int c = 0; c < nao; ++c) acc += chi[base + (size_t)c * npts_pad] * v[base + (size_t)c * npts_pad];
  }
  e[i] = -0.5 * acc;
}

// Partial sums part[chunk][s][a][b] = sum_{r in chunk} g_s(r) phi[r][a] chi[s][b][r]; 32 x 32
// output tile a block (256 threads, 2 x 2 outputs each), 32 points staged per step.
constexpr int KM_T = 32, KM_R = 32;
__global__ __launch_bounds__(256) void hf_kmat_kernel(int npts, int npts_pad, int nao, int nao_pad,
                                                      const double* __restrict__ phi, const double* __restrict__ chi,
                                                      const double* __restrict__ g, long g_spin_stride, int nchunk,
                                                      int chunk_len, double* __restrict__ part) {
  __shared__ double as[KM_R][KM_T + 1];
  __shared__ double bs[KM_R][KM_T + 1];
  const int ntile = (nao + KM_T - 1) / KM_T;
  const int tile = blockIdx.x % (ntile * ntile);
  const int chunk = blockIdx.x / (ntile * ntile);
  const int s = blockIdx.y;
  const int a0 = (tile / ntile) * KM_T, b0 = (tile % ntile) * KM_T;
  const int tid = threadIdx.x, ta = tid >> 4, tb = tid & 15;
  const int rbeg = chunk * chunk_len, rend = min(npts, rbeg + chunk_len);
  const double* chis = chi + (size_t)s * nao * npts_pad;
  const double* gs = g + (size_t)s * g_spin_stride;
  double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int r0 = rbeg; r0 < rend; r0 += KM_R) {
    for (int i = tid; i < KM_R * KM_T; i += 256) {
      const int rr = i / KM_T, aa = i % KM_T;
      const int r = r0 + rr, a = a0 + aa;
      as[rr][aa] = (r < rend && a < nao) ? phi[(size_t)r * nao_pad + a] : 0.0;
    }
    for (int i = tid; i < KM_R * KM_T; i += 256) {
      const int bb = i / KM_R, rr = i % KM_R;
      const int r = r0 + rr, b = b0 + bb;
      bs[rr][bb] = (r < rend && b < nao) ? gs[r] * chis[(size_t)b * npts_pad + r] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < KM_R; ++rr) {
      const double x0 = as[rr][ta], x1 = as[rr][ta + 16];
      const double y0 = bs[rr][tb], y1 = bs[rr][tb + 16];
      acc[0][0] += x0 * y0;
      acc[0][1] += x0 * y1;
      acc[1][0] += x1 * y0;
      acc[1][1] += x1 * y1;
    }
    __syncthreads();
  }
  const int nspin = gridDim.y;
  double* out = part + ((size_t)chunk * nspin + s) * nao * nao;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int a = a0 + ta + 16 * i, b = b0 + tb + 16 * j;
      if (a < nao && b < nao) out[(size_t)a * nao + b] = acc[i][j];
    }
}

__global__ void hf_kmat_reduce_kernel(const double* __restrict__ part, int nchunk, long len, double scale,
                                      double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double acc = 0.0;
  for (int c = 0; c < nchunk; ++c) acc += part[(size_t)c * len + i];
  out[i] = scale * acc;
}

inline unsigned nblk256(long n) { return (unsigned)((n + 255) / 256); }

// ---------------------------------------------------------------- host: plan ----------
struct HFPlan {
  int nshell = 0, nao = 0;
  int nb[HF_LTOT + 1] = {};   // output shells per L
  int4* d_bl[HF_LTOT + 1] = {};
  double* d_ef = nullptr;
  int* d_ei = nullptr;
  long* d_em = nullptr;
  double* d_M = nullptr;
  double* d_boys = nullptr;
  long nentries = 0, msize = 0;
  double screen = 0.0;
  int device = -1;
};

void plan_free(HFPlan* p) {
  if (!p) return;
  for (int L = 0; L <= HF_LTOT; ++L)
    if (p->d_bl[L]) (void)hipFree(p->d_bl[L]);
  if (p->d_ef) (void)hipFree(p->d_ef);
  if (p->d_ei) (void)hipFree(p->d_ei);
  if (p->d_em) (void)hipFree(p->d_em);
  if (p->d_M) (void)hipFree(p->d_M);
  if (p->d_boys) (void)hipFree(p->d_boys);
  delete p;
}

template <int L, int NS>
void launch_chi(hipStream_t st, const HFPlan* P, int npts, int npts_pad, const double* coords, double omega,
                const double* u, double* chi) {
  if (!P->nb[L]) return;
  const dim3 grid(nblk256(npts), (unsigned)P->nb[L]);
  GDX_LAUNCH("hf_chi", (hf_chi_kernel<L, NS>), grid, dim3(256), 0, st, npts, npts_pad, coords, P->d_bl[L], P->d_ef,
             P->d_ei, P->d_em, P->d_M, omega, P->d_boys, P->nao, u, chi);
}

template <int NS>
void launch_chi_all(hipStream_t st, const HFPlan* P, int npts, int npts_pad, const double* coords, double omega,
                    const double* u, double* chi) {
  launch_chi<0, NS>(st, P, npts, npts_pad, coords, omega, u, chi);
  launch_chi<1, NS>(st, P, npts, npts_pad, coords, omega, u, chi);
  launch_chi<2, NS>(st, P, npts, npts_pad, coords, omega, u, chi);
  launch_chi<3, NS>(st, P, npts, npts_pad, coords, omega, u, chi);
  launch_chi<4, NS>(st, P, npts, npts_pad, coords, omega, u, chi);
  launch_chi<5, NS>(st, P, npts, npts_pad, coords, omega, u, chi);
  launch_chi<6, NS>(st, P, npts, npts_pad, coords, omega, u, chi);
}

// kmat split: about 2048 blocks, chunks of whole 32-point steps
void kmat_split(int npts, int nao, int nspin, int* nchunk, int* chunk_len) {
  const int ntile = (nao + KM_T - 1) / KM_T;
  const int tiles = ntile * ntile * nspin;
  int nc = std::max(1, std::min(2048 / std::max(tiles, 1), (npts + 255) / 256));
  int len = (npts + nc - 1) / nc;
  len = (len + KM_R - 1) / KM_R * KM_R;
  nc = (npts + len - 1) / len;
  *nchunk = std::max(nc, 1);
  *chunk_len = len;
}

}  // namespace
}  // namespace gdx

using namespace gdx;

extern "C" {

int gdx_hf_plan_create(int32_t nshell, const int32_t* shell_i32, const double* shell_f64, const double* prim_f64,
                       const double* tmat_f64, int32_t nao, double screen, void** plan_out) {
  GDX_REQUIRE(plan_out && nshell > 0 && nao > 0 && shell_i32 && shell_f64 && prim_f64 && tmat_f64,
              "gdx_hf_plan_create: bad arguments");
  GDX_REQUIRE(screen >= 0.0, "gdx_hf_plan_create: negative screening threshold");
  *plan_out = nullptr;
  std::vector<md::Shell> sh(nshell);
  for (int s = 0; s < nshell; ++s) {
    sh[s] = md::shell_view(shell_i32, shell_f64, s);
    const md::Shell& x = sh[s];
    GDX_REQUIRE(x.l >= 0 && x.l <= HF_LMAX && x.nfunc > 0 && x.nfunc <= HF_MAXF && x.aoff + x.nfunc <= nao,
                "gdx_hf_plan_create: shell with l > 3 or components outside the basis");
  }
  // entries per (L, B): every ordered pair (A, B) and primitive pair above the threshold
  struct Ent {
    double p, P[3];
    int aoff, nfA;
    long moff;
  };
  std::vector<Ent> ents;
  std::vector<double> Mall;
  std::vector<int4> bl[HF_LTOT + 1];
  for (int L = 0; L <= HF_LTOT; ++L) {
    for (int sb = 0; sb < nshell; ++sb) {
      const md::Shell& B = sh[sb];
      const int first = (int)ents.size();
      for (int sa = 0; sa < nshell; ++sa) {
        const md::Shell& A = sh[sa];
        if (A.l + B.l != L) continue;
        const double AB[3] = {A.c[0] - B.c[0], A.c[1] - B.c[1], A.c[2] - B.c[2]};
        for (int i = 0; i < A.nprim; ++i) {
          const double a = prim_f64[2 * (A.poff + i)], ca = prim_f64[2 * (A.poff + i) + 1];
          for (int j = 0; j < B.nprim; ++j) {
            const double b = prim_f64[2 * (B.poff + j)], cb = prim_f64[2 * (B.poff + j) + 1];
            const double p = a + b;
            const double KAB = std::exp(-a * b / p * (AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2]));
            if (std::fabs(2.0 * md::PI / p * ca * cb) * KAB < screen) continue;
            Ent en;
            en.aoff = A.aoff;
            en.nfA = A.nfunc;
            en.moff = (long)Mall.size();
            md::pair_matrix(A, B, a, ca, b, cb, tmat_f64, Mall, en.p, en.P);
            ents.push_back(en);
          }
        }
      }
      if ((int)ents.size() > first) bl[L].push_back(make_int4(B.aoff, B.nfunc, first, (int)ents.size()));
    }
  }
  HFPlan* P = new HFPlan();
  P->nshell = nshell;
  P->nao = nao;
  P->screen = screen;
  P->nentries = (long)ents.size();
  P->msize = (long)Mall.size();
  auto fail = [&](hipError_t e, const char* what) {
    set_error(std::string("gdx_hf_plan_create: ") + what + ": " + hipGetErrorString(e));
    plan_free(P);
    return GDX_ERR_HIP;
  };
  hipError_t e = hipGetDevice(&P->device);
  if (e != hipSuccess) return fail(e, "hipGetDevice");
  std::vector<double> ef(4 * std::max<size_t>(ents.size(), 1));
  std::vector<int> ei(2 * std::max<size_t>(ents.size(), 1));
  std::vector<long> em(std::max<size_t>(ents.size(), 1));
  for (size_t k = 0; k < ents.size(); ++k) {
    ef[4 * k] = ents[k].p;
    for (int d = 0; d < 3; ++d) ef[4 * k + 1 + d] = ents[k].P[d];
    ei[2 * k] = ents[k].aoff;
    ei[2 * k + 1] = ents[k].nfA;
    em[k] = ents[k].moff;
  }
  if (Mall.empty()) Mall.push_back(0.0);
#define HF_UPLOAD(dst, vec)                                                                    \
  do {                                                                                         \
    e = hipMalloc((void**)&(dst), (vec).size() * sizeof((vec)[0]));                            \
    if (e != hipSuccess) return fail(e, "hipMalloc");                                          \
    e = hipMemcpy((dst), (vec).data(), (vec).size() * sizeof((vec)[0]), hipMemcpyHostToDevice); \
    if (e != hipSuccess) return fail(e, "hipMemcpy");                                          \
  } while (0)
  HF_UPLOAD(P->d_ef, ef);
  HF_UPLOAD(P->d_ei, ei);
  HF_UPLOAD(P->d_em, em);
  HF_UPLOAD(P->d_M, Mall);
  for (int L = 0; L <= HF_LTOT; ++L) {
    P->nb[L] = (int)bl[L].size();
    if (P->nb[L]) HF_UPLOAD(P->d_bl[L], bl[L]);
  }
#undef HF_UPLOAD
  e = hipMalloc((void**)&P->d_boys, sizeof(double) * md::BOYS_ROWS * BOYS_NCOL);
  if (e != hipSuccess) return fail(e, "hipMalloc");
  hipLaunchKernelGGL(md::boys_table_kernel<BOYS_NCOL>, dim3((md::BOYS_ROWS + 255) / 256), dim3(256), 0,
                     (hipStream_t)0, P->d_boys);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(e, "boys_table_kernel");
  e = hipStreamSynchronize((hipStream_t)0);
  if (e != hipSuccess) return fail(e, "hipStreamSynchronize");
  *plan_out = P;
  return GDX_OK;
}

int gdx_hf_plan_destroy(void* plan) {
  plan_free((HFPlan*)plan);
  return GDX_OK;
}

int gdx_hf_plan_info(const void* plan, int64_t* nentries, int64_t* mdoubles) {
  GDX_REQUIRE(plan, "gdx_hf_plan_info: null plan");
  const HFPlan* P = (const HFPlan*)plan;
  if (nentries) *nentries = P->nentries;
  if (mdoubles) *mdoubles = P->msize;
  return GDX_OK;
}

int gdx_hf_chi(void* stream, const void* plan, int32_t npts, int32_t npts_pad, const double* coords, int32_t nspin,
               const double* u, double omega, double* chi) {
  GDX_REQUIRE(plan && coords && u && chi, "gdx_hf_chi: null pointer");
  GDX_REQUIRE(npts > 0 && npts_pad >= npts, "gdx_hf_chi: bad grid size");
  GDX_REQUIRE(nspin == 1 || nspin == 2, "gdx_hf_chi: nspin must be 1 or 2");
  GDX_REQUIRE(omega >= 0.0, "gdx_hf_chi: omega must be non-negative");
  const HFPlan* P = (const HFPlan*)plan;
  hipStream_t st = (hipStream_t)stream;
  const long len = (long)nspin * P->nao * npts_pad;
  GDX_LAUNCH("hf_zero", hf_zero_kernel, dim3(nblk256(len)), dim3(256), 0, st, chi, len);
  if (nspin == 1)
    launch_chi_all<1>(st, P, npts, npts_pad, coords, omega, u, chi);
  else
    launch_chi_all<2>(st, P, npts, npts_pad, coords, omega, u, chi);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_ao_matmul(void* stream, int32_t npts, int32_t npts_pad, int32_t nao, int32_t nao_pad, const double* phi,
                  int32_t nspin, const double* M, double* out) {
  GDX_REQUIRE(phi && M && out, "gdx_ao_matmul: null pointer");
  GDX_REQUIRE(npts > 0 && npts_pad >= npts && nao > 0 && nao_pad >= nao && nspin > 0, "gdx_ao_matmul: bad shape");
  const dim3 grid((unsigned)((npts_pad + MM_P - 1) / MM_P), (unsigned)((nao + MM_C - 1) / MM_C), (unsigned)nspin);
  GDX_LAUNCH("ao_matmul", ao_matmul_kernel, grid, dim3(256), 0, (hipStream_t)stream, npts, npts_pad, nao, nao_pad, phi,
             M, out);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_hf_energy_density(void* stream, int32_t npts, int32_t npts_pad, int32_t nao, int32_t nspin, const double* chi,
                          const double* v, double* e) {
  GDX_REQUIRE(chi && v && e, "gdx_hf_energy_density: null pointer");
  GDX_REQUIRE(npts > 0 && npts_pad >= npts && nao > 0 && nspin > 0, "gdx_hf_energy_density: bad shape");
  GDX_LAUNCH("hf_rowdot", hf_rowdot_kernel, dim3(nblk256((long)nspin * npts_pad)), dim3(256), 0, (hipStream_t)stream,
             npts, npts_pad, nao, nspin, chi, v, e);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

int gdx_hf_fock_workspace_size(int32_t npts, int32_t nao, int32_t nspin, size_t* bytes) {
  GDX_REQUIRE(bytes && npts > 0 && nao > 0 && nspin > 0, "gdx_hf_fock_workspace_size: bad arguments");
  int nchunk, len;
  kmat_split(npts, nao, nspin, &nchunk, &len);
  *bytes = (size_t)nchunk * nspin * nao * nao * sizeof(double);
  return GDX_OK;
}

int gdx_hf_fock(void* stream, int32_t npts, int32_t npts_pad, int32_t nao, int32_t nao_pad, const double* phi,
                int32_t nspin, const double* chi, const double* g, int64_t g_spin_stride, double scale, double* out,
                void* workspace) {
  GDX_REQUIRE(phi && chi && g && out && workspace, "gdx_hf_fock: null pointer");
  GDX_REQUIRE(npts > 0 && npts_pad >= npts && nao > 0 && nao_pad >= nao && nspin > 0 && g_spin_stride >= 0,
              "gdx_hf_fock: bad shape");
  hipStream_t st = (hipStream_t)stream;
  int nchunk, len;
  kmat_split(npts, nao, nspin, &nchunk, &len);
  const int ntile = (nao + KM_T - 1) / KM_T;
  double* part = (double*)workspace;
  GDX_LAUNCH("hf_kmat", hf_kmat_kernel, dim3((unsigned)(ntile * ntile * nchunk), (unsigned)nspin), dim3(256), 0, st,
             npts, npts_pad, nao, nao_pad, phi, chi, g, (long)g_spin_stride, nchunk, len, part);
  const long n2 = (long)nspin * nao * nao;
  GDX_LAUNCH("hf_kmat_reduce", hf_kmat_reduce_kernel, dim3(nblk256(n2)), dim3(256), 0, st, part, nchunk, n2, scale,
             out);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

}  // extern "C"
