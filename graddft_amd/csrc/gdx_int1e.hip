// One-electron integrals on the device: overlap S, kinetic T and nuclear attraction V.
//
// Reference computations replaced: mf.get_ovlp() and mf.get_hcore() = T + V as read by
// _package_outputs, grad_dft/interface/pyscf.py:828-829 (s1e, h1e of the Molecule).
//
// For a primitive pair (exponents a on A, b on B; p = a + b, P = (a A + b B) / p, mu = a b / p)
// and one Cartesian direction, the Obara-Saika overlap recursion
//   S_00 = 1,  S_{i+1,j} = X_PA S_ij + (i S_{i-1,j} + j S_{i,j-1}) / (2p)   (likewise j with X_PB)
// gives the 1-D overlaps up to the factor exp(-mu X_AB^2) sqrt(pi / p), applied once for all
// three directions as exp(-mu |AB|^2) (pi / p)^(3/2).  The kinetic integral acts on the B side,
//   K_ij = -2 b^2 S_{i,j+2} + b (2j + 1) S_ij - j (j - 1) / 2 S_{i,j-2},
// and T = Kx Sy Sz + Sx Ky Sz + Sx Sy Kz.  V = -sum_C Z_C sum_tuv M[a][b][tuv] R_tuv(p, P - C)
// with the host-folded McMurchie-Davidson pair matrices M of gdx_md.h (which carry 2 pi / p and
// the contraction coefficients).
//
// One thread per (shell pair A >= B, AO component a of A, b of B); per (l_A, l_B) class one
// launch for S and T and one for V, with their tables in registers.  Each element of the symmetric outputs is written by one
// thread (for A = B only a >= b writes), so nothing needs clearing and results are
// reproducible.
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "gdx_common.h"
#include "gdx_md.h"
#include "gdx_profile.h"

namespace gdx {
namespace {

using md::n_tuv;
using md::tuv_index;

constexpr int I1_LMAX = md::LMAX;                              // f shells
constexpr int I1_NCOL = 2 * I1_LMAX + md::BOYS_TAYLOR;
constexpr int I1_TPB = 128;

__host__ __device__ constexpr int ncart_c(int l) { return (l + 1) * (l + 2) / 2; }

template <int LA, int LB>
__global__ __launch_bounds__(I1_TPB) void int1e_st_kernel(const int* __restrict__ pairs, int npairs, int nfA,
                                                          int nfB, const int4* __restrict__ sp,
                                                          const double* __restrict__ centre,
                                                          const int2* __restrict__ sh_off,
                                                          const double* __restrict__ tmat,
                                                          const double* __restrict__ ef,
                                                          const double* __restrict__ eab, int nao,
                                                          double* __restrict__ S, double* __restrict__ T) {
  constexpr int NCA = ncart_c(LA), NCB = ncart_c(LB), NJ = LB + 3;
  static_assert(NCA <= md::MAXF && NCB <= md::MAXF, "component count");
  const int tid = blockIdx.x * I1_TPB + threadIdx.x;
  const int nf2 = nfA * nfB;
  if (tid >= npairs * nf2) return;
  const int pi = tid / nf2;
  const int fa = (tid - pi * nf2) / nfB, fb = tid - pi * nf2 - fa * nfB;
  const int4 s = sp[pairs[pi]];
  if (s.x == s.y && fa < fb) return;
  const double A[3] = {centre[3 * s.x], centre[3 * s.x + 1], centre[3 * s.x + 2]};
  const double B[3] = {centre[3 * s.y], centre[3 * s.y + 1], centre[3 * s.y + 2]};
  const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
  const double* TA = tmat + sh_off[s.x].y + fa * NCA;
  const double* TB = tmat + sh_off[s.y].y + fb * NCB;
  double accS = 0.0, accT = 0.0;
  for (int e = s.z; e < s.w; ++e) {
    const double p = ef[4 * e], P[3] = {ef[4 * e + 1], ef[4 * e + 2], ef[4 * e + 3]};
    const double a = eab[3 * e], b = eab[3 * e + 1], cab = eab[3 * e + 2];
    const double o2p = 0.5 / p;
    // 1-D overlaps S1[d][i][j], i <= LA, j <= LB + 2
    double S1[3][LA + 1][NJ];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double XPA = P[d] - A[d], XPB = P[d] - B[d];
      S1[d][0][0] = 1.0;
#pragma unroll
      for (int j = 1; j < NJ; ++j)
        S1[d][0][j] = XPB * S1[d][0][j - 1] + (j > 1 ? (j - 1) * o2p * S1[d][0][j - 2] : 0.0);
#pragma unroll
      for (int i = 1; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          S1[d][i][j] = XPA * S1[d][i - 1][j] + (i > 1 ? (i - 1) * o2p * S1[d][i - 2][j] : 0.0) +
                        (j > 0 ? j * o2p * S1[d][i - 1][j - 1] : 0.0);
    }
    // kinetic factor of direction d
    auto kin = [&](int d, int i, int j) {
      return -2.0 * b * b * S1[d][i][j + 2] + b * (2 * j + 1) * S1[d][i][j] -
             (j > 1 ? 0.5 * j * (j - 1) * S1[d][i][j - 2] : 0.0);
    };
    double cs = 0.0, ct = 0.0;
    // Cartesian components in the order of the component matrices: x descending, then y
#pragma unroll
    for (int ix = LA, ia = 0; ix >= 0; --ix)
#pragma unroll
      for (int iy = LA - ix; iy >= 0; --iy, ++ia) {
        const int iz = LA - ix - iy;
        const double ta = TA[ia];
#pragma unroll
        for (int jx = LB, jb = 0; jx >= 0; --jx)
#pragma unroll
          for (int jy = LB - jx; jy >= 0; --jy, ++jb) {
            const int jz = LB - jx - jy;
            const double w = ta * TB[jb];
            const double sx = S1[0][ix][jx], sy = S1[1][iy][jy], sz = S1[2][iz][jz];
            cs += w * (sx * sy * sz);
            ct += w * (kin(0, ix, jx) * sy * sz + sx * kin(1, iy, jy) * sz + sx * sy * kin(2, iz, jz));
          }
      }
    const double pref = cab * exp(-a * b / p * AB2) * (md::PI / p) * sqrt(md::PI / p);
    accS += pref * cs;
    accT += pref * ct;
  }
  const long I = sh_off[s.x].x + fa, J = sh_off[s.y].x + fb;
  S[I * nao + J] = accS;
  S[J * nao + I] = accS;
  T[I * nao + J] = accT;
  T[J * nao + I] = accT;
}

// V for the shell pairs of one L = l_A + l_B: the folded pair matrices carry every angular
// detail, so one kernel per L serves all (l_A, l_B) with that sum.
template <int L>
__global__ __launch_bounds__(I1_TPB) void int1e_v_kernel(const int* __restrict__ pairs, int npairs, int nfA, int nfB,
                                                         const int4* __restrict__ sp, const int2* __restrict__ sh_off,
                                                         const double* __restrict__ ef, const long* __restrict__ em,
                                                         const double* __restrict__ M, int natm,
                                                         const double* __restrict__ nuc,
                                                         const double* __restrict__ boys, int nao,
                                                         double* __restrict__ V) {
  constexpr int NT = n_tuv(L);
  const int tid = blockIdx.x * I1_TPB + threadIdx.x;
  const int nf2 = nfA * nfB;
  if (tid >= npairs * nf2) return;
  const int pi = tid / nf2;
  const int fa = (tid - pi * nf2) / nfB, fb = tid - pi * nf2 - fa * nfB;
  const int4 s = sp[pairs[pi]];
  if (s.x == s.y && fa < fb) return;
  double accV = 0.0;
  for (int e = s.z; e < s.w; ++e) {
    const double p = ef[4 * e], Px = ef[4 * e + 1], Py = ef[4 * e + 2], Pz = ef[4 * e + 3];
    const double* Mab = M + em[e] + (size_t)(fa * nfB + fb) * NT;
    for (int c = 0; c < natm; ++c) {
      const double X = Px - nuc[4 * c], Y = Py - nuc[4 * c + 1], Z = Pz - nuc[4 * c + 2];
      double F[L + 1];
      md::boys_eval<L, I1_NCOL>(p * (X * X + Y * Y + Z * Z), boys, F);
      double R[NT];
      md::hermite_coulomb<L>(p, X, Y, Z, F, R);
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < NT; ++q) v += Mab[q] * R[q];
      accV -= nuc[4 * c + 3] * v;
    }
  }
  const long I = sh_off[s.x].x + fa, J = sh_off[s.y].x + fb;
  V[I * nao + J] = accV;
  V[J * nao + I] = accV;
}

typedef void (*StKernel)(const int*, int, int, int, const int4*, const double*, const int2*, const double*,
                         const double*, const double*, int, double*, double*);
typedef void (*VKernel)(const int*, int, int, int, const int4*, const int2*, const double*, const long*,
                        const double*, int, const double*, const double*, int, double*);

template <int LA>
StKernel st_row(int lb) {
  switch (lb) {
    case 0: return int1e_st_kernel<LA, 0>;
    case 1: return int1e_st_kernel<LA, 1>;
    case 2: return int1e_st_kernel<LA, 2>;
    default: return int1e_st_kernel<LA, 3>;
  }
}
StKernel st_for(int la, int lb) {
  switch (la) {
    case 0: return st_row<0>(lb);
    case 1: return st_row<1>(lb);
    case 2: return st_row<2>(lb);
    default: return st_row<3>(lb);
  }
}
VKernel v_for(int L) {
  switch (L) {
    case 0: return int1e_v_kernel<0>;
    case 1: return int1e_v_kernel<1>;
    case 2: return int1e_v_kernel<2>;
    case 3: return int1e_v_kernel<3>;
    case 4: return int1e_v_kernel<4>;
    case 5: return int1e_v_kernel<5>;
    default: return int1e_v_kernel<6>;
  }
}

struct Int1eClass {
  int la, lb, nfA, nfB, p0, np;   // range in the pair list
};

struct Int1ePlan {
  int nao = 0, natm = 0, nshell = 0;
  std::vector<Int1eClass> classes;
  int* d_pairs = nullptr;
  int4* d_sp = nullptr;
  double* d_centre = nullptr;
  int2* d_off = nullptr;
  double* d_tmat = nullptr;
  double* d_ef = nullptr;
  double* d_eab = nullptr;
  long* d_em = nullptr;
  double* d_M = nullptr;
  double* d_nuc = nullptr;
  double* d_boys = nullptr;
};

void int1e_free(Int1ePlan* P) {
  if (!P) return;
  void* ptrs[] = {P->d_pairs, P->d_sp, P->d_centre, P->d_off, P->d_tmat, P->d_ef,
                  P->d_eab,   P->d_em, P->d_M,      P->d_nuc, P->d_boys};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete P;
}

}  // namespace
}  // namespace gdx

using namespace gdx;

extern "C" {

int gdx_int1e_plan_create(int32_t nshell, const int32_t* shell_i32, const double* shell_f64, const double* prim_f64,
                          const double* tmat_f64, int32_t ntmat, int32_t nao, int32_t natm, const double* charges,
                          const double* atom_xyz, void** plan_out) {
  GDX_REQUIRE(plan_out && nshell > 0 && nao > 0 && natm > 0 && ntmat > 0 && shell_i32 && shell_f64 && prim_f64 &&
                  tmat_f64 && charges && atom_xyz,
              "gdx_int1e_plan_create: bad arguments");
  *plan_out = nullptr;
  std::vector<md::Shell> sh(nshell);
  for (int s = 0; s < nshell; ++s) {
    sh[s] = md::shell_view(shell_i32, shell_f64, s);
    if (sh[s].l < 0 || sh[s].l > I1_LMAX) {
      set_error("gdx_int1e_plan_create: shells up to f (l <= 3); this basis has l = " + std::to_string(sh[s].l));
      return GDX_ERR_UNSUPPORTED;
    }
    GDX_REQUIRE(sh[s].nfunc > 0 && sh[s].aoff >= 0 && sh[s].aoff + sh[s].nfunc <= nao,
                "gdx_int1e_plan_create: shell outside the basis");
    GDX_REQUIRE(sh[s].toff >= 0 && sh[s].toff + sh[s].nfunc * md::ncart(sh[s].l) <= ntmat,
                "gdx_int1e_plan_create: component matrix outside tmat");
  }
  std::vector<int4> sp;
  std::vector<double> ef, eab, Mall;
  std::vector<long> em;
  std::map<std::pair<int, int>, std::vector<int>> groups;
  for (int A = 0; A < nshell; ++A)
    for (int B = 0; B <= A; ++B) {
      const int e0 = (int)em.size();
      const md::Shell &SA = sh[A], &SB = sh[B];
      for (int ia = 0; ia < SA.nprim; ++ia) {
        const double a = prim_f64[2 * (SA.poff + ia)], ca = prim_f64[2 * (SA.poff + ia) + 1];
        for (int jb = 0; jb < SB.nprim; ++jb) {
          const double b = prim_f64[2 * (SB.poff + jb)], cb = prim_f64[2 * (SB.poff + jb) + 1];
          double pp, P[3];
          em.push_back((long)Mall.size());
          md::pair_matrix(SA, SB, a, ca, b, cb, tmat_f64, Mall, pp, P);
          ef.insert(ef.end(), {pp, P[0], P[1], P[2]});
          eab.insert(eab.end(), {a, b, ca * cb});
        }
      }
      groups[std::make_pair(SA.l, SB.l)].push_back((int)sp.size());
      sp.push_back(make_int4(A, B, e0, (int)em.size()));
    }
  Int1ePlan* P = new Int1ePlan();
  P->nao = nao;
  P->natm = natm;
  P->nshell = nshell;
  std::vector<int> pairs;
  for (auto& kv : groups) {
    Int1eClass c;
    c.la = kv.first.first;
    c.lb = kv.first.second;
    c.nfA = sh[sp[kv.second.front()].x].nfunc;
    c.nfB = sh[sp[kv.second.front()].y].nfunc;
    for (int q : kv.second)
      if (sh[sp[q].x].nfunc != c.nfA || sh[sp[q].y].nfunc != c.nfB) {
        int1e_free(P);
        set_error("gdx_int1e_plan_create: shells of one angular momentum with different component counts");
        return GDX_ERR_UNSUPPORTED;
      }
    c.p0 = (int)pairs.size();
    c.np = (int)kv.second.size();
    pairs.insert(pairs.end(), kv.second.begin(), kv.second.end());
    P->classes.push_back(c);
  }
  std::vector<double> centre(3 * (size_t)nshell);
  std::vector<int2> off(nshell);
  for (int s = 0; s < nshell; ++s) {
    for (int k = 0; k < 3; ++k) centre[3 * s + k] = sh[s].c[k];
    off[s] = make_int2(sh[s].aoff, sh[s].toff);
  }
  std::vector<double> nuc(4 * (size_t)natm), tmat(tmat_f64, tmat_f64 + ntmat);
  for (int c = 0; c < natm; ++c) {
    for (int k = 0; k < 3; ++k) nuc[4 * c + k] = atom_xyz[3 * c + k];
    nuc[4 * c + 3] = charges[c];
  }
  auto fail = [&](hipError_t e, const char* what) {
    set_error(std::string("gdx_int1e_plan_create: ") + what + ": " + hipGetErrorString(e));
    int1e_free(P);
    return GDX_ERR_HIP;
  };
  hipError_t e;
#define I1_UPLOAD(dst, vec)                                                                    \
  do {                                                                                         \
    e = hipMalloc((void**)&(dst), (vec).size() * sizeof((vec)[0]));                            \
    if (e != hipSuccess) return fail(e, "hipMalloc");                                          \
    e = hipMemcpy((dst), (vec).data(), (vec).size() * sizeof((vec)[0]), hipMemcpyHostToDevice); \
    if (e != hipSuccess) return fail(e, "hipMemcpy");                                          \
  } while (0)
  I1_UPLOAD(P->d_pairs, pairs);
  I1_UPLOAD(P->d_sp, sp);
  I1_UPLOAD(P->d_centre, centre);
  I1_UPLOAD(P->d_off, off);
  I1_UPLOAD(P->d_tmat, tmat);
  I1_UPLOAD(P->d_ef, ef);
  I1_UPLOAD(P->d_eab, eab);
  I1_UPLOAD(P->d_em, em);
  I1_UPLOAD(P->d_M, Mall);
  I1_UPLOAD(P->d_nuc, nuc);
#undef I1_UPLOAD
  e = hipMalloc((void**)&P->d_boys, sizeof(double) * md::BOYS_ROWS * I1_NCOL);
  if (e != hipSuccess) return fail(e, "hipMalloc");
  hipLaunchKernelGGL(md::boys_table_kernel<I1_NCOL>, dim3((md::BOYS_ROWS + 255) / 256), dim3(256), 0, (hipStream_t)0,
                     P->d_boys);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(e, "boys_table_kernel");
  e = hipStreamSynchronize((hipStream_t)0);
  if (e != hipSuccess) return fail(e, "hipStreamSynchronize");
  *plan_out = P;
  return GDX_OK;
}

int gdx_int1e_plan_destroy(void* plan) {
  int1e_free((Int1ePlan*)plan);
  return GDX_OK;
}

int gdx_int1e_eval(void* stream, const void* plan, double* S, double* T, double* V) {
  GDX_REQUIRE(plan && S && T && V, "gdx_int1e_eval: null pointer");
  const Int1ePlan* P = (const Int1ePlan*)plan;
  hipStream_t st = (hipStream_t)stream;
  for (const Int1eClass& c : P->classes) {
    const long nthr = (long)c.np * c.nfA * c.nfB;
    const dim3 grid((unsigned)((nthr + I1_TPB - 1) / I1_TPB));
    GDX_LAUNCH("int1e_st", st_for(c.la, c.lb), grid, dim3(I1_TPB), 0, st, P->d_pairs + c.p0, c.np, c.nfA, c.nfB,
               P->d_sp, P->d_centre, P->d_off, P->d_tmat, P->d_ef, P->d_eab, P->nao, S, T);
    GDX_LAUNCH("int1e_v", v_for(c.la + c.lb), grid, dim3(I1_TPB), 0, st, P->d_pairs + c.p0, c.np, c.nfA, c.nfB,
               P->d_sp, P->d_off, P->d_ef, P->d_em, P->d_M, P->natm, P->d_nuc, P->d_boys, P->nao, V);
  }
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

}  // extern "C"
