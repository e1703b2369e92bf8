// Electron-repulsion integrals on the device, 8-fold packed.
//
// Reference computation replaced: mol.intor("int2e") as called by calc_eri_with_pyscf,
// grad_dft/interface/pyscf.py:746-803 (the dense (n,n,n,n) tensor the reference keeps in
// Molecule.rep_tensor); here the 8-fold packed form of the host library gdx_int2e_packed:
// pair index I = i(i+1)/2 + j (i >= j), element I(I+1)/2 + K (I >= K).
//
// McMurchie-Davidson: for primitive pairs k (bra, exponent p, centre P) and l (ket, q, Q),
//   (ab|cd) = sum_kl sqrt(pi) / (2 sqrt(p + q))
//             sum_tuv M_k[a][b][tuv] sum_t'u'v' (-1)^(t'+u'+v') M_l[c][d][t'u'v'] R_{t+t',u+u',v+v'}(alpha, P - Q)
// with alpha = p q / (p + q) and M the folded pair matrices of gdx_md.h (each carrying its 2 pi / p).
// Work items: shell-pair quartets (SP1 >= SP2, shell pairs A >= B) grouped by the angular
// momenta (la, lb, lc, ld); one thread per (quartet, a, b, c, d) element, one launch per class,
// templated on la + lb and lc + ld so R and the bra row of M live in registers.  Every packed
// element is written by exactly one thread (the duplicates inside a quartet with A = B, C = D or
// SP1 = SP2 are skipped), so no clearing is needed and the result is reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "gdx_common.h"
#include "gdx_md.h"
#include "gdx_profile.h"

namespace gdx {
namespace {

using md::n_tuv;
using md::tuv_index;

constexpr int ERI_LMAX = 2;                     // shells up to d
constexpr int ERI_LTOT = 4 * ERI_LMAX;          // la + lb + lc + ld
constexpr int ERI_NCOL = ERI_LTOT + md::BOYS_TAYLOR;

template <int LAB, int LCD>
__global__ __launch_bounds__(256) void eri_kernel(const int2* __restrict__ quartets, long nq, int nfA, int nfB,
                                                  int nfC, int nfD, const int4* __restrict__ sp,
                                                  const int* __restrict__ aoff, const double* __restrict__ ef,
                                                  const long* __restrict__ em, const double* __restrict__ M,
                                                  const double* __restrict__ boys, double* __restrict__ eri) {
  constexpr int L = LAB + LCD, NTA = n_tuv(LAB), NTC = n_tuv(LCD);
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nf4 = nfA * nfB * nfC * nfD;
  if (tid >= nq * nf4) return;
  const long qi = tid / nf4;
  int e = (int)(tid - qi * nf4);
  const int d = e % nfD;
  e /= nfD;
  const int c = e % nfC;
  e /= nfC;
  const int b = e % nfB;
  const int a = e / nfB;
  const int2 qq = quartets[qi];
  const int4 s1 = sp[qq.x], s2 = sp[qq.y];
  const long i = aoff[s1.x] + a, j = aoff[s1.y] + b, k = aoff[s2.x] + c, l = aoff[s2.y] + d;
  // one writer per packed element: skip the copies a quartet holds twice
  if (s1.x == s1.y && i < j) return;
  if (s2.x == s2.y && k < l) return;
  const long I = i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
  const long K = k >= l ? k * (k + 1) / 2 + l : l * (l + 1) / 2 + k;
  if (qq.x == qq.y && I < K) return;
  double acc = 0.0;
  for (int e1 = s1.z; e1 < s1.w; ++e1) {
    const double p = ef[4 * e1], Px = ef[4 * e1 + 1], Py = ef[4 * e1 + 2], Pz = ef[4 * e1 + 3];
    const double* Mb = M + em[e1] + (size_t)(a * nfB + b) * NTA;
    double mb[NTA];
#pragma unroll
    for (int t = 0; t < NTA; ++t) mb[t] = Mb[t];
    for (int e2 = s2.z; e2 < s2.w; ++e2) {
      const double q = ef[4 * e2];
      const double X = Px - ef[4 * e2 + 1], Y = Py - ef[4 * e2 + 2], Z = Pz - ef[4 * e2 + 3];
      const double* Mk = M + em[e2] + (size_t)(c * nfD + d) * NTC;
      const double alpha = p * q / (p + q);
      double F[L + 1];
      md::boys_eval<L, ERI_NCOL>(alpha * (X * X + Y * Y + Z * Z), boys, F);
      double R[n_tuv(L)];
      md::hermite_coulomb<L>(alpha, X, Y, Z, F, R);
      double s = 0.0;
#pragma unroll
      for (int t2 = 0; t2 <= LCD; ++t2)
#pragma unroll
        for (int u2 = 0; u2 <= LCD - t2; ++u2)
#pragma unroll
          for (int v2 = 0; v2 <= LCD - t2 - u2; ++v2) {
            double inner = 0.0;
#pragma unroll
            for (int t1 = 0; t1 <= LAB; ++t1)
#pragma unroll
              for (int u1 = 0; u1 <= LAB - t1; ++u1)
#pragma unroll
                for (int v1 = 0; v1 <= LAB - t1 - u1; ++v1)
                  inner += mb[tuv_index(t1, u1, v1)] * R[tuv_index(t1 + t2, u1 + u2, v1 + v2)];
            const double mk = Mk[tuv_index(t2, u2, v2)];
            s += ((t2 + u2 + v2) & 1) ? -mk * inner : mk * inner;
          }
      acc += s * (0.88622692545275801365 / sqrt(p + q));   // sqrt(pi) / 2 / sqrt(p + q)
    }
  }
  const long hi = I >= K ? I : K, lo = I >= K ? K : I;
  eri[hi * (hi + 1) / 2 + lo] = acc;
}

typedef void (*EriKernel)(const int2*, long, int, int, int, int, const int4*, const int*, const double*,
                          const long*, const double*, const double*, double*);

template <int LAB, int LCD>
EriKernel eri_kernel_ptr() {
  return eri_kernel<LAB, LCD>;
}
template <int LAB>
EriKernel eri_kernel_row(int lcd) {
  switch (lcd) {
    case 0: return eri_kernel_ptr<LAB, 0>();
    case 1: return eri_kernel_ptr<LAB, 1>();
    case 2: return eri_kernel_ptr<LAB, 2>();
    case 3: return eri_kernel_ptr<LAB, 3>();
    default: return eri_kernel_ptr<LAB, 4>();
  }
}
EriKernel eri_kernel_for(int lab, int lcd) {
  switch (lab) {
    case 0: return eri_kernel_row<0>(lcd);
    case 1: return eri_kernel_row<1>(lcd);
    case 2: return eri_kernel_row<2>(lcd);
    case 3: return eri_kernel_row<3>(lcd);
    default: return eri_kernel_row<4>(lcd);
  }
}

struct EriClass {
  int la, lb, lc, ld, nfA, nfB, nfC, nfD;
  long q0, nq;   // range in the quartet list
};

struct EriPlan {
  int nshell = 0, nao = 0;
  long npair = 0;
  std::vector<EriClass> classes;
  int2* d_q = nullptr;
  int4* d_sp = nullptr;
  int* d_aoff = nullptr;
  double* d_ef = nullptr;
  long* d_em = nullptr;
  double* d_M = nullptr;
  double* d_boys = nullptr;
  long nquartet = 0, nentries = 0;
};

void eri_free(EriPlan* P) {
  if (!P) return;
  void* ptrs[] = {P->d_q, P->d_sp, P->d_aoff, P->d_ef, P->d_em, P->d_M, P->d_boys};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete P;
}

inline unsigned nblk256(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace gdx

using namespace gdx;

extern "C" {

int gdx_eri_plan_create(int32_t nshell, const int32_t* shell_i32, const double* shell_f64, const double* prim_f64,
                        const double* tmat_f64, int32_t nao, double screen, void** plan_out) {
  GDX_REQUIRE(plan_out && nshell > 0 && nao > 0 && shell_i32 && shell_f64 && prim_f64 && tmat_f64,
              "gdx_eri_plan_create: bad arguments");
  GDX_REQUIRE(screen >= 0.0, "gdx_eri_plan_create: negative screening threshold");
  *plan_out = nullptr;
  std::vector<md::Shell> sh(nshell);
  for (int s = 0; s < nshell; ++s) {
    sh[s] = md::shell_view(shell_i32, shell_f64, s);
    if (sh[s].l > ERI_LMAX) {
      set_error("gdx_eri_plan_create: the device ERI kernels take shells up to d (l <= 2); this basis has l = " +
                std::to_string(sh[s].l));
      return GDX_ERR_UNSUPPORTED;
    }
    GDX_REQUIRE(sh[s].nfunc > 0 && sh[s].aoff + sh[s].nfunc <= nao, "gdx_eri_plan_create: shell outside the basis");
  }
  // shell pairs A >= B with their primitive-pair entries
  std::vector<int4> sp;
  std::vector<double> ef, Mall;
  std::vector<long> em;
  for (int A = 0; A < nshell; ++A)
    for (int B = 0; B <= A; ++B) {
      const int e0 = (int)em.size();
      const md::Shell &SA = sh[A], &SB = sh[B];
      const double R2 = (SA.c[0] - SB.c[0]) * (SA.c[0] - SB.c[0]) + (SA.c[1] - SB.c[1]) * (SA.c[1] - SB.c[1]) +
                        (SA.c[2] - SB.c[2]) * (SA.c[2] - SB.c[2]);
      for (int ia = 0; ia < SA.nprim; ++ia) {
        const double a = prim_f64[2 * (SA.poff + ia)], ca = prim_f64[2 * (SA.poff + ia) + 1];
        for (int jb = 0; jb < SB.nprim; ++jb) {
          const double b = prim_f64[2 * (SB.poff + jb)], cb = prim_f64[2 * (SB.poff + jb) + 1];
          const double p = a + b;
          if (std::fabs(2.0 * md::PI / p * ca * cb) * std::exp(-a * b / p * R2) < screen) continue;
          double pp, P[3];
          em.push_back((long)Mall.size());
          md::pair_matrix(SA, SB, a, ca, b, cb, tmat_f64, Mall, pp, P);
          ef.push_back(pp);
          ef.push_back(P[0]);
          ef.push_back(P[1]);
          ef.push_back(P[2]);
        }
      }
      sp.push_back(make_int4(A, B, e0, (int)em.size()));
    }
  // quartets SP1 >= SP2, grouped by (la, lb, lc, ld)
  std::map<std::tuple<int, int, int, int>, std::vector<int2>> groups;
  const int nsp = (int)sp.size();
  for (int s1 = 0; s1 < nsp; ++s1)
    for (int s2 = 0; s2 <= s1; ++s2) {
      const int4 a = sp[s1], b = sp[s2];
      groups[std::make_tuple(sh[a.x].l, sh[a.y].l, sh[b.x].l, sh[b.y].l)].push_back(make_int2(s1, s2));
    }
  EriPlan* P = new EriPlan();
  P->nshell = nshell;
  P->nao = nao;
  P->npair = (long)nao * (nao + 1) / 2;
  std::vector<int2> qall;
  for (auto& kv : groups) {
    EriClass c;
    std::tie(c.la, c.lb, c.lc, c.ld) = kv.first;
    const int2 q = kv.second.front();
    c.nfA = sh[sp[q.x].x].nfunc;
    c.nfB = sh[sp[q.x].y].nfunc;
    c.nfC = sh[sp[q.y].x].nfunc;
    c.nfD = sh[sp[q.y].y].nfunc;
    c.q0 = (long)qall.size();
    c.nq = (long)kv.second.size();
    qall.insert(qall.end(), kv.second.begin(), kv.second.end());
    P->classes.push_back(c);
  }
  // every shell of one l must have one component count (spherical or Cartesian throughout)
  for (const EriClass& c : P->classes)
    for (long t = c.q0; t < c.q0 + c.nq; ++t) {
      const int2 q = qall[t];
      if (sh[sp[q.x].x].nfunc != c.nfA || sh[sp[q.x].y].nfunc != c.nfB || sh[sp[q.y].x].nfunc != c.nfC ||
          sh[sp[q.y].y].nfunc != c.nfD) {
        eri_free(P);
        set_error("gdx_eri_plan_create: shells of one angular momentum with different component counts");
        return GDX_ERR_UNSUPPORTED;
      }
    }
  std::vector<int> aoff(nshell);
  for (int s = 0; s < nshell; ++s) aoff[s] = sh[s].aoff;
  P->nquartet = (long)qall.size();
  P->nentries = (long)em.size();
  if (Mall.empty()) Mall.push_back(0.0);
  if (ef.empty()) ef.resize(4, 0.0);
  if (em.empty()) em.push_back(0);
  auto fail = [&](hipError_t e, const char* what) {
    set_error(std::string("gdx_eri_plan_create: ") + what + ": " + hipGetErrorString(e));
    eri_free(P);
    return GDX_ERR_HIP;
  };
  hipError_t e;
#define ERI_UPLOAD(dst, vec)                                                                   \
  do {                                                                                         \
    e = hipMalloc((void**)&(dst), (vec).size() * sizeof((vec)[0]));                            \
    if (e != hipSuccess) return fail(e, "hipMalloc");                                          \
    e = hipMemcpy((dst), (vec).data(), (vec).size() * sizeof((vec)[0]), hipMemcpyHostToDevice); \
    if (e != hipSuccess) return fail(e, "hipMemcpy");                                          \
  } while (0)
  ERI_UPLOAD(P->d_q, qall);
  ERI_UPLOAD(P->d_sp, sp);
  ERI_UPLOAD(P->d_aoff, aoff);
  ERI_UPLOAD(P->d_ef, ef);
  ERI_UPLOAD(P->d_em, em);
  ERI_UPLOAD(P->d_M, Mall);
#undef ERI_UPLOAD
  e = hipMalloc((void**)&P->d_boys, sizeof(double) * md::BOYS_ROWS * ERI_NCOL);
  if (e != hipSuccess) return fail(e, "hipMalloc");
  hipLaunchKernelGGL(md::boys_table_kernel<ERI_NCOL>, dim3((md::BOYS_ROWS + 255) / 256), dim3(256), 0,
                     (hipStream_t)0, P->d_boys);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(e, "boys_table_kernel");
  e = hipStreamSynchronize((hipStream_t)0);
  if (e != hipSuccess) return fail(e, "hipStreamSynchronize");
  *plan_out = P;
  return GDX_OK;
}

int gdx_eri_plan_destroy(void* plan) {
  eri_free((EriPlan*)plan);
  return GDX_OK;
}

int gdx_eri_plan_info(const void* plan, int64_t* nquartets, int64_t* nentries) {
  GDX_REQUIRE(plan, "gdx_eri_plan_info: null plan");
  const EriPlan* P = (const EriPlan*)plan;
  if (nquartets) *nquartets = P->nquartet;
  if (nentries) *nentries = P->nentries;
  return GDX_OK;
}

int gdx_eri_packed(void* stream, const void* plan, double* eri) {
  GDX_REQUIRE(plan && eri, "gdx_eri_packed: null pointer");
  const EriPlan* P = (const EriPlan*)plan;
  hipStream_t st = (hipStream_t)stream;
  for (const EriClass& c : P->classes) {
    const long nthr = c.nq * c.nfA * c.nfB * c.nfC * c.nfD;
    EriKernel k = eri_kernel_for(c.la + c.lb, c.lc + c.ld);
    GDX_LAUNCH("eri", k, dim3(nblk256(nthr)), dim3(256), 0, st, P->d_q + c.q0, c.nq, c.nfA, c.nfB, c.nfC, c.nfD,
               P->d_sp, P->d_aoff, P->d_ef, P->d_em, P->d_M, P->d_boys, eri);
  }
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

}  // extern "C"
