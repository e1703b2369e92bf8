// AO evaluation on the grid (value, gradient, laplacian) and the library's error state.
// Replaces what the reference receives from PySCF numint.eval_ao(deriv=1) and
// ao_grads(order=2) (grad_dft/interface/pyscf.py:824-827, 677-705).
//
// A block owns a tile of P consecutive grid points (P chosen so the tile of all five planes,
// 5 x P x nao_pad doubles, fits in 64 KB of LDS).  Its threads take the (point, shell) items
// of the tile: each contracts the shell's primitives into R0 = sum c e^{-a r^2},
// R1 = sum c a e^{-a r^2}, R2 = sum c a^2 e^{-a r^2}, forms every Cartesian monomial P with
// its first and pure second derivatives
//   d/dx (P R) = P_x R0 - 2 x P R1
//   d2/dx2 (P R) = P_xx R0 - 2 (P + 2 x P_x) R1 + 4 x^2 P R2
// maps monomials to AO components with the shell's component matrix T, and writes the values
// into the LDS tile.  The tile (zero where no shell writes: padding columns, padding rows) is
// then stored with 16-byte writes of whole rows, so every plane element is written once, by
// full-line stores, and nothing needs clearing beforehand.
#include "gdx_common.h"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace gdx {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

namespace {

constexpr int MAXL = 3;
constexpr int MAXC = 10;   // Cartesian monomials of an f shell
constexpr int MAXF = 10;   // AO components of a shell (Cartesian f)

__device__ __forceinline__ double ipow(double x, int k) {
  double r = 1.0;
  for (int i = 0; i < k; ++i) r *= x;
  return r;
}

__global__ __launch_bounds__(256) void ao_eval_kernel(
    int nshell, const int* __restrict__ si, const double* __restrict__ sf,
    const double* __restrict__ prim, const double* __restrict__ tmat, int npts, int npts_pad,
    const double* __restrict__ coords, int nao_pad, int P, double* __restrict__ ao) {
  extern __shared__ double tile[];   // [5][P][nao_pad]
  const int tid = threadIdx.x;
  const int p0 = blockIdx.x * P;
  const int tsz = 5 * P * nao_pad;
  for (int i = tid; i < tsz; i += 256) tile[i] = 0.0;
  __syncthreads();
  const int items = P * nshell;
  for (int it = tid; it < items; it += 256) {
    const int pl = it / nshell, s = it - pl * nshell;
    const int p = p0 + pl;
    if (p >= npts) continue;
    const int l = si[6 * s + 0], nprim = si[6 * s + 1], poff = si[6 * s + 2];
    const int aoff = si[6 * s + 3], nfunc = si[6 * s + 4], toff = si[6 * s + 5];
    const double dx = coords[3 * p + 0] - sf[3 * s + 0];
    const double dy = coords[3 * p + 1] - sf[3 * s + 1];
    const double dz = coords[3 * p + 2] - sf[3 * s + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    double R0 = 0.0, R1 = 0.0, R2 = 0.0;
    for (int i = 0; i < nprim; ++i) {
      const double a = prim[2 * (poff + i)];
      const double c = prim[2 * (poff + i) + 1] * exp(-a * r2);
      R0 += c;
      R1 += c * a;
      R2 += c * a * a;
    }
    double out[5][MAXF];
#pragma unroll
    for (int f = 0; f < MAXF; ++f)
#pragma unroll
      for (int q = 0; q < 5; ++q) out[q][f] = 0.0;
    const double d[3] = {dx, dy, dz};
    const int nc = (l + 1) * (l + 2) / 2;
    int k = 0;
    for (int ea = l; ea >= 0; --ea) {
      for (int eb = l - ea; eb >= 0; --eb, ++k) {
        const int e[3] = {ea, eb, l - ea - eb};
        double pw[3], p1[3], p2[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          pw[j] = ipow(d[j], e[j]);
          p1[j] = e[j] > 0 ? e[j] * ipow(d[j], e[j] - 1) : 0.0;
          p2[j] = e[j] > 1 ? e[j] * (e[j] - 1) * ipow(d[j], e[j] - 2) : 0.0;
        }
        const double Pm = pw[0] * pw[1] * pw[2];
        const double Pd[3] = {p1[0] * pw[1] * pw[2], pw[0] * p1[1] * pw[2], pw[0] * pw[1] * p1[2]};
        const double Pdd[3] = {p2[0] * pw[1] * pw[2], pw[0] * p2[1] * pw[2], pw[0] * pw[1] * p2[2]};
        const double v0 = Pm * R0;
        double g[3], lap = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          g[j] = Pd[j] * R0 - 2.0 * d[j] * Pm * R1;
          lap += Pdd[j] * R0 - 2.0 * (Pm + 2.0 * d[j] * Pd[j]) * R1 + 4.0 * d[j] * d[j] * Pm * R2;
        }
#pragma unroll
        for (int f = 0; f < MAXF; ++f) {
          if (f < nfunc) {   // register indices fixed at compile time
            const double t = tmat[toff + f * nc + k];
            out[0][f] += t * v0;
            out[1][f] += t * g[0];
            out[2][f] += t * g[1];
            out[3][f] += t * g[2];
            out[4][f] += t * lap;
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
      for (int f = 0; f < MAXF; ++f)
        if (f < nfunc) tile[((size_t)q * P + pl) * nao_pad + aoff + f] = out[q][f];
  }
  __syncthreads();
  // whole rows of the five planes, 16 bytes a thread per step (nao_pad is even)
  const size_t plane = (size_t)npts_pad * nao_pad;
  const int nv = nao_pad / 2;
  const int nrows = min(P, npts_pad - p0);
  for (int i = tid; i < 5 * nrows * nv; i += 256) {
    const int v = i % nv, r = i / nv;
    const int pl = r % nrows, q = r / nrows;
    const double2 x = *(const double2*)&tile[((size_t)q * P + pl) * nao_pad + 2 * v];
    *(double2*)(ao + q * plane + (size_t)(p0 + pl) * nao_pad + 2 * v) = x;
  }
}

}  // namespace
}  // namespace gdx

extern "C" {

const char* gdx_last_error(void) { return gdx::g_last_error.c_str(); }
int gdx_version(void) { return 1; }

int gdx_ao_eval(void* stream, int32_t nshell, const int32_t* shell_i32, const double* shell_f64,
                const double* prim_f64, const double* tmat_f64, int32_t npts, int32_t npts_pad,
                const double* coords, int32_t nao, int32_t nao_pad, double* ao) {
  GDX_REQUIRE(nshell > 0 && npts > 0 && nao > 0, "gdx_ao_eval: empty basis or grid");
  GDX_REQUIRE(npts_pad >= npts && nao_pad >= nao, "gdx_ao_eval: padding smaller than shape");
  GDX_REQUIRE(shell_i32 && shell_f64 && prim_f64 && tmat_f64 && coords && ao,
              "gdx_ao_eval: null pointer");
  GDX_REQUIRE(nao_pad % 2 == 0 && nao_pad <= 1600, "gdx_ao_eval: nao_pad must be even and at most 1600");
  hipStream_t st = (hipStream_t)stream;
  // points per block: the five-plane tile in at most 64 KB of LDS
  const int P = std::max(1, std::min(32, (int)(65536 / (5 * 8 * (long)nao_pad))));
  const long grid = ((long)npts_pad + P - 1) / P;
  GDX_REQUIRE(grid < (1L << 31), "gdx_ao_eval: grid too large");
  const size_t lds = sizeof(double) * 5 * (size_t)P * nao_pad;
  hipLaunchKernelGGL(gdx::ao_eval_kernel, dim3((unsigned)grid), dim3(256), lds, st, nshell, shell_i32, shell_f64,
                     prim_f64, tmat_f64, npts, npts_pad, coords, nao_pad, P, ao);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}

}  // extern "C"

// ------------------------------------------------------------ kernel timing ----
#include "gdx_profile.h"
#include <map>
#include <vector>

namespace gdx {
namespace {
struct ProfEntry {
  std::string name;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfEntry> g_prof;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t take_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

bool prof_active() { return g_prof_on; }

int prof_start(const char* name, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_prof_on) return -1;
  ProfEntry e{name, take_event(), take_event()};
  if (!e.a || !e.b) return -1;
  (void)hipEventRecord(e.a, st);
  g_prof.push_back(e);
  return (int)g_prof.size() - 1;
}

void prof_stop(int slot, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (slot < 0 || slot >= (int)g_prof.size()) return;
  (void)hipEventRecord(g_prof[slot].b, st);
}

}  // namespace gdx

extern "C" {

int gdx_profile_begin(void) {
  std::lock_guard<std::mutex> lk(gdx::g_prof_mu);
  gdx::g_prof.clear();
  gdx::g_prof_on = true;
  return GDX_OK;
}

// Stops recording, waits for the recorded events and aggregates durations by kernel
// name into names (max_entries x 64 chars), total_ms, counts; *n_out = kernels found.
int gdx_profile_end(int32_t max_entries, char* names, double* total_ms, int32_t* counts,
                    int32_t* n_out) {
  std::lock_guard<std::mutex> lk(gdx::g_prof_mu);
  gdx::g_prof_on = false;
  std::map<std::string, std::pair<double, int>> agg;
  std::vector<std::string> order;
  for (auto& e : gdx::g_prof) {
    GDX_HIP_CHECK(hipEventSynchronize(e.b));
    float ms = 0.f;
    GDX_HIP_CHECK(hipEventElapsedTime(&ms, e.a, e.b));
    if (!agg.count(e.name)) order.push_back(e.name);
    agg[e.name].first += ms;
    agg[e.name].second += 1;
    gdx::g_event_pool.push_back(e.a);
    gdx::g_event_pool.push_back(e.b);
  }
  gdx::g_prof.clear();
  int n = 0;
  for (auto& k : order) {
    if (n >= max_entries) break;
    std::strncpy(names + 64 * n, k.c_str(), 63);
    names[64 * n + 63] = 0;
    total_ms[n] = agg[k].first;
    counts[n] = agg[k].second;
    ++n;
  }
  *n_out = n;
  return GDX_OK;
}

}  // extern "C"
